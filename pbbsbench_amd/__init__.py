"""MI355X (gfx950) deterministic breadth-first search, a drop-in for PBBS
breadthFirstSearch/deterministicBFS, with its siblings on the same graphs
(maximalIndependentSet's incrementalMIS, spanningForest's incrementalST).  The
work runs in libpbbs_bfs.so (HIP kernels + host engine) behind the C ABI of
include/pbbs_bfs.h."""
from ._lib import PbbsError, LIB_PATH  # noqa: F401
from .core import (  # noqa: F401
    BFS,
    Context,
    DeviceEdges,
    DeviceGraph,
    HostGraph,
    Server,
    from_edges,
    grid3d,
    permute,
    randlocal,
    read_edge_array_from_file,
    read_graph_binary,
    read_graph_from_file,
    read_int_seq_from_file,
    rmat,
    write_edge_array_to_file,
    write_graph_binary,
    write_graph_to_file,
    write_sequence_to_file,
)

__all__ = [
    "BFS", "Context", "DeviceEdges", "DeviceGraph", "HostGraph", "PbbsError", "Server",
    "from_edges", "grid3d", "permute", "randlocal", "read_edge_array_from_file",
    "read_graph_binary", "read_graph_from_file", "read_int_seq_from_file", "rmat",
    "write_edge_array_to_file", "write_graph_binary", "write_graph_to_file",
    "write_sequence_to_file",
]
