/*
 * pbbs_bfs.h -- C ABI of the MI355X (gfx950) breadth-first search.
 *
 * Drop-in for PBBS breadthFirstSearch/deterministicBFS.  The reference binds
 * its implementation at link time through one C++ function
 *
 *     parlay::sequence<vertexId> BFS(vertexId start, const Graph& G, bool verbose)
 *     (benchmarks/breadthFirstSearch/bench/BFS.h:4-11,
 *      defined benchmarks/breadthFirstSearch/deterministicBFS/BFS.C:39-78)
 *
 * and reads its graph through readGraphFromFile (common/graphIO.h:191-215)
 * into graph<int,uint> (common/graph.h:104-152).  This header replaces that
 * pair with plain C entry points: a context (one GPU), a graph resident in
 * HBM, and BFS calls whose outputs are bit-identical to deterministicBFS:
 *
 *     parent[s] = s, parent[v] = -1 if v is unreached, otherwise
 *     parent[v] = the LARGEST vertex id u with (u,v) an edge and
 *                 depth(u) = depth(v) - 1
 *     (the write_max claim of deterministicBFS/BFS.C:65-68 with
 *      common/atomics.h:75-81, followed by the winner filter BFS.C:71-74)
 *
 *     depth[s] = 0, depth[v] = n if v is unreached, otherwise the hop count
 *     (the level_number convention of bench/BFSCheck.C:39-56).
 *
 * Conventions: every function returns PBBS_OK (0) or a negative PBBS_ERR_*
 * code; pbbs_last_error() gives a message for the calling thread.  No C++
 * exception crosses this boundary.  A context is used by one host thread at a
 * time (as the reference's BFS() is called from one thread,
 * bench/BFSTime.C:38-41); the server object serialises its own work.
 * Vertex ids are int32 (the reference's signed vertexId, bench/BFS.h:4);
 * offsets are uint32 (the reference's edgeId = uint, bench/BFS.h:6) or int64.
 */
#ifndef PBBS_BFS_H_
#define PBBS_BFS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PBBS_OK              0
#define PBBS_ERR_INVALID    -1  /* bad argument (null pointer, source out of range, ...) */
#define PBBS_ERR_HIP        -2  /* a HIP runtime call failed */
#define PBBS_ERR_NOMEM      -3  /* device or host allocation failed */
#define PBBS_ERR_LIMIT      -4  /* size beyond a documented limit */
#define PBBS_ERR_FORMAT     -5  /* malformed input file */
#define PBBS_ERR_IO         -6  /* file could not be opened / written */
#define PBBS_ERR_STATE      -7  /* call not valid in the object's current state */

/* Message for the last failing call on this thread ("" if none). */
const char* pbbs_last_error(void);

/* Library version string, e.g. "pbbs-bfs-gfx950 0.1". */
const char* pbbs_version(void);

/* ---- context ----------------------------------------------------------- */
typedef struct pbbs_ctx pbbs_ctx;

/* Opens `device` (HIP ordinal), creates the context's stream.  Replaces the
 * implicit global parlay scheduler the reference's BFS() runs on
 * (common/parallelDefs:11-28). */
int pbbs_ctx_create(int device, pbbs_ctx** out);
int pbbs_ctx_destroy(pbbs_ctx* ctx);

/* Direction-optimising switch and launch shape.  Defaults follow ligraLight's
 * edge_map rule (common/ligraLight.h:120-133): go bottom-up when
 * |F| + sum deg(F) > m / td_to_bu_div, stay bottom-up while |F| > n / bu_to_td_div. */
typedef struct {
    int32_t td_to_bu_div;   /* default 20; 0 disables bottom-up entirely */
    int32_t bu_to_td_div;   /* default 20 */
    int32_t grid_blocks;    /* workgroups per level launch; 0 = auto (occupancy x CUs) */
    int32_t bu_serial_depth;/* row entries a lane tests alone (bottom-up) before its
                               wavefront takes over the row; default 128 */
    int32_t chunk_levels;   /* levels per captured HIP graph launch; default 16.  After
                               the first run on a graph, a run's first chunk is that
                               graph's previous launch count plus one (at most 1024) */
    int32_t use_graphs;     /* 1 = replay captured HIP graphs of level launches (default),
                               0 = launch each level kernel on the stream */
    int32_t batch_pull_div; /* batched BFS: pull a level when |F| + sum deg(F) > m / this,
                               else push; default 10; 0 = push only */
    int32_t bu_row_tops;    /* 1 = bottom-up levels read a 32-byte per-vertex record (the
                               six largest neighbours, row end, degree), built in HBM on the
                               first BFS of a sorted symmetric graph (default); 0 = they read
                               offsets and row tops from the CSR (no extra HBM) */
    int32_t queue_div;      /* the single-source frontier queues hold n / queue_div entries
                               (default 4); a level whose next frontier does not fit stores
                               the rest in the frontier bitmap only, and the level after it
                               expands that bitmap (same outputs, bounded workspace) */
    int32_t persist;        /* 1 = a single-source BFS runs every level in one persistent
                               launch whose workgroups (grid_blocks, else one per CU) meet at
                               a grid barrier between levels; one such launch per device at a
                               time, others run one level per launch.  0 = one level per
                               launch.  2 (default) = persistent once the graph's previous run
                               needed 64 levels or more (deep graphs), else one per launch */
    int32_t persist_wait_us;/* longest wait at a barrier of a persistent launch (default
                               200000); a launch that waits longer (its grid could not gather,
                               e.g. on a device shared with other processes) gives up and the
                               BFS runs again one level per launch, with the same outputs */
} pbbs_tuning;

int pbbs_tuning_get(const pbbs_ctx* ctx, pbbs_tuning* out);
int pbbs_tuning_set(pbbs_ctx* ctx, const pbbs_tuning* t);

/* ---- graph resident in HBM --------------------------------------------- */
typedef struct pbbs_graph pbbs_graph;

#define PBBS_GRAPH_OFFSETS_I64      0x1u  /* offsets are int64_t[n+1] (else uint32_t[n+1]);
                                             HBM keeps the width given */
#define PBBS_GRAPH_DEVICE_INPUT     0x2u  /* offsets/adj point to device memory */
#define PBBS_GRAPH_ASSUME_SYMMETRIC 0x8u  /* skip the symmetry check; caller vouches for it */

typedef struct {
    int64_t n;              /* vertices */
    int64_t m;              /* directed CSR entries */
    int32_t offsets_i64;    /* 1 if HBM offsets are int64 */
    int32_t sorted;         /* 1 if every adjacency row is ascending */
    int32_t symmetric;      /* 1 if (u,v) in E <=> (v,u) in E (checked on upload) */
    int32_t pad;
    int64_t isolated;       /* vertices of degree 0 */
    int64_t max_degree;
    int64_t device_bytes;   /* HBM held by the graph */
} pbbs_graph_info;

/* Copies a CSR graph (`offsets[n+1]`, offsets[n] == m, `adj[m]`) into HBM and
 * computes the per-graph metadata (row sortedness, symmetry, isolated-vertex
 * bitmap).  Replaces graph<int,uint>'s construction in readGraphFromFile
 * (common/graphIO.h:191-215, common/graph.h:141-148).  Limits:
 * 0 < n < 2^31, m < 2^32.  Bottom-up steps need a symmetric graph (they scan
 * N(v) as in-neighbours); a non-symmetric graph runs top-down only, as
 * deterministicBFS does on a directed graph. */
int pbbs_graph_upload(pbbs_ctx* ctx, int64_t n, int64_t m, const void* offsets,
                      const int32_t* adj, uint32_t flags, pbbs_graph** out);
int pbbs_graph_info_get(const pbbs_graph* g, pbbs_graph_info* out);
int pbbs_graph_free(pbbs_graph* g);

/* Builds a graph's search data now rather than on its first BFS, so that a
 * first BFS costs what a later one does and the HBM it takes is known:
 *   PBBS_PREPARE_SINGLE  the bottom-up share bounds for ctx's launch shape
 *                        and, on a sorted symmetric graph with bu_row_tops
 *                        on, the 32-byte row-top records (32 n bytes);
 *   PBBS_PREPARE_BATCH   (symmetric graphs) degree ranks, the rank-id copy
 *                        of the adjacency, the list of non-isolated rows and
 *                        the pull slices for ctx's batched launch shape
 *                        (12 n + 4 m bytes).
 * Each piece is built once per graph and kept until pbbs_graph_free;
 * pbbs_graph_info's device_bytes counts it.  *ms (may be NULL) receives the
 * call's wall time.  No reference counterpart (the reference's BFS keeps no
 * per-graph data between calls). */
#define PBBS_PREPARE_SINGLE 0x1u
#define PBBS_PREPARE_BATCH  0x2u
int pbbs_graph_prepare(pbbs_ctx* ctx, pbbs_graph* g, uint32_t what, double* ms);

/* ---- single-source BFS ------------------------------------------------- */
#define PBBS_OUT_DEVICE   0x1u  /* parent/depth are device pointers (no D2H copy) */
/* Any BFS-tree parent, the contract of backForwardBFS/BFS.C:38-61 (a
 * compare-and-swap from -1: whichever frontier neighbour claims v first):
 * parent[v] is some neighbour one level up, depths are exact.  Bottom-up
 * levels then probe each row from its smallest entry up and stop at the first
 * frontier neighbour; top-down levels still take the largest.  The outputs
 * are the same on every run of the same inputs, pass the reference's
 * BFSCheck (bench/BFSCheck.C:59-90), and differ from the deterministic rule
 * only at vertices a bottom-up level reached. */
#define PBBS_BFS_ANY_PARENT 0x4u

typedef struct {
    int32_t levels;         /* BFS levels executed (max depth + 1) */
    int32_t td_levels;      /* levels run top-down */
    int32_t bu_levels;      /* levels run bottom-up */
    int32_t launches;       /* level-kernel launches, including no-op overshoot */
    int64_t reached;        /* n_r: vertices with parent != -1 */
    int64_t edges_reached;  /* m_r: sum of degrees over reached vertices */
    double  device_ms;      /* hipEvent time: init kernel start -> last level done */
    double  copy_ms;        /* D2H time of the outputs (0 with PBBS_OUT_DEVICE) */
} pbbs_bfs_stats;

/* One BFS from `source`.  parent[n] and depth[n] (either may be NULL) receive
 * the results above; with PBBS_OUT_DEVICE they are device pointers and the
 * call returns once they are written in HBM.  Replaces
 * BFS(start, G, verbose) of bench/BFS.h:11 / deterministicBFS/BFS.C:39. */
int pbbs_bfs(pbbs_ctx* ctx, pbbs_graph* g, int32_t source, int32_t* parent,
             int32_t* depth, uint32_t out_flags, pbbs_bfs_stats* stats);

/* Allocation-free variant of pbbs_bfs.  No reference counterpart: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The BFS's working
 * memory (bitmaps, frontier queues, control block) is a static 32 MiB device
 * array of the library, one per GPU, held until the process ends; the call
 * makes no hipMalloc, hipFree or other allocating call.  Limits and rules:
 *   - n <= 2^25 vertices (PBBS_ERR_LIMIT beyond);
 *   - parent and depth are DEVICE buffers of n entries, both required;
 *   - the graph must have been prepared on this context
 *     (pbbs_graph_prepare with PBBS_PREPARE_SINGLE), else PBBS_ERR_STATE;
 *   - every call on one GPU shares that GPU's array: calls that overlap on one
 *     GPU (other threads or contexts) race, so the caller serialises them.
 * Outputs are bit-identical to pbbs_bfs; the call returns once they are in
 * HBM.  Frontier queues are smaller than pbbs_bfs's for large n (the rest of
 * a large frontier is expanded from its bitmap), so timings may differ. */
int pbbs_bfs_noalloc(pbbs_ctx* ctx, pbbs_graph* g, int32_t source, int32_t* parent,
                     int32_t* depth, pbbs_bfs_stats* stats);

/* ---- maximal independent set ------------------------------------------ */
typedef struct {
    int32_t rounds;         /* worklist rounds run (a round that found no work ends the run) */
    int32_t pad;
    int64_t members;        /* vertices in the set */
    double  device_ms;      /* hipEvent time: init kernel start -> output written */
    double  copy_ms;        /* D2H time of the flags (0 with PBBS_OUT_DEVICE) */
} pbbs_mis_stats;

/* Maximal independent set of g, the one PBBS's incrementalMIS computes
 * (benchmarks/maximalIndependentSet/incrementalMIS/MIS.C:35-66): vertex v is
 * in the set unless a neighbour with a lower id is -- the lexicographically
 * first maximal independent set, the same on every run.  flags[n] (may be
 * NULL) receives 1 for members and 2 for the others, the int sequence
 * MISTime.C:38-43 writes; with PBBS_OUT_DEVICE it is a device pointer.  On a
 * non-symmetric graph the rule reads each vertex's row (its out-neighbours),
 * as the reference does.  Replaces maximalIndependentSet(G)
 * (maximalIndependentSet/bench/MIS.h). */
int pbbs_mis(pbbs_ctx* ctx, pbbs_graph* g, int32_t* flags, uint32_t out_flags,
             pbbs_mis_stats* stats);

/* ---- spanning forest ---------------------------------------------------- */
#define PBBS_EDGES_DEVICE 0x2u   /* the edge arrays are device pointers */
/* Any spanning forest (the contract of spanningForest/ndST/ST.C, checked by
 * bench/STCheck): one pass of hook-and-link over the edges instead of
 * incrementalST's round schedule; which forest, and its order, vary from
 * run to run.  Ids are listed by the hooked root, as ndST filters its hooks.
 * No m < 300 refusal (there is no schedule). */
#define PBBS_ST_ANY_FOREST 0x8u

typedef struct {
    int32_t rounds;         /* rounds of the speculative schedule */
    int32_t pad;
    int64_t edges;          /* forest edges */
    double  device_ms;      /* hipEvent time: init kernel start -> forest written */
    double  copy_ms;        /* H2D of host edges + D2H of the forest */
} pbbs_st_stats;

/* Spanning forest of the edge array (u[i], v[i]), i < m, on n vertices: the
 * edge ids PBBS's incrementalST returns (spanningForest/incrementalST/
 * ST.C:35-71), in its order.  That answer is the one its speculative_for
 * schedule with granularity 100 produces (common/speculative_for.h:30-100),
 * and the GPU runs that schedule, so the ids equal the reference's bit for
 * bit.  forest receives *count <= n - 1 edge ids (may be NULL for the count
 * only); with PBBS_OUT_DEVICE it is a device pointer with room for n - 1
 * entries.  PBBS_EDGES_DEVICE: u and v are device pointers.  Edge ends must
 * lie in [0, n).  0 < m < 300 is refused (PBBS_ERR_LIMIT): the reference's
 * schedule starts those at round size 0 and throws.  Replaces st(E)
 * (spanningForest/bench/ST.h:7). */
int pbbs_spanning_forest(pbbs_ctx* ctx, int64_t n, int64_t m, const int32_t* u,
                         const int32_t* v, uint32_t flags, int32_t* forest, int64_t* count,
                         pbbs_st_stats* stats);

/* ---- batched BFS (up to 64 sources in one traversal) ------------------- */
#define PBBS_MAX_BATCH 64

/* BFS from each of sources[0..nsrc) (1 <= nsrc <= 64, repeats allowed) in one
 * traversal that carries a 64-bit source mask per vertex.  parents and depths
 * are [nsrc][n], source-major: row i is exactly what pbbs_bfs(sources[i])
 * returns.  Needs a symmetric graph.  Stats: td_levels / bu_levels count push
 * / pull levels, reached = sum over sources of n_r, edges_reached = sum of m_r.
 * out_flags: PBBS_OUT_DEVICE; PBBS_BFS_ANY_PARENT is accepted and changes
 * nothing (the deterministic rows meet that contract); other bits are
 * PBBS_ERR_INVALID, as for pbbs_bfs and pbbs_bfs_multi.
 * No reference counterpart: it batches the reference's BFS(start, G)
 * (bench/BFS.h:11) over many starts. */
int pbbs_bfs_batch(pbbs_ctx* ctx, pbbs_graph* g, const int32_t* sources, int32_t nsrc,
                   int32_t* parents, int32_t* depths, uint32_t out_flags, pbbs_bfs_stats* stats);

/* ---- several graphs in one launch --------------------------------------- */
#define PBBS_MAX_MULTI 64

/* Single-source BFS on each of njobs graphs (1 <= njobs <= 64; graphs may
 * repeat and may differ in size) in one persistent launch: each job gets a
 * share of the device's resident workgroups in proportion to n + m and runs
 * its levels on them, meeting its own barrier, so jobs of different depths
 * proceed independently.  Jobs are grouped by offset width and row order,
 * one launch per group.  parents[j] / depths[j] (either array, or any entry,
 * may be NULL) receive job j's outputs, n_j entries each; with
 * PBBS_OUT_DEVICE they are device pointers.  stats[j] is job j's (device_ms
 * is the whole launch's).  Every output is bit-identical to pbbs_bfs on
 * graphs[j] from sources[j].  A job on a graph of more than 2^23 edges runs
 * alone, one launch sequence on the whole device (device_ms its own): on a
 * share of a grid it would take longer.  A job whose workgroups could not
 * gather, or all jobs when another persistent launch holds the device, runs
 * alone too, with the same outputs.  No reference counterpart: it runs the
 * reference's BFS(start, G) (bench/BFS.h:11) on several graphs at once. */
int pbbs_bfs_multi(pbbs_ctx* ctx, int32_t njobs, pbbs_graph* const* graphs,
                   const int32_t* sources, int32_t* const* parents, int32_t* const* depths,
                   uint32_t out_flags, pbbs_bfs_stats* stats);

/* Device pointers of the context's own output buffers: those pbbs_bfs
 * (batched = 0; n entries each) or pbbs_bfs_batch (batched = 1; [k][n]) last
 * wrote when the caller passed no output pointers with PBBS_OUT_DEVICE.
 * Valid until the next call of that kind on the context.  No reference
 * counterpart (for callers that consume the outputs in HBM). */
int pbbs_ctx_outputs(const pbbs_ctx* ctx, int batched, const int32_t** parent,
                     const int32_t** depth);

/* ---- checker ----------------------------------------------------------- */
/* Validates a BFS output on the GPU: bench/BFSCheck.C:59-90 (with the source
 * as a parameter; levels from the parents by level_number, BFSCheck.C:39-56),
 * plus the max-parent rule of deterministicBFS and, if depths are given,
 * depth == level.  parent/depth are host arrays of n entries, or device
 * pointers with PBBS_CHECK_DEVICE_INPUT (e.g. one row of a batched output).
 * Replaces running bench/BFSCheck on the written files. */
#define PBBS_CHECK_DEVICE_INPUT 0x1u

typedef struct {
    int32_t valid;         /* 1 iff BFSCheck's rule accepts every vertex */
    int32_t cycle;         /* the parents contain a cycle (no level fixed point) */
    int64_t first_bad;     /* first vertex the rule rejects (or with a parent outside
                              [-1, n)), -1 if none */
    int64_t bad;           /* vertices rejected */
    int64_t max_rule_bad;  /* vertices whose parent is not the largest neighbour one
                              level up (-1 if unreached, itself for the source); -1 on a
                              graph that is not symmetric, where a row does not list the
                              in-neighbours the rule needs (the rule is not checked) */
    int64_t depth_bad;     /* vertices whose depth differs from their level */
    int64_t reached;       /* vertices with a level below n */
    int32_t max_level;
    int32_t pad;
} pbbs_check_result;

int pbbs_check_bfs(pbbs_ctx* ctx, const pbbs_graph* g, int32_t source, const int32_t* parent,
                   const int32_t* depth, uint32_t flags, pbbs_check_result* out);

/* Validates a maximal independent set on the GPU, the rule of
 * maximalIndependentSet/bench/MISCheck.C:37-63 (flags: 1 = member, anything
 * else not): no member has a member neighbour (independent), and every
 * non-member has one (maximal).  flags is a host array of n entries, or a
 * device pointer with PBBS_CHECK_DEVICE_INPUT.  Replaces running
 * bench/MISCheck on the written file. */
typedef struct {
    int32_t valid;          /* 1 iff both rules hold at every vertex */
    int32_t pad;
    int64_t first_bad;      /* smallest vertex breaking a rule, -1 if none */
    int64_t bad_edges;      /* members with a member neighbour */
    int64_t bad_vertices;   /* non-members with no member neighbour */
} pbbs_mis_check_result;

int pbbs_check_mis(pbbs_ctx* ctx, const pbbs_graph* g, const int32_t* flags,
                   uint32_t check_flags, pbbs_mis_check_result* out);

/* ---- serving loop ------------------------------------------------------ */
/* Graphs stay resident; queries arrive from any host thread and complete
 * asynchronously.  A worker thread keeps up to `slots` runs in flight (each
 * with its own stream and workspace, so their kernels overlap on the GPU),
 * admits queries as slots and the HBM budget free up, and merges waiting
 * queries for one graph into one batched traversal of up to `max_batch`
 * sources.  Every result is bit-identical to a lone pbbs_bfs call.  No
 * reference counterpart (the reference runs one BFS per process,
 * bench/BFSTime.C:51-61). */
typedef struct pbbs_server pbbs_server;

typedef struct {
    int32_t slots;        /* concurrent runs; 0 = 4 */
    int32_t max_batch;    /* queries merged per run, 1..64; 0 = 64; 1 = never merge */
    int64_t hbm_budget;   /* HBM the server may hold: resident graphs with their search
                             data, plus every slot's workspace buffers (kept between runs).
                             A run is admitted, or its batch shrunk, to fit; idle slots'
                             buffers are freed when needed.  0 = 80% of free HBM */
    int32_t reserved[4];
} pbbs_server_config;

typedef struct {
    uint64_t ticket;
    int32_t  graph;
    int32_t  source;
    int32_t  status;      /* PBBS_OK or the error of the run that served it */
    int32_t  batch;       /* queries served by that run */
    double   latency_ms;  /* arrival -> result in the caller's buffers */
    double   queue_ms;    /* arrival -> admission to a slot */
    double   device_ms;   /* hipEvent time of the run */
} pbbs_server_result;

int pbbs_server_create(int device, const pbbs_server_config* cfg, pbbs_server** out);
/* Finishes every submitted query, then releases everything. */
int pbbs_server_destroy(pbbs_server* s);
/* Uploads a CSR (as pbbs_graph_upload), builds its search data
 * (pbbs_graph_prepare) and returns its id.  PBBS_ERR_LIMIT if the resident
 * graphs plus one single-source run would exceed hbm_budget. */
int pbbs_server_add_graph(pbbs_server* s, int64_t n, int64_t m, const void* offsets,
                          const int32_t* adj, uint32_t flags, int32_t* graph_id);
/* Queues a BFS; parent/depth (host, n entries each, either may be NULL) must
 * stay valid until its result is polled. */
int pbbs_server_submit(pbbs_server* s, int32_t graph_id, int32_t source, int32_t* parent,
                       int32_t* depth, uint64_t* ticket);
/* Takes up to max finished results; waits up to wait_ms (-1 = until one). */
int pbbs_server_poll(pbbs_server* s, pbbs_server_result* out, int32_t max, int32_t* count,
                     int32_t wait_ms);
/* Queries waiting, queries in runs, and the HBM bytes charged to the budget
 * now (resident graphs with their search data + slots' workspace buffers). */
int pbbs_server_load(pbbs_server* s, int64_t* queued, int64_t* running, int64_t* hbm_in_use);

/* Per-level timing of one BFS, for tuning and measurement: every level kernel
 * is launched on its own with a hipEvent after it (no graph replay), so the
 * times exclude launch gaps.  Writes up to `max_levels` records and the number
 * of levels in *count; *init_ms gets the init kernel's time.  Outputs stay in
 * the context's device buffers.  No reference counterpart (the reference's
 * per-round timers print wall time only, common/ligraLight.h:115-126). */
typedef struct {
    int32_t level;          /* k: this kernel consumed F_k */
    int32_t mode;           /* 0 top-down, 1 bottom-up */
    int64_t frontier;       /* |F_k| (vertices of degree > 0) */
    int64_t frontier_edges; /* sum of degrees over F_k */
    double  ms;             /* kernel time */
} pbbs_level_profile;

int pbbs_bfs_profile(pbbs_ctx* ctx, pbbs_graph* g, int32_t source, pbbs_level_profile* out,
                     int32_t max_levels, int32_t* count, double* init_ms);

/* ---- host-side CSR container, generators and PBBS text I/O ------------- */
/* A host CSR: int64 offsets[n+1], int32 adj[m], rows ascending, no duplicate
 * entries.  Mirrors graph<intV,intE> (common/graph.h:104-152). */
typedef struct {
    int64_t  n;
    int64_t  m;
    int64_t* offsets;
    int32_t* adj;
} pbbs_csr;

void pbbs_csr_free(pbbs_csr* g);

/* Symmetric RMAT graph: 2^scale vertices, `draws` directed edge draws with
 * quadrant probabilities a, b, c (d = 1-a-b-c), seeded; self-loops dropped,
 * reverse edges added, duplicates removed, rows sorted.  The recursion
 * follows testData/graphData/rMatGraph.C:53-79 (one uniform draw per level
 * picks the quadrant) with this library's own counter-based hash in place of
 * parlay::hash64.  `permute_seed` != 0 relabels vertices with a seeded random
 * permutation (the unordered variant, common/graphUtils.h:166-189). */
int pbbs_gen_rmat(int32_t scale, int64_t draws, double a, double b, double c,
                  uint64_t seed, uint64_t permute_seed, pbbs_csr* out);

/* d x d x d torus, degree 6: the same graph as `gridGraph -j -d 3 -o d^3`
 * (testData/graphData/gridGraph.C:53-72, symmetrised). */
int pbbs_gen_grid3d(int32_t d, pbbs_csr* out);

/* Random-local graph: vertex i draws `k` neighbours j = (i +- floor(n^u)) mod n,
 * u ~ U[0,1), random sign; self-loops dropped, symmetrised, deduplicated. */
int pbbs_gen_randlocal(int64_t n, int32_t k, uint64_t seed, pbbs_csr* out);

/* Builds a CSR from an edge list (u[i], v[i]); symmetrise != 0 adds reverse
 * edges.  Self-loops are dropped when symmetrising (common/graphUtils.h:63-71);
 * duplicates are always removed (common/graphUtils.h:54-61). */
int pbbs_csr_from_edges(int64_t n, int64_t num_edges, const int32_t* u,
                        const int32_t* v, int symmetrise, pbbs_csr* out);

/* Relabels g's vertices with a seeded random permutation (the one
 * pbbs_gen_rmat's permute_seed applies) and sorts the rows: the unordered
 * graphs the benchmarks are timed on (3Dgrid_JR, randLocalGraph_JR,
 * rMatGraph_JR) are the ordered ones relabelled this way
 * (common/graphUtils.h:166-189).  seed != 0. */
int pbbs_csr_permute(const pbbs_csr* g, uint64_t seed, pbbs_csr* out);

/* ---- graphs built in HBM ----------------------------------------------- */
/* The generators below, run on the GPU: the resulting graph is bit-identical
 * to pbbs_gen_rmat / pbbs_gen_grid3d (same arguments) followed by
 * pbbs_graph_upload, without the host build or the copy.  flags: 0 (uint32
 * offsets) or PBBS_GRAPH_OFFSETS_I64.  Replaces running rMatGraph
 * (testData/graphData/rMatGraph.C:33-79) / gridGraph (gridGraph.C:53-72) and
 * reading the file back (common/graphIO.h:191-215). */
int pbbs_graph_gen_rmat(pbbs_ctx* ctx, int32_t scale, int64_t draws, double a, double b, double c,
                        uint64_t seed, uint64_t permute_seed, uint32_t flags, pbbs_graph** out);
int pbbs_graph_gen_grid3d(pbbs_ctx* ctx, int32_t d, uint32_t flags, pbbs_graph** out);
/* The undirected edges of a resident symmetric graph as an edge array in
 * HBM: one (u[i], v[i]) with u < v per edge, in row order, or in the order of
 * a seeded random shuffle when shuffle_seed != 0 (the unordered edge files
 * the spanning-forest benchmark reads are written shuffled:
 * common/graphIO.h:319-333).  u and v are device buffers with room for cap
 * entries; *m_out gets the edge count (also when cap is too small, which is
 * PBBS_ERR_LIMIT; u = v = NULL, cap = 0 asks for the count only).  Replaces
 * writing an EdgeArray file (edgesFromGraph + randomShuffle) and reading it
 * back (common/graphIO.h:147-166). */
int pbbs_graph_edges(pbbs_ctx* ctx, const pbbs_graph* g, uint64_t shuffle_seed, int32_t* u,
                     int32_t* v, int64_t cap, int64_t* m_out);
/* Copies a resident graph back into a host CSR (int64 offsets; free it with
 * pbbs_csr_free).  Replaces writeGraphToFile's source of data
 * (common/graphIO.h:78-105). */
int pbbs_graph_download(const pbbs_graph* g, pbbs_csr* out);


/* PBBS AdjacencyGraph text format (docs/fileFormats/graph.md:16-39;
 * reader common/graphIO.h:191-215, writer common/graphIO.h:78-105).  Rows
 * are kept in file order. */
int pbbs_read_adjacency_graph(const char* path, pbbs_csr* out);
int pbbs_write_adjacency_graph(const char* path, const pbbs_csr* g);

/* Binary CSR file, this library's own format (the reference has no binary
 * graph file): 32-byte header (magic "PBBSCSR1", n u64, m u64, 8 reserved
 * zero bytes), then offsets int64[n+1], then adjacency int32[m], all
 * little-endian.  Reading it is a memory copy, so a graph parsed once from
 * text reloads at disk speed.  The reader rejects a size that does not match
 * its header. */
int pbbs_write_graph_binary(const char* path, const pbbs_csr* g);
int pbbs_read_graph_binary(const char* path, pbbs_csr* out);

/* Reads a graph file (AdjacencyGraph text, or the binary format above,
 * told apart by the magic) into pinned host staging and uploads it as
 * pbbs_graph_upload would (flags: PBBS_GRAPH_OFFSETS_I64,
 * PBBS_GRAPH_ASSUME_SYMMETRIC).  Replaces the benchmark driver's
 * readGraphFromFile + graph-to-device step (common/graphIO.h:191-215,
 * benchmarks/breadthFirstSearch/bench/BFSTime.C:57-58). */
int pbbs_graph_read(pbbs_ctx* ctx, const char* path, uint32_t flags, pbbs_graph** out);

/* PBBS sequenceInt text format (docs/fileFormats/sequence.md:10-26; writer
 * common/IO.h:151-166 via common/sequenceIO.h:189-193, reader
 * common/IO.h:217-229).  The reader allocates *out with malloc. */
int pbbs_write_sequence_int(const char* path, const int32_t* a, int64_t n);
/* PBBS EdgeArray text format ("EdgeArray", then u v per edge): reader
 * common/graphIO.h:147-166 (n = the largest id + 1; an odd last token is
 * ignored, as there; ids must lie in [0, 2^31 - 1)), writer :136-138.  The
 * reader's arrays are freed with pbbs_free. */
int pbbs_read_edge_array(const char* path, int64_t* n, int64_t* m, int32_t** u, int32_t** v);
int pbbs_write_edge_array(const char* path, int64_t m, const int32_t* u, const int32_t* v);
int pbbs_read_sequence_int(const char* path, int32_t** out, int64_t* n);
void pbbs_free(void* p);

#ifdef __cplusplus
}
#endif
#endif /* PBBS_BFS_H_ */
