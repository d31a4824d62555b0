// Internal helpers shared by the library's translation units.
#pragma once
#include <stdint.h>

namespace pbbs {
// Records a message for pbbs_last_error() and returns `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
const char* last_error();
// The seeded vertex permutation of the generators (Fisher-Yates over a
// SplitMix64 sequence): perm[0..n) becomes a permutation of 0..n-1.
void make_permutation(int64_t n, uint64_t seed, int32_t* perm);
// Finds the device address of the library's static scratch array on
// `device` (the current device), once per device (noalloc.hip).
int scratch_init(int device);
}  // namespace pbbs
