// hip_util.hpp - error plumbing shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>

#include "pgdg.h"

#define PGDG_HIP_TRY(expr)                                                        \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      fprintf(stderr, "pgdg: %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(e_), \
              __FILE__, __LINE__);                                                \
      return PGDG_ERR_HIP;                                                        \
    }                                                                             \
  } while (0)

#define PGDG_TRY(expr)        \
  do {                        \
    int r_ = (expr);          \
    if (r_ != PGDG_OK) return r_; \
  } while (0)

namespace pgdg {

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Grid size for a 1-D launch over n items with `block` threads, capped so a
// grid-stride loop covers the rest.
inline unsigned grid_for(long long n, int block, long long cap = 1 << 20) {
  long long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// Workgroups are dispatched to the 8 XCDs round-robin by workgroup id, and
// each XCD has its own L2.  xcd_block() renumbers them so that each XCD
// walks runs of kXcdChunk consecutive logical workgroups (neighbouring
// workgroups read overlapping neighbour data and then share one L2), with
// the runs dealt round-robin so that spatially uneven work stays balanced
// across XCDs.  The grid must come from grid_xcd (a multiple of
// 8 * kXcdChunk workgroups); logical ids past the work exit.
constexpr int kXcds = 8;
constexpr int kXcdChunk = 16;
inline unsigned grid_xcd(long long n, int block) {
  long long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  const long long q = (long long)kXcds * kXcdChunk;
  g = (g + q - 1) / q * q;
  return (unsigned)g;
}
__device__ inline unsigned xcd_block() {
  const unsigned x = blockIdx.x % kXcds, s = blockIdx.x / kXcds;
  return ((s / kXcdChunk) * kXcds + x) * kXcdChunk + s % kXcdChunk;
}

}  // namespace pgdg
