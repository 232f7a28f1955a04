// Shared helpers for the fastlpr_amd HIP sources (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/fastlpr_amd.h"

// Last HIP error seen by this thread inside the library (errno-style; read
// with flpr_last_hip_error).
void flpr_set_hip_error(hipError_t e);

// Entry points clear any error left behind by earlier runtime calls of other
// code before launching, then check their own launches.
#define FLPR_CLEAR_ERROR() (void)hipGetLastError()
#define FLPR_CHECK_LAUNCH()                 \
  do {                                      \
    hipError_t e_ = hipGetLastError();      \
    if (e_ != hipSuccess) {                 \
      flpr_set_hip_error(e_);               \
      return -2;                            \
    }                                       \
  } while (0)

// Batch view of one launch: the problem is blockIdx.z.  items (device
// memory) holds the per-problem geometry and point range, or is null when
// the launch serves one problem; s[] are the per-problem strides (elements)
// of the kernel's per-problem pointers, in the order each kernel documents.
// A launch for one problem passes a zeroed FlprBz with gridDim.z = 1.
struct FlprBz {
  const FlprBatchItem* items;
  int64_t s[6];
};

static inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

static inline int flpr_grid_for(int64_t n, int threads, int cap) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

static inline bool flpr_geom_ok(const FlprGeom* g) {
  if (!g || g->d < 1 || g->d > FLPR_MAX_DIM || g->msp < 0 || g->msp > FLPR_MAX_MSP) return false;
  for (int k = 0; k < g->d; ++k)
    if (g->n_grid[k] < 2 || g->tile[k] < 1 || g->ntile[k] < 1 || g->G[k] != 2 * g->E[k] + 1 ||
        (int64_t)g->ntile[k] * g->tile[k] < g->G[k] || g->L[k] < 2)
      return false;
  return true;
}

static inline int64_t flpr_ntiles(const FlprGeom* g) {
  int64_t t = 1;
  for (int k = 0; k < g->d; ++k) t *= g->ntile[k];
  return t;
}

static inline int64_t flpr_nkeys(const FlprGeom* g) {
  int64_t t = flpr_ntiles(g);
  for (int k = 0; k < g->d; ++k) t *= g->tile[k];
  return t;
}

static inline __host__ __device__ int64_t flpr_nodes(const FlprGeom* g) {
  int64_t t = 1;
  for (int k = 0; k < g->d; ++k) t *= g->n_grid[k];
  return t;
}

// Order-preserving encoding of a double into uint64 (for atomicMax).
__host__ __device__ inline uint64_t flpr_encode(double v) {
  union {
    double d;
    uint64_t u;
  } c;
  c.d = v;
  return (c.u >> 63) ? ~c.u : (c.u | 0x8000000000000000ull);
}

__host__ __device__ inline double flpr_decode(uint64_t k) {
  union {
    double d;
    uint64_t u;
  } c;
  c.u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return c.d;
}
