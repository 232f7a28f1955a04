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

// ---------------------------------------------------------------------------
// Cross-block sums in a fixed order (no floating-point atomics)
// ---------------------------------------------------------------------------
//
// A reduction whose result is a sum over the nb blocks of a group (the
// blocks along gridDim.x that share blockIdx.y / blockIdx.z) stores each
// block's partial values and lets the group's last block add them in block
// order, so the result does not depend on the order in which blocks finish.
// Workspace (caller memory, sum_part / sum_part_len of the entry points): the
// first (groups + 1) / 2 doubles hold one 32-bit ticket per group, the rest
// groups * nv * nb partial values.  It must be all zero when a launch starts;
// the last block of each group clears its ticket and the partials it read,
// so every launch leaves it all zero again, whatever layout the next launch
// on it uses.

// Doubles of workspace for `groups` groups of nv values over nb blocks.
static inline int64_t flpr_sum_part_len(int64_t groups, int nv, int64_t nb) {
  return (groups + 1) / 2 + groups * (int64_t)nv * nb;
}

// Blocks of one group at most (the last block reads all partials at once:
// at most four per lane and value).
constexpr int FLPR_SUM_MAX_BLOCKS = 256;

// The largest block count per group (at most want and FLPR_SUM_MAX_BLOCKS, at
// least 1) whose partials fit in len doubles; 0 when not even one block per
// group fits.
static inline int flpr_sum_blocks(int64_t groups, int nv, int want, int64_t len) {
  const int64_t room = len - (groups + 1) / 2;
  if (groups < 1 || nv < 1 || room < groups * (int64_t)nv) return 0;
  int64_t fit = room / (groups * (int64_t)nv);
  if (fit > want) fit = want;
  if (fit > FLPR_SUM_MAX_BLOCKS) fit = FLPR_SUM_MAX_BLOCKS;
  return (int)fit;
}

// Called by every thread of every block of a group (block b of nb <=
// FLPR_SUM_MAX_BLOCKS, group grp; blocks of 256 threads): threads j < nv <=
// 32 hold value j in v.  The group's last block to arrive adds, for each j,
// the nb partials in a fixed order to out[j]: wave w takes the values j = w,
// w + 4, ..., lane l the partials l, l + 64, l + 128, l + 192 (added in that
// order), then a fixed xor-shuffle tree over the lanes; every partial is
// cleared as it is read.  Partials and tickets are written and read with
// device-scope atomic stores, loads and adds, which are coherent across the
// 8 XCDs without any cache maintenance; each thread waits for its partial
// store to complete before the barrier that precedes the ticket, so a block
// that finds itself last sees every partial.  No device-scope fence is used:
// on this GPU one writes back and invalidates the whole L2 of the XCD, under
// every other kernel's working set.
__device__ inline void flpr_group_sum(double v, int nv, int b, int nb, int64_t grp,
                                      int64_t groups, double* sum_part, double* out) {
  constexpr int NW = 4, PER = FLPR_SUM_MAX_BLOCKS / 64, MAXJ = 32 / NW;
  __shared__ int last;
  uint32_t* tickets = reinterpret_cast<uint32_t*>(sum_part);
  double* slots = sum_part + (groups + 1) / 2 + grp * (int64_t)nv * nb;
  const int t = (int)threadIdx.x;
  if (t < nv)
    __hip_atomic_store(&slots[(int64_t)t * nb + b], v, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (t == 0)
    last = __hip_atomic_fetch_add(&tickets[grp], 1u, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT) == (uint32_t)(nb - 1);
  __syncthreads();
  if (!last) return;
  const int lane = t & 63, w = t >> 6;
  double acc[MAXJ][PER];
#pragma unroll
  for (int k = 0; k < MAXJ; ++k) {
    const int j = w + NW * k;
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int i = lane + 64 * r;
      acc[k][r] = 0.0;
      if (j < nv && i < nb)
        acc[k][r] = __hip_atomic_load(&slots[(int64_t)j * nb + i], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
    }
  }
#pragma unroll
  for (int k = 0; k < MAXJ; ++k) {
    const int j = w + NW * k;
    if (j >= nv) break;
    double a = acc[k][0];
#pragma unroll
    for (int r = 1; r < PER; ++r) a += acc[k][r];
#pragma unroll
    for (int r = 0; r < PER; ++r)
      if (lane + 64 * r < nb)
        __hip_atomic_store(&slots[(int64_t)j * nb + lane + 64 * r], 0.0, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
    if (lane == 0) out[j] += a;
  }
  if (t == 0) __hip_atomic_store(&tickets[grp], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
