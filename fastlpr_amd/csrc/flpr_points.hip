// Point-side kernels: x statistics, knot mapping and cell binning, packing of
// the sorted point data, Gaussian-gridding spread, grid centring, adjoint
// interpolation sums for the Hutchinson dots and the residual sums, the
// residual quadratic form, and the final interpolation.  Point arithmetic is
// float64.
//
// Reference dataflow restated here (fastLPR_py/src/fastlpr/...):
//   model.py:58-72            z-score (ddof=1) and centring
//   convolution.py:169-199    knot mapping into the padded frequency window
//   nufftn_type1.py:140-176   scale_knots (recentre, x 2 pi, mod 2 pi)
//   nufftn_type1.py:520-636   round(xmod/hx), (2 Msp + 1)^d taps, exp(-diff^2/(4 tau))
//   spreading_fused.py:67-94  fused per-point spread (2-D form)
//   regression.py:1256-1280   Hutchinson p' * interp(m)
//   api.py:686-723            RegularGridInterpolator 'linear' at the data points, RSS
//
// Spread.  Points are counting-sorted by spread cell (cells grouped by tile).
// One block owns a tile; one work item owns (cell, channel) and accumulates
// the cell's points into the full (2 Msp + 1)^d footprint in float64
// registers, then adds the footprint once into the tile's LDS window.  The
// tap weights factor as exp(-(delta - hx k)^2 c) = A B^k G_k with
// A = exp(-delta^2 c), B = exp(2 delta hx c), G_k = exp(-hx^2 k^2 c), so a
// point needs two exponentials per dimension.
//
// Residual sums.  yhat_i(h) = sum_node w_i(node) m_h(node) is linear in m, so
//   sum_i p_ik phat_ik(h) = sum_node Q_k(node) m_hk(node),  Q_k = sum_i p_ik w_i
//   sum_i (y_i - yhat_i)^2 = sum_i y_i^2 - 2 sum_node Q_y m_h + sum_cell m' W m
// with W the per-cell Gram matrix of the 2^d interpolation weights.  Q and W
// are accumulated once per fit from the sorted points; the per-bandwidth work
// is on the grid only.
#include "flpr_common.h"

#include <mutex>

namespace {

constexpr int XS_BLOCKS = 512;
constexpr int XS_THREADS = 256;
constexpr int MAXY = 32;  // y columns summed by the statistics pass

struct Welford {
  double n, mean, m2, mn, mx;
};

__device__ inline Welford wf_merge(Welford a, Welford b) {
  if (a.n == 0) return b;
  if (b.n == 0) return a;
  Welford r;
  r.n = a.n + b.n;
  const double delta = b.mean - a.mean;
  r.mean = a.mean + delta * (b.n / r.n);
  r.m2 = a.m2 + b.m2 + delta * delta * (a.n * b.n / r.n);
  r.mn = fmin(a.mn, b.mn);
  r.mx = fmax(a.mx, b.mx);
  return r;
}

// Welford merge over a block (valid on thread 0): an xor-shuffle tree inside
// each wave, then the waves' results in wave order (a fixed shape, no LDS
// tree with a barrier per level).
__device__ inline Welford wf_block(Welford w, Welford* sh) {
  for (int off = 32; off > 0; off >>= 1) {
    const Welford o = {__shfl_xor(w.n, off), __shfl_xor(w.mean, off), __shfl_xor(w.m2, off),
                       __shfl_xor(w.mn, off), __shfl_xor(w.mx, off)};
    w = wf_merge(w, o);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = w;
  __syncthreads();
  Welford r = sh[0];
  if (threadIdx.x == 0)
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) r = wf_merge(r, sh[i]);
  return r;
}

__device__ inline double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += sh[w];
  return r;
}

// Per-dimension Welford statistics of x; counts of non-finite values in x and
// y; sums of squares of the y columns.  Partials per block:
//   [d][blocks][5] Welford, then [blocks][2] non-finite counts,
//   then [ny][blocks] sums of squares.
// Batch (blockIdx.z = problem): items give the point range; s[0] = partials.
__global__ __launch_bounds__(XS_THREADS) void k_xstats(const double* __restrict__ x, int64_t n,
                                                       int d, const double* __restrict__ y,
                                                       int ny, double* __restrict__ partials,
                                                       FlprBz bz) {
  __shared__ Welford sh[XS_THREADS / 64];
  if (bz.items) {
    const FlprBatchItem& it = bz.items[blockIdx.z];
    if (x) x += it.pt * d;
    if (y) y += it.pt * ny;
    n = it.n;
  }
  partials += blockIdx.z * bz.s[0];
  __shared__ double red[XS_THREADS / 64];
  double bad_x = 0.0, bad_y = 0.0;
  for (int k = 0; k < d; ++k) {
    // a thread's values as sums of v - s1, s1 the first point's value (when
    // finite): no division per value; Welford form once per thread
    const double s1 = (n > 0 && isfinite(x[k])) ? x[k] : 0.0;
    double cnt = 0.0, t1 = 0.0, t2 = 0.0, mn = INFINITY, mx = -INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
      const double v = x[i * d + k];
      if (!isfinite(v)) {
        bad_x += 1.0;
        continue;
      }
      const double dv = v - s1;
      cnt += 1.0;
      t1 += dv;
      t2 = fma(dv, dv, t2);
      mn = fmin(mn, v);
      mx = fmax(mx, v);
    }
    Welford w = {cnt, 0.0, 0.0, mn, mx};
    if (cnt > 0.0) {
      const double dm = t1 / cnt;
      w.mean = s1 + dm;
      w.m2 = fmax(t2 - t1 * dm, 0.0);
    }
    w = wf_block(w, sh);
    if (threadIdx.x == 0) {
      double* p = partials + ((size_t)k * gridDim.x + blockIdx.x) * 5;
      p[0] = w.n; p[1] = w.mean; p[2] = w.m2; p[3] = w.mn; p[4] = w.mx;
    }
  }
  double* cnt = partials + (size_t)d * gridDim.x * 5;
  double* sq = cnt + (size_t)gridDim.x * 2;
  for (int c = 0; c < ny; ++c) {
    double s2 = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
      const double v = y[i * ny + c];
      if (!isfinite(v)) bad_y += 1.0;
      else s2 += v * v;
    }
    const double t = block_sum(s2, red);
    if (threadIdx.x == 0) sq[(size_t)c * gridDim.x + blockIdx.x] = t;
  }
  const double bx = block_sum(bad_x, red);
  const double by = block_sum(bad_y, red);
  if (threadIdx.x == 0) {
    cnt[2 * blockIdx.x] = bx;
    cnt[2 * blockIdx.x + 1] = by;
  }
}

// out: [d][5] = min, max, mean, M2, count; then bad_x, bad_y; then ny sums of squares.
// Batch: s[0] = partials, s[1] = out.
__global__ __launch_bounds__(XS_THREADS) void k_xstats_final(const double* __restrict__ partials,
                                                             int nparts, int d, int ny,
                                                             double* __restrict__ out, FlprBz bz) {
  __shared__ Welford sh[XS_THREADS / 64];
  partials += blockIdx.z * bz.s[0];
  out += blockIdx.z * bz.s[1];
  __shared__ double red[XS_THREADS / 64];
  for (int k = 0; k < d; ++k) {
    Welford w = {0.0, 0.0, 0.0, INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
      const double* p = partials + ((size_t)k * nparts + b) * 5;
      const Welford o = {p[0], p[1], p[2], p[3], p[4]};
      w = wf_merge(w, o);
    }
    w = wf_block(w, sh);
    if (threadIdx.x == 0) {
      out[5 * k + 0] = w.mn;
      out[5 * k + 1] = w.mx;
      out[5 * k + 2] = w.mean;
      out[5 * k + 3] = w.m2;
      out[5 * k + 4] = w.n;
    }
  }
  const double* cnt = partials + (size_t)d * nparts * 5;
  const double* sq = cnt + (size_t)nparts * 2;
  double bx = 0.0, by = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
    bx += cnt[2 * b];
    by += cnt[2 * b + 1];
  }
  bx = block_sum(bx, red);
  by = block_sum(by, red);
  if (threadIdx.x == 0) {
    out[5 * d] = bx;
    out[5 * d + 1] = by;
  }
  for (int c = 0; c < ny; ++c) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) s += sq[(size_t)c * nparts + b];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[5 * d + 2 + c] = s;
  }
}

// numpy's float modulo for a positive divisor (npy_divmod): fmod, then shift
// negative remainders by the divisor; an exact zero becomes +0.
__device__ inline double np_mod_pos(double a, double b) {
  double m = fmod(a, b);
  if (m != 0.0) {
    if (m < 0.0) m += b;
  } else {
    m = 0.0;
  }
  return m;
}

// Knot position and spread cell of one coordinate, exactly as k_prepare
// computes them (same float64 operation order, no contraction).
__device__ inline double knot_of(const FlprGeom& g, int k, double v, int* m_out) {
#pragma clang fp contract(off)
  const double two_pi = 2.0 * M_PI;
  const double t = (v - g.mean[k]) / g.std_[k];
  const double xs = t - g.centre[k];
  double xn = (xs - g.xmin_s[k]) / g.scale_s[k];
  xn = xn * g.ks[k];
  xn = xn + g.idx0[k];
  xn = xn + 0.5;
  xn = np_mod_pos(xn, 1.0);
  double kn = xn - 0.5;
  kn = kn * 1.0;
  kn = kn + g.shift[k];
  kn = kn * two_pi;
  const double xm = np_mod_pos(kn, two_pi);
  *m_out = (int)rint(xm / g.hx[k]);
  return xm;
}

constexpr int PREP_THREADS = 256;

// Spread cell and sort key (tile-major cell index) of every point, from the
// knot mapping of knot_of.
// Batch: items give the geometry and point range; problem z's keys are
// offset by z * s[0] (its key count).
// (D = g.d: the axis loops unroll, so a geometry copied from the batch table
// stays in registers)
template <int D>
__global__ __launch_bounds__(PREP_THREADS) void k_prepare(const double* __restrict__ x, int64_t n,
                                                          FlprGeom g, int32_t* __restrict__ key,
                                                          int32_t* __restrict__ key_count,
                                                          int32_t* __restrict__ rank, FlprBz bz) {
  int key0 = 0;
  if (bz.items) {
    const FlprBatchItem& it = bz.items[blockIdx.z];
    g = it.g;
    x += it.pt * D;
    key += it.pt;
    if (rank) rank += it.pt;
    n = it.n;
    key0 = (int)(blockIdx.z * bz.s[0]);
  }
  int tc = 1;
#pragma unroll
  for (int k = 0; k < D; ++k) tc *= g.tile[k];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int tid = 0, lid = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      int m;
      knot_of(g, k, x[i * D + k], &m);
      const int pm = (m > g.L[k] / 2) ? m - g.L[k] : m;
      // the data support maps into [1 + msp, G - 2 - msp]; the clamp only keeps
      // a malformed geometry from addressing outside the key table
      const int s = min(max(pm + g.E[k], g.msp), g.G[k] - 1 - g.msp);
      tid = tid * g.ntile[k] + s / g.tile[k];
      lid = lid * g.tile[k] + s % g.tile[k];
    }
    const int kk = key0 + tid * tc + lid;
    key[i] = kk;
    if (rank) rank[i] = atomicAdd(&key_count[kk], 1);
    else atomicAdd(&key_count[kk], 1);
  }
}

// k_prepare for key tables that fit in LDS (nkeys <= PREP_LDS_KEYS): a
// block counts its points per key in LDS (the LDS atomic's return value is
// the point's rank within the block), then takes one global atomic per
// (block, key) for the key's base and adds it to its points' ranks.  With
// few keys (1-D grids: about a thousand cells on a few dozen cache lines)
// this replaces one contended global atomic per point by one per block and
// key.  A block owns a contiguous run of points.
constexpr int PREP_LDS_KEYS = 8192;

template <int D>
__global__ __launch_bounds__(PREP_THREADS) void k_prepare_lds(const double* __restrict__ x,
                                                              int64_t n, FlprGeom g,
                                                              int32_t* __restrict__ key,
                                                              int32_t* __restrict__ key_count,
                                                              int32_t* __restrict__ rank,
                                                              int nkeys, FlprBz bz) {
  extern __shared__ int hist[];
  int key0 = 0;
  if (bz.items) {
    const FlprBatchItem& it = bz.items[blockIdx.z];
    g = it.g;
    x += it.pt * D;
    key += it.pt;
    rank += it.pt;
    n = it.n;
    key0 = (int)(blockIdx.z * bz.s[0]);
  }
  for (int k = threadIdx.x; k < nkeys; k += blockDim.x) hist[k] = 0;
  __syncthreads();
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * per;
  const int64_t i1 = min(n, i0 + per);
  int tc = 1;
#pragma unroll
  for (int k = 0; k < D; ++k) tc *= g.tile[k];
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    int tid = 0, lid = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      int m;
      knot_of(g, k, x[i * D + k], &m);
      const int pm = (m > g.L[k] / 2) ? m - g.L[k] : m;
      const int s = min(max(pm + g.E[k], g.msp), g.G[k] - 1 - g.msp);
      tid = tid * g.ntile[k] + s / g.tile[k];
      lid = lid * g.tile[k] + s % g.tile[k];
    }
    const int kl = tid * tc + lid;
    key[i] = key0 + kl;
    rank[i] = atomicAdd(&hist[kl], 1);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nkeys; k += blockDim.x) {
    const int c = hist[k];
    if (c > 0) hist[k] = atomicAdd(&key_count[key0 + k], c);
  }
  __syncthreads();
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) rank[i] += hist[key[i] - key0];
}

// blocks of k_prepare_lds: about 16 points per key and block, at most 4096
static inline int prep_lds_blocks(int64_t n, int64_t nkeys) {
  int64_t b = n / (16 * nkeys);
  if (b < 1) b = 1;
  if (b > 4096) b = 4096;
  return (int)b;
}

constexpr int SCAN_THREADS = 1024;

// Exclusive scan in three launches: block sums, a scan of the block sums, and a
// block-local scan plus offset.  Every load is coalesced.
__device__ inline int block_exclusive_scan(int v, int* sh, int* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int incl = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(incl, off);
    if (lane >= off) incl += t;
  }
  if (lane == 63) sh[wid] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      const int t = sh[w];
      sh[w] = run;
      run += t;
    }
    sh[blockDim.x >> 6] = run;
  }
  __syncthreads();
  const int out = sh[wid] + incl - v;
  *total = sh[blockDim.x >> 6];
  __syncthreads();
  return out;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_blocks(const int32_t* __restrict__ cnt,
                                                              int nt,
                                                              int32_t* __restrict__ bsum) {
  __shared__ int sh[SCAN_THREADS / 64 + 1];
  const int i = blockIdx.x * SCAN_THREADS + threadIdx.x;
  int total;
  block_exclusive_scan(i < nt ? cnt[i] : 0, sh, &total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_top(int32_t* __restrict__ bsum, int nb) {
  __shared__ int sh[SCAN_THREADS / 64 + 1];
  int run = 0;
  for (int b0 = 0; b0 < nb; b0 += SCAN_THREADS) {
    const int i = b0 + threadIdx.x;
    int total;
    const int ex = block_exclusive_scan(i < nb ? bsum[i] : 0, sh, &total);
    if (i < nb) bsum[i] = run + ex;
    run += total;
  }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const int32_t* __restrict__ cnt,
                                                             int nt,
                                                             int32_t* __restrict__ bsum,
                                                             int32_t* __restrict__ start) {
  __shared__ int sh[SCAN_THREADS / 64 + 1];
  const int i = blockIdx.x * SCAN_THREADS + threadIdx.x;
  const int v = i < nt ? cnt[i] : 0;
  const int base = bsum[blockIdx.x];
  int total;
  const int ex = block_exclusive_scan(v, sh, &total);  // synchronises the block
  if (i < nt) start[i] = base + ex;
  if (i == nt - 1) start[nt] = base + ex + v;
  // every thread has read its block sum: clear it so the scratch is zero again
  if (threadIdx.x == 0) bsum[blockIdx.x] = 0;
}

// Scatter of point indices into key order: at the rank the counting pass
// returned, or with one atomic per point on its cell's cursor (about n / cells
// points share a cursor).
__global__ __launch_bounds__(256) void k_bin(const int32_t* __restrict__ key, int64_t n,
                                             const int32_t* __restrict__ start,
                                             int32_t* __restrict__ cursor,
                                             const int32_t* __restrict__ rank,
                                             int32_t* __restrict__ order) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int k = key[i];
    const int r = rank ? rank[i] : atomicAdd(&cursor[k], 1);
    order[start[k] + r] = (int32_t)i;
  }
}

// Interval index and fraction on the raw grid axis linspace(a, b, N), as
// RegularGridInterpolator does: the interval is clamped to [0, N-2] and the
// fraction is not clamped (linear extrapolation outside the grid).
__device__ inline void raw_locate(double v, double a, double b, int N, int* idx, double* frac) {
  const double step = (b - a) / (double)(N - 1);
  int i = (int)floor((v - a) / step);
  i = max(0, min(i, N - 2));
  // numpy linspace values: t*step + start, last element = stop
  double lo = (double)i * step + a;
  double hi = (i + 1 == N - 1) ? b : (double)(i + 1) * step + a;
  if (v < lo && i > 0) {
    --i;
    hi = lo;
    lo = (double)i * step + a;
  } else if (v >= hi && i < N - 2) {
    ++i;
    lo = hi;
    hi = (i + 1 == N - 1) ? b : (double)(i + 1) * step + a;
  }
  *idx = i;
  *frac = (v - lo) / (hi - lo);
}

// Sorted point data for the grid kernels, point-major so that the work items
// of one point (one per channel) read the same addresses:
//   wab  [n][d][2]  (A_k, B_k): the spread weight of footprint offset j - MSP
//                   (j = 0 .. 2 MSP) along axis k is A_k B_k^j G_k[j], with
//                   A = exp(-(delta^2 + 2 MSP hx delta) c), B = exp(2 delta hx c),
//                   G[j] = exp(-hx^2 (j - MSP)^2 c), c = 1 / (4 tau)
//   vals [n][nv]    y columns then probe columns
//   rawi [n][d]     raw-grid interval, rawf [n][d] fraction in it
// Batch: items give the geometry and the sorted range [pt, pt + n) (the
// counting sort keeps every problem's points in its own range); x and vy are
// indexed by batch row, pr by the row within the problem (one probe matrix
// whose leading rows serve every problem).
template <int D>
__global__ __launch_bounds__(256) void k_pack(FlprGeom g, const double* __restrict__ x,
                                              const int32_t* __restrict__ order, int64_t n,
                                              const double* __restrict__ vy, int ny,
                                              const double* __restrict__ pr, int np,
                                              double* __restrict__ wab, double* __restrict__ vals,
                                              int32_t* __restrict__ rawi,
                                              double* __restrict__ rawf, FlprBz bz) {
  const int nv = ny + np;
  int64_t prow0 = 0;
  if (bz.items) {
    const FlprBatchItem& it = bz.items[blockIdx.z];
    g = it.g;
    n = it.n;
    prow0 = it.pt;
    order += it.pt;
    if (wab) wab += it.pt * D * 2;
    if (vals) vals += it.pt * nv;
    if (rawi) {
      rawi += it.pt * D;
      rawf += it.pt * D;
    }
  }
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n && (wab || rawi);
       p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = order[p];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double xv = x[i * D + k];
      if (wab) {
        int m;
        const double xm = knot_of(g, k, xv, &m);
        const double delta = xm - g.hx[k] * (double)m;
        const double c = g.inv4tau[k];
        wab[(p * D + k) * 2] = exp(-((delta * delta + 2.0 * g.msp * g.hx[k] * delta) * c));
        wab[(p * D + k) * 2 + 1] = exp(2.0 * delta * g.hx[k] * c);
      }
      if (rawi) {
        int ri;
        double rf;
        raw_locate(xv, g.graw0[k], g.graw1[k], g.n_grid[k], &ri, &rf);
        rawi[p * D + k] = ri;
        rawf[p * D + k] = rf;
      }
    }
  }
  // values: one thread per (point, column), so that consecutive lanes write
  // consecutive addresses of the point-major table
  const int64_t nvals = n * nv;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nvals;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = e / nv;
    const int c = (int)(e - p * nv);
    const int64_t i = order[p];
    vals[e] = (c < ny) ? vy[i * ny + c] : pr[(i - prow0) * np + (c - ny)];
  }
}

template <int D>
__device__ inline void tile_coords(const FlprGeom& g, int tid, int* t) {
#pragma unroll
  for (int k = D - 1; k >= 0; --k) {
    t[k] = tid % g.ntile[k];
    tid /= g.ntile[k];
  }
}

template <int D>
__device__ inline void local_coords(const FlprGeom& g, int lc, int* l) {
#pragma unroll
  for (int k = D - 1; k >= 0; --k) {
    l[k] = lc % g.tile[k];
    lc /= g.tile[k];
  }
}

// Footprint weights along one axis: w[j] = A B^j G[j], j = 0 .. 2 MSP.
template <int MSP, typename T>
__device__ inline void tap_weights(T A, T B, const double* G, T* w) {
  T bp = A;
#pragma unroll
  for (int j = 0; j <= 2 * MSP; ++j) {
    w[j] = bp * (T)G[j];
    bp *= B;
  }
}

// Clears words of device memory with vector stores (a kernel node, unlike a
// memset, replays correctly inside a captured HIP graph).
__global__ __launch_bounds__(256) void k_fill_zero(uint32_t* __restrict__ p, int64_t words) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// Points per work item the spread and adjoint aim for when a cell holds many
// points (a cell's points are split over several items), and the number of
// blocks below which channels are split over blocks.
constexpr int ITEM_POINTS_1D = 8;
constexpr int ITEM_POINTS = 16;
constexpr int MIN_BLOCKS = 512;

__host__ __device__ inline int item_splits(const FlprGeom* g, int64_t n) {
  const int64_t nodes = flpr_nodes(g);
  const int target = g->item_points > 0 ? g->item_points
                                        : ((g->d == 1) ? ITEM_POINTS_1D : ITEM_POINTS);
  int64_t s = (n + nodes * target - 1) / (nodes * target);
  if (s < 1) s = 1;
  if (s > 64) s = 64;
  return (int)s;
}

// Splits of a cell's points so that a block's work items cover its threads
// (few channels per block, e.g. a single new y column): items of one split
// number tile cells x rows x channels.
__host__ __device__ inline int fill_splits(int nsplit, int64_t items_per_split, int threads) {
  if (items_per_split < 1) return nsplit;
  int64_t need = (threads + items_per_split - 1) / items_per_split;
  if (need > 64) need = 64;
  return (int)(nsplit > need ? nsplit : need);
}

constexpr int SPREAD_THREADS = 256;
constexpr int SPREAD_MAX_TAPS = 2 * FLPR_MAX_MSP + 1;
// Minimum waves per SIMD.  The 3-D item's 9 x 9 float64 block left alone
// takes just over 256 registers, one wave per SIMD, with nothing to hide the
// point loads and the window adds behind; capped at 256 two waves fit (a few
// registers spill).  The compensated float32 block (sum and compensation)
// would spill far more and keeps one wave.
template <int D, typename A>
constexpr int SPREAD_WAVES = (D == 3 && sizeof(A) == 8) ? 2 : 1;

// G[k][j] = exp(-hx_k^2 (j - MSP)^2 / (4 tau_k)); passed by value so that the
// kernels read it through the scalar unit.
struct SpreadTab {
  double G[FLPR_MAX_DIM][SPREAD_MAX_TAPS];
};

// Register-accumulated spread.  Work item = (split, cell, row, channel), the
// channel fastest so that the lanes sharing a point read the same point data.
// An item accumulates its share (split s of nsplit) of the cell's points into
// a footprint block held in float64 registers: the last F = d - FIX axes of
// the (2 MSP + 1)^d footprint, at fixed offsets ("row") along the first FIX
// axes (FIX = 0 for d = 1 and for d = 2 with MSP = 4, so the whole footprint
// is one block; otherwise enough axes are fixed to keep the block at
// (2 MSP + 1)^2 registers or fewer).  The block is added once into the tile's
// LDS window, and the window is written to the tile's partial slot
// (k_spread_gather sums the overlapping windows per cell, without atomics).
// A = float (the precision form's spread, g.spread_f32): the weights and the
// footprint block are float32 and each accumulator carries a Kahan
// compensation term; the LDS window and the partial slots are float32 (the
// gather adds the windows in float64).
// Batch: s[0] = key_start (the problem's key count), s[1] = part; a
// problem's cell split follows its own point count (items).
template <int D, int MSP, typename A>
__global__ __launch_bounds__(SPREAD_THREADS, (SPREAD_WAVES<D, A>)) void k_spread_cells(
    FlprGeom g, SpreadTab tab, const int32_t* __restrict__ key_start,
    const double* __restrict__ wab, const double* __restrict__ vals, int64_t n, int with_ones,
    int C, int cg, int nsplit, A* __restrict__ part, FlprBz bz) {
  extern __shared__ char lds_raw[];
  A* lds = reinterpret_cast<A*>(lds_raw);
  constexpr bool KAHAN = sizeof(A) == 4;
  key_start += blockIdx.z * bz.s[0];
  part += blockIdx.z * bz.s[1];
  constexpr int ntap = 2 * MSP + 1;
  constexpr int FIX = (D == 1) ? 0 : (D == 2 ? (MSP > 4 ? 1 : 0) : (MSP > 4 ? 2 : 1));
  constexpr int F = D - FIX;
  constexpr int ROWS = (FIX == 0) ? 1 : (FIX == 1 ? ntap : ntap * ntap);
  constexpr int NACC = (F == 1) ? ntap : ntap * ntap;
  const int nv = C - with_ones;
  const int c0 = blockIdx.y * cg;
  const int cn = min(cg, C - c0);
  int t[3] = {0, 0, 0};
  tile_coords<D>(g, blockIdx.x, t);
  int fp[3] = {1, 1, 1};
  int fcells = 1, tc = 1;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    fp[k] = g.tile[k] + 2 * MSP;
    fcells *= fp[k];
    tc *= g.tile[k];
  }
  for (int i = threadIdx.x; i < fcells * cn; i += blockDim.x) lds[i] = A(0);
  __syncthreads();
  const int kbase = blockIdx.x * tc;
  if (bz.items)
    nsplit = fill_splits(item_splits(&g, bz.items[blockIdx.z].n),
                         (int64_t)tc * ROWS * min(cg, C), SPREAD_THREADS);
  const int nitems = tc * nsplit * ROWS * cn;
  for (int item = threadIdx.x; item < nitems; item += blockDim.x) {
    const int c = item % cn;
    int rest = item / cn;
    const int row = rest % ROWS;
    rest /= ROWS;
    // the split is the slowest index: neighbouring lanes flush neighbouring
    // cells' footprints (the splits of one cell would add to the same addresses)
    const int lc = rest % tc;
    const int sp = rest / tc;
    const int p0 = key_start[kbase + lc], p1 = key_start[kbase + lc + 1];
    const int64_t len = p1 - p0;
    const int q0 = p0 + (int)((len * sp) / nsplit);
    const int q1 = p0 + (int)((len * (sp + 1)) / nsplit);
    if (q0 == q1) continue;
    const int ch = c0 + c;
    const bool ones = ch < with_ones;
    const double* vp = vals + (ones ? 0 : ch - with_ones);
    int pos[3] = {0, 0, 0};
    local_coords<D>(g, lc, pos);
    int r[2] = {0, 0};  // fixed offsets along axes 0 .. FIX-1
    if constexpr (FIX == 1) r[0] = row;
    if constexpr (FIX == 2) {
      r[0] = row / ntap;
      r[1] = row % ntap;
    }
    // G[r] of the fixed axes, chosen once per item
    A gfix[FIX > 0 ? FIX : 1];
#pragma unroll
    for (int k = 0; k < FIX; ++k) {
      A sel = A(0);
#pragma unroll
      for (int j = 0; j < ntap; ++j) sel = (j == r[k]) ? (A)tab.G[k][j] : sel;
      gfix[k] = sel;
    }
    A acc[NACC];
    A cmp[KAHAN ? NACC : 1];
#pragma unroll
    for (int a = 0; a < NACC; ++a) acc[a] = A(0);
#pragma unroll
    for (int a = 0; a < (KAHAN ? NACC : 1); ++a) cmp[a] = A(0);
    // acc[a] += v, compensated in float32 (Kahan): cmp holds minus the
    // low-order part lost so far
    auto add = [&](int a, A v) {
      if constexpr (KAHAN) {
        const A yv = v - cmp[a];
        const A t = acc[a] + yv;
        cmp[a] = (t - acc[a]) - yv;
        acc[a] = t;
      } else {
        acc[a] += v;
      }
    };
    // software pipeline, two points deep: the data of points p + 1 and p + 2
    // is in flight while point p's footprint is accumulated
    double nxt[2][2 * D];
    double nxv[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int pj = min(q0 + j, q1 - 1);
#pragma unroll
      for (int k = 0; k < 2 * D; ++k) nxt[j][k] = wab[(int64_t)pj * (2 * D) + k];
      nxv[j] = ones ? 1.0 : vp[(int64_t)pj * nv];
    }
    for (int p = q0; p < q1; ++p) {
      double cur[2 * D];
#pragma unroll
      for (int k = 0; k < 2 * D; ++k) {
        cur[k] = nxt[0][k];
        nxt[0][k] = nxt[1][k];
      }
      A s = (A)nxv[0];
      nxv[0] = nxv[1];
      if (p + 2 < q1) {
#pragma unroll
        for (int k = 0; k < 2 * D; ++k) nxt[1][k] = wab[(int64_t)(p + 2) * (2 * D) + k];
        nxv[1] = ones ? 1.0 : vp[(int64_t)(p + 2) * nv];
      }
      A w[D][ntap];
#pragma unroll
      for (int k = FIX; k < D; ++k)
        tap_weights<MSP, A>((A)cur[2 * k], (A)cur[2 * k + 1], tab.G[k], w[k]);
      // a fixed axis needs only the weight of offset r[k]: A B^r G[r], with
      // B^r by squaring (r <= 2 MSP < 16)
#pragma unroll
      for (int k = 0; k < FIX; ++k) {
        A bb = (A)cur[2 * k + 1], bp = A(1);
#pragma unroll
        for (int bit = 0; (1 << bit) <= 2 * MSP; ++bit) {
          bp *= ((r[k] >> bit) & 1) ? bb : A(1);
          bb *= bb;
        }
        s *= (A)cur[2 * k] * bp * gfix[k];
      }
      if constexpr (F == 1) {
#pragma unroll
        for (int j = 0; j < ntap; ++j) add(j, s * w[D - 1][j]);
      } else {
#pragma unroll
        for (int a = 0; a < ntap; ++a) {
          const A sa = s * w[D - 2][a];
#pragma unroll
          for (int b = 0; b < ntap; ++b) add(a * ntap + b, sa * w[D - 1][b]);
        }
      }
    }
    // window cell of the block's first entry, then the free axes' strides
    int base = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) base = base * fp[k] + pos[k] + (k < FIX ? r[k] : 0);
    // the compensated sum is acc - cmp
    auto total = [&](int a) -> A {
      if constexpr (KAHAN) return acc[a] - cmp[a];
      else return acc[a];
    };
    if constexpr (F == 1) {
#pragma unroll
      for (int j = 0; j < ntap; ++j) atomicAdd(&lds[(base + j) * cn + c], total(j));
    } else {
      const int sa = fp[D - 1];
#pragma unroll
      for (int a = 0; a < ntap; ++a)
#pragma unroll
        for (int b = 0; b < ntap; ++b)
          atomicAdd(&lds[(base + a * sa + b) * cn + c], total(a * ntap + b));
    }
  }
  __syncthreads();
  A* dst = part + ((int64_t)blockIdx.x * C + c0) * fcells;
  for (int e = threadIdx.x; e < fcells * cn; e += blockDim.x) {
    const int c = e / fcells, cell = e % fcells;
    dst[(int64_t)c * fcells + cell] = lds[cell * cn + c];
  }
}


// Three-dimensional spread with footprints wider than the LDS window admits
// (the oversampled grid of accuracies 7..12: up to 25 cells per axis, a
// window of (tile + 2 MSP)^3 doubles per channel would exceed the LDS).  Work
// item = (cell, row, channel) as in k_spread_cells with the first two axes
// fixed (row = offsets r0, r1 of the footprint): the item accumulates its
// cell's points into one footprint row of 2 MSP + 1 float64 registers and
// adds the row straight into the grid with float64 atomics (grid zeroed by
// the caller).  The weights are those of tap_weights (A B^j G[j], the same
// products in the same order); the order of the atomics varies between runs,
// so the grid varies in its last bits.
template <int MSP>
__global__ __launch_bounds__(SPREAD_THREADS) void k_spread_wide3(
    FlprGeom g, SpreadTab tab, const int32_t* __restrict__ key_start,
    const double* __restrict__ wab, const double* __restrict__ vals, int with_ones, int C,
    double* __restrict__ grid, int64_t cells) {
  constexpr int ntap = 2 * MSP + 1;
  constexpr int ROWS = ntap * ntap;
  const int nv = C - with_ones;
  int t[3] = {0, 0, 0};
  tile_coords<3>(g, blockIdx.x, t);
  const int tc = g.tile[0] * g.tile[1] * g.tile[2];
  const int kbase = blockIdx.x * tc;
  const int64_t nitems = (int64_t)tc * ROWS * C;
  for (int64_t item = threadIdx.x; item < nitems; item += blockDim.x) {
    const int c = (int)(item % C);
    const int64_t rest = item / C;
    const int row = (int)(rest % ROWS);
    const int lc = (int)(rest / ROWS);
    const int p0 = key_start[kbase + lc], p1 = key_start[kbase + lc + 1];
    if (p0 == p1) continue;
    const bool ones = c < with_ones;
    const double* vp = vals + (ones ? 0 : c - with_ones);
    int pos[3] = {0, 0, 0};
    local_coords<3>(g, lc, pos);
    const int r0 = row / ntap, r1 = row % ntap;
    // grid cell of footprint offsets (r0, r1, j): tile origin - MSP + position + offset
    const int s0 = t[0] * g.tile[0] - MSP + pos[0] + r0;
    const int s1 = t[1] * g.tile[1] - MSP + pos[1] + r1;
    if (s0 < 0 || s0 >= g.G[0] || s1 < 0 || s1 >= g.G[1]) continue;
    double acc[ntap];
#pragma unroll
    for (int j = 0; j < ntap; ++j) acc[j] = 0.0;
    for (int p = p0; p < p1; ++p) {
      const double* wp = wab + (int64_t)p * 6;
      double b0 = wp[0], b1 = wp[2];
      for (int j = 0; j < r0; ++j) b0 *= wp[1];
      for (int j = 0; j < r1; ++j) b1 *= wp[3];
      const double s = (ones ? 1.0 : vp[(int64_t)p * nv]) * (b0 * tab.G[0][r0]) *
                       (b1 * tab.G[1][r1]);
      double w2[ntap];
      tap_weights<MSP, double>(wp[4], wp[5], tab.G[2], w2);
#pragma unroll
      for (int j = 0; j < ntap; ++j) acc[j] += s * w2[j];
    }
    const int s2b = t[2] * g.tile[2] - MSP + pos[2];
    double* gp = grid + (int64_t)c * cells + ((int64_t)s0 * g.G[1] + s1) * g.G[2];
#pragma unroll
    for (int j = 0; j < ntap; ++j) {
      const int s2 = s2b + j;
      if (s2 >= 0 && s2 < g.G[2] && acc[j] != 0.0) atomicAdd(&gp[s2], acc[j]);
    }
  }
}

// Per-channel totals of a grid, in a fixed order so that the totals (and the
// centring shifts made from them) are the same in every run; sums[c] +=
// total (sums zeroed by the caller; problem p's grid at p * grid_stride, its
// sums at p * C).  With a workspace: block (b, c, p) adds the cells of chunk
// b of channel c (thread t: cells t, t + 256, ... of the chunk, then a fixed
// block tree) into part[(p * C + c) * nb + b], and k_grid_sums_final adds the
// nb partials of each channel (lane l: partials l, l + 64, ...; then a fixed
// xor tree).  Without one (the 3-D oversampled spread): one block per
// channel adds all its cells.
constexpr int CSUM_THREADS = 256;
constexpr int CSUM_CHUNK = 8192;      // cells per block of the partial pass
constexpr int CSUM_MAX_PARTS = 256;   // partials per channel at most

__global__ __launch_bounds__(CSUM_THREADS) void k_grid_sums(const double* __restrict__ grid,
                                                            int64_t cells, int64_t grid_stride,
                                                            int C, int64_t chunk,
                                                            double* __restrict__ dst) {
  __shared__ double red[CSUM_THREADS / 64];
  const int b = blockIdx.x, c = blockIdx.y, p = blockIdx.z;
  const double* gp = grid + (int64_t)p * grid_stride + (int64_t)c * cells;
  const int64_t lo = (int64_t)b * chunk;
  const int64_t hi = lo + chunk < cells ? lo + chunk : cells;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int64_t i = lo + threadIdx.x;
  for (; i + 3 * CSUM_THREADS < hi; i += 4 * CSUM_THREADS) {
    s0 += gp[i];
    s1 += gp[i + CSUM_THREADS];
    s2 += gp[i + 2 * CSUM_THREADS];
    s3 += gp[i + 3 * CSUM_THREADS];
  }
  for (; i < hi; i += CSUM_THREADS) s0 += gp[i];
  const double s = block_sum((s0 + s1) + (s2 + s3), red);
  if (threadIdx.x == 0) {
    if (gridDim.x == 1)
      dst[(int64_t)p * C + c] += s;  // one block per channel: the total
    else
      dst[((int64_t)p * C + c) * gridDim.x + b] = s;
  }
}

__global__ __launch_bounds__(64) void k_grid_sums_final(const double* __restrict__ part, int nb,
                                                        int C, double* __restrict__ sums) {
  const int c = blockIdx.x, p = blockIdx.y;
  const double* pp = part + ((int64_t)p * C + c) * nb;
  double a = 0.0;
  for (int i = threadIdx.x; i < nb; i += 64) a += pp[i];
  for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
  if (threadIdx.x == 0) sums[(int64_t)p * C + c] += a;
}

// part: workspace of at least P * C * CSUM_MAX_PARTS doubles, or NULL.
void launch_grid_sums(hipStream_t st, const double* grid, int64_t cells, int64_t grid_stride,
                      int C, int P, double* sums, double* part) {
  int64_t nb = part ? (cells + CSUM_CHUNK - 1) / CSUM_CHUNK : 1;
  if (nb > CSUM_MAX_PARTS) nb = CSUM_MAX_PARTS;
  if (nb < 1) nb = 1;
  const int64_t chunk = (cells + nb - 1) / nb;
  hipLaunchKernelGGL(k_grid_sums, dim3((unsigned)nb, (unsigned)C, (unsigned)P),
                     dim3(CSUM_THREADS), 0, st, grid, cells, grid_stride, C, chunk,
                     nb == 1 ? sums : part);
  if (nb > 1)
    hipLaunchKernelGGL(k_grid_sums_final, dim3((unsigned)C, (unsigned)P), dim3(64), 0, st, part,
                       (int)nb, C, sums);
}

__device__ inline int floordiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

// Sum of the overlapping tile windows for every spread-grid cell (channel =
// blockIdx.y).  Batch: s[0] = part, s[1] = grid.
template <int D, typename A>
__global__ __launch_bounds__(256) void k_spread_gather(FlprGeom g, const A* __restrict__ part,
                                                       int C, int64_t cells,
                                                       double* __restrict__ grid, FlprBz bz) {
  part += blockIdx.z * bz.s[0];
  grid += blockIdx.z * bz.s[1];
  const int c = blockIdx.y;
  const int msp = g.msp;
  int fp[3] = {1, 1, 1};
  int fcells = 1;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    fp[k] = g.tile[k] + 2 * msp;
    fcells *= fp[k];
  }
  for (int64_t lin = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; lin < cells;
       lin += (int64_t)gridDim.x * blockDim.x) {
    int sc[3] = {0, 0, 0}, tlo[3] = {0, 0, 0}, thi[3] = {0, 0, 0};
    int64_t rem = lin;
#pragma unroll
    for (int k = D - 1; k >= 0; --k) {
      sc[k] = (int)(rem % g.G[k]);
      rem /= g.G[k];
      // windows [t*tile - msp, t*tile + tile + msp) that contain sc
      tlo[k] = max(floordiv(sc[k] - g.tile[k] - msp, g.tile[k]) + 1, 0);
      thi[k] = min(floordiv(sc[k] + msp, g.tile[k]), g.ntile[k] - 1);
    }
    double v = 0.0;
    if constexpr (D == 1) {
      for (int t0 = tlo[0]; t0 <= thi[0]; ++t0)
        v += (double)part[((int64_t)t0 * C + c) * fcells + (sc[0] - (t0 * g.tile[0] - msp))];
    } else if constexpr (D == 2) {
      for (int t0 = tlo[0]; t0 <= thi[0]; ++t0)
        for (int t1 = tlo[1]; t1 <= thi[1]; ++t1) {
          const int64_t tile = (int64_t)t0 * g.ntile[1] + t1;
          const int l0 = sc[0] - (t0 * g.tile[0] - msp), l1 = sc[1] - (t1 * g.tile[1] - msp);
          v += (double)part[(tile * C + c) * fcells + l0 * fp[1] + l1];
        }
    } else {
      for (int t0 = tlo[0]; t0 <= thi[0]; ++t0)
        for (int t1 = tlo[1]; t1 <= thi[1]; ++t1)
          for (int t2 = tlo[2]; t2 <= thi[2]; ++t2) {
            const int64_t tile = ((int64_t)t0 * g.ntile[1] + t1) * g.ntile[2] + t2;
            const int l0 = sc[0] - (t0 * g.tile[0] - msp), l1 = sc[1] - (t1 * g.tile[1] - msp),
                      l2 = sc[2] - (t2 * g.tile[2] - msp);
            v += (double)part[(tile * C + c) * fcells + (l0 * fp[1] + l1) * fp[2] + l2];
          }
    }
    grid[(int64_t)c * cells + lin] = v;
  }
}

// Centre every channel c >= 1 on the grid: F_c -= shift_c F_0 with
// shift_c = sum F_c / sum F_0 (the kernel-weighted mean of the channel).  The
// per-node solve adds shift_c times the ones-channel moments back, which is
// exact by linearity; the float32 convolution then only carries y - shift.
// Batch: s[0] = grid and grid32, s[1] = sums and shift.
__global__ __launch_bounds__(256) void k_grid_center(double* __restrict__ grid, int64_t cells,
                                                     const double* __restrict__ sums,
                                                     double* __restrict__ shift,
                                                     float* __restrict__ grid32, FlprBz bz) {
  const int c = blockIdx.y;
  grid += blockIdx.z * bz.s[0];
  if (grid32) grid32 += blockIdx.z * bz.s[0];
  sums += blockIdx.z * bz.s[1];
  if (shift) shift += blockIdx.z * bz.s[1];
  const double s0 = sums[0];
  const double sh = (c == 0 || s0 == 0.0) ? 0.0 : sums[c] / s0;
  if (blockIdx.x == 0 && threadIdx.x == 0 && shift) shift[c] = sh;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells;
       i += (int64_t)gridDim.x * blockDim.x) {
    double v = grid[(int64_t)c * cells + i];
    if (c > 0) {
      v = v - sh * grid[i];
      grid[(int64_t)c * cells + i] = v;
    }
    if (grid32) grid32[(int64_t)c * cells + i] = (float)v;
  }
}

// Periodic folding of the unwrapped spread grid for grids narrower than their
// footprint.  The reference spreads onto an L-periodic grid and convolves
// circularly (nufftn_type1.py:565-636, convolution.py:904-951); the unwrapped
// grid (signed cells -E..E, G = 2E+1 of them) equals it only while no two data
// cells alias, i.e. G <= L.  For small N (G > L on some axis) every cell is set
// to the sum over its residue class modulo L: each output's N-tap window then
// sees every residue once, with the periodic grid's value, which is the
// circular convolution.  Pass 1 gathers each residue's sum into per (channel,
// residue) storage, pass 2 writes the sums back to every cell of the class.
template <int D>
__global__ __launch_bounds__(256) void k_periodic_sum(FlprGeom g, const double* __restrict__ grid,
                                                      int64_t cells, int64_t pcells,
                                                      double* __restrict__ per) {
  const int c = blockIdx.y;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < pcells;
       r += (int64_t)gridDim.x * blockDim.x) {
    int res[3] = {0, 0, 0}, cnt[3] = {1, 1, 1}, first[3] = {0, 0, 0};
    int64_t rem = r;
#pragma unroll
    for (int k = D - 1; k >= 0; --k) {
      res[k] = (int)(rem % g.L[k]);
      rem /= g.L[k];
      // cells s with (s - E) mod L == res: the first is s0 = the smallest s >= 0
      int s0 = ((res[k] + g.E[k]) % g.L[k] + g.L[k]) % g.L[k];
      first[k] = s0;
      cnt[k] = (s0 < g.G[k]) ? (g.G[k] - 1 - s0) / g.L[k] + 1 : 0;
    }
    double v = 0.0;
    for (int a = 0; a < cnt[0]; ++a) {
      const int s0 = first[0] + a * g.L[0];
      if constexpr (D == 1) {
        v += grid[(int64_t)c * cells + s0];
      } else {
        for (int b = 0; b < cnt[1]; ++b) {
          const int s1 = first[1] + b * g.L[1];
          if constexpr (D == 2) {
            v += grid[(int64_t)c * cells + (int64_t)s0 * g.G[1] + s1];
          } else {
            for (int e = 0; e < cnt[2]; ++e) {
              const int s2 = first[2] + e * g.L[2];
              v += grid[(int64_t)c * cells + ((int64_t)s0 * g.G[1] + s1) * g.G[2] + s2];
            }
          }
        }
      }
    }
    per[(int64_t)c * pcells + r] = v;
  }
}

template <int D>
__global__ __launch_bounds__(256) void k_periodic_put(FlprGeom g, double* __restrict__ grid,
                                                      int64_t cells, int64_t pcells,
                                                      const double* __restrict__ per) {
  const int c = blockIdx.y;
  for (int64_t lin = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; lin < cells;
       lin += (int64_t)gridDim.x * blockDim.x) {
    int64_t rem = lin, r = 0, mul = 1;
#pragma unroll
    for (int k = D - 1; k >= 0; --k) {
      const int s = (int)(rem % g.G[k]);
      rem /= g.G[k];
      const int res = ((s - g.E[k]) % g.L[k] + g.L[k]) % g.L[k];
      r += res * mul;
      mul *= g.L[k];
    }
    grid[(int64_t)c * cells + lin] = per[(int64_t)c * pcells + r];
  }
}

// Corner pair (qa <= qb) of the e-th entry of a cell's interpolation Gram matrix.
template <int D>
__device__ inline void gram_pair(int e, int* qa, int* qb) {
  constexpr int nq = 1 << D;
  int a = 0;
  while (a < nq - 1 && e >= nq - a) {
    e -= nq - a;
    ++a;
  }
  *qa = a;
  *qb = a + e;
}

constexpr int ADJ_THREADS = 256;

// Adjoint interpolation sums on the raw grid, per spread tile:
//   mode 0: q[c][node]   += v_ic w_i(node)          (value channels)
//   mode 1: W[e][rcell]  += w_i(qa) w_i(qb)          (Gram entries)
// Work item = (split, spread cell, channel) with the channel fastest; the
// points of a spread cell touch at most three raw intervals per dimension, so
// an item accumulates 4^d node (mode 0) or 3^d cell (mode 1) sums in
// registers, then adds them into the tile's LDS window, which is written to
// the tile's partial slot (k_adjoint_gather adds the overlapping windows).
// Batch: s[0] = key_start (the problem's key count), s[1] = out, s[2] =
// part; a problem's cell split follows its own point count (items).
template <int D>
__global__ __launch_bounds__(ADJ_THREADS) void k_adjoint(
    FlprGeom g, const int32_t* __restrict__ key_start, const int32_t* __restrict__ rawi,
    const double* __restrict__ rawf, const double* __restrict__ vals, int64_t n, int nch,
    int mode, int cg, int nsplit, double* __restrict__ out, int64_t out_stride,
    double* __restrict__ part, FlprBz bz) {
  extern __shared__ double lds[];
  key_start += blockIdx.z * bz.s[0];
  out += blockIdx.z * bz.s[1];
  part += blockIdx.z * bz.s[2];
  constexpr int NA = (D == 1) ? 4 : (D == 2 ? 16 : 64);   // 4^D node accumulators
  constexpr int NC = (D == 1) ? 3 : (D == 2 ? 9 : 27);    // 3^D cell accumulators
  const int c0 = blockIdx.y * cg;
  const int cn = min(cg, nch - c0);
  int t[3] = {0, 0, 0};
  tile_coords<D>(g, blockIdx.x, t);
  int ext[3] = {1, 1, 1}, lb[3] = {0, 0, 0};
  int wcells = 1, tc = 1;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    ext[k] = g.tile[k] + 5;
    lb[k] = t[k] * g.tile[k] - g.E[k] + (g.n_grid[k] - 1) / 2 - 2;
    wcells *= ext[k];
    tc *= g.tile[k];
  }
  for (int i = threadIdx.x; i < wcells * cn; i += blockDim.x) lds[i] = 0.0;
  __syncthreads();
  const int kbase = blockIdx.x * tc;
  if (bz.items)
    nsplit = fill_splits(item_splits(&g, bz.items[blockIdx.z].n), (int64_t)tc * min(cg, nch),
                         ADJ_THREADS);
  const int nitems = tc * nsplit * cn;
  // With an even number of nodes on every axis the raw nodes sit half a cell
  // off the spread cell centres, so a cell's points share one raw interval
  // (rel = 1 on every axis; only rounding at the data's ends differs): those
  // points take a short path that accumulates the 2^D corner products only.
  bool aligned = true;
#pragma unroll
  for (int k = 0; k < D; ++k) aligned = aligned && !(g.n_grid[k] & 1);
  for (int item = threadIdx.x; item < nitems; item += blockDim.x) {
    const int c = item % cn;
    const int rest = item / cn;
    const int lc = rest % tc;   // the split slowest, as in k_spread_cells
    const int sp = rest / tc;
    const int p0 = key_start[kbase + lc], p1 = key_start[kbase + lc + 1];
    const int64_t len = p1 - p0;
    const int q0 = p0 + (int)((len * sp) / nsplit);
    const int q1 = p0 + (int)((len * (sp + 1)) / nsplit);
    if (q0 == q1) continue;
    int l[3] = {0, 0, 0};
    local_coords<D>(g, lc, l);
    int ib[3] = {0, 0, 0};  // lowest raw interval the cell's points can fall in
#pragma unroll
    for (int k = 0; k < D; ++k)
      ib[k] = (t[k] * g.tile[k] + l[k] - g.E[k]) + (g.n_grid[k] - 1) / 2 - 1;
    const int ch = c0 + c;
    int qa = 0, qb = 0;
    if (mode == 1) gram_pair<D>(ch, &qa, &qb);
    double acc[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[a] = 0.0;
    // software pipeline: the next point's data is loaded before the current
    // point is accumulated
    int nri[3] = {0, 0, 0};
    double nf[3] = {0, 0, 0};
    double nv = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      nri[k] = rawi[(int64_t)q0 * D + k];
      nf[k] = rawf[(int64_t)q0 * D + k];
    }
    if (mode == 0) nv = vals[(int64_t)q0 * nch + ch];
    for (int p = q0; p < q1; ++p) {
      int ri[3] = {0, 0, 0}, rel[3] = {0, 0, 0};
      double f[3] = {0, 0, 0};
      const double v = nv;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        ri[k] = nri[k];
        f[k] = nf[k];
      }
      if (p + 1 < q1) {
        const int64_t pn = p + 1;
#pragma unroll
        for (int k = 0; k < D; ++k) {
          nri[k] = rawi[pn * D + k];
          nf[k] = rawf[pn * D + k];
        }
        if (mode == 0) nv = vals[pn * nch + ch];
      }
      bool inside = true, main = aligned;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        rel[k] = ri[k] - ib[k];
        if (rel[k] < 0 || rel[k] > 2) inside = false;
        main = main && rel[k] == 1;
      }
      if (main) {
        // nodes ib + 1 + bit per axis: accumulator (1 + b0, 1 + b1, ..) of the
        // 4^D node window, or the centre of the 3^D cell window
        if (mode == 0) {
          if constexpr (D == 1) {
            acc[1] += v * (1.0 - f[0]);
            acc[2] += v * f[0];
          } else if constexpr (D == 2) {
            const double va = v * (1.0 - f[0]), vb = v * f[0];
            const double g1 = 1.0 - f[1];
            acc[5] += va * g1;
            acc[6] += va * f[1];
            acc[9] += vb * g1;
            acc[10] += vb * f[1];
          } else {
            const double va = v * (1.0 - f[0]), vb = v * f[0];
            const double g1 = 1.0 - f[1], g2 = 1.0 - f[2];
            const double w00 = va * g1, w01 = va * f[1], w10 = vb * g1, w11 = vb * f[1];
            acc[21] += w00 * g2;
            acc[22] += w00 * f[2];
            acc[25] += w01 * g2;
            acc[26] += w01 * f[2];
            acc[37] += w10 * g2;
            acc[38] += w10 * f[2];
            acc[41] += w11 * g2;
            acc[42] += w11 * f[2];
          }
        } else {
          double wa = 1.0, wb = 1.0;
#pragma unroll
          for (int k = 0; k < D; ++k) {
            const int ba = (qa >> (D - 1 - k)) & 1, bb = (qb >> (D - 1 - k)) & 1;
            wa *= ba ? f[k] : 1.0 - f[k];
            wb *= bb ? f[k] : 1.0 - f[k];
          }
          constexpr int centre = (D == 1) ? 1 : (D == 2 ? 4 : 13);
          acc[centre] += wa * wb;
        }
        continue;
      }
      if (mode == 0) {
        if (!inside) {
#pragma unroll
          for (int q = 0; q < (1 << D); ++q) {
            double w = v;
            int64_t node = 0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
              const int bit = (q >> (D - 1 - k)) & 1;
              w *= bit ? f[k] : 1.0 - f[k];
              node = node * g.n_grid[k] + ri[k] + bit;
            }
            atomicAdd(&out[(int64_t)ch * out_stride + node], w);
          }
          continue;
        }
        double wn[3][4];
#pragma unroll
        for (int k = 0; k < D; ++k)
#pragma unroll
          for (int a = 0; a < 4; ++a)
            wn[k][a] = (a == rel[k]) ? 1.0 - f[k] : ((a == rel[k] + 1) ? f[k] : 0.0);
        if constexpr (D == 1) {
#pragma unroll
          for (int a = 0; a < 4; ++a) acc[a] += v * wn[0][a];
        } else if constexpr (D == 2) {
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            const double va = v * wn[0][a];
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a * 4 + b] += va * wn[1][b];
          }
        } else {
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
              const double vab = v * wn[0][a] * wn[1][b];
#pragma unroll
              for (int e = 0; e < 4; ++e) acc[(a * 4 + b) * 4 + e] += vab * wn[2][e];
            }
        }
      } else {
        double wa = 1.0, wb = 1.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
          const int ba = (qa >> (D - 1 - k)) & 1, bb = (qb >> (D - 1 - k)) & 1;
          wa *= ba ? f[k] : 1.0 - f[k];
          wb *= bb ? f[k] : 1.0 - f[k];
        }
        const double val = wa * wb;
        if (!inside) {
          int64_t rc = 0;
#pragma unroll
          for (int k = 0; k < D; ++k) rc = rc * (g.n_grid[k] - 1) + ri[k];
          atomicAdd(&out[(int64_t)ch * out_stride + rc], val);
          continue;
        }
        int pos = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) pos = pos * 3 + rel[k];
#pragma unroll
        for (int a = 0; a < NC; ++a) acc[a] += (a == pos) ? val : 0.0;
      }
    }
    // flush into the LDS window (node window for mode 0, cell window for mode 1)
    const int na = (mode == 0) ? NA : NC;
    const int side = (mode == 0) ? 4 : 3;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      if (a >= na) break;
      int rem = a, widx = 0;
      int off[3] = {0, 0, 0};
#pragma unroll
      for (int k = D - 1; k >= 0; --k) {
        off[k] = rem % side;
        rem /= side;
      }
#pragma unroll
      for (int k = 0; k < D; ++k) widx = widx * ext[k] + (ib[k] + off[k] - lb[k]);
      if (acc[a] != 0.0) atomicAdd(&lds[widx * cn + c], acc[a]);
    }
  }
  __syncthreads();
  double* dst = part + ((int64_t)blockIdx.x * nch + c0) * wcells;
  for (int e = threadIdx.x; e < wcells * cn; e += blockDim.x) {
    const int c = e / wcells, cell = e % wcells;
    dst[(int64_t)c * wcells + cell] = lds[cell * cn + c];
  }
}

// Adjoint sums (k_adjoint's two modes) for grids with an even node count on
// every axis.  There a spread cell's points lie in one raw interval (the
// cell's "main" interval; only rounding at the data's ends puts a point into
// a neighbour).  A work item is (cell, channel) for the value channels
// (MODE 0, channel fastest) or a cell for the Gram entries (MODE 1), and
// WALK_LANES consecutive lanes take one item, each walking every
// WALK_LANES-th of the cell's points in sorted order and summing their 2^D
// corner products v_ic w_i(corner) (MODE 0) or nw corner-weight products
// (MODE 1) in registers; the lanes' sums are added with shuffles.  MODE 0
// stores the item's 2^D corner sums to part[item][corner] (zeros for a cell
// without points), which k_adjoint_walk_gather adds into the nodes in a fixed
// order; MODE 1 adds each entry to its raw cell with one float64 atomic (a
// raw cell has one main spread cell).  A point outside the main interval
// adds its own products to out with atomics.  out must be zeroed.
constexpr int WALK_LANES = 4;

template <int D, int MODE>
__global__ __launch_bounds__(256) void k_adjoint_walk(FlprGeom g,
                                                      const int32_t* __restrict__ key_start,
                                                      const int32_t* __restrict__ rawi,
                                                      const double* __restrict__ rawf,
                                                      const double* __restrict__ vals, int nch,
                                                      int64_t nkeys, double* __restrict__ out,
                                                      int64_t out_stride,
                                                      double* __restrict__ part, FlprBz bz) {
  // batch: problem blockIdx.z (one geometry shape for all; the point data
  // through its key range); s[0] = key_start, s[1] = out, s[2] = part
  key_start += blockIdx.z * bz.s[0];
  out += blockIdx.z * bz.s[1];
  part += blockIdx.z * bz.s[2];
  constexpr int NQ2 = 1 << D;
  constexpr int NACC = (MODE == 0) ? NQ2 : NQ2 * (NQ2 + 1) / 2;
  int tc = 1;
#pragma unroll
  for (int k = 0; k < D; ++k) tc *= g.tile[k];
  const int per = (MODE == 0) ? nch : 1;
  const int64_t nitems = nkeys * per;
  const int sub = threadIdx.x % WALK_LANES;
  // every lane of a WALK_LANES group has the same item, so a group is either
  // wholly inside the loop body or wholly out of it (the shuffles stay in groups)
  for (int64_t it = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WALK_LANES; it < nitems;
       it += (int64_t)gridDim.x * blockDim.x / WALK_LANES) {
    const int c = (int)(it % per);
    const int64_t key = it / per;
    const int p0 = key_start[key], p1 = key_start[key + 1];
    if (MODE == 1 && p0 == p1) continue;
    int t[3] = {0, 0, 0}, l[3] = {0, 0, 0}, im[3] = {0, 0, 0};
    tile_coords<D>(g, (int)(key / tc), t);
    local_coords<D>(g, (int)(key % tc), l);
#pragma unroll
    for (int k = 0; k < D; ++k) im[k] = (t[k] * g.tile[k] + l[k] - g.E[k]) + (g.n_grid[k] - 1) / 2;
    double acc[NACC];
#pragma unroll
    for (int e = 0; e < NACC; ++e) acc[e] = 0.0;
    int any_main = 0;
    for (int p = p0 + sub; p < p1; p += WALK_LANES) {
      int ri[3] = {0, 0, 0};
      double f[3] = {0, 0, 0};
      bool main = true;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        ri[k] = rawi[(int64_t)p * D + k];
        f[k] = rawf[(int64_t)p * D + k];
        main = main && ri[k] == im[k];
      }
      const double v = (MODE == 0) ? vals[(int64_t)p * nch + c] : 1.0;
      double w[NQ2];
#pragma unroll
      for (int q = 0; q < NQ2; ++q) {
        double wq = 1.0;
#pragma unroll
        for (int k = 0; k < D; ++k) wq *= ((q >> (D - 1 - k)) & 1) ? f[k] : 1.0 - f[k];
        w[q] = wq;
      }
      any_main |= main ? 1 : 0;
      if constexpr (MODE == 0) {
#pragma unroll
        for (int q = 0; q < NQ2; ++q) {
          const double val = v * w[q];
          if (main) {
            acc[q] += val;
          } else {
            int64_t node = 0;
#pragma unroll
            for (int k = 0; k < D; ++k)
              node = node * g.n_grid[k] + ri[k] + ((q >> (D - 1 - k)) & 1);
            atomicAdd(&out[(int64_t)c * out_stride + node], val);
          }
        }
      } else {
        int64_t rc = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) rc = rc * (g.n_grid[k] - 1) + ri[k];
        int e = 0;
#pragma unroll
        for (int qa = 0; qa < NQ2; ++qa)
#pragma unroll
          for (int qb = qa; qb < NQ2; ++qb, ++e) {
            const double val = w[qa] * w[qb];
            if (main) acc[e] += val;
            else atomicAdd(&out[(int64_t)e * out_stride + rc], val);
          }
      }
    }
#pragma unroll
    for (int off = 1; off < WALK_LANES; off <<= 1) {
      any_main |= __shfl_xor(any_main, off);
#pragma unroll
      for (int e = 0; e < NACC; ++e) acc[e] += __shfl_xor(acc[e], off);
    }
    if constexpr (MODE == 0) {
      // corner q of the item in lane q of the group (NQ2 <= 8 lanes: a
      // 3-D item writes its upper four corners from the same lanes)
#pragma unroll
      for (int q = 0; q < NQ2; ++q)
        if (sub == (q % WALK_LANES)) part[it * NQ2 + q] = acc[q];
    } else {
      if (sub != 0 || !any_main) continue;
      int64_t rc = 0;
#pragma unroll
      for (int k = 0; k < D; ++k) rc = rc * (g.n_grid[k] - 1) + im[k];
#pragma unroll
      for (int e = 0; e < NACC; ++e) atomicAdd(&out[(int64_t)e * out_stride + rc], acc[e]);
    }
  }
}

// q[c][node] += the corner sums of the (up to 2^D) cells around the node,
// from k_adjoint_walk's part, in a fixed order; one thread per (node,
// channel = blockIdx.y).
template <int D>
__global__ __launch_bounds__(256) void k_adjoint_walk_gather(FlprGeom g,
                                                             const double* __restrict__ part,
                                                             int nch, int64_t nodes,
                                                             double* __restrict__ q, FlprBz bz) {
  // batch: s[0] = part, s[1] = q
  part += blockIdx.z * bz.s[0];
  q += blockIdx.z * bz.s[1];
  constexpr int NQ2 = 1 << D;
  const int c = blockIdx.y;
  int tc = 1, tstride[3] = {1, 1, 1}, lstride[3] = {1, 1, 1};
#pragma unroll
  for (int k = D - 1; k >= 0; --k) {
    if (k < D - 1) {
      tstride[k] = tstride[k + 1] * g.ntile[k + 1];
      lstride[k] = lstride[k + 1] * g.tile[k + 1];
    }
    tc *= g.tile[k];
  }
  for (int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; node < nodes;
       node += (int64_t)gridDim.x * blockDim.x) {
    int r[3] = {0, 0, 0};
    int64_t rem = node;
#pragma unroll
    for (int k = D - 1; k >= 0; --k) {
      r[k] = (int)(rem % g.n_grid[k]);
      rem /= g.n_grid[k];
    }
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < NQ2; ++q) {
      // the interval im = r - bit along each axis holds this node as its corner q
      bool ok = true;
      int64_t key_t = 0, key_l = 0;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const int bit = (q >> (D - 1 - k)) & 1;
        const int im = r[k] - bit;
        ok = ok && im >= 0 && im <= g.n_grid[k] - 2;
        const int sidx = im - (g.n_grid[k] - 1) / 2 + g.E[k];
        key_t += (int64_t)(sidx / g.tile[k]) * tstride[k];
        key_l += (int64_t)(sidx % g.tile[k]) * lstride[k];
      }
      if (ok) v += part[((key_t * tc + key_l) * nch + c) * NQ2 + q];
    }
    q[(int64_t)c * nodes + node] += v;
  }
}

// Adds the overlapping adjoint windows into out (nodes for mode 0, raw cells
// for mode 1), one thread per (entry, channel = blockIdx.y), no atomics.
// Batch: s[0] = part, s[1] = out.
template <int D>
__global__ __launch_bounds__(256) void k_adjoint_gather(FlprGeom g, const double* __restrict__ part,
                                                        int nch, int mode, int64_t count,
                                                        double* __restrict__ out, FlprBz bz) {
  const int c = blockIdx.y;
  part += blockIdx.z * bz.s[0];
  out += blockIdx.z * bz.s[1];
  int ext[3] = {1, 1, 1}, off[3] = {0, 0, 0}, lim[3] = {1, 1, 1};
  int wcells = 1;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    ext[k] = g.tile[k] + 5;
    off[k] = -g.E[k] + (g.n_grid[k] - 1) / 2 - 2;
    lim[k] = (mode == 0) ? g.n_grid[k] : g.n_grid[k] - 1;
    wcells *= ext[k];
  }
  for (int64_t lin = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; lin < count;
       lin += (int64_t)gridDim.x * blockDim.x) {
    int r[3] = {0, 0, 0}, tlo[3] = {0, 0, 0}, thi[3] = {0, 0, 0};
    int64_t rem = lin;
#pragma unroll
    for (int k = D - 1; k >= 0; --k) {
      r[k] = (int)(rem % lim[k]);
      rem /= lim[k];
      // windows [t*tile + off, t*tile + off + ext) that contain r
      tlo[k] = max(floordiv(r[k] - off[k] - ext[k], g.tile[k]) + 1, 0);
      thi[k] = min(floordiv(r[k] - off[k], g.tile[k]), g.ntile[k] - 1);
    }
    double v = 0.0;
    if constexpr (D == 1) {
      for (int t0 = tlo[0]; t0 <= thi[0]; ++t0)
        v += part[((int64_t)t0 * nch + c) * wcells + (r[0] - (t0 * g.tile[0] + off[0]))];
    } else if constexpr (D == 2) {
      for (int t0 = tlo[0]; t0 <= thi[0]; ++t0)
        for (int t1 = tlo[1]; t1 <= thi[1]; ++t1) {
          const int64_t tile = (int64_t)t0 * g.ntile[1] + t1;
          const int l0 = r[0] - (t0 * g.tile[0] + off[0]), l1 = r[1] - (t1 * g.tile[1] + off[1]);
          v += part[(tile * nch + c) * wcells + l0 * ext[1] + l1];
        }
    } else {
      for (int t0 = tlo[0]; t0 <= thi[0]; ++t0)
        for (int t1 = tlo[1]; t1 <= thi[1]; ++t1)
          for (int t2 = tlo[2]; t2 <= thi[2]; ++t2) {
            const int64_t tile = ((int64_t)t0 * g.ntile[1] + t1) * g.ntile[2] + t2;
            const int l0 = r[0] - (t0 * g.tile[0] + off[0]), l1 = r[1] - (t1 * g.tile[1] + off[1]),
                      l2 = r[2] - (t2 * g.tile[2] + off[2]);
            v += part[(tile * nch + c) * wcells + (l0 * ext[1] + l1) * ext[2] + l2];
          }
    }
    out[(int64_t)c * count + lin] += v;  // the kernel's rare out-of-window atomics landed first
  }
}

constexpr int QF_THREADS = 256;
static_assert(QF_THREADS == 256, "flpr_group_sum takes blocks of 256 threads");
// Cell blocks per (problem, h, c): each block's last step waits for its
// ticket, so a few long blocks per sum cost less than many short ones.
constexpr int QF_BLOCKS = 16;

// quad[h][c] = sum_cell sum_{a,b} W_ab(cell) m(h,c,node_a) m(h,c,node_b)
// Batch: s[0] = m, s[1] = gram, s[2] = quad.
template <int D>
__global__ __launch_bounds__(QF_THREADS) void k_quadform(FlprGeom g, const double* __restrict__ m,
                                                         int64_t nodes,
                                                         const double* __restrict__ gram,
                                                         int64_t rcells,
                                                         double* __restrict__ quad,
                                                         double* __restrict__ sum_part, FlprBz bz) {
  __shared__ double red[QF_THREADS / 64];
  m += blockIdx.z * bz.s[0];
  gram += blockIdx.z * bz.s[1];
  quad += blockIdx.z * bz.s[2];
  constexpr int NQ = 1 << D;
  const int hc = blockIdx.y;  // h * ncol + c
  const double* mm = m + (int64_t)hc * nodes;
  double s = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rcells;
       r += (int64_t)gridDim.x * blockDim.x) {
    int id[3] = {0, 0, 0};
    int64_t rem = r;
#pragma unroll
    for (int k = D - 1; k >= 0; --k) {
      id[k] = (int)(rem % (g.n_grid[k] - 1));
      rem /= (g.n_grid[k] - 1);
    }
    double mv[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      int64_t node = 0;
#pragma unroll
      for (int k = 0; k < D; ++k) node = node * g.n_grid[k] + id[k] + ((q >> (D - 1 - k)) & 1);
      mv[q] = mm[node];
    }
    int e = 0;
#pragma unroll
    for (int a = 0; a < NQ; ++a)
#pragma unroll
      for (int b = a; b < NQ; ++b) {
        const double w = gram[(int64_t)e * rcells + r];
        s += (a == b ? 1.0 : 2.0) * w * mv[a] * mv[b];
        ++e;
      }
  }
  s = block_sum(s, red);
  // the cell blocks of (problem, h, c) add their sums in block order
  flpr_group_sum(s, 1, blockIdx.x, gridDim.x, (int64_t)blockIdx.z * gridDim.y + hc,
                 (int64_t)gridDim.z * gridDim.y, sum_part, quad + hc);
}

template <int D>
struct CornersT {
  int64_t node[1 << D];
  double w[1 << D];
};

template <int D>
__device__ inline CornersT<D> corners(const FlprGeom& g, const double* __restrict__ xp) {
  CornersT<D> c;
  int id[D];
  double fr[D];
#pragma unroll
  for (int k = 0; k < D; ++k) raw_locate(xp[k], g.graw0[k], g.graw1[k], g.n_grid[k], &id[k], &fr[k]);
#pragma unroll
  for (int q = 0; q < (1 << D); ++q) {
    int64_t lin = 0;
    double w = 1.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const int bit = (q >> (D - 1 - k)) & 1;
      lin = lin * g.n_grid[k] + id[k] + bit;
      w *= bit ? fr[k] : (1.0 - fr[k]);
    }
    c.node[q] = lin;
    c.w[q] = w;
  }
  return c;
}

constexpr int EVAL_THREADS = 256;

template <int D>
__global__ __launch_bounds__(EVAL_THREADS) void k_interp_eval(FlprGeom g,
                                                              const double* __restrict__ x,
                                                              int64_t n,
                                                              const double* __restrict__ m,
                                                              int ncol, int64_t nodes,
                                                              double* __restrict__ out,
                                                              const double* __restrict__ yvar,
                                                              double* __restrict__ expsum) {
  __shared__ double sh[EVAL_THREADS / 64];
  double es = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const CornersT<D> c = corners<D>(g, x + i * D);
    for (int col = 0; col < ncol; ++col) {
      const double* mm = m + (int64_t)col * nodes;
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < (1 << D); ++q) v += c.w[q] * mm[c.node[q]];
      out[i * ncol + col] = v;
      if (col == 0 && yvar) es += yvar[i] * exp(-v);
    }
  }
  if (yvar) {
    const double s = block_sum(es, sh);
    if (threadIdx.x == 0) atomicAdd(expsum, s);
  }
}

// The fit's outputs at the selected bandwidth (index *sel on the device, or 0):
// yhat at the points by multilinear interpolation of m_all[sel] (api.py:760-770)
// and the residuals yref - yhat.  variance: yhat first holds the log-variance
// fit; each block writes its share of sum_i yref_i exp(-fit_i) to part, and
// k_fit_finish applies d = n / sum (api.py:771-786): yhat = exp(fit) / d, the
// grid values exp(m) / d into grid, resid = yref - yhat.
constexpr int FIT_BLOCKS = 1024;

// Batch: items give the geometry and the point rows; s[0] = m_all, s[1] =
// part; sel holds one index per problem.
template <int D>
__global__ __launch_bounds__(EVAL_THREADS) void k_fit_points(
    FlprGeom g, const double* __restrict__ x, int64_t n, const double* __restrict__ m_all,
    int ncol, int64_t nodes, const int32_t* __restrict__ sel, const double* __restrict__ yref,
    int variance, double* __restrict__ yhat, double* __restrict__ resid,
    double* __restrict__ part, FlprBz bz) {
  __shared__ double sh[EVAL_THREADS / 64];
  if (bz.items) {
    const FlprBatchItem& it = bz.items[blockIdx.z];
    g = it.g;
    n = it.n;
    x += it.pt * D;
    if (yref) yref += it.pt * ncol;
    yhat += it.pt * ncol;
    if (resid) resid += it.pt * ncol;
    m_all += blockIdx.z * bz.s[0];
    if (sel) sel += blockIdx.z;
    if (part) part += blockIdx.z * bz.s[1];
  }
  const double* m = m_all + (sel ? (int64_t)sel[0] : 0) * ncol * nodes;
  double es = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const CornersT<D> c = corners<D>(g, x + i * D);
    for (int col = 0; col < ncol; ++col) {
      const double* mm = m + (int64_t)col * nodes;
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < (1 << D); ++q) v += c.w[q] * mm[c.node[q]];
      yhat[i * ncol + col] = v;
      if (variance) {
        es += yref[i] * exp(-v);
      } else if (resid) {
        resid[i * ncol + col] = yref[i * ncol + col] - v;
      }
    }
  }
  if (variance) {
    const double t = block_sum(es, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

// Batch: items give the point rows; s[0] = m_all, s[1] = part, s[2] = grid,
// s[3] = fac_out; sel holds one index per problem.
__global__ __launch_bounds__(EVAL_THREADS) void k_fit_finish(
    int64_t n, const double* __restrict__ m_all, int64_t nodes, const int32_t* __restrict__ sel,
    const double* __restrict__ yref, const double* __restrict__ part, int nparts,
    double* __restrict__ yhat, double* __restrict__ resid, double* __restrict__ grid,
    double* __restrict__ fac_out, FlprBz bz) {
  __shared__ double sh[EVAL_THREADS / 64];
  if (bz.items) {
    const FlprBatchItem& it = bz.items[blockIdx.z];
    n = it.n;
    yref += it.pt;
    yhat += it.pt;
    if (resid) resid += it.pt;
    m_all += blockIdx.z * bz.s[0];
    if (sel) sel += blockIdx.z;
    part += blockIdx.z * bz.s[1];
    grid += blockIdx.z * bz.s[2];
    if (fac_out) fac_out += blockIdx.z * bz.s[3];
  }
  __shared__ double fac_s;
  double p = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) p += part[b];
  p = block_sum(p, sh);
  if (threadIdx.x == 0) fac_s = p * (1.0 / (double)n);  // 1 / d
  __syncthreads();
  const double fac = fac_s;
  if (blockIdx.x == 0 && threadIdx.x == 0 && fac_out) fac_out[0] = fac;
  const double* m = m_all + (sel ? (int64_t)sel[0] : 0) * nodes;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double v = exp(yhat[i]) * fac;
    yhat[i] = v;
    if (resid) resid[i] = yref[i] - v;
  }
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nodes;
       j += (int64_t)gridDim.x * blockDim.x)
    grid[j] = exp(m[j]) * fac;
}

// Likelihood cross-validation of the density grids (api.py:1111-1132): for
// every bandwidth b and data point i
//   f_i = linear interpolation of dens[b] at x_i
//   l_i = log(max((n / (n - 1)) (f_i - k0h[b] / n), 1e-10))
// and per block the count, sum and centred sum of squares of l_i, written to
// part[(block * dh + b) * 3].  A block owns a contiguous run of points, taken
// in the order of `order` when given (the points grouped by spread cell, so
// that neighbouring lanes interpolate neighbouring cells and the density
// reads stay in cache), and passes over it once per group of LCV_GROUP
// bandwidths: a point's interpolation corners serve the whole group.  A
// thread keeps sum l and the sums of (l - k) and (l - k)^2 with k its first
// l, so no division per point; the centred sum of squares is
// sum (l - k)^2 - (sum (l - k))^2 / count, merged across threads with Chan's
// formula.
constexpr int LCV_THREADS = 256;
constexpr int LCV_BLOCKS = 1024;
constexpr int LCV_GROUP = 8;

struct Moments3 {
  double n, s, m2;  // count, sum, sum of squared deviations from the mean
};

__device__ inline Moments3 m3_merge(Moments3 a, Moments3 b) {
  if (a.n == 0) return b;
  if (b.n == 0) return a;
  Moments3 r;
  r.n = a.n + b.n;
  const double d = b.s / b.n - a.s / a.n;
  r.s = a.s + b.s;
  r.m2 = a.m2 + b.m2 + d * d * (a.n * b.n / r.n);
  return r;
}

template <int D>
__global__ __launch_bounds__(LCV_THREADS) void k_lcv(FlprGeom g, const double* __restrict__ x,
                                                     const int32_t* __restrict__ order,
                                                     int64_t n, const double* __restrict__ dens,
                                                     int dh, int64_t nodes,
                                                     const double* __restrict__ k0h,
                                                     double* __restrict__ part) {
  __shared__ Moments3 sh[LCV_THREADS / 64][LCV_GROUP];
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * per;
  const int64_t i1 = min(n, i0 + per);
  const double nd = (double)n;
  const double ratio = nd / (nd - 1.0);
  for (int b0 = 0; b0 < dh; b0 += LCV_GROUP) {
    const int nb = min(LCV_GROUP, dh - b0);
    double c[LCV_GROUP], k[LCV_GROUP], sl[LCV_GROUP], s1[LCV_GROUP], s2[LCV_GROUP];
#pragma unroll
    for (int j = 0; j < LCV_GROUP; ++j) {
      c[j] = j < nb ? k0h[b0 + j] / nd : 0.0;
      k[j] = sl[j] = s1[j] = s2[j] = 0.0;
    }
    double cnt = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
      const int64_t p = order ? (int64_t)order[i] : i;
      const CornersT<D> cr = corners<D>(g, x + p * D);
#pragma unroll
      for (int j = 0; j < LCV_GROUP; ++j) {
        if (j < nb) {
          const double* f = dens + (int64_t)(b0 + j) * nodes;
          double fa = 0.0;
#pragma unroll
          for (int q = 0; q < (1 << D); ++q) fa += cr.w[q] * f[cr.node[q]];
          double lo = ratio * (fa - c[j]);
          lo = (lo != lo) ? lo : (lo > 1e-10 ? lo : 1e-10);
          const double l = log(lo);
          if (cnt == 0.0) k[j] = l;
          const double dl = l - k[j];
          sl[j] += l;
          s1[j] += dl;
          s2[j] += dl * dl;
        }
      }
      cnt += 1.0;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < LCV_GROUP; ++j) {
      if (j >= nb) break;
      Moments3 acc = {cnt, sl[j], cnt > 0.0 ? fmax(s2[j] - s1[j] * s1[j] / cnt, 0.0) : 0.0};
      // s2 - s1^2 / cnt is NaN whenever an l is (as Welford's would be)
      if (s2[j] != s2[j]) acc.m2 = s2[j];
      for (int off = 32; off > 0; off >>= 1) {
        Moments3 o;
        o.n = __shfl_xor(acc.n, off);
        o.s = __shfl_xor(acc.s, off);
        o.m2 = __shfl_xor(acc.m2, off);
        acc = m3_merge(acc, o);
      }
      if (lane == 0) sh[w][j] = acc;
    }
    __syncthreads();
    if (threadIdx.x < nb) {
      const int j = threadIdx.x;
      Moments3 r = sh[0][j];
      for (int v = 1; v < LCV_THREADS / 64; ++v) r = m3_merge(r, sh[v][j]);
      double* o = part + ((int64_t)blockIdx.x * dh + b0 + j) * 3;
      o[0] = r.n;
      o[1] = r.s;
      o[2] = r.m2;
    }
    __syncthreads();
  }
}

// out[b] = sum_i l_i (the LCV score), out[dh + b] = std(l) sqrt(n) (its
// standard error, np.std with ddof = 0).  One block per bandwidth: thread t
// merges block partials t, t + 256, ... in order, then a fixed xor-shuffle
// tree and the four waves in order (Chan's merge is symmetric, so every lane
// of the tree holds the same value).
__global__ __launch_bounds__(256) void k_lcv_final(const double* __restrict__ part, int nblocks,
                                                   int dh, double* __restrict__ out) {
  __shared__ Moments3 sh[4];
  const int b = blockIdx.x;
  Moments3 r = {0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < nblocks; k += blockDim.x) {
    const double* o = part + ((int64_t)k * dh + b) * 3;
    r = m3_merge(r, Moments3{o[0], o[1], o[2]});
  }
  for (int off = 32; off > 0; off >>= 1) {
    Moments3 o;
    o.n = __shfl_xor(r.n, off);
    o.s = __shfl_xor(r.s, off);
    o.m2 = __shfl_xor(r.m2, off);
    r = m3_merge(r, o);
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    r = sh[0];
    for (int w = 1; w < 4; ++w) r = m3_merge(r, sh[w]);
    out[b] = r.s;
    out[dh + b] = sqrt(r.m2 / r.n) * sqrt(r.n);
  }
}

// Cubic tensor-product spline evaluation (RegularGridInterpolator 'cubic',
// i.e. scipy's make_interp_spline(k=3) with not-a-knot ends along every axis,
// api.py:588-594, 651-657).  The caller supplies B-spline coefficients
// c (ngrid, N_0, ..., N_{d-1}); along axis k the knot vector is
//   x_0 x_0 x_0 x_0 x_2 x_3 ... x_{N-3} x_{N-1} x_{N-1} x_{N-1} x_{N-1}
// over the raw axis x_j = linspace(graw0, graw1, N)[j].  Outside the grid the
// end polynomial pieces extrapolate (BSpline extrapolate=True).
__device__ inline double np_linspace_at(double a, double b, int N, int j) {
  if (j <= 0) return a;
  if (j >= N - 1) return b;
  const double step = (b - a) / (double)(N - 1);
  return (double)j * step + a;
}

// knot t[i] of the not-a-knot cubic knot vector (length N + 4)
__device__ inline double nak_knot(double a, double b, int N, int i) {
  if (i <= 3) return a;
  if (i >= N) return b;
  return np_linspace_at(a, b, N, i - 2);
}

// The four nonzero cubic B-splines at v: basis[r] = B_{l-3+r}(v), l the knot
// interval (de Boor's triangular recurrence).
__device__ inline int cubic_basis(double v, double a, double b, int N, double* basis) {
  // interval l in [3, N-1] with t[l] <= v < t[l+1], clamped at the ends
  int l;
  if (!(v >= a)) {
    l = 3;
  } else if (!(v < b)) {
    l = N - 1;
  } else {
    const double step = (b - a) / (double)(N - 1);
    int j = (int)floor((v - a) / step);  // raw interval, then knot interval
    j = max(0, min(j, N - 2));
    // move to the interval that contains v exactly (linspace rounding)
    while (j > 0 && v < np_linspace_at(a, b, N, j)) --j;
    while (j < N - 2 && v >= np_linspace_at(a, b, N, j + 1)) ++j;
    l = min(max(j + 2, 3), N - 1);  // raw interval j lies in knot interval j+2 (ends merged)
  }
  double left[4], right[4];
  basis[0] = 1.0;
#pragma unroll
  for (int j = 1; j <= 3; ++j) {
    left[j] = v - nak_knot(a, b, N, l + 1 - j);
    right[j] = nak_knot(a, b, N, l + j) - v;
    double saved = 0.0;
#pragma unroll
    for (int r = 0; r < j; ++r) {
      const double temp = basis[r] / (right[r + 1] + left[j - r]);
      basis[r] = saved + right[r + 1] * temp;
      saved = left[j - r] * temp;
    }
    basis[j] = saved;
  }
  return l - 3;  // first coefficient index
}

constexpr int SPL_THREADS = 256;
constexpr int SPL_GROUP = 8;  // grids per work item

// out[i][gidx] = sum over the 4^d coefficients around x_i of c[gidx] times the
// tensor B-spline weights.  Work item = (point, group of SPL_GROUP grids).
template <int D>
__global__ __launch_bounds__(SPL_THREADS) void k_spline_eval(FlprGeom g,
                                                             const double* __restrict__ x,
                                                             int64_t n,
                                                             const double* __restrict__ c,
                                                             int ngrid, int64_t nodes,
                                                             double* __restrict__ out) {
  const int ngroups = (ngrid + SPL_GROUP - 1) / SPL_GROUP;
  const int64_t items = n * ngroups;
  for (int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; it < items;
       it += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = it / ngroups;
    const int gq = (int)(it - i * ngroups);
    double bw[3][4];
    int first[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < D; ++k)
      first[k] = cubic_basis(x[i * D + k], g.graw0[k], g.graw1[k], g.n_grid[k], bw[k]);
    const int g0 = gq * SPL_GROUP, g1 = min(ngrid, g0 + SPL_GROUP);
    for (int gi = g0; gi < g1; ++gi) {
      const double* cg = c + (int64_t)gi * nodes;
      double v = 0.0;
      if constexpr (D == 1) {
#pragma unroll
        for (int a = 0; a < 4; ++a) v += bw[0][a] * cg[first[0] + a];
      } else if constexpr (D == 2) {
        const int N1 = g.n_grid[1];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          double r = 0.0;
#pragma unroll
          for (int b = 0; b < 4; ++b) r += bw[1][b] * cg[(int64_t)(first[0] + a) * N1 + first[1] + b];
          v += bw[0][a] * r;
        }
      } else {
        const int N1 = g.n_grid[1], N2 = g.n_grid[2];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          double r = 0.0;
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            double q = 0.0;
            const int64_t row = ((int64_t)(first[0] + a) * N1 + first[1] + b) * N2 + first[2];
#pragma unroll
            for (int e = 0; e < 4; ++e) q += bw[2][e] * cg[row + e];
            r += bw[1][b] * q;
          }
          v += bw[0][a] * r;
        }
      }
      out[i * ngrid + gi] = v;
    }
  }
}


// ---------------------------------------------------------------------------
// GCV / DOF sums at the points for cubic (or mixed) interpolation
// ---------------------------------------------------------------------------

// One axis of the B-spline coefficient transform, batched over (problem,
// grid) = blockIdx.z: out[(o Nk + i) inner + c] = sum_s A[i][s] in[(o Nk + s)
// inner + c] with A the problem's inverse collocation matrix of that axis
// (a + (z / ngrid) * a_stride, row-major Nk x Nk).  A block owns 64 rows i x
// 64 columns j = (o, c); each thread 4 x 4 of them; A and in are staged
// through LDS in chunks of 16 s.
constexpr int AG_T = 64, AG_K = 16;

__global__ __launch_bounds__(256) void k_axis_gemm(const double* __restrict__ a, int64_t a_stride,
                                                   int ngrid, const double* __restrict__ in,
                                                   double* __restrict__ out, int64_t nodes,
                                                   int Nk, int64_t outer, int64_t inner) {
  __shared__ double As[AG_K][AG_T + 1];
  __shared__ double Bs[AG_K][AG_T + 1];
  const int64_t z = blockIdx.z;
  const double* A = a + (z / ngrid) * a_stride;
  const double* src = in + z * nodes;
  double* dst = out + z * nodes;
  const int64_t ncol = outer * inner;
  const int i0 = blockIdx.y * AG_T;
  const int64_t j0 = (int64_t)blockIdx.x * AG_T;
  const int t = threadIdx.x;
  const int ti = t % 16, tj = t / 16;
  double acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
  for (int s0 = 0; s0 < Nk; s0 += AG_K) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = t + 256 * q;
      {
        const int i = e / AG_K, sl = e % AG_K;
        const int gi = i0 + i, gs = s0 + sl;
        As[sl][i] = (gi < Nk && gs < Nk) ? A[(int64_t)gi * Nk + gs] : 0.0;
      }
      {
        int sl, j;
        if (inner == 1) {
          j = e / AG_K;
          sl = e % AG_K;
        } else {
          sl = e / AG_T;
          j = e % AG_T;
        }
        const int64_t gj = j0 + j;
        const int gs = s0 + sl;
        double v = 0.0;
        if (gj < ncol && gs < Nk) {
          const int64_t o = gj / inner, c = gj - o * inner;
          v = src[(o * Nk + gs) * inner + c];
        }
        Bs[sl][j] = v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int sl = 0; sl < AG_K; ++sl) {
      double av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) av[u] = As[sl][ti * 4 + u];
#pragma unroll
      for (int v = 0; v < 4; ++v) bv[v] = Bs[sl][tj * 4 + v];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = fma(av[u], bv[v], acc[u][v]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int gi = i0 + ti * 4 + u;
    if (gi >= Nk) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t gj = j0 + tj * 4 + v;
      if (gj >= ncol) continue;
      const int64_t o = gj / inner, c = gj - o * inner;
      dst[(o * Nk + gi) * inner + c] = acc[u][v];
    }
  }
}

constexpr int PS_THREADS = 256;
constexpr int PS_GROUP = 8;  // grids per block

// Per-grid sums at the data points of one problem (blockIdx.z) for the grid
// solutions m (dh, nval, nodes): value channels c < ny give the residual sums
// sum_i (y_ic - yhat_ic)^2, the probe channels the Hutchinson dots sum_i
// p_ik phat_ik (regression.py:1256-1280, api.py:661-723).  yhat / phat are
// the cubic spline of the coefficients coef (gcv_cubic / dof_cubic) or the
// multilinear interpolation of m.  blockIdx.y = group of PS_GROUP grids,
// blockIdx.x = a share of the points; each block writes its PS_GROUP sums to
// part, k_point_sums_final adds them in block order.
template <int D>
__global__ __launch_bounds__(PS_THREADS) void k_point_sums(
    FlprGeom g, const double* __restrict__ x, int64_t n, const double* __restrict__ y, int ny,
    const double* __restrict__ pr, int np, const double* __restrict__ m,
    const double* __restrict__ coef, int64_t nodes, int dh, int gcv_cubic, int dof_cubic,
    double* __restrict__ part, FlprBz bz) {
  __shared__ double red[PS_THREADS / 64];
  const int nval = ny + np;
  const int ngrid = dh * nval;
  int64_t pt = 0;
  if (bz.items) {
    const FlprBatchItem& it = bz.items[blockIdx.z];
    g = it.g;
    n = it.n;
    pt = it.pt;
  }
  x += pt * D;
  y += pt * ny;
  m += blockIdx.z * bz.s[0];
  coef += blockIdx.z * bz.s[0];
  const int gq = blockIdx.y;
  const int gl0 = gq * PS_GROUP;
  const bool any_cubic = gcv_cubic || dof_cubic;
  const bool any_linear = !gcv_cubic || !dof_cubic;
  double acc[PS_GROUP];
#pragma unroll
  for (int q = 0; q < PS_GROUP; ++q) acc[q] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double bw[3][4];
    int first[3] = {0, 0, 0};
    if (any_cubic) {
#pragma unroll
      for (int k = 0; k < D; ++k)
        first[k] = cubic_basis(x[i * D + k], g.graw0[k], g.graw1[k], g.n_grid[k], bw[k]);
    }
    CornersT<D> cr;
    if (any_linear) cr = corners<D>(g, x + i * D);
#pragma unroll
    for (int q = 0; q < PS_GROUP; ++q) {
      const int gi = gl0 + q;
      if (gi >= ngrid) break;
      const int c = gi % nval;
      const bool cubic = (c < ny) ? gcv_cubic : dof_cubic;
      double v = 0.0;
      if (cubic) {
        const double* cg = coef + (int64_t)gi * nodes;
        if constexpr (D == 1) {
#pragma unroll
          for (int a = 0; a < 4; ++a) v += bw[0][a] * cg[first[0] + a];
        } else if constexpr (D == 2) {
          const int N1 = g.n_grid[1];
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            double r = 0.0;
#pragma unroll
            for (int b = 0; b < 4; ++b)
              r += bw[1][b] * cg[(int64_t)(first[0] + a) * N1 + first[1] + b];
            v += bw[0][a] * r;
          }
        } else {
          const int N1 = g.n_grid[1], N2 = g.n_grid[2];
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            double r = 0.0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
              double qv = 0.0;
              const int64_t row = ((int64_t)(first[0] + a) * N1 + first[1] + b) * N2 + first[2];
#pragma unroll
              for (int e = 0; e < 4; ++e) qv += bw[2][e] * cg[row + e];
              r += bw[1][b] * qv;
            }
            v += bw[0][a] * r;
          }
        }
      } else {
        const double* mg = m + (int64_t)gi * nodes;
#pragma unroll
        for (int qq = 0; qq < (1 << D); ++qq) v += cr.w[qq] * mg[cr.node[qq]];
      }
      if (c < ny) {
        const double r = y[i * ny + c] - v;
        acc[q] += r * r;
      } else {
        acc[q] += pr[i * np + (c - ny)] * v;
      }
    }
  }
  double* dst = part + (((int64_t)blockIdx.z * gridDim.y + gq) * gridDim.x + blockIdx.x) * PS_GROUP;
#pragma unroll
  for (int q = 0; q < PS_GROUP; ++q) {
    const double t = block_sum(acc[q], red);
    if (threadIdx.x == 0) dst[q] = t;
  }
}

// Adds the blocks' sums in block order and writes them where flpr_select
// reads them: the probe dots at red[h nval + c], the residual sums at
// red[dh nval + h ny + c] (the y dots stay zero: with residual sums in place
// of the quadratic forms and zero sums of squares, select forms RSS itself).
__global__ __launch_bounds__(256) void k_point_sums_final(const double* __restrict__ part,
                                                          int nblk, int ngroups, int dh, int ny,
                                                          int np, double* __restrict__ red,
                                                          int64_t red_stride) {
  const int nval = ny + np;
  const int ngrid = dh * nval;
  const int64_t z = blockIdx.z;
  for (int gi = blockIdx.x * blockDim.x + threadIdx.x; gi < ngrid; gi += gridDim.x * blockDim.x) {
    const int gq = gi / PS_GROUP, q = gi % PS_GROUP;
    const double* src = part + ((z * ngroups + gq) * nblk) * PS_GROUP + q;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += src[(int64_t)b * PS_GROUP];
    const int h = gi / nval, c = gi % nval;
    double* r = red + z * red_stride;
    if (c < ny) r[(int64_t)dh * nval + h * ny + c] = s;
    else r[(int64_t)h * nval + c] = s;
  }
}

int lds_group(int64_t cells_per_channel, int channels, int cap, int64_t budget = 64 * 1024) {
  const int64_t per = cells_per_channel * (int64_t)sizeof(double);
  int cg = (int)(budget / per);
  if (cg < 1) cg = 1;
  if (cg > cap) cg = cap;
  if (cg > channels) cg = channels;
  return cg;
}

template <int D, typename A>
void launch_spread_t(int msp, dim3 grd, size_t lds, hipStream_t st, const FlprGeom& g,
                     const SpreadTab& tab, const int32_t* ks, const double* wab,
                     const double* vals, int64_t n, int wo, int C, int cg, int nsplit, A* part,
                     const FlprBz& bz) {
#define FLPR_SPREAD_CASE(M)                                                                    \
  case M:                                                                                      \
    hipLaunchKernelGGL((k_spread_cells<D, M, A>), grd, dim3(SPREAD_THREADS), lds, st, g, tab,  \
                       ks, wab, vals, n, wo, C, cg, nsplit, part, bz);                         \
    break;
  switch (msp) {
    FLPR_SPREAD_CASE(0)
    FLPR_SPREAD_CASE(1)
    FLPR_SPREAD_CASE(2)
    FLPR_SPREAD_CASE(3)
    FLPR_SPREAD_CASE(4)
    FLPR_SPREAD_CASE(5)
    FLPR_SPREAD_CASE(6)
    default:
      break;
  }
  if constexpr (D < 3) {
    switch (msp) {
      FLPR_SPREAD_CASE(7)
      FLPR_SPREAD_CASE(8)
      FLPR_SPREAD_CASE(9)
      FLPR_SPREAD_CASE(10)
      FLPR_SPREAD_CASE(11)
      FLPR_SPREAD_CASE(12)
      default:
        break;
    }
  }
#undef FLPR_SPREAD_CASE
}

// The spread of g (float32 compensated accumulation when g.spread_f32; the
// partial windows are then float32, half the bytes of part).
template <int D>
void launch_spread(int msp, dim3 grd, size_t lds, hipStream_t st, const FlprGeom& g,
                   const SpreadTab& tab, const int32_t* ks, const double* wab, const double* vals,
                   int64_t n, int wo, int C, int cg, int nsplit, double* part, const FlprBz& bz) {
  if (g.spread_f32)
    launch_spread_t<D, float>(msp, grd, lds / 2, st, g, tab, ks, wab, vals, n, wo, C, cg, nsplit,
                              reinterpret_cast<float*>(part), bz);
  else
    launch_spread_t<D, double>(msp, grd, lds, st, g, tab, ks, wab, vals, n, wo, C, cg, nsplit,
                               part, bz);
}

template <int D>
void launch_gather(dim3 gg, hipStream_t st, const FlprGeom& g, const double* part, int C,
                   int64_t cells, double* grid, const FlprBz& bz) {
  if (g.spread_f32)
    hipLaunchKernelGGL((k_spread_gather<D, float>), gg, dim3(256), 0, st, g,
                       reinterpret_cast<const float*>(part), C, cells, grid, bz);
  else
    hipLaunchKernelGGL((k_spread_gather<D, double>), gg, dim3(256), 0, st, g, part, C, cells,
                       grid, bz);
}

// Channels per block: all of them when the tiles alone fill the device, else
// split so that tiles x groups reaches MIN_BLOCKS; capped by the LDS budget.
int channel_group(int64_t ntiles, int64_t cells_per_channel, int channels) {
  int64_t want = (MIN_BLOCKS + ntiles - 1) / ntiles;
  if (want < 1) want = 1;
  int cg = (int)((channels + want - 1) / want);
  return lds_group(cells_per_channel, channels, cg < 1 ? 1 : cg);
}

}  // namespace

static thread_local int g_hip_error = 0;

void flpr_set_hip_error(hipError_t e) { g_hip_error = (int)e; }

extern "C" {

int flpr_last_hip_error(char* buf, int32_t len) {
  const int e = g_hip_error;
  if (buf && len > 0) {
    const char* msg = e ? hipGetErrorString((hipError_t)e) : "no error";
    int i = 0;
    for (; msg[i] && i < len - 1; ++i) buf[i] = msg[i];
    buf[i] = 0;
  }
  return e;
}

int flpr_xstats(const double* x, int64_t n, int32_t d, const double* y, int32_t ny,
                double* partials, double* out, void* stream) {
  FLPR_CLEAR_ERROR();
  if ((d > 0 && !x) || !partials || !out || n < 1 || d < 0 || d > FLPR_MAX_DIM || ny < 0 ||
      ny > MAXY || (ny > 0 && !y))
    return -1;
  hipLaunchKernelGGL(k_xstats, dim3(XS_BLOCKS), dim3(XS_THREADS), 0, as_stream(stream), x, n, d,
                     y, ny, partials, FlprBz{});
  FLPR_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_xstats_final, dim3(1), dim3(XS_THREADS), 0, as_stream(stream), partials,
                     XS_BLOCKS, d, ny, out, FlprBz{});
  FLPR_CHECK_LAUNCH();
  return 0;
}

// ---------------------------------------------------------------------------
// flpr_xstats without allocation: the block partials live in a file-scope
// device array instead of a caller-provided buffer.  Host code never passes
// the array's name to a kernel (on the host it is only a stand-in for the
// symbol): its device address is looked up once per device with
// hipGetSymbolAddress and kept in g_scratch_ptr.
// ---------------------------------------------------------------------------

namespace {
constexpr int XS_SCRATCH_DOUBLES = XS_BLOCKS * (5 * FLPR_MAX_DIM + 2 + MAXY);  // 196 KiB
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};
}  // namespace

static __device__ __align__(256) double scratch[XS_SCRATCH_DOUBLES];

// Device address of this device's copy of scratch (every GPU holds its own
// copy at its own address); looked up on the first call on each device.
static int scratch_on_current_device(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    flpr_set_hip_error(e);
    return -2;
  }
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return -1;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      flpr_set_hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol);
      return -2;
    }
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

int flpr_xstats_noalloc(const double* x, int64_t n, int32_t d, const double* y, int32_t ny,
                        double* out, void* stream) {
  FLPR_CLEAR_ERROR();
  if ((d > 0 && !x) || !out || n < 1 || d < 0 || ny < 0 || (ny > 0 && !y)) return -1;
  // the partials of d dimensions and ny columns must fit the scratch array
  if (d > FLPR_MAX_DIM || ny > MAXY ||
      (int64_t)XS_BLOCKS * (5 * d + 2 + ny) > (int64_t)XS_SCRATCH_DOUBLES)
    return -3;
  double* part = nullptr;
  const int rc = scratch_on_current_device(&part);
  if (rc != 0) return rc;
  // k_xstats writes every partial slot before k_xstats_final reads it, so
  // the contents left by an earlier call need no clearing
  hipLaunchKernelGGL(k_xstats, dim3(XS_BLOCKS), dim3(XS_THREADS), 0, as_stream(stream), x, n, d,
                     y, ny, part, FlprBz{});
  FLPR_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_xstats_final, dim3(1), dim3(XS_THREADS), 0, as_stream(stream), part,
                     XS_BLOCKS, d, ny, out, FlprBz{});
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_points_prepare(const double* x, int64_t n, const FlprGeom* g, int32_t* key,
                        int32_t* key_count, int32_t* rank, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!x || !key || !key_count || n < 1 || !flpr_geom_ok(g)) return -1;
  if (flpr_nkeys(g) > ((int64_t)1 << 30)) return -1;
  hipStream_t st = as_stream(stream);
  const int64_t nk = flpr_nkeys(g);
  if (rank && nk <= PREP_LDS_KEYS) {
    dim3 gl(prep_lds_blocks(n, nk));
    const size_t lds = (size_t)nk * sizeof(int);
    switch (g->d) {
      case 1: hipLaunchKernelGGL(k_prepare_lds<1>, gl, dim3(PREP_THREADS), lds, st, x, n, *g, key, key_count, rank, (int)nk, FlprBz{}); break;
      case 2: hipLaunchKernelGGL(k_prepare_lds<2>, gl, dim3(PREP_THREADS), lds, st, x, n, *g, key, key_count, rank, (int)nk, FlprBz{}); break;
      default: hipLaunchKernelGGL(k_prepare_lds<3>, gl, dim3(PREP_THREADS), lds, st, x, n, *g, key, key_count, rank, (int)nk, FlprBz{}); break;
    }
    FLPR_CHECK_LAUNCH();
    return 0;
  }
  dim3 grd(flpr_grid_for(n, PREP_THREADS, 4096));
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_prepare<1>, grd, dim3(PREP_THREADS), 0, st, x, n, *g, key, key_count, rank, FlprBz{}); break;
    case 2: hipLaunchKernelGGL(k_prepare<2>, grd, dim3(PREP_THREADS), 0, st, x, n, *g, key, key_count, rank, FlprBz{}); break;
    default: hipLaunchKernelGGL(k_prepare<3>, grd, dim3(PREP_THREADS), 0, st, x, n, *g, key, key_count, rank, FlprBz{}); break;
  }
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_points_bin(const int32_t* key, int64_t n, int32_t nkeys, const int32_t* key_count,
                    int32_t* key_start, int32_t* key_cursor, const int32_t* rank,
                    int32_t* order, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!key || !key_count || !key_start || !key_cursor || !order || n < 1 || nkeys < 1 ||
      n > INT32_MAX)
    return -1;
  // key_cursor doubles as the block-sum scratch of the scan; k_scan_apply
  // clears each entry again before the scatter uses the cursors
  const int nb = (nkeys + SCAN_THREADS - 1) / SCAN_THREADS;
  if (nb > nkeys) return -3;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(SCAN_THREADS), 0, st, key_count, nkeys,
                     key_cursor);
  FLPR_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_THREADS), 0, st, key_cursor, nb);
  FLPR_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_THREADS), 0, st, key_count, nkeys,
                     key_cursor, key_start);
  FLPR_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_bin, dim3(flpr_grid_for(n, 256, 4096)), dim3(256), 0, as_stream(stream),
                     key, n, key_start, key_cursor, rank, order);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_points_pack(const FlprGeom* g, const double* x, const int32_t* order, int64_t n,
                     const double* vals_y, int32_t ny, const double* probes, int32_t nprobe,
                     double* wab, double* vals, int32_t* rawi, double* rawf, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !x || !order || (!rawi != !rawf) || n < 1 ||
      ny < 0 || nprobe < 0 || ((ny + nprobe) > 0 && !vals) || (ny > 0 && !vals_y) ||
      (nprobe > 0 && !probes))
    return -1;
  dim3 grd(flpr_grid_for(n, 256, 4096));
  hipStream_t st = as_stream(stream);
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_pack<1>, grd, dim3(256), 0, st, *g, x, order, n, vals_y, ny, probes, nprobe, wab, vals, rawi, rawf, FlprBz{}); break;
    case 2: hipLaunchKernelGGL(k_pack<2>, grd, dim3(256), 0, st, *g, x, order, n, vals_y, ny, probes, nprobe, wab, vals, rawi, rawf, FlprBz{}); break;
    default: hipLaunchKernelGGL(k_pack<3>, grd, dim3(256), 0, st, *g, x, order, n, vals_y, ny, probes, nprobe, wab, vals, rawi, rawf, FlprBz{}); break;
  }
  FLPR_CHECK_LAUNCH();
  return 0;
}

// G[k][j] = exp(-hx_k^2 (j - msp)^2 / (4 tau_k)) of a geometry.
static SpreadTab spread_tab(const FlprGeom& g) {
  SpreadTab tab;
  for (int k = 0; k < FLPR_MAX_DIM; ++k)
    for (int j = 0; j < SPREAD_MAX_TAPS; ++j) {
      const double c2 = (k < g.d) ? g.hx[k] * g.hx[k] * g.inv4tau[k] : 0.0;
      const double o = (double)(j - g.msp);
      tab.G[k][j] = exp(-c2 * o * o);
    }
  return tab;
}

int flpr_spread_wide(const FlprGeom* g, const int32_t* key_start, const double* wab,
                     const double* vals, int64_t n, int32_t with_ones, int32_t nval,
                     double* grid, double* sums, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !key_start || !wab || !grid || !sums || n < 1 || nval < 0 ||
      (nval > 0 && !vals))
    return -1;
  if (g->d != 3 || g->msp < 7 || g->msp > FLPR_MAX_MSP) return -1;
  const int wo = with_ones ? 1 : 0;
  const int C = wo + nval;
  if (C < 1 || C > FLPR_MAX_CHANNELS) return -1;
  int64_t cells = 1;
  for (int k = 0; k < 3; ++k) cells *= g->G[k];
  const SpreadTab tab = spread_tab(*g);
  const int64_t ntiles = flpr_ntiles(g);
  if (ntiles > INT32_MAX) return -1;
  hipStream_t st = as_stream(stream);
#define FLPR_WIDE_CASE(M)                                                                     \
  case M:                                                                                     \
    hipLaunchKernelGGL((k_spread_wide3<M>), dim3((unsigned)ntiles), dim3(SPREAD_THREADS), 0, \
                       st, *g, tab, key_start, wab, vals, wo, C, grid, cells);                \
    break;
  switch (g->msp) {
    FLPR_WIDE_CASE(7)
    FLPR_WIDE_CASE(8)
    FLPR_WIDE_CASE(9)
    FLPR_WIDE_CASE(10)
    FLPR_WIDE_CASE(11)
    default:
      FLPR_WIDE_CASE(12)
  }
#undef FLPR_WIDE_CASE
  FLPR_CHECK_LAUNCH();
  launch_grid_sums(st, grid, cells, (int64_t)C * cells, C, 1, sums, nullptr);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_spread(const FlprGeom* g, const int32_t* key_start, const double* wab,
                const double* vals, int64_t n, int32_t with_ones, int32_t nval, double* part,
                double* grid, double* sums, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !key_start || !wab || !part || !grid || !sums || n < 1 || nval < 0 ||
      (nval > 0 && !vals))
    return -1;
  // footprints wider than 13 cells only in 1-D and 2-D (the 3-D window would
  // not fit the LDS)
  if (g->msp < 0 || g->msp > (g->d == 3 ? 6 : FLPR_MAX_MSP)) return -1;
  const int wo = with_ones ? 1 : 0;
  const int C = wo + nval;
  if (C < 1 || C > FLPR_MAX_CHANNELS) return -1;
  int64_t fcells = 1, cells = 1;
  for (int k = 0; k < g->d; ++k) {
    fcells *= g->tile[k] + 2 * g->msp;
    cells *= g->G[k];
  }
  const int64_t ntiles = flpr_ntiles(g);
  const int cg = channel_group(ntiles, fcells, C);
  const size_t lds = (size_t)fcells * cg * sizeof(double);
  if (lds > 64 * 1024) return -1;
  const SpreadTab tab = spread_tab(*g);
  int64_t tc = 1;
  for (int k = 0; k < g->d; ++k) tc *= g->tile[k];
  // footprint rows per item (k_spread_cells' ROWS)
  const int ntap = 2 * g->msp + 1;
  const int rows = (g->d == 1) ? 1 : (g->d == 2 ? (g->msp > 4 ? ntap : 1)
                                               : (g->msp > 4 ? ntap * ntap : ntap));
  const int nsplit = fill_splits(item_splits(g, n), tc * rows * std::min(cg, C), SPREAD_THREADS);
  dim3 grd((unsigned)ntiles, (unsigned)((C + cg - 1) / cg));
  hipStream_t st = as_stream(stream);
  const int msp = g->msp;
  switch (g->d) {
    case 1: launch_spread<1>(msp, grd, lds, st, *g, tab, key_start, wab, vals, n, wo, C, cg, nsplit, part, FlprBz{}); break;
    case 2: launch_spread<2>(msp, grd, lds, st, *g, tab, key_start, wab, vals, n, wo, C, cg, nsplit, part, FlprBz{}); break;
    default: launch_spread<3>(msp, grd, lds, st, *g, tab, key_start, wab, vals, n, wo, C, cg, nsplit, part, FlprBz{}); break;
  }
  FLPR_CHECK_LAUNCH();
  dim3 gg(flpr_grid_for(cells, 256, 256), (unsigned)C);
  switch (g->d) {
    case 1: launch_gather<1>(gg, st, *g, part, C, cells, grid, FlprBz{}); break;
    case 2: launch_gather<2>(gg, st, *g, part, C, cells, grid, FlprBz{}); break;
    default: launch_gather<3>(gg, st, *g, part, C, cells, grid, FlprBz{}); break;
  }
  FLPR_CHECK_LAUNCH();
  // the tile windows in part are dead once gathered: the channel totals'
  // partials go there (ntiles * C * fcells >= C * CSUM_MAX_PARTS doubles, or
  // one block per channel)
  launch_grid_sums(st, grid, cells, (int64_t)C * cells, C, 1, sums,
                   ntiles * fcells >= CSUM_MAX_PARTS ? part : nullptr);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_grid_periodize(const FlprGeom* g, double* grid, int32_t C, double* work,
                        void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !grid || !work || C < 1) return -1;
  int64_t cells = 1, pcells = 1;
  for (int k = 0; k < g->d; ++k) {
    cells *= g->G[k];
    pcells *= g->L[k];
  }
  hipStream_t st = as_stream(stream);
  dim3 g1(flpr_grid_for(pcells, 256, 256), (unsigned)C), g2(flpr_grid_for(cells, 256, 256), (unsigned)C);
  switch (g->d) {
    case 1:
      hipLaunchKernelGGL(k_periodic_sum<1>, g1, dim3(256), 0, st, *g, grid, cells, pcells, work);
      hipLaunchKernelGGL(k_periodic_put<1>, g2, dim3(256), 0, st, *g, grid, cells, pcells, work);
      break;
    case 2:
      hipLaunchKernelGGL(k_periodic_sum<2>, g1, dim3(256), 0, st, *g, grid, cells, pcells, work);
      hipLaunchKernelGGL(k_periodic_put<2>, g2, dim3(256), 0, st, *g, grid, cells, pcells, work);
      break;
    default:
      hipLaunchKernelGGL(k_periodic_sum<3>, g1, dim3(256), 0, st, *g, grid, cells, pcells, work);
      hipLaunchKernelGGL(k_periodic_put<3>, g2, dim3(256), 0, st, *g, grid, cells, pcells, work);
      break;
  }
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_grid_center(double* grid, int32_t C, int64_t cells, const double* sums, double* shift,
                     float* grid32, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!grid || !sums || C < 1 || cells < 1) return -1;
  dim3 grd(flpr_grid_for(cells, 256, 256), (unsigned)C);
  hipLaunchKernelGGL(k_grid_center, grd, dim3(256), 0, as_stream(stream), grid, cells, sums, shift,
                     grid32, FlprBz{});
  FLPR_CHECK_LAUNCH();
  return 0;
}

// The walk form of the adjoint sums (every axis with an even node count) for
// P problems of one geometry shape, problem p's key table at p * key_stride,
// its q, gram and walk workspace at p * q_stride, gram_stride and
// part_stride (P = 1: strides unused).  part: nkeys * nval * 2^d doubles per
// problem.
static int launch_adjoint_walk(const FlprGeom* g, int P, hipStream_t st, const int32_t* key_start,
                        const int32_t* rawi, const double* rawf, const double* vals, int nval,
                        double* part, double* q, double* gram, int64_t key_stride,
                        int64_t q_stride, int64_t gram_stride, int64_t part_stride) {
  int64_t tcells = 1, nodes = 1, rcells = 1;
  for (int k = 0; k < g->d; ++k) {
    tcells *= g->tile[k];
    nodes *= g->n_grid[k];
    rcells *= g->n_grid[k] - 1;
  }
  const int64_t nkeys = flpr_ntiles(g) * tcells;
#define FLPR_WALK(M, out_, stride_, nch_, bz_)                                                   \
    do {                                                                                        \
      dim3 grd(flpr_grid_for(nkeys * (nch_) * WALK_LANES, 256, 16384), 1, (unsigned)P);        \
      switch (g->d) {                                                                           \
        case 1: hipLaunchKernelGGL((k_adjoint_walk<1, M>), grd, dim3(256), 0, st, *g, key_start, \
                                   rawi, rawf, vals, nval, nkeys, out_, stride_, part, bz_);    \
          break;                                                                                \
        case 2: hipLaunchKernelGGL((k_adjoint_walk<2, M>), grd, dim3(256), 0, st, *g, key_start, \
                                   rawi, rawf, vals, nval, nkeys, out_, stride_, part, bz_);    \
          break;                                                                                \
        default: hipLaunchKernelGGL((k_adjoint_walk<3, M>), grd, dim3(256), 0, st, *g,           \
                                    key_start, rawi, rawf, vals, nval, nkeys, out_, stride_,    \
                                    part, bz_);                                                 \
          break;                                                                                \
      }                                                                                         \
      FLPR_CHECK_LAUNCH();                                                                      \
    } while (0)
  if (gram) {
    FlprBz bw{};
    bw.s[0] = key_stride;
    bw.s[1] = gram_stride;
    FLPR_WALK(1, gram, rcells, 1, bw);
  }
  if (nval > 0) {
    // points outside their cell's main interval add into q directly
    FlprBz bw{};
    bw.s[0] = key_stride;
    bw.s[1] = q_stride;
    bw.s[2] = part_stride;
    FLPR_WALK(0, q, nodes, nval, bw);
    FlprBz bg{};
    bg.s[0] = part_stride;
    bg.s[1] = q_stride;
    dim3 gg(flpr_grid_for(nodes, 256, 1024), (unsigned)nval, (unsigned)P);
    switch (g->d) {
      case 1: hipLaunchKernelGGL(k_adjoint_walk_gather<1>, gg, dim3(256), 0, st, *g, part, nval, nodes, q, bg); break;
      case 2: hipLaunchKernelGGL(k_adjoint_walk_gather<2>, gg, dim3(256), 0, st, *g, part, nval, nodes, q, bg); break;
      default: hipLaunchKernelGGL(k_adjoint_walk_gather<3>, gg, dim3(256), 0, st, *g, part, nval, nodes, q, bg); break;
    }
    FLPR_CHECK_LAUNCH();
  }
#undef FLPR_WALK
  return 0;
}

int flpr_adjoint(const FlprGeom* g, const int32_t* key_start, const int32_t* rawi,
                 const double* rawf, const double* vals, int64_t n, int32_t nval, double* part,
                 double* q, double* gram, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !key_start || !rawi || !rawf || !part || n < 1 || nval < 0 ||
      (nval > 0 && (!vals || !q)))
    return -1;
  int64_t wcells = 1, nodes = 1, rcells = 1;
  for (int k = 0; k < g->d; ++k) {
    wcells *= g->tile[k] + 5;
    nodes *= g->n_grid[k];
    rcells *= g->n_grid[k] - 1;
  }
  const int nq = 1 << g->d;
  const int nw = nq * (nq + 1) / 2;
  const int nsplit0 = item_splits(g, n);
  hipStream_t st = as_stream(stream);
  // even node counts: both sums by k_adjoint_walk
  bool aligned = true;
  for (int k = 0; k < g->d; ++k) aligned = aligned && !(g->n_grid[k] & 1);
  if (aligned)
    return launch_adjoint_walk(g, 1, st, key_start, rawi, rawf, vals, nval, part, q, gram, 0, 0,
                               0, 0);
  for (int mode = 0; mode < 2; ++mode) {
    const int nch = (mode == 0) ? nval : (gram ? nw : 0);
    if (nch == 0) continue;
    double* out = (mode == 0) ? q : gram;
    const int64_t ostride = (mode == 0) ? nodes : rcells;
    const int cg = channel_group(flpr_ntiles(g), wcells, nch);
    int64_t tc = 1;
    for (int k = 0; k < g->d; ++k) tc *= g->tile[k];
    const int nsplit = fill_splits(nsplit0, tc * std::min(cg, nch), ADJ_THREADS);
    const size_t lds = (size_t)wcells * cg * sizeof(double);
    if (lds > 64 * 1024) return -1;
    dim3 grd((unsigned)flpr_ntiles(g), (unsigned)((nch + cg - 1) / cg));
    switch (g->d) {
      case 1: hipLaunchKernelGGL(k_adjoint<1>, grd, dim3(ADJ_THREADS), lds, st, *g, key_start, rawi, rawf, vals, n, nch, mode, cg, nsplit, out, ostride, part, FlprBz{}); break;
      case 2: hipLaunchKernelGGL(k_adjoint<2>, grd, dim3(ADJ_THREADS), lds, st, *g, key_start, rawi, rawf, vals, n, nch, mode, cg, nsplit, out, ostride, part, FlprBz{}); break;
      default: hipLaunchKernelGGL(k_adjoint<3>, grd, dim3(ADJ_THREADS), lds, st, *g, key_start, rawi, rawf, vals, n, nch, mode, cg, nsplit, out, ostride, part, FlprBz{}); break;
    }
    FLPR_CHECK_LAUNCH();
    dim3 gg(flpr_grid_for(ostride, 256, 256), (unsigned)nch);
    switch (g->d) {
      case 1: hipLaunchKernelGGL(k_adjoint_gather<1>, gg, dim3(256), 0, st, *g, part, nch, mode, ostride, out, FlprBz{}); break;
      case 2: hipLaunchKernelGGL(k_adjoint_gather<2>, gg, dim3(256), 0, st, *g, part, nch, mode, ostride, out, FlprBz{}); break;
      default: hipLaunchKernelGGL(k_adjoint_gather<3>, gg, dim3(256), 0, st, *g, part, nch, mode, ostride, out, FlprBz{}); break;
    }
    FLPR_CHECK_LAUNCH();
  }
  return 0;
}

int flpr_quadform(const FlprGeom* g, const double* m, int32_t ncol, int32_t dh,
                  const double* gram, double* quad, double* sum_part, int64_t sum_part_len,
                  void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !m || !gram || !quad || ncol < 1 || dh < 1) return -1;
  const int64_t nodes = flpr_nodes(g);
  int64_t rcells = 1;
  for (int k = 0; k < g->d; ++k) rcells *= g->n_grid[k] - 1;
  const int nb = sum_part ? flpr_sum_blocks((int64_t)dh * ncol, 1,
                                            flpr_grid_for(rcells, QF_THREADS, QF_BLOCKS), sum_part_len)
                          : 0;
  if (nb < 1) return -3;
  dim3 grd((unsigned)nb, (unsigned)(dh * ncol));
  hipStream_t st = as_stream(stream);
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_quadform<1>, grd, dim3(QF_THREADS), 0, st, *g, m, nodes, gram, rcells, quad, sum_part, FlprBz{}); break;
    case 2: hipLaunchKernelGGL(k_quadform<2>, grd, dim3(QF_THREADS), 0, st, *g, m, nodes, gram, rcells, quad, sum_part, FlprBz{}); break;
    default: hipLaunchKernelGGL(k_quadform<3>, grd, dim3(QF_THREADS), 0, st, *g, m, nodes, gram, rcells, quad, sum_part, FlprBz{}); break;
  }
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_interp_eval(const FlprGeom* g, const double* x, int64_t n, const double* m, int32_t ncol,
                     double* out, const double* yvar, double* expsum, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !x || !m || !out || n < 1 || ncol < 1) return -1;
  if (yvar && !expsum) return -1;
  dim3 grd((unsigned)flpr_grid_for(n, EVAL_THREADS, 2048));
  const int64_t nodes = flpr_nodes(g);
  hipStream_t st = as_stream(stream);
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_interp_eval<1>, grd, dim3(EVAL_THREADS), 0, st, *g, x, n, m, ncol, nodes, out, yvar, expsum); break;
    case 2: hipLaunchKernelGGL(k_interp_eval<2>, grd, dim3(EVAL_THREADS), 0, st, *g, x, n, m, ncol, nodes, out, yvar, expsum); break;
    default: hipLaunchKernelGGL(k_interp_eval<3>, grd, dim3(EVAL_THREADS), 0, st, *g, x, n, m, ncol, nodes, out, yvar, expsum); break;
  }
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_fit_output(const FlprGeom* g, const double* x, int64_t n, const double* m_all,
                    int32_t ncol, const int32_t* sel, const double* yref, int32_t variance,
                    double* yhat, double* resid, double* grid, double* part, double* fac,
                    void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !x || !m_all || !yhat || n < 1 || ncol < 1 || (resid && !yref) ||
      (variance && (ncol != 1 || !yref || !grid || !part)))
    return -1;
  dim3 grd((unsigned)flpr_grid_for(n, EVAL_THREADS, FIT_BLOCKS));
  const int64_t nodes = flpr_nodes(g);
  hipStream_t st = as_stream(stream);
  const int var = variance ? 1 : 0;
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_fit_points<1>, grd, dim3(EVAL_THREADS), 0, st, *g, x, n, m_all, ncol, nodes, sel, yref, var, yhat, resid, part, FlprBz{}); break;
    case 2: hipLaunchKernelGGL(k_fit_points<2>, grd, dim3(EVAL_THREADS), 0, st, *g, x, n, m_all, ncol, nodes, sel, yref, var, yhat, resid, part, FlprBz{}); break;
    default: hipLaunchKernelGGL(k_fit_points<3>, grd, dim3(EVAL_THREADS), 0, st, *g, x, n, m_all, ncol, nodes, sel, yref, var, yhat, resid, part, FlprBz{}); break;
  }
  FLPR_CHECK_LAUNCH();
  if (variance) {
    hipLaunchKernelGGL(k_fit_finish, grd, dim3(EVAL_THREADS), 0, st, n, m_all, nodes, sel, yref,
                       part, (int)grd.x, yhat, resid, grid, fac, FlprBz{});
    FLPR_CHECK_LAUNCH();
  }
  return 0;
}

int flpr_lcv(const FlprGeom* g, const double* x, const int32_t* order, int64_t n,
             const double* dens, int32_t dh, const double* k0h, double* part, double* out,
             void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !x || !dens || !k0h || !part || !out || n < 2 || dh < 1) return -1;
  const int64_t nodes = flpr_nodes(g);
  hipStream_t st = as_stream(stream);
  dim3 grd(LCV_BLOCKS);
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_lcv<1>, grd, dim3(LCV_THREADS), 0, st, *g, x, order, n, dens, dh, nodes, k0h, part); break;
    case 2: hipLaunchKernelGGL(k_lcv<2>, grd, dim3(LCV_THREADS), 0, st, *g, x, order, n, dens, dh, nodes, k0h, part); break;
    default: hipLaunchKernelGGL(k_lcv<3>, grd, dim3(LCV_THREADS), 0, st, *g, x, order, n, dens, dh, nodes, k0h, part); break;
  }
  FLPR_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_lcv_final, dim3(dh), dim3(256), 0, st, part, LCV_BLOCKS, dh, out);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_spline_eval(const FlprGeom* g, const double* x, int64_t n, const double* coef,
                     int32_t ngrid, double* out, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !x || !coef || !out || n < 1 || ngrid < 1) return -1;
  for (int k = 0; k < g->d; ++k)
    if (g->n_grid[k] < 4) return -1;
  const int64_t nodes = flpr_nodes(g);
  const int64_t items = n * ((ngrid + SPL_GROUP - 1) / SPL_GROUP);
  dim3 grd(flpr_grid_for(items, SPL_THREADS, 4096));
  hipStream_t st = as_stream(stream);
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_spline_eval<1>, grd, dim3(SPL_THREADS), 0, st, *g, x, n, coef, ngrid, nodes, out); break;
    case 2: hipLaunchKernelGGL(k_spline_eval<2>, grd, dim3(SPL_THREADS), 0, st, *g, x, n, coef, ngrid, nodes, out); break;
    default: hipLaunchKernelGGL(k_spline_eval<3>, grd, dim3(SPL_THREADS), 0, st, *g, x, n, coef, ngrid, nodes, out); break;
  }
  FLPR_CHECK_LAUNCH();
  return 0;
}

// ---------------------------------------------------------------------------
// batches of independent problems: the single-problem kernels with the
// problem on blockIdx.z
// ---------------------------------------------------------------------------

static bool batch_ok(const FlprGeom* g, const FlprBatchItem* items, int32_t P, int64_t nmax) {
  return flpr_geom_ok(g) && items && P >= 1 && P <= 65535 && nmax >= 1;
}

int flpr_fill_zero(void* p, int64_t bytes, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!p || bytes < 0 || (bytes & 3)) return -1;
  if (bytes == 0) return 0;
  const int64_t words = bytes / 4;
  hipLaunchKernelGGL(k_fill_zero, dim3(flpr_grid_for(words, 256, 4096)), dim3(256), 0,
                     as_stream(stream), (uint32_t*)p, words);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_batch_xstats(const FlprBatchItem* items, int32_t P, const double* x, int32_t d,
                      const double* y, int32_t ny, double* partials, double* out, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!items || P < 1 || P > 65535 || (d > 0 && !x) || !partials || !out || d < 0 ||
      d > FLPR_MAX_DIM || ny < 0 || ny > MAXY || (ny > 0 && !y) || d + ny < 1)
    return -1;
  const int per = 5 * d + 2 + ny;
  FlprBz bz{};
  bz.items = items;
  bz.s[0] = (int64_t)XS_BLOCKS * per;
  hipLaunchKernelGGL(k_xstats, dim3(XS_BLOCKS, 1, P), dim3(XS_THREADS), 0, as_stream(stream), x,
                     (int64_t)0, d, y, ny, partials, bz);
  FLPR_CHECK_LAUNCH();
  FlprBz bf{};
  bf.s[0] = (int64_t)XS_BLOCKS * per;
  bf.s[1] = per;
  hipLaunchKernelGGL(k_xstats_final, dim3(1, 1, P), dim3(XS_THREADS), 0, as_stream(stream),
                     partials, XS_BLOCKS, d, ny, out, bf);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_batch_points_prepare(const FlprGeom* g, const FlprBatchItem* items, int32_t P,
                              int64_t nmax, const double* x, int32_t* key, int32_t* key_count,
                              int32_t* rank, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!batch_ok(g, items, P, nmax) || !x || !key || !key_count) return -1;
  if (flpr_nkeys(g) * P > ((int64_t)1 << 30)) return -1;
  FlprBz bz{};
  bz.items = items;
  bz.s[0] = flpr_nkeys(g);
  hipStream_t st = as_stream(stream);
  const int64_t nk = flpr_nkeys(g);
  if (rank && nk <= PREP_LDS_KEYS) {
    dim3 gl(prep_lds_blocks(nmax, nk), 1, P);
    const size_t lds = (size_t)nk * sizeof(int);
    switch (g->d) {
      case 1: hipLaunchKernelGGL(k_prepare_lds<1>, gl, dim3(PREP_THREADS), lds, st, x, nmax, *g, key, key_count, rank, (int)nk, bz); break;
      case 2: hipLaunchKernelGGL(k_prepare_lds<2>, gl, dim3(PREP_THREADS), lds, st, x, nmax, *g, key, key_count, rank, (int)nk, bz); break;
      default: hipLaunchKernelGGL(k_prepare_lds<3>, gl, dim3(PREP_THREADS), lds, st, x, nmax, *g, key, key_count, rank, (int)nk, bz); break;
    }
    FLPR_CHECK_LAUNCH();
    return 0;
  }
  dim3 grd(flpr_grid_for(nmax, PREP_THREADS, 4096), 1, P);
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_prepare<1>, grd, dim3(PREP_THREADS), 0, st, x, nmax, *g, key, key_count, rank, bz); break;
    case 2: hipLaunchKernelGGL(k_prepare<2>, grd, dim3(PREP_THREADS), 0, st, x, nmax, *g, key, key_count, rank, bz); break;
    default: hipLaunchKernelGGL(k_prepare<3>, grd, dim3(PREP_THREADS), 0, st, x, nmax, *g, key, key_count, rank, bz); break;
  }
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_batch_points_pack(const FlprGeom* g, const FlprBatchItem* items, int32_t P, int64_t nmax,
                           const double* x, const int32_t* order, const double* vals_y, int32_t ny,
                           const double* probes, int32_t nprobe, double* wab, double* vals,
                           int32_t* rawi, double* rawf, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!batch_ok(g, items, P, nmax) || !x || !order || (!rawi != !rawf) || ny < 0 ||
      nprobe < 0 || ((ny + nprobe) > 0 && !vals) || (ny > 0 && !vals_y) ||
      (nprobe > 0 && !probes))
    return -1;
  FlprBz bz{};
  bz.items = items;
  dim3 grd(flpr_grid_for(nmax * std::max(1, ny + nprobe), 256, 4096), 1, P);
  hipStream_t st = as_stream(stream);
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_pack<1>, grd, dim3(256), 0, st, *g, x, order, nmax, vals_y, ny, probes, nprobe, wab, vals, rawi, rawf, bz); break;
    case 2: hipLaunchKernelGGL(k_pack<2>, grd, dim3(256), 0, st, *g, x, order, nmax, vals_y, ny, probes, nprobe, wab, vals, rawi, rawf, bz); break;
    default: hipLaunchKernelGGL(k_pack<3>, grd, dim3(256), 0, st, *g, x, order, nmax, vals_y, ny, probes, nprobe, wab, vals, rawi, rawf, bz); break;
  }
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_batch_spread(const FlprGeom* g, const FlprBatchItem* items, int32_t P, int64_t nmax,
                      const int32_t* key_start, const double* wab, const double* vals,
                      int32_t with_ones, int32_t nval, double* part, double* grid,
                      int64_t grid_stride, double* sums, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!batch_ok(g, items, P, nmax) || !key_start || !wab || !part || !grid || !sums ||
      nval < 0 || (nval > 0 && !vals))
    return -1;
  if (g->msp > (g->d == 3 ? 6 : FLPR_MAX_MSP)) return -1;
  const int wo = with_ones ? 1 : 0;
  const int C = wo + nval;
  if (C < 1 || C > FLPR_MAX_CHANNELS) return -1;
  int64_t fcells = 1, cells = 1;
  for (int k = 0; k < g->d; ++k) {
    fcells *= g->tile[k] + 2 * g->msp;
    cells *= g->G[k];
  }
  if (grid_stride < C * cells) return -1;
  const int64_t ntiles = flpr_ntiles(g);
  // the channel grouping of one problem: every problem of the batch is
  // spread exactly as a single call would spread it
  const int cg = channel_group(ntiles, fcells, C);
  const size_t lds = (size_t)fcells * cg * sizeof(double);
  if (lds > 64 * 1024) return -1;
  SpreadTab tab;
  for (int k = 0; k < FLPR_MAX_DIM; ++k)
    for (int j = 0; j < SPREAD_MAX_TAPS; ++j) {
      const double c2 = (k < g->d) ? g->hx[k] * g->hx[k] * g->inv4tau[k] : 0.0;
      const double o = (double)(j - g->msp);
      tab.G[k][j] = exp(-c2 * o * o);
    }
  FlprBz bz{};
  bz.items = items;
  bz.s[0] = flpr_nkeys(g);
  bz.s[1] = ntiles * C * fcells;
  dim3 grd((unsigned)ntiles, (unsigned)((C + cg - 1) / cg), (unsigned)P);
  hipStream_t st = as_stream(stream);
  const int msp = g->msp;
  switch (g->d) {
    case 1: launch_spread<1>(msp, grd, lds, st, *g, tab, key_start, wab, vals, nmax, wo, C, cg, 1, part, bz); break;
    case 2: launch_spread<2>(msp, grd, lds, st, *g, tab, key_start, wab, vals, nmax, wo, C, cg, 1, part, bz); break;
    default: launch_spread<3>(msp, grd, lds, st, *g, tab, key_start, wab, vals, nmax, wo, C, cg, 1, part, bz); break;
  }
  FLPR_CHECK_LAUNCH();
  FlprBz bg{};
  bg.s[0] = ntiles * C * fcells;
  bg.s[1] = grid_stride;
  dim3 gg(flpr_grid_for(cells, 256, 256), (unsigned)C, (unsigned)P);
  switch (g->d) {
    case 1: launch_gather<1>(gg, st, *g, part, C, cells, grid, bg); break;
    case 2: launch_gather<2>(gg, st, *g, part, C, cells, grid, bg); break;
    default: launch_gather<3>(gg, st, *g, part, C, cells, grid, bg); break;
  }
  FLPR_CHECK_LAUNCH();
  launch_grid_sums(st, grid, cells, grid_stride, C, P, sums,
                   ntiles * fcells >= CSUM_MAX_PARTS ? part : nullptr);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_batch_grid_center(int32_t P, double* grid, int64_t grid_stride, int32_t C, int64_t cells,
                           const double* sums, double* shift, float* grid32, void* stream) {
  FLPR_CLEAR_ERROR();
  if (P < 1 || P > 65535 || !grid || !sums || C < 1 || cells < 1 || grid_stride < C * cells)
    return -1;
  FlprBz bz{};
  bz.s[0] = grid_stride;
  bz.s[1] = C;
  dim3 grd(flpr_grid_for(cells, 256, 256), (unsigned)C, (unsigned)P);
  hipLaunchKernelGGL(k_grid_center, grd, dim3(256), 0, as_stream(stream), grid, cells, sums, shift,
                     grid32, bz);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_batch_adjoint(const FlprGeom* g, const FlprBatchItem* items, int32_t P, int64_t nmax,
                       const int32_t* key_start, const int32_t* rawi, const double* rawf,
                       const double* vals, int32_t nval, double* part, double* q, double* gram,
                       void* stream) {
  FLPR_CLEAR_ERROR();
  if (!batch_ok(g, items, P, nmax) || !key_start || !rawi || !rawf || !part || nval < 0 ||
      (nval > 0 && (!vals || !q)))
    return -1;
  int64_t wcells = 1, nodes = 1, rcells = 1;
  for (int k = 0; k < g->d; ++k) {
    wcells *= g->tile[k] + 5;
    nodes *= g->n_grid[k];
    rcells *= g->n_grid[k] - 1;
  }
  const int nq = 1 << g->d;
  const int nw = nq * (nq + 1) / 2;
  const int64_t ntiles = flpr_ntiles(g);
  int64_t tcells = 1;
  for (int k = 0; k < g->d; ++k) tcells *= g->tile[k];
  // per problem: the LDS windows of k_adjoint or the corner sums of the walk
  const int64_t part_stride = ntiles * std::max(nval, nw) * std::max(wcells, tcells << g->d);
  hipStream_t st = as_stream(stream);
  // even node counts: both sums by the walk, as a single problem's
  bool aligned = true;
  for (int k = 0; k < g->d; ++k) aligned = aligned && !(g->n_grid[k] & 1);
  if (aligned)
    return launch_adjoint_walk(g, P, st, key_start, rawi, rawf, vals, nval, part, q, gram,
                               flpr_nkeys(g), (int64_t)nval * nodes, (int64_t)nw * rcells,
                               part_stride);
  for (int mode = 0; mode < 2; ++mode) {
    const int nch = (mode == 0) ? nval : (gram ? nw : 0);
    if (nch == 0) continue;
    double* out = (mode == 0) ? q : gram;
    const int64_t ostride = (mode == 0) ? nodes : rcells;
    const int cg = channel_group(ntiles, wcells, nch);
    const size_t lds = (size_t)wcells * cg * sizeof(double);
    if (lds > 64 * 1024) return -1;
    FlprBz bz{};
    bz.items = items;
    bz.s[0] = flpr_nkeys(g);
    bz.s[1] = (int64_t)nch * ostride;
    bz.s[2] = part_stride;
    dim3 grd((unsigned)ntiles, (unsigned)((nch + cg - 1) / cg), (unsigned)P);
    switch (g->d) {
      case 1: hipLaunchKernelGGL(k_adjoint<1>, grd, dim3(ADJ_THREADS), lds, st, *g, key_start, rawi, rawf, vals, nmax, nch, mode, cg, 1, out, ostride, part, bz); break;
      case 2: hipLaunchKernelGGL(k_adjoint<2>, grd, dim3(ADJ_THREADS), lds, st, *g, key_start, rawi, rawf, vals, nmax, nch, mode, cg, 1, out, ostride, part, bz); break;
      default: hipLaunchKernelGGL(k_adjoint<3>, grd, dim3(ADJ_THREADS), lds, st, *g, key_start, rawi, rawf, vals, nmax, nch, mode, cg, 1, out, ostride, part, bz); break;
    }
    FLPR_CHECK_LAUNCH();
    FlprBz bg{};
    bg.s[0] = part_stride;
    bg.s[1] = (int64_t)nch * ostride;
    dim3 gg(flpr_grid_for(ostride, 256, 256), (unsigned)nch, (unsigned)P);
    switch (g->d) {
      case 1: hipLaunchKernelGGL(k_adjoint_gather<1>, gg, dim3(256), 0, st, *g, part, nch, mode, ostride, out, bg); break;
      case 2: hipLaunchKernelGGL(k_adjoint_gather<2>, gg, dim3(256), 0, st, *g, part, nch, mode, ostride, out, bg); break;
      default: hipLaunchKernelGGL(k_adjoint_gather<3>, gg, dim3(256), 0, st, *g, part, nch, mode, ostride, out, bg); break;
    }
    FLPR_CHECK_LAUNCH();
  }
  return 0;
}

int flpr_batch_quadform(const FlprGeom* g, int32_t P, const double* m, int64_t m_stride,
                        int32_t ncol, int32_t dh, const double* gram, double* quad,
                        int64_t quad_stride, double* sum_part, int64_t sum_part_len,
                        void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || P < 1 || P > 65535 || !m || !gram || !quad || ncol < 1 || dh < 1)
    return -1;
  const int64_t nodes = flpr_nodes(g);
  int64_t rcells = 1;
  for (int k = 0; k < g->d; ++k) rcells *= g->n_grid[k] - 1;
  const int nq = 1 << g->d;
  FlprBz bz{};
  bz.s[0] = m_stride;
  bz.s[1] = (int64_t)(nq * (nq + 1) / 2) * rcells;
  bz.s[2] = quad_stride;
  const int nb = sum_part ? flpr_sum_blocks((int64_t)P * dh * ncol, 1,
                                            flpr_grid_for(rcells, QF_THREADS, QF_BLOCKS), sum_part_len)
                          : 0;
  if (nb < 1) return -3;
  dim3 grd((unsigned)nb, (unsigned)(dh * ncol), (unsigned)P);
  hipStream_t st = as_stream(stream);
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_quadform<1>, grd, dim3(QF_THREADS), 0, st, *g, m, nodes, gram, rcells, quad, sum_part, bz); break;
    case 2: hipLaunchKernelGGL(k_quadform<2>, grd, dim3(QF_THREADS), 0, st, *g, m, nodes, gram, rcells, quad, sum_part, bz); break;
    default: hipLaunchKernelGGL(k_quadform<3>, grd, dim3(QF_THREADS), 0, st, *g, m, nodes, gram, rcells, quad, sum_part, bz); break;
  }
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_batch_fit_output(const FlprGeom* g, const FlprBatchItem* items, int32_t P, int64_t nmax,
                          const double* x, const double* m_all, int64_t m_stride, int32_t ncol,
                          const int32_t* sel, const double* yref, int32_t variance, double* yhat,
                          double* resid, double* grid, double* part, double* fac,
                          int64_t fac_stride, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!batch_ok(g, items, P, nmax) || !x || !m_all || !yhat || ncol < 1 || (resid && !yref) ||
      (variance && (ncol != 1 || !yref || !grid || !part)))
    return -1;
  const unsigned gx = (unsigned)flpr_grid_for(nmax, EVAL_THREADS, FIT_BLOCKS);
  const int64_t nodes = flpr_nodes(g);
  hipStream_t st = as_stream(stream);
  const int var = variance ? 1 : 0;
  FlprBz bz{};
  bz.items = items;
  bz.s[0] = m_stride;
  bz.s[1] = gx;
  dim3 grd(gx, 1, (unsigned)P);
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_fit_points<1>, grd, dim3(EVAL_THREADS), 0, st, *g, x, nmax, m_all, ncol, nodes, sel, yref, var, yhat, resid, part, bz); break;
    case 2: hipLaunchKernelGGL(k_fit_points<2>, grd, dim3(EVAL_THREADS), 0, st, *g, x, nmax, m_all, ncol, nodes, sel, yref, var, yhat, resid, part, bz); break;
    default: hipLaunchKernelGGL(k_fit_points<3>, grd, dim3(EVAL_THREADS), 0, st, *g, x, nmax, m_all, ncol, nodes, sel, yref, var, yhat, resid, part, bz); break;
  }
  FLPR_CHECK_LAUNCH();
  if (variance) {
    FlprBz bf{};
    bf.items = items;
    bf.s[0] = m_stride;
    bf.s[1] = gx;
    bf.s[2] = nodes;
    bf.s[3] = fac_stride;
    hipLaunchKernelGGL(k_fit_finish, grd, dim3(EVAL_THREADS), 0, st, nmax, m_all, nodes, sel, yref,
                       part, (int)gx, yhat, resid, grid, fac, bf);
    FLPR_CHECK_LAUNCH();
  }
  return 0;
}

int flpr_batch_spline_coef(const FlprGeom* g, int32_t P, const double* inv, int64_t inv_stride,
                           const double* m, int32_t ngrid, double* work, double* coef,
                           void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || P < 1 || !inv || !m || !work || !coef || ngrid < 1 ||
      (int64_t)P * ngrid > 65535)
    return -1;
  const int d = g->d;
  int64_t nodes = 1, aoff[FLPR_MAX_DIM] = {0, 0, 0}, acc = 0;
  for (int k = 0; k < d; ++k) {
    nodes *= g->n_grid[k];
    aoff[k] = acc;
    acc += (int64_t)g->n_grid[k] * g->n_grid[k];
  }
  if (inv_stride < acc) return -1;
  hipStream_t st = as_stream(stream);
  // d passes that end in coef: m -> (work | coef) -> ... -> coef
  const double* src = m;
  for (int k = 0; k < d; ++k) {
    double* dst = ((d - 1 - k) % 2 == 0) ? coef : work;
    int64_t outer = 1, inner = 1;
    for (int j = 0; j < k; ++j) outer *= g->n_grid[j];
    for (int j = k + 1; j < d; ++j) inner *= g->n_grid[j];
    const int Nk = g->n_grid[k];
    dim3 grd((unsigned)((outer * inner + AG_T - 1) / AG_T), (unsigned)((Nk + AG_T - 1) / AG_T),
             (unsigned)(P * ngrid));
    hipLaunchKernelGGL(k_axis_gemm, grd, dim3(256), 0, st, inv + aoff[k], inv_stride, ngrid, src,
                       dst, nodes, Nk, outer, inner);
    FLPR_CHECK_LAUNCH();
    src = dst;
  }
  return 0;
}

int flpr_batch_point_sums(const FlprGeom* g, const FlprBatchItem* items, int32_t P, int64_t nmax,
                          const double* x, const double* y, int32_t ny, const double* probes,
                          int32_t nprobe, const double* m, const double* coef, int32_t dh,
                          int32_t gcv_cubic, int32_t dof_cubic, double* part, double* red,
                          int64_t red_stride, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!batch_ok(g, items, P, nmax) || !x || !y || ny < 1 || nprobe < 1 || !probes || !m ||
      ((gcv_cubic || dof_cubic) && !coef) || dh < 1 || !part || !red)
    return -1;
  for (int k = 0; k < g->d; ++k)
    if ((gcv_cubic || dof_cubic) && g->n_grid[k] < 4) return -1;
  const int nval = ny + nprobe;
  const int ngroups = (dh * nval + PS_GROUP - 1) / PS_GROUP;
  if (ngroups > 65535) return -1;
  const int nblk = flpr_grid_for(nmax, PS_THREADS, 64);
  const int64_t nodes = flpr_nodes(g);
  FlprBz bz{};
  bz.items = items;
  bz.s[0] = (int64_t)dh * nval * nodes;
  dim3 grd((unsigned)nblk, (unsigned)ngroups, (unsigned)P);
  hipStream_t st = as_stream(stream);
  const int gc = gcv_cubic ? 1 : 0, dc = dof_cubic ? 1 : 0;
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_point_sums<1>, grd, dim3(PS_THREADS), 0, st, *g, x, nmax, y, ny, probes, nprobe, m, coef, nodes, dh, gc, dc, part, bz); break;
    case 2: hipLaunchKernelGGL(k_point_sums<2>, grd, dim3(PS_THREADS), 0, st, *g, x, nmax, y, ny, probes, nprobe, m, coef, nodes, dh, gc, dc, part, bz); break;
    default: hipLaunchKernelGGL(k_point_sums<3>, grd, dim3(PS_THREADS), 0, st, *g, x, nmax, y, ny, probes, nprobe, m, coef, nodes, dh, gc, dc, part, bz); break;
  }
  FLPR_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_point_sums_final, dim3((unsigned)((dh * nval + 255) / 256), 1, (unsigned)P),
                     dim3(256), 0, st, part, nblk, ngroups, dh, ny, nprobe, red, red_stride);
  FLPR_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
