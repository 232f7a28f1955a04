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
__global__ __launch_bounds__(XS_THREADS) void k_xstats(const double* __restrict__ x, int64_t n,
                                                       int d, const double* __restrict__ y,
                                                       int ny, double* __restrict__ partials) {
  __shared__ Welford sh[XS_THREADS];
  __shared__ double red[XS_THREADS / 64];
  double bad_x = 0.0, bad_y = 0.0;
  for (int k = 0; k < d; ++k) {
    Welford w = {0.0, 0.0, 0.0, INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
      const double v = x[i * d + k];
      if (!isfinite(v)) {
        bad_x += 1.0;
        continue;
      }
      w.n += 1.0;
      const double delta = v - w.mean;
      w.mean += delta / w.n;
      w.m2 += delta * (v - w.mean);
      w.mn = fmin(w.mn, v);
      w.mx = fmax(w.mx, v);
    }
    sh[threadIdx.x] = w;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] = wf_merge(sh[threadIdx.x], sh[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      double* p = partials + ((size_t)k * gridDim.x + blockIdx.x) * 5;
      p[0] = sh[0].n; p[1] = sh[0].mean; p[2] = sh[0].m2; p[3] = sh[0].mn; p[4] = sh[0].mx;
    }
    __syncthreads();
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
__global__ __launch_bounds__(XS_THREADS) void k_xstats_final(const double* __restrict__ partials,
                                                             int nparts, int d, int ny,
                                                             double* __restrict__ out) {
  __shared__ Welford sh[XS_THREADS];
  __shared__ double red[XS_THREADS / 64];
  for (int k = 0; k < d; ++k) {
    Welford w = {0.0, 0.0, 0.0, INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
      const double* p = partials + ((size_t)k * nparts + b) * 5;
      const Welford o = {p[0], p[1], p[2], p[3], p[4]};
      w = wf_merge(w, o);
    }
    sh[threadIdx.x] = w;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] = wf_merge(sh[threadIdx.x], sh[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      out[5 * k + 0] = sh[0].mn;
      out[5 * k + 1] = sh[0].mx;
      out[5 * k + 2] = sh[0].mean;
      out[5 * k + 3] = sh[0].m2;
      out[5 * k + 4] = sh[0].n;
    }
    __syncthreads();
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
__global__ __launch_bounds__(PREP_THREADS) void k_prepare(const double* __restrict__ x, int64_t n,
                                                          FlprGeom g, int32_t* __restrict__ key,
                                                          int32_t* __restrict__ key_count) {
  int tc = 1;
  for (int k = 0; k < g.d; ++k) tc *= g.tile[k];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int tid = 0, lid = 0;
    for (int k = 0; k < g.d; ++k) {
      int m;
      knot_of(g, k, x[i * g.d + k], &m);
      const int pm = (m > g.L[k] / 2) ? m - g.L[k] : m;
      // the data support maps into [1 + msp, G - 2 - msp]; the clamp only keeps
      // a malformed geometry from addressing outside the key table
      const int s = min(max(pm + g.E[k], g.msp), g.G[k] - 1 - g.msp);
      tid = tid * g.ntile[k] + s / g.tile[k];
      lid = lid * g.tile[k] + s % g.tile[k];
    }
    const int kk = tid * tc + lid;
    key[i] = kk;
    atomicAdd(&key_count[kk], 1);
  }
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

// Scatter of point indices into key order (one atomic per point on its cell's
// cursor; about n / cells points share a cursor).
__global__ __launch_bounds__(256) void k_bin(const int32_t* __restrict__ key, int64_t n,
                                             const int32_t* __restrict__ start,
                                             int32_t* __restrict__ cursor,
                                             int32_t* __restrict__ order) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int k = key[i];
    order[start[k] + atomicAdd(&cursor[k], 1)] = (int32_t)i;
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

// Sorted structure-of-arrays copy of what the grid kernels need per point:
//   wab  [2d][n]  A, B of the tap-weight factorisation per dimension
//   vals [nv][n]  y columns then probe columns
//   rawi [d][n]   raw-grid interval, rawf [d][n] fraction in it
template <int D>
__global__ __launch_bounds__(256) void k_pack(FlprGeom g, const double* __restrict__ x,
                                              const int32_t* __restrict__ order, int64_t n,
                                              const double* __restrict__ vy, int ny,
                                              const double* __restrict__ pr, int np,
                                              double* __restrict__ wab, double* __restrict__ vals,
                                              int32_t* __restrict__ rawi,
                                              double* __restrict__ rawf) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = order[p];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double xv = x[i * D + k];
      int m;
      const double xm = knot_of(g, k, xv, &m);
      const double delta = xm - g.hx[k] * (double)m;
      wab[(2 * k) * n + p] = exp(-(delta * delta * g.inv4tau[k]));
      wab[(2 * k + 1) * n + p] = exp(2.0 * delta * g.hx[k] * g.inv4tau[k]);
      int ri;
      double rf;
      raw_locate(xv, g.graw0[k], g.graw1[k], g.n_grid[k], &ri, &rf);
      rawi[k * n + p] = ri;
      rawf[k * n + p] = rf;
    }
    for (int c = 0; c < ny; ++c) vals[c * n + p] = vy[i * ny + c];
    for (int c = 0; c < np; ++c) vals[(int64_t)(ny + c) * n + p] = pr[i * np + c];
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

// w[j] = exp(-(delta - hx (j - MSP))^2 c) from A, B and the constants G[j].
template <int MSP>
__device__ inline void tap_weights(double A, double B, const double* G, double* w) {
  w[MSP] = A * G[MSP];
  double bp = 1.0;
#pragma unroll
  for (int k = 1; k <= MSP; ++k) {
    bp *= B;
    w[MSP + k] = A * bp * G[MSP + k];
  }
  const double binv = 1.0 / B;
  bp = 1.0;
#pragma unroll
  for (int k = 1; k <= MSP; ++k) {
    bp *= binv;
    w[MSP - k] = A * bp * G[MSP - k];
  }
}

constexpr int SPREAD_THREADS = 256;

// Register-accumulated spread (d = 1, 2).  Item = (cell, channel[, row j0]);
// SPLIT items carry one footprint row (used when (2 MSP + 1)^2 float64
// accumulators would not fit in registers).
template <int D, int MSP, bool SPLIT>
__global__ __launch_bounds__(SPREAD_THREADS) void k_spread_cells(
    FlprGeom g, const int32_t* __restrict__ key_start, const double* __restrict__ wab,
    const double* __restrict__ vals, int64_t n, int with_ones, int C, int cg,
    double* __restrict__ part) {
  extern __shared__ double lds[];
  constexpr int ntap = 2 * MSP + 1;
  const int c0 = blockIdx.y * cg;
  const int cn = min(cg, C - c0);
  int t[3] = {0, 0, 0};
  tile_coords<D>(g, blockIdx.x, t);
  int fp[3] = {1, 1, 1}, org[3] = {0, 0, 0};
  int fcells = 1, tc = 1;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    fp[k] = g.tile[k] + 2 * MSP;
    org[k] = t[k] * g.tile[k] - MSP;
    fcells *= fp[k];
    tc *= g.tile[k];
  }
  for (int i = threadIdx.x; i < fcells * cn; i += blockDim.x) lds[i] = 0.0;
  double G[D][ntap];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const double c2 = g.hx[k] * g.hx[k] * g.inv4tau[k];
#pragma unroll
    for (int j = 0; j < ntap; ++j) G[k][j] = exp(-c2 * (double)((j - MSP) * (j - MSP)));
  }
  __syncthreads();
  const int nsplit = SPLIT ? ntap : 1;
  const int nitems = tc * cn * nsplit;
  const int kbase = blockIdx.x * tc;
  for (int item = threadIdx.x; item < nitems; item += blockDim.x) {
    const int lc = item % tc;
    const int rest = item / tc;
    const int c = rest % cn;
    const int js = rest / cn;  // footprint row of a SPLIT item
    const int p0 = key_start[kbase + lc], p1 = key_start[kbase + lc + 1];
    if (p0 == p1) continue;
    const int ch = c0 + c;
    const double* vch = (ch >= with_ones) ? vals + (int64_t)(ch - with_ones) * n : nullptr;
    int l[3] = {0, 0, 0};
    local_coords<D>(g, lc, l);
    if constexpr (D == 1) {
      double acc[ntap];
#pragma unroll
      for (int j = 0; j < ntap; ++j) acc[j] = 0.0;
      for (int p = p0; p < p1; ++p) {
        const double v = vch ? vch[p] : 1.0;
        double w[ntap];
        tap_weights<MSP>(wab[p], wab[n + p], G[0], w);
#pragma unroll
        for (int j = 0; j < ntap; ++j) acc[j] += v * w[j];
      }
#pragma unroll
      for (int j = 0; j < ntap; ++j) atomicAdd(&lds[(l[0] + j) * cn + c], acc[j]);
    } else if constexpr (!SPLIT) {
      double acc[ntap][ntap];
#pragma unroll
      for (int a = 0; a < ntap; ++a)
#pragma unroll
        for (int b = 0; b < ntap; ++b) acc[a][b] = 0.0;
      for (int p = p0; p < p1; ++p) {
        const double v = vch ? vch[p] : 1.0;
        double w0[ntap], w1[ntap];
        tap_weights<MSP>(wab[p], wab[n + p], G[0], w0);
        tap_weights<MSP>(wab[2 * n + p], wab[3 * n + p], G[1], w1);
#pragma unroll
        for (int a = 0; a < ntap; ++a) {
          const double vw = v * w0[a];
#pragma unroll
          for (int b = 0; b < ntap; ++b) acc[a][b] += vw * w1[b];
        }
      }
#pragma unroll
      for (int a = 0; a < ntap; ++a)
#pragma unroll
        for (int b = 0; b < ntap; ++b)
          atomicAdd(&lds[((l[0] + a) * fp[1] + (l[1] + b)) * cn + c], acc[a][b]);
    } else {
      double acc[ntap];
#pragma unroll
      for (int b = 0; b < ntap; ++b) acc[b] = 0.0;
      for (int p = p0; p < p1; ++p) {
        const double v = vch ? vch[p] : 1.0;
        double w0[ntap], w1[ntap];
        tap_weights<MSP>(wab[p], wab[n + p], G[0], w0);
        tap_weights<MSP>(wab[2 * n + p], wab[3 * n + p], G[1], w1);
        double sel = 0.0;
#pragma unroll
        for (int a = 0; a < ntap; ++a) sel = (a == js) ? w0[a] : sel;
        const double vw = v * sel;
#pragma unroll
        for (int b = 0; b < ntap; ++b) acc[b] += vw * w1[b];
      }
#pragma unroll
      for (int b = 0; b < ntap; ++b)
        atomicAdd(&lds[((l[0] + js) * fp[1] + (l[1] + b)) * cn + c], acc[b]);
    }
  }
  __syncthreads();
  // the tile's window goes to its own partial slot; k_spread_gather adds the
  // overlapping windows per cell without atomics
  double* dst = part + ((int64_t)blockIdx.x * C + c0) * fcells;
  for (int e = threadIdx.x; e < fcells * cn; e += blockDim.x) {
    const int c = e / fcells, cell = e % fcells;
    dst[(int64_t)c * fcells + cell] = lds[cell * cn + c];
  }
}

// 3-D spread: one thread per point scatters its (2 MSP + 1)^3 footprint into
// the tile's LDS window with float64 atomics.
template <int MSP>
__global__ __launch_bounds__(SPREAD_THREADS) void k_spread_points3(
    FlprGeom g, const int32_t* __restrict__ key_start, const double* __restrict__ wab,
    const double* __restrict__ vals, int64_t n, int with_ones, int C, int cg,
    double* __restrict__ part) {
  extern __shared__ double lds[];
  constexpr int D = 3;
  constexpr int ntap = 2 * MSP + 1;
  const int c0 = blockIdx.y * cg;
  const int cn = min(cg, C - c0);
  int t[3];
  tile_coords<D>(g, blockIdx.x, t);
  int fp[3], org[3];
  int fcells = 1, tc = 1;
  for (int k = 0; k < D; ++k) {
    fp[k] = g.tile[k] + 2 * MSP;
    org[k] = t[k] * g.tile[k] - MSP;
    fcells *= fp[k];
    tc *= g.tile[k];
  }
  for (int i = threadIdx.x; i < fcells * cn; i += blockDim.x) lds[i] = 0.0;
  double G[D][ntap];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const double c2 = g.hx[k] * g.hx[k] * g.inv4tau[k];
#pragma unroll
    for (int j = 0; j < ntap; ++j) G[k][j] = exp(-c2 * (double)((j - MSP) * (j - MSP)));
  }
  __syncthreads();
  const int kbase = blockIdx.x * tc;
  const int p_begin = key_start[kbase], p_end = key_start[kbase + tc];
  for (int p = p_begin + threadIdx.x; p < p_end; p += blockDim.x) {
    // cell of this point: binary search of the tile's key table
    int lo = 0, hi = tc - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (key_start[kbase + mid] <= p) lo = mid;
      else hi = mid - 1;
    }
    int l[3];
    local_coords<D>(g, lo, l);
    double w0[ntap], w1[ntap], w2[ntap];
    tap_weights<MSP>(wab[p], wab[n + p], G[0], w0);
    tap_weights<MSP>(wab[2 * n + p], wab[3 * n + p], G[1], w1);
    tap_weights<MSP>(wab[4 * n + p], wab[5 * n + p], G[2], w2);
    for (int c = 0; c < cn; ++c) {
      const int ch = c0 + c;
      const double v = (ch >= with_ones) ? vals[(int64_t)(ch - with_ones) * n + p] : 1.0;
#pragma unroll
      for (int a = 0; a < ntap; ++a) {
        const double va = v * w0[a];
#pragma unroll
        for (int b = 0; b < ntap; ++b) {
          const double vab = va * w1[b];
          const int base = (((l[0] + a) * fp[1] + (l[1] + b)) * fp[2] + l[2]) * cn + c;
#pragma unroll
          for (int e = 0; e < ntap; ++e) atomicAdd(&lds[base + e * cn], vab * w2[e]);
        }
      }
    }
  }
  __syncthreads();
  // the tile's window goes to its own partial slot; k_spread_gather adds the
  // overlapping windows per cell without atomics
  double* dst = part + ((int64_t)blockIdx.x * C + c0) * fcells;
  for (int e = threadIdx.x; e < fcells * cn; e += blockDim.x) {
    const int c = e / fcells, cell = e % fcells;
    dst[(int64_t)c * fcells + cell] = lds[cell * cn + c];
  }
}


__device__ inline int floordiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

// Sum of the overlapping tile windows for every spread-grid cell (channel =
// blockIdx.y); also the per-channel grid sums used by the centring step.
template <int D>
__global__ __launch_bounds__(256) void k_spread_gather(FlprGeom g, const double* __restrict__ part,
                                                       int C, int64_t cells,
                                                       double* __restrict__ grid,
                                                       double* __restrict__ sums) {
  __shared__ double red[4];
  const int c = blockIdx.y;
  const int msp = g.msp;
  int fp[3] = {1, 1, 1};
  int fcells = 1;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    fp[k] = g.tile[k] + 2 * msp;
    fcells *= fp[k];
  }
  double csum = 0.0;
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
        v += part[((int64_t)t0 * C + c) * fcells + (sc[0] - (t0 * g.tile[0] - msp))];
    } else if constexpr (D == 2) {
      for (int t0 = tlo[0]; t0 <= thi[0]; ++t0)
        for (int t1 = tlo[1]; t1 <= thi[1]; ++t1) {
          const int64_t tile = (int64_t)t0 * g.ntile[1] + t1;
          const int l0 = sc[0] - (t0 * g.tile[0] - msp), l1 = sc[1] - (t1 * g.tile[1] - msp);
          v += part[(tile * C + c) * fcells + l0 * fp[1] + l1];
        }
    } else {
      for (int t0 = tlo[0]; t0 <= thi[0]; ++t0)
        for (int t1 = tlo[1]; t1 <= thi[1]; ++t1)
          for (int t2 = tlo[2]; t2 <= thi[2]; ++t2) {
            const int64_t tile = ((int64_t)t0 * g.ntile[1] + t1) * g.ntile[2] + t2;
            const int l0 = sc[0] - (t0 * g.tile[0] - msp), l1 = sc[1] - (t1 * g.tile[1] - msp),
                      l2 = sc[2] - (t2 * g.tile[2] - msp);
            v += part[(tile * C + c) * fcells + (l0 * fp[1] + l1) * fp[2] + l2];
          }
    }
    grid[(int64_t)c * cells + lin] = v;
    csum += v;
  }
  csum = block_sum(csum, red);
  if (threadIdx.x == 0) atomicAdd(&sums[c], csum);
}

// Centre every channel c >= 1 on the grid: F_c -= shift_c F_0 with
// shift_c = sum F_c / sum F_0 (the kernel-weighted mean of the channel).  The
// per-node solve adds shift_c times the ones-channel moments back, which is
// exact by linearity; the float32 convolution then only carries y - shift.
__global__ __launch_bounds__(256) void k_grid_center(double* __restrict__ grid, int64_t cells,
                                                     const double* __restrict__ sums,
                                                     double* __restrict__ shift,
                                                     float* __restrict__ grid32) {
  const int c = blockIdx.y;
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

// Adjoint interpolation sums on the raw grid for one channel group, per
// spread tile:
//   mode 0: q[c][node]   += v_ic w_i(node)          (value channels)
//   mode 1: W[e][rcell]  += w_i(qa) w_i(qb)          (Gram entries)
// Item = (spread cell, channel); the points of a spread cell touch at most
// three raw intervals per dimension.  Accumulated in registers, then in the
// tile's LDS window, which is written to the tile's partial slot
// (k_adjoint_gather adds the overlapping windows).
template <int D>
__global__ __launch_bounds__(ADJ_THREADS) void k_adjoint(
    FlprGeom g, const int32_t* __restrict__ key_start, const int32_t* __restrict__ rawi,
    const double* __restrict__ rawf, const double* __restrict__ vals, int64_t n, int nch,
    int mode, int cg, double* __restrict__ out, int64_t out_stride, double* __restrict__ part) {
  extern __shared__ double lds[];
  constexpr int NA = (D == 1) ? 4 : (D == 2 ? 16 : 64);   // 4^D node accumulators
  constexpr int NC = (D == 1) ? 3 : (D == 2 ? 9 : 27);    // 3^D cell accumulators
  const int c0 = blockIdx.y * cg;
  const int cn = min(cg, nch - c0);
  int t[3] = {0, 0, 0};
  tile_coords<D>(g, blockIdx.x, t);
  int ext[3] = {1, 1, 1}, lb[3] = {0, 0, 0}, lim[3] = {1, 1, 1};
  int wcells = 1, tc = 1;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    ext[k] = g.tile[k] + 5;
    lb[k] = t[k] * g.tile[k] - g.E[k] + (g.n_grid[k] - 1) / 2 - 2;
    lim[k] = (mode == 0) ? g.n_grid[k] : g.n_grid[k] - 1;
    wcells *= ext[k];
    tc *= g.tile[k];
  }
  for (int i = threadIdx.x; i < wcells * cn; i += blockDim.x) lds[i] = 0.0;
  __syncthreads();
  const int kbase = blockIdx.x * tc;
  const int nitems = tc * cn;
  for (int item = threadIdx.x; item < nitems; item += blockDim.x) {
    const int lc = item % tc;
    const int c = item / tc;
    const int p0 = key_start[kbase + lc], p1 = key_start[kbase + lc + 1];
    if (p0 == p1) continue;
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
    for (int p = p0; p < p1; ++p) {
      int rel[3] = {0, 0, 0};
      double f[3] = {0, 0, 0};
      bool inside = true;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        rel[k] = rawi[k * n + p] - ib[k];
        f[k] = rawf[k * n + p];
        if (rel[k] < 0 || rel[k] > 2) inside = false;
      }
      if (mode == 0) {
        const double v = vals[(int64_t)ch * n + p];
        if (!inside) {
#pragma unroll
          for (int q = 0; q < (1 << D); ++q) {
            double w = v;
            int64_t node = 0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
              const int bit = (q >> (D - 1 - k)) & 1;
              w *= bit ? f[k] : 1.0 - f[k];
              node = node * g.n_grid[k] + rawi[k * n + p] + bit;
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
          for (int k = 0; k < D; ++k) rc = rc * (g.n_grid[k] - 1) + rawi[k * n + p];
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

// Adds the overlapping adjoint windows into out (nodes for mode 0, raw cells
// for mode 1), one thread per (entry, channel = blockIdx.y), no atomics.
template <int D>
__global__ __launch_bounds__(256) void k_adjoint_gather(FlprGeom g, const double* __restrict__ part,
                                                        int nch, int mode, int64_t count,
                                                        double* __restrict__ out) {
  const int c = blockIdx.y;
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

// quad[h][c] = sum_cell sum_{a,b} W_ab(cell) m(h,c,node_a) m(h,c,node_b)
template <int D>
__global__ __launch_bounds__(QF_THREADS) void k_quadform(FlprGeom g, const double* __restrict__ m,
                                                         int64_t nodes,
                                                         const double* __restrict__ gram,
                                                         int64_t rcells,
                                                         double* __restrict__ quad) {
  __shared__ double red[QF_THREADS / 64];
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
  if (threadIdx.x == 0) atomicAdd(&quad[hc], s);
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

int lds_group(int64_t cells_per_channel, int channels, int cap, int64_t budget = 64 * 1024) {
  const int64_t per = cells_per_channel * (int64_t)sizeof(double);
  int cg = (int)(budget / per);
  if (cg < 1) cg = 1;
  if (cg > cap) cg = cap;
  if (cg > channels) cg = channels;
  return cg;
}

template <int D, int MSP, bool SPLIT>
void launch_cells(dim3 grd, size_t lds, hipStream_t st, const FlprGeom& g, const int32_t* ks,
                  const double* wab, const double* vals, int64_t n, int wo, int C, int cg,
                  double* part) {
  hipLaunchKernelGGL((k_spread_cells<D, MSP, SPLIT>), grd, dim3(SPREAD_THREADS), lds, st, g, ks,
                     wab, vals, n, wo, C, cg, part);
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
  if (!x || !partials || !out || n < 1 || d < 1 || d > FLPR_MAX_DIM || ny < 0 || ny > MAXY ||
      (ny > 0 && !y))
    return -1;
  hipLaunchKernelGGL(k_xstats, dim3(XS_BLOCKS), dim3(XS_THREADS), 0, as_stream(stream), x, n, d,
                     y, ny, partials);
  FLPR_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_xstats_final, dim3(1), dim3(XS_THREADS), 0, as_stream(stream), partials,
                     XS_BLOCKS, d, ny, out);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_points_prepare(const double* x, int64_t n, const FlprGeom* g, int32_t* key,
                        int32_t* key_count, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!x || !key || !key_count || n < 1 || !flpr_geom_ok(g)) return -1;
  if (flpr_nkeys(g) > ((int64_t)1 << 30)) return -1;
  hipLaunchKernelGGL(k_prepare, dim3(flpr_grid_for(n, PREP_THREADS, 4096)), dim3(PREP_THREADS), 0,
                     as_stream(stream), x, n, *g, key, key_count);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_points_bin(const int32_t* key, int64_t n, int32_t nkeys, const int32_t* key_count,
                    int32_t* key_start, int32_t* key_cursor, int32_t* order, void* stream) {
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
                     key, n, key_start, key_cursor, order);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_points_pack(const FlprGeom* g, const double* x, const int32_t* order, int64_t n,
                     const double* vals_y, int32_t ny, const double* probes, int32_t nprobe,
                     double* wab, double* vals, int32_t* rawi, double* rawf, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !x || !order || !wab || !rawi || !rawf || n < 1 ||
      ny < 0 || nprobe < 0 || ((ny + nprobe) > 0 && !vals) || (ny > 0 && !vals_y) ||
      (nprobe > 0 && !probes))
    return -1;
  dim3 grd(flpr_grid_for(n, 256, 4096));
  hipStream_t st = as_stream(stream);
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_pack<1>, grd, dim3(256), 0, st, *g, x, order, n, vals_y, ny, probes, nprobe, wab, vals, rawi, rawf); break;
    case 2: hipLaunchKernelGGL(k_pack<2>, grd, dim3(256), 0, st, *g, x, order, n, vals_y, ny, probes, nprobe, wab, vals, rawi, rawf); break;
    default: hipLaunchKernelGGL(k_pack<3>, grd, dim3(256), 0, st, *g, x, order, n, vals_y, ny, probes, nprobe, wab, vals, rawi, rawf); break;
  }
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
  if (g->msp < 4 || g->msp > 6) return -1;
  const int wo = with_ones ? 1 : 0;
  const int C = wo + nval;
  if (C < 1 || C > FLPR_MAX_CHANNELS) return -1;
  int64_t fcells = 1, cells = 1;
  for (int k = 0; k < g->d; ++k) {
    fcells *= g->tile[k] + 2 * g->msp;
    cells *= g->G[k];
  }
  const int cg = lds_group(fcells, C, 16);
  const size_t lds = (size_t)fcells * cg * sizeof(double);
  if (lds > 64 * 1024) return -1;
  dim3 grd((unsigned)flpr_ntiles(g), (unsigned)((C + cg - 1) / cg));
  hipStream_t st = as_stream(stream);
  const int msp = g->msp;
  if (g->d == 1) {
    if (msp == 4) launch_cells<1, 4, false>(grd, lds, st, *g, key_start, wab, vals, n, wo, C, cg, part);
    else if (msp == 5) launch_cells<1, 5, false>(grd, lds, st, *g, key_start, wab, vals, n, wo, C, cg, part);
    else launch_cells<1, 6, false>(grd, lds, st, *g, key_start, wab, vals, n, wo, C, cg, part);
  } else if (g->d == 2) {
    if (msp == 4) launch_cells<2, 4, false>(grd, lds, st, *g, key_start, wab, vals, n, wo, C, cg, part);
    else if (msp == 5) launch_cells<2, 5, true>(grd, lds, st, *g, key_start, wab, vals, n, wo, C, cg, part);
    else launch_cells<2, 6, true>(grd, lds, st, *g, key_start, wab, vals, n, wo, C, cg, part);
  } else {
    if (msp == 4) hipLaunchKernelGGL(k_spread_points3<4>, grd, dim3(SPREAD_THREADS), lds, st, *g, key_start, wab, vals, n, wo, C, cg, part);
    else if (msp == 5) hipLaunchKernelGGL(k_spread_points3<5>, grd, dim3(SPREAD_THREADS), lds, st, *g, key_start, wab, vals, n, wo, C, cg, part);
    else hipLaunchKernelGGL(k_spread_points3<6>, grd, dim3(SPREAD_THREADS), lds, st, *g, key_start, wab, vals, n, wo, C, cg, part);
  }
  FLPR_CHECK_LAUNCH();
  dim3 gg(flpr_grid_for(cells, 256, 256), (unsigned)C);
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_spread_gather<1>, gg, dim3(256), 0, st, *g, part, C, cells, grid, sums); break;
    case 2: hipLaunchKernelGGL(k_spread_gather<2>, gg, dim3(256), 0, st, *g, part, C, cells, grid, sums); break;
    default: hipLaunchKernelGGL(k_spread_gather<3>, gg, dim3(256), 0, st, *g, part, C, cells, grid, sums); break;
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
                     grid32);
  FLPR_CHECK_LAUNCH();
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
  hipStream_t st = as_stream(stream);
  for (int mode = 0; mode < 2; ++mode) {
    const int nch = (mode == 0) ? nval : (gram ? nw : 0);
    if (nch == 0) continue;
    double* out = (mode == 0) ? q : gram;
    const int64_t ostride = (mode == 0) ? nodes : rcells;
    const int cg = lds_group(wcells, nch, 64);
    const size_t lds = (size_t)wcells * cg * sizeof(double);
    if (lds > 64 * 1024) return -1;
    dim3 grd((unsigned)flpr_ntiles(g), (unsigned)((nch + cg - 1) / cg));
    switch (g->d) {
      case 1: hipLaunchKernelGGL(k_adjoint<1>, grd, dim3(ADJ_THREADS), lds, st, *g, key_start, rawi, rawf, vals, n, nch, mode, cg, out, ostride, part); break;
      case 2: hipLaunchKernelGGL(k_adjoint<2>, grd, dim3(ADJ_THREADS), lds, st, *g, key_start, rawi, rawf, vals, n, nch, mode, cg, out, ostride, part); break;
      default: hipLaunchKernelGGL(k_adjoint<3>, grd, dim3(ADJ_THREADS), lds, st, *g, key_start, rawi, rawf, vals, n, nch, mode, cg, out, ostride, part); break;
    }
    FLPR_CHECK_LAUNCH();
    dim3 gg(flpr_grid_for(ostride, 256, 256), (unsigned)nch);
    switch (g->d) {
      case 1: hipLaunchKernelGGL(k_adjoint_gather<1>, gg, dim3(256), 0, st, *g, part, nch, mode, ostride, out); break;
      case 2: hipLaunchKernelGGL(k_adjoint_gather<2>, gg, dim3(256), 0, st, *g, part, nch, mode, ostride, out); break;
      default: hipLaunchKernelGGL(k_adjoint_gather<3>, gg, dim3(256), 0, st, *g, part, nch, mode, ostride, out); break;
    }
    FLPR_CHECK_LAUNCH();
  }
  return 0;
}

int flpr_quadform(const FlprGeom* g, const double* m, int32_t ncol, int32_t dh,
                  const double* gram, double* quad, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !m || !gram || !quad || ncol < 1 || dh < 1) return -1;
  const int64_t nodes = flpr_nodes(g);
  int64_t rcells = 1;
  for (int k = 0; k < g->d; ++k) rcells *= g->n_grid[k] - 1;
  dim3 grd(flpr_grid_for(rcells, QF_THREADS, 64), (unsigned)(dh * ncol));
  hipStream_t st = as_stream(stream);
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_quadform<1>, grd, dim3(QF_THREADS), 0, st, *g, m, nodes, gram, rcells, quad); break;
    case 2: hipLaunchKernelGGL(k_quadform<2>, grd, dim3(QF_THREADS), 0, st, *g, m, nodes, gram, rcells, quad); break;
    default: hipLaunchKernelGGL(k_quadform<3>, grd, dim3(QF_THREADS), 0, st, *g, m, nodes, gram, rcells, quad); break;
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

}  // extern "C"
