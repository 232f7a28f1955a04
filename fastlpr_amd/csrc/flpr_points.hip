// Point-side kernels: x statistics, knot mapping + spread-tile binning,
// Gaussian-gridding spread, adjoint interpolation of the probes, residual sums
// and final interpolation.  All point arithmetic is float64.
//
// Reference dataflow restated here (fastLPR_py/src/fastlpr/...):
//   model.py:58-72            z-score (ddof=1) and centring
//   convolution.py:169-199    knot mapping into the padded frequency window
//   nufftn_type1.py:140-176   scale_knots (recentre, x 2 pi, mod 2 pi)
//   nufftn_type1.py:520-636   round(xmod/hx), (2 Msp + 1)^d taps, exp(-diff^2/(4 tau))
//   spreading_fused.py:67-94  fused per-point spread (2-D form)
//   regression.py:1256-1280   Hutchinson p' * interp(m)
//   api.py:686-723            RegularGridInterpolator 'linear' at the data points
#include "flpr_common.h"

namespace {

constexpr int XS_BLOCKS = 512;
constexpr int XS_THREADS = 256;

struct Welford {
  double n, mean, m2, mn, mx;
};

__device__ inline Welford wf_merge(Welford a, Welford b) {
  if (a.n == 0) return b;
  if (b.n == 0) return a;
  Welford r;
  r.n = a.n + b.n;
  double delta = b.mean - a.mean;
  r.mean = a.mean + delta * (b.n / r.n);
  r.m2 = a.m2 + b.m2 + delta * delta * (a.n * b.n / r.n);
  r.mn = fmin(a.mn, b.mn);
  r.mx = fmax(a.mx, b.mx);
  return r;
}

// Chan/Welford statistics per dimension; one partial per block per dimension.
__global__ __launch_bounds__(XS_THREADS) void k_xstats(const double* __restrict__ x, int64_t n,
                                                       int d, double* __restrict__ partials) {
  __shared__ Welford sh[XS_THREADS];
  for (int k = 0; k < d; ++k) {
    Welford w = {0.0, 0.0, 0.0, INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
      double v = x[i * d + k];
      w.n += 1.0;
      double delta = v - w.mean;
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
}

__global__ __launch_bounds__(XS_THREADS) void k_xstats_final(const double* __restrict__ partials,
                                                             int nparts, int d,
                                                             double* __restrict__ out) {
  __shared__ Welford sh[XS_THREADS];
  for (int k = 0; k < d; ++k) {
    Welford w = {0.0, 0.0, 0.0, INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
      const double* p = partials + ((size_t)k * nparts + b) * 5;
      Welford o = {p[0], p[1], p[2], p[3], p[4]};
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

constexpr int PREP_THREADS = 256;
constexpr int PREP_LDS_TILES = 8192;

// Knot mapping of every point and the spread-tile histogram.
// The operation order mirrors the reference line by line so that the knot
// positions are the same float64 values (no FMA contraction here).
__global__ __launch_bounds__(PREP_THREADS) void k_prepare(const double* __restrict__ x, int64_t n,
                                                          FlprGeom g, double* __restrict__ xmod,
                                                          int32_t* __restrict__ cell,
                                                          int32_t* __restrict__ tile,
                                                          int32_t* __restrict__ tile_count,
                                                          int ntiles) {
#pragma clang fp contract(off)
  __shared__ int32_t hist[PREP_LDS_TILES];
  const bool use_lds = ntiles <= PREP_LDS_TILES;
  if (use_lds) {
    for (int t = threadIdx.x; t < ntiles; t += blockDim.x) hist[t] = 0;
    __syncthreads();
  }
  const double two_pi = 2.0 * M_PI;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int tid = 0;
    for (int k = 0; k < g.d; ++k) {
      double v = x[i * g.d + k];
      double t = (v - g.mean[k]) / g.std_[k];         // model.py:65
      double xs = t - g.centre[k];                     // model.py:72
      double xn = (xs - g.xmin_s[k]) / g.scale_s[k];  // convolution.py:186-188
      xn = xn * g.ks[k];
      xn = xn + g.idx0[k];
      xn = xn + 0.5;                                   // convolution.py:198
      xn = np_mod_pos(xn, 1.0);                        // convolution.py:199
      double kn = xn - 0.5;                            // convolution.py:67
      kn = kn * 1.0;                                   // nufftn_type1.py:159 (Fs/N = 1)
      kn = kn + g.shift[k];                            // nufftn_type1.py:163-165
      kn = kn * two_pi;                                // nufftn_type1.py:169
      double xm = np_mod_pos(kn, two_pi);              // nufftn_type1.py:170
      xmod[i * g.d + k] = xm;
      int m = (int)rint(xm / g.hx[k]);                 // round half to even, as np.round
      cell[i * g.d + k] = m;
      int pm = (m > g.L[k] / 2) ? m - g.L[k] : m;
      // the data support maps into [1 + msp, G - 2 - msp]; the clamp only
      // keeps a malformed geometry from addressing outside the tile table
      int s = min(max(pm + g.E[k], g.msp), g.G[k] - 1 - g.msp);
      int tk = s / g.tile[k];
      tid = tid * g.ntile[k] + tk;
    }
    tile[i] = tid;
    if (use_lds) atomicAdd(&hist[tid], 1);
    else atomicAdd(&tile_count[tid], 1);
  }
  if (use_lds) {
    __syncthreads();
    for (int t = threadIdx.x; t < ntiles; t += blockDim.x)
      if (hist[t]) atomicAdd(&tile_count[t], hist[t]);
  }
}

constexpr int SCAN_THREADS = 1024;

__global__ __launch_bounds__(SCAN_THREADS) void k_scan(const int32_t* __restrict__ cnt, int nt,
                                                       int32_t* __restrict__ start) {
  __shared__ int32_t part[SCAN_THREADS];
  int per = (nt + SCAN_THREADS - 1) / SCAN_THREADS;
  int b = threadIdx.x * per;
  int e = min(b + per, nt);
  int s = 0;
  for (int i = b; i < e; ++i) s += cnt[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < SCAN_THREADS; off <<= 1) {
    int v = (int)threadIdx.x >= off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int run = part[threadIdx.x] - s;
  for (int i = b; i < e; ++i) {
    start[i] = run;
    run += cnt[i];
  }
  if (threadIdx.x == SCAN_THREADS - 1) start[nt] = part[SCAN_THREADS - 1];
}

constexpr int BIN_THREADS = 256;
constexpr int BIN_PER_THREAD = 8;

// Scatter of point indices into tile order.  Each block ranks its points per
// tile in LDS and reserves one range per (block, tile) with a single global
// atomic, so global contention is per block, not per point.
__global__ __launch_bounds__(BIN_THREADS) void k_bin(const int32_t* __restrict__ tile, int64_t n,
                                                     int ntiles,
                                                     const int32_t* __restrict__ start,
                                                     int32_t* __restrict__ cursor,
                                                     int32_t* __restrict__ order) {
  __shared__ int32_t cnt[PREP_LDS_TILES];
  __shared__ int32_t base[PREP_LDS_TILES];
  const bool use_lds = ntiles <= PREP_LDS_TILES;
  int64_t b0 = (int64_t)blockIdx.x * BIN_THREADS * BIN_PER_THREAD;
  if (!use_lds) {
    for (int j = 0; j < BIN_PER_THREAD; ++j) {
      int64_t i = b0 + (int64_t)j * BIN_THREADS + threadIdx.x;
      if (i < n) {
        int t = tile[i];
        order[start[t] + atomicAdd(&cursor[t], 1)] = (int32_t)i;
      }
    }
    return;
  }
  for (int t = threadIdx.x; t < ntiles; t += blockDim.x) cnt[t] = 0;
  __syncthreads();
  int rank[BIN_PER_THREAD];
  int tl[BIN_PER_THREAD];
  for (int j = 0; j < BIN_PER_THREAD; ++j) {
    int64_t i = b0 + (int64_t)j * BIN_THREADS + threadIdx.x;
    tl[j] = -1;
    if (i < n) {
      tl[j] = tile[i];
      rank[j] = atomicAdd(&cnt[tl[j]], 1);
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < ntiles; t += blockDim.x)
    base[t] = cnt[t] ? start[t] + atomicAdd(&cursor[t], cnt[t]) : 0;
  __syncthreads();
  for (int j = 0; j < BIN_PER_THREAD; ++j) {
    int64_t i = b0 + (int64_t)j * BIN_THREADS + threadIdx.x;
    if (tl[j] >= 0) order[base[tl[j]] + rank[j]] = (int32_t)i;
  }
}

constexpr int SPREAD_THREADS = 256;
constexpr int MAX_CG = 16;  // channels per LDS group

// statically unrolled channel loop: keeps the per-point channel values in registers
#define CH_LOOP(stmt)                                    \
  _Pragma("unroll") for (int c = 0; c < MAX_CG; ++c) {   \
    if (c < cn) {                                        \
      stmt;                                              \
    }                                                    \
  }

template <int D>
__device__ inline void tile_coords(const FlprGeom& g, int tid, int* t) {
#pragma unroll
  for (int k = D - 1; k >= 0; --k) {
    t[k] = tid % g.ntile[k];
    tid /= g.ntile[k];
  }
}

// Gaussian-gridding spread of one channel group for one spread tile.  The
// tile's footprint (tile + 2 msp per dimension) is accumulated in LDS in
// float64 and flushed to the global grid with float64 atomics (only halo cells
// are shared with neighbouring tiles).
template <int D, int MSP>
__global__ __launch_bounds__(SPREAD_THREADS) void k_spread(
    FlprGeom g, const double* __restrict__ xmod, const int32_t* __restrict__ cell,
    const int32_t* __restrict__ order, const int32_t* __restrict__ tile_start, int with_ones,
    const double* __restrict__ vy, int ny, const double* __restrict__ pr, int np, int cg,
    double* __restrict__ grid, int64_t cstride) {
  extern __shared__ double acc[];
  constexpr int ntap = 2 * MSP + 1;
  const int C = with_ones + ny + np;
  const int c0 = blockIdx.y * cg;
  const int cn = min(cg, C - c0);
  int t[3] = {0, 0, 0};
  tile_coords<D>(g, blockIdx.x, t);
  int fp[3] = {1, 1, 1};
  int org[3] = {0, 0, 0};  // storage index of the footprint origin
  int fcells = 1;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    fp[k] = g.tile[k] + 2 * MSP;
    org[k] = t[k] * g.tile[k] - MSP;
    fcells *= fp[k];
  }
  for (int i = threadIdx.x; i < fcells * cn; i += blockDim.x) acc[i] = 0.0;
  __syncthreads();
  const int p0 = tile_start[blockIdx.x], p1 = tile_start[blockIdx.x + 1];
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int i = order[p];
    double w[D][ntap];
    int lc[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double xm = xmod[(int64_t)i * D + k];
      const int m = cell[(int64_t)i * D + k];
#pragma unroll
      for (int j = 0; j < ntap; ++j) {
        const double diff = xm - g.hx[k] * (double)(m + j - MSP);
        w[k][j] = exp(-(diff * diff * g.inv4tau[k]));
      }
      const int pm = (m > g.L[k] / 2) ? m - g.L[k] : m;
      const int sk = min(max(pm + g.E[k], MSP), g.G[k] - 1 - MSP);
      lc[k] = sk - t[k] * g.tile[k];  // in [0, tile); footprint index lc + j
    }
    double v[MAX_CG];
#pragma unroll
    for (int c = 0; c < MAX_CG; ++c) {
      const int ch = c0 + c;
      v[c] = 0.0;
      if (c < cn) {
        if (ch < with_ones) v[c] = 1.0;
        else if (ch < with_ones + ny) v[c] = vy[(int64_t)i * ny + (ch - with_ones)];
        else v[c] = pr[(int64_t)i * np + (ch - with_ones - ny)];
      }
    }
    if constexpr (D == 1) {
#pragma unroll
      for (int j0 = 0; j0 < ntap; ++j0) {
        const int idx = (lc[0] + j0) * cn;
        CH_LOOP(atomicAdd(&acc[idx + c], w[0][j0] * v[c]));
      }
    } else if constexpr (D == 2) {
#pragma unroll
      for (int j0 = 0; j0 < ntap; ++j0) {
#pragma unroll
        for (int j1 = 0; j1 < ntap; ++j1) {
          const double wv = w[0][j0] * w[1][j1];
          const int idx = ((lc[0] + j0) * fp[1] + (lc[1] + j1)) * cn;
          CH_LOOP(atomicAdd(&acc[idx + c], wv * v[c]));
        }
      }
    } else {
      for (int j0 = 0; j0 < ntap; ++j0) {
#pragma unroll
        for (int j1 = 0; j1 < ntap; ++j1) {
          const double w01 = w[0][j0] * w[1][j1];
#pragma unroll
          for (int j2 = 0; j2 < ntap; ++j2) {
            const double wv = w01 * w[2][j2];
            const int idx = (((lc[0] + j0) * fp[1] + (lc[1] + j1)) * fp[2] + (lc[2] + j2)) * cn;
            CH_LOOP(atomicAdd(&acc[idx + c], wv * v[c]));
          }
        }
      }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < fcells; e += blockDim.x) {
    int rem = e;
    int64_t lin = 0;
    bool ok = true;
    int sc[3] = {0, 0, 0};
#pragma unroll
    for (int k = D - 1; k >= 0; --k) {
      sc[k] = org[k] + rem % fp[k];
      rem /= fp[k];
      if (sc[k] < 0 || sc[k] >= g.G[k]) ok = false;
    }
    if (!ok) continue;
#pragma unroll
    for (int k = 0; k < D; ++k) lin = lin * g.G[k] + sc[k];
    for (int c = 0; c < cn; ++c) {
      const double a = acc[e * cn + c];
      if (a != 0.0) atomicAdd(&grid[(int64_t)(c0 + c) * cstride + lin], a);
    }
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

template <int D>
struct Corners {
  int64_t node[1 << D];
  double w[1 << D];
};

template <int D>
__device__ inline Corners<D> corners(const FlprGeom& g, const double* __restrict__ xp) {
  Corners<D> c;
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

constexpr int ADJ_THREADS = 256;
constexpr int ADJ_MARGIN = 5;

// Adjoint of the multilinear interpolation for the probe columns, one block
// per spread tile and channel group.  Raw nodes near the tile are accumulated
// in LDS; a point whose nodes fall outside the LDS window goes to global atomics.
template <int D>
__global__ __launch_bounds__(ADJ_THREADS) void k_adjoint(FlprGeom g, const double* __restrict__ x,
                                                         const int32_t* __restrict__ order,
                                                         const int32_t* __restrict__ tile_start,
                                                         const double* __restrict__ pr, int np,
                                                         int cg, double* __restrict__ q,
                                                         int64_t nodes) {
  extern __shared__ double acc[];
  const int c0 = blockIdx.y * cg;
  const int cn = min(cg, np - c0);
  int t[3] = {0, 0, 0};
  tile_coords<D>(g, blockIdx.x, t);
  int ext[3] = {1, 1, 1}, lb[3] = {0, 0, 0};
  int cells = 1;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    ext[k] = g.tile[k] + ADJ_MARGIN;
    lb[k] = t[k] * g.tile[k] - g.E[k] + (g.n_grid[k] - 1) / 2 - 2;
    cells *= ext[k];
  }
  for (int i = threadIdx.x; i < cells * cn; i += blockDim.x) acc[i] = 0.0;
  __syncthreads();
  const int p0 = tile_start[blockIdx.x], p1 = tile_start[blockIdx.x + 1];
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int i = order[p];
    const Corners<D> cr = corners<D>(g, x + (int64_t)i * D);
    // the lowest corner decides whether the whole cell lies in the window
    bool inside = true;
    int base_l[3] = {0, 0, 0};
    {
      int64_t rem = cr.node[0];
#pragma unroll
      for (int k = D - 1; k >= 0; --k) {
        const int id = (int)(rem % g.n_grid[k]);
        rem /= g.n_grid[k];
        base_l[k] = id - lb[k];
        if (base_l[k] < 0 || base_l[k] + 1 >= ext[k]) inside = false;
      }
    }
#pragma unroll
    for (int qq = 0; qq < (1 << D); ++qq) {
      int llin = 0;
#pragma unroll
      for (int k = 0; k < D; ++k) llin = llin * ext[k] + base_l[k] + ((qq >> (D - 1 - k)) & 1);
      for (int c = 0; c < cn; ++c) {
        const double v = cr.w[qq] * pr[(int64_t)i * np + c0 + c];
        if (inside) atomicAdd(&acc[llin * cn + c], v);
        else atomicAdd(&q[(int64_t)(c0 + c) * nodes + cr.node[qq]], v);
      }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < cells; e += blockDim.x) {
    int rem = e;
    int64_t lin = 0;
    bool ok = true;
    int sc[3] = {0, 0, 0};
#pragma unroll
    for (int k = D - 1; k >= 0; --k) {
      sc[k] = lb[k] + rem % ext[k];
      rem /= ext[k];
      if (sc[k] < 0 || sc[k] >= g.n_grid[k]) ok = false;
    }
    if (!ok) continue;
#pragma unroll
    for (int k = 0; k < D; ++k) lin = lin * g.n_grid[k] + sc[k];
    for (int c = 0; c < cn; ++c) {
      const double a = acc[e * cn + c];
      if (a != 0.0) atomicAdd(&q[(int64_t)(c0 + c) * nodes + lin], a);
    }
  }
}

constexpr int RSS_THREADS = 256;
constexpr int RSS_HC = 8;  // bandwidths per block row

__device__ inline double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += sh[w];
  __syncthreads();
  return r;
}

// Residual sums per (bandwidth, output column).  grid.y = bandwidth chunk,
// grid.z = output column.  Complex columns are (re, im) pairs of m and y.
template <int D>
__global__ __launch_bounds__(RSS_THREADS) void k_rss(FlprGeom g, const double* __restrict__ x,
                                                     int64_t n, const double* __restrict__ y,
                                                     int ncol_m, int cplx,
                                                     const double* __restrict__ m, int dh,
                                                     int64_t nodes, double* __restrict__ rss) {
  __shared__ double sh[RSS_THREADS / 64];
  const int h0 = blockIdx.y * RSS_HC;
  const int nh = min(RSS_HC, dh - h0);
  const int col = blockIdx.z;  // output column
  const int ncol_out = cplx ? ncol_m / 2 : ncol_m;
  double acc[RSS_HC];
#pragma unroll
  for (int j = 0; j < RSS_HC; ++j) acc[j] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const Corners<D> c = corners<D>(g, x + i * D);
    if (cplx) {
      const double yr = y[i * ncol_m + 2 * col], yi = y[i * ncol_m + 2 * col + 1];
#pragma unroll
      for (int j = 0; j < RSS_HC; ++j) {
        if (j < nh) {
          const double* mr = m + ((int64_t)(h0 + j) * ncol_m + 2 * col) * nodes;
          const double* mi = mr + nodes;
          double vr = 0.0, vi = 0.0;
#pragma unroll
          for (int q = 0; q < (1 << D); ++q) {
            vr += c.w[q] * mr[c.node[q]];
            vi += c.w[q] * mi[c.node[q]];
          }
          acc[j] += (yr - vr) * (yr - vr) + (yi - vi) * (yi - vi);
        }
      }
    } else {
      const double yv = y[i * ncol_m + col];
#pragma unroll
      for (int j = 0; j < RSS_HC; ++j) {
        if (j < nh) {
          const double* mm = m + ((int64_t)(h0 + j) * ncol_m + col) * nodes;
          double v = 0.0;
#pragma unroll
          for (int q = 0; q < (1 << D); ++q) v += c.w[q] * mm[c.node[q]];
          acc[j] += (yv - v) * (yv - v);
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < RSS_HC; ++j) {
    const double s = block_sum(acc[j], sh);
    if (threadIdx.x == 0 && j < nh) atomicAdd(&rss[(h0 + j) * ncol_out + col], s);
  }
}

template <int D>
__global__ __launch_bounds__(RSS_THREADS) void k_interp_eval(FlprGeom g,
                                                             const double* __restrict__ x,
                                                             int64_t n,
                                                             const double* __restrict__ m,
                                                             int ncol, int64_t nodes,
                                                             double* __restrict__ out,
                                                             const double* __restrict__ yvar,
                                                             double* __restrict__ expsum) {
  __shared__ double sh[RSS_THREADS / 64];
  double es = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const Corners<D> c = corners<D>(g, x + i * D);
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

__global__ void k_f64_to_f32(const double* __restrict__ s, float* __restrict__ d, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (float)s[i];
}

int lds_group(int64_t cells_per_channel, int channels, int64_t budget = 64 * 1024) {
  const int64_t per = cells_per_channel * (int64_t)sizeof(double);
  int cg = (int)(budget / per);
  if (cg < 1) cg = 1;
  if (cg > MAX_CG) cg = MAX_CG;
  if (cg > channels) cg = channels;
  return cg;
}

template <int D, int MSP>
void launch_spread(dim3 grd, size_t lds, hipStream_t st, const FlprGeom& g, const double* xmod,
                   const int32_t* cell, const int32_t* order, const int32_t* ts, int wo,
                   const double* vy, int ny, const double* pr, int np, int cg, double* grid,
                   int64_t cstride) {
  hipLaunchKernelGGL((k_spread<D, MSP>), grd, dim3(SPREAD_THREADS), lds, st, g, xmod, cell, order,
                     ts, wo, vy, ny, pr, np, cg, grid, cstride);
}

template <int D>
void launch_spread_d(int msp, dim3 grd, size_t lds, hipStream_t st, const FlprGeom& g,
                     const double* xmod, const int32_t* cell, const int32_t* order,
                     const int32_t* ts, int wo, const double* vy, int ny, const double* pr, int np,
                     int cg, double* grid, int64_t cstride) {
  switch (msp) {
    case 4: launch_spread<D, 4>(grd, lds, st, g, xmod, cell, order, ts, wo, vy, ny, pr, np, cg, grid, cstride); break;
    case 5: launch_spread<D, 5>(grd, lds, st, g, xmod, cell, order, ts, wo, vy, ny, pr, np, cg, grid, cstride); break;
    default: launch_spread<D, 6>(grd, lds, st, g, xmod, cell, order, ts, wo, vy, ny, pr, np, cg, grid, cstride); break;
  }
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

int flpr_xstats(const double* x, int64_t n, int32_t d, double* partials, double* out,
                void* stream) {
  FLPR_CLEAR_ERROR();
  if (!x || !partials || !out || n < 1 || d < 1 || d > FLPR_MAX_DIM) return -1;
  hipLaunchKernelGGL(k_xstats, dim3(XS_BLOCKS), dim3(XS_THREADS), 0, as_stream(stream), x, n, d,
                     partials);
  FLPR_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_xstats_final, dim3(1), dim3(XS_THREADS), 0, as_stream(stream), partials,
                     XS_BLOCKS, d, out);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_points_prepare(const double* x, int64_t n, const FlprGeom* g, double* xmod,
                        int32_t* cell, int32_t* tile, int32_t* tile_count, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!x || !xmod || !cell || !tile || !tile_count || n < 1 || !flpr_geom_ok(g)) return -1;
  const int64_t nt = flpr_ntiles(g);
  if (nt > ((int64_t)1 << 30)) return -1;
  hipLaunchKernelGGL(k_prepare, dim3(flpr_grid_for(n, PREP_THREADS, 2048)), dim3(PREP_THREADS), 0,
                     as_stream(stream), x, n, *g, xmod, cell, tile, tile_count, (int)nt);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_points_bin(const int32_t* tile, int64_t n, int32_t ntiles, const int32_t* tile_count,
                    int32_t* tile_start, int32_t* tile_cursor, int32_t* order, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!tile || !tile_count || !tile_start || !tile_cursor || !order || n < 1 || ntiles < 1 ||
      n > INT32_MAX)
    return -1;
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(SCAN_THREADS), 0, as_stream(stream), tile_count,
                     ntiles, tile_start);
  FLPR_CHECK_LAUNCH();
  const int64_t per_block = (int64_t)BIN_THREADS * BIN_PER_THREAD;
  const int blocks = (int)((n + per_block - 1) / per_block);
  hipLaunchKernelGGL(k_bin, dim3(blocks), dim3(BIN_THREADS), 0, as_stream(stream), tile, n,
                     ntiles, tile_start, tile_cursor, order);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_spread(const FlprGeom* g, const double* xmod, const int32_t* cell, const int32_t* order,
                const int32_t* tile_start, int32_t with_ones, const double* vals_y, int32_t ny,
                const double* probes, int32_t nprobe, double* grid, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !xmod || !cell || !order || !tile_start || !grid) return -1;
  if (g->msp < 4) return -1;
  if ((ny > 0 && !vals_y) || (nprobe > 0 && !probes) || ny < 0 || nprobe < 0) return -1;
  const int wo = with_ones ? 1 : 0;
  const int C = wo + ny + nprobe;
  if (C < 1 || C > FLPR_MAX_CHANNELS) return -1;
  int64_t fcells = 1, cstride = 1;
  for (int k = 0; k < g->d; ++k) {
    fcells *= g->tile[k] + 2 * g->msp;
    cstride *= g->G[k];
  }
  const int cg = lds_group(fcells, C);
  const size_t lds = (size_t)fcells * cg * sizeof(double);
  if (lds > 64 * 1024) return -1;
  const int ngroups = (C + cg - 1) / cg;
  dim3 grd((unsigned)flpr_ntiles(g), (unsigned)ngroups);
  hipStream_t st = as_stream(stream);
  switch (g->d) {
    case 1: launch_spread_d<1>(g->msp, grd, lds, st, *g, xmod, cell, order, tile_start, wo, vals_y, ny, probes, nprobe, cg, grid, cstride); break;
    case 2: launch_spread_d<2>(g->msp, grd, lds, st, *g, xmod, cell, order, tile_start, wo, vals_y, ny, probes, nprobe, cg, grid, cstride); break;
    default: launch_spread_d<3>(g->msp, grd, lds, st, *g, xmod, cell, order, tile_start, wo, vals_y, ny, probes, nprobe, cg, grid, cstride); break;
  }
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_adjoint_interp(const FlprGeom* g, const double* x, int64_t n, const int32_t* order,
                        const int32_t* tile_start, const double* probes, int32_t nprobe,
                        double* q, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !x || !order || !tile_start || !probes || !q || nprobe < 1 || n < 1)
    return -1;
  int64_t cells = 1;
  for (int k = 0; k < g->d; ++k) cells *= g->tile[k] + ADJ_MARGIN;
  const int cg = lds_group(cells, nprobe);
  const size_t lds = (size_t)cells * cg * sizeof(double);
  if (lds > 64 * 1024) return -1;
  const int ngroups = (nprobe + cg - 1) / cg;
  dim3 grd((unsigned)flpr_ntiles(g), (unsigned)ngroups);
  const int64_t nodes = flpr_nodes(g);
  hipStream_t st = as_stream(stream);
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_adjoint<1>, grd, dim3(ADJ_THREADS), lds, st, *g, x, order, tile_start, probes, nprobe, cg, q, nodes); break;
    case 2: hipLaunchKernelGGL(k_adjoint<2>, grd, dim3(ADJ_THREADS), lds, st, *g, x, order, tile_start, probes, nprobe, cg, q, nodes); break;
    default: hipLaunchKernelGGL(k_adjoint<3>, grd, dim3(ADJ_THREADS), lds, st, *g, x, order, tile_start, probes, nprobe, cg, q, nodes); break;
  }
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_interp_rss(const FlprGeom* g, const double* x, int64_t n, const int32_t* order,
                    const double* y, int32_t ncol_m, int32_t cplx, const double* m, int32_t dh,
                    double* rss, void* stream) {
  FLPR_CLEAR_ERROR();
  (void)order;
  if (!flpr_geom_ok(g) || !x || !y || !m || !rss || n < 1 || ncol_m < 1 || dh < 1) return -1;
  if (cplx && (ncol_m % 2)) return -1;
  const int ncol_out = cplx ? ncol_m / 2 : ncol_m;
  dim3 grd((unsigned)flpr_grid_for(n, RSS_THREADS, 1024), (unsigned)((dh + RSS_HC - 1) / RSS_HC),
           (unsigned)ncol_out);
  const int64_t nodes = flpr_nodes(g);
  hipStream_t st = as_stream(stream);
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_rss<1>, grd, dim3(RSS_THREADS), 0, st, *g, x, n, y, ncol_m, cplx, m, dh, nodes, rss); break;
    case 2: hipLaunchKernelGGL(k_rss<2>, grd, dim3(RSS_THREADS), 0, st, *g, x, n, y, ncol_m, cplx, m, dh, nodes, rss); break;
    default: hipLaunchKernelGGL(k_rss<3>, grd, dim3(RSS_THREADS), 0, st, *g, x, n, y, ncol_m, cplx, m, dh, nodes, rss); break;
  }
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_interp_eval(const FlprGeom* g, const double* x, int64_t n, const double* m, int32_t ncol,
                     double* out, const double* yvar, double* expsum, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !x || !m || !out || n < 1 || ncol < 1) return -1;
  if (yvar && !expsum) return -1;
  dim3 grd((unsigned)flpr_grid_for(n, RSS_THREADS, 2048));
  const int64_t nodes = flpr_nodes(g);
  hipStream_t st = as_stream(stream);
  switch (g->d) {
    case 1: hipLaunchKernelGGL(k_interp_eval<1>, grd, dim3(RSS_THREADS), 0, st, *g, x, n, m, ncol, nodes, out, yvar, expsum); break;
    case 2: hipLaunchKernelGGL(k_interp_eval<2>, grd, dim3(RSS_THREADS), 0, st, *g, x, n, m, ncol, nodes, out, yvar, expsum); break;
    default: hipLaunchKernelGGL(k_interp_eval<3>, grd, dim3(RSS_THREADS), 0, st, *g, x, n, m, ncol, nodes, out, yvar, expsum); break;
  }
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_f64_to_f32(const double* src, float* dst, int64_t count, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!src || !dst || count < 0) return -1;
  if (count == 0) return 0;
  hipLaunchKernelGGL(k_f64_to_f32, dim3(flpr_grid_for(count, 256, 4096)), dim3(256), 0,
                     as_stream(stream), src, dst, count);
  FLPR_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
