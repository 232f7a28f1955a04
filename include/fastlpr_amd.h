/*
 * fastlpr_amd.h -- C-ABI of the MI355X engine for the fastLPR cv_fastlpr /
 * cv_fastkde hot path.
 *
 * Every entry point takes device pointers (hipMalloc'd or torch-allocated
 * memory on the current HIP device), plain sizes, an explicit HIP stream
 * passed as void* (NULL = default stream), and returns an int status:
 *   0  ok
 *  -1  invalid argument (sizes, dimensions, null pointers)
 *  -2  HIP launch / runtime error (hipGetLastError after the launch)
 *  -3  workspace or table too small
 * No C++ exceptions cross this boundary and the library holds no mutable
 * global state: every buffer is caller-owned, calls on distinct streams are
 * independent.
 *
 * Citations are to the reference checkout (fastLPR_py/src/fastlpr/...).
 */
#ifndef FASTLPR_AMD_H
#define FASTLPR_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLPR_MAX_DIM 3
#define FLPR_MAX_TERMS 10   /* 3-D order 2 */
#define FLPR_MAX_MOMENTS 35 /* 3-D order 4 monomials */
#define FLPR_MAX_CHANNELS 64

/* Geometry of one fit.  Filled on the host from the x statistics; mirrors
 * the reference's knot mapping (convolution.py:169-199, nufftn_type1.py:140-176),
 * grid parameters (kernel.py:97-156) and padded size / offsets
 * (convolution.py:390-409). */
typedef struct FlprGeom {
  int32_t d;                      /* dimensions 1..3 */
  int32_t msp;                    /* spread half-width in cells (= accuracy) */
  int32_t n_grid[FLPR_MAX_DIM];   /* output grid nodes per dimension (N) */
  int32_t L[FLPR_MAX_DIM];        /* padded FFT length of the reference (power of 2) */
  int32_t E[FLPR_MAX_DIM];        /* unwrapped spread grid covers signed cells [-E, E] */
  int32_t G[FLPR_MAX_DIM];        /* 2E+1 */
  int32_t tile[FLPR_MAX_DIM];     /* spread tile edge in cells */
  int32_t ntile[FLPR_MAX_DIM];    /* tiles per dimension = ceil(G / tile) */
  double hx[FLPR_MAX_DIM];        /* 2 pi / L */
  double inv4tau[FLPR_MAX_DIM];   /* 1 / (4 tau) */
  double mean[FLPR_MAX_DIM];      /* z-score mean (ddof=1 std below) */
  double std_[FLPR_MAX_DIM];
  double centre[FLPR_MAX_DIM];    /* midpoint of the z-scored extremes */
  double xmin_s[FLPR_MAX_DIM];    /* min of the centred scaled data */
  double scale_s[FLPR_MAX_DIM];   /* max - min of the centred scaled data */
  double xmax_s[FLPR_MAX_DIM];    /* max of the centred scaled data */
  double ks[FLPR_MAX_DIM];        /* knot scale idx_range[qin1] - idx_range[qin0] */
  double idx0[FLPR_MAX_DIM];      /* idx_range[qin0] */
  double shift[FLPR_MAX_DIM];     /* -0.5 - kmin of scale_knots */
  double graw0[FLPR_MAX_DIM];     /* raw grid axis: linspace(graw0, graw1, N) */
  double graw1[FLPR_MAX_DIM];
} FlprGeom;

/* Last HIP error code seen by the calling thread inside the library (0 if
 * none); its message is copied into buf (len bytes, NUL-terminated). */
int flpr_last_hip_error(char* buf, int32_t len);

/* x statistics (model.py:58-68, core.py:83-99): per dimension
 * out[5*k+0..4] = min, max, mean, sum of squared deviations, count.
 * x is (n, d) row-major float64.  partials must hold 5*d*512 doubles. */
int flpr_xstats(const double* x, int64_t n, int32_t d, double* partials, double* out,
                void* stream);

/* Knot positions, spread cells and spread-tile histogram for every point
 * (convolution.py:169-199 + nufftn_type1.py:140-176, 520-543).
 *   xmod  (n, d) float64   knot position in [0, 2 pi)
 *   cell  (n, d) int32     round(xmod / hx) in [0, L]
 *   tile  (n)    int32     linear spread-tile id
 *   tile_count (ntiles) int32, zeroed by the caller, incremented here. */
int flpr_points_prepare(const double* x, int64_t n, const FlprGeom* g, double* xmod,
                        int32_t* cell, int32_t* tile, int32_t* tile_count, void* stream);

/* Counting sort of the points by spread tile.  tile_start (ntiles+1) gets the
 * exclusive scan of tile_count; tile_cursor (ntiles) is scratch zeroed by the
 * caller; order (n) receives point indices grouped by tile. */
int flpr_points_bin(const int32_t* tile, int64_t n, int32_t ntiles, const int32_t* tile_count,
                    int32_t* tile_start, int32_t* tile_cursor, int32_t* order, void* stream);

/* Gaussian-gridding spread of C channels onto the unwrapped spread grid
 * (nufftn_type1.py:565-636, spreading_fused.py:67-94, _nufft_accel.pyx:32-76).
 * Channel 0 is the constant 1 when with_ones != 0, then ny columns of vals_y
 * (n, ny) float64, then nprobe columns of probes (n, nprobe) float64.
 * grid (C, G0[, G1[, G2]]) float64 must be zeroed by the caller; values are
 * accumulated in float64. */
int flpr_spread(const FlprGeom* g, const double* xmod, const int32_t* cell,
                const int32_t* order, const int32_t* tile_start, int32_t with_ones,
                const double* vals_y, int32_t ny, const double* probes, int32_t nprobe,
                double* grid, void* stream);

/* Kernel-moment taps (convolution.py:412-512, kernel.py:330-357,
 * design_matrix_optimized.py:216-298): for bandwidth row b (h is (dh, d)),
 * dimension k and power a in [0, npow): tap[((b*d + k)*npow + a)*nmax + t] =
 * exp(-u^2/2) u^a with u = linspace(xmin_s, xmin_s+scale_s, N_k)[t] / h[b,k];
 * dimension 0 is further scaled by 1/(sqrt(2 pi) prod_k h[b,k] * norm).
 * Written in float64 (tap64) and float32 (tap32, may be NULL). */
int flpr_taps(const FlprGeom* g, const double* h, int32_t dh, int32_t npow, int32_t nmax,
              double norm, double* tap64, float* tap32, void* stream);

/* One separable convolution pass along one axis for a batch of jobs: the
 * reference's spectrum product + inverse FFT + qout extraction
 * (convolution.py:904-951) restated as a banded Toeplitz product on the
 * unwrapped grid, computed with MFMA.
 *   mode 0 (ROW, contiguous axis):  out[r][i] = sum_s in[r][s] * tap[i + base - s]
 *   mode 1 (COL, strided axis):     out[o][i][c] = sum_s tap[i + base - s] * in[o][s][c]
 * Job j (device array of nj records of 8 int64: in_off, out_off, tap_off,
 * band, rows_or_outer, 0, 0, 0) reads from in + in_off and writes
 * out + out_off; taps outside [bands[2*band], bands[2*band+1]] (device int32
 * array) are treated as zero.  K = input axis length, nout = output axis
 * length, inner = contiguous extent for mode 1, rows_or_outer = rows (mode 0)
 * or outer batch count (mode 1) of the job, max_rows_or_outer its maximum.
 * fp64 != 0 selects the float64 MFMA path (in/out/tap are double*). */
int flpr_conv(int32_t mode, int32_t fp64, const void* in, void* out, const void* tap,
              const int64_t* jobs, const int32_t* bands, int32_t nj, int32_t max_rows_or_outer,
              int32_t K, int32_t nout, int32_t inner, int32_t base, void* stream);

/* Largest |S_kk| over all nodes, bandwidths and diagonal moments
 * (regression.py:243-267) when absval != 0, or the largest S (order-0
 * threshold, regression.py:217) when absval == 0.  The moment slabs are
 * src + h*per_h + offs[k] for k < nslab, each of length nodes; offs is a
 * HOST array (copied into the kernel arguments, nslab <= FLPR_MAX_MOMENTS).
 * Result stored as an order-preserving uint64 key in *key (caller zeroes);
 * decode with flpr_key_to_double on the host. */
int flpr_max_moment(int32_t fp64, const void* src, int64_t per_h, int32_t dh,
                    const int64_t* offs, int32_t nslab, int64_t nodes, int32_t absval,
                    uint64_t* key, void* stream);
double flpr_key_to_double(uint64_t key);

/* Per-node solve m = e1' (S + lam I)^-1 T (order >= 1) or T / max(S, thr)
 * (order 0) for every node and bandwidth (regression.py:162-277,
 * lwp_formulas.py:133-380), in float64.
 *   mom       stage-0 moment slabs (fp64 ? double : float)
 *   per_h     elements per bandwidth
 *   s_off[ns] slab offset of S pair (i,j>=i) (row-major pairs), channel 0
 *   t_off[nt] slab offset of term i; channel c adds c*nodes
 *             (s_off and t_off are HOST arrays copied into the kernel arguments)
 *   key       max key from flpr_max_moment; lam_scale multiplies it
 *             (1e-6 for order >= 1, regularization for order 0)
 *   eps_add   added to S for order 0 (regression.py:127)
 *   chan_first, nchan   T channels solved (absolute channel ids)
 *   m_out     (dh, n_out_chan, nodes) float64 for channels
 *             [chan_first, chan_first + n_out_chan)
 *   q         (nprobe, nodes) float64 adjoint-interpolated probes or NULL;
 *             channels [probe_first, probe_first + nprobe) then accumulate
 *             dot[h*nprobe + k] += sum_node q[k][node] * m_k(node). */
int flpr_solve(int32_t fp64, const void* mom, int64_t per_h, int32_t dh, int64_t nodes,
               int32_t order, int32_t nt, const int64_t* s_off, const int64_t* t_off,
               const uint64_t* key, double lam_scale, double eps_add, int32_t n_out_chan,
               int32_t chan_first, double* m_out, int32_t probe_first, int32_t nprobe,
               const double* q, double* dot, void* stream);

/* Adjoint of multilinear interpolation on the raw grid for the probe
 * columns: q[k][node] = sum_i p[i][k] w_i(node) (regression.py:1256-1280
 * computes p' * interp(m); this is the same sum reordered).  q (nprobe, nodes)
 * float64 must be zeroed by the caller. */
int flpr_adjoint_interp(const FlprGeom* g, const double* x, int64_t n, const int32_t* order,
                        const int32_t* tile_start, const double* probes, int32_t nprobe,
                        double* q, void* stream);

/* Residual sums for GCV (api.py:686-723, regression.py:384-396): for every
 * bandwidth h and output column c, rss[h*ncol + c] += sum_i |y_ic - yhat_ic(h)|^2
 * with yhat the multilinear interpolant (RegularGridInterpolator 'linear') of
 * m (dh, ncol_m, nodes).  Complex columns are passed as (re, im) pairs with
 * cplx != 0.  rss must be zeroed by the caller. */
int flpr_interp_rss(const FlprGeom* g, const double* x, int64_t n, const int32_t* order,
                    const double* y, int32_t ncol_m, int32_t cplx, const double* m, int32_t dh,
                    double* rss, void* stream);

/* Multilinear interpolation of one grid (ncol channels, (ncol, nodes)) at n
 * points (api.py:760-763, interp_optimized.py / scipy linear with linear
 * extrapolation outside the grid).  out (n, ncol) float64.  When
 * yvar != NULL also accumulates *expsum += sum_i yvar[i] exp(-out[i,0])
 * (api.py:771-778). */
int flpr_interp_eval(const FlprGeom* g, const double* x, int64_t n, const double* m,
                     int32_t ncol, double* out, const double* yvar, double* expsum,
                     void* stream);

/* Float64 -> float32 copy (spread grid for the float32 convolution path). */
int flpr_f64_to_f32(const double* src, float* dst, int64_t count, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FASTLPR_AMD_H */
