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
#define FLPR_MAX_MSP 12     /* spread half-width (accuracy 12; 3-D: 6) */

/* Geometry of one fit.  Filled on the host from the x statistics; mirrors
 * the reference's knot mapping (convolution.py:169-199, nufftn_type1.py:140-176),
 * grid parameters (kernel.py:97-156) and padded size / offsets
 * (convolution.py:390-409). */
typedef struct FlprGeom {
  int32_t d;                      /* dimensions 1..3 */
  int32_t msp;                    /* spread half-width in cells (= accuracy) */
  int32_t n_grid[FLPR_MAX_DIM];   /* output grid nodes per dimension (N) */
  int32_t L[FLPR_MAX_DIM];        /* periodic length of the spread grid: the reference's
                                     padded FFT length (power of 2), or R times it for
                                     the oversampled grid of accuracy > 6 */
  int32_t E[FLPR_MAX_DIM];        /* unwrapped spread grid covers signed cells [-E, E] */
  int32_t G[FLPR_MAX_DIM];        /* 2E+1 */
  int32_t tile[FLPR_MAX_DIM];     /* spread tile edge in cells */
  int32_t ntile[FLPR_MAX_DIM];    /* tiles per dimension = ceil(G / tile) */
  int32_t item_points;            /* points per spread / adjoint work item when a cell
                                     holds many (its points are split over items);
                                     0 = the library default (8 in 1-D, 16 otherwise) */
  int32_t spread_f32;             /* 1: flpr_spread accumulates the footprints in float32
                                     with compensated (Kahan) summation and float32
                                     partial windows (the precision form); 0: float64 */
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

/* One problem of a batch (the flpr_batch_* entry points).  The problems of
 * one batch share the grid structure of the geometry passed beside the
 * table (d, msp, n_grid, L, E, G, tile, ntile); each has its own knot
 * mapping and raw grid (the rest of its FlprGeom) and its own points.  Point
 * arrays of a batch (x, y, the sorted point data) hold the problems' rows one
 * after the other: problem p owns rows [pt, pt + n). */
typedef struct FlprBatchItem {
  FlprGeom g;
  int64_t n;   /* points of the problem */
  int64_t pt;  /* its first row in the batch's point arrays */
} FlprBatchItem;

/* Last HIP error code seen by the calling thread inside the library (0 if
 * none); its message is copied into buf (len bytes, NUL-terminated). */
int flpr_last_hip_error(char* buf, int32_t len);

/* x statistics (model.py:58-68, core.py:83-99) and input checks
 * (api.py:307-323).  x is (n, d) row-major float64; y (n, ny) float64 or NULL.
 * out[5*k+0..4] = min, max, mean, sum of squared deviations, count (finite
 * values of dimension k); out[5*d] = non-finite values in x; out[5*d+1] =
 * non-finite values in y; out[5*d+2+c] = sum_i y_ic^2 (finite values).
 * partials must hold 512*(5*d + 2 + ny) doubles; ny <= 32.  d = 0 gives the
 * y part only (x may be NULL). */
int flpr_xstats(const double* x, int64_t n, int32_t d, const double* y, int32_t ny,
                double* partials, double* out, void* stream);

/* flpr_xstats without allocation.  No counterpart in the reference: this is
 * for callers that cannot allocate device memory or pass a workspace (code
 * under HIP graph capture, serving loops and harnesses that forbid
 * hipMalloc).  Same inputs and out as flpr_xstats; the 512 block partials
 * live in a library-owned static device array of 196 KiB per GPU
 * (512 * (5*3 + 2 + 32) doubles), held until the process ends.  Returns -3
 * when d > 3 or ny > 32 (the partials would not fit).  Allocates nothing
 * (its first call on a device looks up that device's copy of the array).
 * Every call on one GPU shares that array: calls that overlap in time on one
 * GPU (different streams or threads) race and must be serialised by the
 * caller. */
int flpr_xstats_noalloc(const double* x, int64_t n, int32_t d, const double* y, int32_t ny,
                        double* out, void* stream);

/* Spread cell sort keys of every point, from the knot mapping of the
 * reference (convolution.py:169-199 + nufftn_type1.py:140-176, 520-543):
 *   key   (n)    int32     tile-major spread-cell index (tile, then cell in tile)
 *   key_count (nkeys = ntiles * prod(tile)) int32, zeroed by the caller
 *   rank  (n)    int32     the point's place among its key's points (may be NULL). */
int flpr_points_prepare(const double* x, int64_t n, const FlprGeom* g, int32_t* key,
                        int32_t* key_count, int32_t* rank, void* stream);

/* Counting sort of the points by key.  key_start (nkeys+1) receives the
 * exclusive scan of key_count; key_cursor (nkeys) is scratch zeroed by the
 * caller (left zeroed only up to the scatter's own use); order (n) receives
 * point indices grouped by key.  With the ranks of flpr_points_prepare (rank
 * non-NULL) the scatter needs no atomics. */
int flpr_points_bin(const int32_t* key, int64_t n, int32_t nkeys, const int32_t* key_count,
                    int32_t* key_start, int32_t* key_cursor, const int32_t* rank,
                    int32_t* order, void* stream);

/* Sorted point data for the grid kernels, point-major (point p's entries are
 * contiguous):
 *   wab  (n, d, 2) float64  (A, B) per dimension with delta = xmod - hx*round(xmod/hx),
 *                         c = 1/(4 tau): A = exp(-(delta^2 + 2 msp hx delta) c),
 *                         B = exp(2 delta hx c), so that the spread weight of footprint
 *                         offset j - msp (j = 0..2 msp) is
 *                         A B^j exp(-hx^2 (j-msp)^2 c) = exp(-(delta - hx (j-msp))^2 c)
 *                         (nufftn_type1.py:226-249, spreading_fused.py:80-94)
 *   vals (n, ny + nprobe) float64  y columns then probe columns
 *   rawi (n, d) int32, rawf (n, d) float64  interval and fraction on the raw grid
 *                         axis linspace(graw0, graw1, N) (RegularGridInterpolator);
 *                         both may be NULL (not written); wab may be NULL too (a
 *                         second set of values for points already sorted). */
int flpr_points_pack(const FlprGeom* g, const double* x, const int32_t* order, int64_t n,
                     const double* vals_y, int32_t ny, const double* probes, int32_t nprobe,
                     double* wab, double* vals, int32_t* rawi, double* rawf, void* stream);

/* Gaussian-gridding spread onto the unwrapped spread grid
 * (nufftn_type1.py:565-636, spreading_fused.py:67-94, _nufft_accel.pyx:32-76).
 * Channel 0 is the constant 1 when with_ones != 0, then the nval columns of vals.
 * Every tile accumulates its footprint window in float64 and writes it to the
 * workspace part (ntiles * C * prod(tile + 2 msp) doubles); a second pass
 * adds the overlapping windows per cell (no atomics) into grid
 * (C, G0[, G1[, G2]]) float64, and a third adds each channel's cells in a
 * fixed order (one block per channel) into sums (C doubles, zeroed by the
 * caller), so the totals are the same in every run. */
int flpr_spread(const FlprGeom* g, const int32_t* key_start, const double* wab,
                const double* vals, int64_t n, int32_t with_ones, int32_t nval, double* part,
                double* grid, double* sums, void* stream);

/* flpr_spread for three-dimensional footprints wider than its LDS window
 * (the oversampled grid of accuracy 7..12, nufftn_type1.py:520-636 with
 * kernel.py:97-156): d == 3 and 7 <= msp <= 12 only.  Same inputs; no
 * partial windows: every (cell, footprint row, channel) adds its row into
 * grid with float64 atomics, so grid (C, G0, G1, G2) and sums (C) must be
 * zeroed by the caller, and the grid values vary between runs in their last
 * bits (the channel totals are added in a fixed order, as in flpr_spread). */
int flpr_spread_wide(const FlprGeom* g, const int32_t* key_start, const double* wab,
                     const double* vals, int64_t n, int32_t with_ones, int32_t nval,
                     double* grid, double* sums, void* stream);

/* Accuracy > 6 (nufftn_type1.py:693-770, kernel.py:97-156,
 * convolution.py:235-236): the nch channels of fine_grid, spread with the
 * oversampled geometry `fine` (L = R * coarse L, R = 2 or 3, msp = accuracy,
 * tau = pi msp / (L^2 R (R - 1/2)) of the coarse L), are Fourier-transformed,
 * cut to the central L frequencies, divided by the Gaussian's transform and
 * scaled by 1/R^d, then transformed back onto the coarse unwrapped grid
 * (coarse E >= N on every axis; cells equal modulo L hold equal values) -- all
 * as one separable linear map in float64.  With pairs, channels (2q, 2q+1)
 * are the real and imaginary parts of one complex channel; otherwise every
 * channel is real and only the real part of its result is kept.  Writes the
 * nch channels of grid (coarse cells each) and their totals into sums.  work
 * holds flpr_resample_work(...) doubles. */
int flpr_resample_work(const FlprGeom* fine, const FlprGeom* coarse, int32_t nch, int32_t pairs,
                       int64_t* doubles);
int flpr_resample(const FlprGeom* fine, const FlprGeom* coarse, const double* fine_grid,
                  int32_t nch, int32_t pairs, double* work, double* grid, double* sums,
                  void* stream);

/* Grids narrower than their spread footprint (G > L on some axis; small N):
 * the reference's spread grid is L-periodic and its convolution circular
 * (nufftn_type1.py:565-636, convolution.py:904-951).  Sets every cell of the
 * C channels of grid (the unwrapped grid, G0 x .. cells per channel) to the
 * sum over its residue class modulo L, after which the convolution passes
 * compute the circular convolution.  work holds C * prod(L) doubles.  Apply
 * once to freshly spread channels (a second application would multiply). */
int flpr_grid_periodize(const FlprGeom* g, double* grid, int32_t C, double* work, void* stream);

/* Centre channels 1..C-1 of the spread grid on their kernel-weighted mean:
 * shift[c] = sums[c] / sums[0], F_c -= shift[c] F_0 (float64, in place), and
 * optionally write a float32 copy of the whole grid.  shift (C doubles) may be
 * NULL.  The solve adds the shift back exactly (flpr_solve). */
int flpr_grid_center(double* grid, int32_t C, int64_t cells, const double* sums, double* shift,
                     float* grid32, void* stream);

/* Kernel-moment taps (convolution.py:412-512, kernel.py:330-357,
 * design_matrix_optimized.py:216-298): for bandwidth row b (h is (dh, d)),
 * dimension k and power a in [0, npow): tap[((b*d + k)*npow + a)*nmax + t] =
 * exp(-u^2/2) u^a with u = linspace(xmin_s, xmin_s+scale_s, N_k)[t] / h[b,k];
 * dimension 0 is further scaled by 1/(sqrt(2 pi) prod_k h[b,k] * norm).
 * Written in float64 (tap64) and float32 (tap32, may be NULL). */
int flpr_taps(const FlprGeom* g, const double* h, int32_t dh, int32_t npow, int32_t nmax,
              double norm, double* tap64, float* tap32, void* stream);

/* The float32 form's tap operand for flpr_conv: every float32 tap row
 * (ntap = dh * d * npow rows of nmax, the tap32 layout of flpr_taps; row A
 * uses band A / npow) masked to its band, reversed and split exactly into
 * three bfloat16 planes (hi + mid + lo), in eight copies shifted by one
 * element each: ntap * 24 * flpr_conv_strip_len(nmax) uint16 values. */
int64_t flpr_conv_strip_len(int32_t nmax);
int flpr_conv_strips(const float* tap32, const int32_t* bands, int32_t ntap, int32_t npow,
                     int32_t nmax, uint16_t* strips, void* stream);

/* One separable convolution pass along one axis for a batch of jobs: the
 * reference's spectrum product + inverse FFT + qout extraction
 * (convolution.py:904-951) restated as a banded Toeplitz product on the
 * unwrapped grid, computed with MFMA.
 *   mode 0 (ROW, contiguous axis):  out[r][i] = sum_s in[r][s] * tap[i + base - s]
 *   mode 1 (COL, strided axis):     out[o][i][c] = sum_s tap[i + base - s] * in[o][s][c]
 * Job j (device array of nj records of 8 int64: in_off, out_off, tap_off,
 * band, rows_or_outer, powers, out_stride, 0) reads from in + in_off and
 * writes out + out_off; taps outside [bands[2*band], bands[2*band+1]] (device
 * int32 array) are treated as zero.  K = input axis length, nout = output axis
 * length, inner = contiguous extent for mode 1, rows_or_outer = rows (mode 0)
 * or outer batch count (mode 1) of the job, max_rows_or_outer its maximum.
 * fp64 != 0: float64 MFMA, in/out/tap are double* (tap in the flpr_taps
 * layout, row length nmax); one tap row per job (powers 0 or 1, max_powers
 * 1).  fp64 == 0: float32 in/out, products of the three-term bfloat16 splits
 * on bf16 MFMA (float32 accuracy), tap = the flpr_conv_strips table of the
 * float32 taps: a job applies the powers (1 to max_powers <= 3; 0 counts as
 * 1) consecutive rows tap_off / nmax + p of that table to the same input and
 * writes row p's output at out + out_off + p * out_stride, loading and
 * splitting the input once for all of them.
 * Tiles are read whole: in must stay readable for 256*K + 64*inner + 512
 * elements past the end of every job's input (those values are discarded or
 * zeroed, never used). */
int flpr_conv(int32_t mode, int32_t fp64, const void* in, void* out, const void* tap,
              int32_t nmax, const int64_t* jobs, const int32_t* bands, int32_t nj,
              int32_t max_rows_or_outer, int32_t K, int32_t nout, int32_t inner, int32_t base,
              int32_t max_powers, const int32_t* blocks, int32_t nblocks, void* stream);

/* Tiles of one flpr_conv job (rows_or_outer of the job, max_rows_or_outer of
 * the launch) and the stride of the tile numbering.  With a block table
 * (blocks, device int32 pairs, nblocks of them) block b runs tile
 * blocks[2b+1] in [0, tiles) of job blocks[2b] (a negative job idles): the
 * caller chooses the order and the placement of the tiles (the device hands
 * consecutive blocks to its XCDs in turn).  Without one (blocks NULL) the
 * launch runs every tile of every job, largest job first in x. */
int flpr_conv_tiles(int32_t mode, int32_t fp64, int32_t rows_or_outer, int32_t max_rows_or_outer,
                    int32_t nout, int32_t inner, int32_t* tiles, int32_t* tile_stride);

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
 *   pivot     nonzero: nodes where S + lam I is not positive definite (the
 *             deconvolved grids of accuracy > 6) are solved by Gaussian
 *             elimination with partial pivoting instead of giving NaN
 *   s_off[ns] slab offset of S pair (i,j>=i) (row-major pairs), channel 0
 *   t_off[nt] slab offset of term i; channel c adds c*nodes
 *             (s_off and t_off are HOST arrays copied into the kernel arguments)
 *   key       max key from flpr_max_moment; lam_scale multiplies it
 *             (1e-6 for order >= 1, regularization for order 0)
 *   eps_add   added to S for order 0 (regression.py:127)
 *   shift     per-channel centring of flpr_grid_center (C doubles) or NULL
 *   nval      value channels 1..nval are solved (<= 32)
 *   nout      the first nout of them are written to m_out (dh, nout, nodes) float64
 *   q, dot    when q (nval, nodes) float64 is given, dot[h*nval + c] accumulates
 *             sum_node q[c][node] m_c(h, node) (dot zeroed by the caller)
 *   sum_part  with q: workspace of sum_part_len doubles for the dots' block
 *             partials, added in block order (no floating-point atomics, so
 *             dot is the same in every run); it must be all zero on entry
 *             (each call leaves it all zero).  dh * nval * 256 +
 *             (dh + 1) / 2 doubles always suffice (at most 256 node blocks
 *             per bandwidth); with less the launch uses fewer node blocks
 *             (-3 when not even one per bandwidth fits).
 *             May be NULL without q. */
int flpr_solve(int32_t fp64, const void* mom, int64_t per_h, int32_t dh, int64_t nodes,
               int32_t order, int32_t nt, int32_t pivot, const int64_t* s_off,
               const int64_t* t_off, const uint64_t* key, double lam_scale, double eps_add,
               const double* shift, int32_t nval, int32_t nout, double* m_out, const double* q,
               double* dot, double* sum_part, int64_t sum_part_len, void* stream);

/* flpr_solve for float32 moments in the precision form (no pivoting): for
 * order >= 1 each node's S + lam I is factored in float32 (Cholesky) and z =
 * (S + lam I)^-1 e1 takes one refinement step with the residual formed in
 * float64; nodes whose float32 factorisation meets a non-positive pivot or
 * whose pivot ratio max/min (a cheap lower estimate of the condition
 * number) exceeds 1e6 are solved in float64.  The rest as flpr_solve
 * (regression.py:162-277). */
int flpr_solve_refined(const void* mom, int64_t per_h, int32_t dh, int64_t nodes,
                       int32_t order, int32_t nt, const int64_t* s_off, const int64_t* t_off,
                       const uint64_t* key, double lam_scale, double eps_add,
                       const double* shift, int32_t nval, int32_t nout, double* m_out,
                       const double* q, double* dot, double* sum_part, int64_t sum_part_len,
                       void* stream);

/* Adjoint interpolation sums on the raw grid, from the sorted point data:
 *   q (nval, nodes)        q[c][node] = sum_i vals[i][c] w_i(node)
 *   gram (nw, rcells)      per raw cell the Gram entries sum_i w_i(a) w_i(b) of
 *                          its 2^d corner weights, a <= b row-major (nw =
 *                          2^d (2^d + 1) / 2; rcells = prod(N_k - 1)); may be NULL
 * with w_i the multilinear weights of RegularGridInterpolator 'linear'.  Then
 *   sum_i v_i yhat_i = sum_node q m   and   sum_i yhat_i^2 = sum_cell m' W m,
 * which turn the Hutchinson dots p' phat and the residual sums of
 * regression.py:1256-1280 / api.py:686-723 into grid-sized work per bandwidth.
 * part is workspace of ntiles * max(nval, nw) * max(prod(tile + 5),
 * 2^d prod(tile)) doubles; q and gram must be zeroed by the caller. */
int flpr_adjoint(const FlprGeom* g, const int32_t* key_start, const int32_t* rawi,
                 const double* rawf, const double* vals, int64_t n, int32_t nval, double* part,
                 double* q, double* gram, void* stream);

/* quad[h*ncol + c] += sum_cell m' W m for grids m (dh, ncol, nodes) float64.
 * The cell blocks' partial sums are added in block order through the
 * workspace sum_part (sum_part_len doubles, all zero on entry and left zero;
 * dh * ncol * 16 + (dh*ncol + 1) / 2 suffice, less uses fewer blocks; as
 * flpr_solve), so quad is the same in every run. */
int flpr_quadform(const FlprGeom* g, const double* m, int32_t ncol, int32_t dh,
                  const double* gram, double* quad, double* sum_part, int64_t sum_part_len,
                  void* stream);

/* Multilinear interpolation of one grid (ncol channels, (ncol, nodes)) at n
 * points (api.py:760-763, interp_optimized.py / scipy linear with linear
 * extrapolation outside the grid).  out (n, ncol) float64.  When
 * yvar != NULL also accumulates *expsum += sum_i yvar[i] exp(-out[i,0])
 * (api.py:771-778). */
int flpr_interp_eval(const FlprGeom* g, const double* x, int64_t n, const double* m,
                     int32_t ncol, double* out, const double* yvar, double* expsum,
                     void* stream);

/* The fit's outputs at the selected bandwidth (api.py:760-786): m_all is
 * (dh, ncol, nodes) float64, sel a device int32 index into it (NULL: 0).
 *   yhat  (n, ncol)  multilinear interpolation of m_all[sel] at x
 *   resid (n, ncol)  yref - yhat (resid may be NULL; yref (n, ncol))
 * variance != 0 (ncol = 1; yref = the raw response y of the variance fit, n):
 * with d = n / sum_i yref_i exp(-fit_i), yhat = exp(fit) / d, grid (nodes)
 * = exp(m_all[sel]) / d, resid = yref - yhat, fac[0] = 1 / d (may be NULL);
 * part holds 1024 doubles of partial sums. */
int flpr_fit_output(const FlprGeom* g, const double* x, int64_t n, const double* m_all,
                    int32_t ncol, const int32_t* sel, const double* yref, int32_t variance,
                    double* yhat, double* resid, double* grid, double* part, double* fac,
                    void* stream);

/* Cubic spline interpolation (RegularGridInterpolator 'cubic', i.e. scipy's
 * make_interp_spline(k=3) with not-a-knot ends along every axis; the
 * reference's GCV / DOF interpolation for 2-D and 3-D data with fewer than
 * 5000 points, api.py:588-594, 651-657, regression.py:929-1092).  coef
 * (ngrid, N_0, ..., N_{d-1}) float64 B-spline coefficients on the raw grid
 * axes linspace(graw0, graw1, N), knots x_0 (four times), x_2 .. x_{N-3}, x_{N-1}
 * (four times);
 * out (n, ngrid) float64 values at x (n, d); N_k >= 4. */
int flpr_spline_eval(const FlprGeom* g, const double* x, int64_t n, const double* coef,
                     int32_t ngrid, double* out, void* stream);

/* Likelihood cross-validation of density grids (api.py:1111-1132): dens
 * (dh, nodes) float64 densities on the raw grid, x (n, d) float64 data points,
 * k0h (dh) K(0) / prod_k h_bk.  For every bandwidth b,
 *   l_i = log(max((n/(n-1)) (f_b(x_i) - k0h[b]/n), 1e-10))
 * with f_b the linear interpolant (RegularGridInterpolator 'linear');
 * out[b] = sum_i l_i and out[dh + b] = std(l) sqrt(n) (ddof = 0).
 * order (n) int32, may be NULL: the order in which the points are visited,
 * e.g. flpr_points_bin's grouping by spread cell, which keeps the density
 * reads of neighbouring threads in cache; NULL visits them as stored.
 * part is workspace of 1024 * dh * 3 doubles. */
int flpr_lcv(const FlprGeom* g, const double* x, const int32_t* order, int64_t n,
             const double* dens, int32_t dh, const double* k0h, double* part, double* out,
             void* stream);

/* Bandwidth selection from the per-bandwidth sums, on the device
 * (regression.py:1276-1354 Hutchinson DOF and penalties, regression.py:378-421
 * GCV with aggregation agg = 0 first / 1 average / 2 sum, api.py:73-164 1-SE
 * rule with dstd, preferring the larger sum(h), ties to the smaller score).
 *   red   (dh, nval) dots [y columns (ny), probes (nprobe)] then (dh, ny) sums
 *         sum_cell m' W m, as flpr_solve / flpr_quadform leave them
 *   ysq   (ny) sum_i y_ic^2; pp (nprobe) p'p per probe; hsum (dh) sum_k h_bk
 *   cplx  y columns are (re, im) pairs: residual sums pair up, penalties use
 *         2 pdof - 1 (regression.py:1310-1317)
 * out (4 dh + 3) float64: GCV scores, their standard errors, dof, dof standard
 * errors (dh each), then the selected index, the index of the smallest score
 * (np.argmin) and the number of finite scores; *sel (device int32, may be
 * NULL) also receives the selected index, so work at that bandwidth can be
 * queued behind this call without a host round trip. */
int flpr_select(const double* red, int32_t dh, int32_t nval, int32_t ny, int32_t nprobe,
                const double* ysq, const double* pp, const double* hsum, int64_t n,
                int32_t cplx, int32_t agg, double dstd, double* out, int32_t* sel,
                void* stream);

/* ------------------------------------------------------------------------
 * Batches of independent problems (the reference fits one problem per
 * cv_fastlpr call, api.py:167-832; these run the same stages for P problems
 * of one grid structure in one launch each, problem = blockIdx.z).  items is
 * a DEVICE array of P FlprBatchItem; g (host) is the shared grid structure
 * (items[p].g agrees with it on d, msp, n_grid, L, E, G, tile, ntile); nmax
 * is the largest n of the batch.  Per-problem buffers lie at the stated
 * strides (elements); every problem computes exactly what the single-problem
 * entry point computes for it.
 * ------------------------------------------------------------------------ */

/* Clears bytes (a multiple of 4) of device memory with a kernel: a node that
 * a captured HIP graph replays correctly, where a memset node may not. */
int flpr_fill_zero(void* p, int64_t bytes, void* stream);

/* flpr_xstats per problem: partials P * 512 * (5d + 2 + ny), out P * (5d + 2
 * + ny) doubles; x (rows, d) and y (rows, ny) hold the batch's rows (d = 0:
 * the y part only, x may be NULL). */
int flpr_batch_xstats(const FlprBatchItem* items, int32_t P, const double* x, int32_t d,
                      const double* y, int32_t ny, double* partials, double* out, void* stream);

/* flpr_points_prepare per problem with its own knot mapping; problem p's keys
 * are offset by p * nkeys(g), so one flpr_points_bin over all rows and P *
 * nkeys keys sorts every problem into its own row range [pt, pt + n). */
int flpr_batch_points_prepare(const FlprGeom* g, const FlprBatchItem* items, int32_t P,
                              int64_t nmax, const double* x, int32_t* key, int32_t* key_count,
                              int32_t* rank, void* stream);

/* flpr_points_pack per problem (its sorted rows); probes (nmax rows) are read
 * at the row within the problem: one probe matrix whose leading rows serve
 * every problem (the reference's MT19937 matrix of n rows is the leading n
 * rows of any longer one, regression.py:1193-1194). */
int flpr_batch_points_pack(const FlprGeom* g, const FlprBatchItem* items, int32_t P, int64_t nmax,
                           const double* x, const int32_t* order, const double* vals_y, int32_t ny,
                           const double* probes, int32_t nprobe, double* wab, double* vals,
                           int32_t* rawi, double* rawf, void* stream);

/* flpr_spread per problem: part P * (ntiles * C * window cells) doubles, the
 * grid of problem p at grid + p * grid_stride (C * cells used), sums P * C
 * (zeroed by the caller). */
int flpr_batch_spread(const FlprGeom* g, const FlprBatchItem* items, int32_t P, int64_t nmax,
                      const int32_t* key_start, const double* wab, const double* vals,
                      int32_t with_ones, int32_t nval, double* part, double* grid,
                      int64_t grid_stride, double* sums, void* stream);

/* flpr_grid_center per problem (grid and grid32 at grid_stride, sums and
 * shift at C). */
int flpr_batch_grid_center(int32_t P, double* grid, int64_t grid_stride, int32_t C, int64_t cells,
                           const double* sums, double* shift, float* grid32, void* stream);

/* flpr_taps per problem (its geometry; norm = its point count): tap tables at
 * a stride of dh * d * npow * nmax. */
int flpr_batch_taps(const FlprGeom* g, const FlprBatchItem* items, int32_t P, const double* h,
                    int32_t dh, int32_t npow, int32_t nmax, double* tap64, float* tap32,
                    void* stream);

/* flpr_max_moment per problem: slabs at src + p * prob_stride, key[p]. */
int flpr_batch_max_moment(int32_t fp64, int32_t P, const void* src, int64_t prob_stride,
                          int64_t per_h, int32_t dh, const int64_t* offs, int32_t nslab,
                          int64_t nodes, int32_t absval, uint64_t* key, void* stream);

/* flpr_solve per problem (no pivoting): moments at mom_stride, key[p], shift
 * at nval + 1, m_out at m_stride, q at nval * nodes, dot at dot_stride; the
 * dot workspace as flpr_solve's with P * dh groups. */
int flpr_batch_solve(int32_t fp64, int32_t P, const void* mom, int64_t mom_stride, int64_t per_h,
                     int32_t dh, int64_t nodes, int32_t order, int32_t nt, const int64_t* s_off,
                     const int64_t* t_off, const uint64_t* key, double lam_scale, double eps_add,
                     const double* shift, int32_t nval, int32_t nout, double* m_out,
                     int64_t m_stride, const double* q, double* dot, int64_t dot_stride,
                     double* sum_part, int64_t sum_part_len, void* stream);

/* flpr_adjoint per problem: part P * (ntiles * max(nval, nw) *
 * max(prod(tile + 5), 2^d prod(tile))) doubles, q at nval * nodes, gram at
 * nw * rcells (both zeroed by the caller).  Even node counts on every axis
 * use the cell walk, as flpr_adjoint does. */
int flpr_batch_adjoint(const FlprGeom* g, const FlprBatchItem* items, int32_t P, int64_t nmax,
                       const int32_t* key_start, const int32_t* rawi, const double* rawf,
                       const double* vals, int32_t nval, double* part, double* q, double* gram,
                       void* stream);

/* flpr_quadform per problem: m at m_stride, gram at nw * rcells, quad at
 * quad_stride (zeroed by the caller); the workspace as flpr_quadform's with
 * P * dh * ncol groups. */
int flpr_batch_quadform(const FlprGeom* g, int32_t P, const double* m, int64_t m_stride,
                        int32_t ncol, int32_t dh, const double* gram, double* quad,
                        int64_t quad_stride, double* sum_part, int64_t sum_part_len,
                        void* stream);

/* flpr_select per problem (n from items): red at red_stride, ysq at ny, pp at
 * nprobe, out at out_stride (>= 4 dh + 3), sel[p]. */
int flpr_batch_select(const FlprBatchItem* items, int32_t P, const double* red,
                      int64_t red_stride, int32_t dh, int32_t nval, int32_t ny, int32_t nprobe,
                      const double* ysq, const double* pp, const double* hsum, int32_t cplx,
                      int32_t agg, double dstd, double* out, int64_t out_stride, int32_t* sel,
                      void* stream);

/* flpr_fit_output per problem: m_all at m_stride, sel[p], yref / yhat / resid
 * at the problem's rows, grid at nodes, part P * 1024, fac at fac_stride. */
int flpr_batch_fit_output(const FlprGeom* g, const FlprBatchItem* items, int32_t P, int64_t nmax,
                          const double* x, const double* m_all, int64_t m_stride, int32_t ncol,
                          const int32_t* sel, const double* yref, int32_t variance, double* yhat,
                          double* resid, double* grid, double* part, double* fac,
                          int64_t fac_stride, void* stream);

/* B-spline coefficients for the cubic GCV / DOF interpolation
 * (RegularGridInterpolator 'cubic', api.py:588-594, 651-657): for every
 * problem p and grid of m (P * ngrid grids of prod(N) values, row-major on
 * the raw grid), coef = A_0^-1 x_0 A_1^-1 x_1 ... m, the inverse not-a-knot
 * collocation matrix of every axis applied along that axis.  inv holds the
 * problem's matrices (A_0^-1 then A_1^-1 ..., N_k x N_k row-major) at
 * inv_stride; work is a scratch of the size of m (unused for d = 1). */
int flpr_batch_spline_coef(const FlprGeom* g, int32_t P, const double* inv, int64_t inv_stride,
                           const double* m, int32_t ngrid, double* work, double* coef,
                           void* stream);

/* Residual sums and Hutchinson dots at the data points for the grid
 * solutions m (P x (dh, ny + nprobe, nodes)): yhat of the y channels and phat
 * of the probe channels are the cubic spline of coef (gcv_cubic, dof_cubic)
 * or the multilinear interpolation of m; writes sum_i (y_ic - yhat_ic)^2 at
 * red[dh (ny + nprobe) + h ny + c] and sum_i p_ik phat_ik at red[h (ny +
 * nprobe) + ny + k] (problem p at red + p red_stride; the y dots are left
 * as they are), summed in a fixed order (regression.py:1256-1280,
 * api.py:661-723).  y (rows, ny) at the problems' rows, probes (nmax rows)
 * at the row within the problem.  part holds P * ceil(dh (ny + nprobe) / 8)
 * * 8 * min(64, ceil(nmax / 256)) doubles. */
int flpr_batch_point_sums(const FlprGeom* g, const FlprBatchItem* items, int32_t P, int64_t nmax,
                          const double* x, const double* y, int32_t ny, const double* probes,
                          int32_t nprobe, const double* m, const double* coef, int32_t dh,
                          int32_t gcv_cubic, int32_t dof_cubic, double* part, double* red,
                          int64_t red_stride, void* stream);

/* ------------------------------------------------------------------------
 * The learned bandwidth bracket.  No counterpart in the reference, which
 * sweeps every bandwidth (api.py:664-740): a small MLP proposes where the
 * GCV minimum lies so that only a few bandwidths need full solves; the full
 * sweep remains the fallback.  The MLP is 64 -> 128 -> 128 -> 2 with GELU
 * (0.5 z (1 + erf(z / sqrt 2))); parameters packed as W1 (128 x 64), b1
 * (128), W2 (128 x 128), b2, W3 (2 x 128), b3 (2), float32.  GEMM operands
 * are rounded to bfloat16, products accumulated in float32 (bf16 MFMA).
 * ------------------------------------------------------------------------ */
int flpr_mlp_param_count(void);

/* C (M x N, ldc) = op(A) op(B) with op(A)[i][k] = A[i lda + k] (ta = 0) or
 * A[k lda + i], op(B)[k][j] = B[k ldb + j] (tb = 0) or B[j ldb + k];
 * mode 0: C = acc (+ bias[j] if bias); mode 1: aux = acc + bias[j], C =
 * gelu(aux); mode 2: C = acc * gelu'(aux) (aux read, ldc layout). */
int flpr_mlp_gemm(int32_t M, int32_t N, int32_t K, const float* A, int32_t lda, int32_t ta,
                  const float* B, int32_t ldb, int32_t tb, const float* bias, int32_t mode,
                  float* aux, float* C, int32_t ldc, void* stream);

/* Forward pass of M rows X (M x 64): pre-activations Z1, Z2 and activations
 * H1, H2 (M x 128 each) for the backward pass, outputs O (M x 2). */
int flpr_mlp_forward(const float* params, int32_t M, const float* X, float* Z1, float* H1,
                     float* Z2, float* H2, float* O, void* stream);

/* Gradient of sum(O * dO) with respect to the packed parameters (grads, the
 * parameter layout), from a forward pass's X, Z1, H1, Z2, H2; work holds
 * 2 * M * 128 floats. */
int flpr_mlp_backward(const float* params, int32_t M, const float* X, const float* Z1,
                      const float* H1, const float* Z2, const float* H2, const float* dO,
                      float* grads, float* work, void* stream);

/* loss[0] = sum_i (w0 (O_i0 - T_i0)^2 + w1 (O_i1 - T_i1)^2) / M and its
 * gradient dO (M x 2). */
int flpr_mlp_loss(int32_t M, const float* O, const float* T, float w0, float w1, float* dO,
                  float* loss, void* stream);

/* One Adam step (bias-corrected, step >= 1, L2 weight decay added to the
 * gradient) on n float32 parameters with moments m, v. */
int flpr_mlp_adam(int32_t n, float* params, const float* grads, float* m, float* v, float lr,
                  float beta1, float beta2, float eps, int32_t step, float weight_decay,
                  void* stream);

/* The 64 bracket features of P spread grids (problem p at grid + p
 * grid_stride: channels F0 = sum w and F1 = sum w y of G0 x G1 cells each,
 * before any centring), with ysq[p] = sum_i y_i^2 and npts[p] points:
 * 0..55 log weighted mean squared increments of m = F1 / F0 over 14 scales
 * (1 .. 48 cells) along (1,0), (0,1), (1,1), (1,-1), with weight
 * F0(c) F0(c') / (F0(c) + F0(c')) where both exceed floor_[p]; 56 log10
 * npts; 57 log variance of y; 58 log local variance (ysq / n minus the
 * F0-weighted mean of m^2); 59 log mean F0; 60 log squared coefficient of
 * variation of F0; 61 share of cells below the floor; 62 log(variance /
 * local variance); 63 log10 dx[p].  part holds P * 57 * 64 * 8 doubles; feat
 * (P x 64) float32. */
int flpr_bracket_features(int32_t P, const double* grid, int64_t grid_stride, int32_t G0,
                          int32_t G1, const double* floor_, const double* npts, const double* ysq,
                          const double* dx, double* part, float* feat, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FASTLPR_AMD_H */
