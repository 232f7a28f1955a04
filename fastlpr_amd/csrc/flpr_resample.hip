// Accuracy > 6: the oversampled spread grid, deconvolved and resampled onto
// the unwrapped grid of the convolution passes.
//
// Reference dataflow restated here (fastLPR_py/src/fastlpr/...):
//   kernel.py:97-156          R = 2 (accuracy <= 11) or 3, Mr = R L,
//                             tau = pi Msp / (L^2 R (R - 1/2))
//   nufftn_type1.py:693-770   FFT of the Mr^d spread grid, fftshift, / R^d,
//                             the central L frequencies, times 1/Kn with
//                             Kn = sqrt(tau/pi) exp(-tau kappa^2) per axis
//   convolution.py:235-236    ifftshift; then convolution.py:904-951 multiplies
//                             by the kernel spectrum and inverse-FFTs over L.
// Writing g = IFFT_L(deconvolved spectrum), the result of the reference is the
// circular convolution of g with the kernel samples, which is what the
// convolution passes compute from the unwrapped grid.  g is a linear map of
// the fine grid that only depends on the phase difference:
//   g[j] = sum_c F[c] a[(R j - c) mod Mr],
//   a[delta] = 1/(L R) sum_{kappa=-L/2}^{L/2-1} sqrt(pi/tau) exp(tau kappa^2)
//              exp(2 pi i kappa delta / Mr),
// applied one axis at a time (the factors are separable).  j runs over the
// signed coarse cells [-E, E] of the unwrapped grid (periodic in L: cells that
// coincide modulo L get equal values), c over the signed fine cells.  g is not
// compactly supported, so the coarse grid must reach E >= N to hold every cell
// a node's kernel window sees.
#include "flpr_common.h"

namespace {

constexpr int RS_THREADS = 256;

// a[delta] for one axis, delta in [0, Mr).
__global__ __launch_bounds__(RS_THREADS) void k_deconv_table(double2* __restrict__ tab, int Mr,
                                                             int L, int R, double tau) {
  const double w0 = sqrt(M_PI / tau);
  const double scale = 1.0 / ((double)L * (double)R);
  for (int delta = blockIdx.x * blockDim.x + threadIdx.x; delta < Mr;
       delta += gridDim.x * blockDim.x) {
    double re = 0.0, im = 0.0;
    for (int kap = -L / 2; kap < L / 2; ++kap) {
      const double w = w0 * exp(tau * (double)kap * (double)kap);
      int64_t r = ((int64_t)kap * delta) % Mr;
      if (r < 0) r += Mr;
      double sn, cs;
      sincospi(2.0 * (double)r / (double)Mr, &sn, &cs);
      re += w * cs;
      im += w * sn;
    }
    tab[delta] = make_double2(re * scale, im * scale);
  }
}

// One axis: out[ch][p][j][q] = sum_c a[(R (j - Eout) - (c - Ein)) mod Mr] in[ch][p][c][q].
// IN_REAL: the input is the real fine grid (channel stride in_cells); with
// pairs, real channels 2ch and 2ch+1 are the real and imaginary parts of
// complex channel ch, else channel ch is real.  OUT_REAL: the output is written
// to the real coarse grid (channel stride out_cells), real part only unless
// pairs.  Otherwise the arrays are complex, (nchc, pre, G, post).
template <bool IN_REAL, bool OUT_REAL>
__global__ __launch_bounds__(RS_THREADS) void k_resample(
    const double* __restrict__ in, double* __restrict__ out, const double2* __restrict__ tab,
    int nchc, int pairs, int64_t pre, int Gin, int Ein, int64_t post, int Gout, int Eout, int R,
    int Mr, int64_t in_cells, int64_t out_cells) {
  const int64_t total = (int64_t)nchc * pre * Gout * post;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = idx % post;
    int64_t t = idx / post;
    const int j = (int)(t % Gout);
    t /= Gout;
    const int64_t p = t % pre;
    const int ch = (int)(t / pre);
    int dl = (int)(((int64_t)R * (j - Eout) + Ein) % Mr);
    if (dl < 0) dl += Mr;
    double re = 0.0, im = 0.0;
    const int64_t base = p * Gin * post + q;
    for (int c = 0; c < Gin; ++c) {
      const double2 a = tab[dl];
      const int64_t off = base + (int64_t)c * post;
      double vr, vi;
      if constexpr (IN_REAL) {
        if (pairs) {
          vr = in[(int64_t)(2 * ch) * in_cells + off];
          vi = in[(int64_t)(2 * ch + 1) * in_cells + off];
        } else {
          vr = in[(int64_t)ch * in_cells + off];
          vi = 0.0;
        }
      } else {
        const double2 v = reinterpret_cast<const double2*>(in)[(int64_t)ch * pre * Gin * post + off];
        vr = v.x;
        vi = v.y;
      }
      re += a.x * vr - a.y * vi;
      im += a.x * vi + a.y * vr;
      dl = (dl == 0) ? Mr - 1 : dl - 1;
    }
    const int64_t o = (p * Gout + j) * post + q;
    if constexpr (OUT_REAL) {
      if (pairs) {
        out[(int64_t)(2 * ch) * out_cells + o] = re;
        out[(int64_t)(2 * ch + 1) * out_cells + o] = im;
      } else {
        out[(int64_t)ch * out_cells + o] = re;
      }
    } else {
      reinterpret_cast<double2*>(out)[(int64_t)ch * pre * Gout * post + o] = make_double2(re, im);
    }
  }
}

// sums[c] = sum of channel c of the coarse grid (one block per channel).
__global__ __launch_bounds__(RS_THREADS) void k_channel_sums(const double* __restrict__ grid,
                                                             int64_t cells,
                                                             double* __restrict__ sums) {
  __shared__ double red[RS_THREADS / 64];
  const double* g = grid + (int64_t)blockIdx.x * cells;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < cells; i += blockDim.x) s += g[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < RS_THREADS / 64; ++w) r += red[w];
    sums[blockIdx.x] = r;
  }
}

struct Plan {
  int d, R, nchc;
  int Mr[FLPR_MAX_DIM];
  int64_t table;      // doubles of the tables (2 per complex entry)
  int64_t inter;      // doubles of one complex intermediate
};

bool make_plan(const FlprGeom* f, const FlprGeom* c, int32_t nch, int32_t pairs, Plan* p) {
  if (!flpr_geom_ok(f) || !flpr_geom_ok(c) || f->d != c->d || nch < 1 || (pairs && (nch & 1)))
    return false;
  const int d = c->d;
  int R = 0;
  for (int k = 0; k < d; ++k) {
    if (f->L[k] % c->L[k] != 0 || (c->L[k] & 1)) return false;
    const int r = f->L[k] / c->L[k];
    if ((r != 2 && r != 3) || (R && r != R)) return false;
    R = r;
    if (f->n_grid[k] != c->n_grid[k] || c->E[k] < c->n_grid[k]) return false;
  }
  p->d = d;
  p->R = R;
  p->nchc = pairs ? nch / 2 : nch;
  p->table = 0;
  for (int k = 0; k < d; ++k) {
    p->Mr[k] = f->L[k];
    p->table += 2 * (int64_t)f->L[k];
  }
  // after axes d-1..k are resampled: fine sizes below k, coarse from k on
  p->inter = 0;
  for (int k = d - 1; k >= 1; --k) {
    int64_t sz = p->nchc;
    for (int j = 0; j < d; ++j) sz *= (j < k) ? f->G[j] : c->G[j];
    if (2 * sz > p->inter) p->inter = 2 * sz;
  }
  return true;
}

}  // namespace

extern "C" {

int flpr_resample_work(const FlprGeom* fine, const FlprGeom* coarse, int32_t nch, int32_t pairs,
                       int64_t* doubles) {
  Plan p;
  if (!doubles || !make_plan(fine, coarse, nch, pairs, &p)) return -1;
  *doubles = p.table + 2 * p.inter;
  return 0;
}

int flpr_resample(const FlprGeom* fine, const FlprGeom* coarse, const double* fine_grid,
                  int32_t nch, int32_t pairs, double* work, double* grid, double* sums,
                  void* stream) {
  FLPR_CLEAR_ERROR();
  Plan p;
  if (!fine_grid || !work || !grid || !sums || !make_plan(fine, coarse, nch, pairs, &p))
    return -1;
  hipStream_t st = as_stream(stream);
  const int d = p.d;
  double2* tab[FLPR_MAX_DIM];
  int64_t off = 0;
  for (int k = 0; k < d; ++k) {
    tab[k] = reinterpret_cast<double2*>(work + off);
    off += 2 * (int64_t)p.Mr[k];
    const int Lc = coarse->L[k];
    // kernel.py:97-156, in the reference's operation order
    const double tau = (M_PI * fine->msp) / ((double)((int64_t)Lc * Lc * p.R) * (p.R - 0.5));
    hipLaunchKernelGGL(k_deconv_table, dim3(flpr_grid_for(p.Mr[k], RS_THREADS, 1024)),
                       dim3(RS_THREADS), 0, st, tab[k], p.Mr[k], Lc, p.R, tau);
    FLPR_CHECK_LAUNCH();
  }
  double* buf[2] = {work + p.table, work + p.table + p.inter};
  int64_t cells_f = 1, cells_c = 1;
  for (int k = 0; k < d; ++k) {
    cells_f *= fine->G[k];
    cells_c *= coarse->G[k];
  }
  const double* src = fine_grid;
  for (int k = d - 1; k >= 0; --k) {
    int64_t pre = 1, post = 1;
    for (int j = 0; j < k; ++j) pre *= fine->G[j];
    for (int j = k + 1; j < d; ++j) post *= coarse->G[j];
    const bool first = (k == d - 1), last = (k == 0);
    double* dst = last ? grid : buf[(d - 1 - k) & 1];
    const int64_t total = (int64_t)p.nchc * pre * coarse->G[k] * post;
    dim3 grd(flpr_grid_for(total, RS_THREADS, 65536));
#define FLPR_RS_ARGS                                                                          \
  src, dst, tab[k], p.nchc, pairs ? 1 : 0, pre, fine->G[k], fine->E[k], post, coarse->G[k],   \
      coarse->E[k], p.R, p.Mr[k], cells_f, cells_c
    if (first && last)
      hipLaunchKernelGGL((k_resample<true, true>), grd, dim3(RS_THREADS), 0, st, FLPR_RS_ARGS);
    else if (first)
      hipLaunchKernelGGL((k_resample<true, false>), grd, dim3(RS_THREADS), 0, st, FLPR_RS_ARGS);
    else if (last)
      hipLaunchKernelGGL((k_resample<false, true>), grd, dim3(RS_THREADS), 0, st, FLPR_RS_ARGS);
    else
      hipLaunchKernelGGL((k_resample<false, false>), grd, dim3(RS_THREADS), 0, st, FLPR_RS_ARGS);
#undef FLPR_RS_ARGS
    FLPR_CHECK_LAUNCH();
    src = dst;
  }
  hipLaunchKernelGGL(k_channel_sums, dim3((unsigned)nch), dim3(RS_THREADS), 0, st, grid, cells_c,
                     sums);
  FLPR_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
