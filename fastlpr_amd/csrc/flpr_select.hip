// Bandwidth selection on the device: Hutchinson degrees of freedom, GCV
// scores and the 1-SE rule, from the per-bandwidth sums the solve and the
// quadratic form leave behind.  Running it on the device lets the final
// interpolation at the selected bandwidth be queued behind it, so one fit
// needs no host round trip between the solve and the result.
//
// Reference dataflow restated here (fastLPR_py/src/fastlpr/...):
//   regression.py:1276-1354  tr(H) ~ n (p'phat)/(p'p), NaN -> n, clip [0, n],
//                            penalty 1/max(pdof, .01)^2, mean / std(ddof=1)
//   regression.py:378-421    GCV = mse * penalty, stderr = mse * penalty_std,
//                            aggregation first / average / sum over columns
//   api.py:73-164            1-SE rule, larger sum(h) preferred, ties broken
//                            by the smaller score
// The arithmetic follows NumPy's operation order (sequential sums over the
// probe axis, NumPy's pairwise sum over columns) without contraction, so the
// scores and the selected index equal a NumPy evaluation of the same sums.
#include "flpr_common.h"

namespace {

constexpr int SEL_THREADS = 256;

__device__ inline double clip(double v, double lo, double hi) {
  // np.clip: NaN stays NaN
  return (v != v) ? v : (v < lo ? lo : (v > hi ? hi : v));
}

__device__ inline double np_maximum(double a, double b) {
  // np.maximum propagates NaN
  if (a != a) return a;
  if (b != b) return b;
  return a > b ? a : b;
}

// NumPy's pairwise summation of f(0..n-1) (n <= 128: eight strided partial
// sums, combined pairwise, then the remainder); n < 8 is a plain loop.  The
// terms are produced on demand, so no per-thread array (scratch memory) is
// needed.
template <typename F>
__device__ inline double np_sum(F f, int n) {
#pragma clang fp contract(off)
  if (n < 8) {
    double r = -0.0;
    for (int i = 0; i < n; ++i) r += f(i);
    return r;
  }
  double r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = f(j);
  int i = 8;
  for (; i < n - (n % 8); i += 8)
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] += f(i + j);
  double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; ++i) res += f(i);
  return res;
}

// Scores: one block per bandwidth; lane k computes the terms of probe k
// (trace, pdof, penalty) into LDS, then lane 0 forms every sum sequentially in
// probe order (as NumPy over axis 0), so the values are NumPy's bit for bit.
//   red     (dh, nval) dots [y columns, probes] then (dh, ny) quadratic forms
//   ysq     (ny) sum_i y_ic^2; pp (K) p'p per probe
//   out     scores, errs, dof, dof_stderr (dh each)
// Batch (blockIdx.z = problem): items give the point count; s[0] = red,
// s[1] = ysq, s[2] = pp, s[3] = out.
__global__ __launch_bounds__(64) void k_select_scores(
    const double* __restrict__ red, int dh, int nval, int ny, int K,
    const double* __restrict__ ysq, const double* __restrict__ pp, double n, int cplx, int agg,
    double* __restrict__ out, FlprBz bz) {
#pragma clang fp contract(off)
  if (bz.items) n = (double)bz.items[blockIdx.z].n;
  red += blockIdx.z * bz.s[0];
  ysq += blockIdx.z * bz.s[1];
  pp += blockIdx.z * bz.s[2];
  out += blockIdx.z * bz.s[3];
  __shared__ double tr[FLPR_MAX_CHANNELS], pdv[FLPR_MAX_CHANNELS], penv[FLPR_MAX_CHANNELS];
  const int b = blockIdx.x;
  const double* dot = red + (int64_t)b * nval;
  const double* quad = red + (int64_t)dh * nval + (int64_t)b * ny;
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    double t = n * (dot[ny + k] / pp[k]);
    if (t != t) t = n;
    t = clip(t, 0.0, n);
    tr[k] = t;
    double pd = 1.0 - t / n;
    if (cplx) pd = 2.0 * pd - 1.0;
    pd = clip(pd, 0.0, 1.0);
    pdv[k] = pd;
    const double m = np_maximum(pd, 0.01);
    penv[k] = 1.0 / (m * m);
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int ncol = cplx ? ny / 2 : ny;
  // --- Hutchinson over the K probes (sequential sums, as NumPy over axis 0)
  double s = 0.0;
  for (int k = 0; k < K; ++k) s += tr[k];
  double dof = s / (double)K;
  double dse = 0.0, pen_std = 0.0, pen_mean = 0.0;
  const bool many = K > 1;
  if (many) {
    double ss = 0.0;
    for (int k = 0; k < K; ++k) {
      const double e = tr[k] - dof;
      ss += e * e;
    }
    dse = sqrt(ss / (double)(K - 1)) / sqrt((double)K);
    double pcs = 0.0, ps = 0.0;
    for (int k = 0; k < K; ++k) {
      pcs += pdv[k];
      ps += penv[k];
    }
    pen_mean = ps / (double)K;
    double es = 0.0;
    for (int k = 0; k < K; ++k) {
      const double e = penv[k] - pen_mean;
      es += e * e;
    }
    pen_std = sqrt(es / (double)(K - 1));
    if (cplx) dof = n * (1.0 - pcs / (double)K);
  }
  // --- residual sums per real column, complex columns paired
  double pen;
  if (many) {
    pen = pen_mean;
  } else {
    const double pd = clip(1.0 - dof / n, 0.0, 1.0);
    const double m = np_maximum(pd, 0.01);
    pen = 1.0 / (m * m);
  }
  auto mse = [&](int c) {
    double rss;
    if (cplx) {
      const double r0 = (ysq[2 * c] - 2.0 * dot[2 * c]) + quad[2 * c];
      const double r1 = (ysq[2 * c + 1] - 2.0 * dot[2 * c + 1]) + quad[2 * c + 1];
      rss = r0 + r1;
    } else {
      rss = (ysq[c] - 2.0 * dot[c]) + quad[c];
    }
    return rss / n;
  };
  auto gcv = [&](int c) { return mse(c) * pen; };
  double score, m;
  if (agg == 0) {
    score = gcv(0);
    m = mse(0);
  } else if (agg == 1) {
    score = np_sum(gcv, ncol) / (double)ncol;
    m = np_sum(mse, ncol) / (double)ncol;
  } else {
    score = np_sum(gcv, ncol);
    m = np_sum(mse, ncol);
  }
  out[b] = score;
  out[dh + b] = (pen_std > 0.0) ? m * pen_std : 0.0;
  out[2 * dh + b] = dof;
  out[3 * dh + b] = many ? dse : 0.0;
}

constexpr int SEL_LDS = 2048;  // bandwidths held in LDS for the selection

// (value, index) minimum over the wave, the lower index on ties: the first
// minimum of a sequential scan.
__device__ inline void wave_argmin(double& v, int& i) {
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off);
    const int oi = __shfl_xor(i, off);
    if (ov < v || (ov == v && oi < i)) {
      v = ov;
      i = oi;
    }
  }
}

// The selection of k_select_pick for dh <= 64 with one bandwidth per lane:
// the same choices (first NaN or first minimum for idmin, first minimum over
// the finite scores, the 1-SE rule preferring the larger sum(h), then the
// first lowest score) from wave reductions instead of a serial walk.
__device__ inline void select_pick_wave(int dh, const double* __restrict__ hsum_g, double dstd,
                                        double* __restrict__ out, int32_t* __restrict__ sel_out) {
  const int b = threadIdx.x;
  const bool on = b < dh;
  const double sc = on ? out[b] : 0.0;
  const double er = on ? out[dh + b] : 0.0;
  const double hs = on ? hsum_g[b] : 0.0;
  const bool fin = on && isfinite(sc);
  const uint64_t nan_mask = __ballot(on && sc != sc);
  const int nvalid = __popcll(__ballot(fin));
  // idmin: np.argmin, the first NaN when there is one
  int idmin;
  if (nan_mask) {
    idmin = __ffsll((unsigned long long)nan_mask) - 1;
  } else {
    double v = on ? sc : INFINITY;
    int i = on ? b : 1 << 30;
    wave_argmin(v, i);
    idmin = (i >= dh) ? 0 : i;   // all scores +inf: the first one
  }
  int sel = 0;
  if (nvalid > 0) {
    double v = fin ? sc : INFINITY;
    int best = fin ? b : 1 << 30;
    wave_argmin(v, best);
    const double thr = __shfl(sc, best) + dstd * __shfl(er, best);
    const bool cand = fin && sc <= thr;
    double smax = cand ? hs : -INFINITY;
    for (int off = 32; off > 0; off >>= 1) smax = fmax(smax, __shfl_xor(smax, off));
    const bool c2 = cand && hs == smax;
    double v2 = c2 ? sc : INFINITY;
    int i2 = c2 ? b : 1 << 30;
    wave_argmin(v2, i2);
    sel = (i2 >= dh) ? best : i2;
  }
  if (b == 0) {
    out[4 * dh] = (double)sel;
    out[4 * dh + 1] = (double)idmin;
    out[4 * dh + 2] = (double)nvalid;
    if (sel_out) sel_out[0] = sel;
  }
}

// The selection over the scores (one block; thread 0 walks the bandwidths in
// LDS when they fit): out[4 dh] = sel, out[4 dh + 1] = idmin, out[4 dh + 2] =
// number of finite scores.
// Batch: s[3] = out; one selected index per problem.
__global__ __launch_bounds__(SEL_THREADS) void k_select_pick(int dh,
                                                             const double* __restrict__ hsum_g,
                                                             double dstd,
                                                             double* __restrict__ out,
                                                             int32_t* __restrict__ sel_out,
                                                             FlprBz bz) {
  __shared__ double lsc[SEL_LDS], ler[SEL_LDS], lhs[SEL_LDS];
  out += blockIdx.z * bz.s[3];
  if (sel_out) sel_out += blockIdx.z;
  if (dh <= 64) {
    if (threadIdx.x < 64) select_pick_wave(dh, hsum_g, dstd, out, sel_out);
    return;
  }
  const bool in_lds = dh <= SEL_LDS;
  if (in_lds) {
    for (int b = threadIdx.x; b < dh; b += blockDim.x) {
      lsc[b] = out[b];
      ler[b] = out[dh + b];
      lhs[b] = hsum_g[b];
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double* scores = in_lds ? lsc : out;
  const double* errs = in_lds ? ler : out + dh;
  const double* hsum = in_lds ? lhs : hsum_g;
  // --- idmin: np.argmin (the first NaN when there is one)
  int idmin = 0;
  for (int b = 0; b < dh; ++b) {
    const double v = scores[b], w = scores[idmin];
    if (v != v) {
      idmin = b;
      break;
    }
    if (v < w) idmin = b;
  }
  // --- 1-SE rule over the finite scores, prefer the larger sum(h)
  int best = -1, nvalid = 0;
  for (int b = 0; b < dh; ++b) {
    if (!isfinite(scores[b])) continue;
    ++nvalid;
    if (best < 0 || scores[b] < scores[best]) best = b;
  }
  int sel = 0;
  if (best >= 0) {
    const double thr = scores[best] + dstd * errs[best];
    double smax = -INFINITY;
    for (int b = 0; b < dh; ++b)
      if (isfinite(scores[b]) && scores[b] <= thr && hsum[b] > smax) smax = hsum[b];
    sel = -1;
    for (int b = 0; b < dh; ++b) {
      if (!isfinite(scores[b]) || !(scores[b] <= thr) || hsum[b] != smax) continue;
      if (sel < 0 || scores[b] < scores[sel]) sel = b;
    }
    if (sel < 0) sel = best;
  }
  out[4 * dh] = (double)sel;
  out[4 * dh + 1] = (double)idmin;
  out[4 * dh + 2] = (double)nvalid;
  if (sel_out) sel_out[0] = sel;
}

}  // namespace

extern "C" {

int flpr_select(const double* red, int32_t dh, int32_t nval, int32_t ny, int32_t nprobe,
                const double* ysq, const double* pp, const double* hsum, int64_t n,
                int32_t cplx, int32_t agg, double dstd, double* out, int32_t* sel,
                void* stream) {
  FLPR_CLEAR_ERROR();
  if (!red || !ysq || !pp || !hsum || !out || dh < 1 || ny < 1 || nprobe < 1 ||
      nprobe > FLPR_MAX_CHANNELS || ny > FLPR_MAX_CHANNELS || nval != ny + nprobe || n < 1 ||
      agg < 0 || agg > 2 || (cplx && (ny % 2)))
    return -1;
  hipLaunchKernelGGL(k_select_scores, dim3((unsigned)dh), dim3(64), 0, as_stream(stream), red, dh,
                     nval, ny, nprobe, ysq, pp, (double)n, cplx, agg, out, FlprBz{});
  FLPR_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_select_pick, dim3(1), dim3(SEL_THREADS), 0, as_stream(stream), dh, hsum,
                     dstd, out, sel, FlprBz{});
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_batch_select(const FlprBatchItem* items, int32_t P, const double* red,
                      int64_t red_stride, int32_t dh, int32_t nval, int32_t ny, int32_t nprobe,
                      const double* ysq, const double* pp, const double* hsum, int32_t cplx,
                      int32_t agg, double dstd, double* out, int64_t out_stride, int32_t* sel,
                      void* stream) {
  FLPR_CLEAR_ERROR();
  if (!items || P < 1 || P > 65535 || !red || !ysq || !pp || !hsum || !out || dh < 1 || ny < 1 ||
      nprobe < 1 || nprobe > FLPR_MAX_CHANNELS || ny > FLPR_MAX_CHANNELS || nval != ny + nprobe ||
      agg < 0 || agg > 2 || (cplx && (ny % 2)) || out_stride < 4 * dh + 3)
    return -1;
  FlprBz bz{};
  bz.items = items;
  bz.s[0] = red_stride;
  bz.s[1] = ny;
  bz.s[2] = nprobe;
  bz.s[3] = out_stride;
  hipLaunchKernelGGL(k_select_scores, dim3((unsigned)dh, 1, (unsigned)P), dim3(64), 0,
                     as_stream(stream), red, dh, nval, ny, nprobe, ysq, pp, 1.0, cplx, agg, out,
                     bz);
  FLPR_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_select_pick, dim3(1, 1, (unsigned)P), dim3(SEL_THREADS), 0,
                     as_stream(stream), dh, hsum, dstd, out, sel, bz);
  FLPR_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
