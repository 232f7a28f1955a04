// Grid-side kernels: kernel-moment taps, the separable banded convolution on
// MFMA, the global-regularisation reduction and the per-node solves.
//
// The reference convolves every spread grid with every kernel-moment grid
// K_h(u) u1^a u2^b ... through zero-padded FFTs (convolution.py:747-794,
// 904-951).  Because the padded circular convolution never wraps on the data
// support (L >= 2N) and the Gaussian kernel is a product of 1-D factors, the
// same numbers are a sequence of 1-D banded Toeplitz products along each axis
// of the unwrapped spread grid.  Each pass is a batched GEMM in which one
// operand is the Toeplitz band, generated in LDS from a 1-D tap strip; the
// GEMM runs on v_mfma_f64_16x16x4f64 (float64 form) or on bf16 MFMA with
// exact three-term splits of the float32 operands (float32 form).
#include <type_traits>

#include "flpr_common.h"

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

// v_mfma_f64_16x16x4f64: lane l holds A[l & 15][l >> 4] and B[l >> 4][l & 15];
// result register r of lane l is D[(l >> 4) + 4 r][l & 15] (checked on gfx950
// with a probe of known operands).
struct Mfma64 {
  typedef v4d acc;
  __device__ static inline acc run(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

// ---------------------------------------------------------------------------
// taps
// ---------------------------------------------------------------------------

constexpr int TAP_THREADS = 256;

// Batch: items give the geometry and the point count (the norm); s[0] =
// tap64 and tap32.
__global__ __launch_bounds__(TAP_THREADS) void k_taps(FlprGeom g, const double* __restrict__ h,
                                                      int dh, int npow, int nmax, double norm,
                                                      double* __restrict__ tap64,
                                                      float* __restrict__ tap32, FlprBz bz) {
#pragma clang fp contract(off)
  // the per-axis fields in registers (a copied FlprGeom indexed by a runtime
  // axis would live in scratch memory)
  int ngr[FLPR_MAX_DIM];
  double glo[FLPR_MAX_DIM], ghi[FLPR_MAX_DIM];
  {
    const FlprGeom* src = bz.items ? &bz.items[blockIdx.z].g : &g;
#pragma unroll
    for (int k = 0; k < FLPR_MAX_DIM; ++k) {
      ngr[k] = src->n_grid[k];
      glo[k] = src->xmin_s[k];
      ghi[k] = src->xmax_s[k];
    }
  }
  if (bz.items) {
    norm = (double)bz.items[blockIdx.z].n;
    tap64 += blockIdx.z * bz.s[0];
    if (tap32) tap32 += blockIdx.z * bz.s[0];
  }
  const int64_t total = (int64_t)dh * g.d * nmax;
  const double inv_sqrt_2pi = 1.0 / sqrt(2.0 * M_PI);
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int t = (int)(idx % nmax);
    const int64_t r = idx / nmax;
    const int k = (int)(r % g.d);
    const int b = (int)(r / g.d);
    const int N = k == 0 ? ngr[0] : (k == 1 ? ngr[1] : ngr[2]);
    const int64_t o = ((int64_t)(b * g.d + k) * npow) * nmax + t;
    if (t >= N) {
      for (int a = 0; a < npow; ++a) {
        tap64[o + (int64_t)a * nmax] = 0.0;
        if (tap32) tap32[o + (int64_t)a * nmax] = 0.0f;
      }
      continue;
    }
    // numpy linspace(lo, hi, N): t*step + lo, last element exactly hi
    const double lo = k == 0 ? glo[0] : (k == 1 ? glo[1] : glo[2]);
    const double hi = k == 0 ? ghi[0] : (k == 1 ? ghi[1] : ghi[2]);
    double xg = lo;
    if (N > 1) {
      const double step = (hi - lo) / (double)(N - 1);
      xg = (double)t * step;
      xg = xg + lo;
      if (t == N - 1) xg = hi;
    }
    const double u = xg / h[b * g.d + k];  // convolution.py:443-447 (x / h)
    double e = exp(-0.5 * (u * u));        // kernel.py:351-356
    if (k == 0) {
      double ph = 1.0;
      for (int kk = 0; kk < g.d; ++kk) ph *= h[b * g.d + kk];
      e = e * inv_sqrt_2pi;
      e = e / ph;    // kernel.py:301-304 (divide by det h)
      e = e / norm;  // nufftn_type1.py:709-711 (divide by the number of points)
    }
    double pw = 1.0;
    for (int a = 0; a < npow; ++a) {
      const double v = e * pw;
      tap64[o + (int64_t)a * nmax] = v;
      if (tap32) tap32[o + (int64_t)a * nmax] = (float)v;
      pw = pw * u;  // design_matrix_optimized.py:268-274 (cumulative powers)
    }
  }
}

// ---------------------------------------------------------------------------
// banded Toeplitz GEMM on MFMA
// ---------------------------------------------------------------------------

constexpr int CT = 256;        // 4 waves
constexpr int LT = 64;         // Toeplitz extent of a block tile (four 16-wide blocks)
constexpr int NDB = 1;         // dense 16-blocks of one wave
constexpr int WD = NDB * 16;   // dense extent of one wave
constexpr int LD = 4 * WD;     // dense extent of a block tile
constexpr int NQ = LT / 16;    // Toeplitz 16-blocks, shared by the four waves

// Banded Toeplitz GEMM of the float64 form.  A block owns a 64-wide stretch
// of the convolved (Toeplitz) axis and a 64-wide stretch of the other (dense)
// axis; each of its four waves owns 16 of the dense extent and all 64
// Toeplitz outputs, as 4 MFMA 16x16 blocks.  The dense operand of chunk k+1
// is fetched into registers while chunk k is multiplied from LDS
// (double-buffered, one barrier per chunk; K chunks of 16).  The Toeplitz
// operand is generated from a 1-D strip of taps: element (i, s) is
// tap[i + base - s].
//
// Dense rows (ROW pass) or columns (COL pass) past the job only feed outputs
// that are never stored, and inputs past the chunk's band end are replaced by
// zero, so the loads carry no bounds branches; the input buffer must stay
// readable LD*K + KC*inner + LD elements past each job's input (the caller
// pads its allocations).
//
// Block schedule: with a block table (blocks != nullptr) block b runs tile
// blocks[2b+1] of job blocks[2b] (a negative job is an idle block); without
// one, the grid is (tiles of the largest job, jobs).
template <int MODE>
__global__ __launch_bounds__(CT) void k_conv(const double* __restrict__ in,
                                             double* __restrict__ out,
                                             const double* __restrict__ tap,
                                             const int64_t* __restrict__ jobs,
                                             const int32_t* __restrict__ bands, int K, int nout,
                                             int inner, int base, int ntd, int ntt,
                                             const int32_t* __restrict__ blocks) {
  typedef double T;
  constexpr int KC = 16;
  // LDS pitch of the dense tile: a K step reads 16 rows x 4 K (ROW pass) or
  // 4 K rows x 16 columns (COL pass); the pads spread each half-wave over 32 banks
  constexpr int PD = (MODE == 0) ? KC + 2 : LD + 16;
  static_assert(CT % LD == 0 && CT % KC == 0, "staging layout");
  constexpr int DENSE = (MODE == 0) ? LD * PD : KC * PD;
  constexpr int NSTRIP = LT + KC - 1;
  constexpr int PER = LD * KC / CT;   // dense elements staged per thread and chunk
  constexpr int RSTEP = CT / KC;      // ROW pass: row stride between a thread's elements
  constexpr int KSTEP = CT / LD;      // COL pass: K-row stride between a thread's elements
  __shared__ T dense[2][DENSE];
  __shared__ T strip[2][NSTRIP + 1];
  int job = blockIdx.y, bx = blockIdx.x;
  if (blocks) {
    job = blocks[2 * blockIdx.x];
    bx = blocks[2 * blockIdx.x + 1];
    if (job < 0) return;
  }
  const int64_t* jb = jobs + 8 * (int64_t)job;
  const int64_t in_off = jb[0], out_off = jb[1], tap_off = jb[2];
  const int band = (int)jb[3], ro = (int)jb[4];
  const int t_lo = bands[2 * band], t_hi = bands[2 * band + 1];
  const int td = bx % ntd;
  bx /= ntd;
  const int tt = bx % ntt;
  const int o = bx / ntt;
  const int d0 = td * LD, i0 = tt * LT;
  int ndense, ntoe;
  if (MODE == 0) {
    ndense = ro;
    ntoe = nout;
  } else {
    ndense = inner;
    ntoe = nout;
    if (o >= ro) return;
  }
  if (d0 >= ndense || i0 >= ntoe) return;
  const int s_begin = max(i0 + base - t_hi, 0);
  const int s_end = min(i0 + LT - 1 + base - t_lo, K - 1);
  const int nchunks = (s_end >= s_begin) ? (s_end - s_begin) / KC + 1 : 0;
  const T* src = in + in_off + (MODE == 0 ? 0 : (int64_t)o * K * inner);
  T* dst = out + out_off + (MODE == 0 ? 0 : (int64_t)o * nout * inner);
  const T* tp = tap + tap_off;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int dw = wave * WD;  // the wave's offset along the dense extent
  // staging: ROW pass thread owns K column kc0 of rows r0 + i*RSTEP;
  // COL pass thread owns dense column cc0 of K rows k0 + i*KSTEP
  const int kc0 = tid % KC, r0 = tid / KC;
  const int cc0 = tid % LD, k0 = tid / LD;
  const T* sbase = (MODE == 0) ? src + (int64_t)(d0 + r0) * K + kc0
                               : src + (int64_t)k0 * inner + d0 + cc0;
  const int64_t sstep = (MODE == 0) ? (int64_t)RSTEP * K : (int64_t)KSTEP * inner;

  Mfma64::acc acc[NQ][NDB];
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int db = 0; db < NDB; ++db)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[q][db][j] = T(0);

  // Two register sets of prefetched raw values keep two chunks in flight;
  // the zeroing past the band end is applied when a set is stored to LDS, so
  // the loads are not waited for before the MFMAs.
  T preA[PER], preB[PER];
  T pstA = T(0), pstB = T(0);
  auto gload = [&](T (&pre)[PER], T& pst, int s0) {
    const T* p = sbase + (MODE == 0 ? (int64_t)s0 : (int64_t)s0 * inner);
#pragma unroll
    for (int i = 0; i < PER; ++i) pre[i] = p[i * sstep];
    if (tid < NSTRIP) {
      const int t = i0 + base - s0 - (KC - 1) + tid;
      pst = tp[max(min(t, t_hi), 0)];  // any in-strip address; masked when stored
    }
  };
  auto sstore = [&](const T (&pre)[PER], T pst, int s0, int buf) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      if (MODE == 0) dense[buf][(r0 + i * RSTEP) * PD + kc0] = (s0 + kc0 <= s_end) ? pre[i] : T(0);
      else dense[buf][(k0 + i * KSTEP) * PD + cc0] = (s0 + k0 + i * KSTEP <= s_end) ? pre[i] : T(0);
    }
    if (tid < NSTRIP) {
      const int t = i0 + base - s0 - (KC - 1) + tid;
      strip[buf][tid] = (t >= t_lo && t <= t_hi) ? pst : T(0);
    }
  };
  // Multiply chunk ch (staged in LDS buffer buf) and fold it.  All operands
  // of the chunk are read into registers first.  The Toeplitz operand of K
  // step st and block q is strip[16 q + (lane & 15) - 4 st - (lane >> 4) + KC - 1],
  // which depends on 4 q - st only: NTV distinct values per lane per chunk.
  constexpr int NST = KC / 4;
  constexpr int NTV = 4 * (NQ - 1) + NST;
  auto compute = [&](int ch, int buf) {
    const T* dn = dense[buf];
    const T* sp = strip[buf];
    const int sc0 = s_begin + ch * KC;  // absolute input index of the chunk's first K
    // Toeplitz blocks that meet the band anywhere in this chunk: a block of
    // outputs starting at i needs inputs in [i + base - t_hi, i + 15 + base - t_lo].
    // A scalar per chunk; MFMAs on the band's zero taps add zeros.
    int live = 0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int ib = i0 + q * 16;
      if (!(sc0 + KC - 1 < ib + base - t_hi || sc0 > ib + 15 + base - t_lo)) live |= 1 << q;
    }
    live = __builtin_amdgcn_readfirstlane(live);
    if (live != 0) {
      const int sb = (lane & 15) - (lane >> 4) + KC - 1 - 4 * (NST - 1);
      T tv[NTV], dv[NST][NDB];
#pragma unroll
      for (int m = 0; m < NTV; ++m) tv[m] = sp[sb + 4 * m];  // m = 4 q - st + NST - 1
#pragma unroll
      for (int st = 0; st < NST; ++st) {
        const int kl = st * 4 + (lane >> 4);
#pragma unroll
        for (int db = 0; db < NDB; ++db) {
          const int dd = dw + db * 16 + (lane & 15);
          dv[st][db] = (MODE == 0) ? dn[dd * PD + kl] : dn[kl * PD + dd];
        }
      }
      // one Toeplitz block q over the whole chunk
      auto mac_q = [&](auto qc) {
        constexpr int q = decltype(qc)::value;
#pragma unroll
        for (int st = 0; st < NST; ++st) {
          const T tq = tv[4 * q - st + NST - 1];
#pragma unroll
          for (int db = 0; db < NDB; ++db)
            acc[q][db] = (MODE == 0) ? Mfma64::run(dv[st][db], tq, acc[q][db])
                                     : Mfma64::run(tq, dv[st][db], acc[q][db]);
        }
      };
      if (live == (1 << NQ) - 1) {
        // all four blocks, interleaved per K step
#pragma unroll
        for (int st = 0; st < NST; ++st)
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            const T tq = tv[4 * q - st + NST - 1];
#pragma unroll
            for (int db = 0; db < NDB; ++db)
              acc[q][db] = (MODE == 0) ? Mfma64::run(dv[st][db], tq, acc[q][db])
                                       : Mfma64::run(tq, dv[st][db], acc[q][db]);
          }
      } else {
        // the live blocks are contiguous (the band is): one scalar test per block
        if (live & 1) mac_q(std::integral_constant<int, 0>{});
        if (live & 2) mac_q(std::integral_constant<int, 1>{});
        if (live & 4) mac_q(std::integral_constant<int, 2>{});
        if (live & 8) mac_q(std::integral_constant<int, 3>{});
      }
    }
  };

  // chunk c lives in LDS buffer c & 1 and, before that, in register set A
  // (odd c) or B (even c, c >= 2)
  if (nchunks > 0) {
    gload(preA, pstA, s_begin);
    sstore(preA, pstA, s_begin, 0);
    if (nchunks > 1) gload(preA, pstA, s_begin + KC);
    if (nchunks > 2) gload(preB, pstB, s_begin + 2 * KC);
  }
  __syncthreads();
  for (int ch = 0; ch < nchunks; ch += 2) {
    compute(ch, 0);
    if (ch + 1 < nchunks) {
      sstore(preA, pstA, s_begin + (ch + 1) * KC, 1);
      if (ch + 3 < nchunks) gload(preA, pstA, s_begin + (ch + 3) * KC);
    }
    __syncthreads();
    if (ch + 1 < nchunks) {
      compute(ch + 1, 1);
      if (ch + 2 < nchunks) {
        sstore(preB, pstB, s_begin + (ch + 2) * KC, 0);
        if (ch + 4 < nchunks) gload(preB, pstB, s_begin + (ch + 4) * KC);
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int db = 0; db < NDB; ++db)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        // MFMA result (row(lane, j), lane & 15): rows follow the first operand
        const int rr = Mfma64::row(lane, j), cc = lane & 15;
        int dd, ii;
        if (MODE == 0) {
          dd = d0 + dw + db * 16 + rr;
          ii = i0 + q * 16 + cc;
        } else {
          ii = i0 + q * 16 + rr;
          dd = d0 + dw + db * 16 + cc;
        }
        if (dd < ndense && ii < ntoe) {
          if (MODE == 0) dst[(int64_t)dd * nout + ii] = acc[q][db][j];
          else dst[(int64_t)ii * inner + dd] = acc[q][db][j];
        }
      }
}

// ---------------------------------------------------------------------------
// float32 form: the same banded Toeplitz GEMM on bf16 MFMA with three-term splits
// ---------------------------------------------------------------------------
//
// Every float32 operand x is written exactly as x = hi + mid + lo with three
// bfloat16 values (hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid);
// 3 x 8 significand bits hold the 24 of a float32).  A product of two split
// operands keeps the six terms whose levels add up to at most two
// (hi*hi, hi*mid, mid*hi, hi*lo, lo*hi, mid*mid); the three dropped ones are
// below 2^-26 of the product.  Each term is an exact float32 product summed
// by v_mfma_f32_16x16x32_bf16 (16 cycles for 16x16x32, against 32 cycles for
// 16x16x4 in float32), so the six terms cost 2.67x less MFMA time than the
// float32 product at float32 accuracy.  The products accumulate in the
// float32 MFMA accumulators over the tile's whole K range (float32 GEMM
// accuracy; a float64 fold per chunk measured 2.5x smaller errors and 11 %
// more time).
//
// Operand layout of v_mfma_f32_16x16x32_bf16 (checked on gfx950 with known
// operands): lane l holds A[l & 15][8 (l >> 4) + j] and B[8 (l >> 4) + j][l & 15],
// j = 0..7; result register r of lane l is D[4 (l >> 4) + r][l & 15].
//
// The Toeplitz operand of a chunk comes from a reversed tap window
// R[v] = tap[t_top - v], v in [0, 95), t_top = i0 + base - s0 + 63: lane
// (i, g) of output block q needs R[63 - 16 q - i + 8 g + j], 8 consecutive
// values at an alignment set by i.  LDS holds eight copies of the window
// shifted by r = 0..7 (copy r, element m = R[m + r]), so every lane reads one
// aligned 16-byte unit of copy 7 - (i & 7).  The copies are cut from a table
// built once per fit (flpr_conv_strips): per tap row, the reversed, band-masked,
// split row in eight shifted copies, so that the per-chunk staging is aligned
// 16-byte loads.

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));   // one 16-byte unit

constexpr int SX_LD = 128;          // dense extent of a block tile (2 x 16 per wave)
constexpr int SX_KC = 32;           // K chunk
constexpr int SX_NDB = 2;           // dense 16-blocks per wave
constexpr int SX_COPY = 88;         // elements of one shifted window copy in LDS (11 units)
constexpr int SX_UNITS = SX_COPY / 8;
constexpr int SX_PITCH = 12;        // LDS units per copy: with 12 the fragment reads of the
                                    // eight copies measure conflict-free on gfx950 (11: 1.25x)
constexpr int SX_MAX_POWERS = 3;   // tap rows one job of the float32 form may apply
constexpr int SX_TOP_PAD = 62;      // reversed table index v = (nmax + 62) - t

__host__ __device__ inline int64_t split_strip_len(int nmax) {
  // v runs over [0, nmax + 62 + 95) for the windows a kernel can ask for
  return ((int64_t)nmax + SX_TOP_PAD + 95 + 7) / 8 * 8;
}

__device__ inline void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = (__bf16)x;
  const float r1 = x - (float)h;
  m = (__bf16)r1;
  const float r2 = r1 - (float)m;
  l = (__bf16)r2;
}

// The same split for 8 values at once, two per conversion (v_cvt_pk_bf16_f32)
// and the exact bf16 -> f32 widening done with shifts and masks.
__device__ inline void split3x8(const float (&x)[8], u32x4& hv, u32x4& mv, u32x4& lv) {
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    float a = x[2 * p], b = x[2 * p + 1];
    uint32_t w[3];
#pragma unroll
    for (int lev = 0; lev < 3; ++lev) {
      typedef float v2f __attribute__((ext_vector_type(2)));
      typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
      v2f ab = {a, b};
      const uint32_t pk = __builtin_bit_cast(uint32_t, __builtin_convertvector(ab, v2bf));
      w[lev] = pk;
      if (lev < 2) {
        a -= __builtin_bit_cast(float, pk << 16);
        b -= __builtin_bit_cast(float, pk & 0xffff0000u);
      }
    }
    hv[p] = w[0];
    mv[p] = w[1];
    lv[p] = w[2];
  }
}

// Table row A (tap row (b*d + k)*npow + a, band A / npow), plane p, copy r:
// strips[((A*3 + p)*8 + r)*V + e] = split plane p of R[e + r], with
// R[v] = tap[nmax + 62 - v] inside the band and zero elsewhere.
__global__ __launch_bounds__(256) void k_split_taps(const float* __restrict__ tap,
                                                    const int32_t* __restrict__ bands, int ntap,
                                                    int npow, int nmax, int64_t V,
                                                    uint16_t* __restrict__ strips) {
  const int64_t units = V / 8;
  const int64_t total = (int64_t)ntap * 8 * units;
  const int top = nmax + SX_TOP_PAD;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t u = idx % units;
    const int r = (int)((idx / units) % 8);
    const int A = (int)(idx / (units * 8));
    const int band = A / npow;
    const int t_lo = max(bands[2 * band], 0), t_hi = min(bands[2 * band + 1], nmax - 1);
    v8bf hv, mv, lv;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int t = top - (int)(u * 8 + j + r);
      const float x = (t >= t_lo && t <= t_hi) ? tap[(int64_t)A * nmax + t] : 0.0f;
      __bf16 h, m, l;
      split3(x, h, m, l);
      hv[j] = h;
      mv[j] = m;
      lv[j] = l;
    }
    u32x4* dst = reinterpret_cast<u32x4*>(strips);
    const int64_t row = (int64_t)A * 3 * 8 + r;
    dst[((row + 0 * 8) * V) / 8 + u] = __builtin_bit_cast(u32x4, hv);
    dst[((row + 1 * 8) * V) / 8 + u] = __builtin_bit_cast(u32x4, mv);
    dst[((row + 2 * 8) * V) / 8 + u] = __builtin_bit_cast(u32x4, lv);
  }
}

__device__ inline v4f mfma_bf(const u32x4& a, const u32x4& b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a),
                                                 __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
}

// The six split products of X (first operand) and Y, added into c.
__device__ inline v4f mfma_split6(const u32x4 (&x)[3], const u32x4 (&y)[3], v4f c) {
  c = mfma_bf(x[1], y[1], c);
  c = mfma_bf(x[0], y[2], c);
  c = mfma_bf(x[2], y[0], c);
  c = mfma_bf(x[0], y[1], c);
  c = mfma_bf(x[1], y[0], c);
  c = mfma_bf(x[0], y[0], c);
  return c;
}

// Banded Toeplitz GEMM of the float32 form (jobs, bands, tiles and the block
// schedule as in k_conv).  A block owns 64 Toeplitz outputs x 128 dense; wave
// w owns dense [32 w, 32 w + 32) and all 64 outputs (4 x 2 MFMA blocks).  The
// dense operand is not shared between waves, so each wave loads its own
// fragments straight into registers (lane (n, g) of dense block db holds the
// 8 K values 8 g .. 8 g + 7 of dense index 16 db + n, two chunks ahead) and
// splits them there; only the Toeplitz windows, which all four waves read,
// go through LDS (double-buffered, one barrier per chunk).
// Powers: a job applies np <= NP consecutive tap rows (the polynomial powers
// of one axis) to the same input, writing power p's output at
// out_off + p * out_stride; every dense fragment is loaded and split once for
// all of them, and each power keeps its own accumulators and LDS window.
template <int MODE, int NP>
__global__ __launch_bounds__(CT) void k_conv_split(const float* __restrict__ in,
                                                   float* __restrict__ out,
                                                   const uint16_t* __restrict__ strips, int nmax,
                                                   const int64_t* __restrict__ jobs,
                                                   const int32_t* __restrict__ bands, int K,
                                                   int nout, int inner, int base, int ntd, int ntt,
                                                   const int32_t* __restrict__ blocks) {
  constexpr int KC = SX_KC;
  __shared__ u32x4 strip[2][NP][3][8][SX_PITCH];
  int job = blockIdx.y, bx = blockIdx.x;
  if (blocks) {
    job = blocks[2 * blockIdx.x];
    bx = blocks[2 * blockIdx.x + 1];
    if (job < 0) return;
  }
  const int64_t* jb = jobs + 8 * (int64_t)job;
  const int64_t in_off = jb[0], out_off = jb[1], tap_off = jb[2];
  const int band = (int)jb[3], ro = (int)jb[4];
  const int jnp = min(max((int)jb[5], 1), NP);
  const int64_t ostride = jb[6];
  const int t_lo = bands[2 * band], t_hi = bands[2 * band + 1];
  const int td = bx % ntd;
  bx /= ntd;
  const int tt = bx % ntt;
  const int o = bx / ntt;
  const int d0 = td * SX_LD, i0 = tt * LT;
  const int ndense = (MODE == 0) ? ro : inner;
  if (MODE == 1 && o >= ro) return;
  if (d0 >= ndense || i0 >= nout) return;
  const int s_begin = max(i0 + base - t_hi, 0);
  const int s_end = min(i0 + LT - 1 + base - t_lo, K - 1);
  const int nchunks = (s_end >= s_begin) ? (s_end - s_begin) / KC + 1 : 0;
  const float* src = in + in_off + (MODE == 0 ? 0 : (int64_t)o * K * inner);
  float* dst = out + out_off + (MODE == 0 ? 0 : (int64_t)o * nout * inner);
  const int64_t V = split_strip_len(nmax);
  const int64_t rowu = 3 * 8 * (V / 8);  // 16-byte units per table row
  const u32x4* sg = reinterpret_cast<const u32x4*>(strips) + (tap_off / nmax) * rowu;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int dw = wave * (SX_NDB * 16);
  const int n = lane & 15, g = lane >> 4;
  const int cr = 7 - (n & 7);

  // this lane's dense data: index d0 + dw + 16 db + n, K = s0 + 8 g + j.  Rows
  // (ROW pass) or columns (COL pass) past the job feed outputs that are never
  // stored; reads stay inside the caller's padding.
  const float* dsrc[SX_NDB];
#pragma unroll
  for (int db = 0; db < SX_NDB; ++db) {
    const int dd = d0 + dw + 16 * db + n;
    dsrc[db] = (MODE == 0) ? src + (int64_t)dd * K + 8 * g : src + (int64_t)(8 * g) * inner + dd;
  }
  // window staging: unit (power p, plane pl, copy r, unit mc) of a chunk's
  // windows, NP * 3 * 8 * SX_UNITS of them, thread tid owning units tid + CT u.
  // Copy r of the window of chunk c is table copy (vb + r) & 7 from unit
  // ((vb + r) >> 3) + mc, vb = nmax - 1 - i0 - base + s_begin + KC c; KC is a
  // multiple of 8, so the copy is the same for every chunk and the unit
  // advances by KC / 8 per chunk.
  constexpr int SITEMS = NP * 3 * 8 * SX_UNITS;
  constexpr int SPT = (SITEMS + CT - 1) / CT;
  int soff[SPT], sdst[SPT];
  {
    const int vb0 = nmax - 1 - i0 - base + s_begin;
#pragma unroll
    for (int u = 0; u < SPT; ++u) {
      const int idx = tid + CT * u;
      const int p = idx / (3 * 8 * SX_UNITS);
      const int rem = idx - p * (3 * 8 * SX_UNITS);
      const int pl = rem / (8 * SX_UNITS);
      const int r = (rem - pl * 8 * SX_UNITS) / SX_UNITS;
      const int mc = rem - pl * 8 * SX_UNITS - r * SX_UNITS;
      const bool on = idx < SITEMS && p < jnp;
      soff[u] = on ? (int)(p * rowu + ((int64_t)(pl * 8 + ((vb0 + r) & 7)) * V) / 8 +
                           ((vb0 + r) >> 3) + mc)
                   : 0;
      sdst[u] = on ? ((p * 3 + pl) * 8 + r) * SX_PITCH + mc : -1;
    }
  }

  // float32 MFMA accumulators over the whole K range of the tile
  v4f acc[NP][NQ][SX_NDB];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int db = 0; db < SX_NDB; ++db) acc[p][q][db] = v4f{0.0f, 0.0f, 0.0f, 0.0f};

  float preA[SX_NDB][8], preB[SX_NDB][8];
  u32x4 spA[SPT], spB[SPT];
  // the window units first: a wait for them then leaves the dense loads
  // issued after them in flight
  auto gload = [&](float (&pre)[SX_NDB][8], u32x4 (&sp)[SPT], int c) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < SPT; ++u)
      if (sdst[u] >= 0) sp[u] = sg[soff[u] + (KC / 8) * c];
    const int s0 = s_begin + c * KC;
#pragma unroll
    for (int db = 0; db < SX_NDB; ++db) {
      const float* p = (MODE == 0) ? dsrc[db] + s0 : dsrc[db] + (int64_t)s0 * inner;
#pragma unroll
      for (int j = 0; j < 8; ++j) pre[db][j] = (MODE == 0) ? p[j] : p[(int64_t)j * inner];
    }
  };
  auto sstore = [&](const u32x4 (&sp)[SPT], int buf) __attribute__((always_inline)) {
    u32x4* sb = &strip[buf][0][0][0][0];
#pragma unroll
    for (int u = 0; u < SPT; ++u)
      if (sdst[u] >= 0) sb[sdst[u]] = sp[u];
  };
  // Toeplitz fragments: block q of chunk ch + 1 equals block q - 2 of chunk ch
  // (i - s is what counts and the chunk advances by 32 = two blocks).  With
  // one power, a chunk reads blocks 0 and 1 into tvn and takes blocks 2 and 3
  // from the previous chunk's blocks 0 and 1 (tvo); the first chunk reads all
  // four.  With several powers every fragment is read from LDS where it is
  // used (holding two blocks per power would cost 24 registers a power).
  auto compute = [&](int ch, int buf, const float (&pre)[SX_NDB][8], u32x4 (&tvn)[2][3],
                     u32x4 (&tvo)[2][3]) __attribute__((always_inline)) {
    const int sc0 = s_begin + ch * KC;
    if constexpr (NP == 1) {
      if (ch == 0) {
#pragma unroll
        for (int q = 2; q < 4; ++q)
#pragma unroll
          for (int pl = 0; pl < 3; ++pl)
            tvo[q - 2][pl] = strip[buf][0][pl][cr][7 - 2 * q - (n >> 3) + g];
      }
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int pl = 0; pl < 3; ++pl)
          tvn[q][pl] = strip[buf][0][pl][cr][7 - 2 * q - (n >> 3) + g];
    }
    // dense fragments: only the last chunk reaches past the band end (inputs
    // there may be another slab's values or padding: zeroed)
    float x[SX_NDB][8];
#pragma unroll
    for (int db = 0; db < SX_NDB; ++db)
#pragma unroll
      for (int j = 0; j < 8; ++j) x[db][j] = pre[db][j];
    if (sc0 + KC - 1 > s_end) {
      __asm__ volatile("" ::: "memory");  // a branch, not eight selects per fragment
      const int nv = s_end - (sc0 + 8 * g);
#pragma unroll
      for (int db = 0; db < SX_NDB; ++db)
#pragma unroll
        for (int j = 0; j < 8; ++j) x[db][j] = (j <= nv) ? x[db][j] : 0.0f;
    }
    u32x4 dv[SX_NDB][3];
#pragma unroll
    for (int db = 0; db < SX_NDB; ++db) split3x8(x[db], dv[db][0], dv[db][1], dv[db][2]);
    int live = 0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int ib = i0 + q * 16;
      if (!(sc0 + KC - 1 < ib + base - t_hi || sc0 > ib + 15 + base - t_lo)) live |= 1 << q;
    }
    live = __builtin_amdgcn_readfirstlane(live);
    if constexpr (NP == 1) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        if (live & (1 << q)) {
          const u32x4 (&tq)[3] = (q < 2) ? tvn[q & 1] : tvo[q & 1];
#pragma unroll
          for (int db = 0; db < SX_NDB; ++db)
            acc[0][q][db] = (MODE == 0) ? mfma_split6(dv[db], tq, acc[0][q][db])
                                        : mfma_split6(tq, dv[db], acc[0][q][db]);
        }
      }
    } else {
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        if (p < jnp) {
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            if (live & (1 << q)) {
              u32x4 tq[3];
#pragma unroll
              for (int pl = 0; pl < 3; ++pl)
                tq[pl] = strip[buf][p][pl][cr][7 - 2 * q - (n >> 3) + g];
#pragma unroll
              for (int db = 0; db < SX_NDB; ++db)
                acc[p][q][db] = (MODE == 0) ? mfma_split6(dv[db], tq, acc[p][q][db])
                                            : mfma_split6(tq, dv[db], acc[p][q][db]);
            }
          }
        }
      }
    }
  };

  // Chunk c: dense values in register set A (even c) or B (odd c), loaded
  // after chunk c - 2; Toeplitz windows in LDS buffer c & 1, their units
  // loaded with the same set and stored one chunk later (after chunk c - 1),
  // so that every load has a chunk of work to complete behind.
  if (nchunks > 0) {
    gload(preA, spA, 0);
    sstore(spA, 0);
    if (nchunks > 1) gload(preB, spB, 1);
  }
  __syncthreads();
  u32x4 tvA[2][3], tvB[2][3];
  for (int ch = 0; ch < nchunks; ch += 2) {
    compute(ch, 0, preA, tvA, tvB);
    if (ch + 2 < nchunks) gload(preA, spA, ch + 2);
    if (ch + 1 < nchunks) sstore(spB, 1);
    __syncthreads();
    if (ch + 1 < nchunks) {
      compute(ch + 1, 1, preB, tvB, tvA);
      if (ch + 3 < nchunks) gload(preB, spB, ch + 3);
      if (ch + 2 < nchunks) sstore(spA, 0);
      __syncthreads();
    }
  }
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    if (p >= jnp) break;
    float* dp = dst + p * ostride;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int db = 0; db < SX_NDB; ++db)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int rr = 4 * (lane >> 4) + j, cc = lane & 15;
          int dd, ii;
          if (MODE == 0) {
            dd = d0 + dw + db * 16 + rr;
            ii = i0 + q * 16 + cc;
          } else {
            ii = i0 + q * 16 + rr;
            dd = d0 + dw + db * 16 + cc;
          }
          if (dd < ndense && ii < nout) {
            if (MODE == 0) dp[(int64_t)dd * nout + ii] = acc[p][q][db][j];
            else dp[(int64_t)ii * inner + dd] = acc[p][q][db][j];
          }
        }
  }
}

// ---------------------------------------------------------------------------
// largest (|) moment (|) over nodes and bandwidths
// ---------------------------------------------------------------------------

struct OffTab {
  int64_t v[FLPR_MAX_MOMENTS + 1];
};

constexpr int RED_THREADS = 256;

// Batch: s[0] = src; one key per problem.
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_maxmom(const T* __restrict__ src, int64_t per_h,
                                                        OffTab offs, int nslab, int64_t nodes,
                                                        int absval, uint64_t* __restrict__ key,
                                                        FlprBz bz) {
  __shared__ double sh[RED_THREADS / 64];
  src += blockIdx.z * bz.s[0];
  key += blockIdx.z;
  const int hs = blockIdx.y;  // h * nslab + slab
  const int h = hs / nslab, k = hs % nslab;
  const T* p = src + (int64_t)h * per_h + offs.v[k];
  double mx = -INFINITY;
  int64_t head = 0;
  if constexpr (sizeof(T) == 4) {
    // 16-byte loads over the aligned body of the slab
    if ((reinterpret_cast<uintptr_t>(p) & 15) == 0) {
      const int64_t n4 = nodes / 4;
      const float4* p4 = reinterpret_cast<const float4*>(p);
      for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
           i += (int64_t)gridDim.x * blockDim.x) {
        const float4 q = p4[i];
        double a0 = q.x, a1 = q.y, a2 = q.z, a3 = q.w;
        if (absval) {
          a0 = fabs(a0);
          a1 = fabs(a1);
          a2 = fabs(a2);
          a3 = fabs(a3);
        }
        mx = fmax(mx, fmax(fmax(a0, a1), fmax(a2, a3)));
      }
      head = n4 * 4;
    }
  }
  for (int64_t node = head + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; node < nodes;
       node += (int64_t)gridDim.x * blockDim.x) {
    double v = (double)p[node];
    if (absval) v = fabs(v);
    mx = fmax(mx, v);
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = sh[0];
    for (int w = 1; w < RED_THREADS / 64; ++w) r = fmax(r, sh[w]);
    atomicMax((unsigned long long*)key, (unsigned long long)flpr_encode(r));
  }
}

// ---------------------------------------------------------------------------
// per-node solves
// ---------------------------------------------------------------------------

struct SolveTab {
  int64_t s_off[FLPR_MAX_TERMS * (FLPR_MAX_TERMS + 1) / 2];
  int64_t t_off[FLPR_MAX_TERMS];
};

constexpr int SOLVE_THREADS = 256;
static_assert(SOLVE_THREADS == 256, "flpr_group_sum takes blocks of 256 threads");
constexpr int MAXV = 32;  // value channels reduced per solve launch
// value channels whose moments are loaded together (a divisor of MAXV)
template <typename T>
constexpr int solve_group() { return 4; }

// One thread per (node, bandwidth); block row blockIdx.y = bandwidth.
// Order >= 1: Cholesky of S + lam I in float64 (symmetric positive definite
// for lam > 0 when the spread weights are positive), z = (S + lam I)^-1 e1,
// and m_c = z . T_c for every value channel (the intercept e1' S^-1 T that the
// reference's Cramer formulas evaluate).  The deconvolved grids of accuracy
// > 6 can make S indefinite at a few boundary nodes: with PIV, a node whose
// Cholesky meets a non-positive pivot is solved again by Gaussian elimination
// with partial pivoting (LAPACK gesv's method); without PIV such a node gives
// NaN (a separate instantiation keeps the fallback's registers out of the
// positive-definite case).  Order 0: m_c = T_c / max(S + eps, thr).  Channels were centred on
// the grid by shift_c; since T_c(y) = T_c(y - shift_c) + shift_c T_0, the
// solve adds shift_c z . T_0 (shift_c S_00 for order 0) back exactly.
// The first nout value channels are written to m_out; for channels with an
// adjoint row in q, dot[h][c] += sum_node q_c(node) m_c(node) is reduced in
// registers over the block's nodes and then once per block.
// BUF (float32 moments whose per-bandwidth block is below 2 GiB): the
// moments are read through one buffer resource per bandwidth, with the
// per-thread part of each address (channel plane + node) in a 32-bit vector
// offset and the moment's plane offset in a scalar one, so that the compiler
// keeps no 64-bit base per (moment, channel).  Reads past the block return 0.
// Batch (blockIdx.z = problem): s[0] = mom, s[1] = shift, s[2] = m_out,
// s[3] = q, s[4] = dot; one key per problem.
// REF (the precision form, float32 moments): S + lam I is factored in
// float32 and z = (S + lam I)^-1 e1 takes one refinement step with the
// residual e1 - (S + lam I) z formed in float64; a node whose float32
// Cholesky meets a non-positive pivot, or whose pivot ratio max/min (a cheap
// lower estimate of the condition number) exceeds REF_COND, is solved in
// float64 as without REF.
constexpr float REF_COND = 1e6f;

template <typename T, int NT, bool PIV, bool BUF, bool REF = false>
__global__ __launch_bounds__(SOLVE_THREADS) void k_solve(
    const T* __restrict__ mom, int64_t per_h, int64_t nodes, SolveTab tab,
    const uint64_t* __restrict__ key, double lam_scale, double eps_add,
    const double* __restrict__ shift, int nval, int nout, double* __restrict__ m_out,
    const double* __restrict__ q, double* __restrict__ dot, double* __restrict__ sum_part,
    FlprBz bz) {
  __shared__ double red[SOLVE_THREADS / 64][MAXV];
  {
    const int64_t z = blockIdx.z;
    mom += z * bz.s[0];
    key += z;
    if (shift) shift += z * bz.s[1];
    if (m_out) m_out += z * bz.s[2];
    if (q) q += z * bz.s[3];
    if (dot) dot += z * bz.s[4];
  }
  constexpr int CGR = solve_group<T>();
  static_assert(MAXV % CGR == 0, "channel groups");
  const double kmax = flpr_decode(*key);
  const int h = blockIdx.y;
  const T* b = mom + (int64_t)h * per_h;
  // (unused, and removed, without BUF)
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<T*>(b), (short)0, BUF ? (int)(per_h * (int64_t)sizeof(T)) : 0, 0x00020000);
  // moment element uni + thr of this bandwidth's block (uni uniform)
  auto mload = [&](int64_t uni, int64_t thr) -> T {
    if constexpr (BUF)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(
                                       rs, (int)(thr * (int64_t)sizeof(T)),
                                       (int)(uni * (int64_t)sizeof(T)), 0));
    else
      return b[uni + thr];
  };
  double dacc[MAXV];
#pragma unroll
  for (int c = 0; c < MAXV; ++c) dacc[c] = 0.0;
  for (int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; node < nodes;
       node += (int64_t)gridDim.x * blockDim.x) {
    double z[NT];
    double den = 1.0, s00 = 0.0, zt0 = 0.0;
    // the value channels' moments are read in groups of CGR channels, each
    // group's loads issued together (and the first group's before the
    // factorisation), so that a node waits for memory a few times, not once
    // per channel
    T tnext[CGR][NT];
    double qnext[CGR];
    auto load_group = [&](int g0) {
#pragma unroll
      for (int j = 0; j < CGR; ++j) {
        const bool on = g0 + j < nval;
        const int64_t c = g0 + j + 1;  // channel 0 is the ones channel
#pragma unroll
        for (int i = 0; i < NT; ++i) tnext[j][i] = on ? mload(tab.t_off[i], c * nodes + node) : T(0);
        qnext[j] = (q && on) ? q[(int64_t)(g0 + j) * nodes + node] : 0.0;
      }
    };
    load_group(0);
    T t0v[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) t0v[i] = mload(tab.t_off[i], node);
    if (NT == 1) {
      s00 = (double)mload(tab.s_off[0], node);
      const double s = s00 + eps_add;                    // regression.py:127
      const double thr = (kmax + eps_add) * lam_scale;  // regression.py:217
      den = fmax(s, thr);
      z[0] = 1.0;
    } else {
      const double lam = kmax * lam_scale;  // regression.py:243-267
      bool done = false;
      if constexpr (REF) {
        float F[NT][NT];
        int q = 0;
#pragma unroll
        for (int i = 0; i < NT; ++i)
#pragma unroll
          for (int j = i; j < NT; ++j) {
            const float v = (float)mload(tab.s_off[q], node);
            F[i][j] = v;
            F[j][i] = v;
            ++q;
          }
        const float lamf = (float)lam;
#pragma unroll
        for (int i = 0; i < NT; ++i) F[i][i] += lamf;
        float dmin = INFINITY, dmax = 0.0f;
        bool indef = false;
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          float dsum = F[j][j];
#pragma unroll
          for (int k = 0; k < j; ++k) dsum -= F[j][k] * F[j][k];
          indef |= !(dsum > 0.0f);
          dmin = fminf(dmin, dsum);
          dmax = fmaxf(dmax, dsum);
          const float inv = 1.0f / sqrtf(dsum);
          F[j][j] = inv;
#pragma unroll
          for (int i = j + 1; i < NT; ++i) {
            float sm = F[i][j];
#pragma unroll
            for (int k = 0; k < j; ++k) sm -= F[i][k] * F[j][k];
            F[i][j] = sm * inv;
          }
        }
        if (!indef && dmax <= REF_COND * dmin) {
          // (L L')^-1 b in float32
          auto chol_solve = [&](const float (&b)[NT], float (&x)[NT]) {
            float yv[NT];
#pragma unroll
            for (int i = 0; i < NT; ++i) {
              float sm = b[i];
#pragma unroll
              for (int k = 0; k < i; ++k) sm -= F[i][k] * yv[k];
              yv[i] = sm * F[i][i];
            }
#pragma unroll
            for (int i = NT - 1; i >= 0; --i) {
              float sm = yv[i];
#pragma unroll
              for (int k = i + 1; k < NT; ++k) sm -= F[k][i] * x[k];
              x[i] = sm * F[i][i];
            }
          };
          float e[NT], zf[NT], dz[NT];
#pragma unroll
          for (int i = 0; i < NT; ++i) e[i] = (i == 0) ? 1.0f : 0.0f;
          chol_solve(e, zf);
          // residual in float64 from the moments
          double r[NT];
#pragma unroll
          for (int i = 0; i < NT; ++i) r[i] = (i == 0) ? 1.0 : 0.0;
          q = 0;
#pragma unroll
          for (int i = 0; i < NT; ++i)
#pragma unroll
            for (int j = i; j < NT; ++j) {
              const double v = (double)mload(tab.s_off[q], node) + (i == j ? lam : 0.0);
              r[i] -= v * (double)zf[j];
              if (i != j) r[j] -= v * (double)zf[i];
              ++q;
            }
          float rf[NT];
#pragma unroll
          for (int i = 0; i < NT; ++i) rf[i] = (float)r[i];
          chol_solve(rf, dz);
#pragma unroll
          for (int i = 0; i < NT; ++i) z[i] = (double)zf[i] + (double)dz[i];
          done = true;
        }
      }
      if (!done) {
      double A[NT][NT];
      int p = 0;
#pragma unroll
      for (int i = 0; i < NT; ++i)
#pragma unroll
        for (int j = i; j < NT; ++j) {
          const double v = (double)mload(tab.s_off[p], node);
          A[i][j] = v;
          A[j][i] = v;
          ++p;
        }
#pragma unroll
      for (int i = 0; i < NT; ++i) A[i][i] += lam;
      bool indefinite = false;
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        double dsum = A[j][j];
#pragma unroll
        for (int k = 0; k < j; ++k) dsum -= A[j][k] * A[j][k];
        indefinite |= !(dsum > 0.0);
        // the diagonal keeps 1 / L_jj: the triangular solves multiply by it
        const double inv = 1.0 / sqrt(dsum);
        A[j][j] = inv;
#pragma unroll
        for (int i = j + 1; i < NT; ++i) {
          double s = A[i][j];
#pragma unroll
          for (int k = 0; k < j; ++k) s -= A[i][k] * A[j][k];
          A[i][j] = s * inv;
        }
      }
      double y[NT];
#pragma unroll
      for (int i = 0; i < NT; ++i) {
        double s = (i == 0) ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < i; ++k) s -= A[i][k] * y[k];
        y[i] = s * A[i][i];
      }
#pragma unroll
      for (int i = NT - 1; i >= 0; --i) {
        double s = y[i];
#pragma unroll
        for (int k = i + 1; k < NT; ++k) s -= A[k][i] * z[k];
        z[i] = s * A[i][i];
      }
      if (PIV && indefinite) {
        // reload S + lam I; rows are exchanged with selects so that every
        // register index stays static
        p = 0;
#pragma unroll
        for (int i = 0; i < NT; ++i)
#pragma unroll
          for (int j = i; j < NT; ++j) {
            const double v = (double)b[tab.s_off[p] + node] + (i == j ? lam : 0.0);
            A[i][j] = v;
            A[j][i] = v;
            ++p;
          }
#pragma unroll
        for (int i = 0; i < NT; ++i) y[i] = (i == 0) ? 1.0 : 0.0;
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          int piv = j;
          double best = fabs(A[j][j]);
#pragma unroll
          for (int i = j + 1; i < NT; ++i) {
            const double a = fabs(A[i][j]);
            if (a > best) {
              best = a;
              piv = i;
            }
          }
#pragma unroll
          for (int i = j + 1; i < NT; ++i) {
            const bool sw = (i == piv);
#pragma unroll
            for (int k = j; k < NT; ++k) {
              const double u = A[j][k], v = A[i][k];
              A[j][k] = sw ? v : u;
              A[i][k] = sw ? u : v;
            }
            const double u = y[j], v = y[i];
            y[j] = sw ? v : u;
            y[i] = sw ? u : v;
          }
          const double inv = 1.0 / A[j][j];
#pragma unroll
          for (int i = j + 1; i < NT; ++i) {
            const double f = A[i][j] * inv;
#pragma unroll
            for (int k = j + 1; k < NT; ++k) A[i][k] -= f * A[j][k];
            y[i] -= f * y[j];
          }
        }
#pragma unroll
        for (int i = NT - 1; i >= 0; --i) {
          double s = y[i];
#pragma unroll
          for (int k = i + 1; k < NT; ++k) s -= A[i][k] * z[k];
          z[i] = s / A[i][i];
        }
      }
      }  // !done
      // z . T_0 (ones channel) for the centring correction
#pragma unroll
      for (int i = 0; i < NT; ++i) zt0 += z[i] * (double)t0v[i];
    }
#pragma unroll
    for (int g0 = 0; g0 < MAXV; g0 += CGR) {
      if (g0 >= nval) break;
      T tcur[CGR][NT];
      double qcur[CGR];
#pragma unroll
      for (int j = 0; j < CGR; ++j) {
        qcur[j] = qnext[j];
#pragma unroll
        for (int i = 0; i < NT; ++i) tcur[j][i] = tnext[j][i];
      }
      if (g0 + CGR < nval) load_group(g0 + CGR);
#pragma unroll
      for (int j = 0; j < CGR; ++j) {
        const int cc = g0 + j;
        if (cc >= nval) break;
        const int c = cc + 1;
        const double sh = shift ? shift[c] : 0.0;
        double m;
        if (NT == 1) {
          m = ((double)tcur[j][0] + sh * s00) / den;
        } else {
          m = sh * zt0;
#pragma unroll
          for (int i = 0; i < NT; ++i) m += z[i] * (double)tcur[j][i];
        }
        if (cc < nout) m_out[((int64_t)h * nout + cc) * nodes + node] = m;
        if (q) dacc[cc] += qcur[j] * m;
      }
    }
  }
  if (q && nval > 0) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int cc = 0; cc < MAXV; ++cc) {
      if (cc >= nval) continue;
      double v = dacc[cc];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0) red[wid][cc] = v;
    }
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x < (unsigned)nval)
      for (int w = 0; w < SOLVE_THREADS / 64; ++w) s += red[w][threadIdx.x];
    // the node blocks of (problem, bandwidth) add their dots in block order
    flpr_group_sum(s, nval, blockIdx.x, gridDim.x, (int64_t)blockIdx.z * gridDim.y + h,
                   (int64_t)gridDim.z * gridDim.y, sum_part, dot + (int64_t)h * nval);
  }
}

template <typename T, int NT>
void launch_solve(hipStream_t st, dim3 grd, bool piv, const T* mom, int64_t per_h,
                  int64_t nodes, const SolveTab& tab, const uint64_t* key, double lam_scale,
                  double eps_add, const double* shift, int nval, int nout, double* m_out,
                  const double* q, double* dot, double* sum_part, const FlprBz& bz,
                  bool ref = false) {
  const bool buf = sizeof(T) == 4 && per_h * (int64_t)sizeof(T) <= (int64_t)INT32_MAX;
  if (ref && sizeof(T) == 4 && NT > 1) {
    if (buf)
      hipLaunchKernelGGL((k_solve<T, NT, false, sizeof(T) == 4, true>), grd, dim3(SOLVE_THREADS),
                         0, st, mom, per_h, nodes, tab, key, lam_scale, eps_add, shift, nval,
                         nout, m_out, q, dot, sum_part, bz);
    else
      hipLaunchKernelGGL((k_solve<T, NT, false, false, true>), grd, dim3(SOLVE_THREADS), 0, st,
                         mom, per_h, nodes, tab, key, lam_scale, eps_add, shift, nval, nout,
                         m_out, q, dot, sum_part, bz);
    return;
  }
  if (piv && NT > 1)
    hipLaunchKernelGGL((k_solve<T, NT, true, false>), grd, dim3(SOLVE_THREADS), 0, st, mom,
                       per_h, nodes, tab, key, lam_scale, eps_add, shift, nval, nout, m_out, q, dot, sum_part, bz);
  else if (buf)
    hipLaunchKernelGGL((k_solve<T, NT, false, sizeof(T) == 4>), grd, dim3(SOLVE_THREADS), 0, st,
                       mom, per_h, nodes, tab, key, lam_scale, eps_add, shift, nval, nout, m_out,
                       q, dot, sum_part, bz);
  else
    hipLaunchKernelGGL((k_solve<T, NT, false, false>), grd, dim3(SOLVE_THREADS), 0, st, mom,
                       per_h, nodes, tab, key, lam_scale, eps_add, shift, nval, nout, m_out, q, dot, sum_part, bz);
}

template <typename T>
int dispatch_solve(int nt, hipStream_t st, dim3 grd, bool piv, const T* mom, int64_t per_h,
                   int64_t nodes,
                   const SolveTab& tab, const uint64_t* key, double lam_scale, double eps_add,
                   const double* shift, int nval, int nout, double* m_out, const double* q,
                   double* dot, double* sum_part, const FlprBz& bz, bool ref = false) {
#define FLPR_SOLVE_CASE(N)                                                                   \
  case N:                                                                                    \
    launch_solve<T, N>(st, grd, piv, mom, per_h, nodes, tab, key, lam_scale, eps_add, shift, \
                       nval, nout, m_out, q, dot, sum_part, bz, ref);                        \
    return 0;
  switch (nt) {
    FLPR_SOLVE_CASE(1)
    FLPR_SOLVE_CASE(2)
    FLPR_SOLVE_CASE(3)
    FLPR_SOLVE_CASE(4)
    FLPR_SOLVE_CASE(6)
    FLPR_SOLVE_CASE(10)
    default:
      return -1;
  }
#undef FLPR_SOLVE_CASE
}

// Node blocks per bandwidth of a solve launch: enough to fill the device,
// and, with probe dots (nval > 0), no more than the dot workspace holds
// (groups = P * dh of nval sums each); 0 when the workspace is missing or
// too small for one block per group.
int solve_blocks(int64_t nodes, int dh, int P, int nval, const double* sum_part,
                 int64_t sum_part_len) {
  const int bx = flpr_grid_for(nodes, SOLVE_THREADS, 1024 / (dh > 64 ? 64 : dh) + 1);
  if (nval < 1) return bx;
  if (!sum_part) return 0;
  return flpr_sum_blocks((int64_t)P * dh, nval, bx, sum_part_len);
}

}  // namespace

extern "C" {

int flpr_taps(const FlprGeom* g, const double* h, int32_t dh, int32_t npow, int32_t nmax,
              double norm, double* tap64, float* tap32, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !h || !tap64 || dh < 1 || npow < 1 || nmax < 1 || norm <= 0) return -1;
  for (int k = 0; k < g->d; ++k)
    if (g->n_grid[k] > nmax) return -3;
  const int64_t total = (int64_t)dh * g->d * nmax;
  hipLaunchKernelGGL(k_taps, dim3(flpr_grid_for(total, TAP_THREADS, 4096)), dim3(TAP_THREADS), 0,
                     as_stream(stream), *g, h, dh, npow, nmax, norm, tap64, tap32, FlprBz{});
  FLPR_CHECK_LAUNCH();
  return 0;
}

int64_t flpr_conv_strip_len(int32_t nmax) { return nmax < 1 ? -1 : split_strip_len(nmax); }

int flpr_conv_strips(const float* tap32, const int32_t* bands, int32_t ntap, int32_t npow,
                     int32_t nmax, uint16_t* strips, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!tap32 || !bands || !strips || ntap < 1 || npow < 1 || nmax < 1 || ntap % npow) return -1;
  const int64_t V = split_strip_len(nmax);
  const int64_t total = (int64_t)ntap * 8 * (V / 8);
  hipLaunchKernelGGL(k_split_taps, dim3(flpr_grid_for(total, 256, 4096)), dim3(256), 0,
                     as_stream(stream), tap32, bands, ntap, npow, nmax, V, strips);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_conv_tiles(int32_t mode, int32_t fp64, int32_t rows_or_outer, int32_t max_rows_or_outer,
                    int32_t nout, int32_t inner, int32_t* tiles, int32_t* tile_stride) {
  if (!tiles || !tile_stride || rows_or_outer < 1 || max_rows_or_outer < rows_or_outer ||
      nout < 1 || (mode != 0 && mode != 1) || (mode == 1 && inner < 1))
    return -1;
  const int ld = fp64 ? LD : SX_LD;
  const int ntt = (nout + LT - 1) / LT;
  if (mode == 0) {
    *tile_stride = (max_rows_or_outer + ld - 1) / ld;
    *tiles = ((rows_or_outer + ld - 1) / ld) * ntt;
  } else {
    *tile_stride = (inner + ld - 1) / ld;
    *tiles = *tile_stride * ntt * rows_or_outer;
  }
  return 0;
}

int flpr_conv(int32_t mode, int32_t fp64, const void* in, void* out, const void* tap,
              int32_t nmax, const int64_t* jobs, const int32_t* bands, int32_t nj,
              int32_t max_rows_or_outer, int32_t K, int32_t nout, int32_t inner, int32_t base,
              int32_t max_powers, const int32_t* blocks, int32_t nblocks, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!in || !out || !tap || !jobs || !bands || nj < 1 || nj > 65535 || max_rows_or_outer < 1 || K < 1 ||
      nout < 1 || (mode != 0 && mode != 1) || (mode == 1 && inner < 1) ||
      (blocks && nblocks < 1) || nmax < 1 || max_powers < 1 ||
      max_powers > (fp64 ? 1 : SX_MAX_POWERS))
    return -1;
  const int ld = fp64 ? LD : SX_LD;
  const int ntd = (mode == 0) ? (max_rows_or_outer + ld - 1) / ld : (inner + ld - 1) / ld;
  const int ntt = (nout + LT - 1) / LT;
  const int64_t gx = (mode == 0) ? (int64_t)ntd * ntt : (int64_t)ntd * ntt * max_rows_or_outer;
  if (gx > INT32_MAX) return -1;
  dim3 grd = blocks ? dim3((unsigned)nblocks) : dim3((unsigned)gx, (unsigned)nj);
  hipStream_t st = as_stream(stream);
  if (fp64) {
    if (mode == 0)
      hipLaunchKernelGGL((k_conv<0>), grd, dim3(CT), 0, st, (const double*)in,
                         (double*)out, (const double*)tap, jobs, bands, K, nout, inner, base, ntd, ntt,
                         blocks);
    else
      hipLaunchKernelGGL((k_conv<1>), grd, dim3(CT), 0, st, (const double*)in,
                         (double*)out, (const double*)tap, jobs, bands, K, nout, inner, base, ntd, ntt,
                         blocks);
  } else {
#define FLPR_SPLIT_LAUNCH(M, P)                                                                 \
  hipLaunchKernelGGL((k_conv_split<M, P>), grd, dim3(CT), 0, st, (const float*)in, (float*)out, \
                     (const uint16_t*)tap, nmax, jobs, bands, K, nout, inner, base, ntd, ntt,   \
                     blocks)
    if (mode == 0) {
      if (max_powers == 1) FLPR_SPLIT_LAUNCH(0, 1);
      else if (max_powers == 2) FLPR_SPLIT_LAUNCH(0, 2);
      else FLPR_SPLIT_LAUNCH(0, 3);
    } else {
      if (max_powers == 1) FLPR_SPLIT_LAUNCH(1, 1);
      else if (max_powers == 2) FLPR_SPLIT_LAUNCH(1, 2);
      else FLPR_SPLIT_LAUNCH(1, 3);
    }
#undef FLPR_SPLIT_LAUNCH
  }
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_max_moment(int32_t fp64, const void* src, int64_t per_h, int32_t dh, const int64_t* offs,
                    int32_t nslab, int64_t nodes, int32_t absval, uint64_t* key, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!src || !offs || !key || dh < 1 || nslab < 1 || nslab > FLPR_MAX_MOMENTS || nodes < 1)
    return -1;
  OffTab t;
  for (int k = 0; k < nslab; ++k) t.v[k] = offs[k];
  if ((int64_t)dh * nslab > 65535) return -1;
  // about a thousand blocks in total: every block ends in one atomic on *key
  const int per = 1024 / (dh * nslab) > 1 ? 1024 / (dh * nslab) : 1;
  dim3 grd(flpr_grid_for(nodes, RED_THREADS, per), (unsigned)(dh * nslab));
  if (fp64)
    hipLaunchKernelGGL((k_maxmom<double>), grd, dim3(RED_THREADS), 0, as_stream(stream),
                       (const double*)src, per_h, t, nslab, nodes, absval, key, FlprBz{});
  else
    hipLaunchKernelGGL((k_maxmom<float>), grd, dim3(RED_THREADS), 0, as_stream(stream),
                       (const float*)src, per_h, t, nslab, nodes, absval, key, FlprBz{});
  FLPR_CHECK_LAUNCH();
  return 0;
}

double flpr_key_to_double(uint64_t key) { return flpr_decode(key); }

int flpr_solve(int32_t fp64, const void* mom, int64_t per_h, int32_t dh, int64_t nodes,
               int32_t order, int32_t nt, int32_t pivot, const int64_t* s_off,
               const int64_t* t_off, const uint64_t* key, double lam_scale, double eps_add,
               const double* shift, int32_t nval, int32_t nout, double* m_out, const double* q,
               double* dot, double* sum_part, int64_t sum_part_len, void* stream) {
  FLPR_CLEAR_ERROR();
  if (!mom || !s_off || !t_off || !key || dh < 1 || dh > 65535 || nodes < 1 || nt < 1 ||
      nt > FLPR_MAX_TERMS || nval < 0 || nval > MAXV || nout < 0 || nout > nval)
    return -1;
  if ((order == 0) != (nt == 1)) return -1;
  if (nout > 0 && !m_out) return -1;
  if (q && !dot) return -1;
  SolveTab tab;
  const int ns = nt * (nt + 1) / 2;
  for (int i = 0; i < ns; ++i) tab.s_off[i] = s_off[i];
  for (int i = 0; i < nt; ++i) tab.t_off[i] = t_off[i];
  const int bx = solve_blocks(nodes, dh, 1, q ? nval : 0, sum_part, sum_part_len);
  if (bx < 1) return -3;
  dim3 grd((unsigned)bx, (unsigned)dh);
  hipStream_t st = as_stream(stream);
  int rc;
  if (fp64)
    rc = dispatch_solve<double>(nt, st, grd, pivot != 0, (const double*)mom, per_h, nodes, tab, key,
                                lam_scale, eps_add, shift, nval, nout, m_out, q, dot, sum_part,
                                FlprBz{});
  else
    rc = dispatch_solve<float>(nt, st, grd, pivot != 0, (const float*)mom, per_h, nodes, tab, key,
                               lam_scale, eps_add, shift, nval, nout, m_out, q, dot, sum_part,
                               FlprBz{});
  if (rc) return rc;
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_solve_refined(const void* mom, int64_t per_h, int32_t dh, int64_t nodes,
                       int32_t order, int32_t nt, const int64_t* s_off, const int64_t* t_off,
                       const uint64_t* key, double lam_scale, double eps_add,
                       const double* shift, int32_t nval, int32_t nout, double* m_out,
                       const double* q, double* dot, double* sum_part, int64_t sum_part_len,
                       void* stream) {
  FLPR_CLEAR_ERROR();
  if (!mom || !s_off || !t_off || !key || dh < 1 || dh > 65535 || nodes < 1 || nt < 1 ||
      nt > FLPR_MAX_TERMS || nval < 0 || nval > MAXV || nout < 0 || nout > nval)
    return -1;
  if ((order == 0) != (nt == 1)) return -1;
  if (nout > 0 && !m_out) return -1;
  if (q && !dot) return -1;
  SolveTab tab;
  const int ns = nt * (nt + 1) / 2;
  for (int i = 0; i < ns; ++i) tab.s_off[i] = s_off[i];
  for (int i = 0; i < nt; ++i) tab.t_off[i] = t_off[i];
  const int bx = solve_blocks(nodes, dh, 1, q ? nval : 0, sum_part, sum_part_len);
  if (bx < 1) return -3;
  dim3 grd((unsigned)bx, (unsigned)dh);
  const int rc = dispatch_solve<float>(nt, as_stream(stream), grd, false, (const float*)mom, per_h,
                                       nodes, tab, key, lam_scale, eps_add, shift, nval, nout,
                                       m_out, q, dot, sum_part, FlprBz{}, true);
  if (rc) return rc;
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_batch_taps(const FlprGeom* g, const FlprBatchItem* items, int32_t P, const double* h,
                    int32_t dh, int32_t npow, int32_t nmax, double* tap64, float* tap32,
                    void* stream) {
  FLPR_CLEAR_ERROR();
  if (!flpr_geom_ok(g) || !items || P < 1 || P > 65535 || !h || !tap64 || dh < 1 || npow < 1 ||
      nmax < 1)
    return -1;
  for (int k = 0; k < g->d; ++k)
    if (g->n_grid[k] > nmax) return -3;
  const int64_t total = (int64_t)dh * g->d * nmax;
  FlprBz bz{};
  bz.items = items;
  bz.s[0] = total * npow;
  hipLaunchKernelGGL(k_taps, dim3(flpr_grid_for(total, TAP_THREADS, 4096), 1, P),
                     dim3(TAP_THREADS), 0, as_stream(stream), *g, h, dh, npow, nmax, 1.0, tap64,
                     tap32, bz);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_batch_max_moment(int32_t fp64, int32_t P, const void* src, int64_t prob_stride,
                          int64_t per_h, int32_t dh, const int64_t* offs, int32_t nslab,
                          int64_t nodes, int32_t absval, uint64_t* key, void* stream) {
  FLPR_CLEAR_ERROR();
  if (P < 1 || P > 65535 || !src || !offs || !key || dh < 1 || nslab < 1 ||
      nslab > FLPR_MAX_MOMENTS || nodes < 1 || (int64_t)dh * nslab > 65535)
    return -1;
  OffTab t;
  for (int k = 0; k < nslab; ++k) t.v[k] = offs[k];
  const int per = 1024 / (dh * nslab) > 1 ? 1024 / (dh * nslab) : 1;
  dim3 grd(flpr_grid_for(nodes, RED_THREADS, per), (unsigned)(dh * nslab), (unsigned)P);
  FlprBz bz{};
  bz.s[0] = prob_stride;
  if (fp64)
    hipLaunchKernelGGL((k_maxmom<double>), grd, dim3(RED_THREADS), 0, as_stream(stream),
                       (const double*)src, per_h, t, nslab, nodes, absval, key, bz);
  else
    hipLaunchKernelGGL((k_maxmom<float>), grd, dim3(RED_THREADS), 0, as_stream(stream),
                       (const float*)src, per_h, t, nslab, nodes, absval, key, bz);
  FLPR_CHECK_LAUNCH();
  return 0;
}

int flpr_batch_solve(int32_t fp64, int32_t P, const void* mom, int64_t mom_stride, int64_t per_h,
                     int32_t dh, int64_t nodes, int32_t order, int32_t nt, const int64_t* s_off,
                     const int64_t* t_off, const uint64_t* key, double lam_scale, double eps_add,
                     const double* shift, int32_t nval, int32_t nout, double* m_out,
                     int64_t m_stride, const double* q, double* dot, int64_t dot_stride,
                     double* sum_part, int64_t sum_part_len, void* stream) {
  FLPR_CLEAR_ERROR();
  if (P < 1 || P > 65535 || !mom || !s_off || !t_off || !key || dh < 1 || dh > 65535 ||
      nodes < 1 || nt < 1 || nt > FLPR_MAX_TERMS || nval < 0 || nval > MAXV || nout < 0 ||
      nout > nval)
    return -1;
  if ((order == 0) != (nt == 1)) return -1;
  if (nout > 0 && !m_out) return -1;
  if (q && !dot) return -1;
  SolveTab tab;
  const int ns = nt * (nt + 1) / 2;
  for (int i = 0; i < ns; ++i) tab.s_off[i] = s_off[i];
  for (int i = 0; i < nt; ++i) tab.t_off[i] = t_off[i];
  const int bx = solve_blocks(nodes, dh, P, q ? nval : 0, sum_part, sum_part_len);
  if (bx < 1) return -3;
  dim3 grd((unsigned)bx, (unsigned)dh, (unsigned)P);
  FlprBz bz{};
  bz.s[0] = mom_stride;
  bz.s[1] = nval + 1;
  bz.s[2] = m_stride;
  bz.s[3] = (int64_t)nval * nodes;
  bz.s[4] = dot_stride;
  hipStream_t st = as_stream(stream);
  int rc;
  if (fp64)
    rc = dispatch_solve<double>(nt, st, grd, false, (const double*)mom, per_h, nodes, tab, key,
                                lam_scale, eps_add, shift, nval, nout, m_out, q, dot, sum_part, bz);
  else
    rc = dispatch_solve<float>(nt, st, grd, false, (const float*)mom, per_h, nodes, tab, key,
                               lam_scale, eps_add, shift, nval, nout, m_out, q, dot, sum_part, bz);
  if (rc) return rc;
  FLPR_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
