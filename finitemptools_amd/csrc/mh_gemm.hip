// Strided GEMM on CDNA4 matrix cores (gfx950).
//
//   C[m,n] = alpha * sum_k op(A)[m,k] * B[k,n] + beta * C[m,n]
//
// Every operand is addressed through two strides, so the contraction steps of
// H_eff and of the environment updates read the column-major tensors in place
// (no permute pass).  One of the two strides of A and of B is 1; the template
// flags AMC ("A is m-contiguous") and BKC ("B is k-contiguous") pick the
// coalesced global->LDS mapping.
//
// Tiling: a 256-thread block (4 waves of 64) owns a BM x BN tile of C; each
// wave owns a (BM/WGM) x (BN/WGN) sub-tile made of 16x16 MFMA tiles.  K is
// walked in steps of BK = 16 with a double-buffered LDS stage: global loads of
// step k+1 are issued into registers before the MFMAs of step k, and written to
// the other LDS buffer afterwards, so one barrier per step suffices.
//
// FP64: v_mfma_f64_16x16x4f64.  Lane l holds A[l%16][l/16], B[l/16][l%16];
// accumulator register r of lane l is C[l/16 + 4*r][l%16] (measured).
// Complex FP64: four real MFMAs per 16x16x4 step on split re/im LDS planes.
#include "mh_internal.h"

namespace mh {

typedef double d4_t __attribute__((ext_vector_type(4)));

struct GemmP {
  int M, N, K;
  const double *A;
  int64_t sam, sak;
  const double *B;
  int64_t sbk, sbn;
  double *C;
  int64_t scm, scn;
  double alpha_re, alpha_im, beta_re, beta_im;
  int conjA;
};

constexpr int BK = 16;

// ---------------------------------------------------------------- real FP64
template <int BM, int BN, int WGM, int WGN, bool AMC, bool BKC>
__global__ void __launch_bounds__(256)
    k_gemm_f64(const GemmP p) {
  constexpr int WTM = BM / WGM, WTN = BN / WGN;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  constexpr int PA = BM + 16, PB = BN + 16;
  constexpr int LA = BM * BK / 256, LB = BN * BK / 256;
  static_assert(WGM * WGN == 4, "4 waves per block");
  static_assert(TM >= 1 && TN >= 1, "wave tile");
  __shared__ double sA[2][BK * PA];
  __shared__ double sB[2][BK * PB];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WGM, wn = wave / WGM;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int lr = lane & 15, lk = lane >> 4;

  double ra[LA], rb[LB];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      int e = tid + 256 * i;
      int mm = AMC ? (e % BM) : (e / BK);
      int kk = AMC ? (e / BM) : (e % BK);
      int gm = m0 + mm, gk = k0 + kk;
      ra[i] = (gm < p.M && gk < p.K) ? p.A[gm * p.sam + gk * p.sak] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      int e = tid + 256 * i;
      int nn = BKC ? (e / BK) : (e % BN);
      int kk = BKC ? (e % BK) : (e / BN);
      int gn = n0 + nn, gk = k0 + kk;
      rb[i] = (gn < p.N && gk < p.K) ? p.B[gk * p.sbk + gn * p.sbn] : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      int e = tid + 256 * i;
      int mm = AMC ? (e % BM) : (e / BK);
      int kk = AMC ? (e / BM) : (e % BK);
      sA[buf][kk * PA + mm] = ra[i];
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      int e = tid + 256 * i;
      int nn = BKC ? (e / BK) : (e % BN);
      int kk = BKC ? (e % BK) : (e / BN);
      sB[buf][kk * PB + nn] = rb[i];
    }
  };

  d4_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};

  const int nk = (p.K + BK - 1) / BK;
  load(0);
  store(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BK);
#pragma unroll
    for (int k4 = 0; k4 < BK; k4 += 4) {
      double a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i)
        a[i] = sA[buf][(k4 + lk) * PA + wm * WTM + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < TN; ++j)
        b[j] = sB[buf][(k4 + lk) * PB + wn * WTN + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j],
                                                          0, 0, 0);
    }
    if (kt + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }

  const double alpha = p.alpha_re, beta = p.beta_re;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int n = n0 + wn * WTN + j * 16 + lr;
      if (n >= p.N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * WTM + i * 16 + lk + 4 * r;
        if (m >= p.M) continue;
        double *c = p.C + m * p.scm + n * p.scn;
        double v = alpha * acc[i][j][r];
        if (beta != 0.0) v += beta * (*c);
        *c = v;
      }
    }
}

// ------------------------------------------------------------- complex FP64
template <int BM, int BN, int WGM, int WGN, bool AMC, bool BKC>
__global__ void __launch_bounds__(256)
    k_gemm_c64(const GemmP p) {
  constexpr int WTM = BM / WGM, WTN = BN / WGN;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  constexpr int PA = BM + 16, PB = BN + 16;
  constexpr int LA = BM * BK / 256, LB = BN * BK / 256;
  __shared__ double sAr[2][BK * PA], sAi[2][BK * PA];
  __shared__ double sBr[2][BK * PB], sBi[2][BK * PB];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WGM, wn = wave / WGM;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int lr = lane & 15, lk = lane >> 4;
  const double2 *A = reinterpret_cast<const double2 *>(p.A);
  const double2 *B = reinterpret_cast<const double2 *>(p.B);
  const double csign = p.conjA ? -1.0 : 1.0;

  double2 ra[LA], rb[LB];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      int e = tid + 256 * i;
      int mm = AMC ? (e % BM) : (e / BK);
      int kk = AMC ? (e / BM) : (e % BK);
      int gm = m0 + mm, gk = k0 + kk;
      ra[i] = (gm < p.M && gk < p.K) ? A[gm * p.sam + gk * p.sak]
                                     : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      int e = tid + 256 * i;
      int nn = BKC ? (e / BK) : (e % BN);
      int kk = BKC ? (e % BK) : (e / BN);
      int gn = n0 + nn, gk = k0 + kk;
      rb[i] = (gn < p.N && gk < p.K) ? B[gk * p.sbk + gn * p.sbn]
                                     : make_double2(0.0, 0.0);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      int e = tid + 256 * i;
      int mm = AMC ? (e % BM) : (e / BK);
      int kk = AMC ? (e / BM) : (e % BK);
      sAr[buf][kk * PA + mm] = ra[i].x;
      sAi[buf][kk * PA + mm] = csign * ra[i].y;
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      int e = tid + 256 * i;
      int nn = BKC ? (e / BK) : (e % BN);
      int kk = BKC ? (e % BK) : (e / BN);
      sBr[buf][kk * PB + nn] = rb[i].x;
      sBi[buf][kk * PB + nn] = rb[i].y;
    }
  };

  d4_t accr[TM][TN], acci[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      accr[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
      acci[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
    }

  const int nk = (p.K + BK - 1) / BK;
  load(0);
  store(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BK);
#pragma unroll
    for (int k4 = 0; k4 < BK; k4 += 4) {
      double ar[TM], ai[TM], nai[TM], br[TN], bi[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        int o = (k4 + lk) * PA + wm * WTM + i * 16 + lr;
        ar[i] = sAr[buf][o];
        ai[i] = sAi[buf][o];
        nai[i] = -ai[i];
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        int o = (k4 + lk) * PB + wn * WTN + j * 16 + lr;
        br[j] = sBr[buf][o];
        bi[j] = sBi[buf][o];
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], accr[i][j], 0, 0, 0);
          acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], acci[i][j], 0, 0, 0);
          accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[i], bi[j], accr[i][j], 0, 0, 0);
          acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], acci[i][j], 0, 0, 0);
        }
    }
    if (kt + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }

  double2 *C = reinterpret_cast<double2 *>(p.C);
  const bool has_beta = (p.beta_re != 0.0 || p.beta_im != 0.0);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int n = n0 + wn * WTN + j * 16 + lr;
      if (n >= p.N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * WTM + i * 16 + lk + 4 * r;
        if (m >= p.M) continue;
        double2 *c = C + m * p.scm + n * p.scn;
        double xr = accr[i][j][r], xi = acci[i][j][r];
        double vr = p.alpha_re * xr - p.alpha_im * xi;
        double vi = p.alpha_re * xi + p.alpha_im * xr;
        if (has_beta) {
          double2 o = *c;
          vr += p.beta_re * o.x - p.beta_im * o.y;
          vi += p.beta_re * o.y + p.beta_im * o.x;
        }
        *c = make_double2(vr, vi);
      }
    }
}

// ------------------------------------------------------------------ dispatch
struct TileCfg {
  int bm, bn;
};

// Pick the tile that minimises (tiles per CU, rounded up) x (tile area): the
// kernels are MFMA-bound, so a CU's time is proportional to the C area it owns.
static TileCfg pick_tile(int M, int N) {
  const TileCfg cands[3] = {{64, 64}, {128, 64}, {128, 128}};
  const int ncu = 256;
  TileCfg best = cands[0];
  double best_cost = 1e300;
  for (const auto &c : cands) {
    int64_t tiles = (int64_t)((M + c.bm - 1) / c.bm) * ((N + c.bn - 1) / c.bn);
    double rounds = (double)((tiles + ncu - 1) / ncu);
    // Smaller tiles move more bytes per flop; weigh them slightly.
    double penalty = (c.bm * c.bn == 64 * 64) ? 1.08 : (c.bm * c.bn == 128 * 64 ? 1.03 : 1.0);
    double cost = rounds * c.bm * c.bn * penalty;
    if (cost < best_cost) {
      best_cost = cost;
      best = c;
    }
  }
  return best;
}

template <template <int, int, int, int, bool, bool> class K, bool AMC, bool BKC>
struct Launcher;

template <int BM, int BN, bool AMC, bool BKC>
static void launch_f64(const GemmP &p, dim3 grid, hipStream_t s, bool cplx) {
  if (cplx)
    hipLaunchKernelGGL((k_gemm_c64<BM, BN, 2, 2, AMC, BKC>), grid, dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((k_gemm_f64<BM, BN, 2, 2, AMC, BKC>), grid, dim3(256), 0, s, p);
}

template <bool AMC, bool BKC>
static void launch_tile(const TileCfg &t, const GemmP &p, hipStream_t s, bool cplx) {
  dim3 grid((p.M + t.bm - 1) / t.bm, (p.N + t.bn - 1) / t.bn);
  if (t.bm == 64 && t.bn == 64)
    launch_f64<64, 64, AMC, BKC>(p, grid, s, cplx);
  else if (t.bm == 128 && t.bn == 64)
    launch_f64<128, 64, AMC, BKC>(p, grid, s, cplx);
  else
    launch_f64<128, 128, AMC, BKC>(p, grid, s, cplx);
}

int gemm_mixed(mh_handle h, int dtype, int prec, const GemmDesc &g);

int gemm(mh_handle h, int dtype, int prec, const GemmDesc &g) {
  if (g.M <= 0 || g.N <= 0) return MH_OK;
  if (prec != MH_PREC_F64) return gemm_mixed(h, dtype, prec, g);
  GemmP p;
  p.M = g.M;
  p.N = g.N;
  p.K = g.K;
  p.A = (const double *)g.A;
  p.sam = g.sam;
  p.sak = g.sak;
  p.B = (const double *)g.B;
  p.sbk = g.sbk;
  p.sbn = g.sbn;
  p.C = (double *)g.C;
  p.scm = g.scm;
  p.scn = g.scn;
  p.alpha_re = g.alpha_re;
  p.alpha_im = g.alpha_im;
  p.beta_re = g.beta_re;
  p.beta_im = g.beta_im;
  p.conjA = g.conjA;
  const bool cplx = dtype == MH_C64;
  const bool amc = (g.sam == 1) || (g.M == 1);
  const bool bkc = (g.sbk == 1) || (g.K == 1);
  TileCfg t = pick_tile(g.M, g.N);
  double flops = 2.0 * g.M * (double)g.N * g.K * (cplx ? 4.0 : 1.0);
  int slot;
  prof_begin(h, MH_K_GEMM, flops, &slot);
  if (amc && bkc)
    launch_tile<true, true>(t, p, h->stream, cplx);
  else if (amc && !bkc)
    launch_tile<true, false>(t, p, h->stream, cplx);
  else if (!amc && bkc)
    launch_tile<false, true>(t, p, h->stream, cplx);
  else
    launch_tile<false, false>(t, p, h->stream, cplx);
  prof_end(h, slot);
  return check_launch(h, "gemm");
}

}  // namespace mh

extern "C" int mh_gemm(mh_handle h, int dtype, int prec, int M, int N, int K,
                       const void *A, int64_t sam, int64_t sak, int conjA,
                       const void *B, int64_t sbk, int64_t sbn, void *C,
                       int64_t scm, int64_t scn, double alpha_re,
                       double alpha_im, double beta_re, double beta_im) {
  using namespace mh;
  if (!h) return MH_ERR_ARG;
  if (dtype != MH_F64 && dtype != MH_C64)
    return fail(h, MH_ERR_DTYPE, "mh_gemm: unknown dtype %d", dtype);
  if (M < 0 || N < 0 || K < 0)
    return fail(h, MH_ERR_SHAPE, "mh_gemm: negative extent");
  if ((M && N && K) && (!A || !B || !C))
    return fail(h, MH_ERR_ARG, "mh_gemm: null operand");
  if (sam != 1 && sak != 1 && M > 1 && K > 1)
    return fail(h, MH_ERR_ARG, "mh_gemm: A needs a unit stride");
  if (sbk != 1 && sbn != 1 && K > 1 && N > 1)
    return fail(h, MH_ERR_ARG, "mh_gemm: B needs a unit stride");
  if (dtype == MH_F64 && (alpha_im != 0.0 || beta_im != 0.0))
    return fail(h, MH_ERR_DTYPE, "mh_gemm: complex scalar with F64 data");
  GemmDesc g{M, N, K, A, sam, sak, B, sbk, sbn, C, scm, scn,
             alpha_re, alpha_im, beta_re, beta_im, conjA};
  if (K == 0) {
    // C = beta C: express as a GEMM with a zero column of A.
    return fail(h, MH_ERR_SHAPE, "mh_gemm: K = 0 is not supported");
  }
  return gemm(h, dtype, prec, g);
}
