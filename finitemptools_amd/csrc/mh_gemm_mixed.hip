// Mixed-precision strided GEMM on the CDNA4 matrix cores.
//
// Operands are "packed" 32-bit words per real component:
//   MH_PREC_F32    : the FP32 value (round to nearest even from FP64);
//   MH_PREC_BF16X3 : a BF16 pair (hi, lo) with hi = bf16(x), lo = bf16(x - hi),
//                    hi in the low 16 bits, lo in the high 16 bits.
// Complex operands store (re, im) word pairs.  Accumulation is FP32 in the
// matrix cores.
//
//   F32    : v_mfma_f32_16x16x4f32, one product per 16x16x4 step.
//   BF16X3 : v_mfma_f32_16x16x32_bf16, three products per 16x16x32 step,
//            hi*hi + hi*lo + lo*hi (the lo*lo term, 2^-16 relative, is dropped).
//
// C is FP64, or (GemmP::c_f32 = 1) "packed" words of the same precision as
// the operands, so one GEMM's output can feed the next without a separate
// conversion pass.
//
// Both operands are staged k-major in LDS ([r][k], r = m for A, n for B), so
// a lane's fragment (4 consecutive k for FP32, 8 for BF16) is one 16-byte
// LDS read.  Threads move pairs of adjacent k; an operand contiguous along r
// is read with two coalesced 4-byte loads per pair, one contiguous along k
// with one 8-byte load.  Tiling, K loop, split-K and the XCD-grouped order
// are those of the FP64 family (mh_gemm.hip).
#include <cstdint>

#include "mh_gemm_common.h"

namespace mh {

typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef short s8_t __attribute__((ext_vector_type(8)));

constexpr int PM_F32 = 1, PM_BF16 = 2;

__device__ __forceinline__ uint32_t pack_word(double x, int pm) {
  const float f = (float)x;
  if (pm == PM_F32) return __builtin_bit_cast(uint32_t, f);
  const __bf16 hi = (__bf16)f;
  const __bf16 lo = (__bf16)(f - (float)hi);
  return (uint32_t)__builtin_bit_cast(uint16_t, hi) |
         ((uint32_t)__builtin_bit_cast(uint16_t, lo) << 16);
}

__device__ __forceinline__ float unpack_word(uint32_t w, int pm) {
  if (pm == PM_F32) return __builtin_bit_cast(float, w);
  return __builtin_bit_cast(float, w << 16) + __builtin_bit_cast(float, w & 0xffff0000u);
}

__device__ __forceinline__ uint32_t pack_word_f(float f, int pm) {
  if (pm == PM_F32) return __builtin_bit_cast(uint32_t, f);
  const __bf16 hi = (__bf16)f;
  const __bf16 lo = (__bf16)(f - (float)hi);
  return (uint32_t)__builtin_bit_cast(uint16_t, hi) |
         ((uint32_t)__builtin_bit_cast(uint16_t, lo) << 16);
}

// dst[i] = pack(src[i]) over n doubles (a complex array is 2n doubles).
template <int PM>
__global__ void k_pack(const double *__restrict__ src, uint32_t *__restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = pack_word(src[i], PM);
}

template <int PM, bool CPLX, int BR, bool RMAJ, int NT = 256>
struct MxStage {
  static constexpr int BK = 32;
  static constexpr int NC = CPLX ? 2 : 1;        // components
  static constexpr int NS = PM == PM_BF16 ? 2 : 1;  // planes per component (hi, lo)
  static constexpr int PITCH = PM == PM_BF16 ? BK + 8 : BK + 4;  // LDS elements per row
  static constexpr int NPAIR = BR * BK / (2 * NT);
  static_assert(NPAIR >= 1, "tile");
  __device__ static void coord(int t, int i, int &r, int &k) {
    const int e = t + NT * i;
    if (RMAJ) {
      r = e % BR;
      k = 2 * (e / BR);
    } else {
      k = 2 * (e % (BK / 2));
      r = e / (BK / 2);
    }
  }
};

// NW = 4 or 8 waves (2 x NW/2 wave grid); eight give two waves per SIMD to
// GEMMs with about one output tile per CU.
template <int PM, bool CPLX, int BM, int BN, bool AMC, bool BKC, bool VEC, bool OUTF32, int NW = 4>
__global__ void __launch_bounds__(64 * NW)
    k_gemm_mx(const GemmP p) {
  constexpr int BK = 32;
  constexpr int WTM = BM / 2, WTN = BN / (NW / 2);
  constexpr int TM = WTM / 16, TN = WTN / 16;
  static_assert(TM >= 1 && TN >= 1, "tile");
  constexpr int NC = CPLX ? 2 : 1;
  constexpr int NS = PM == PM_BF16 ? 2 : 1;
  using SA = MxStage<PM, CPLX, BM, AMC, 64 * NW>;
  using SB = MxStage<PM, CPLX, BN, !BKC, 64 * NW>;
  using E = typename std::conditional<PM == PM_BF16, uint16_t, float>::type;
  constexpr int PITCH = SA::PITCH;
  __shared__ __align__(16) E sA[NC * NS][2][BM * PITCH];
  __shared__ __align__(16) E sB[NC * NS][2][BN * PITCH];

  int tm, tn;
  if (!tile_of_block(p, tm, tn)) return;
  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int lr = lane & 15, lk = lane >> 4;
  const int kbeg = blockIdx.z * p.kchunk;
  const int kend = min(p.K, kbeg + p.kchunk);
  const uint32_t *A = reinterpret_cast<const uint32_t *>(p.A);
  const uint32_t *B = reinterpret_cast<const uint32_t *>(p.B);
  const int64_t a_sr = p.sam, a_sk = p.sak, b_sr = p.sbn, b_sk = p.sbk;
  const uint32_t conj_mask = p.conjA ? (PM == PM_BF16 ? 0x80008000u : 0x80000000u) : 0u;

  // Per-pair clamped row offsets and row validity.
  int64_t aoff[SA::NPAIR], boff[SB::NPAIR];
  bool aok[SA::NPAIR], bok[SB::NPAIR];
#pragma unroll
  for (int i = 0; i < SA::NPAIR; ++i) {
    int r, k;
    SA::coord(tid, i, r, k);
    aok[i] = m0 + r < p.M;
    aoff[i] = (int64_t)min(m0 + r, p.M - 1) * a_sr;
  }
#pragma unroll
  for (int i = 0; i < SB::NPAIR; ++i) {
    int r, k;
    SB::coord(tid, i, r, k);
    bok[i] = n0 + r < p.N;
    boff[i] = (int64_t)min(n0 + r, p.N - 1) * b_sr;
  }

  uint32_t ra[NC][SA::NPAIR][2], rb[NC][SB::NPAIR][2];

  // Global -> registers (branch-free; k clamped into the matrix).
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < SA::NPAIR; ++i) {
      int r, k;
      SA::coord(tid, i, r, k);
      if (VEC && !AMC) {  // pair contiguous along k
        const int kc = min(k0 + k, p.K - 2);
        const int64_t off = aoff[i] + (int64_t)kc;
        if constexpr (CPLX) {
          const uint4 v = *reinterpret_cast<const uint4 *>(A + 2 * off);
          ra[0][i][0] = v.x; ra[1][i][0] = v.y; ra[0][i][1] = v.z; ra[1][i][1] = v.w;
        } else {
          const uint2 v = *reinterpret_cast<const uint2 *>(A + off);
          ra[0][i][0] = v.x; ra[0][i][1] = v.y;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int kc = min(k0 + k + e, p.K - 1);
          const int64_t off = aoff[i] + (int64_t)kc * a_sk;
          if constexpr (CPLX) {
            const uint2 v = *reinterpret_cast<const uint2 *>(A + 2 * off);
            ra[0][i][e] = v.x;
            ra[1][i][e] = v.y;
          } else {
            ra[0][i][e] = A[off];
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < SB::NPAIR; ++i) {
      int r, k;
      SB::coord(tid, i, r, k);
      if (VEC && BKC) {
        const int kc = min(k0 + k, p.K - 2);
        const int64_t off = boff[i] + (int64_t)kc;
        if constexpr (CPLX) {
          const uint4 v = *reinterpret_cast<const uint4 *>(B + 2 * off);
          rb[0][i][0] = v.x; rb[1][i][0] = v.y; rb[0][i][1] = v.z; rb[1][i][1] = v.w;
        } else {
          const uint2 v = *reinterpret_cast<const uint2 *>(B + off);
          rb[0][i][0] = v.x; rb[0][i][1] = v.y;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int kc = min(k0 + k + e, p.K - 1);
          const int64_t off = boff[i] + (int64_t)kc * b_sk;
          if constexpr (CPLX) {
            const uint2 v = *reinterpret_cast<const uint2 *>(B + 2 * off);
            rb[0][i][e] = v.x;
            rb[1][i][e] = v.y;
          } else {
            rb[0][i][e] = B[off];
          }
        }
      }
    }
  };

  // Registers -> LDS (masked), k-major [r][k].
  auto store = [&](int buf, int k0) {
#pragma unroll
    for (int i = 0; i < SA::NPAIR; ++i) {
      int r, k;
      SA::coord(tid, i, r, k);
      const bool ok0 = aok[i] && (k0 + k < kend), ok1 = aok[i] && (k0 + k + 1 < kend);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const uint32_t cm = (c == 1) ? conj_mask : 0u;
        const uint32_t w0 = ok0 ? (ra[c][i][0] ^ cm) : 0u, w1 = ok1 ? (ra[c][i][1] ^ cm) : 0u;
        if constexpr (PM == PM_BF16) {
          *reinterpret_cast<uint32_t *>(&sA[c * 2][buf][r * PITCH + k]) = (w0 & 0xffffu) | (w1 << 16);
          *reinterpret_cast<uint32_t *>(&sA[c * 2 + 1][buf][r * PITCH + k]) = (w0 >> 16) | (w1 & 0xffff0000u);
        } else {
          *reinterpret_cast<uint2 *>(&sA[c][buf][r * PITCH + k]) = make_uint2(w0, w1);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < SB::NPAIR; ++i) {
      int r, k;
      SB::coord(tid, i, r, k);
      const bool ok0 = bok[i] && (k0 + k < kend), ok1 = bok[i] && (k0 + k + 1 < kend);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const uint32_t w0 = ok0 ? rb[c][i][0] : 0u, w1 = ok1 ? rb[c][i][1] : 0u;
        if constexpr (PM == PM_BF16) {
          *reinterpret_cast<uint32_t *>(&sB[c * 2][buf][r * PITCH + k]) = (w0 & 0xffffu) | (w1 << 16);
          *reinterpret_cast<uint32_t *>(&sB[c * 2 + 1][buf][r * PITCH + k]) = (w0 >> 16) | (w1 & 0xffff0000u);
        } else {
          *reinterpret_cast<uint2 *>(&sB[c][buf][r * PITCH + k]) = make_uint2(w0, w1);
        }
      }
    }
  };

  f4_t acc[NC][TM][TN];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[c][i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

  auto compute = [&](int buf) {
    if constexpr (PM == PM_F32) {
#pragma unroll
      for (int kb = 0; kb < BK; kb += 16) {
        f4_t a[NC][TM], b[NC][TN];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
#pragma unroll
          for (int i = 0; i < TM; ++i)
            a[c][i] = *reinterpret_cast<const f4_t *>(&sA[c][buf][(wm * WTM + i * 16 + lr) * PITCH + kb + 4 * lk]);
#pragma unroll
          for (int j = 0; j < TN; ++j)
            b[c][j] = *reinterpret_cast<const f4_t *>(&sB[c][buf][(wn * WTN + j * 16 + lr) * PITCH + kb + 4 * lk]);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) {
              if constexpr (CPLX) {
                acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0][i][s], b[0][j][s], acc[0][i][j], 0, 0, 0);
                acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(-a[1][i][s], b[1][j][s], acc[0][i][j], 0, 0, 0);
                acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0][i][s], b[1][j][s], acc[1][i][j], 0, 0, 0);
                acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1][i][s], b[0][j][s], acc[1][i][j], 0, 0, 0);
              } else {
                acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0][i][s], b[0][j][s], acc[0][i][j], 0, 0, 0);
              }
            }
      }
    } else {
      bf8_t ah[NC][TM], al[NC][TM], bh[NC][TN], bl[NC][TN];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          const int o = (wm * WTM + i * 16 + lr) * PITCH + 8 * lk;
          ah[c][i] = *reinterpret_cast<const bf8_t *>(&sA[2 * c][buf][o]);
          al[c][i] = *reinterpret_cast<const bf8_t *>(&sA[2 * c + 1][buf][o]);
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const int o = (wn * WTN + j * 16 + lr) * PITCH + 8 * lk;
          bh[c][j] = *reinterpret_cast<const bf8_t *>(&sB[2 * c][buf][o]);
          bl[c][j] = *reinterpret_cast<const bf8_t *>(&sB[2 * c + 1][buf][o]);
        }
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[0][i], bh[0][j], acc[0][i][j], 0, 0, 0);
          acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[0][i], bl[0][j], acc[0][i][j], 0, 0, 0);
          acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[0][i], bh[0][j], acc[0][i][j], 0, 0, 0);
          if constexpr (CPLX) {
            // re -= im(A) im(B): negate the A imaginary fragments (sign bits)
            const bf8_t nah = __builtin_bit_cast(bf8_t, __builtin_bit_cast(s8_t, ah[1][i]) ^ (short)0x8000);
            const bf8_t nal = __builtin_bit_cast(bf8_t, __builtin_bit_cast(s8_t, al[1][i]) ^ (short)0x8000);
            acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nah, bh[1][j], acc[0][i][j], 0, 0, 0);
            acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nah, bl[1][j], acc[0][i][j], 0, 0, 0);
            acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nal, bh[1][j], acc[0][i][j], 0, 0, 0);
            acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[0][i], bh[1][j], acc[1][i][j], 0, 0, 0);
            acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[0][i], bl[1][j], acc[1][i][j], 0, 0, 0);
            acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[0][i], bh[1][j], acc[1][i][j], 0, 0, 0);
            acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[1][i], bh[0][j], acc[1][i][j], 0, 0, 0);
            acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[1][i], bl[0][j], acc[1][i][j], 0, 0, 0);
            acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[1][i], bh[0][j], acc[1][i][j], 0, 0, 0);
          }
        }
    }
  };

  const int nk = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
  if (nk > 0) {
    load(kbeg);
    store(0, kbeg);
    __syncthreads();
    for (int kt = 0; kt + 1 < nk; ++kt) {
      const int k1 = kbeg + (kt + 1) * BK;
      load(k1);
      compute(kt & 1);
      store((kt + 1) & 1, k1);
      __syncthreads();
    }
    compute((nk - 1) & 1);
  }

  // f32 MFMA accumulator layout: register r of lane l is C[4*(l/16) + r][l%16].
  const bool split = gridDim.z > 1;
  const bool has_beta = (p.beta_re != 0.0 || p.beta_im != 0.0);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int n = n0 + wn * WTN + j * 16 + lr;
      if (n >= p.N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * WTM + i * 16 + 4 * lk + r;
        if (m >= p.M) continue;
        const int64_t off = (int64_t)m * p.scm + (int64_t)n * p.scn;
        const double xr = (double)acc[0][i][j][r];
        const double xi = CPLX ? (double)acc[NC - 1][i][j][r] : 0.0;
        double vr = p.alpha_re * xr - p.alpha_im * xi;
        double vi = p.alpha_re * xi + p.alpha_im * xr;
        if constexpr (OUTF32) {
          // packed output words of this family's precision
          uint32_t *cw = reinterpret_cast<uint32_t *>(p.C);
          const int64_t ow = CPLX ? 2 * off : off;
          if (has_beta) {
            const double orr = unpack_word(cw[ow], PM);
            const double oi = CPLX ? unpack_word(cw[ow + 1], PM) : 0.0;
            vr += p.beta_re * orr - p.beta_im * oi;
            vi += p.beta_re * oi + p.beta_im * orr;
          }
          if (split) {  // F32 family only (see cand_ok_mx)
            atomicAdd(reinterpret_cast<float *>(cw + ow), (float)vr);
            if (CPLX) atomicAdd(reinterpret_cast<float *>(cw + ow + 1), (float)vi);
          } else {
            cw[ow] = pack_word_f((float)vr, PM);
            if (CPLX) cw[ow + 1] = pack_word_f((float)vi, PM);
          }
        } else {
          if constexpr (CPLX) {
            double2 *c = reinterpret_cast<double2 *>(p.C) + off;
            if (split) {
              atomicAdd(&c->x, vr);
              atomicAdd(&c->y, vi);
            } else {
              if (has_beta) {
                const double2 o = *c;
                vr += p.beta_re * o.x - p.beta_im * o.y;
                vi += p.beta_re * o.y + p.beta_im * o.x;
              }
              *c = make_double2(vr, vi);
            }
          } else {
            double *c = p.C + off;
            if (split) atomicAdd(c, vr);
            else *c = has_beta ? vr + p.beta_re * (*c) : vr;
          }
        }
      }
    }
}

// ------------------------------------------------------------------ dispatch

static const Cand kMxCands[] = {
    {64, 64, 32, 1, 0}, {128, 64, 32, 1, 0}, {64, 128, 32, 1, 0}, {128, 128, 32, 1, 0},
    {64, 64, 32, 2, 0}, {128, 64, 32, 2, 0}, {128, 128, 32, 2, 0},
    {64, 64, 32, 1, 8}, {128, 64, 32, 1, 8}, {128, 128, 32, 1, 8},
    // eight waves per block
    {64, 64, 32, 1, 0, 8}, {64, 64, 32, 2, 0, 8}, {128, 64, 32, 1, 0, 8}, {128, 64, 32, 2, 0, 8},
    {64, 64, 32, 1, 8, 8},
};
constexpr int kNMx = sizeof(kMxCands) / sizeof(kMxCands[0]);

template <int PM, bool CPLX, bool AMC, bool BKC, bool VEC>
static void launch_mx_cfg(const Cand &c, GemmP p, hipStream_t s) {
  dim3 grid = setup_grid(c, p);
#define MH_L(BM, BN)                                                                      \
  do {                                                                                    \
    if (p.c_f32)                                                                          \
      hipLaunchKernelGGL((k_gemm_mx<PM, CPLX, BM, BN, AMC, BKC, VEC, true>), grid, dim3(256), 0, s, p); \
    else                                                                                  \
      hipLaunchKernelGGL((k_gemm_mx<PM, CPLX, BM, BN, AMC, BKC, VEC, false>), grid, dim3(256), 0, s, p); \
  } while (0)
#define MH_L8(BM, BN)                                                                     \
  do {                                                                                    \
    if (p.c_f32)                                                                          \
      hipLaunchKernelGGL((k_gemm_mx<PM, CPLX, BM, BN, AMC, BKC, VEC, true, 8>), grid, dim3(512), 0, s, p); \
    else                                                                                  \
      hipLaunchKernelGGL((k_gemm_mx<PM, CPLX, BM, BN, AMC, BKC, VEC, false, 8>), grid, dim3(512), 0, s, p); \
  } while (0)
  if (c.waves == 8) {
    if (c.bm == 128) MH_L8(128, 64);
    else MH_L8(64, 64);
  }
  else if (c.bm == 64 && c.bn == 64) MH_L(64, 64);
  else if (c.bm == 128 && c.bn == 64) MH_L(128, 64);
  else if (c.bm == 64 && c.bn == 128) MH_L(64, 128);
  else if constexpr (!CPLX) MH_L(128, 128);
#undef MH_L
#undef MH_L8
}

// 8-byte pair loads along k need an even K, an even stride along rows and an
// 8-byte (complex 16-byte) aligned base; only k-major operands use them.
static bool mx_vec_ok(const GemmP &p, bool amc, bool bkc, bool cplx) {
  const uintptr_t al = cplx ? 15 : 7;
  const bool a = amc || (((uintptr_t)p.A & al) == 0 && p.K % 2 == 0 && (p.sam % 2 == 0 || p.M == 1));
  const bool b = !bkc || (((uintptr_t)p.B & al) == 0 && p.K % 2 == 0 && (p.sbn % 2 == 0 || p.N == 1));
  return a && b;
}

template <int PM, bool CPLX, bool VEC>
static void launch_mx_layout(const Cand &c, const GemmP &p, hipStream_t s, bool amc, bool bkc) {
  if (amc && bkc) launch_mx_cfg<PM, CPLX, true, true, VEC>(c, p, s);
  else if (amc) launch_mx_cfg<PM, CPLX, true, false, VEC>(c, p, s);
  else if (bkc) launch_mx_cfg<PM, CPLX, false, true, VEC>(c, p, s);
  else launch_mx_cfg<PM, CPLX, false, false, VEC>(c, p, s);
}

template <int PM>
static void launch_mx(const Cand &c, const GemmP &p, hipStream_t s, bool cplx, bool amc, bool bkc) {
  const bool vec = mx_vec_ok(p, amc, bkc, cplx);
  if (cplx) {
    if (vec) launch_mx_layout<PM, true, true>(c, p, s, amc, bkc);
    else launch_mx_layout<PM, true, false>(c, p, s, amc, bkc);
  } else {
    if (vec) launch_mx_layout<PM, false, true>(c, p, s, amc, bkc);
    else launch_mx_layout<PM, false, false>(c, p, s, amc, bkc);
  }
}

static bool cand_ok_mx_f32(const Cand &c, const GemmP &p, bool cplx) {
  if (c.split > 1 && (p.beta_re != 0.0 || p.beta_im != 0.0)) return false;
  if (c.split > 1 && p.K < 512 * c.split) return false;
  if (cplx && c.bm * c.bn > 64 * 128) return false;
  return true;
}

// BF16-pair output words cannot be accumulated atomically: no split-K there.
static bool cand_ok_mx_bf16(const Cand &c, const GemmP &p, bool cplx) {
  if (c.split > 1 && p.c_f32) return false;
  return cand_ok_mx_f32(c, p, cplx);
}

static const GemmFamily kF32Family = {"f32", kMxCands, kNMx, 32, cand_ok_mx_f32, launch_mx<PM_F32>};
static const GemmFamily kBF16Family = {"bf16x3", kMxCands, kNMx, 32, cand_ok_mx_bf16, launch_mx<PM_BF16>};

const GemmFamily &mixed_family(int prec) {
  return prec == MH_PREC_F32 ? kF32Family : kBF16Family;
}

int pack_operand(mh_handle h, int prec, const double *src, int64_t ndouble, uint32_t *dst) {
  if (ndouble <= 0) return MH_OK;
  dim3 grid((unsigned)((ndouble + 255) / 256));
  if (prec == MH_PREC_F32)
    hipLaunchKernelGGL(k_pack<PM_F32>, grid, dim3(256), 0, h->stream, src, dst, ndouble);
  else
    hipLaunchKernelGGL(k_pack<PM_BF16>, grid, dim3(256), 0, h->stream, src, dst, ndouble);
  return check_launch(h, "pack");
}

// out[p, j, q] = sum_i in[p, i, q] Mt[i, j] on packed words (FP32 arithmetic);
// the FP64 Mt chunk is rounded to FP32 in LDS.  Same thread mapping and
// zero-row skipping as the FP64 mode product.
template <int PM, bool CPLX, int JC>
__global__ void __launch_bounds__(256)
    k_mode_p(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
             const double *__restrict__ Mt, int P, int ni, int nj, int64_t PQ, int64_t sIq,
             int64_t sOq) {
  constexpr int EW = CPLX ? 2 : 1;
  extern __shared__ float smem_f[];
  float *sM = smem_f;                              // [ni][JC][EW]
  int *nz = reinterpret_cast<int *>(sM + ni * JC * EW);
  const int tid = threadIdx.x;
  const int j0 = blockIdx.y * JC;
  const int jn = min(JC, nj - j0);
  for (int t = tid; t < ni * JC; t += 256) {
    const int i = t / JC, jj = t % JC;
#pragma unroll
    for (int c = 0; c < EW; ++c)
      sM[t * EW + c] = (jj < jn) ? (float)Mt[(i + (int64_t)ni * (j0 + jj)) * EW + c] : 0.f;
  }
  __syncthreads();
  for (int i = tid; i < ni; i += 256) {
    int f = 0;
    for (int q = 0; q < JC * EW; ++q) f |= sM[i * JC * EW + q] != 0.f;
    nz[i] = f;
  }
  __syncthreads();
  const int64_t idx = (int64_t)blockIdx.x * 256 + tid;
  if (idx >= PQ) return;
  const int64_t pp = idx % P, q = idx / P;
  float ar[JC], ai[JC];
#pragma unroll
  for (int jj = 0; jj < JC; ++jj) ar[jj] = ai[jj] = 0.f;
  const int64_t base = pp + q * sIq;
  for (int i = 0; i < ni; ++i) {
    if (!nz[i]) continue;
    const int64_t o = (base + (int64_t)i * P) * EW;
    const float vr = unpack_word(in[o], PM);
    const float vi = CPLX ? unpack_word(in[o + 1], PM) : 0.f;
#pragma unroll
    for (int jj = 0; jj < JC; ++jj) {
      const float mr = sM[(i * JC + jj) * EW];
      if constexpr (CPLX) {
        const float mi = sM[(i * JC + jj) * EW + 1];
        ar[jj] += vr * mr - vi * mi;
        ai[jj] += vr * mi + vi * mr;
      } else {
        ar[jj] += vr * mr;
      }
    }
  }
  const int64_t ob = pp + q * sOq;
#pragma unroll
  for (int jj = 0; jj < JC; ++jj)
    if (jj < jn) {
      const int64_t o = (ob + (int64_t)(j0 + jj) * P) * EW;
      out[o] = pack_word_f(ar[jj], PM);
      if (CPLX) out[o + 1] = pack_word_f(ai[jj], PM);
    }
}

template <int PM, bool CPLX, int JC>
static void launch_mode_p(mh_handle h, int P, int ni, int nj, int64_t Q, const uint32_t *in,
                          int64_t sIq, const double *Mt, uint32_t *out, int64_t sOq) {
  const int64_t PQ = (int64_t)P * Q;
  dim3 grid((unsigned)((PQ + 255) / 256), (nj + JC - 1) / JC);
  const size_t smem = sizeof(float) * ni * JC * (CPLX ? 2 : 1) + sizeof(int) * ni;
  hipLaunchKernelGGL((k_mode_p<PM, CPLX, JC>), grid, dim3(256), smem, h->stream, in, out, Mt, P,
                     ni, nj, PQ, sIq, sOq);
}

template <int PM, bool CPLX>
static void mode_p_dispatch(mh_handle h, int P, int ni, int nj, int64_t Q, const uint32_t *in,
                            int64_t sIq, const double *Mt, uint32_t *out, int64_t sOq) {
  if (nj <= 8) launch_mode_p<PM, CPLX, 8>(h, P, ni, nj, Q, in, sIq, Mt, out, sOq);
  else if (nj <= 16) launch_mode_p<PM, CPLX, 16>(h, P, ni, nj, Q, in, sIq, Mt, out, sOq);
  else launch_mode_p<PM, CPLX, 24>(h, P, ni, nj, Q, in, sIq, Mt, out, sOq);
}

int mode_product_packed(mh_handle h, int dtype, int prec, int P, int ni, int nj, int Q,
                        const void *in, int64_t sIq, const void *Mt, void *out, int64_t sOq) {
  if ((int64_t)P * Q <= 0) return MH_OK;
  const bool c = dtype == MH_C64;
  double flops = 2.0 * P * (double)Q * ni * nj * (c ? 4.0 : 1.0);
  int slot;
  prof_begin(h, MH_K_WSTAGE, flops, &slot);
  const uint32_t *pin = (const uint32_t *)in;
  uint32_t *pout = (uint32_t *)out;
  const double *m = (const double *)Mt;
  if (prec == MH_PREC_F32) {
    if (c) mode_p_dispatch<PM_F32, true>(h, P, ni, nj, Q, pin, sIq, m, pout, sOq);
    else mode_p_dispatch<PM_F32, false>(h, P, ni, nj, Q, pin, sIq, m, pout, sOq);
  } else {
    if (c) mode_p_dispatch<PM_BF16, true>(h, P, ni, nj, Q, pin, sIq, m, pout, sOq);
    else mode_p_dispatch<PM_BF16, false>(h, P, ni, nj, Q, pin, sIq, m, pout, sOq);
  }
  prof_end(h, slot);
  return check_launch(h, "mode_product_packed");
}

// ------------------------------------------------ fused first H_eff step, packed operands
//
// Packed-operand form of the fused FP64 step (k_heff_step1, mh_gemm.hip), for
// FP64 real operators:
//   T1[a, w, s, b'] = sum_a' L[a, w, a'] x[a', s, b']    (MFMA, FP32 accumulators)
//   T2[a, j, b']    = sum_i T1[a, i, b'] Mt[i, j]         (FP32, epilogue)
// with L, x and T2 packed words of the family's precision (i = w + WL*s,
// j = s' + DS*w').  T1 never leaves the block and is not rounded to packed
// words (the unfused path rounds it once more).  A block owns a = a0..a0+31
// for every channel w (BM = 32*WL rows: row r is a0 + r%32 of channel r/32)
// and 64 columns c = s + DS*b'; four waves 2 x 2, BK = 32, double-buffered
// k-major LDS staging as in k_gemm_mx.  The epilogue stages the 64-column
// tile in LDS ([column][row], pitch BM + 4, so an accumulator's four rows are
// one conflict-free 16-byte write), then each thread forms T2 for 2-4 (a, b')
// pairs with Mt read from LDS as FP32.  Needs Dl even (8-byte x pairs).
template <int PM, int WL, int DS, int WR>
__global__ void __launch_bounds__(256)
    k_heff_step1_mx(const uint32_t *__restrict__ L, const uint32_t *__restrict__ X,
                    const double *__restrict__ Mt, uint32_t *__restrict__ T2, int Dl, int Dr,
                    int tiles_a, int tiles_b, int xcd_chunk, int group) {
  constexpr int BA = 32, BM = BA * WL, BN = 64, BK = 32;
  constexpr int NI = WL * DS, NJ = DS * WR;
  constexpr int WTM = BM / 2, WTN = BN / 2, TM = WTM / 16, TN = WTN / 16;
  constexpr int BB = BN / DS;  // b' values per block
  constexpr int NPQ = BA * BB / 256;  // (a, b') pairs per thread
  static_assert(BN % DS == 0 && (BA * BB) % 256 == 0 && WTM % 16 == 0, "tile");
  using SA = MxStage<PM, false, BM, true>;   // L rows (contiguous along a), [r][k]
  using SB = MxStage<PM, false, BN, false>;  // x columns (contiguous along a'), [c][k]
  constexpr int NS = PM == PM_BF16 ? 2 : 1;
  using E = typename std::conditional<PM == PM_BF16, uint16_t, float>::type;
  constexpr int PITCH = SA::PITCH;
  constexpr int SZA = BM * PITCH, SZB = BN * PITCH;  // elements per plane and buffer
  constexpr int STAGE_BYTES = NS * 2 * (SZA + SZB) * (int)sizeof(E);
  constexpr int CP = BM + 4;
  constexpr int EPI_BYTES = BN * CP * (int)sizeof(float);
  constexpr int SMEM = STAGE_BYTES > EPI_BYTES ? STAGE_BYTES : EPI_BYTES;
  __shared__ __align__(16) char smem[SMEM];
  __shared__ __align__(16) float sM[NI * NJ];
  E *const sA = reinterpret_cast<E *>(smem);  // [plane][buf][BM * PITCH]
  E *const sB = sA + NS * 2 * SZA;            // [plane][buf][BN * PITCH]

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int local = blockIdx.x >> 3;
  const int t = (blockIdx.x & 7) * xcd_chunk + local;
  if (local >= xcd_chunk || t >= tiles_a * tiles_b) return;
  const int span = group * tiles_b, first = (t / span) * group;
  const int gsz = min(tiles_a - first, group), wt = t % span;
  const int a0 = (first + wt % gsz) * BA, c0 = (wt / gsz) * BN;
  const int K = Dl, NC = DS * Dr;
  const int64_t ldL = (int64_t)Dl * WL;
  const int lr = lane & 15, lk = lane >> 4;

  for (int e = tid; e < NI * NJ; e += 256) sM[e] = (float)Mt[e];

  int64_t aoff[SA::NPAIR], boff[SB::NPAIR];
  bool aok[SA::NPAIR], bok[SB::NPAIR];
#pragma unroll
  for (int i = 0; i < SA::NPAIR; ++i) {
    int r, k;
    SA::coord(tid, i, r, k);
    const int a = a0 + r % BA, w = r / BA;
    aok[i] = a < Dl;
    aoff[i] = (int64_t)min(a, Dl - 1) + (int64_t)Dl * w;
  }
#pragma unroll
  for (int i = 0; i < SB::NPAIR; ++i) {
    int r, k;
    SB::coord(tid, i, r, k);
    const int c = c0 + r;
    bok[i] = c < NC;
    boff[i] = (int64_t)min(c, NC - 1) * Dl;
  }

  uint32_t ra[SA::NPAIR][2], rb[SB::NPAIR][2];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < SA::NPAIR; ++i) {
      int r, k;
      SA::coord(tid, i, r, k);
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int kc = min(k0 + k + e, K - 1);
        ra[i][e] = L[aoff[i] + (int64_t)kc * ldL];
      }
    }
#pragma unroll
    for (int i = 0; i < SB::NPAIR; ++i) {
      int r, k;
      SB::coord(tid, i, r, k);
      const int kc = min(k0 + k, K - 2);  // even: k0, k even and K even
      const uint2 v = *reinterpret_cast<const uint2 *>(X + boff[i] + kc);
      rb[i][0] = v.x;
      rb[i][1] = v.y;
    }
  };
  auto put = [&](E *dst, int buf, int sz, int off, uint32_t w0, uint32_t w1) {
    if constexpr (PM == PM_BF16) {
      *reinterpret_cast<uint32_t *>(&dst[buf * sz + off]) = (w0 & 0xffffu) | (w1 << 16);
      *reinterpret_cast<uint32_t *>(&dst[(2 + buf) * sz + off]) = (w0 >> 16) | (w1 & 0xffff0000u);
    } else {
      *reinterpret_cast<uint2 *>(&dst[buf * sz + off]) = make_uint2(w0, w1);
    }
  };
  auto store = [&](int buf, int k0) {
#pragma unroll
    for (int i = 0; i < SA::NPAIR; ++i) {
      int r, k;
      SA::coord(tid, i, r, k);
      const uint32_t w0 = (aok[i] && k0 + k < K) ? ra[i][0] : 0u;
      const uint32_t w1 = (aok[i] && k0 + k + 1 < K) ? ra[i][1] : 0u;
      put(sA, buf, SZA, r * PITCH + k, w0, w1);
    }
#pragma unroll
    for (int i = 0; i < SB::NPAIR; ++i) {
      int r, k;
      SB::coord(tid, i, r, k);
      const uint32_t w0 = (bok[i] && k0 + k < K) ? rb[i][0] : 0u;
      const uint32_t w1 = (bok[i] && k0 + k + 1 < K) ? rb[i][1] : 0u;
      put(sB, buf, SZB, r * PITCH + k, w0, w1);
    }
  };

  f4_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

  auto compute = [&](int buf) {
    const E *A = sA + buf * SZA;
    const E *B = sB + buf * SZB;
    if constexpr (PM == PM_F32) {
#pragma unroll
      for (int kb = 0; kb < BK; kb += 16) {
        f4_t a[TM], b[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i)
          a[i] = *reinterpret_cast<const f4_t *>(&A[(wm * WTM + i * 16 + lr) * PITCH + kb + 4 * lk]);
#pragma unroll
        for (int j = 0; j < TN; ++j)
          b[j] = *reinterpret_cast<const f4_t *>(&B[(wn * WTN + j * 16 + lr) * PITCH + kb + 4 * lk]);
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][s], b[j][s], acc[i][j], 0, 0, 0);
      }
    } else {
      bf8_t ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int o = (wm * WTM + i * 16 + lr) * PITCH + 8 * lk;
        ah[i] = *reinterpret_cast<const bf8_t *>(&A[o]);
        al[i] = *reinterpret_cast<const bf8_t *>(&A[2 * SZA + o]);
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int o = (wn * WTN + j * 16 + lr) * PITCH + 8 * lk;
        bh[j] = *reinterpret_cast<const bf8_t *>(&B[o]);
        bl[j] = *reinterpret_cast<const bf8_t *>(&B[2 * SZB + o]);
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
  };

  const int nk = (K + BK - 1) / BK;
  load(0);
  store(0, 0);
  __syncthreads();
  for (int kt = 0; kt + 1 < nk; ++kt) {
    const int k1 = (kt + 1) * BK;
    load(k1);
    compute(kt & 1);
    store((kt + 1) & 1, k1);
    __syncthreads();
  }
  compute((nk - 1) & 1);

  // Epilogue.  f32 accumulator layout: register r of lane l is
  // C[4*(l/16) + r][l%16], so a lane's four values are four adjacent rows.
  __syncthreads();  // every wave is done with the staging buffers
  float *const Cs = reinterpret_cast<float *>(smem);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int m = wm * WTM + i * 16 + 4 * lk;
      const int n = wn * WTN + j * 16 + lr;
      *reinterpret_cast<f4_t *>(&Cs[n * CP + m]) = acc[i][j];
    }
  __syncthreads();
  const int bq0 = c0 / DS;  // first b' of the block
  float v[NPQ][NI];
#pragma unroll
  for (int q = 0; q < NPQ; ++q) {
    const int pr = tid + 256 * q;
    const int al = pr % BA, bl = pr / BA;
#pragma unroll
    for (int s = 0; s < DS; ++s)
#pragma unroll
      for (int w = 0; w < WL; ++w) v[q][w + WL * s] = Cs[(bl * DS + s) * CP + al + BA * w];
  }
  const int a = a0 + (tid % BA);
#pragma unroll 1
  for (int j = 0; j < NJ; ++j) {
    float o[NPQ];
#pragma unroll
    for (int q = 0; q < NPQ; ++q) o[q] = 0.f;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const float m = sM[i + NI * j];
#pragma unroll
      for (int q = 0; q < NPQ; ++q) o[q] = __builtin_fmaf(v[q][i], m, o[q]);
    }
#pragma unroll
    for (int q = 0; q < NPQ; ++q) {
      const int bq = bq0 + (tid + 256 * q) / BA;
      if (a < Dl && bq < Dr) T2[a + (int64_t)Dl * (j + (int64_t)NJ * bq)] = pack_word_f(o[q], PM);
    }
  }
}

template <int PM, int WL, int DS>
static void launch_step1_mx(int Dl, int Dr, const uint32_t *L, const uint32_t *X,
                            const double *Mt, uint32_t *T2, hipStream_t s) {
  const int tiles_a = (Dl + 31) / 32, tiles_b = (DS * Dr + 63) / 64;
  const int chunk = (tiles_a * tiles_b + 7) / 8;
  const int group = min(tiles_a, 8);
  hipLaunchKernelGGL((k_heff_step1_mx<PM, WL, DS, WL>), dim3(8 * chunk), dim3(256), 0, s, L, X,
                     Mt, T2, Dl, Dr, tiles_a, tiles_b, chunk, group);
}

// Packed T2 = mode(L x) in one launch (FP64 real operators, packed L and x,
// wl = wr in {3,4,5,6,8}, physical extent 2 or 4, Dl even);
// MH_ERR_UNSUPPORTED otherwise (the caller runs GEMM + packed mode product).
int heff_step1_fused_mx(mh_handle h, int prec, int Dl, int Dr, int wl, int ds, int wr,
                        const void *L, const void *x, const void *Mt, void *T2) {
  auto al8 = [](const void *q) { return ((uintptr_t)q & 7) == 0; };
  if (prec != MH_PREC_F32 && prec != MH_PREC_BF16X3) return MH_ERR_UNSUPPORTED;
  if (wl != wr || Dl < 2 || Dl % 2 != 0 || !al8(x)) return MH_ERR_UNSUPPORTED;
  const uint32_t *Lp = (const uint32_t *)L, *Xp = (const uint32_t *)x;
  const double *Mp = (const double *)Mt;
  uint32_t *Tp = (uint32_t *)T2;
  hipStream_t s = h->stream;
  const double flops = 2.0 * Dl * (double)wl * ds * Dr * Dl;
  int slot;
#define MH_S1(W, D)                                                        \
  if (wl == W && ds == D) {                                                \
    prof_begin(h, MH_K_GEMM, flops, &slot);                                \
    if (prec == MH_PREC_F32) launch_step1_mx<PM_F32, W, D>(Dl, Dr, Lp, Xp, Mp, Tp, s);  \
    else launch_step1_mx<PM_BF16, W, D>(Dl, Dr, Lp, Xp, Mp, Tp, s);        \
    prof_end(h, slot);                                                     \
    return check_launch(h, "fused packed H_eff step 1");                   \
  }
  MH_S1(5, 4) MH_S1(5, 2) MH_S1(3, 4) MH_S1(3, 2) MH_S1(4, 4) MH_S1(4, 2)
  MH_S1(6, 4) MH_S1(6, 2) MH_S1(8, 4) MH_S1(8, 2)
#undef MH_S1
  return MH_ERR_UNSUPPORTED;
}

// GEMM on operands that are already packed; c_packed selects packed words
// (1) or FP64 (0) for C.
int gemm_packed(mh_handle h, int dtype, int prec, const GemmDesc &g, int c_packed) {
  if (g.M <= 0 || g.N <= 0) return MH_OK;
  GemmP p = make_params(g);
  p.c_f32 = c_packed;
  return gemm_dispatch(h, mixed_family(prec), p, dtype == MH_C64,
                       prec == MH_PREC_F32 ? "f32" : "bf16x3");
}

// Element span of a strided matrix: the largest offset reached plus one.
static int64_t span(int rows, int cols, int64_t sr, int64_t sc) {
  return (int64_t)(rows - 1) * sr + (int64_t)(cols - 1) * sc + 1;
}

// mh_gemm with a mixed precision: pack both FP64 operands (their whole
// element span, so the strides carry over unchanged) into the handle's
// packing buffer, then run the mixed family with an FP64 C.
int gemm_mixed(mh_handle h, int dtype, int prec, const GemmDesc &g) {
  if (prec != MH_PREC_F32 && prec != MH_PREC_BF16X3)
    return fail(h, MH_ERR_DTYPE, "unknown precision %d", prec);
  const bool cplx = dtype == MH_C64;
  const int64_t ew = cplx ? 2 : 1;
  const int64_t na = span(g.M, g.K, g.sam, g.sak) * ew;
  const int64_t nb = span(g.K, g.N, g.sbk, g.sbn) * ew;
  const size_t abytes = ((size_t)na * 4 + 255) & ~size_t(255);
  const size_t bytes = abytes + (size_t)nb * 4;
  if (h->pack_bytes < bytes) {
    MH_HIP(h, hipStreamSynchronize(h->stream));
    if (h->pack) MH_HIP(h, hipFree(h->pack));
    h->pack = nullptr;
    h->pack_bytes = 0;
    if (hipMalloc(&h->pack, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(h, MH_ERR_OOM, "packing buffer of %zu bytes", bytes);
    }
    h->pack_bytes = bytes;
  }
  uint32_t *pa = (uint32_t *)h->pack;
  uint32_t *pb = (uint32_t *)((char *)h->pack + abytes);
  MH_TRY(pack_operand(h, prec, (const double *)g.A, na, pa));
  MH_TRY(pack_operand(h, prec, (const double *)g.B, nb, pb));
  GemmP p = make_params(g);
  p.A = (const double *)pa;
  p.B = (const double *)pb;
  return gemm_dispatch(h, mixed_family(prec), p, cplx, prec == MH_PREC_F32 ? "f32" : "bf16x3");
}

}  // namespace mh
