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
#include <type_traits>

#include "mh_gemm_common.h"

namespace mh {



// Operand staging.  One operand tile is BR rows ("r" = m for A, n for B) by
// BK values of k.  An operand contiguous along r ("r-major": A m-contiguous,
// B n-contiguous) is staged in LDS as [k][r] with pitch BR + 4; an operand
// contiguous along k ("k-major") as [r][k] with pitch BK + 2.  Either way a
// thread moves pairs of elements adjacent along the contiguous dimension
// (one 16-byte global load when the operand is suitably aligned, one 16-byte
// LDS store always), so the LDS stores are conflict-free for both layouts.
//
// MFMA K order: within each 16-deep slice, sub-step s (0..3) of lane group
// lk = lane/16 uses k = 4*lk + s.  The four sub-steps still cover the slice
// once, A and B agree on the order, and a k-major fragment then reads four
// consecutive k of one row with two 16-byte LDS reads.
template <bool CPLX, int BR, int BK, bool RMAJ, int NT = 256>
struct Stage {
  static constexpr int NP = CPLX ? 2 : 1;
  static constexpr int PITCH = RMAJ ? BR + 4 : BK + 2;
  static constexpr int SIZE = RMAJ ? BK * PITCH : BR * PITCH;
  static constexpr int NPAIR = BR * BK / (2 * NT);
  static_assert(NPAIR >= 1, "tile too small for the thread count");
  __device__ static void coord(int t, int i, int &r, int &k) {
    const int e = t + NT * i;
    if (RMAJ) {
      r = 2 * (e % (BR / 2));
      k = e / (BR / 2);
    } else {
      k = 2 * (e % (BK / 2));
      r = e / (BK / 2);
    }
  }
};

// One kernel for real (CPLX = false) and complex FP64.  Complex operands are
// split into re/im LDS planes at the LDS store.  Wave tiles small enough for
// a third accumulator plane (M3: at most two 16-row tiles, TM * TN * BK <= 64)
// run three real MFMAs per 16x16x4 step (Gauss's three-multiplication form,
// as in zgemm3m):
// P1 += Ar Br, P2 += Ai Bi, P3 += (Ar + Ai)(Br + Bi), and after the K loop
// Re C = P1 - P2, Im C = P3 - P1 - P2.  The error is normwise of the same
// order as the four-product form (Higham, "Stability of a method for
// multiplying complex matrices with three real matrix multiplications",
// SIAM J. Matrix Anal. Appl. 13 (1992)); the imaginary part's bound carries
// (|Ar| + |Ai|)(|Br| + |Bi|) in place of |Ar||Bi| + |Ai||Br|.  Larger wave
// tiles keep the four-product form (Re += Ar Br - Ai Bi, Im += Ar Bi + Ai Br),
// whose two accumulator planes fit their registers without spilling.
//
// Split-K: blockIdx.z selects the K range [z*kchunk, (z+1)*kchunk).  With
// gridDim.z == 2 the partial tiles are added with FP64 atomics onto a C that
// was zeroed beforehand: 0 + a + b == 0 + b + a exactly, so the result does
// not depend on the order the two adds land (only used for beta == 0).
// NW = 4 or 8 waves per block (2 x NW/2 wave grid): eight waves give two
// waves per SIMD to GEMMs with few output tiles (one block per CU).
// KG = 2 (with NW = 8, BK a multiple of 32): the waves form two 2x2 grids;
// grid g computes the 16-deep K slices with (slice index % 2) == g of every
// staged step, so each wave keeps a 2x-larger tile (more MFMAs per LDS
// fragment) while two waves still share each SIMD; the second grid's tile is
// added to the first's through LDS after the K loop (a fixed order).
// PF = 2: the global -> register prefetch runs two K steps ahead (two
// register sets, the K loop unrolled by two), so a tile's loads have two
// compute phases to land instead of one.
template <bool CPLX, int BM, int BN, int BK, bool AMC, bool BKC, bool VEC, int NW = 4, int KG = 1,
          int PF = 1>
__global__ void __launch_bounds__(64 * NW)
    k_gemm(const GemmP p) {
  static_assert(KG == 1 || (KG == 2 && NW == 8 && BK % 32 == 0), "K-grouped waves");
  static_assert(PF == 1 || PF == 2, "prefetch depth");
  constexpr int WGM = 2, WGN = NW / (2 * KG);
  constexpr int WTM = BM / WGM, WTN = BN / WGN;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  constexpr int NP = CPLX ? 2 : 1;
  constexpr int EW = CPLX ? 2 : 1;  // doubles per element in global memory
  using SA = Stage<CPLX, BM, BK, AMC, 64 * NW>;
  using SB = Stage<CPLX, BN, BK, !BKC, 64 * NW>;
  static_assert(TM >= 1 && TN >= 1 && BK % 16 == 0, "tile");
  __shared__ __align__(16) double sA[NP][2][SA::SIZE];
  __shared__ __align__(16) double sB[NP][2][SB::SIZE];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int kgrp = wave / (NW / KG), wv = wave % (NW / KG);
  const int wm = wv % WGM, wn = wv / WGM;
  int tm, tn;
  if (!tile_of_block(p, tm, tn)) return;  // whole block: padding of the grouped grid
  const int m0 = tm * BM, n0 = tn * BN;
  const int lr = lane & 15, lk = lane >> 4;
  const int kbeg = blockIdx.z * p.kchunk;
  const int kend = min(p.K, kbeg + p.kchunk);
  const double csign = p.conjA ? -1.0 : 1.0;
  // operand descriptors: r-stride / k-stride in elements
  const int64_t a_sr = p.sam, a_sk = p.sak, b_sr = p.sbn, b_sk = p.sbk;

  // Per-pair row offsets (clamped into the matrix) and row validity.
  int64_t aoff[SA::NPAIR][2], boff[SB::NPAIR][2];
  bool aok[SA::NPAIR][2], bok[SB::NPAIR][2];
#pragma unroll
  for (int i = 0; i < SA::NPAIR; ++i) {
    int r, k;
    SA::coord(tid, i, r, k);
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int gr = m0 + r + (AMC ? e : 0);
      const int kk = k + (AMC ? 0 : e);
      aok[i][e] = gr < p.M;
      // A 16-byte pair along rows starts at most at row M-2 (M is even when
      // VEC), so both of its elements lie inside the matrix.
      const int rmax = (VEC && AMC && e == 0) ? p.M - 2 : p.M - 1;
      aoff[i][e] = (int64_t)min(gr, rmax) * a_sr + (int64_t)kk * a_sk;
    }
  }
#pragma unroll
  for (int i = 0; i < SB::NPAIR; ++i) {
    int r, k;
    SB::coord(tid, i, r, k);
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int gr = n0 + r + (BKC ? 0 : e);
      const int kk = k + (BKC ? e : 0);
      bok[i][e] = gr < p.N;
      const int rmax = (VEC && !BKC && e == 0) ? p.N - 2 : p.N - 1;
      boff[i][e] = (int64_t)min(gr, rmax) * b_sr + (int64_t)kk * b_sk;
    }
  }

  double ra0[NP][SA::NPAIR][2], rb0[NP][SB::NPAIR][2];
  double ra1[NP][SA::NPAIR][2], rb1[NP][SB::NPAIR][2];

  // Global -> registers, branch-free: k is clamped into [0, K) (a pair along
  // k into [0, K-1)), rows were clamped once above, so every address is
  // inside the matrix; out-of-range values are zeroed at the LDS store.
  auto load = [&](double (&ra)[NP][SA::NPAIR][2], double (&rb)[NP][SB::NPAIR][2], int k0) {
#pragma unroll
    for (int i = 0; i < SA::NPAIR; ++i) {
      int r, k;
      SA::coord(tid, i, r, k);
      if (VEC) {
        const int kc = AMC ? min(k0 + k, p.K - 1) : min(k0 + k, p.K - 2);
        const double *src = p.A + EW * (aoff[i][0] + (int64_t)(kc - k) * a_sk);
        if constexpr (CPLX) {
          const double4 v = *reinterpret_cast<const double4 *>(src);
          ra[0][i][0] = v.x; ra[1][i][0] = v.y; ra[0][i][1] = v.z; ra[1][i][1] = v.w;
        } else {
          const double2 v = *reinterpret_cast<const double2 *>(src);
          ra[0][i][0] = v.x; ra[0][i][1] = v.y;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int kk = k + (AMC ? 0 : e);
          const int kc = min(k0 + kk, p.K - 1);
          const int64_t off = aoff[i][e] + (int64_t)(kc - kk) * a_sk;
          if constexpr (CPLX) {
            const double2 v = reinterpret_cast<const double2 *>(p.A)[off];
            ra[0][i][e] = v.x;
            ra[1][i][e] = v.y;
          } else {
            ra[0][i][e] = p.A[off];
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < SB::NPAIR; ++i) {
      int r, k;
      SB::coord(tid, i, r, k);
      if (VEC) {
        const int kc = BKC ? min(k0 + k, p.K - 2) : min(k0 + k, p.K - 1);
        const double *src = p.B + EW * (boff[i][0] + (int64_t)(kc - k) * b_sk);
        if constexpr (CPLX) {
          const double4 v = *reinterpret_cast<const double4 *>(src);
          rb[0][i][0] = v.x; rb[1][i][0] = v.y; rb[0][i][1] = v.z; rb[1][i][1] = v.w;
        } else {
          const double2 v = *reinterpret_cast<const double2 *>(src);
          rb[0][i][0] = v.x; rb[0][i][1] = v.y;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int kk = k + (BKC ? e : 0);
          const int kc = min(k0 + kk, p.K - 1);
          const int64_t off = boff[i][e] + (int64_t)(kc - kk) * b_sk;
          if constexpr (CPLX) {
            const double2 v = reinterpret_cast<const double2 *>(p.B)[off];
            rb[0][i][e] = v.x;
            rb[1][i][e] = v.y;
          } else {
            rb[0][i][e] = p.B[off];
          }
        }
      }
    }
  };

  // Registers -> LDS (masked: rows outside the matrix and k outside the
  // split's range become zero).
  auto store = [&](const double (&ra)[NP][SA::NPAIR][2], const double (&rb)[NP][SB::NPAIR][2],
                   int buf, int k0) {
#pragma unroll
    for (int i = 0; i < SA::NPAIR; ++i) {
      int r, k;
      SA::coord(tid, i, r, k);
      const bool k0ok = k0 + k < kend, k1ok = k0 + k + (AMC ? 0 : 1) < kend;
      const bool ok0 = aok[i][0] && k0ok, ok1 = aok[i][1] && k1ok;
      const int o = AMC ? k * SA::PITCH + r : r * SA::PITCH + k;
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        const double sg = (q == 1) ? csign : 1.0;
        double2 v;
        v.x = ok0 ? sg * ra[q][i][0] : 0.0;
        v.y = ok1 ? sg * ra[q][i][1] : 0.0;
        *reinterpret_cast<double2 *>(&sA[q][buf][o]) = v;
      }
    }
#pragma unroll
    for (int i = 0; i < SB::NPAIR; ++i) {
      int r, k;
      SB::coord(tid, i, r, k);
      const bool k0ok = k0 + k < kend, k1ok = k0 + k + (BKC ? 1 : 0) < kend;
      const bool ok0 = bok[i][0] && k0ok, ok1 = bok[i][1] && k1ok;
      const int o = BKC ? r * SB::PITCH + k : k * SB::PITCH + r;
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        double2 v;
        v.x = ok0 ? rb[q][i][0] : 0.0;
        v.y = ok1 ? rb[q][i][1] : 0.0;
        *reinterpret_cast<double2 *>(&sB[q][buf][o]) = v;
      }
    }
  };

  constexpr bool M3 = CPLX && TM <= 2 && TM * TN * BK <= 64;
  constexpr int NQ = M3 ? 3 : NP;  // accumulator planes (P1, P2, P3 for M3)
  d4_t acc[NQ][TM][TN];
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[q][i][j] = d4_t{0.0, 0.0, 0.0, 0.0};

  // Fragments of one 16-deep slice: a[q][i][s], b[q][j][s] (sub-step s).
  auto frag = [&](int buf, int kb, double (&a)[NP][TM][4], double (&b)[NP][TN][4]) {
    if constexpr (AMC && !BKC) {
      // both operands staged [k][r]: sub-step by sub-step, A and B together,
      // so the first MFMAs wait only for the first reads
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int q = 0; q < NP; ++q) {
#pragma unroll
          for (int i = 0; i < TM; ++i)
            a[q][i][s] = sA[q][buf][(kb + 4 * lk + s) * SA::PITCH + wm * WTM + i * 16 + lr];
#pragma unroll
          for (int j = 0; j < TN; ++j)
            b[q][j][s] = sB[q][buf][(kb + 4 * lk + s) * SB::PITCH + wn * WTN + j * 16 + lr];
        }
    } else {
#pragma unroll
      for (int q = 0; q < NP; ++q) {
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          const int r = wm * WTM + i * 16 + lr;
          if (AMC) {
#pragma unroll
            for (int s = 0; s < 4; ++s) a[q][i][s] = sA[q][buf][(kb + 4 * lk + s) * SA::PITCH + r];
          } else {
            const double2 *src = reinterpret_cast<const double2 *>(&sA[q][buf][r * SA::PITCH + kb + 4 * lk]);
            const double2 u = src[0], v = src[1];
            a[q][i][0] = u.x; a[q][i][1] = u.y; a[q][i][2] = v.x; a[q][i][3] = v.y;
          }
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const int r = wn * WTN + j * 16 + lr;
          if (!BKC) {
#pragma unroll
            for (int s = 0; s < 4; ++s) b[q][j][s] = sB[q][buf][(kb + 4 * lk + s) * SB::PITCH + r];
          } else {
            const double2 *src = reinterpret_cast<const double2 *>(&sB[q][buf][r * SB::PITCH + kb + 4 * lk]);
            const double2 u = src[0], v = src[1];
            b[q][j][0] = u.x; b[q][j][1] = u.y; b[q][j][2] = v.x; b[q][j][3] = v.y;
          }
        }
      }
    }
  };
  auto mma = [&](const double (&a)[NP][TM][4], const double (&b)[NP][TN][4]) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if constexpr (M3) {
        double sa[TM], sb[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) sa[i] = a[0][i][s] + a[1][i][s];
#pragma unroll
        for (int j = 0; j < TN; ++j) sb[j] = b[0][j][s] + b[1][j][s];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0][i][s], b[0][j][s], acc[0][i][j], 0, 0, 0);
            acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1][i][s], b[1][j][s], acc[1][i][j], 0, 0, 0);
            acc[2][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(sa[i], sb[j], acc[2][i][j], 0, 0, 0);
          }
      } else if constexpr (CPLX) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            const double nai = -a[1][i][s];
            acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0][i][s], b[0][j][s], acc[0][i][j], 0, 0, 0);
            acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0][i][s], b[1][j][s], acc[1][i][j], 0, 0, 0);
            acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, b[1][j][s], acc[0][i][j], 0, 0, 0);
            acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1][i][s], b[0][j][s], acc[1][i][j], 0, 0, 0);
          }
      } else {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
            acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0][i][s], b[0][j][s], acc[0][i][j], 0, 0, 0);
      }
    }
  };
  // One K step from LDS buffer buf.  A wave runs the slices of its K group
  // (all of them, or every other one for KG = 2), software-pipelined: the
  // fragments of slice t+1 are read before the MFMAs of slice t are issued,
  // and the scheduling barrier keeps the compiler from sinking those reads
  // back to their first use, so LDS latency hides behind the MFMAs.
  auto compute = [&](int buf) {
    constexpr int NSL = BK / 16 / KG;  // slices per wave and K step
    double fa[2][NP][TM][4], fb[2][NP][TN][4];
    frag(buf, (KG == 2 ? kgrp : 0) * 16, fa[0], fb[0]);
#pragma unroll
    for (int t = 0; t < NSL; ++t) {
      if (t + 1 < NSL) frag(buf, ((t + 1) * KG + (KG == 2 ? kgrp : 0)) * 16, fa[(t + 1) & 1], fb[(t + 1) & 1]);
      __builtin_amdgcn_sched_barrier(0);
      mma(fa[t & 1], fb[t & 1]);
    }
  };

  const int nk = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
  if (nk > 0 && PF == 1) {
    load(ra0, rb0, kbeg);
    store(ra0, rb0, 0, kbeg);
    __syncthreads();
    // Steady state: prefetch step kt+1 into registers, run step kt from LDS,
    // then write the prefetched tile to the other LDS buffer.  The last
    // compute is peeled so the loop body has no branches.
    for (int kt = 0; kt + 1 < nk; ++kt) {
      const int k1 = kbeg + (kt + 1) * BK;
      load(ra0, rb0, k1);
      compute(kt & 1);
      store(ra0, rb0, (kt + 1) & 1, k1);
      __syncthreads();
    }
    compute((nk - 1) & 1);
  } else if (nk > 0) {
    // Two register sets: at the top of each pair of steps LDS buffer 0 holds
    // step kt and set 1 holds step kt+1 (in flight).  Loads of steps past the
    // end read clamped addresses and are never stored.
    load(ra0, rb0, kbeg);
    store(ra0, rb0, 0, kbeg);
    load(ra1, rb1, kbeg + BK);
    __syncthreads();
    int kt = 0;
    for (; kt + 2 < nk; kt += 2) {
      load(ra0, rb0, kbeg + (kt + 2) * BK);
      compute(0);
      store(ra1, rb1, 1, kbeg + (kt + 1) * BK);
      __syncthreads();
      load(ra1, rb1, kbeg + (kt + 3) * BK);
      compute(1);
      store(ra0, rb0, 0, kbeg + (kt + 2) * BK);
      __syncthreads();
    }
    compute(0);
    if (kt + 1 < nk) {
      store(ra1, rb1, 1, kbeg + (kt + 1) * BK);
      __syncthreads();
      compute(1);
    }
  }

  if constexpr (M3) {
    // (P1, P2, P3) -> (Re, Im) in planes 0 and 1
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double p1 = acc[0][i][j][r], p2 = acc[1][i][j][r];
          acc[0][i][j][r] = p1 - p2;
          acc[1][i][j][r] = acc[2][i][j][r] - p1 - p2;
        }
  }

  if constexpr (KG == 2) {
    // grid 1 hands its tile to grid 0 through LDS (the staging buffers are
    // free once every wave is past its last compute)
    constexpr int NACC = NP * TM * TN * 4;
    static_assert((size_t)NACC * 256 <= sizeof(sA) / sizeof(double), "LDS for the K-group reduction");
    double *red = &sA[0][0][0];
    __syncthreads();
    if (kgrp == 1) {
#pragma unroll
      for (int q = 0; q < NP; ++q)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              red[((((q * TM + i) * TN + j) * 4 + r) * 4 + wv) * 64 + lane] = acc[q][i][j][r];
    }
    __syncthreads();
    if (kgrp == 1) return;
#pragma unroll
    for (int q = 0; q < NP; ++q)
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            acc[q][i][j][r] += red[((((q * TM + i) * TN + j) * 4 + r) * 4 + wv) * 64 + lane];
  }

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
        const int m = m0 + wm * WTM + i * 16 + lk + 4 * r;
        if (m >= p.M) continue;
        const int64_t off = (int64_t)m * p.scm + (int64_t)n * p.scn;
        if constexpr (CPLX) {
          const double xr = acc[0][i][j][r], xi = acc[1][i][j][r];
          double vr = p.alpha_re * xr - p.alpha_im * xi;
          double vi = p.alpha_re * xi + p.alpha_im * xr;
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
          double v = p.alpha_re * acc[0][i][j][r];
          double *c = p.C + off;
          if (split) {
            atomicAdd(c, v);
          } else {
            if (has_beta) v += p.beta_re * (*c);
            *c = v;
          }
        }
      }
    }
}

// ------------------------------------------------------------------ dispatch

// Candidate configurations the tuner measures (wave grid 2x2, 256 threads).
static const Cand kCands[] = {
    {64, 64, 16, 1, 0}, {64, 64, 32, 1, 0}, {128, 64, 16, 1, 0}, {64, 128, 16, 1, 0},
    {128, 128, 16, 1, 0}, {64, 64, 16, 2, 0}, {64, 64, 32, 2, 0}, {128, 64, 16, 2, 0},
    {64, 64, 16, 1, 8}, {64, 64, 32, 1, 8}, {64, 64, 16, 2, 8}, {64, 64, 32, 2, 8},
    {128, 64, 16, 1, 8}, {64, 128, 16, 1, 8},
    // eight waves per block: two waves per SIMD when a GEMM has about one
    // output tile per CU
    {64, 64, 16, 1, 0, 8}, {64, 64, 32, 1, 0, 8}, {128, 64, 16, 1, 0, 8},
    {64, 64, 16, 1, 8, 8}, {64, 64, 32, 2, 0, 8},
    // eight waves as two K-slice groups of a 2x2 wave grid
    {64, 64, 32, 1, 0, 8, 2}, {64, 64, 32, 1, 8, 8, 2}, {128, 64, 32, 1, 0, 8, 2},
    {64, 64, 64, 1, 0, 8, 2},
    // the same eight-wave forms with the two-step register prefetch (real)
    {64, 64, 32, 1, 0, 8, 2, 2}, {64, 64, 64, 1, 0, 8, 2, 2}, {128, 64, 32, 1, 0, 8, 2, 2},
    {64, 64, 32, 1, 8, 8, 2, 2}, {64, 64, 32, 1, 0, 8, 0, 2}, {64, 64, 16, 1, 0, 8, 0, 2},
};
constexpr int kNCand = sizeof(kCands) / sizeof(kCands[0]);

template <bool CPLX, bool AMC, bool BKC, bool VEC>
static void launch_cfg(const Cand &c, GemmP p, hipStream_t s) {
  dim3 grid = setup_grid(c, p);
#define MH_L(BM, BN, BK) \
  hipLaunchKernelGGL((k_gemm<CPLX, BM, BN, BK, AMC, BKC, VEC>), grid, dim3(256), 0, s, p)
#define MH_L8(BM, BN, BK) \
  hipLaunchKernelGGL((k_gemm<CPLX, BM, BN, BK, AMC, BKC, VEC, 8>), grid, dim3(512), 0, s, p)
#define MH_L8K(BM, BN, BK) \
  hipLaunchKernelGGL((k_gemm<CPLX, BM, BN, BK, AMC, BKC, VEC, 8, 2>), grid, dim3(512), 0, s, p)
#define MH_L8KP(BM, BN, BK) \
  hipLaunchKernelGGL((k_gemm<CPLX, BM, BN, BK, AMC, BKC, VEC, 8, 2, 2>), grid, dim3(512), 0, s, p)
#define MH_L8P(BM, BN, BK) \
  hipLaunchKernelGGL((k_gemm<CPLX, BM, BN, BK, AMC, BKC, VEC, 8, 1, 2>), grid, dim3(512), 0, s, p)
  if (c.prefetch == 2) {
    if constexpr (!CPLX) {  // real only (registers)
      if (c.kgroups == 2) {
        if (c.bm == 128) MH_L8KP(128, 64, 32);
        else if (c.bk == 64) MH_L8KP(64, 64, 64);
        else MH_L8KP(64, 64, 32);
      } else {
        if (c.bk == 32) MH_L8P(64, 64, 32);
        else MH_L8P(64, 64, 16);
      }
    }
  } else if (c.kgroups == 2) {
    if constexpr (!CPLX) {  // real only (LDS and registers)
      if (c.bm == 128) MH_L8K(128, 64, 32);
      else if (c.bk == 64) MH_L8K(64, 64, 64);
      else MH_L8K(64, 64, 32);
    }
  } else if (c.waves == 8) {
    if (c.bm == 128) MH_L8(128, 64, 16);
    else if (c.bk == 32) MH_L8(64, 64, 32);
    else MH_L8(64, 64, 16);
  }
  else if (c.bm == 64 && c.bn == 64 && c.bk == 16) MH_L(64, 64, 16);
  else if (c.bm == 64 && c.bn == 64) MH_L(64, 64, 32);
  else if (c.bm == 128 && c.bn == 64) MH_L(128, 64, 16);
  else if (c.bm == 64 && c.bn == 128) MH_L(64, 128, 16);
  else if constexpr (!CPLX) MH_L(128, 128, 16);  // complex 128x128 is never admitted
#undef MH_L
#undef MH_L8
#undef MH_L8K
#undef MH_L8KP
#undef MH_L8P
}

template <bool CPLX, bool VEC>
static void launch_layout(const Cand &c, const GemmP &p, hipStream_t s, bool amc, bool bkc) {
  if (amc && bkc) launch_cfg<CPLX, true, true, VEC>(c, p, s);
  else if (amc) launch_cfg<CPLX, true, false, VEC>(c, p, s);
  else if (bkc) launch_cfg<CPLX, false, true, VEC>(c, p, s);
  else launch_cfg<CPLX, false, false, VEC>(c, p, s);
}

// 16-byte (complex: 32-byte) pair loads need the pair's first element at an
// even index of the contiguous dimension, an even extent along it, an even
// stride along the other dimension and a 16-byte aligned base.
static bool vec_ok(const GemmP &p, bool amc, bool bkc) {
  auto al = [](const void *q) { return ((uintptr_t)q & 15) == 0; };
  const bool a = al(p.A) && (amc ? (p.sam == 1 && p.M % 2 == 0 && (p.sak % 2 == 0 || p.K == 1))
                                 : (p.sak == 1 && p.K % 2 == 0 && (p.sam % 2 == 0 || p.M == 1)));
  const bool b = al(p.B) && (bkc ? (p.sbk == 1 && p.K % 2 == 0 && (p.sbn % 2 == 0 || p.N == 1))
                                 : (p.sbn == 1 && p.N % 2 == 0 && (p.sbk % 2 == 0 || p.K == 1)));
  return a && b;
}

static void launch_f64(const Cand &c, const GemmP &p, hipStream_t s, bool cplx, bool amc, bool bkc) {
  const bool vec = vec_ok(p, amc, bkc);
  if (cplx) {
    if (vec) launch_layout<true, true>(c, p, s, amc, bkc);
    else launch_layout<true, false>(c, p, s, amc, bkc);
  } else {
    if (vec) launch_layout<false, true>(c, p, s, amc, bkc);
    else launch_layout<false, false>(c, p, s, amc, bkc);
  }
}

static bool cand_ok_f64(const Cand &c, const GemmP &p, bool cplx) {
  if (c.split > 1 && (p.beta_re != 0.0 || p.beta_im != 0.0)) return false;
  if (c.split > 1 && p.K < 256 * c.split) return false;
  if (cplx && c.bk == 32) return false;              // LDS: keep <= 80 KB
  if (cplx && c.kgroups == 2) return false;
  if (cplx && c.prefetch == 2) return false;
  if (cplx && c.bm * c.bn > 64 * 128) return false;  // register pressure
  return true;
}

static const GemmFamily kF64Family = {"f64", kCands, kNCand, 16, cand_ok_f64, launch_f64};

// ------------------------------------------------ fused first H_eff step
//
// One- and two-site H_eff, FP64 real:
//   T1[a, w, s, b'] = sum_a' L[a, w, a'] x[a', s, b']           (MFMA)
//   T2[a, j, b']    = sum_i T1[a, i, b'] Mt[i, j],  i = w + WL*s (epilogue)
// (s: the physical index of x, DS values; j = s' + DS*w', NJ = DS*WR.)
// A block owns a = a0..a0+31 for EVERY channel w (BM = 32*WL logical rows:
// row r is a = a0 + r%32 of channel r/32) and 64/DS values of b' for every s
// (BN = 64 columns c = s + DS*b', contiguous in x), so the whole T1[a, :, b']
// of each of its (a, b') pairs is inside the block and T1 never goes to HBM.
// The K loop (a') is k_gemm's (double-buffered LDS, one barrier per step,
// v_mfma_f64_16x16x4f64).  Epilogue in two halves of 32 columns: the waves
// owning a half write it to LDS ([column][row], pitch BM + 1), then every
// thread forms T2 for one or more (a, b') pairs from NI values; Mt entries
// are wave-uniform (scalar loads).  Needs Dl even and 16-byte aligned L, x.
template <int WL, int DS, int WR>
__global__ void __launch_bounds__(256)
    k_heff_step1(const double *__restrict__ L, const double *__restrict__ X,
                 const double *__restrict__ Mt, double *__restrict__ T2, int Dl, int Dr,
                 int tiles_a, int tiles_b, int xcd_chunk, int group) {
  constexpr int BA = 32, BM = BA * WL, BN = 64, BK = 16;
  constexpr int NI = WL * DS, NJ = DS * WR;
  constexpr int WTM = BM / 2, WTN = BN / 2, TM = WTM / 16, TN = WTN / 16;
  constexpr int HALF = BN / 2, BBH = HALF / DS;  // columns / b' values per half
  constexpr int NPAIRS = BA * BBH;               // (a, b') pairs per half
  static_assert(BN % DS == 0 && HALF % DS == 0 && NPAIRS % 256 == 0, "tile");
  using SA = Stage<false, BM, BK, true>;   // L rows, staged [k][r]
  using SB = Stage<false, BN, BK, false>;  // x columns, staged [c][k]
  constexpr int CP = BM + 1;
  constexpr int NSTAGE = 2 * SA::SIZE + 2 * SB::SIZE;
  constexpr int NEPI = HALF * CP;
  __shared__ __align__(16) double smem[NSTAGE > NEPI ? NSTAGE : NEPI];
  double *const sA = smem;
  double *const sB = smem + 2 * SA::SIZE;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  // XCD-contiguous tile order: block id -> XCD id % 8 -> a contiguous range
  // of a grouped order (group a-tiles swept together along b'), so the tiles
  // of one XCD form a compact block of (a, b') whose L rows and x columns its
  // own L2 can hold.
  const int local = blockIdx.x >> 3;
  const int t = (blockIdx.x & 7) * xcd_chunk + local;
  if (local >= xcd_chunk || t >= tiles_a * tiles_b) return;
  const int span = group * tiles_b, first = (t / span) * group;
  const int gsz = min(tiles_a - first, group), wt = t % span;
  const int a0 = (first + wt % gsz) * BA, c0 = (wt / gsz) * BN;
  const int K = Dl, NC = DS * Dr;
  const int64_t ldL = (int64_t)Dl * WL;
  const int lr = lane & 15, lk = lane >> 4;

  int64_t aoff[SA::NPAIR], boff[SB::NPAIR];
  bool aok[SA::NPAIR][2], bok[SB::NPAIR];
#pragma unroll
  for (int i = 0; i < SA::NPAIR; ++i) {
    int r, k;
    SA::coord(tid, i, r, k);
    const int a = a0 + r % BA, w = r / BA;
    aok[i][0] = a < Dl;
    aok[i][1] = a + 1 < Dl;
    // a is even and Dl is even: the clamped pair stays inside channel w
    aoff[i] = (int64_t)min(a, Dl - 2) + (int64_t)Dl * w + (int64_t)k * ldL;
  }
#pragma unroll
  for (int i = 0; i < SB::NPAIR; ++i) {
    int r, k;
    SB::coord(tid, i, r, k);
    const int c = c0 + r;
    bok[i] = c < NC;
    boff[i] = (int64_t)min(c, NC - 1) * Dl + k;
  }

  double ra[SA::NPAIR][2], rb[SB::NPAIR][2];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < SA::NPAIR; ++i) {
      int r, k;
      SA::coord(tid, i, r, k);
      const int kc = min(k0 + k, K - 1);
      const double2 v = *reinterpret_cast<const double2 *>(L + aoff[i] + (int64_t)(kc - k) * ldL);
      ra[i][0] = v.x;
      ra[i][1] = v.y;
    }
#pragma unroll
    for (int i = 0; i < SB::NPAIR; ++i) {
      int r, k;
      SB::coord(tid, i, r, k);
      const int kc = min(k0 + k, K - 2);
      const double2 v = *reinterpret_cast<const double2 *>(X + boff[i] + (kc - k));
      rb[i][0] = v.x;
      rb[i][1] = v.y;
    }
  };
  auto store = [&](int buf, int k0) {
#pragma unroll
    for (int i = 0; i < SA::NPAIR; ++i) {
      int r, k;
      SA::coord(tid, i, r, k);
      const bool kok = k0 + k < K;
      double2 v;
      v.x = (aok[i][0] && kok) ? ra[i][0] : 0.0;
      v.y = (aok[i][1] && kok) ? ra[i][1] : 0.0;
      *reinterpret_cast<double2 *>(&sA[buf * SA::SIZE + k * SA::PITCH + r]) = v;
    }
#pragma unroll
    for (int i = 0; i < SB::NPAIR; ++i) {
      int r, k;
      SB::coord(tid, i, r, k);
      double2 v;
      v.x = (bok[i] && k0 + k < K) ? rb[i][0] : 0.0;
      v.y = (bok[i] && k0 + k + 1 < K) ? rb[i][1] : 0.0;
      *reinterpret_cast<double2 *>(&sB[buf * SB::SIZE + r * SB::PITCH + k]) = v;
    }
  };

  d4_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};

  auto compute = [&](int buf) {
    const double *A = sA + buf * SA::SIZE;
    const double *B = sB + buf * SB::SIZE;
#pragma unroll
    for (int kb = 0; kb < BK; kb += 16) {
      double a[TM][4], b[TN][4];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int r = wm * WTM + i * 16 + lr;
#pragma unroll
        for (int s = 0; s < 4; ++s) a[i][s] = A[(kb + 4 * lk + s) * SA::PITCH + r];
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int r = wn * WTN + j * 16 + lr;
        const double2 *src = reinterpret_cast<const double2 *>(&B[r * SB::PITCH + kb + 4 * lk]);
        const double2 u = src[0], v = src[1];
        b[j][0] = u.x; b[j][1] = u.y; b[j][2] = v.x; b[j][3] = v.y;
      }
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][s], b[j][s], acc[i][j], 0, 0, 0);
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

  // Epilogue: half h holds the columns of the waves with wn == h.
  double *const Cs = smem;
  const int bq0 = c0 / DS;  // first b' of the block
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    __syncthreads();  // LDS free: the K loop (h = 0) or the previous half is done
    if (wn == h) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int m = wm * WTM + i * 16 + lk + 4 * r;
            const int n = j * 16 + lr;  // column inside the half
            Cs[n * CP + m] = acc[i][j][r];
          }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NPAIRS / 256; ++q) {
      const int pr = tid + 256 * q;
      const int al = pr % BA, bl = pr / BA;
      const int a = a0 + al, bq = bq0 + h * BBH + bl;
      double v[NI];
#pragma unroll
      for (int s = 0; s < DS; ++s)
#pragma unroll
        for (int w = 0; w < WL; ++w) v[w + WL * s] = Cs[(bl * DS + s) * CP + al + BA * w];
      if (a < Dl && bq < Dr) {
        double *dst = T2 + a + (int64_t)Dl * NJ * bq;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          double o = 0.0;
#pragma unroll
          for (int i = 0; i < NI; ++i) o = __builtin_fma(v[i], Mt[i + NI * j], o);
          dst[(int64_t)Dl * j] = o;
        }
      }
    }
  }
}

template <int WL, int DS>
static void launch_step1(int Dl, int Dr, const double *L, const double *X, const double *Mt,
                         double *T2, hipStream_t s) {
  const int tiles_a = (Dl + 31) / 32, tiles_b = (DS * Dr + 63) / 64;
  const int chunk = (tiles_a * tiles_b + 7) / 8;
  const int group = min(tiles_a, 8);
  hipLaunchKernelGGL((k_heff_step1<WL, DS, WL>), dim3(8 * chunk), dim3(256), 0, s, L, X, Mt, T2,
                     Dl, Dr, tiles_a, tiles_b, chunk, group);
}

// ComplexF64 form of k_heff_step1: L, x, Mt and T2 hold interleaved (re, im)
// pairs.  The tile is BM = 32 WL rows (a, w) by BN = 32 complex columns
// (s, b'); both operands go to LDS as re / im planes (A [k][r], B [c][k]);
// each 16x16x4 step runs the three real MFMAs of k_gemm's complex form.  The epilogue stages the whole
// 32-column tile ([plane][column][row], pitch BM + 1) and each thread forms
// T2[a, j, b'] = sum_i T1[a, i, b'] Mt[i, j] in complex arithmetic for its
// (a, b') pairs.  Tile order as k_heff_step1 (XCD-contiguous groups).
template <int WL, int DS, int WR>
__global__ void __launch_bounds__(256)
    k_heff_step1_c(const double *__restrict__ L, const double *__restrict__ X,
                   const double *__restrict__ Mt, double *__restrict__ T2, int Dl, int Dr,
                   int tiles_a, int tiles_b, int xcd_chunk, int group) {
  constexpr int BA = 32, BM = BA * WL, BN = 32, BK = 16;
  constexpr int NI = WL * DS, NJ = DS * WR;
  constexpr int WTM = BM / 2, WTN = BN / 2, TM = WTM / 16, TN = WTN / 16;
  constexpr int BB = BN / DS;            // b' values per tile
  constexpr int NPAIRS = BA * BB;        // (a, b') pairs per tile
  static_assert(BN % DS == 0 && NPAIRS % 256 == 0, "tile");
  using SA = Stage<false, BM, BK, true>;   // L rows, staged [k][r]
  using SB = Stage<false, BN, BK, false>;  // x columns, staged [c][k]
  constexpr int CP = BM + 1;
  constexpr int NSTAGE = 2 * 2 * SA::SIZE + 2 * 2 * SB::SIZE;  // buffers x planes
  constexpr int NEPI = 2 * BN * CP;
  __shared__ __align__(16) double smem[NSTAGE > NEPI ? NSTAGE : NEPI];
  // plane q of LDS buffer b: A at smem + (2 b + q) SA::SIZE, B after all of A
  auto sA = [&](int b, int q) { return smem + (2 * b + q) * SA::SIZE; };
  auto sB = [&](int b, int q) { return smem + 4 * SA::SIZE + (2 * b + q) * SB::SIZE; };

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

  int64_t aoff[SA::NPAIR], boff[SB::NPAIR];
  bool aok[SA::NPAIR][2], bok[SB::NPAIR];
#pragma unroll
  for (int i = 0; i < SA::NPAIR; ++i) {
    int r, k;
    SA::coord(tid, i, r, k);
    const int a = a0 + r % BA, w = r / BA;
    aok[i][0] = a < Dl;
    aok[i][1] = a + 1 < Dl;
    // a is even and Dl is even: the clamped pair stays inside channel w
    aoff[i] = (int64_t)min(a, Dl - 2) + (int64_t)Dl * w + (int64_t)k * ldL;
  }
#pragma unroll
  for (int i = 0; i < SB::NPAIR; ++i) {
    int r, k;
    SB::coord(tid, i, r, k);
    const int c = c0 + r;
    bok[i] = c < NC;
    boff[i] = (int64_t)min(c, NC - 1) * Dl + k;
  }

  // [plane][pair][element]
  double ra[2][SA::NPAIR][2], rb[2][SB::NPAIR][2];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < SA::NPAIR; ++i) {
      int r, k;
      SA::coord(tid, i, r, k);
      const int kc = min(k0 + k, K - 1);
      const double4 v = *reinterpret_cast<const double4 *>(L + 2 * (aoff[i] + (int64_t)(kc - k) * ldL));
      ra[0][i][0] = v.x; ra[1][i][0] = v.y; ra[0][i][1] = v.z; ra[1][i][1] = v.w;
    }
#pragma unroll
    for (int i = 0; i < SB::NPAIR; ++i) {
      int r, k;
      SB::coord(tid, i, r, k);
      const int kc = min(k0 + k, K - 2);
      const double4 v = *reinterpret_cast<const double4 *>(X + 2 * (boff[i] + (kc - k)));
      rb[0][i][0] = v.x; rb[1][i][0] = v.y; rb[0][i][1] = v.z; rb[1][i][1] = v.w;
    }
  };
  auto store = [&](int buf, int k0) {
#pragma unroll
    for (int i = 0; i < SA::NPAIR; ++i) {
      int r, k;
      SA::coord(tid, i, r, k);
      const bool kok = k0 + k < K;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        double2 v;
        v.x = (aok[i][0] && kok) ? ra[q][i][0] : 0.0;
        v.y = (aok[i][1] && kok) ? ra[q][i][1] : 0.0;
        *reinterpret_cast<double2 *>(sA(buf, q) + k * SA::PITCH + r) = v;
      }
    }
#pragma unroll
    for (int i = 0; i < SB::NPAIR; ++i) {
      int r, k;
      SB::coord(tid, i, r, k);
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        double2 v;
        v.x = (bok[i] && k0 + k < K) ? rb[q][i][0] : 0.0;
        v.y = (bok[i] && k0 + k + 1 < K) ? rb[q][i][1] : 0.0;
        *reinterpret_cast<double2 *>(sB(buf, q) + r * SB::PITCH + k) = v;
      }
    }
  };

  d4_t acc[3][TM][TN];  // P1, P2, P3 of the three-product form (as k_gemm)
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[q][i][j] = d4_t{0.0, 0.0, 0.0, 0.0};

  auto compute = [&](int buf) {
#pragma unroll
    for (int kb = 0; kb < BK; kb += 16) {
      double a[2][TM][4], b[2][TN][4];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const double *A = sA(buf, q);
        const double *B = sB(buf, q);
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          const int r = wm * WTM + i * 16 + lr;
#pragma unroll
          for (int s = 0; s < 4; ++s) a[q][i][s] = A[(kb + 4 * lk + s) * SA::PITCH + r];
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const int r = wn * WTN + j * 16 + lr;
          const double2 *src = reinterpret_cast<const double2 *>(&B[r * SB::PITCH + kb + 4 * lk]);
          const double2 u = src[0], v = src[1];
          b[q][j][0] = u.x; b[q][j][1] = u.y; b[q][j][2] = v.x; b[q][j][3] = v.y;
        }
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        double sa[TM], sb[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) sa[i] = a[0][i][s] + a[1][i][s];
#pragma unroll
        for (int j = 0; j < TN; ++j) sb[j] = b[0][j][s] + b[1][j][s];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0][i][s], b[0][j][s], acc[0][i][j], 0, 0, 0);
            acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1][i][s], b[1][j][s], acc[1][i][j], 0, 0, 0);
            acc[2][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(sa[i], sb[j], acc[2][i][j], 0, 0, 0);
          }
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

  // Epilogue: the whole tile to LDS, then the complex mode product.
  double *const Cr = smem, *const Ci = smem + BN * CP;
  __syncthreads();  // every wave is past its last compute: LDS is free
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = wm * WTM + i * 16 + lk + 4 * r;
        const int n = wn * WTN + j * 16 + lr;
        Cr[n * CP + m] = acc[0][i][j][r] - acc[1][i][j][r];
        Ci[n * CP + m] = acc[2][i][j][r] - acc[0][i][j][r] - acc[1][i][j][r];
      }
  __syncthreads();
  const int bq0 = c0 / DS;
#pragma unroll
  for (int q = 0; q < NPAIRS / 256; ++q) {
    const int pr = tid + 256 * q;
    const int al = pr % BA, bl = pr / BA;
    const int a = a0 + al, bq = bq0 + bl;
    double vr[NI], vi[NI];
#pragma unroll
    for (int s2 = 0; s2 < DS; ++s2)
#pragma unroll
      for (int w = 0; w < WL; ++w) {
        vr[w + WL * s2] = Cr[(bl * DS + s2) * CP + al + BA * w];
        vi[w + WL * s2] = Ci[(bl * DS + s2) * CP + al + BA * w];
      }
    if (a < Dl && bq < Dr) {
      double *dst = T2 + 2 * (a + (int64_t)Dl * NJ * bq);
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        double orr = 0.0, oi = 0.0;
#pragma unroll
        for (int i = 0; i < NI; ++i) {
          const double mr = Mt[2 * (i + NI * j)], mi = Mt[2 * (i + NI * j) + 1];
          orr = __builtin_fma(vr[i], mr, orr);
          orr = __builtin_fma(-vi[i], mi, orr);
          oi = __builtin_fma(vr[i], mi, oi);
          oi = __builtin_fma(vi[i], mr, oi);
        }
        *reinterpret_cast<double2 *>(dst + 2 * (int64_t)Dl * j) = make_double2(orr, oi);
      }
    }
  }
}

template <int WL, int DS>
static void launch_step1_c(int Dl, int Dr, const double *L, const double *X, const double *Mt,
                           double *T2, hipStream_t s) {
  const int tiles_a = (Dl + 31) / 32, tiles_b = (DS * Dr + 31) / 32;
  const int chunk = (tiles_a * tiles_b + 7) / 8;
  const int group = min(tiles_a, 8);
  hipLaunchKernelGGL((k_heff_step1_c<WL, DS, WL>), dim3(8 * chunk), dim3(256), 0, s, L, X, Mt, T2,
                     Dl, Dr, tiles_a, tiles_b, chunk, group);
}

// T2 = mode(L x) in one launch for ComplexF64 operators with wl = wr in
// {3, 4, 5, 6} and a physical extent of 2 or 4; MH_ERR_UNSUPPORTED otherwise.
int heff_step1_fused_c(mh_handle h, int Dl, int Dr, int wl, int ds, int wr, const void *L,
                       const void *x, const void *Mt, void *T2) {
  auto al16 = [](const void *q) { return ((uintptr_t)q & 15) == 0; };
  if (wl != wr || Dl < 2 || Dl % 2 != 0 || !al16(L) || !al16(x) || !al16(T2) || !al16(Mt))
    return MH_ERR_UNSUPPORTED;
  const double *Lp = (const double *)L, *Xp = (const double *)x, *Mp = (const double *)Mt;
  double *Tp = (double *)T2;
  hipStream_t s = h->stream;
  const double flops = 8.0 * Dl * (double)wl * ds * Dr * Dl;  // complex GEMM flops
  int slot;
#define MH_S1C(W, D)                                  \
  if (wl == W && ds == D) {                           \
    prof_begin(h, MH_K_GEMM, flops, &slot);           \
    launch_step1_c<W, D>(Dl, Dr, Lp, Xp, Mp, Tp, s);  \
    prof_end(h, slot);                                \
    return check_launch(h, "fused complex H_eff step 1"); \
  }
  MH_S1C(3, 4) MH_S1C(3, 2) MH_S1C(4, 4) MH_S1C(4, 2) MH_S1C(5, 4) MH_S1C(5, 2)
  MH_S1C(6, 4) MH_S1C(6, 2)
#undef MH_S1C
  return MH_ERR_UNSUPPORTED;
}

// T2 = mode(L x) in one launch for FP64 real operators with wl = wr in
// {3,4,5,6,8} and a physical extent of 2 or 4 (one-site d = 2, 4; two-site
// d = 2); MH_ERR_UNSUPPORTED otherwise (the caller runs GEMM + mode product).
int heff_step1_fused(mh_handle h, int Dl, int Dr, int wl, int ds, int wr, const void *L,
                     const void *x, const void *Mt, void *T2) {
  auto al16 = [](const void *q) { return ((uintptr_t)q & 15) == 0; };
  if (wl != wr || Dl < 2 || Dl % 2 != 0 || !al16(L) || !al16(x)) return MH_ERR_UNSUPPORTED;
  const double *Lp = (const double *)L, *Xp = (const double *)x, *Mp = (const double *)Mt;
  double *Tp = (double *)T2;
  hipStream_t s = h->stream;
  // profiled as a GEMM launch with the GEMM's 2 M N K flops
  const double flops = 2.0 * Dl * (double)wl * ds * Dr * Dl;
  int slot;
#define MH_S1(W, D)                                   \
  if (wl == W && ds == D) {                           \
    prof_begin(h, MH_K_GEMM, flops, &slot);           \
    launch_step1<W, D>(Dl, Dr, Lp, Xp, Mp, Tp, s);    \
    prof_end(h, slot);                                \
    return check_launch(h, "fused H_eff step 1");     \
  }
  MH_S1(5, 4) MH_S1(5, 2) MH_S1(3, 4) MH_S1(3, 2) MH_S1(4, 4) MH_S1(4, 2)
  MH_S1(6, 4) MH_S1(6, 2) MH_S1(8, 4) MH_S1(8, 2)
#undef MH_S1
  return MH_ERR_UNSUPPORTED;
}

// Register-only FP64 MFMA throughput probe: four independent accumulator
// chains per wave, no memory traffic in the loop.  Measures the matrix-core
// rate this device actually sustains.
__global__ void __launch_bounds__(256) k_mfma_probe(double *out, int iters) {
  d4_t acc0 = {0, 0, 0, 0}, acc1 = acc0, acc2 = acc0, acc3 = acc0;
  const double a = 1.0 + 1e-6 * threadIdx.x, b = 1.0 - 1e-6 * threadIdx.x;
  for (int i = 0; i < iters; ++i) {
    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, acc1, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, a, acc2, 0, 0, 0);
    acc3 = __builtin_amdgcn_mfma_f64_16x16x4f64(b, b, acc3, 0, 0, 0);
  }
  const d4_t s = acc0 + acc1 + acc2 + acc3;
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s[0] + s[1] + s[2] + s[3];
}

}  // namespace mh

extern "C" int mh_probe_mfma_f64(mh_handle h, double *tflops) {
  using namespace mh;
  if (!h || !tflops) return MH_ERR_ARG;
  int ncu = 0;
  MH_HIP(h, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device));
  const int blocks = ncu * 4, iters = 4000;
  MH_TRY(ensure_red(h, (size_t)blocks * 256));
  hipEvent_t e0, e1;
  MH_HIP(h, hipEventCreate(&e0));
  MH_HIP(h, hipEventCreate(&e1));
  hipLaunchKernelGGL(k_mfma_probe, dim3(blocks), dim3(256), 0, h->stream, h->red, iters);
  MH_HIP(h, hipEventRecord(e0, h->stream));
  hipLaunchKernelGGL(k_mfma_probe, dim3(blocks), dim3(256), 0, h->stream, h->red, iters);
  MH_HIP(h, hipEventRecord(e1, h->stream));
  MH_HIP(h, hipEventSynchronize(e1));
  float ms = 0.f;
  MH_HIP(h, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const double flops = (double)blocks * 4 /*waves*/ * iters * 4 /*chains*/ * 2.0 * 16 * 16 * 4;
  *tflops = flops / (ms * 1e-3) / 1e12;
  return check_launch(h, "mfma probe");
}

namespace mh {

int gemm_mixed(mh_handle h, int dtype, int prec, const GemmDesc &g);

int gemm(mh_handle h, int dtype, int prec, const GemmDesc &g) {
  if (g.M <= 0 || g.N <= 0) return MH_OK;
  if (prec != MH_PREC_F64) return gemm_mixed(h, dtype, prec, g);
  GemmP p = make_params(g);
  return gemm_dispatch(h, kF64Family, p, dtype == MH_C64, "f64");
}

// ------------------------------------------------------------ shared tuner

static int run_cand(mh_handle h, const GemmFamily &fam, const Cand &c, const GemmP &p, bool cplx,
                    bool amc, bool bkc) {
  if (c.split > 1) {
    int64_t n = (int64_t)p.M * p.N;
    hipLaunchKernelGGL(k_zero_c, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                       p.C, p.M, p.N, p.scm, p.scn, cplx ? 1 : 0, p.c_f32);
  }
  fam.launch(c, p, h->stream, cplx, amc, bkc);
  return check_launch(h, "gemm");
}

// Heuristic when no measurement is available: minimise (tiles per CU, rounded
// up) x (tile area) -- the kernels are MFMA-bound -- and split K in two when
// the grid leaves CUs idle.
static int heuristic(const GemmFamily &fam, const GemmP &p, bool cplx) {
  const int ncu = 256;
  int best = -1;
  double best_cost = 1e300;
  for (int i = 0; i < fam.n; ++i) {
    const Cand &c = fam.cands[i];
    if (!fam.ok(c, p, cplx) || c.bk != fam.default_bk || c.group != 0) continue;
    int64_t tiles = (int64_t)((p.M + c.bm - 1) / c.bm) * ((p.N + c.bn - 1) / c.bn) * c.split;
    double rounds = (double)((tiles + ncu - 1) / ncu);
    double area = (double)c.bm * c.bn / c.split;
    double penalty = (c.bm * c.bn == 64 * 64) ? 1.08 : (c.bm * c.bn == 128 * 64 ? 1.03 : 1.0);
    if (c.split > 1) penalty *= 1.05;
    double cost = rounds * area * penalty;
    if (cost < best_cost) {
      best_cost = cost;
      best = i;
    }
  }
  return best < 0 ? 0 : best;
}

// Measure every admissible candidate on this shape (the output is recomputed
// each time, so C ends up holding a correct result) and remember the fastest.
static int tune(mh_handle h, const GemmFamily &fam, const GemmP &p, bool cplx, bool amc,
                bool bkc, int *choice) {
  hipEvent_t e0, e1;
  MH_HIP(h, hipEventCreate(&e0));
  MH_HIP(h, hipEventCreate(&e1));
  float best = 1e30f;
  int bi = heuristic(fam, p, cplx);
  for (int i = 0; i < fam.n; ++i) {
    const Cand &c = fam.cands[i];
    if (!fam.ok(c, p, cplx)) continue;
    MH_TRY(run_cand(h, fam, c, p, cplx, amc, bkc));  // warm-up
    MH_HIP(h, hipEventRecord(e0, h->stream));
    for (int r = 0; r < 3; ++r) MH_TRY(run_cand(h, fam, c, p, cplx, amc, bkc));
    MH_HIP(h, hipEventRecord(e1, h->stream));
    MH_HIP(h, hipEventSynchronize(e1));
    float ms = 0.f;
    MH_HIP(h, hipEventElapsedTime(&ms, e0, e1));
    ms /= 3.f;
    if (h->tune_log)
      fprintf(stderr, "[mh tune] %s%s M=%d N=%d K=%d amc=%d bkc=%d tile=%dx%dx%d split=%d group=%d waves=%d kgroups=%d prefetch=%d: %.1f us (%.2f TFLOP/s)\n",
              fam.name, cplx ? "/c" : "", p.M, p.N, p.K, (int)amc, (int)bkc, c.bm, c.bn, c.bk,
              c.split, c.group, c.waves ? c.waves : 4, c.kgroups ? c.kgroups : 1,
              c.prefetch == 2 ? 2 : 1, ms * 1e3,
              2.0 * p.M * (double)p.N * p.K * (cplx ? 4 : 1) / (ms * 1e-3) / 1e12);
    if (ms < best) {
      best = ms;
      bi = i;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // Re-run the winner so the final state does not depend on which candidate
  // happened to be measured last.
  MH_TRY(run_cand(h, fam, fam.cands[bi], p, cplx, amc, bkc));
  *choice = bi;
  return MH_OK;
}

int gemm_dispatch(mh_handle h, const GemmFamily &fam, const GemmP &p, bool cplx,
                  const char *tag) {
  const bool amc = (p.sam == 1) || (p.M == 1);
  const bool bkc = (p.sbk == 1) || (p.K == 1);
  const double flops = 2.0 * p.M * (double)p.N * p.K * (cplx ? 4.0 : 1.0);
  const bool has_beta = p.beta_re != 0.0 || p.beta_im != 0.0;
  int choice = -1;
  // a forced configuration (tests), if this family has it and admits it
  if (h->gemm_force >= 0 && h->gemm_force < fam.n && fam.ok(fam.cands[h->gemm_force], p, cplx)) {
    int slot;
    prof_begin(h, MH_K_GEMM, flops, &slot);
    int st = run_cand(h, fam, fam.cands[h->gemm_force], p, cplx, amc, bkc);
    prof_end(h, slot);
    return st;
  }
  // Allocation-free callers: the heuristic configuration, no table lookup
  // (the key string would be a host allocation) and no tuning runs.
  if (h->noalloc) return run_cand(h, fam, fam.cands[heuristic(fam, p, cplx)], p, cplx, amc, bkc);
  auto make_key = [&](int beta_flag) {
    char key[192];
    snprintf(key, sizeof(key), "%s:%d:%d:%d:%d:%d%d:%d:%d", tag, cplx ? 1 : 0, p.M, p.N, p.K,
             (int)amc, (int)bkc, beta_flag, p.c_f32);
    return std::string(key);
  };
  const std::string key = make_key(has_beta ? 1 : 0);
  auto it = h->tune.find(key);
  if (it != h->tune.end()) {
    choice = it->second;
  } else {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(h->stream, &cs);
    // Tune shapes worth more than ~50 us of work; small ones use the
    // heuristic.  Tuning runs every candidate on the caller's C, which is
    // only valid when C is pure output (beta = 0): with beta != 0 each run
    // would fold the previous run's result back in.  Such shapes take the
    // configuration tuned for the same shape with beta = 0 when there is
    // one that admits beta, else the heuristic.
    if (!has_beta && h->autotune && cs == hipStreamCaptureStatusNone && flops > 2e9 &&
        !h->profiling) {
      MH_TRY(tune(h, fam, p, cplx, amc, bkc, &choice));
      h->tune[key] = choice;
      return MH_OK;
    }
    if (has_beta) {
      auto it0 = h->tune.find(make_key(0));
      if (it0 != h->tune.end() && fam.ok(fam.cands[it0->second], p, cplx)) choice = it0->second;
    }
    if (choice < 0) choice = heuristic(fam, p, cplx);
    h->tune[key] = choice;
  }
  int slot;
  prof_begin(h, MH_K_GEMM, flops, &slot);
  int st = run_cand(h, fam, fam.cands[choice], p, cplx, amc, bkc);
  prof_end(h, slot);
  return st;
}

}  // namespace mh

namespace mh {
const GemmFamily &mixed_family(int prec);
}

extern "C" int mh_gemm_candidates(int prec, int *n) {
  if (!n) return MH_ERR_ARG;
  if (prec == MH_PREC_F64) *n = mh::kNCand;
  else if (prec == MH_PREC_F32 || prec == MH_PREC_BF16X3) *n = mh::mixed_family(prec).n;
  else return MH_ERR_ARG;
  return MH_OK;
}

extern "C" int mh_gemm_force_candidate(mh_handle h, int cand) {
  if (!h) return MH_ERR_ARG;
  if (cand < -1 || cand >= 64) return mh::fail(h, MH_ERR_ARG, "no GEMM configuration %d", cand);
  h->gemm_force = cand;
  return MH_OK;
}

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
