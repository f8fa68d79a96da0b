// Block-sparse (abelian-symmetric) contractions: grouped GEMM and block
// mode product, driven by a plan built on the host from the charge sectors.
//
// A symmetric tensor is stored as a list of dense column-major blocks in one
// buffer.  A contraction such as the two-site H_eff becomes three stages:
//   1. grouped GEMM         T1 blocks = environment blocks x vector blocks
//   2. block mode product   T2 blocks = sum_c coef_c * T1 block_c
//                           (the MPO entries that conserve the charges)
//   3. grouped GEMM         output blocks = T2 blocks x environment blocks
// A GEMM task is one output block C (M x N) = sum over its K-segments s of
// opA_s (M x K_s) * opB_s (K_s x N), every operand addressed by an offset and
// a leading dimension inside one of six buffers (vector in/out, two
// environments or site tensors, two intermediates).  All tiles of all tasks
// of a stage run in ONE launch, so sectors of uneven size share the machine;
// tiles are dealt to the XCDs by work in row bands and run longest first,
// so the launch does not end on a few long tiles.  The tile shape (64x64, 32x64, 128x64, 64x128; K depth
// 16 or 32) is picked per shape class of the stage by timing the variants on
// first use.
//
// Precisions: FP64 (v_mfma_f64_16x16x4f64), and the mixed forms FP32 and
// split-BF16 (hi*hi + hi*lo + lo*hi, v_mfma_f32_16x16x32_bf16) with FP32
// accumulation.  The mixed kernels read the FP64 operands and round them while
// staging to LDS, so the plan, the buffers and the Krylov vectors are the same
// for every precision; only the matrix-core arithmetic changes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <map>

#include "mh_gemm_common.h"

namespace mh {

typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

struct GTask {      // one output block
  int M, N, seg0, nseg;
  int64_t c_off;    // element offset of C in its buffer
  int ldc, pad;
};
struct GSeg {       // one K-segment of a task
  int64_t a_off, b_off;
  int K, lda, ldb, pad;
};
struct GTile {      // task < 0: padding entry of the XCD-ordered grid
  int task, m0, n0, pad;   // task: a stack; m0: first row within the stack
};
// Tasks that share N and the B side of every K-segment (offset, K, ldb) are
// stacked along M: one GEMM of sum(M) rows whose row r comes from task
// rows[row0 + r].task, local row rows[row0 + r].row.  A tile then covers the
// rows of several small blocks instead of padding each to the tile height.
struct GStack {
  int M, N, nseg, task0;   // task0: the task whose segments give the B side
  int row0, pad;
};
struct GRow {
  int task, row;
};
struct MChunk {     // up to 1024 consecutive elements of one destination block
  int64_t dst_off;
  int rows, ld_dst;   // destination block rows and leading dimension
  int e0, ne;         // element range (column-major index within the block)
  int c0, nc;         // contributions
};
struct MContrib {
  int64_t src_off;
  int ld_src, pad;
  double coef;
};

constexpr int NSRC = 6;

// Tile shapes of the grouped GEMM: every stage carries one tile list per
// shape; the variants below pick a shape and a K depth.
constexpr int NTS = 4;
constexpr int kTileBM[NTS] = {64, 32, 128, 64};
constexpr int kTileBN[NTS] = {64, 64, 64, 128};
struct GgVariant {
  int bm, bn, bk, list;
};
constexpr GgVariant kGgF64[] = {{64, 64, 16, 0}, {64, 64, 32, 0}, {32, 64, 16, 1}, {128, 64, 16, 2},
                                {64, 128, 16, 3}};
constexpr GgVariant kGgMx[] = {{64, 64, 32, 0}, {32, 64, 32, 1}, {128, 64, 32, 2}};
constexpr int kNF64 = 5, kNMx = 3;

struct BsStage {
  int kind;                  // 0 grouped GEMM, 1 block mode product
  int a_src, b_src, c_src;   // mode: a_src = input, c_src = output
  int amc, bkc;
  int nchunk;
  int ntile_v[NTS];          // tiles per shape (0: nothing to do)
  mutable int choice;        // tuned variant (-1: not yet chosen)
  uint64_t key;              // shape class for the tuner
  double flops;
  // Device copies of the tables (set when the plan is uploaded).
  GTile *tiles_v[NTS];
  GStack *stacks;
  GRow *rows;
  GTask *tasks;
  GSeg *segs;
  MChunk *chunks;
  MContrib *contribs;
  // Host tables: the plan is built on the host and uploaded on first use.
  std::vector<GTile> h_tiles[NTS];
  std::vector<GStack> h_stacks;
  std::vector<GRow> h_rows;
  std::vector<GTask> h_tasks;
  std::vector<GSeg> h_segs;
  std::vector<MChunk> h_chunks;
  std::vector<MContrib> h_contribs;
  std::vector<int64_t> h_doff;   // mode destinations as added (for inspection)
  std::vector<int32_t> h_di;
};

}  // namespace mh

struct mh_bs_plan_s {
  int prec = MH_PREC_F64;
  int64_t size[mh::NSRC] = {0, 0, 0, 0, 0, 0};  // elements of each buffer
  std::vector<mh::BsStage> stages;
  double flops = 0;
  // One device allocation holds every table; made by the first apply on a
  // handle, which fixes the plan's device.
  void *dev = nullptr;
  int device = -1;
};

namespace mh {

// ------------------------------------------------------------------ FP64
// Staging as in the strided FP64 GEMM: an operand contiguous along its row
// index r (A m-contiguous, B n-contiguous) goes to LDS as [k][r] (pitch
// BR + 4), one contiguous along k as [r][k] (pitch BK + 2); each thread moves
// pairs adjacent along the contiguous dimension.  MFMA sub-step s of lane
// group lk uses k = 4*lk + s inside each 16-deep slice.
template <int BR, int BK, bool RMAJ>
struct GStage {
  static constexpr int PITCH = RMAJ ? BR + 4 : BK + 2;
  static constexpr int SIZE = RMAJ ? BK * PITCH : BR * PITCH;
  static constexpr int NPAIR = BR * BK / 512;
  __device__ static void coord(int t, int i, int &r, int &k) {
    const int e = t + 256 * i;
    if (RMAJ) {
      r = 2 * (e % (BR / 2));
      k = e / (BR / 2);
    } else {
      k = 2 * (e % (BK / 2));
      r = e / (BK / 2);
    }
  }
};

template <int BM, int BN, int BK, bool AMC, bool BKC>
__global__ void __launch_bounds__(256)
    k_ggemm(const GTile *__restrict__ tiles, const GStack *__restrict__ stacks,
            const GRow *__restrict__ rows, const GTask *__restrict__ tasks,
            const GSeg *__restrict__ segs, const double *__restrict__ Abase,
            const double *__restrict__ Bbase, double *__restrict__ Cbase) {
  constexpr int WTM = BM / 2, WTN = BN / 2;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  using SA = GStage<BM, BK, AMC>;
  using SB = GStage<BN, BK, !BKC>;
  __shared__ __align__(16) double sA[2][SA::SIZE];
  __shared__ __align__(16) double sB[2][SB::SIZE];
  const GTile tl = tiles[blockIdx.x];
  if (tl.task < 0) return;
  const GStack S = stacks[tl.task];
  const GTask t = tasks[S.task0];  // B side and K of every segment
  const int m0 = tl.m0, n0 = tl.n0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int lr = lane & 15, lk = lane >> 4;
  // the task (its first segment) and local row of each staged A row,
  // clamped into the stack (values of rows past it are masked at the store)
  int aseg[GStage<BM, BK, AMC>::NPAIR][2], alrow[GStage<BM, BK, AMC>::NPAIR][2];
#pragma unroll
  for (int i = 0; i < GStage<BM, BK, AMC>::NPAIR; ++i) {
    int r, k;
    GStage<BM, BK, AMC>::coord(tid, i, r, k);
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const GRow rw = rows[S.row0 + min(m0 + r + (AMC ? e : 0), S.M - 1)];
      aseg[i][e] = tasks[rw.task].seg0;
      alrow[i][e] = rw.row;
    }
  }

  d4_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};

  for (int sg = 0; sg < S.nseg; ++sg) {
    const GSeg g = segs[t.seg0 + sg];
    const double *B = Bbase + g.b_off;
    const int64_t b_sr = BKC ? g.ldb : 1, b_sk = BKC ? 1 : g.ldb;
    const int K = g.K;
    // each staged A row reads its own task's segment (offset, lda); rows
    // clamped into the stack and into the block are masked at the LDS store
    const double *aptr[SA::NPAIR][2];
    int64_t aks[SA::NPAIR][2], boff[SB::NPAIR][2];
    bool aok[SA::NPAIR][2], bok[SB::NPAIR][2];
#pragma unroll
    for (int i = 0; i < SA::NPAIR; ++i) {
      int r, k;
      SA::coord(tid, i, r, k);
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const GSeg ga = segs[aseg[i][e] + sg];
        aok[i][e] = m0 + r + (AMC ? e : 0) < S.M;
        aptr[i][e] = Abase + ga.a_off + (int64_t)alrow[i][e] * (AMC ? 1 : ga.lda);
        aks[i][e] = AMC ? ga.lda : 1;
      }
    }
#pragma unroll
    for (int i = 0; i < SB::NPAIR; ++i) {
      int r, k;
      SB::coord(tid, i, r, k);
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int gr = n0 + r + (BKC ? 0 : e);
        bok[i][e] = gr < S.N;
        boff[i][e] = (int64_t)min(gr, S.N - 1) * b_sr;
      }
    }
    double ra[SA::NPAIR][2], rb[SB::NPAIR][2];
    auto load = [&](int k0) {
#pragma unroll
      for (int i = 0; i < SA::NPAIR; ++i) {
        int r, k;
        SA::coord(tid, i, r, k);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int kc = min(k0 + k + (AMC ? 0 : e), K - 1);
          ra[i][e] = aptr[i][e][(int64_t)kc * aks[i][e]];
        }
      }
#pragma unroll
      for (int i = 0; i < SB::NPAIR; ++i) {
        int r, k;
        SB::coord(tid, i, r, k);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int kc = min(k0 + k + (BKC ? e : 0), K - 1);
          rb[i][e] = B[boff[i][e] + (int64_t)kc * b_sk];
        }
      }
    };
    auto store = [&](int buf, int k0) {
#pragma unroll
      for (int i = 0; i < SA::NPAIR; ++i) {
        int r, k;
        SA::coord(tid, i, r, k);
        const bool ok0 = aok[i][0] && (k0 + k < K);
        const bool ok1 = aok[i][1] && (k0 + k + (AMC ? 0 : 1) < K);
        const int o = AMC ? k * SA::PITCH + r : r * SA::PITCH + k;
        double2 v;
        v.x = ok0 ? ra[i][0] : 0.0;
        v.y = ok1 ? ra[i][1] : 0.0;
        *reinterpret_cast<double2 *>(&sA[buf][o]) = v;
      }
#pragma unroll
      for (int i = 0; i < SB::NPAIR; ++i) {
        int r, k;
        SB::coord(tid, i, r, k);
        const bool ok0 = bok[i][0] && (k0 + k < K);
        const bool ok1 = bok[i][1] && (k0 + k + (BKC ? 1 : 0) < K);
        const int o = BKC ? r * SB::PITCH + k : k * SB::PITCH + r;
        double2 v;
        v.x = ok0 ? rb[i][0] : 0.0;
        v.y = ok1 ? rb[i][1] : 0.0;
        *reinterpret_cast<double2 *>(&sB[buf][o]) = v;
      }
    };
    auto compute = [&](int buf) {
#pragma unroll
      for (int kb = 0; kb < BK; kb += 16) {
        double a[TM][4], b[TN][4];
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          const int r = wm * WTM + i * 16 + lr;
          if (AMC) {
#pragma unroll
            for (int s = 0; s < 4; ++s) a[i][s] = sA[buf][(kb + 4 * lk + s) * SA::PITCH + r];
          } else {
            const double2 *p = reinterpret_cast<const double2 *>(&sA[buf][r * SA::PITCH + kb + 4 * lk]);
            const double2 u = p[0], v = p[1];
            a[i][0] = u.x; a[i][1] = u.y; a[i][2] = v.x; a[i][3] = v.y;
          }
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const int r = wn * WTN + j * 16 + lr;
          if (!BKC) {
#pragma unroll
            for (int s = 0; s < 4; ++s) b[j][s] = sB[buf][(kb + 4 * lk + s) * SB::PITCH + r];
          } else {
            const double2 *p = reinterpret_cast<const double2 *>(&sB[buf][r * SB::PITCH + kb + 4 * lk]);
            const double2 u = p[0], v = p[1];
            b[j][0] = u.x; b[j][1] = u.y; b[j][2] = v.x; b[j][3] = v.y;
          }
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
    __syncthreads();  // the previous segment's last compute is done with LDS
    load(0);
    store(0, 0);
    __syncthreads();
    for (int kt = 0; kt + 1 < nk; ++kt) {
      load((kt + 1) * BK);
      compute(kt & 1);
      store((kt + 1) & 1, (kt + 1) * BK);
      __syncthreads();
    }
    compute((nk - 1) & 1);
  }

  // FP64 accumulator register r of lane l holds C[l/16 + 4r][l%16]; stack
  // row m goes to its task's block.
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + wm * WTM + i * 16 + lk + 4 * r;
      if (m >= S.M) continue;
      const GRow rw = rows[S.row0 + m];
      const GTask tj = tasks[rw.task];
      double *C = Cbase + tj.c_off + rw.row;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int n = n0 + wn * WTN + j * 16 + lr;
        if (n < S.N) C[(int64_t)tj.ldc * n] = acc[i][j][r];
      }
    }
}

// ------------------------------------------------------------- mixed
// Both operands are staged k-major ([r][k]) as 32-bit words: the FP32 value,
// or a BF16 (hi, lo) pair split into two planes.  A lane's fragment (4
// consecutive k for FP32, 8 for BF16) is one 16-byte LDS read.
constexpr int PM_F32 = 1, PM_BF16 = 2;

__device__ __forceinline__ uint32_t round_word(double x, int pm) {
  const float f = (float)x;
  if (pm == PM_F32) return __builtin_bit_cast(uint32_t, f);
  const __bf16 hi = (__bf16)f;
  const __bf16 lo = (__bf16)(f - (float)hi);
  return (uint32_t)__builtin_bit_cast(uint16_t, hi) |
         ((uint32_t)__builtin_bit_cast(uint16_t, lo) << 16);
}

template <int BR, bool RMAJ>
struct XStage {
  static constexpr int BK = 32;
  static constexpr int NPAIR = BR * BK / 512;
  __device__ static void coord(int t, int i, int &r, int &k) {
    const int e = t + 256 * i;
    if (RMAJ) {
      r = e % BR;
      k = 2 * (e / BR);
    } else {
      k = 2 * (e % (BK / 2));
      r = e / (BK / 2);
    }
  }
};

template <int PM, int BM, int BN, bool AMC, bool BKC>
__global__ void __launch_bounds__(256)
    k_ggemm_mx(const GTile *__restrict__ tiles, const GStack *__restrict__ stacks,
               const GRow *__restrict__ rows, const GTask *__restrict__ tasks,
               const GSeg *__restrict__ segs, const double *__restrict__ Abase,
               const double *__restrict__ Bbase, double *__restrict__ Cbase) {
  constexpr int BK = 32;
  constexpr int WTM = BM / 2, WTN = BN / 2;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  constexpr int NS = PM == PM_BF16 ? 2 : 1;
  constexpr int PITCH = PM == PM_BF16 ? BK + 8 : BK + 4;
  using E = typename std::conditional<PM == PM_BF16, uint16_t, float>::type;
  using SA = XStage<BM, AMC>;
  using SB = XStage<BN, !BKC>;
  __shared__ __align__(16) E sA[NS][2][BM * PITCH];
  __shared__ __align__(16) E sB[NS][2][BN * PITCH];
  const GTile tl = tiles[blockIdx.x];
  if (tl.task < 0) return;
  const GStack S = stacks[tl.task];
  const GTask t = tasks[S.task0];  // B side and K of every segment
  const int m0 = tl.m0, n0 = tl.n0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int lr = lane & 15, lk = lane >> 4;
  int aseg[SA::NPAIR], alrow[SA::NPAIR];
#pragma unroll
  for (int i = 0; i < SA::NPAIR; ++i) {
    int r, k;
    SA::coord(tid, i, r, k);
    const GRow rw = rows[S.row0 + min(m0 + r, S.M - 1)];
    aseg[i] = tasks[rw.task].seg0;
    alrow[i] = rw.row;
  }

  f4_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

  for (int sg = 0; sg < S.nseg; ++sg) {
    const GSeg g = segs[t.seg0 + sg];
    const double *B = Bbase + g.b_off;
    const int64_t b_sr = BKC ? g.ldb : 1, b_sk = BKC ? 1 : g.ldb;
    const int K = g.K;
    const double *aptr[SA::NPAIR];
    int64_t aks[SA::NPAIR], boff[SB::NPAIR];
    bool aok[SA::NPAIR], bok[SB::NPAIR];
#pragma unroll
    for (int i = 0; i < SA::NPAIR; ++i) {
      int r, k;
      SA::coord(tid, i, r, k);
      const GSeg ga = segs[aseg[i] + sg];
      aok[i] = m0 + r < S.M;
      aptr[i] = Abase + ga.a_off + (int64_t)alrow[i] * (AMC ? 1 : ga.lda);
      aks[i] = AMC ? ga.lda : 1;
    }
#pragma unroll
    for (int i = 0; i < SB::NPAIR; ++i) {
      int r, k;
      SB::coord(tid, i, r, k);
      bok[i] = n0 + r < S.N;
      boff[i] = (int64_t)min(n0 + r, S.N - 1) * b_sr;
    }
    double ra[SA::NPAIR][2], rb[SB::NPAIR][2];
    auto load = [&](int k0) {
#pragma unroll
      for (int i = 0; i < SA::NPAIR; ++i) {
        int r, k;
        SA::coord(tid, i, r, k);
#pragma unroll
        for (int e = 0; e < 2; ++e) ra[i][e] = aptr[i][(int64_t)min(k0 + k + e, K - 1) * aks[i]];
      }
#pragma unroll
      for (int i = 0; i < SB::NPAIR; ++i) {
        int r, k;
        SB::coord(tid, i, r, k);
#pragma unroll
        for (int e = 0; e < 2; ++e) rb[i][e] = B[boff[i] + (int64_t)min(k0 + k + e, K - 1) * b_sk];
      }
    };
    auto store = [&](int buf, int k0) {
#pragma unroll
      for (int i = 0; i < SA::NPAIR; ++i) {
        int r, k;
        SA::coord(tid, i, r, k);
        const uint32_t w0 = (aok[i] && k0 + k < K) ? round_word(ra[i][0], PM) : 0u;
        const uint32_t w1 = (aok[i] && k0 + k + 1 < K) ? round_word(ra[i][1], PM) : 0u;
        if constexpr (PM == PM_BF16) {
          *reinterpret_cast<uint32_t *>(&sA[0][buf][r * PITCH + k]) = (w0 & 0xffffu) | (w1 << 16);
          *reinterpret_cast<uint32_t *>(&sA[1][buf][r * PITCH + k]) = (w0 >> 16) | (w1 & 0xffff0000u);
        } else {
          *reinterpret_cast<uint2 *>(&sA[0][buf][r * PITCH + k]) = make_uint2(w0, w1);
        }
      }
#pragma unroll
      for (int i = 0; i < SB::NPAIR; ++i) {
        int r, k;
        SB::coord(tid, i, r, k);
        const uint32_t w0 = (bok[i] && k0 + k < K) ? round_word(rb[i][0], PM) : 0u;
        const uint32_t w1 = (bok[i] && k0 + k + 1 < K) ? round_word(rb[i][1], PM) : 0u;
        if constexpr (PM == PM_BF16) {
          *reinterpret_cast<uint32_t *>(&sB[0][buf][r * PITCH + k]) = (w0 & 0xffffu) | (w1 << 16);
          *reinterpret_cast<uint32_t *>(&sB[1][buf][r * PITCH + k]) = (w0 >> 16) | (w1 & 0xffff0000u);
        } else {
          *reinterpret_cast<uint2 *>(&sB[0][buf][r * PITCH + k]) = make_uint2(w0, w1);
        }
      }
    };
    auto compute = [&](int buf) {
      if constexpr (PM == PM_F32) {
#pragma unroll
        for (int kb = 0; kb < BK; kb += 16) {
          f4_t a[TM], b[TN];
#pragma unroll
          for (int i = 0; i < TM; ++i)
            a[i] = *reinterpret_cast<const f4_t *>(&sA[0][buf][(wm * WTM + i * 16 + lr) * PITCH + kb + 4 * lk]);
#pragma unroll
          for (int j = 0; j < TN; ++j)
            b[j] = *reinterpret_cast<const f4_t *>(&sB[0][buf][(wn * WTN + j * 16 + lr) * PITCH + kb + 4 * lk]);
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
          ah[i] = *reinterpret_cast<const bf8_t *>(&sA[0][buf][o]);
          al[i] = *reinterpret_cast<const bf8_t *>(&sA[1][buf][o]);
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const int o = (wn * WTN + j * 16 + lr) * PITCH + 8 * lk;
          bh[j] = *reinterpret_cast<const bf8_t *>(&sB[0][buf][o]);
          bl[j] = *reinterpret_cast<const bf8_t *>(&sB[1][buf][o]);
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
    __syncthreads();
    load(0);
    store(0, 0);
    __syncthreads();
    for (int kt = 0; kt + 1 < nk; ++kt) {
      load((kt + 1) * BK);
      compute(kt & 1);
      store((kt + 1) & 1, (kt + 1) * BK);
      __syncthreads();
    }
    compute((nk - 1) & 1);
  }

  // FP32 accumulator register r of lane l holds C[4*(l/16) + r][l%16];
  // stack row m goes to its task's block.
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + wm * WTM + i * 16 + 4 * lk + r;
      if (m >= S.M) continue;
      const GRow rw = rows[S.row0 + m];
      const GTask tj = tasks[rw.task];
      double *C = Cbase + tj.c_off + rw.row;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int n = n0 + wn * WTN + j * 16 + lr;
        if (n < S.N) C[(int64_t)tj.ldc * n] = (double)acc[i][j][r];
      }
    }
}

// ---------------------------------------------------- block mode product
// One workgroup per chunk: dst[e] = sum_c coef_c * src_c[e] over the chunk's
// element range (column-major index e = r + rows * c within the block).
__global__ void __launch_bounds__(256)
    k_bmode(const MChunk *__restrict__ chunks, const MContrib *__restrict__ contribs,
            const double *__restrict__ src, double *__restrict__ dst) {
  const MChunk ch = chunks[blockIdx.x];
  for (int e = ch.e0 + threadIdx.x; e < ch.e0 + ch.ne; e += 256) {
    const int r = e % ch.rows, c = e / ch.rows;
    double acc = 0.0;
    for (int q = 0; q < ch.nc; ++q) {
      const MContrib cb = contribs[ch.c0 + q];
      acc += cb.coef * src[cb.src_off + r + (int64_t)cb.ld_src * c];
    }
    dst[ch.dst_off + r + (int64_t)ch.ld_dst * c] = acc;
  }
}

// ------------------------------------------------- coupled-sector copies
// Block (rows x cols, column-major at blk_off) <-> region of a coupled
// matrix (column-major, leading dimension ld, first element at mat_off).
// One workgroup per 4096-element chunk of a piece.
struct CPiece {
  int64_t blk_off, mat_off;
  int rows, cols, ld, e0;
};

__global__ void __launch_bounds__(256)
    k_coupled_copy(const CPiece *__restrict__ pieces, const double *__restrict__ src,
                   double *__restrict__ dst, int to_coupled) {
  const CPiece p = pieces[blockIdx.x];
  const int n = min(4096, p.rows * p.cols - p.e0);
  for (int t = threadIdx.x; t < n; t += 256) {
    const int e = p.e0 + t;
    const int i = e % p.rows, j = e / p.rows;
    const int64_t b = p.blk_off + e, m = p.mat_off + i + (int64_t)p.ld * j;
    if (to_coupled) dst[m] = src[b];
    else dst[b] = src[m];
  }
}

// pieces: {blk_off, mat_off, rows, cols, ld} per piece (host); the table goes
// to the device through the handle's small buffer.
int bs_coupled_copy(mh_handle h, const std::vector<int64_t> &pc, const double *src, double *dst,
                    int to_coupled) {
  std::vector<CPiece> tab;
  for (size_t k = 0; k + 5 <= pc.size(); k += 5) {
    const int rows = (int)pc[k + 2], cols = (int)pc[k + 3];
    for (int e0 = 0; e0 < rows * cols; e0 += 4096)
      tab.push_back(CPiece{pc[k], pc[k + 1], rows, cols, (int)pc[k + 4], e0});
  }
  if (tab.empty()) return MH_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  (void)hipStreamIsCapturing(h->stream, &cs);
  if (cs != hipStreamCaptureStatusNone)
    return fail(h, MH_ERR_ARG, "coupled-sector copies cannot be captured (they upload a table)");
  const size_t bytes = tab.size() * sizeof(CPiece);
  MH_TRY(ensure_small(h, bytes));
  MH_HIP(h, hipMemcpyAsync(h->small, tab.data(), bytes, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_coupled_copy, dim3((unsigned)tab.size()), dim3(256), 0, h->stream,
                     (const CPiece *)h->small, src, dst, to_coupled);
  MH_TRY(check_launch(h, "coupled-sector copy"));
  // the table lives in the handle's small buffer, which a later call may reuse
  MH_HIP(h, hipStreamSynchronize(h->stream));
  return MH_OK;
}

// One launch of variant v of a grouped-GEMM stage.  The layout with neither
// operand contiguous along its row index is served by the 64x64 FP64 /
// 64x64 mixed variant only (no plan of this library produces it).
template <int PM, bool AMC, bool BKC>
static void launch_gg(const BsStage &st, int v, const double *A, const double *B, double *C,
                      hipStream_t s) {
  const GgVariant g = PM == 0 ? kGgF64[v] : kGgMx[v];
  const dim3 grid(st.ntile_v[g.list]);
  const GTile *tl = st.tiles_v[g.list];
#define MH_GG(BM, BN, BK)                                                                    \
  hipLaunchKernelGGL((k_ggemm<BM, BN, BK, AMC, BKC>), grid, dim3(256), 0, s, tl, st.stacks, \
                     st.rows, st.tasks, st.segs, A, B, C)
#define MH_GX(BM, BN)                                                                         \
  hipLaunchKernelGGL((k_ggemm_mx<PM, BM, BN, AMC, BKC>), grid, dim3(256), 0, s, tl, \
                     st.stacks, st.rows, st.tasks, st.segs, A, B, C)
  if constexpr (PM == 0) {
    if constexpr (AMC || BKC) {
      if (v == 1) MH_GG(64, 64, 32);
      else if (v == 2) MH_GG(32, 64, 16);
      else if (v == 3) MH_GG(128, 64, 16);
      else if (v == 4) MH_GG(64, 128, 16);
      else MH_GG(64, 64, 16);
    } else {
      MH_GG(64, 64, 16);
    }
  } else {
    if constexpr (AMC || BKC) {
      if (v == 1) MH_GX(32, 64);
      else if (v == 2) MH_GX(128, 64);
      else MH_GX(64, 64);
    } else {
      MH_GX(64, 64);
    }
  }
#undef MH_GG
#undef MH_GX
}

template <int PM>
static void launch_pm(const BsStage &st, int v, const double *A, const double *B, double *C,
                      hipStream_t s) {
  if (st.amc && st.bkc) launch_gg<PM, true, true>(st, v, A, B, C, s);
  else if (st.amc) launch_gg<PM, true, false>(st, v, A, B, C, s);
  else if (st.bkc) launch_gg<PM, false, true>(st, v, A, B, C, s);
  else launch_gg<PM, false, false>(st, 0, A, B, C, s);
}

static void launch_stage(mh_handle h, int prec, const BsStage &st, int v, const double *A,
                         const double *B, double *C) {
  if (prec == MH_PREC_F32) launch_pm<PM_F32>(st, v, A, B, C, h->stream);
  else if (prec == MH_PREC_BF16X3) launch_pm<PM_BF16>(st, v, A, B, C, h->stream);
  else launch_pm<0>(st, v, A, B, C, h->stream);
}

// Tile-shape autotuner: the first run of a stage whose shape class is not yet
// in the handle's table times every variant (twice each, the faster counts)
// on the stage's own buffers and records the fastest; later stages of the
// same class launch it directly.  Every variant computes the same result, so
// the stage output is valid after tuning.  Not used while a stream is being
// captured or while profiling (both need plain launches).
static int pick_variant(mh_handle h, int prec, const BsStage &st, const double *A, const double *B,
                        double *C) {
  if (st.choice >= 0) return st.choice;
  const uint64_t key = st.key ^ ((uint64_t)prec << 60);
  auto it = h->bs_tune.find(key);
  if (it != h->bs_tune.end()) return st.choice = it->second;
  const int nv = (st.amc || st.bkc) ? (prec == MH_PREC_F64 ? kNF64 : kNMx) : 1;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  (void)hipStreamIsCapturing(h->stream, &cs);
  if (nv == 1 || !h->autotune || h->profiling || cs != hipStreamCaptureStatusNone)
    return st.choice = 0;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return st.choice = 0;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return st.choice = 0;
  }
  int best = 0;
  float best_ms = 1e30f;
  for (int v = 0; v < nv; ++v) {
    for (int rep = 0; rep < 2; ++rep) {
      (void)hipEventRecord(e0, h->stream);
      launch_stage(h, prec, st, v, A, B, C);
      (void)hipEventRecord(e1, h->stream);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best_ms) {
        best_ms = ms;
        best = v;
      }
    }
    if (h->tune_log)
      fprintf(stderr, "[mh tune] grouped gemm class %016llx prec %d variant %d: %.1f us\n",
              (unsigned long long)key, prec, v, 1e3 * best_ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  h->bs_tune[key] = best;
  return st.choice = best;
}

// The GEMM stage as the per-sector loop of a plain port: one strided FP64
// GEMM launch per output block and K-segment (segments after the first
// accumulate with beta = 1; a block without segments is written as zeros by
// a K = 0 launch).  The configuration comes from the GEMM heuristic: these
// products are far below the size the tuner times.
static int run_stage_plain(mh_handle h, int prec, const BsStage &st, void *const *srcs) {
  if (prec != MH_PREC_F64)
    return fail(h, MH_ERR_UNSUPPORTED, "the per-sector loop runs FP64 plans only");
  const double *A = (const double *)srcs[st.a_src];
  const double *B = (const double *)srcs[st.b_src];
  double *C = (double *)srcs[st.c_src];
  int slot;
  prof_begin(h, MH_K_GEMM, st.flops, &slot);
  for (const GTask &t : st.h_tasks) {
    const int nseg = t.nseg > 0 ? t.nseg : 1;
    for (int sg = 0; sg < nseg; ++sg) {
      GemmDesc d{};
      d.M = t.M;
      d.N = t.N;
      d.C = C + t.c_off;
      d.scm = 1;
      d.scn = t.ldc;
      d.alpha_re = 1.0;
      d.beta_re = sg > 0 ? 1.0 : 0.0;
      if (t.nseg > 0) {
        const GSeg &g = st.h_segs[t.seg0 + sg];
        d.K = g.K;
        d.A = A + g.a_off;
        d.sam = st.amc ? 1 : g.lda;
        d.sak = st.amc ? g.lda : 1;
        d.B = B + g.b_off;
        d.sbk = st.bkc ? 1 : g.ldb;
        d.sbn = st.bkc ? g.ldb : 1;
      } else {
        d.K = 0;
        d.A = A;
        d.B = B;
        d.sam = d.sbk = 1;
        d.sak = d.sbn = 1;
      }
      MH_TRY(gemm(h, MH_F64, MH_PREC_F64, d));
    }
  }
  prof_end(h, slot);
  return check_launch(h, "per-sector gemm");
}

static int run_stage(mh_handle h, int prec, const BsStage &st, void *const *srcs) {
  if (st.kind == 0 && h->bs_plain) return run_stage_plain(h, prec, st, srcs);
  if (st.kind == 0) {
    if (st.ntile_v[0] == 0) return MH_OK;
    const double *A = (const double *)srcs[st.a_src];
    const double *B = (const double *)srcs[st.b_src];
    double *C = (double *)srcs[st.c_src];
    const int v = pick_variant(h, prec, st, A, B, C);
    MH_TRY(check_launch(h, "grouped gemm tuning"));
    int slot;
    prof_begin(h, MH_K_GEMM, st.flops, &slot);
    launch_stage(h, prec, st, v, A, B, C);
    prof_end(h, slot);
    return check_launch(h, "grouped gemm");
  }
  if (st.nchunk == 0) return MH_OK;
  int slot;
  prof_begin(h, MH_K_WSTAGE, st.flops, &slot);
  hipLaunchKernelGGL(k_bmode, dim3(st.nchunk), dim3(256), 0, h->stream, st.chunks, st.contribs,
                     (const double *)srcs[st.a_src], (double *)srcs[st.c_src]);
  prof_end(h, slot);
  return check_launch(h, "block mode product");
}

static size_t al256(int64_t elems) { return ((size_t)elems * 8 + 255) & ~size_t(255); }

size_t bs_ws_bytes(const mh_bs_plan_s &p) {
  return al256(p.size[MH_BS_T1]) + al256(p.size[MH_BS_T2]);
}

size_t bs_vec_len(const mh_bs_plan_s &p) { return (size_t)p.size[MH_BS_X]; }

int bs_apply_ws(mh_handle h, const mh_bs_plan_s &p, const void *x, const void *p1,
                const void *p2, void *y, void *ws) {
  void *srcs[NSRC];
  srcs[MH_BS_X] = const_cast<void *>(x);
  srcs[MH_BS_Y] = y;
  srcs[MH_BS_P1] = const_cast<void *>(p1);
  srcs[MH_BS_P2] = const_cast<void *>(p2);
  srcs[MH_BS_T1] = ws;
  srcs[MH_BS_T2] = (char *)ws + al256(p.size[MH_BS_T1]);
  for (const auto &st : p.stages) MH_TRY(run_stage(h, p.prec, st, srcs));
  return MH_OK;
}

int bs_check_buffers(mh_handle h, const mh_bs_plan_s &p, const void *x, const void *p1,
                     const void *p2, const void *y) {
  const void *b[4] = {x, y, p1, p2};
  static const char *names[4] = {"x", "y", "p1", "p2"};
  for (int i = 0; i < 4; ++i)
    if (p.size[i] > 0 && !b[i]) return fail(h, MH_ERR_ARG, "plan needs buffer %s", names[i]);
  return MH_OK;
}

int bs_check_square(mh_handle h, const mh_bs_plan_s &p, const void *p1, const void *p2) {
  if (p.size[MH_BS_X] <= 0 || p.size[MH_BS_X] != p.size[MH_BS_Y])
    return fail(h, MH_ERR_SHAPE, "plan is not an operator (x has %lld, y %lld elements)",
                (long long)p.size[MH_BS_X], (long long)p.size[MH_BS_Y]);
  if ((p.size[MH_BS_P1] > 0 && !p1) || (p.size[MH_BS_P2] > 0 && !p2))
    return fail(h, MH_ERR_ARG, "plan needs its environment buffers");
  return MH_OK;
}

// Byte offsets of the device tables inside the plan's one allocation.
static size_t al_bytes(size_t b) { return (b + 255) & ~size_t(255); }

template <typename T>
static void lay_table(const std::vector<T> &v, size_t &at, std::vector<size_t> &offs) {
  offs.push_back(at);
  at += al_bytes(v.size() * sizeof(T));
}

// Upload the tables of a host-built plan to the handle's device (once); a
// plan stays on the device of its first use.
int bs_prepare(mh_handle h, mh_bs_plan_s &p) {
  if (p.dev) {
    if (p.device != h->device)
      return fail(h, MH_ERR_ARG, "plan lives on device %d, handle on %d", p.device, h->device);
    return MH_OK;
  }
  size_t total = 0;
  std::vector<size_t> offs;
  for (const auto &st : p.stages) {
    for (int sh = 0; sh < NTS; ++sh) lay_table(st.h_tiles[sh], total, offs);
    lay_table(st.h_stacks, total, offs);
    lay_table(st.h_rows, total, offs);
    lay_table(st.h_tasks, total, offs);
    lay_table(st.h_segs, total, offs);
    lay_table(st.h_chunks, total, offs);
    lay_table(st.h_contribs, total, offs);
  }
  if (total == 0) total = 256;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  (void)hipStreamIsCapturing(h->stream, &cs);
  if (cs != hipStreamCaptureStatusNone)
    return fail(h, MH_ERR_ARG, "a plan's first use cannot be captured (it uploads its tables)");
  std::vector<unsigned char> host(total, 0);
  size_t k = 0;
  auto put = [&](const auto &v) {
    if (!v.empty()) std::memcpy(host.data() + offs[k], v.data(), v.size() * sizeof(v[0]));
    ++k;
  };
  for (const auto &st : p.stages) {
    for (int sh = 0; sh < NTS; ++sh) put(st.h_tiles[sh]);
    put(st.h_stacks);
    put(st.h_rows);
    put(st.h_tasks);
    put(st.h_segs);
    put(st.h_chunks);
    put(st.h_contribs);
  }
  void *d = nullptr;
  if (hipMalloc(&d, total) != hipSuccess) {
    (void)hipGetLastError();
    return fail(h, MH_ERR_OOM, "plan allocation of %zu bytes", total);
  }
  const hipError_t e1 = hipMemcpyAsync(d, host.data(), total, hipMemcpyHostToDevice, h->stream);
  const hipError_t e2 = e1 == hipSuccess ? hipStreamSynchronize(h->stream) : e1;
  if (e2 != hipSuccess) {
    (void)hipFree(d);
    return fail(h, MH_ERR_HIP, "plan upload: %s", hipGetErrorString(e2));
  }
  p.dev = d;
  p.device = h->device;
  unsigned char *base = (unsigned char *)d;
  k = 0;
  auto at = [&](const auto &v) -> void * {
    void *r = v.empty() ? nullptr : (void *)(base + offs[k]);
    ++k;
    return r;
  };
  for (auto &st : p.stages) {
    for (int sh = 0; sh < NTS; ++sh) st.tiles_v[sh] = (GTile *)at(st.h_tiles[sh]);
    st.stacks = (GStack *)at(st.h_stacks);
    st.rows = (GRow *)at(st.h_rows);
    st.tasks = (GTask *)at(st.h_tasks);
    st.segs = (GSeg *)at(st.h_segs);
    st.chunks = (MChunk *)at(st.h_chunks);
    st.contribs = (MContrib *)at(st.h_contribs);
  }
  return MH_OK;
}

// Largest element offset touched by a (rows x cols) operand at `off` with
// leading dimension ld, contiguous along rows (col-major) or along cols.
static int64_t last_elem(int64_t off, int rows, int cols, int ld, bool contig_rows) {
  return contig_rows ? off + (rows - 1) + (int64_t)(cols - 1) * ld
                     : off + (int64_t)(rows - 1) * ld + (cols - 1);
}

}  // namespace mh

using namespace mh;

extern "C" {

int mh_bs_plan_create(mh_handle h, int prec, const int64_t *sizes, mh_bs_plan *out) {
  if (!out || !sizes) return MH_ERR_ARG;
  *out = nullptr;
  if (prec != MH_PREC_F64 && prec != MH_PREC_F32 && prec != MH_PREC_BF16X3)
    return fail(h, MH_ERR_ARG, "unknown precision %d", prec);
  for (int i = 0; i < NSRC; ++i)
    if (sizes[i] < 0) return fail(h, MH_ERR_SHAPE, "negative buffer size");
  auto *p = new mh_bs_plan_s();
  p->prec = prec;
  for (int i = 0; i < NSRC; ++i) p->size[i] = sizes[i];
  *out = p;
  return MH_OK;
}

int mh_bs_plan_add_gemm(mh_handle h, mh_bs_plan plan, int a_src, int b_src, int c_src,
                        int amc, int bkc, int ntask, const int32_t *task_i,
                        const int64_t *task_coff, int nseg, const int64_t *seg_off,
                        const int32_t *seg_i) {
  if (!plan) return MH_ERR_ARG;
  if (ntask < 0 || nseg < 0) return fail(h, MH_ERR_SHAPE, "negative task count");
  if ((ntask > 0 && (!task_i || !task_coff)) || (nseg > 0 && (!seg_off || !seg_i)))
    return fail(h, MH_ERR_ARG, "null task or segment table");
  for (int s : {a_src, b_src, c_src})
    if (s < 0 || s >= NSRC) return fail(h, MH_ERR_ARG, "bad operand buffer %d", s);
  if (c_src == a_src || c_src == b_src)
    return fail(h, MH_ERR_ARG, "gemm output aliases an input");
  const int64_t asz = plan->size[a_src], bsz = plan->size[b_src], csz = plan->size[c_src];
  std::vector<GSeg> segs(nseg);
  for (int s = 0; s < nseg; ++s) {
    segs[s] = GSeg{seg_off[2 * s], seg_off[2 * s + 1], seg_i[3 * s], seg_i[3 * s + 1],
                   seg_i[3 * s + 2], 0};
    if (segs[s].K <= 0 || segs[s].lda <= 0 || segs[s].ldb <= 0 || segs[s].a_off < 0 ||
        segs[s].b_off < 0)
      return fail(h, MH_ERR_SHAPE, "bad segment %d", s);
  }
  std::vector<GTask> tasks(ntask);
  double flops = 0;
  for (int i = 0; i < ntask; ++i) {
    // task_i rows: M, N, seg0, nseg, ldc
    GTask t{task_i[5 * i], task_i[5 * i + 1], task_i[5 * i + 2], task_i[5 * i + 3],
            task_coff[i], task_i[5 * i + 4], 0};
    if (t.M <= 0 || t.N <= 0 || t.nseg <= 0 || t.seg0 < 0 || t.seg0 + t.nseg > nseg ||
        t.ldc < t.M || t.c_off < 0 || last_elem(t.c_off, t.M, t.N, t.ldc, true) >= csz)
      return fail(h, MH_ERR_SHAPE, "bad task %d (M %d N %d ldc %d)", i, t.M, t.N, t.ldc);
    for (int s = t.seg0; s < t.seg0 + t.nseg; ++s) {
      const GSeg &g = segs[s];
      // A is M x K (m-contiguous when amc), B is K x N (k-contiguous when bkc)
      if ((amc ? g.lda < t.M : g.lda < g.K) || (bkc ? g.ldb < g.K : g.ldb < t.N) ||
          last_elem(g.a_off, t.M, g.K, g.lda, amc) >= asz ||
          last_elem(g.b_off, g.K, t.N, g.ldb, bkc) >= bsz)
        return fail(h, MH_ERR_SHAPE, "segment %d of task %d lies outside its buffers", s, i);
      flops += 2.0 * t.M * (double)t.N * g.K;
    }
    tasks[i] = t;
  }
  BsStage st{};
  st.kind = 0;
  st.a_src = a_src;
  st.b_src = b_src;
  st.c_src = c_src;
  st.amc = amc ? 1 : 0;
  st.bkc = bkc ? 1 : 0;
  st.flops = flops;
  st.choice = -1;
  // Stacks: tasks with the same N and the same B side of every K-segment
  // (b_off, K, ldb) in the same order, in order of first appearance.
  std::vector<GStack> stacks;
  std::vector<GRow> rowmap;
  {
    std::map<std::vector<int64_t>, int> key2stack;
    std::vector<std::vector<int>> members;
    for (int i = 0; i < ntask; ++i) {
      std::vector<int64_t> key{tasks[i].N, tasks[i].nseg};
      for (int s = tasks[i].seg0; s < tasks[i].seg0 + tasks[i].nseg; ++s) {
        key.push_back(segs[s].b_off);
        key.push_back(segs[s].K);
        key.push_back(segs[s].ldb);
      }
      auto it = key2stack.find(key);
      if (it == key2stack.end()) {
        key2stack.emplace(std::move(key), (int)members.size());
        members.push_back({i});
      } else {
        members[it->second].push_back(i);
      }
    }
    for (const auto &mem : members) {
      GStack sk{0, tasks[mem[0]].N, tasks[mem[0]].nseg, mem[0], (int)rowmap.size(), 0};
      for (int i : mem) {
        for (int r = 0; r < tasks[i].M; ++r) rowmap.push_back(GRow{i, r});
        sk.M += tasks[i].M;
      }
      stacks.push_back(sk);
    }
  }
  std::vector<std::vector<GTile>> lists(NTS);
  for (int sh = 0; sh < NTS; ++sh) {
    const int bm = kTileBM[sh], bn = kTileBN[sh];
    // Tiles of one stack all cost the same K loop (the sum over its segments
    // of 16-deep steps, plus one for the tile's set-up); stacks differ by up
    // to an order of magnitude.  Workgroups are dispatched in list order,
    // round-robin over the 8 XCDs (block b runs on XCD b % 8), each with its
    // own L2.  The unit of assignment is a row band (all tiles of one stack
    // at one m0, which share their A rows): bands are dealt longest tile
    // first, each to the XCD with the least work so far, and every XCD runs
    // its tiles in that order.  Against stack order this takes each stage of
    // a D = 2048 Hubbard bond 15-30% faster (no launch ends on a few long
    // tiles) while a band's A rows stay in one L2.
    std::vector<std::vector<GTile>> per_xcd(8);
    {
      struct Band {
        int q, m0, ntn, k;
      };
      std::vector<Band> bands;
      for (size_t q = 0; q < stacks.size(); ++q) {
        int k = 0;
        for (int sg = 0; sg < stacks[q].nseg; ++sg)
          k += (segs[tasks[stacks[q].task0].seg0 + sg].K + 15) / 16 + 1;
        const int ntn = (stacks[q].N + bn - 1) / bn;
        for (int m0 = 0; m0 < stacks[q].M; m0 += bm) bands.push_back(Band{(int)q, m0, ntn, k});
      }
      std::stable_sort(bands.begin(), bands.end(), [](const Band &x, const Band &y) { return x.k > y.k; });
      int64_t load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (const Band &b : bands) {
        int x = 0;
        for (int i = 1; i < 8; ++i)
          if (load[i] < load[x]) x = i;
        load[x] += (int64_t)b.k * b.ntn;
        for (int t = 0; t < b.ntn; ++t) per_xcd[x].push_back(GTile{b.q, b.m0, t * bn, 0});
      }
    }
    size_t longest = 0;
    for (auto &v : per_xcd) longest = std::max(longest, v.size());
    const int nl = (int)longest * 8;
    lists[sh].assign((size_t)nl, GTile{-1, 0, 0, 0});
    for (int x = 0; x < 8; ++x)
      for (size_t i = 0; i < per_xcd[x].size(); ++i) lists[sh][i * 8 + x] = per_xcd[x][i];
    st.ntile_v[sh] = longest > 0 ? nl : 0;
    st.tiles_v[sh] = nullptr;
  }
  // Shape class for the tuner: layout, and log2 buckets of the mean task
  // extents and of the task count.
  if (ntask > 0) {
    double sm = 0, sn = 0, sk = 0;
    for (int i = 0; i < ntask; ++i) {
      sm += tasks[i].M;
      sn += tasks[i].N;
      for (int s = tasks[i].seg0; s < tasks[i].seg0 + tasks[i].nseg; ++s) sk += segs[s].K;
    }
    auto lb = [](double v) { return (uint64_t)std::max(0, (int)std::lround(std::log2(std::max(v, 1.0)))); };
    st.key = (uint64_t)(st.amc | (st.bkc << 1)) | (lb(sm / ntask) << 2) | (lb(sn / ntask) << 12) |
             (lb(sk / ntask) << 22) | (lb((double)ntask) << 32);
  }
  st.h_tasks = std::move(tasks);
  st.h_segs = std::move(segs);
  st.h_stacks = std::move(stacks);
  st.h_rows = std::move(rowmap);
  for (int sh = 0; sh < NTS; ++sh)
    if (st.ntile_v[sh] > 0) st.h_tiles[sh] = std::move(lists[sh]);
  plan->stages.push_back(std::move(st));
  plan->flops += flops;
  return MH_OK;
}

int mh_bs_plan_add_mode(mh_handle h, mh_bs_plan plan, int in_src, int out_src, int ndst,
                        const int64_t *dst_off, const int32_t *dst_i, int ncontrib,
                        const int64_t *src_off, const int32_t *src_ld, const double *coef) {
  if (!plan) return MH_ERR_ARG;
  if (ndst < 0 || ncontrib < 0) return fail(h, MH_ERR_SHAPE, "negative count");
  if ((ndst > 0 && (!dst_off || !dst_i)) || (ncontrib > 0 && (!src_off || !src_ld || !coef)))
    return fail(h, MH_ERR_ARG, "null block or contribution table");
  if (in_src < 0 || in_src >= NSRC || out_src < 0 || out_src >= NSRC || in_src == out_src)
    return fail(h, MH_ERR_ARG, "bad mode-product buffers");
  const int64_t isz = plan->size[in_src], osz = plan->size[out_src];
  std::vector<MChunk> chunks;
  double flops = 0;
  for (int d = 0; d < ndst; ++d) {
    // dst_i rows: rows, cols, ld_dst, c0, nc
    const int rows = dst_i[5 * d], cols = dst_i[5 * d + 1], ld = dst_i[5 * d + 2];
    const int c0 = dst_i[5 * d + 3], nc = dst_i[5 * d + 4];
    if (rows <= 0 || cols <= 0 || ld < rows || c0 < 0 || nc < 0 || c0 + nc > ncontrib ||
        dst_off[d] < 0 || last_elem(dst_off[d], rows, cols, ld, true) >= osz)
      return fail(h, MH_ERR_SHAPE, "bad destination block %d", d);
    for (int c = c0; c < c0 + nc; ++c)
      if (src_ld[c] < rows || src_off[c] < 0 ||
          last_elem(src_off[c], rows, cols, src_ld[c], true) >= isz)
        return fail(h, MH_ERR_SHAPE, "contribution %d of block %d lies outside its buffer", c, d);
    const int n = rows * cols;
    for (int e0 = 0; e0 < n; e0 += 1024)
      chunks.push_back(MChunk{dst_off[d], rows, ld, e0, std::min(1024, n - e0), c0, nc});
    flops += 2.0 * n * nc;
  }
  std::vector<MContrib> cbs(ncontrib);
  for (int c = 0; c < ncontrib; ++c) cbs[c] = MContrib{src_off[c], src_ld[c], 0, coef[c]};
  BsStage st{};
  st.kind = 1;
  st.a_src = in_src;
  st.c_src = out_src;
  st.nchunk = (int)chunks.size();
  st.flops = flops;
  st.h_chunks = std::move(chunks);
  st.h_contribs = std::move(cbs);
  st.h_doff.assign(dst_off, dst_off + ndst);
  st.h_di.assign(dst_i, dst_i + 5 * (size_t)ndst);
  plan->stages.push_back(std::move(st));
  plan->flops += flops;
  return MH_OK;
}

int mh_bs_plan_destroy(mh_handle h, mh_bs_plan plan) {
  if (!plan) return MH_ERR_ARG;
  if (plan->dev) {
    // work queued on the handle may still read the tables
    if (h) (void)hipStreamSynchronize(h->stream);
    (void)hipFree(plan->dev);
  }
  delete plan;
  return MH_OK;
}

int mh_bs_plan_variants(mh_bs_plan plan, int *n) {
  if (!plan || !n) return MH_ERR_ARG;
  *n = plan->prec == MH_PREC_F64 ? kNF64 : kNMx;
  return MH_OK;
}

int mh_bs_plan_force_variant(mh_bs_plan plan, int variant) {
  if (!plan) return MH_ERR_ARG;
  const int nv = plan->prec == MH_PREC_F64 ? kNF64 : kNMx;
  if (variant < -1 || variant >= nv) return MH_ERR_ARG;
  for (auto &st : plan->stages)
    if (st.kind == 0) st.choice = (st.amc || st.bkc) ? variant : (variant < 0 ? -1 : 0);
  return MH_OK;
}

int mh_bs_plan_flops(mh_bs_plan plan, double *flops) {
  if (!plan || !flops) return MH_ERR_ARG;
  *flops = plan->flops;
  return MH_OK;
}

int mh_bs_plan_sizes(mh_bs_plan plan, int64_t *sizes) {
  if (!plan || !sizes) return MH_ERR_ARG;
  for (int i = 0; i < NSRC; ++i) sizes[i] = plan->size[i];
  return MH_OK;
}

int mh_bs_plan_stage_count(mh_bs_plan plan, int *n) {
  if (!plan || !n) return MH_ERR_ARG;
  *n = (int)plan->stages.size();
  return MH_OK;
}

int mh_bs_plan_stage_info(mh_bs_plan plan, int stage, int32_t *info) {
  if (!plan || !info || stage < 0 || stage >= (int)plan->stages.size()) return MH_ERR_ARG;
  const BsStage &st = plan->stages[stage];
  if (st.kind == 0) {
    const int32_t v[8] = {0, st.a_src, st.b_src, st.c_src, st.amc, st.bkc,
                          (int32_t)st.h_tasks.size(), (int32_t)st.h_segs.size()};
    std::memcpy(info, v, sizeof(v));
  } else {
    const int32_t v[8] = {1, st.a_src, st.c_src, 0, 0, 0, (int32_t)st.h_doff.size(),
                          (int32_t)st.h_contribs.size()};
    std::memcpy(info, v, sizeof(v));
  }
  return MH_OK;
}

int mh_bs_plan_stage_gemm(mh_bs_plan plan, int stage, int32_t *task_i, int64_t *task_coff,
                          int64_t *seg_off, int32_t *seg_i) {
  if (!plan || stage < 0 || stage >= (int)plan->stages.size()) return MH_ERR_ARG;
  const BsStage &st = plan->stages[stage];
  if (st.kind != 0) return MH_ERR_ARG;
  for (size_t i = 0; i < st.h_tasks.size(); ++i) {
    const GTask &t = st.h_tasks[i];
    if (task_i) {
      const int32_t v[5] = {t.M, t.N, t.seg0, t.nseg, t.ldc};
      std::memcpy(task_i + 5 * i, v, sizeof(v));
    }
    if (task_coff) task_coff[i] = t.c_off;
  }
  for (size_t i = 0; i < st.h_segs.size(); ++i) {
    const GSeg &g = st.h_segs[i];
    if (seg_off) {
      seg_off[2 * i] = g.a_off;
      seg_off[2 * i + 1] = g.b_off;
    }
    if (seg_i) {
      const int32_t v[3] = {g.K, g.lda, g.ldb};
      std::memcpy(seg_i + 3 * i, v, sizeof(v));
    }
  }
  return MH_OK;
}

int mh_bs_plan_stage_mode(mh_bs_plan plan, int stage, int64_t *dst_off, int32_t *dst_i,
                          int64_t *src_off, int32_t *src_ld, double *coef) {
  if (!plan || stage < 0 || stage >= (int)plan->stages.size()) return MH_ERR_ARG;
  const BsStage &st = plan->stages[stage];
  if (st.kind != 1) return MH_ERR_ARG;
  if (dst_off) std::copy(st.h_doff.begin(), st.h_doff.end(), dst_off);
  if (dst_i) std::copy(st.h_di.begin(), st.h_di.end(), dst_i);
  for (size_t c = 0; c < st.h_contribs.size(); ++c) {
    if (src_off) src_off[c] = st.h_contribs[c].src_off;
    if (src_ld) src_ld[c] = st.h_contribs[c].ld_src;
    if (coef) coef[c] = st.h_contribs[c].coef;
  }
  return MH_OK;
}

int mh_bs_apply(mh_handle h, mh_bs_plan plan, const void *x, const void *p1, const void *p2,
                void *y) {
  if (!h || !plan) return MH_ERR_ARG;
  MH_TRY(bs_check_buffers(h, *plan, x, p1, p2, y));
  MH_TRY(bs_prepare(h, *plan));
  MH_TRY(ensure_ws(h, bs_ws_bytes(*plan)));
  return bs_apply_ws(h, *plan, x, p1, p2, y, h->ws);
}

}  // extern "C"
