// Batched strided FP64 GEMMs of unrelated shapes in one launch: the per-sector
// products of the graded two-site split (density matrices, projections of the
// two-site matrix on eigenvector blocks, the second level of the exact cut),
// which a per-sector loop would issue as dozens of small launches per bond.
//
// A task is C = alpha op(A) op(B) + beta C with arbitrary element strides on
// every operand (a transposed operand is a stride swap), given by device
// pointers; the host uploads the task table once and the grid covers every
// task's 64 x 64 output tiles.  Each workgroup stages 64 x 16 panels of A and
// B in LDS and runs them on FP64 MFMA.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "mh_internal.h"

namespace mh {

namespace {

struct BgTask {
  const double *A, *B;
  double *C;
  int64_t sam, sak, sbk, sbn, scm, scn;
  int M, N, K, tile0;  // tile0: first grid tile of this task
  int tm, pad;         // tiles along M
};

typedef double bg_d4 __attribute__((ext_vector_type(4)));

// 64 x 64 output tile per workgroup on FP64 MFMA (16x16x4): wave w owns the
// 32 x 32 quarter (w % 2, w / 2) as 2 x 2 MFMA tiles; 16-deep K panels of A
// and B are staged in LDS from any strides.  MFMA operand lane l holds
// A[l%16][l/16] and B[l/16][l%16]; accumulator register q of lane l is
// C[l/16 + 4q][l%16].
__global__ __launch_bounds__(256) void k_gemm_batched(const BgTask *__restrict__ tab, int nt,
                                                      double alpha, double beta) {
  __shared__ double As[2][16][64 + 2];
  __shared__ double Bs[2][16][64 + 2];
  int lo = 0, hi = nt - 1;
  while (lo < hi) {
    const int md = (lo + hi + 1) >> 1;
    if (tab[md].tile0 <= (int)blockIdx.x) lo = md; else hi = md - 1;
  }
  const BgTask t = tab[lo];
  const int tile = (int)blockIdx.x - t.tile0;
  const int r0 = (tile % t.tm) * 64, c0 = (tile / t.tm) * 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wm = (wv & 1) * 32, wn = (wv >> 1) * 32;
  bg_d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = bg_d4{0.0, 0.0, 0.0, 0.0};
  // staging order per operand: consecutive threads walk the unit-stride
  // dimension (k when |stride along k| == 1, else the row), so a wavefront's
  // loads fall on few cache lines; the next panel is loaded into registers
  // while the current one is multiplied (two LDS buffers, one barrier per
  // panel)
  const bool a_k = (t.sak == 1 || t.sak == -1), b_k = (t.sbk == 1 || t.sbk == -1);
  double ra[4], rb[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = threadIdx.x + 256 * q;
      {
        const int kk = a_k ? (e & 15) : (e >> 6), rr = a_k ? (e >> 4) : (e & 63);
        const int gk = k0 + kk, gm = r0 + rr;
        ra[q] = (gk < t.K && gm < t.M) ? t.A[gm * t.sam + gk * t.sak] : 0.0;
      }
      {
        const int kk = b_k ? (e & 15) : (e >> 6), rr = b_k ? (e >> 4) : (e & 63);
        const int gk = k0 + kk, gn = c0 + rr;
        rb[q] = (gk < t.K && gn < t.N) ? t.B[gk * t.sbk + gn * t.sbn] : 0.0;
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = threadIdx.x + 256 * q;
      As[buf][a_k ? (e & 15) : (e >> 6)][a_k ? (e >> 4) : (e & 63)] = ra[q];
      Bs[buf][b_k ? (e & 15) : (e >> 6)][b_k ? (e >> 4) : (e & 63)] = rb[q];
    }
  };
  const int nk = (t.K + 15) / 16;
  if (nk > 0) {
    load(0);
    store(0);
    __syncthreads();
  }
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load(16 * (kt + 1));
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      double a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a[i] = As[buf][4 * s + lk][wm + 16 * i + li];
        b[i] = Bs[buf][4 * s + lk][wn + 16 * i + li];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gm = r0 + wm + 16 * i + lk + 4 * q, gn = c0 + wn + 16 * j + li;
        if (gm < t.M && gn < t.N) {
          double *c = t.C + gm * t.scm + gn * t.scn;
          *c = (beta == 0.0) ? alpha * acc[i][j][q] : alpha * acc[i][j][q] + beta * *c;
        }
      }
}

struct BgState {
  void *buf = nullptr;
  size_t bytes = 0;
};

}  // namespace

void bg_release(mh_handle h) {
  BgState *s = (BgState *)h->bg_state;
  if (s) {
    if (s->buf) (void)hipFree(s->buf);
    delete s;
  }
  h->bg_state = nullptr;
}

}  // namespace mh

using namespace mh;

extern "C" int mh_gemm_batched(mh_handle h, int count, const int64_t *desc, double alpha,
                               double beta) {
  if (!h || count < 0 || (count > 0 && !desc))
    return fail(h, MH_ERR_ARG, "mh_gemm_batched: null argument");
  std::vector<BgTask> tab;
  tab.reserve(count);
  int tiles = 0;
  for (int i = 0; i < count; ++i) {
    const int64_t *d = desc + 12 * (int64_t)i;
    BgTask t{};
    t.A = (const double *)d[0];
    t.B = (const double *)d[1];
    t.C = (double *)d[2];
    t.M = (int)d[3];
    t.N = (int)d[4];
    t.K = (int)d[5];
    t.sam = d[6];
    t.sak = d[7];
    t.sbk = d[8];
    t.sbn = d[9];
    t.scm = d[10];
    t.scn = d[11];
    if (d[3] < 0 || d[4] < 0 || d[5] < 0 || d[3] > INT32_MAX || d[4] > INT32_MAX ||
        d[5] > INT32_MAX)
      return fail(h, MH_ERR_SHAPE, "mh_gemm_batched: task %d has extents %lld x %lld x %lld", i,
                  (long long)d[3], (long long)d[4], (long long)d[5]);
    if (t.M == 0 || t.N == 0) continue;  // nothing to write
    if (!t.C || (t.K > 0 && (!t.A || !t.B)))
      return fail(h, MH_ERR_ARG, "mh_gemm_batched: task %d has a null operand", i);
    t.tm = (t.M + 63) / 64;
    t.tile0 = tiles;
    tiles += t.tm * ((t.N + 63) / 64);
    tab.push_back(t);
  }
  if (tab.empty()) return MH_OK;
  MH_HIP(h, hipSetDevice(h->device));
  if (!h->bg_state) h->bg_state = new BgState();
  BgState *st = (BgState *)h->bg_state;
  const size_t need = tab.size() * sizeof(BgTask);
  if (st->bytes < need) {
    MH_HIP(h, hipStreamSynchronize(h->stream));
    if (st->buf) MH_HIP(h, hipFree(st->buf));
    st->buf = nullptr;
    st->bytes = 0;
    const size_t b = std::max(need * 2, (size_t)64 * 1024);
    hipError_t e = hipMalloc(&st->buf, b);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      st->buf = nullptr;
      return fail(h, MH_ERR_OOM, "mh_gemm_batched: %zu bytes: %s", b, hipGetErrorString(e));
    }
    st->bytes = b;
  }
  // the table upload is ordered on the stream before the launch; a later call
  // reuses the buffer only behind this launch on the same stream
  MH_HIP(h, hipMemcpyAsync(st->buf, tab.data(), need, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_gemm_batched, dim3(tiles), dim3(256), 0, h->stream,
                     (const BgTask *)st->buf, (int)tab.size(), alpha, beta);
  return check_launch(h, "k_gemm_batched");
}
