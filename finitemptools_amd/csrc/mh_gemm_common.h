// Shared GEMM plumbing: launch parameters, candidate tile configurations and
// the per-shape tuner/dispatcher used by the FP64 and mixed-precision
// kernel families.
#pragma once

#include <type_traits>

#include "mh_internal.h"

namespace mh {

typedef double d4_t __attribute__((ext_vector_type(4)));

struct GemmP {
  int M, N, K;
  int kchunk;  // K extent per split (multiple of the tile depth)
  // XCD-aware tile order (group > 0): the 1-D grid holds 8 * xcd_chunk
  // blocks; blocks are dispatched round-robin over the 8 XCDs, so XCD x
  // receives ids x, x+8, ... and is given the contiguous range
  // [x*xcd_chunk, (x+1)*xcd_chunk) of a grouped tile order (group tile-rows
  // swept column by column), i.e. a compact region of C whose operands its
  // own L2 can hold.
  int tiles_m, tiles_n, xcd_chunk, group;
  // Operand element types are fixed by the kernel family: FP64 (complex:
  // interleaved pairs) for the FP64 family, packed 32-bit words for the mixed
  // family.  Strides count elements (complex elements for complex data).
  const double *A;
  int64_t sam, sak;
  const double *B;
  int64_t sbk, sbn;
  double *C;  // FP64, or FP32 when c_f32 = 1 (mixed family only)
  int64_t scm, scn;
  double alpha_re, alpha_im, beta_re, beta_im;
  int conjA;
  int c_f32;
};

struct Cand {
  int bm, bn, bk, split, group;
  int waves;  // waves per block; 0 means the family default (4)
  int kgroups;  // 2: the 8 waves form two 2x2 wave grids that take alternate
                // 16-deep K slices and add their tiles at the end (0/1: off)
  int prefetch; // 2: global loads run two K steps ahead (0/1: one step)
};

struct GemmFamily {
  const char *name;
  const Cand *cands;
  int n;
  int default_bk;
  bool (*ok)(const Cand &, const GemmP &, bool cplx);
  void (*launch)(const Cand &, const GemmP &, hipStream_t, bool cplx, bool amc, bool bkc);
};

inline GemmP make_params(const GemmDesc &g) {
  GemmP p;
  p.M = g.M;
  p.N = g.N;
  p.K = g.K;
  p.kchunk = g.K;
  p.tiles_m = p.tiles_n = p.xcd_chunk = p.group = 0;
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
  p.c_f32 = 0;
  return p;
}

// Fill the tile-order fields of p for candidate c and return the grid.
inline dim3 setup_grid(const Cand &c, GemmP &p) {
  p.tiles_m = (p.M + c.bm - 1) / c.bm;
  p.tiles_n = (p.N + c.bn - 1) / c.bn;
  p.group = c.group;
  p.xcd_chunk = (p.tiles_m * p.tiles_n + 7) / 8;
  int kc = (p.K + c.split - 1) / c.split;
  p.kchunk = (kc + c.bk - 1) / c.bk * c.bk;
  if (c.group > 0) return dim3(8 * p.xcd_chunk, 1, c.split);
  return dim3(p.tiles_m, p.tiles_n, c.split);
}

// Block -> (tile row, tile column), honouring the XCD-grouped order.
// Returns false for padding blocks of the grouped grid (the whole block exits).
__device__ __forceinline__ bool tile_of_block(const GemmP &p, int &tm, int &tn) {
  tm = blockIdx.x;
  tn = blockIdx.y;
  if (p.group > 0) {
    const int id = blockIdx.x;
    const int local = id >> 3;
    const int t = (id & 7) * p.xcd_chunk + local;
    if (local >= p.xcd_chunk || t >= p.tiles_m * p.tiles_n) return false;
    const int span = p.group * p.tiles_n;
    const int first = (t / span) * p.group;
    const int gsz = min(p.tiles_m - first, p.group);
    const int w = t % span;
    tm = first + w % gsz;
    tn = w / gsz;
  }
  return true;
}

// C[m, n] = 0 before a split-K accumulation (a kernel, not a memset, so the
// sequence stays valid inside captured graphs).
static __global__ void k_zero_c(double *C, int M, int N, int64_t scm, int64_t scn, int cplx,
                                int f32) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)M * N) return;
  const int m = (int)(idx % M), n = (int)(idx / M);
  const int64_t off = (int64_t)m * scm + (int64_t)n * scn;
  if (f32) {
    float *c = reinterpret_cast<float *>(C);
    if (cplx) reinterpret_cast<float2 *>(c)[off] = make_float2(0.f, 0.f);
    else c[off] = 0.f;
  } else if (cplx) {
    reinterpret_cast<double2 *>(C)[off] = make_double2(0.0, 0.0);
  } else {
    C[off] = 0.0;
  }
}

// Tune (first call per shape key) or look up, then launch; profiled as one
// MH_K_GEMM launch with 2*M*N*K (x4 complex) algorithmic flops.
int gemm_dispatch(mh_handle h, const GemmFamily &fam, const GemmP &p, bool cplx, const char *tag);

}  // namespace mh
