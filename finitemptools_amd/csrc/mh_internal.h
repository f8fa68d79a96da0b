// Internal declarations shared by the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mpsheff.h"

namespace mh {

// Per-launch timing record (profiling mode only).
struct ProfRec {
  hipEvent_t start, stop;
  int kclass;
  double flops;
};

}  // namespace mh

struct mh_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  // Grow-only workspace for contraction intermediates.
  void *ws = nullptr;
  size_t ws_bytes = 0;
  // Packed (FP32 / BF16-pair) operand copies for the mixed-precision GEMMs.
  void *pack = nullptr;
  size_t pack_bytes = 0;
  // Small device buffers: packed MPO mode matrices and reduction scratch.
  void *small = nullptr;
  size_t small_bytes = 0;
  double *red = nullptr;       // device reduction scratch (doubles)
  size_t red_len = 0;
  double *red_host = nullptr;  // pinned host mirror for scalar reads
  // Profiling.
  int profiling = 0;
  std::vector<mh::ProfRec> prof_pending;
  double prof_ms[MH_K_NCLASS] = {0, 0, 0};
  int64_t prof_n[MH_K_NCLASS] = {0, 0, 0};
  double prof_flops[MH_K_NCLASS] = {0, 0, 0};
  std::vector<hipEvent_t> event_pool;
  // GEMM autotuner: shape key -> index of the fastest measured configuration.
  std::unordered_map<std::string, int> tune;
  // Grouped-GEMM tuner: stage shape class -> fastest tile variant.
  std::unordered_map<uint64_t, int> bs_tune;
  // Krylov recurrences as hipGraphs: captured on a private stream (a
  // non-blocking stream at the default priority), launched on `stream`.
  int graphs = 1;
  // Fused first H_eff step (GEMM with the mode product in its epilogue).
  int fusion = 1;
  // Krylov steps on short vectors as one single-workgroup launch.
  int krylov_small = 1;
  // Block plans as one GEMM launch per sector product (the plain per-sector
  // loop) instead of one grouped launch per stage.
  int bs_plain = 0;
  hipStream_t cap_stream = nullptr;
  // Batched eigensolver (mh_eigen.hip): exchange rows, barrier counters and
  // error flag of the tridiagonalisation; side streams and rocSOLVER handles.
  void *eig = nullptr;
  size_t eig_bytes = 0;
  int *eig_err = nullptr;
  void *eig_state = nullptr;
  // Divide-and-conquer tridiagonal eigensolver (mh_tridc.hip): work buffer
  // and leaf convergence flag.
  void *dc_state = nullptr;
  // Batched strided GEMMs (mh_batched.hip): task table buffer.
  void *bg_state = nullptr;
  int autotune = 1;
  // Set for the duration of an allocation-free entry point: no tuning runs,
  // no tuner-table insertions and no profiling events (host or device).
  int noalloc = 0;
  int gemm_force = -1;  // FP64 GEMM configuration forced for every call (-1: tuned)
  int tune_log = 0;
};

namespace mh {

int fail(mh_handle h, int code, const char *fmt, ...);

#define MH_HIP(h, expr)                                                        \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess)                                                      \
      return ::mh::fail((h), MH_ERR_HIP, "%s failed: %s (%s:%d)", #expr,        \
                        hipGetErrorString(_e), __FILE__, __LINE__);            \
  } while (0)

#define MH_TRY(expr)                                                           \
  do {                                                                         \
    int _s = (expr);                                                           \
    if (_s != MH_OK) return _s;                                                \
  } while (0)

// Ensure the workspace holds at least `bytes`; returns MH_OK or an error.
int ensure_ws(mh_handle h, size_t bytes);
int ensure_small(mh_handle h, size_t bytes);
int ensure_red(mh_handle h, size_t n);

// Profiling brackets around one kernel launch.
void prof_begin(mh_handle h, int kclass, double flops, int *slot);
void prof_end(mh_handle h, int slot);
int check_launch(mh_handle h, const char *what);

// Strided GEMM: C = alpha * op(A) B + beta * C (see mh_gemm in the header).
struct GemmDesc {
  int M, N, K;
  const void *A;
  int64_t sam, sak;
  const void *B;
  int64_t sbk, sbn;
  void *C;
  int64_t scm, scn;
  double alpha_re, alpha_im, beta_re, beta_im;
  int conjA;
};
int gemm(mh_handle h, int dtype, int prec, const GemmDesc &g);
// Mixed precisions on packed operands (4-byte words per real component: FP32,
// or a BF16 hi/lo pair); c_packed = 1 writes packed words, 0 writes FP64.
int gemm_packed(mh_handle h, int dtype, int prec, const GemmDesc &g, int c_packed);
int pack_operand(mh_handle h, int prec, const double *src, int64_t ndouble, uint32_t *dst);
int mode_product_packed(mh_handle h, int dtype, int prec, int P, int ni, int nj, int Q,
                        const void *in, int64_t sIq, const void *Mt, void *out, int64_t sOq);

// Mode-matrix kinds used to fold MPO tensors into a (ni x nj) matrix Mt so
// that out[p, j, q] = sum_i in[p, i, q] Mt[i, j].
enum ModeKind {
  MODE_SITE = 0,   // i = w + wl*s,              j = s' + d*w'       (W[w,s',w',s])
  MODE_TWO = 1,    // i = w + wl*(s1 + d1*s2),   j = s1' + d1*(s2' + d2*w')
  MODE_RIGHT = 2   // i = s' + d*w',             j = w + wl*s
};
int build_mode_matrix(mh_handle h, int dtype, int kind, int wl, int wm, int wr,
                      int d1, int d2, const void *W1, const void *W2, void *Mt);
// T2 = mode(L x) in one launch (FP64 real, supported (wl, ds) only);
// MH_ERR_UNSUPPORTED when the shape has no fused kernel.
int heff_step1_fused(mh_handle h, int Dl, int Dr, int wl, int ds, int wr, const void *L,
                     const void *x, const void *Mt, void *T2);
// The same for ComplexF64 operators (interleaved re, im): wl = wr in
// {3, 4, 5, 6} and a physical extent of 2 or 4.
int heff_step1_fused_c(mh_handle h, int Dl, int Dr, int wl, int ds, int wr, const void *L,
                       const void *x, const void *Mt, void *T2);
// Packed-operand form for the mixed precisions (mh_gemm_mixed.hip): L, x and
// T2 are packed words of prec's family; same support and return convention.
int heff_step1_fused_mx(mh_handle h, int prec, int Dl, int Dr, int wl, int ds, int wr,
                        const void *L, const void *x, const void *Mt, void *T2);
int mode_product(mh_handle h, int dtype, int P, int ni, int nj, int Q,
                 const void *in, int64_t sIq, const void *Mt, void *out,
                 int64_t sOq);

int heff_apply(mh_handle h, const mh_heff_op &op, const void *x, void *y,
               void *mode_cache);
size_t heff_ws_bytes(const mh_heff_op &op);
size_t heff_vec_len(const mh_heff_op &op);
int heff_check(mh_handle h, const mh_heff_op &op);

// Vector primitives on the handle's stream (results left on device).
int vec_dot_dev(mh_handle h, int dtype, size_t n, const void *x, const void *y,
                double *dres);
int vec_nrm2sq_dev(mh_handle h, int dtype, size_t n, const void *x,
                   double *dres);

// Block-sparse plans (mh_blocksparse.hip).
size_t bs_ws_bytes(const mh_bs_plan_s &p);
size_t bs_vec_len(const mh_bs_plan_s &p);
int bs_apply_ws(mh_handle h, const mh_bs_plan_s &p, const void *x, const void *p1,
                const void *p2, void *y, void *ws);
// Upload a host-built plan's tables to the handle's device (first use only).
int bs_prepare(mh_handle h, mh_bs_plan_s &p);
// An operator plan (x and y of one size) with the environment buffers it needs.
int bs_check_square(mh_handle h, const mh_bs_plan_s &p, const void *p1, const void *p2);

// Block <-> coupled-matrix element copies; pieces: {blk_off, mat_off, rows,
// cols, ld} per piece (mh_blocksparse.hip).
int bs_coupled_copy(mh_handle h, const std::vector<int64_t> &pieces, const double *src,
                    double *dst, int to_coupled);

// Release the eigensolver's streams, library handles and scratch.
void eig_release(mh_handle h);
void dc_release(mh_handle h);
void bg_release(mh_handle h);
int *dc_error_flag(mh_handle h);

// Device address of the allocation-free entry points' static scratch array
// on h's device (looked up once per device when the first handle of that
// device is created).
int noalloc_setup(mh_handle h);

inline size_t dtype_bytes(int dtype) { return dtype == MH_C64 ? 16 : 8; }

}  // namespace mh
