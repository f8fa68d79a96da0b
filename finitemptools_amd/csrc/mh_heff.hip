// Effective Hamiltonians (1-site, 2-site, bond) and environment updates.
//
// Each contraction is evaluated as  GEMM -> mode product -> GEMM  (bond form:
// GEMM -> GEMM).  The two GEMMs carry the O(D^3) work on the matrix cores; the
// mode product folds the MPO tensor(s) into the middle index of the first
// intermediate:
//
//   out[p, j, q] = sum_i in[p, i, q] * Mt[i, j]
//
// where p runs over the fastest bond index (coalesced), q over the slowest,
// and Mt is a small (ni x nj) matrix assembled from W (two-site: W1*W2 summed
// over the shared MPO bond).  The MPO matrices are sparse; rows of Mt that are
// zero for a block's column chunk are skipped (the test is uniform across a
// wave, so it costs no divergence).
#include <mutex>

#include "mh_internal.h"

namespace mh {

// ------------------------------------------- allocation-free entry point
// Scratch of mh_ac2_prime_noalloc: the intermediates and the folded MPO
// matrix of one two-site H_eff live in this file-scope device array instead
// of the handle's grow-only workspace.  Every device the library runs on
// holds its own copy, at its own address, until the process ends.  Host
// code never passes the array's name to a kernel (on the host it is only a
// stand-in symbol); it passes the address hipGetSymbolAddress returned for
// the device, kept in g_scratch_ptr.
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];
static void *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

int noalloc_setup(mh_handle h) {
  if (h->device < 0 || h->device >= kMaxDevices)
    return fail(h, MH_ERR_ARG, "device %d beyond the scratch table (%d devices)", h->device,
                kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[h->device]) return MH_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    (void)hipGetLastError();
    return fail(h, MH_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", h->device,
                hipGetErrorString(e));
  }
  g_scratch_ptr[h->device] = ptr;
  return MH_OK;
}

__device__ __forceinline__ double cmul_re(double2 a, double2 b) { return a.x * b.x - a.y * b.y; }
__device__ __forceinline__ double cmul_im(double2 a, double2 b) { return a.x * b.y + a.y * b.x; }

template <typename T> struct Ar;
template <> struct Ar<double> {
  __device__ static double zero() { return 0.0; }
  __device__ static double mul(double a, double b) { return a * b; }
  __device__ static double fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
  __device__ static bool nz(double a) { return a != 0.0; }
};
template <> struct Ar<double2> {
  __device__ static double2 zero() { return make_double2(0.0, 0.0); }
  __device__ static double2 mul(double2 a, double2 b) { return make_double2(cmul_re(a, b), cmul_im(a, b)); }
  __device__ static double2 fma(double2 a, double2 b, double2 c) {
    return make_double2(c.x + a.x * b.x - a.y * b.y, c.y + a.x * b.y + a.y * b.x);
  }
  __device__ static bool nz(double2 a) { return a.x != 0.0 || a.y != 0.0; }
};

// Mt[i + ni*j] from the MPO tensor(s); W[wl, s', wr, s] column-major.
template <typename T>
__global__ void k_build_mode(int kind, int wl, int wm, int wr, int d1, int d2,
                             const T *__restrict__ W1, const T *__restrict__ W2,
                             T *__restrict__ Mt, int ni, int nj) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ni * nj) return;
  int i = idx % ni, j = idx / ni;
  T v = Ar<T>::zero();
  if (kind == MODE_SITE) {  // i = w + wl*s, j = s' + d*w'
    int w = i % wl, s = i / wl, sp = j % d1, wp = j / d1;
    v = W1[w + wl * (sp + d1 * (wp + wr * s))];
  } else if (kind == MODE_RIGHT) {  // i = s' + d*w', j = w + wl*s
    int sp = i % d1, wp = i / d1, w = j % wl, s = j / wl;
    v = W1[w + wl * (sp + d1 * (wp + wr * s))];
  } else {  // MODE_TWO: i = w + wl*(s1 + d1*s2), j = s1' + d1*(s2' + d2*w')
    int w = i % wl, s1 = (i / wl) % d1, s2 = i / (wl * d1);
    int s1p = j % d1, s2p = (j / d1) % d2, wp = j / (d1 * d2);
    for (int m = 0; m < wm; ++m) {
      T a = W1[w + wl * (s1p + d1 * (m + wm * s1))];
      T b = W2[m + wm * (s2p + d2 * (wp + wr * s2))];
      v = Ar<T>::fma(a, b, v);
    }
  }
  Mt[i + (int64_t)ni * j] = v;
}

template <typename T, int JC>
__global__ void __launch_bounds__(256)
    k_mode(const T *__restrict__ in, T *__restrict__ out,
           const T *__restrict__ Mt, int P, int ni, int nj, int64_t PQ,
           int64_t sIq, int64_t sOq) {
  extern __shared__ double smem_d[];
  T *sM = reinterpret_cast<T *>(smem_d);          // [ni][JC]
  int *nz = reinterpret_cast<int *>(sM + ni * JC);  // [ni]
  const int tid = threadIdx.x;
  const int j0 = blockIdx.y * JC;
  const int jn = min(JC, nj - j0);
  for (int t = tid; t < ni * JC; t += 256) {
    int i = t / JC, jj = t % JC;
    sM[t] = (jj < jn) ? Mt[i + (int64_t)ni * (j0 + jj)] : Ar<T>::zero();
  }
  __syncthreads();
  for (int i = tid; i < ni; i += 256) {
    int f = 0;
    for (int jj = 0; jj < JC; ++jj) f |= Ar<T>::nz(sM[i * JC + jj]) ? 1 : 0;
    nz[i] = f;
  }
  __syncthreads();
  const int64_t idx = (int64_t)blockIdx.x * 256 + tid;
  if (idx >= PQ) return;
  const int64_t p = idx % P, q = idx / P;
  T acc[JC];
#pragma unroll
  for (int jj = 0; jj < JC; ++jj) acc[jj] = Ar<T>::zero();
  const T *src = in + p + q * sIq;
  // Rows in groups of 8: the (independent) loads of a group are issued
  // before its FMAs, so each thread has up to 8 loads in flight.  Rows whose
  // Mt row is zero are neither loaded nor used (a uniform branch).
  for (int i0 = 0; i0 < ni; i0 += 8) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u;
      v[u] = (i < ni && nz[i]) ? src[(int64_t)i * P] : Ar<T>::zero();
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u;
      if (i >= ni || !nz[i]) continue;
#pragma unroll
      for (int jj = 0; jj < JC; ++jj) acc[jj] = Ar<T>::fma(v[u], sM[i * JC + jj], acc[jj]);
    }
  }
  T *dst = out + p + q * sOq;
#pragma unroll
  for (int jj = 0; jj < JC; ++jj)
    if (jj < jn) dst[(int64_t)(j0 + jj) * P] = acc[jj];
}

int build_mode_matrix(mh_handle h, int dtype, int kind, int wl, int wm, int wr,
                      int d1, int d2, const void *W1, const void *W2, void *Mt) {
  int ni, nj;
  if (kind == MODE_SITE) {
    ni = wl * d1;
    nj = d1 * wr;
  } else if (kind == MODE_RIGHT) {
    ni = d1 * wr;
    nj = wl * d1;
  } else {
    ni = wl * d1 * d2;
    nj = d1 * d2 * wr;
  }
  int n = ni * nj;
  dim3 grid((n + 255) / 256);
  if (dtype == MH_C64)
    hipLaunchKernelGGL(k_build_mode<double2>, grid, dim3(256), 0, h->stream, kind, wl, wm, wr, d1, d2,
                       (const double2 *)W1, (const double2 *)W2, (double2 *)Mt, ni, nj);
  else
    hipLaunchKernelGGL(k_build_mode<double>, grid, dim3(256), 0, h->stream, kind, wl, wm, wr, d1, d2,
                       (const double *)W1, (const double *)W2, (double *)Mt, ni, nj);
  return check_launch(h, "build_mode");
}

template <typename T, int JC>
static void launch_mode(mh_handle h, int P, int ni, int nj, int64_t Q, const void *in,
                        int64_t sIq, const void *Mt, void *out, int64_t sOq) {
  int64_t PQ = (int64_t)P * Q;
  dim3 grid((unsigned)((PQ + 255) / 256), (nj + JC - 1) / JC);
  size_t smem = sizeof(T) * ni * JC + sizeof(int) * ni;
  hipLaunchKernelGGL((k_mode<T, JC>), grid, dim3(256), smem, h->stream, (const T *)in,
                     (T *)out, (const T *)Mt, P, ni, nj, PQ, sIq, sOq);
}

int mode_product(mh_handle h, int dtype, int P, int ni, int nj, int Q,
                 const void *in, int64_t sIq, const void *Mt, void *out,
                 int64_t sOq) {
  if ((int64_t)P * Q <= 0) return MH_OK;
  // Count only the stored Mt entries; the sparse skip makes it an upper bound.
  double flops = 2.0 * P * (double)Q * ni * nj * (dtype == MH_C64 ? 4.0 : 1.0);
  int slot;
  prof_begin(h, MH_K_WSTAGE, flops, &slot);
  const bool c = dtype == MH_C64;
  if (nj <= 8) {
    if (c) launch_mode<double2, 8>(h, P, ni, nj, Q, in, sIq, Mt, out, sOq);
    else launch_mode<double, 8>(h, P, ni, nj, Q, in, sIq, Mt, out, sOq);
  } else if (nj <= 16) {
    if (c) launch_mode<double2, 16>(h, P, ni, nj, Q, in, sIq, Mt, out, sOq);
    else launch_mode<double, 16>(h, P, ni, nj, Q, in, sIq, Mt, out, sOq);
  } else {
    if (c) launch_mode<double2, 24>(h, P, ni, nj, Q, in, sIq, Mt, out, sOq);
    else launch_mode<double, 24>(h, P, ni, nj, Q, in, sIq, Mt, out, sOq);
  }
  prof_end(h, slot);
  return check_launch(h, "mode_product");
}

// GemmDesc with alpha = 1, beta = 0.
static inline GemmDesc gd(int M, int N, int K, const void *A, int64_t sam,
                          int64_t sak, int conjA, const void *B, int64_t sbk,
                          int64_t sbn, void *C, int64_t scm, int64_t scn) {
  return GemmDesc{M, N, K, A, sam, sak, B, sbk, sbn, C, scm, scn, 1.0, 0.0, 0.0, 0.0, conjA};
}

// ------------------------------------------------------------- H_eff . x

size_t heff_vec_len(const mh_heff_op &op) {
  switch (op.kind) {
    case MH_HEFF0: return (size_t)op.Dl * op.Dr;
    case MH_HEFF1: return (size_t)op.Dl * op.d1 * op.Dr;
    default: return (size_t)op.Dl * op.d1 * op.d2 * op.Dr;
  }
}

static size_t mode_elems(const mh_heff_op &op) {
  switch (op.kind) {
    case MH_HEFF0: return 0;
    case MH_HEFF1: return (size_t)op.wl * op.d1 * op.d1 * op.wr;
    default: return (size_t)op.wl * op.d1 * op.d2 * op.d1 * op.d2 * op.wr;
  }
}

// Elements of the two intermediates T1, T2.
static void heff_tmp_elems(const mh_heff_op &op, size_t *t1, size_t *t2) {
  switch (op.kind) {
    case MH_HEFF0:
      *t1 = (size_t)op.Dl * op.wl * op.Dr;
      *t2 = 0;
      break;
    case MH_HEFF1:
      *t1 = (size_t)op.Dl * op.wl * op.d1 * op.Dr;
      *t2 = (size_t)op.Dl * op.d1 * op.wr * op.Dr;
      break;
    default:
      *t1 = (size_t)op.Dl * op.wl * op.d1 * op.d2 * op.Dr;
      *t2 = (size_t)op.Dl * op.d1 * op.d2 * op.wr * op.Dr;
  }
}

// Mixed-precision workspace: packed (4-byte per real component) T1, T2, x,
// the FP64 mode matrix, and packed copies of L and R.
struct MixedLayout {
  size_t t1, t2, xp, mt, lp, rp, total;
  explicit MixedLayout(const mh_heff_op &op) {
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    size_t e1, e2;
    heff_tmp_elems(op, &e1, &e2);
    const size_t w = 4 * (op.dtype == MH_C64 ? 2 : 1);
    t1 = 0;
    t2 = t1 + al(e1 * w);
    xp = t2 + al(e2 * w);
    mt = xp + al(heff_vec_len(op) * w);
    lp = mt + al(mode_elems(op) * dtype_bytes(op.dtype));
    rp = lp + al((size_t)op.Dl * op.wl * op.Dl * w);
    total = rp + al((size_t)op.Dr * op.wr * op.Dr * w);
  }
};

size_t heff_ws_bytes(const mh_heff_op &op) {
  if (op.prec != MH_PREC_F64) return MixedLayout(op).total;
  size_t t1, t2;
  heff_tmp_elems(op, &t1, &t2);
  size_t b = dtype_bytes(op.dtype);
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  return al(t1 * b) + al(t2 * b) + al(mode_elems(op) * b);
}

// Mixed forms: pack L and R into their workspace slots (once per operator).
int heff_prepare_mixed(mh_handle h, const mh_heff_op &op, void *ws) {
  const MixedLayout lay(op);
  const int64_t ew = op.dtype == MH_C64 ? 2 : 1;
  char *base = (char *)ws;
  MH_TRY(pack_operand(h, op.prec, (const double *)op.L, (int64_t)op.Dl * op.wl * op.Dl * ew,
                      (uint32_t *)(base + lay.lp)));
  return pack_operand(h, op.prec, (const double *)op.R, (int64_t)op.Dr * op.wr * op.Dr * ew,
                      (uint32_t *)(base + lay.rp));
}

// Mixed forms of H_eff.x: x is packed, step 1 writes packed T1, the mode
// product reads and writes packed words (FP32 arithmetic), step 3 writes the
// FP64 result.  For FP64 real operators of the supported shapes the first
// two run as one launch (heff_step1_fused_mx), with T1 kept in FP32 inside
// the block.  L, R must already be packed (heff_prepare_mixed).
static int heff_apply_mixed(mh_handle h, const mh_heff_op &op, const void *x, void *y, void *ws,
                            const void *Mt_prebuilt) {
  const MixedLayout lay(op);
  char *base = (char *)ws;
  void *T1 = base + lay.t1, *T2 = base + lay.t2, *xp = base + lay.xp;
  const void *Lp = base + lay.lp, *Rp = base + lay.rp;
  void *Mt = Mt_prebuilt ? (void *)Mt_prebuilt : (void *)(base + lay.mt);
  const int64_t ew = op.dtype == MH_C64 ? 2 : 1;
  MH_TRY(pack_operand(h, op.prec, (const double *)x, (int64_t)heff_vec_len(op) * ew, (uint32_t *)xp));
  const int Dl = op.Dl, Dr = op.Dr, wl = op.wl, wr = op.wr;
  if (op.kind == MH_HEFF0) {
    GemmDesc g1 = gd(Dl * wl, Dr, Dl, Lp, 1, (int64_t)Dl * wl, 0, xp, 1, Dl, T1, 1, (int64_t)Dl * wl);
    MH_TRY(gemm_packed(h, op.dtype, op.prec, g1, 1));
    GemmDesc g2 = gd(Dl, Dr, wl * Dr, T1, 1, Dl, 0, Rp, Dr, 1, y, 1, Dl);
    return gemm_packed(h, op.dtype, op.prec, g2, 0);
  }
  const int d1 = op.d1;
  const int ds = (op.kind == MH_HEFF2) ? d1 * op.d2 : d1;
  const int ni = wl * ds, nj = ds * wr;
  if (!Mt_prebuilt)
    MH_TRY(build_mode_matrix(h, op.dtype, op.kind == MH_HEFF2 ? MODE_TWO : MODE_SITE, wl, op.wm,
                             wr, d1, op.d2, op.W1, op.W2, Mt));
  int fused = MH_ERR_UNSUPPORTED;
  if (h->fusion && op.dtype == MH_F64) {
    fused = heff_step1_fused_mx(h, op.prec, Dl, Dr, wl, ds, wr, Lp, xp, Mt, T2);
    if (fused != MH_OK && fused != MH_ERR_UNSUPPORTED) return fused;
  }
  if (fused != MH_OK) {
    GemmDesc g1 = gd(Dl * wl, ds * Dr, Dl, Lp, 1, (int64_t)Dl * wl, 0, xp, 1, Dl, T1, 1,
                     (int64_t)Dl * wl);
    MH_TRY(gemm_packed(h, op.dtype, op.prec, g1, 1));
    MH_TRY(mode_product_packed(h, op.dtype, op.prec, Dl, ni, nj, Dr, T1, (int64_t)Dl * ni, Mt,
                               T2, (int64_t)Dl * nj));
  }
  GemmDesc g3 = gd(Dl * ds, Dr, wr * Dr, T2, 1, (int64_t)Dl * ds, 0, Rp, Dr, 1, y, 1, (int64_t)Dl * ds);
  return gemm_packed(h, op.dtype, op.prec, g3, 0);
}

int heff_check(mh_handle h, const mh_heff_op &op) {
  if (op.dtype != MH_F64 && op.dtype != MH_C64)
    return fail(h, MH_ERR_DTYPE, "unknown dtype %d", op.dtype);
  if (op.prec < MH_PREC_F64 || op.prec > MH_PREC_BF16X3)
    return fail(h, MH_ERR_DTYPE, "unknown precision %d", op.prec);
  if (op.kind < MH_HEFF0 || op.kind > MH_HEFF2)
    return fail(h, MH_ERR_ARG, "unknown operator kind %d", op.kind);
  if (op.Dl <= 0 || op.Dr <= 0 || op.wl <= 0)
    return fail(h, MH_ERR_SHAPE, "non-positive extent (Dl=%d Dr=%d wl=%d)", op.Dl, op.Dr, op.wl);
  if (op.kind >= MH_HEFF1 && (op.d1 <= 0 || op.wr <= 0))
    return fail(h, MH_ERR_SHAPE, "non-positive extent (d1=%d wr=%d)", op.d1, op.wr);
  if (op.kind == MH_HEFF2 && (op.d2 <= 0 || op.wm <= 0))
    return fail(h, MH_ERR_SHAPE, "non-positive extent (d2=%d wm=%d)", op.d2, op.wm);
  if (!op.L || !op.R) return fail(h, MH_ERR_ARG, "null environment");
  if (op.kind >= MH_HEFF1 && !op.W1) return fail(h, MH_ERR_ARG, "null MPO tensor");
  if (op.kind == MH_HEFF2 && !op.W2) return fail(h, MH_ERR_ARG, "null second MPO tensor");
  const int64_t lim = (int64_t)1 << 31;
  size_t t1, t2;
  heff_tmp_elems(op, &t1, &t2);
  if ((int64_t)t1 >= lim || (int64_t)t2 >= lim)
    return fail(h, MH_ERR_SHAPE, "intermediate exceeds 2^31 elements");
  if (op.kind == MH_HEFF2 && (int64_t)op.wl * op.d1 * op.d2 > 4096)
    return fail(h, MH_ERR_SHAPE, "MPO mode too large");
  return MH_OK;
}

// Workspace layout: [T1 | T2 | Mt]; Mt may be supplied prebuilt.  Mixed
// precisions use MixedLayout; lr_ready says L and R are already packed.
int heff_apply_ws(mh_handle h, const mh_heff_op &op, const void *x, void *y,
                  void *ws, const void *Mt_prebuilt, bool lr_ready) {
  if (op.prec != MH_PREC_F64) {
    if (!lr_ready) MH_TRY(heff_prepare_mixed(h, op, ws));
    return heff_apply_mixed(h, op, x, y, ws, Mt_prebuilt);
  }
  const size_t b = dtype_bytes(op.dtype);
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  size_t t1, t2;
  heff_tmp_elems(op, &t1, &t2);
  char *T1 = (char *)ws;
  char *T2 = T1 + al(t1 * b);
  char *Mt = T2 + al(t2 * b);
  const int Dl = op.Dl, Dr = op.Dr, wl = op.wl, wr = op.wr;
  if (op.kind == MH_HEFF0) {
    const int w = wl;
    // T1[(a,w), b'] = L[(a,w), a'] x[a', b']
    GemmDesc g1 = gd(Dl * w, Dr, Dl, op.L, 1, (int64_t)Dl * w, 0, x, 1, Dl, T1, 1,
                (int64_t)Dl * w);
    MH_TRY(gemm(h, op.dtype, op.prec, g1));
    // y[a, b] = sum_{(w,b')} T1[a, (w,b')] R[b, (w,b')]
    GemmDesc g2 = gd(Dl, Dr, w * Dr, T1, 1, Dl, 0, op.R, Dr, 1, y, 1, Dl);
    return gemm(h, op.dtype, op.prec, g2);
  }
  const int d1 = op.d1;
  const int ds = (op.kind == MH_HEFF2) ? d1 * op.d2 : d1;  // physical extent of x
  const int ni = wl * ds, nj = ds * wr;
  if (Mt_prebuilt) {
    Mt = (char *)Mt_prebuilt;
  } else {
    MH_TRY(build_mode_matrix(h, op.dtype, op.kind == MH_HEFF2 ? MODE_TWO : MODE_SITE,
                             wl, op.wm, wr, d1, op.d2, op.W1, op.W2, Mt));
  }
  int fused = MH_ERR_UNSUPPORTED;
  if (h->fusion) {
    // T2 = mode(L x) without T1 in HBM
    fused = op.dtype == MH_F64 ? heff_step1_fused(h, Dl, Dr, wl, ds, wr, op.L, x, Mt, T2)
                               : heff_step1_fused_c(h, Dl, Dr, wl, ds, wr, op.L, x, Mt, T2);
    if (fused != MH_OK && fused != MH_ERR_UNSUPPORTED) return fused;
  }
  if (fused != MH_OK) {
    // T1[(a,w), (s,b')] = L[(a,w), a'] x[a', (s,b')]
    GemmDesc g1 = gd(Dl * wl, ds * Dr, Dl, op.L, 1, (int64_t)Dl * wl, 0, x, 1, Dl, T1, 1,
                (int64_t)Dl * wl);
    MH_TRY(gemm(h, op.dtype, op.prec, g1));
    // T2[a, (s',w'), b'] = sum_{(w,s)} T1[a, (w,s), b'] Mt[(w,s), (s',w')]
    MH_TRY(mode_product(h, op.dtype, Dl, ni, nj, Dr, T1, (int64_t)Dl * ni, Mt, T2,
                        (int64_t)Dl * nj));
  }
  // y[(a,s'), b] = sum_{(w',b')} T2[(a,s'), (w',b')] R[b, (w',b')]
  GemmDesc g3 = gd(Dl * ds, Dr, wr * Dr, T2, 1, (int64_t)Dl * ds, 0, op.R, Dr, 1, y, 1,
              (int64_t)Dl * ds);
  return gemm(h, op.dtype, op.prec, g3);
}

int heff_apply(mh_handle h, const mh_heff_op &op, const void *x, void *y,
               void *mode_cache) {
  MH_TRY(ensure_ws(h, heff_ws_bytes(op)));
  return heff_apply_ws(h, op, x, y, h->ws, mode_cache, false);
}

// ------------------------------------------------------- environment updates

static int env_check(mh_handle h, int dtype, int prec, int Dl, int d, int Dr,
                     int wl, int wr, const void *hold, const void *A,
                     const void *W, const void *B, const void *out) {
  if (!h) return MH_ERR_ARG;
  if (dtype != MH_F64 && dtype != MH_C64)
    return fail(h, MH_ERR_DTYPE, "unknown dtype %d", dtype);
  if (prec < MH_PREC_F64 || prec > MH_PREC_BF16X3)
    return fail(h, MH_ERR_DTYPE, "unknown precision %d", prec);
  if (Dl <= 0 || d <= 0 || Dr <= 0 || wl <= 0 || wr <= 0)
    return fail(h, MH_ERR_SHAPE, "non-positive extent");
  if (!hold || !A || !W || !B || !out) return fail(h, MH_ERR_ARG, "null tensor");
  const int64_t lim = (int64_t)1 << 31;
  if ((int64_t)Dl * wl * d * Dr >= lim || (int64_t)Dl * wr * d * Dr >= lim)
    return fail(h, MH_ERR_SHAPE, "intermediate exceeds 2^31 elements");
  return MH_OK;
}

static size_t env_ws_bytes(int dtype, int Dl, int d, int Dr, int wl, int wr) {
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  size_t b = dtype_bytes(dtype);
  size_t t = (size_t)Dl * d * Dr * (wl > wr ? wl : wr);
  return 2 * al(t * b) + al((size_t)wl * d * d * wr * b);
}

int env_left_ws(mh_handle h, int dtype, int prec, int Dl, int d, int Dr, int wl,
                int wr, const void *hold, const void *A, const void *W,
                const void *B, void *hnew, void *ws) {
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  size_t b = dtype_bytes(dtype);
  size_t t = (size_t)Dl * d * Dr * (wl > wr ? wl : wr);
  char *T1 = (char *)ws, *T2 = T1 + al(t * b), *Mt = T2 + al(t * b);
  MH_TRY(build_mode_matrix(h, dtype, MODE_SITE, wl, 0, wr, d, 1, W, nullptr, Mt));
  // T1[(a,w), (s,b')] = hold[(a,w), a'] B[a', (s,b')]
  GemmDesc g1 = gd(Dl * wl, d * Dr, Dl, hold, 1, (int64_t)Dl * wl, 0, B, 1, Dl, T1, 1,
              (int64_t)Dl * wl);
  MH_TRY(gemm(h, dtype, prec, g1));
  MH_TRY(mode_product(h, dtype, Dl, wl * d, d * wr, Dr, T1, (int64_t)Dl * wl * d, Mt,
                      T2, (int64_t)Dl * d * wr));
  // hnew[b, (w',b')] = sum_{(a,s')} conj(A[(a,s'), b]) T2[(a,s'), (w',b')]
  GemmDesc g3 = gd(Dr, wr * Dr, Dl * d, A, (int64_t)Dl * d, 1, 1, T2, 1, (int64_t)Dl * d,
              hnew, 1, Dr);
  return gemm(h, dtype, prec, g3);
}

int env_right_ws(mh_handle h, int dtype, int prec, int Dl, int d, int Dr, int wl,
                 int wr, const void *hold, const void *A, const void *W,
                 const void *B, void *hnew, void *ws) {
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  size_t b = dtype_bytes(dtype);
  size_t t = (size_t)Dl * d * Dr * (wl > wr ? wl : wr);
  char *T1 = (char *)ws, *T2 = T1 + al(t * b), *Mt = T2 + al(t * b);
  MH_TRY(build_mode_matrix(h, dtype, MODE_RIGHT, wl, 0, wr, d, 1, W, nullptr, Mt));
  // U[(a,s'), (w',b')] = sum_b conj(A[(a,s'), b]) hold[b, (w',b')]
  GemmDesc g1 = gd(Dl * d, wr * Dr, Dr, A, 1, (int64_t)Dl * d, 1, hold, 1, Dr, T1, 1,
              (int64_t)Dl * d);
  MH_TRY(gemm(h, dtype, prec, g1));
  // V[a, (w,s), b'] = sum_{(s',w')} U[a, (s',w'), b'] Mt[(s',w'), (w,s)]
  MH_TRY(mode_product(h, dtype, Dl, d * wr, wl * d, Dr, T1, (int64_t)Dl * d * wr, Mt,
                      T2, (int64_t)Dl * wl * d));
  // hnew[(a,w), a'] = sum_{(s,b')} V[(a,w), (s,b')] B[a', (s,b')]
  GemmDesc g3 = gd(Dl * wl, Dl, d * Dr, T2, 1, (int64_t)Dl * wl, 0, B, Dl, 1, hnew, 1,
              (int64_t)Dl * wl);
  return gemm(h, dtype, prec, g3);
}

}  // namespace mh

using namespace mh;

extern "C" {

int mh_heff_size(const mh_heff_op *op, size_t *n) {
  if (!op || !n) return MH_ERR_ARG;
  *n = heff_vec_len(*op);
  return MH_OK;
}

int mh_heff_apply(mh_handle h, const mh_heff_op *op, const void *x, void *y) {
  if (!h || !op) return MH_ERR_ARG;
  MH_TRY(heff_check(h, *op));
  if (!x || !y) return fail(h, MH_ERR_ARG, "null vector");
  if (x == y) return fail(h, MH_ERR_ARG, "x and y must not alias");
  return heff_apply(h, *op, x, y, nullptr);
}

int mh_ac_prime(mh_handle h, int dtype, int prec, int Dl, int d, int Dr, int wl,
                int wr, const void *x, const void *W, const void *L,
                const void *R, void *y) {
  mh_heff_op op{MH_HEFF1, dtype, prec, Dl, d, 1, Dr, wl, 0, wr, W, nullptr, L, R};
  return mh_heff_apply(h, &op, x, y);
}

int mh_ac2_prime(mh_handle h, int dtype, int prec, int Dl, int d1, int d2,
                 int Dr, int wl, int wm, int wr, const void *x, const void *W1,
                 const void *W2, const void *L, const void *R, void *y) {
  mh_heff_op op{MH_HEFF2, dtype, prec, Dl, d1, d2, Dr, wl, wm, wr, W1, W2, L, R};
  return mh_heff_apply(h, &op, x, y);
}

int mh_ac2_prime_noalloc(mh_handle h, int dtype, int prec, int Dl, int d1, int d2, int Dr,
                         int wl, int wm, int wr, const void *x, const void *W1, const void *W2,
                         const void *L, const void *R, void *y) {
  if (!h) return MH_ERR_ARG;
  mh_heff_op op{MH_HEFF2, dtype, prec, Dl, d1, d2, Dr, wl, wm, wr, W1, W2, L, R};
  MH_TRY(heff_check(h, op));
  if (!x || !y) return fail(h, MH_ERR_ARG, "null vector");
  if (x == y) return fail(h, MH_ERR_ARG, "x and y must not alias");
  if (prec != MH_PREC_F64)
    return fail(h, MH_ERR_UNSUPPORTED, "mh_ac2_prime_noalloc computes in FP64 only (prec %d)",
                prec);
  const size_t need = heff_ws_bytes(op);
  if (need > kNoallocScratchBytes)
    return fail(h, MH_ERR_SHAPE,
                "mh_ac2_prime_noalloc needs %zu bytes of scratch; its static scratch holds %zu",
                need, kNoallocScratchBytes);
  void *ws = (h->device >= 0 && h->device < kMaxDevices) ? g_scratch_ptr[h->device] : nullptr;
  if (!ws) return fail(h, MH_ERR_HIP, "no scratch address for device %d", h->device);
  // Nothing here needs to start at zero: the folded MPO matrix and the
  // intermediates are written in full before they are read (a split-K GEMM
  // clears its output with a kernel).
  h->noalloc = 1;
  const int st = heff_apply_ws(h, op, x, y, ws, nullptr, false);
  h->noalloc = 0;
  return st;
}

int mh_mode_product(mh_handle h, int dtype, int P, int ni, int nj, int Q, const void *in,
                    const void *Mt, void *out) {
  if (!h) return MH_ERR_ARG;
  if (dtype != MH_F64 && dtype != MH_C64) return fail(h, MH_ERR_DTYPE, "unknown dtype %d", dtype);
  if (P < 1 || Q < 1 || ni < 1 || nj < 1 || ni > 4096 || nj > 4096)
    return fail(h, MH_ERR_SHAPE, "mode product extents P=%d ni=%d nj=%d Q=%d", P, ni, nj, Q);
  if ((int64_t)P * ni * Q >= ((int64_t)1 << 31) || (int64_t)P * nj * Q >= ((int64_t)1 << 31))
    return fail(h, MH_ERR_SHAPE, "mode product exceeds 2^31 elements");
  if (!in || !Mt || !out) return fail(h, MH_ERR_ARG, "null tensor");
  if (in == out) return fail(h, MH_ERR_ARG, "in and out must not alias");
  return mode_product(h, dtype, P, ni, nj, Q, in, (int64_t)P * ni, Mt, out, (int64_t)P * nj);
}

int mh_ac_prime_open(mh_handle h, int dtype, int Dl, int d, int Dr, int wl, int wr,
                     const void *x, const void *W, const void *L, void *P) {
  if (!h) return MH_ERR_ARG;
  // the checks of ac_prime with a placeholder right environment
  mh_heff_op op{MH_HEFF1, dtype, MH_PREC_F64, Dl, d, 1, Dr, wl, 0, wr, W, nullptr, L, L};
  MH_TRY(heff_check(h, op));
  if (!x || !P) return fail(h, MH_ERR_ARG, "null tensor");
  if (x == P) return fail(h, MH_ERR_ARG, "x and P must not alias");
  const size_t b = dtype_bytes(dtype);
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  const size_t t1 = (size_t)Dl * wl * d * Dr, mt = (size_t)wl * d * d * wr;
  MH_TRY(ensure_ws(h, al(t1 * b) + al(mt * b)));
  char *T1 = (char *)h->ws;
  char *Mt = T1 + al(t1 * b);
  MH_TRY(build_mode_matrix(h, dtype, MODE_SITE, wl, 0, wr, d, 1, W, nullptr, Mt));
  if (h->fusion) {
    const int st = dtype == MH_F64 ? heff_step1_fused(h, Dl, Dr, wl, d, wr, L, x, Mt, P)
                                   : heff_step1_fused_c(h, Dl, Dr, wl, d, wr, L, x, Mt, P);
    if (st != MH_ERR_UNSUPPORTED) return st;
  }
  GemmDesc g1 = gd(Dl * wl, d * Dr, Dl, L, 1, (int64_t)Dl * wl, 0, x, 1, Dl, T1, 1,
                   (int64_t)Dl * wl);
  MH_TRY(gemm(h, dtype, MH_PREC_F64, g1));
  return mode_product(h, dtype, Dl, wl * d, d * wr, Dr, T1, (int64_t)Dl * wl * d, Mt, P,
                      (int64_t)Dl * d * wr);
}

int mh_c_prime(mh_handle h, int dtype, int prec, int Dl, int Dr, int w,
               const void *x, const void *L, const void *R, void *y) {
  mh_heff_op op{MH_HEFF0, dtype, prec, Dl, 1, 1, Dr, w, 0, w, nullptr, nullptr, L, R};
  return mh_heff_apply(h, &op, x, y);
}

int mh_updateleft(mh_handle h, int dtype, int prec, int Dl, int d, int Dr,
                  int wl, int wr, const void *hold, const void *psiA,
                  const void *W, const void *psiB, void *hnew) {
  MH_TRY(env_check(h, dtype, prec, Dl, d, Dr, wl, wr, hold, psiA, W, psiB, hnew));
  MH_TRY(ensure_ws(h, env_ws_bytes(dtype, Dl, d, Dr, wl, wr)));
  return env_left_ws(h, dtype, prec, Dl, d, Dr, wl, wr, hold, psiA, W, psiB, hnew, h->ws);
}

int mh_updateright(mh_handle h, int dtype, int prec, int Dl, int d, int Dr,
                   int wl, int wr, const void *hold, const void *psiA,
                   const void *W, const void *psiB, void *hnew) {
  MH_TRY(env_check(h, dtype, prec, Dl, d, Dr, wl, wr, hold, psiA, W, psiB, hnew));
  MH_TRY(ensure_ws(h, env_ws_bytes(dtype, Dl, d, Dr, wl, wr)));
  return env_right_ws(h, dtype, prec, Dl, d, Dr, wl, wr, hold, psiA, W, psiB, hnew, h->ws);
}

}  // extern "C"
