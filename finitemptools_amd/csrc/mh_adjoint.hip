// Adjoint of the one-site H_eff / environment contraction with respect to the
// MPO tensor:
//
//   G[w, s', w', s] = sum_{a,a',b,b'} L[a,w,a'] conj(A[a,s',b]) B[a',s,b'] R[b,w',b']
//
// so that <A | H_eff(W) | B> = sum G . W (derivatives.jl:2-4 with the MPO
// leg left open).  G is the pullback of ac_prime / updateleft / updateright
// onto W, and dE/dtheta = sum_i <G_i, dW_i/dtheta> for an MPO H(theta).
//
// Order (cost 4 D^3 w d + 2 D^2 w^2 d^2, the cost of one one-site H_eff):
//   T1[(a,w),(s,b')]  = L[(a,w),a'] * B[a',(s,b')]            (MFMA GEMM)
//   U [(a,s'),(w',b')] = conj(A)[(a,s'),b] * R[b,(w',b')]       (MFMA GEMM)
//   G[(w,s),(s',w')]  = sum_{a,b'} T1[a,(w,s),b'] U[a,(s',w'),b']  (pair reduce)
// The pair reduction has a small output (w d x d w) and a long reduction
// (D^2): each workgroup reduces a range of b' into its own partial G through
// LDS tiles, and a second kernel adds the partials in a fixed order, so the
// result is deterministic.
#include <hip/hip_runtime.h>

#include "mh_internal.h"

namespace mh {

constexpr int PR_MAXPQ = 64;   // max of w*d on either side
constexpr int PR_MAXOUT = 1024; // max (w d)^2: four outputs per thread

template <bool CPLX>
__global__ void __launch_bounds__(256)
    k_pair_partial(const double *__restrict__ T1, const double *__restrict__ U, int Dl, int P,
                   int Q, int Dr, int chunk, double *__restrict__ part) {
  constexpr int EW = CPLX ? 2 : 1;
  constexpr int PR_TA = CPLX ? 32 : 64;  // a-rows per LDS tile
  __shared__ double sT[PR_TA * PR_MAXPQ * EW];
  __shared__ double sU[PR_TA * PR_MAXPQ * EW];
  const int tid = threadIdx.x;
  const int b0 = blockIdx.x * chunk, b1 = min(Dr, b0 + chunk);
  const int PQ = P * Q;
  // each thread owns outputs o = tid, tid + 256, ... (o = p + P*q)
  double acc[4][EW];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < EW; ++c) acc[i][c] = 0.0;
  for (int bp = b0; bp < b1; ++bp) {
    const double *t = T1 + (int64_t)EW * Dl * P * bp;  // T1[:, :, bp] is Dl x P
    const double *u = U + (int64_t)EW * Dl * Q * bp;
    for (int a0 = 0; a0 < Dl; a0 += PR_TA) {
      const int na = min(PR_TA, Dl - a0);
      __syncthreads();
      for (int e = tid; e < na * P; e += 256) {
        const int a = e % na, p = e / na;
#pragma unroll
        for (int c = 0; c < EW; ++c) sT[(p * PR_TA + a) * EW + c] = t[(int64_t)EW * (a0 + a + (int64_t)Dl * p) + c];
      }
      for (int e = tid; e < na * Q; e += 256) {
        const int a = e % na, q = e / na;
#pragma unroll
        for (int c = 0; c < EW; ++c) sU[(q * PR_TA + a) * EW + c] = u[(int64_t)EW * (a0 + a + (int64_t)Dl * q) + c];
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int o = tid + 256 * i;
        if (o >= PQ) break;
        const int p = o % P, q = o / P;
        const double *x = sT + p * PR_TA * EW;
        const double *y = sU + q * PR_TA * EW;
        if constexpr (CPLX) {
          double re = acc[i][0], im = acc[i][1];
          for (int a = 0; a < na; ++a) {
            const double xr = x[2 * a], xi = x[2 * a + 1], yr = y[2 * a], yi = y[2 * a + 1];
            re += xr * yr - xi * yi;
            im += xr * yi + xi * yr;
          }
          acc[i][0] = re;
          acc[i][1] = im;
        } else {
          double s = acc[i][0];
          for (int a = 0; a < na; ++a) s += x[a] * y[a];
          acc[i][0] = s;
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int o = tid + 256 * i;
    if (o >= PQ) break;
#pragma unroll
    for (int c = 0; c < EW; ++c) part[((int64_t)blockIdx.x * PQ + o) * EW + c] = acc[i][c];
  }
}

// G[w, s', w', s] = sum over chunks of part[chunk][(w,s)][(s',w')], fixed order.
template <bool CPLX>
__global__ void k_pair_finish(const double *__restrict__ part, int nchunk, int wl, int d,
                              int wr, double *__restrict__ G) {
  constexpr int EW = CPLX ? 2 : 1;
  const int P = wl * d, Q = d * wr, PQ = P * Q;
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= PQ) return;
  const int p = o % P, q = o / P;
  const int w = p % wl, s = p / wl;     // p = w + wl*s
  const int so = q % d, wo = q / d;     // q = s' + d*w'
  const int64_t g = w + (int64_t)wl * (so + (int64_t)d * (wo + (int64_t)wr * s));
#pragma unroll
  for (int c = 0; c < EW; ++c) {
    double v = 0.0;
    for (int k = 0; k < nchunk; ++k) v += part[((int64_t)k * PQ + o) * EW + c];
    G[g * EW + c] = v;
  }
}

}  // namespace mh

using namespace mh;

extern "C" {

int mh_mpo_grad(mh_handle h, int dtype, int prec, int Dl, int d, int Dr, int wl, int wr,
                const void *L, const void *psiA, const void *psiB, const void *R, void *G) {
  if (!h) return MH_ERR_ARG;
  if (dtype != MH_F64 && dtype != MH_C64) return fail(h, MH_ERR_DTYPE, "unknown dtype %d", dtype);
  if (prec < MH_PREC_F64 || prec > MH_PREC_BF16X3)
    return fail(h, MH_ERR_DTYPE, "unknown precision %d", prec);
  if (Dl <= 0 || d <= 0 || Dr <= 0 || wl <= 0 || wr <= 0)
    return fail(h, MH_ERR_SHAPE, "non-positive extent");
  if (!L || !psiA || !psiB || !R || !G) return fail(h, MH_ERR_ARG, "null tensor");
  if (wl * d > PR_MAXPQ || wr * d > PR_MAXPQ || wl * d * d * wr > PR_MAXOUT)
    return fail(h, MH_ERR_UNSUPPORTED, "MPO gradient needs w*d <= %d and (w d)^2 <= %d",
                PR_MAXPQ, PR_MAXOUT);
  const int64_t lim = (int64_t)1 << 31;
  if ((int64_t)Dl * wl * d * Dr >= lim || (int64_t)Dl * wr * d * Dr >= lim)
    return fail(h, MH_ERR_SHAPE, "intermediate exceeds 2^31 elements");
  const size_t eb = dtype_bytes(dtype);
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  const int P = wl * d, Q = d * wr;
  const int chunk = Dr >= 1024 ? (Dr + 1023) / 1024 : 1;
  const int nchunk = (Dr + chunk - 1) / chunk;
  const size_t t1b = al((size_t)Dl * P * Dr * eb), ub = al((size_t)Dl * Q * Dr * eb);
  const size_t pb = al((size_t)nchunk * P * Q * eb);
  MH_TRY(ensure_ws(h, t1b + ub + pb));
  char *ws = (char *)h->ws;
  void *T1 = ws, *U = ws + t1b;
  double *part = (double *)(ws + t1b + ub);
  // T1[(a,w),(s,b')] = L[(a,w),a'] B[a',(s,b')]
  MH_TRY(gemm(h, dtype, prec,
              GemmDesc{Dl * wl, d * Dr, Dl, L, 1, (int64_t)Dl * wl, psiB, 1, Dl, T1, 1,
                       (int64_t)Dl * wl, 1.0, 0.0, 0.0, 0.0, 0}));
  // U[(a,s'),(w',b')] = conj(A)[(a,s'),b] R[b,(w',b')]
  MH_TRY(gemm(h, dtype, prec,
              GemmDesc{Dl * d, wr * Dr, Dr, psiA, 1, (int64_t)Dl * d, R, 1, Dr, U, 1,
                       (int64_t)Dl * d, 1.0, 0.0, 0.0, 0.0, 1}));
  int slot;
  prof_begin(h, MH_K_WSTAGE, 2.0 * Dl * (double)Dr * P * Q * (dtype == MH_C64 ? 4 : 1), &slot);
  if (dtype == MH_C64) {
    hipLaunchKernelGGL(k_pair_partial<true>, dim3(nchunk), dim3(256), 0, h->stream,
                       (const double *)T1, (const double *)U, Dl, P, Q, Dr, chunk, part);
    hipLaunchKernelGGL(k_pair_finish<true>, dim3((P * Q + 255) / 256), dim3(256), 0, h->stream,
                       part, nchunk, wl, d, wr, (double *)G);
  } else {
    hipLaunchKernelGGL(k_pair_partial<false>, dim3(nchunk), dim3(256), 0, h->stream,
                       (const double *)T1, (const double *)U, Dl, P, Q, Dr, chunk, part);
    hipLaunchKernelGGL(k_pair_finish<false>, dim3((P * Q + 255) / 256), dim3(256), 0, h->stream,
                       part, nchunk, wl, d, wr, (double *)G);
  }
  prof_end(h, slot);
  return check_launch(h, "mpo gradient");
}

}  // extern "C"
