// One-workgroup Hermitian eigensolver for the small reduced density matrices
// of the dense two-site split (the FP64 step behind the truncated SVD of
// auxiliary/stable_svd.jl:4-18 when the bond is small).  At these orders
// (n <= ~100) a library syevd is a chain of launches whose latency (about
// 1.5 ms at n = 128) is many times the arithmetic; here one workgroup per
// matrix does the whole solve with the matrix in LDS:
//
//   1. Householder tridiagonalisation (Golub & Van Loan Alg. 8.3.1, complex
//      reflectors chosen so the off-diagonal comes out real: H = I - tau v v^H
//      with H^H x = beta e1, beta real; the LAPACK zhetd2 / dsytd2 recurrence
//      p = tau A v, w = p - (tau/2)(p^H v) v, A <- A - v w^H - w v^H).
//   2. Implicit QL iterations with Wilkinson shift on the real tridiagonal
//      (Bowdler, Martin, Reinsch & Wilkinson, Numer. Math. 11 (1968)): one
//      thread runs the scalar recurrence of a sweep and records its plane
//      rotations; every thread then applies them to its own row of Z.
//   3. Sort ascending (recorded swaps applied per row), then back-transform
//      the requested columns, z <- H_0 H_1 ... H_{n-2} z, one thread per column.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "mh_internal.h"

namespace mh {

namespace {

constexpr int HV_THREADS = 256;
constexpr int HV_ITERS = 60;  // QL iterations allowed per eigenvalue

__device__ inline double hv_re(double a) { return a; }
__device__ inline double hv_re(double2 a) { return a.x; }
__device__ inline double hv_im(double) { return 0.0; }
__device__ inline double hv_im(double2 a) { return a.y; }
__device__ inline double hv_abs2(double a) { return a * a; }
__device__ inline double hv_abs2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ inline double hv_conj(double a) { return a; }
__device__ inline double2 hv_conj(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline double hv_mul(double a, double b) { return a * b; }
__device__ inline double2 hv_mul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double hv_add(double a, double b) { return a + b; }
__device__ inline double2 hv_add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double hv_sub(double a, double b) { return a - b; }
__device__ inline double2 hv_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
template <class T> __device__ inline T hv_make(double r, double i);
template <> __device__ inline double hv_make<double>(double r, double) { return r; }
template <> __device__ inline double2 hv_make<double2>(double r, double i) { return make_double2(r, i); }

// Sum over the workgroup in a fixed order (deterministic); every thread gets
// the total.  Contains two barriers: call it uniformly.
__device__ inline double hv_sum(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < HV_THREADS / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

template <class T>
__device__ inline T hv_sum(T v, double *red) {
  return hv_make<T>(hv_sum(hv_re(v), red), hv_sum(hv_im(v), red));
}

// LDS layout of one workgroup (n = order, T = element type):
//   A[n*n] (T, column-major; reflector k kept in column k below the diagonal)
//   tau[n], pw[n] (T); d[n], e[n], cs[n], sn[n] (double); sw[n] (int)
template <class T>
__host__ __device__ inline size_t hv_lds_bytes(int n) {
  return (size_t)n * n * sizeof(T) + (size_t)2 * n * sizeof(T) + (size_t)4 * n * sizeof(double) +
         (size_t)n * sizeof(int) + 16;
}

template <class T>
__global__ __launch_bounds__(HV_THREADS) void k_heev_small(int n, int col0, const T *__restrict__ Ain,
                                                           double *__restrict__ W, T *__restrict__ Z,
                                                           int *__restrict__ info) {
  extern __shared__ double2 hv_lds[];
  __shared__ double red[HV_THREADS / 64];
  __shared__ T bscal;
  __shared__ int ctl[4];  // 0: sweep done flag, 1: lowest rotation, 2: highest, 3: failure
  T *A = (T *)hv_lds;
  T *tau = A + (size_t)n * n;
  T *pw = tau + n;
  double *d = (double *)(pw + n);
  double *e = d + n;
  double *cs = e + n;
  double *sn = cs + n;
  int *sw = (int *)(sn + n);

  const int tid = threadIdx.x;
  const size_t b = blockIdx.x;
  const T *Ab = Ain + b * n * n;
  T *Zb = Z + b * n * n;
  double *Wb = W + b * n;
  const T zero = hv_make<T>(0.0, 0.0), one = hv_make<T>(1.0, 0.0);

  for (int i = tid; i < n * n; i += HV_THREADS) A[i] = Ab[i];
  __syncthreads();

  // ---- 1. tridiagonalisation ---------------------------------------------
  for (int k = 0; k + 1 < n; ++k) {
    const int m = n - k - 1;
    T *x = A + (k + 1) + (size_t)n * k;  // column k below the diagonal
    double ss = 0.0;
    for (int i = 1 + tid; i < m; i += HV_THREADS) ss += hv_abs2(x[i]);
    const double xn2 = hv_sum(ss, red);
    if (tid == 0) {
      const T alpha = x[0];
      const double ar = hv_re(alpha), ai = hv_im(alpha);
      if (xn2 == 0.0 && ai == 0.0) {
        tau[k] = zero;
        e[k] = ar;
      } else {
        const double beta = -copysign(sqrt(ar * ar + ai * ai + xn2), ar);
        tau[k] = hv_make<T>((beta - ar) / beta, -ai / beta);
        e[k] = beta;
        // 1 / (alpha - beta)
        const double zr = ar - beta, zi = ai, den = zr * zr + zi * zi;
        bscal = hv_make<T>(zr / den, -zi / den);
        x[0] = one;
      }
    }
    __syncthreads();
    const T tk = tau[k];
    if (hv_re(tk) == 0.0 && hv_im(tk) == 0.0) continue;  // uniform: H_k = I
    const T sc = bscal;
    for (int i = 1 + tid; i < m; i += HV_THREADS) x[i] = hv_mul(x[i], sc);
    __syncthreads();
    T *A22 = A + (k + 1) + (size_t)n * (k + 1);
    for (int i = tid; i < m; i += HV_THREADS) {
      T acc = zero;
      for (int j = 0; j < m; ++j) acc = hv_add(acc, hv_mul(A22[i + (size_t)n * j], x[j]));
      pw[i] = hv_mul(tk, acc);
    }
    __syncthreads();
    T part = zero;
    for (int i = tid; i < m; i += HV_THREADS) part = hv_add(part, hv_mul(hv_conj(pw[i]), x[i]));
    const T pv = hv_sum(part, red);
    const T a2 = hv_mul(hv_make<T>(-0.5, 0.0), hv_mul(tk, pv));
    for (int i = tid; i < m; i += HV_THREADS) pw[i] = hv_add(pw[i], hv_mul(a2, x[i]));
    __syncthreads();
    for (int idx = tid; idx < m * m; idx += HV_THREADS) {
      const int i = idx % m, j = idx / m;
      const T u = hv_add(hv_mul(x[i], hv_conj(pw[j])), hv_mul(pw[i], hv_conj(x[j])));
      A22[i + (size_t)n * j] = hv_sub(A22[i + (size_t)n * j], u);
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += HV_THREADS) d[i] = hv_re(A[i + (size_t)n * i]);
  if (tid == 0) {
    e[n - 1] = 0.0;
    tau[n - 1] = zero;
    ctl[0] = 0;
    ctl[3] = 0;
  }
  for (int i = tid; i < n * n; i += HV_THREADS) Zb[i] = (i % (n + 1) == 0) ? one : zero;
  __syncthreads();

  // ---- 2. implicit QL on (d, e); e[i] couples i and i + 1 -----------------
  // A block splits where e[m] is negligible against its neighbours or against
  // the norm of T: the second test keeps the noise-level blocks of a
  // rank-deficient or strongly graded matrix (a density matrix) from stalling,
  // at the normwise accuracy eps * ||T|| the library solvers give.
  int l = 0, iter = 0;  // used by thread 0 only
  double small = 0.0;
  if (tid == 0) {
    double tn = 0.0;
    for (int i = 0; i < n; ++i)
      tn = fmax(tn, fabs(d[i]) + fabs(e[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0));
    small = DBL_EPSILON * tn;
  }
  for (;;) {
    if (tid == 0) {
      ctl[1] = 1;
      ctl[2] = 0;  // empty rotation range unless a sweep is recorded
      for (;;) {
        if (l >= n) {
          ctl[0] = 1;
          break;
        }
        int mm = l;
        for (; mm < n - 1; ++mm) {
          const double dd = fabs(d[mm]) + fabs(d[mm + 1]);
          if (fabs(e[mm]) <= DBL_EPSILON * dd || fabs(e[mm]) <= small) break;
        }
        if (mm == l) {
          ++l;
          iter = 0;
          continue;
        }
        if (++iter > HV_ITERS) {
          ctl[3] = 1;
          ctl[0] = 1;
          break;
        }
        // Wilkinson shift from the leading 2 x 2 block, folded into g
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = hypot(g, 1.0);
        g = d[mm] - d[l] + e[l] / (g + copysign(r, g));
        double s = 1.0, c = 1.0, p = 0.0;
        int i = mm - 1;
        bool deflated = false;
        for (; i >= l; --i) {
          const double f = s * e[i], bb = c * e[i];
          r = hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {  // the sweep split the matrix: restart from l
            d[i + 1] -= p;
            e[mm] = 0.0;
            deflated = true;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * bb;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - bb;
          cs[i] = c;
          sn[i] = s;
        }
        if (!deflated) {
          d[l] -= p;
          e[l] = g;
          e[mm] = 0.0;
        }
        ctl[1] = deflated ? i + 1 : l;  // rotations ctl[1] .. ctl[2] were recorded
        ctl[2] = mm - 1;
        break;
      }
    }
    __syncthreads();
    const int done = ctl[0], lo = ctl[1], hi = ctl[2];
    if (lo <= hi) {
      for (int r = tid; r < n; r += HV_THREADS) {
        double zc = hv_re(Zb[r + (size_t)n * (hi + 1)]);
        for (int i = hi; i >= lo; --i) {
          const double zi = hv_re(Zb[r + (size_t)n * i]);
          Zb[r + (size_t)n * (i + 1)] = hv_make<T>(sn[i] * zi + cs[i] * zc, 0.0);
          zc = cs[i] * zi - sn[i] * zc;
        }
        Zb[r + (size_t)n * lo] = hv_make<T>(zc, 0.0);
      }
    }
    __syncthreads();
    if (done) break;
  }
  const bool failed = ctl[3] != 0;

  // ---- 3. ascending order, then z <- Q z for the requested columns --------
  if (tid == 0) {
    for (int i = 0; i + 1 < n; ++i) {
      int j = i;
      for (int q = i + 1; q < n; ++q)
        if (d[q] < d[j]) j = q;
      sw[i] = j;
      const double t = d[i];
      d[i] = d[j];
      d[j] = t;
    }
    if (info) info[b] = failed ? 1 : 0;
  }
  __syncthreads();
  for (int r = tid; r < n; r += HV_THREADS)
    for (int i = 0; i + 1 < n; ++i) {
      const int j = sw[i];
      if (j == i) continue;
      const T t = Zb[r + (size_t)n * i];
      Zb[r + (size_t)n * i] = Zb[r + (size_t)n * j];
      Zb[r + (size_t)n * j] = t;
    }
  for (int i = tid; i < n; i += HV_THREADS) Wb[i] = failed ? NAN : d[i];
  __syncthreads();
  for (int j = col0 + tid; j < n; j += HV_THREADS) {
    T *z = Zb + (size_t)n * j;
    for (int k = n - 2; k >= 0; --k) {
      const T tk = tau[k];
      if (hv_re(tk) == 0.0 && hv_im(tk) == 0.0) continue;
      const T *v = A + (k + 1) + (size_t)n * k;
      const int m = n - k - 1;
      T s = zero;
      for (int i = 0; i < m; ++i) s = hv_add(s, hv_mul(hv_conj(v[i]), z[k + 1 + i]));
      s = hv_mul(tk, s);
      for (int i = 0; i < m; ++i) z[k + 1 + i] = hv_sub(z[k + 1 + i], hv_mul(v[i], s));
    }
  }
}

int hv_lds_limit(mh_handle h) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device) != hipSuccess)
    return 65536;
  return v;
}

int hv_max_order(mh_handle h, int dtype) {
  // static shared memory of the kernel (reduction slots, control words) beside
  // the dynamic part
  const size_t lim = (size_t)hv_lds_limit(h) - 256;
  int n = 1;
  if (dtype == MH_C64)
    while (n < 1024 && hv_lds_bytes<double2>(n + 1) <= lim) ++n;
  else
    while (n < 1024 && hv_lds_bytes<double>(n + 1) <= lim) ++n;
  return n;
}

}  // namespace

}  // namespace mh

using namespace mh;

extern "C" int mh_heev_small_max(mh_handle h, int dtype, int *nmax) {
  if (!h || !nmax) return MH_ERR_ARG;
  if (dtype != MH_F64 && dtype != MH_C64) return fail(h, MH_ERR_DTYPE, "dtype %d", dtype);
  MH_HIP(h, hipSetDevice(h->device));
  *nmax = hv_max_order(h, dtype);
  return MH_OK;
}

extern "C" int mh_heev_small(mh_handle h, int dtype, int n, int batch, const void *A, double *W,
                             void *Z, int col0, int *info) {
  if (!h) return MH_ERR_ARG;
  if (dtype != MH_F64 && dtype != MH_C64) return fail(h, MH_ERR_DTYPE, "dtype %d", dtype);
  if (!A || !W || !Z) return fail(h, MH_ERR_ARG, "null matrix, eigenvalue or vector buffer");
  MH_HIP(h, hipSetDevice(h->device));
  const int nmax = hv_max_order(h, dtype);
  if (n < 1 || n > nmax || batch < 1)
    return fail(h, MH_ERR_SHAPE, "order %d outside [1, %d] or batch %d < 1", n, nmax, batch);
  if (col0 < 0 || col0 > n) return fail(h, MH_ERR_SHAPE, "col0 %d outside [0, %d]", col0, n);
  if (batch > 65535) return fail(h, MH_ERR_SHAPE, "batch %d above 65535", batch);
  if (dtype == MH_C64) {
    const size_t lds = hv_lds_bytes<double2>(n);
    if (lds > 65536)
      MH_HIP(h, hipFuncSetAttribute((const void *)k_heev_small<double2>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_heev_small<double2>, dim3(batch), dim3(HV_THREADS), lds, h->stream, n, col0,
                       (const double2 *)A, W, (double2 *)Z, info);
  } else {
    const size_t lds = hv_lds_bytes<double>(n);
    if (lds > 65536)
      MH_HIP(h, hipFuncSetAttribute((const void *)k_heev_small<double>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_heev_small<double>, dim3(batch), dim3(HV_THREADS), lds, h->stream, n, col0,
                       (const double *)A, W, (double *)Z, info);
  }
  MH_HIP(h, hipGetLastError());
  return MH_OK;
}
