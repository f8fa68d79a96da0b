// Krylov drivers with the basis resident in HBM.
//
// The Lanczos recurrence of simple_lanczos_solver (auxiliary/simple_lanczos.jl)
// runs as a fixed sequence of kernels per step; alpha, beta, the accepted
// dimension k and the early-exit flag live in device memory, so the host does
// not synchronise inside the loop.  After beta < btol the next basis vector is
// written as zeros, which keeps every later (discarded) step finite.  One copy
// of the tridiagonal coefficients to the host ends the loop; the k x k
// eigenproblem (or exponential) is solved there in FP64 and the Ritz/result
// vector is combined on the device.
//
// Reductions are two-pass (fixed-order block partials, then one block), so
// results are bitwise reproducible run to run.
#include <cmath>
#include <complex>
#include <cstring>
#include <functional>

#include "mh_internal.h"

namespace mh {

int heff_apply_ws(mh_handle h, const mh_heff_op &op, const void *x, void *y,
                  void *ws, const void *Mt_prebuilt, bool lr_ready);
int heff_prepare_mixed(mh_handle h, const mh_heff_op &op, void *ws);
int build_mode_matrix(mh_handle h, int dtype, int kind, int wl, int wm, int wr,
                      int d1, int d2, const void *W1, const void *W2, void *Mt);

constexpr int RB = 256;      // threads per reduction block
constexpr int MAXPART = 1024;

static int nblocks(size_t ndbl) {
  size_t g = (ndbl + RB * 8 - 1) / (RB * 8);
  if (g < 1) g = 1;
  if (g > MAXPART) g = MAXPART;
  return (int)g;
}

__device__ __forceinline__ double block_sum(double v, double *sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = RB / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// part[2*b + {0,1}] = block partials of <x, y> (conjugate-linear in x).
// n counts scalars (complex: pairs of doubles).
template <bool CPLX>
__global__ void __launch_bounds__(RB) k_dot_part(const double *__restrict__ x,
                                                 const double *__restrict__ y,
                                                 size_t n, double *__restrict__ part) {
  __shared__ double sh[RB];
  double re = 0.0, im = 0.0;
  for (size_t i = (size_t)blockIdx.x * RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) {
    if (CPLX) {
      double2 a = reinterpret_cast<const double2 *>(x)[i];
      double2 b = reinterpret_cast<const double2 *>(y)[i];
      re += a.x * b.x + a.y * b.y;
      im += a.x * b.y - a.y * b.x;
    } else {
      re += x[i] * y[i];
    }
  }
  re = block_sum(re, sh);
  if (CPLX) im = block_sum(im, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = re;
    part[2 * blockIdx.x + 1] = im;
  }
}

// Reduce G partial pairs.  mode 0: out = (re, im); mode 1: out = sqrt(re).
__global__ void __launch_bounds__(RB) k_reduce(const double *__restrict__ part, int G,
                                               double *__restrict__ out, int mode) {
  __shared__ double sh[RB];
  double re = 0.0, im = 0.0;
  for (int i = threadIdx.x; i < G; i += RB) {
    re += part[2 * i];
    im += part[2 * i + 1];
  }
  re = block_sum(re, sh);
  im = block_sum(im, sh);
  if (threadIdx.x == 0) {
    if (mode == 0) {
      out[0] = re;
      out[1] = im;
    } else {
      out[0] = sqrt(re);
    }
  }
}

// |x|^2 block partials.
template <bool CPLX>
__global__ void __launch_bounds__(RB) k_nrm_part(const double *__restrict__ x, size_t n,
                                                 double *__restrict__ part) {
  __shared__ double sh[RB];
  double s = 0.0;
  const size_t nd = CPLX ? 2 * n : n;
  for (size_t i = (size_t)blockIdx.x * RB + threadIdx.x; i < nd; i += (size_t)gridDim.x * RB)
    s += x[i] * x[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s;
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

// y = a x + b y (complex scalars; real data uses the real parts).
template <bool CPLX>
__global__ void k_axpby(size_t n, double ar, double ai, const double *__restrict__ x,
                        double br, double bi, double *__restrict__ y) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (CPLX) {
    double2 u = reinterpret_cast<const double2 *>(x)[i];
    double2 v = reinterpret_cast<double2 *>(y)[i];
    double2 r;
    r.x = ar * u.x - ai * u.y + br * v.x - bi * v.y;
    r.y = ar * u.y + ai * u.x + br * v.y + bi * v.x;
    reinterpret_cast<double2 *>(y)[i] = r;
  } else {
    y[i] = ar * x[i] + br * y[i];
  }
}

// ------------------------------------------------------------ Lanczos state
// st[0] done flag, st[1] accepted dimension k, st[2] max |Im alpha|,
// st[3] initial norm, st[4..5] scratch (re, im of the last dot),
// st[8 .. 8+kmax) alpha, st[8+kmax .. 8+2kmax) beta,
// st[8+2kmax .. 8+2kmax+2*(kmax+1)) combination coefficients (re, im),
// then 2*MAXPART partials, (kmax+1)*MAXPART*2 multi-dot partials and the
// Arnoldi Hessenberg columns H[j, i] (re, im) at o_h + 2*(j + (kmax+1)*i).
// Projected operators (excited states): up to NPMAX projector vectors, their
// multi-dot partials at o_ppart and coefficients at o_pcoef.
constexpr int NPMAX = 8;
struct LanLayout {
  int kmax;
  int o_alpha, o_beta, o_coef, o_part, o_mpart, o_h, o_ppart, o_pcoef, total;
  explicit LanLayout(int k) : kmax(k) {
    o_alpha = 8;
    o_beta = o_alpha + k;
    o_coef = o_beta + k;
    o_part = o_coef + 2 * (k + 1);
    o_mpart = o_part + 2 * MAXPART;
    o_h = o_mpart + 2 * MAXPART * (k + 1);
    o_ppart = o_h + 2 * (k + 1) * k;
    o_pcoef = o_ppart + 2 * MAXPART * NPMAX;
    total = o_pcoef + 2 * NPMAX;
  }
};

__global__ void k_fill(double *p, int n, double v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// dst = done ? 0 : src / (*s)  (s == 0 also writes zeros)
template <bool CPLX>
__global__ void k_div(size_t n, const double *__restrict__ src, double *__restrict__ dst,
                      const double *__restrict__ s, const double *__restrict__ done) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nd = CPLX ? 2 * n : n;
  if (i >= nd) return;
  const double sv = *s;
  const bool z = (done && *done != 0.0) || sv == 0.0;
  dst[i] = z ? 0.0 : src[i] / sv;
}

// alpha_i = Re <q, w>, tracked |Im| ; finalises the dot of step i.
__global__ void k_take_alpha(const double *__restrict__ dotres, double *__restrict__ st,
                             int o_alpha, int i) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (st[0] == 0.0) {
      st[o_alpha + i] = dotres[0];
      double ai = fabs(dotres[1]);
      if (ai > st[2]) st[2] = ai;
    }
  }
}

// w -= alpha_i q_i + beta_{i-1} q_{i-1}; accumulate |w|^2 partials.
template <bool CPLX>
__global__ void __launch_bounds__(RB) k_lan_update(size_t n, double *__restrict__ w,
                                                   const double *__restrict__ q,
                                                   const double *__restrict__ qm,
                                                   const double *__restrict__ st,
                                                   int o_alpha, int o_beta, int i,
                                                   double *__restrict__ part) {
  __shared__ double sh[RB];
  const double a = st[o_alpha + i];
  const double b = (i > 0) ? st[o_beta + i - 1] : 0.0;
  const size_t nd = CPLX ? 2 * n : n;
  double s = 0.0;
  for (size_t j = (size_t)blockIdx.x * RB + threadIdx.x; j < nd; j += (size_t)gridDim.x * RB) {
    double v = w[j] - a * q[j];
    if (i > 0) v -= b * qm[j];
    w[j] = v;
    s += v * v;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s;
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

// beta_i = sqrt(sum partials); update k and the done flag.
__global__ void __launch_bounds__(RB) k_take_beta(const double *__restrict__ part, int G,
                                                  double *__restrict__ st, int o_beta,
                                                  int i, double btol) {
  __shared__ double sh[RB];
  double s = 0.0;
  for (int j = threadIdx.x; j < G; j += RB) s += part[2 * j];
  s = block_sum(s, sh);
  if (threadIdx.x == 0 && st[0] == 0.0) {
    double b = sqrt(s);
    st[o_beta + i] = b;
    st[1] = (double)(i + 1);
    if (b < btol) st[0] = 1.0;
  }
}

// Multi-dot partials: part[(j*G + blk)*2 + {0,1}] = partial <V_j, w>, j < nv.
template <bool CPLX>
__global__ void __launch_bounds__(RB) k_mdot_part(const double *__restrict__ V, size_t n,
                                                  int nv, const double *__restrict__ w,
                                                  double *__restrict__ part) {
  __shared__ double sh[RB];
  const int G = gridDim.x;
  for (int j = 0; j < nv; ++j) {
    const double *x = V + (size_t)j * (CPLX ? 2 * n : n);
    double re = 0.0, im = 0.0;
    for (size_t i = (size_t)blockIdx.x * RB + threadIdx.x; i < n; i += (size_t)G * RB) {
      if (CPLX) {
        double2 a = reinterpret_cast<const double2 *>(x)[i];
        double2 b = reinterpret_cast<const double2 *>(w)[i];
        re += a.x * b.x + a.y * b.y;
        im += a.x * b.y - a.y * b.x;
      } else {
        re += x[i] * w[i];
      }
    }
    re = block_sum(re, sh);
    if (CPLX) im = block_sum(im, sh);
    if (threadIdx.x == 0) {
      part[2 * ((size_t)j * G + blockIdx.x)] = re;
      part[2 * ((size_t)j * G + blockIdx.x) + 1] = im;
    }
  }
}

// coef[j] = sum of partials for vector j (one block per j).
__global__ void __launch_bounds__(RB) k_mdot_reduce(const double *__restrict__ part, int G,
                                                    double *__restrict__ coef) {
  __shared__ double sh[RB];
  const int j = blockIdx.x;
  double re = 0.0, im = 0.0;
  for (int b = threadIdx.x; b < G; b += RB) {
    re += part[2 * ((size_t)j * G + b)];
    im += part[2 * ((size_t)j * G + b) + 1];
  }
  re = block_sum(re, sh);
  im = block_sum(im, sh);
  if (threadIdx.x == 0) {
    coef[2 * j] = re;
    coef[2 * j + 1] = im;
  }
}

// coef_j -> conj(coef_j): the projection of dmrgexcited.jl:1-6 subtracts
// dot(y, p) p with dot conjugate-linear in y, i.e. conj(<p, y>) p.
__global__ void k_conj_coef(double *coef, int nv) {
  const int j = threadIdx.x;
  if (j < nv) coef[2 * j + 1] = -coef[2 * j + 1];
}

// Arnoldi: H[0..nv, i] = coef (pass 0) or += coef (pass 1), frozen once done.
__global__ void k_hcol(double *__restrict__ st, int o_h, int ld, int i, int nv,
                       const double *__restrict__ coef, int pass) {
  const int j = threadIdx.x;
  if (j >= nv || st[0] != 0.0) return;
  double *hc = st + o_h + 2 * ((size_t)j + (size_t)ld * i);
  if (pass == 0) {
    hc[0] = coef[2 * j];
    hc[1] = coef[2 * j + 1];
  } else {
    hc[0] += coef[2 * j];
    hc[1] += coef[2 * j + 1];
  }
}

// w -= sum_j coef_j V_j (skipped once done); optional |w|^2 partials.
template <bool CPLX>
__global__ void __launch_bounds__(RB) k_msub(size_t n, int nv, const double *__restrict__ V,
                                             const double *__restrict__ coef,
                                             double *__restrict__ w,
                                             const double *__restrict__ done,
                                             double *__restrict__ part) {
  __shared__ double sh[RB];
  const bool skip = done && *done != 0.0;
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) {
    if (CPLX) {
      double2 v = reinterpret_cast<double2 *>(w)[i];
      if (!skip)
        for (int j = 0; j < nv; ++j) {
          double2 x = reinterpret_cast<const double2 *>(V + (size_t)j * 2 * n)[i];
          double cr = coef[2 * j], ci = coef[2 * j + 1];
          v.x -= cr * x.x - ci * x.y;
          v.y -= cr * x.y + ci * x.x;
        }
      reinterpret_cast<double2 *>(w)[i] = v;
      s += v.x * v.x + v.y * v.y;
    } else {
      double v = w[i];
      if (!skip)
        for (int j = 0; j < nv; ++j) v -= coef[2 * j] * V[(size_t)j * n + i];
      w[i] = v;
      s += v * v;
    }
  }
  if (part) {
    s = block_sum(s, sh);
    if (threadIdx.x == 0) {
      part[2 * blockIdx.x] = s;
      part[2 * blockIdx.x + 1] = 0.0;
    }
  }
}

// out = scale * sum_j coef_j V_j (complex coefficients).
template <bool CPLX>
__global__ void k_combine(size_t n, int nv, const double *__restrict__ V,
                          const double *__restrict__ coef, double *__restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (CPLX) {
    double2 acc = make_double2(0.0, 0.0);
    for (int j = nv - 1; j >= 0; --j) {
      double2 x = reinterpret_cast<const double2 *>(V + (size_t)j * 2 * n)[i];
      double cr = coef[2 * j], ci = coef[2 * j + 1];
      acc.x += cr * x.x - ci * x.y;
      acc.y += cr * x.y + ci * x.x;
    }
    reinterpret_cast<double2 *>(out)[i] = acc;
  } else {
    double acc = 0.0;
    for (int j = nv - 1; j >= 0; --j) acc += coef[2 * j] * V[(size_t)j * n + i];
    out[i] = acc;
  }
}

// ------------------------------------------------ one-workgroup Krylov step
// For short vectors (at most SMALL_ND doubles) the vector part of a Lanczos
// or Arnoldi step -- alpha, the three-term update, the Gram-Schmidt passes,
// beta and the next basis vector -- is ONE launch of one 1024-thread
// workgroup instead of ten to fourteen launches, which at these sizes cost
// launch latency, not bandwidth (sixteen waves keep enough loads in flight
// to cover the L2 latency of each pass).  Same recurrence, state layout and
// early-exit rules as the multi-block kernels above.  Every thread owns the
// vector elements e = t, t + 1024, ... in every pass, so w needs no barrier
// between passes; sums are per thread in index order, then a fixed
// LDS/shuffle tree, so results are reproducible run to run.
constexpr int SB_T = 1024;
constexpr int SB_W = SB_T / 64;
constexpr size_t SMALL_ND = 16384;
constexpr int SB_J = 8;      // basis vectors per multi-dot batch
constexpr int SB_VMAX = 128; // basis vectors per step (kmax limit)

// out[q] = workgroup-wide sum of v[q], q < NV (a shared array), for NV = 1,
// 2 (wave shuffles, then the SB_W wave sums in order) or NV = 2*SB_J (the
// partials transposed through LDS: SB_T/NV consecutive threads per value add
// NV entries each, then shuffle within their group).  sh holds NV*SB_T doubles.
template <int NV>
__device__ __forceinline__ void sb_sum(double (&v)[NV], double *sh, double *out) {
  const int t = threadIdx.x;
  if constexpr (NV <= 2) {
    const int lane = t & 63, wv = t >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double x = v[q];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
      if (lane == 0) sh[wv * NV + q] = x;
    }
    __syncthreads();
    if (t < NV) {
      double x = 0.0;
      for (int k = 0; k < SB_W; ++k) x += sh[k * NV + t];
      out[t] = x;
    }
  } else {
    constexpr int TPV = SB_T / NV;  // threads per value (64 for NV = 16)
    static_assert(SB_T % NV == 0 && TPV <= 64 && (TPV & (TPV - 1)) == 0, "reduction shape");
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[q * SB_T + t] = v[q];
    __syncthreads();
    const int q = t / TPV, r = t % TPV;
    double x = 0.0;
#pragma unroll
    for (int k = 0; k < SB_T / TPV; ++k) x += sh[q * SB_T + r + k * TPV];
#pragma unroll
    for (int o = TPV / 2; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if (r == 0) out[q] = x;
  }
  __syncthreads();
}

// mode 0: Lanczos step; 1: Lanczos step with two Gram-Schmidt passes
// against V_0..V_i; 2: Arnoldi step (the passes give the Hessenberg column
// H[0..i, i] at st[o_h + 2*(j + ldh*i)]).  next (may be null) receives
// V_{i+1}.  n counts scalars (complex: pairs of doubles).
template <bool CPLX>
__global__ void __launch_bounds__(SB_T)
    k_step_small(size_t n, const double *__restrict__ V, double *__restrict__ w,
                 double *__restrict__ next, double *__restrict__ st, int o_alpha, int o_beta,
                 int o_h, int ldh, int i, int mode, double btol) {
  __shared__ double sh[2 * SB_J * SB_T];
  __shared__ double res[2 * SB_J];
  __shared__ double coef[2 * SB_VMAX];
  __shared__ double scal[4];
  const int t = threadIdx.x;
  const size_t nd = CPLX ? 2 * n : n;
  const bool done = st[0] != 0.0;  // written (by thread 0) only after the last read below
  const double *vi = V + (size_t)i * nd;
  double s2 = 0.0;
  if (mode != 2) {
    double p[2] = {0.0, 0.0};
    for (size_t e = t; e < n; e += SB_T) {
      if (CPLX) {
        const double2 a = reinterpret_cast<const double2 *>(vi)[e];
        const double2 b = reinterpret_cast<const double2 *>(w)[e];
        p[0] += a.x * b.x + a.y * b.y;
        p[1] += a.x * b.y - a.y * b.x;
      } else {
        p[0] += vi[e] * w[e];
      }
    }
    sb_sum<2>(p, sh, res);
    if (t == 0) {
      double a = st[o_alpha + i];
      if (!done) {
        a = res[0];
        st[o_alpha + i] = a;
        const double ai = fabs(res[1]);
        if (ai > st[2]) st[2] = ai;
      }
      scal[0] = a;
      scal[1] = i > 0 ? st[o_beta + i - 1] : 0.0;
    }
    __syncthreads();
    const double a = scal[0], b = scal[1];
    const double *vm = i > 0 ? V + (size_t)(i - 1) * nd : vi;
    for (size_t e = t; e < nd; e += SB_T) {
      double v = w[e] - a * vi[e];
      if (i > 0) v -= b * vm[e];
      w[e] = v;
      s2 += v * v;
    }
  }
  if (mode >= 1) {
    const int nv = i + 1;
    for (int pass = 0; pass < 2; ++pass) {
      for (int j0 = 0; j0 < nv; j0 += SB_J) {
        double p[2 * SB_J];
#pragma unroll
        for (int q = 0; q < 2 * SB_J; ++q) p[q] = 0.0;
        for (size_t e = t; e < n; e += SB_T) {
          if (CPLX) {
            const double2 b = reinterpret_cast<const double2 *>(w)[e];
#pragma unroll
            for (int q = 0; q < SB_J; ++q)
              if (j0 + q < nv) {
                const double2 a = reinterpret_cast<const double2 *>(V + (size_t)(j0 + q) * nd)[e];
                p[2 * q] += a.x * b.x + a.y * b.y;
                p[2 * q + 1] += a.x * b.y - a.y * b.x;
              }
          } else {
            const double b = w[e];
#pragma unroll
            for (int q = 0; q < SB_J; ++q)
              if (j0 + q < nv) p[2 * q] += V[(size_t)(j0 + q) * nd + e] * b;
          }
        }
        sb_sum<2 * SB_J>(p, sh, res);
        if (t < 2 * SB_J && j0 + t / 2 < nv) coef[2 * j0 + t] = res[t];
      }
      __syncthreads();
      if (mode == 2 && t < nv && !done) {
        double *hc = st + o_h + 2 * ((size_t)t + (size_t)ldh * i);
        if (pass == 0) {
          hc[0] = coef[2 * t];
          hc[1] = coef[2 * t + 1];
        } else {
          hc[0] += coef[2 * t];
          hc[1] += coef[2 * t + 1];
        }
      }
      double s = 0.0;
      for (size_t e = t; e < n; e += SB_T) {
        if (CPLX) {
          double2 v = reinterpret_cast<double2 *>(w)[e];
          if (!done)
            for (int j = 0; j < nv; ++j) {
              const double2 x = reinterpret_cast<const double2 *>(V + (size_t)j * nd)[e];
              const double cr = coef[2 * j], ci = coef[2 * j + 1];
              v.x -= cr * x.x - ci * x.y;
              v.y -= cr * x.y + ci * x.x;
            }
          reinterpret_cast<double2 *>(w)[e] = v;
          s += v.x * v.x + v.y * v.y;
        } else {
          double v = w[e];
          if (!done)
            for (int j = 0; j < nv; ++j) v -= coef[2 * j] * V[(size_t)j * nd + e];
          w[e] = v;
          s += v * v;
        }
      }
      s2 = s;
      __syncthreads();  // coef is rewritten by the next pass
    }
  }
  double pb[1] = {s2};
  sb_sum<1>(pb, sh, res);
  if (t == 0) {
    if (!done) {
      const double b = sqrt(res[0]);
      st[o_beta + i] = b;
      st[1] = (double)(i + 1);
      if (b < btol) st[0] = 1.0;
    }
    scal[2] = st[0];
    scal[3] = st[o_beta + i];
  }
  __syncthreads();
  if (next) {
    const double sv = scal[3];
    const bool z = scal[2] != 0.0 || sv == 0.0;
    for (size_t e = t; e < nd; e += SB_T) next[e] = z ? 0.0 : w[e] / sv;
  }
}

// Start vector for short vectors: st[3] = |x0|, V_0 = x0 / |x0| (zeros for
// a zero vector), one launch.
__global__ void __launch_bounds__(SB_T)
    k_start_small(size_t nd, const double *__restrict__ x0, double *__restrict__ v0,
                  double *__restrict__ st) {
  __shared__ double sh[SB_W];
  __shared__ double res[1];  // sb_sum<1> uses SB_W entries of sh
  const int t = threadIdx.x;
  double p[1] = {0.0};
  for (size_t e = t; e < nd; e += SB_T) p[0] += x0[e] * x0[e];
  sb_sum<1>(p, sh, res);
  const double sv = sqrt(res[0]);
  if (t == 0) st[3] = sv;
  for (size_t e = t; e < nd; e += SB_T) v0[e] = sv == 0.0 ? 0.0 : x0[e] / sv;
}

// ---------------------------------------------------------- host helpers

// Eigen-decomposition of a real symmetric k x k matrix (column-major) by the
// cyclic Jacobi method; eigenvalues ascending, eigenvectors in columns of Z.
void sym_eig(int k, const double *A, double *evals, double *Z) {
  std::vector<double> a(A, A + (size_t)k * k);
  for (int i = 0; i < k * k; ++i) Z[i] = 0.0;
  for (int i = 0; i < k; ++i) Z[i + k * i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) {
        double v = a[i + k * j] * a[i + k * j];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-300 || off <= tot * 1e-34) break;
    for (int p = 0; p < k - 1; ++p)
      for (int q = p + 1; q < k; ++q) {
        double apq = a[p + k * q];
        if (apq == 0.0) continue;
        double app = a[p + k * p], aqq = a[q + k * q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int r = 0; r < k; ++r) {  // columns p, q
          double arp = a[r + k * p], arq = a[r + k * q];
          a[r + k * p] = c * arp - s * arq;
          a[r + k * q] = s * arp + c * arq;
        }
        for (int r = 0; r < k; ++r) {  // rows p, q
          double apr = a[p + k * r], aqr = a[q + k * r];
          a[p + k * r] = c * apr - s * aqr;
          a[q + k * r] = s * apr + c * aqr;
        }
        for (int r = 0; r < k; ++r) {
          double zrp = Z[r + k * p], zrq = Z[r + k * q];
          Z[r + k * p] = c * zrp - s * zrq;
          Z[r + k * q] = s * zrp + c * zrq;
        }
      }
  }
  std::vector<int> idx(k);
  for (int i = 0; i < k; ++i) idx[i] = i;
  for (int i = 0; i < k; ++i)
    for (int j = i + 1; j < k; ++j)
      if (a[idx[j] + k * idx[j]] < a[idx[i] + k * idx[i]]) std::swap(idx[i], idx[j]);
  std::vector<double> Zs((size_t)k * k);
  for (int c = 0; c < k; ++c) {
    evals[c] = a[idx[c] + k * idx[c]];
    // Sign convention: largest-magnitude component positive.
    int im = 0;
    for (int r = 0; r < k; ++r)
      if (fabs(Z[r + k * idx[c]]) > fabs(Z[im + k * idx[c]])) im = r;
    double sg = Z[im + k * idx[c]] < 0 ? -1.0 : 1.0;
    for (int r = 0; r < k; ++r) Zs[r + k * c] = sg * Z[r + k * idx[c]];
  }
  std::memcpy(Z, Zs.data(), sizeof(double) * k * k);
}

// exp(A) of a small complex k x k matrix (column-major) by scaling and
// squaring with the diagonal [8/8] Pade approximant (Moler & Van Loan,
// "Nineteen dubious ways", method 3): A / 2^s with |A/2^s|_1 <= 1/2,
// N = sum c_j X^j, D = sum (-1)^j c_j X^j, exp(X) ~ D^-1 N (Gaussian
// elimination with partial pivoting), then s squarings.
typedef std::complex<double> cplx_t;
static void expm_small(int k, const std::vector<cplx_t> &A, std::vector<cplx_t> &E) {
  auto idx = [k](int r, int c) { return r + (size_t)k * c; };
  double nrm = 0.0;
  for (int c = 0; c < k; ++c) {
    double s = 0.0;
    for (int r = 0; r < k; ++r) s += std::abs(A[idx(r, c)]);
    nrm = std::max(nrm, s);
  }
  int sq = 0;
  while (nrm > 0.5 && sq < 200) {
    nrm *= 0.5;
    ++sq;
  }
  const double scale = std::ldexp(1.0, -sq);
  std::vector<cplx_t> X(A.size()), P((size_t)k * k), N((size_t)k * k, 0.0), D((size_t)k * k, 0.0);
  for (size_t i = 0; i < A.size(); ++i) X[i] = A[i] * scale;
  auto mul = [&](const std::vector<cplx_t> &a, const std::vector<cplx_t> &b, std::vector<cplx_t> &o) {
    std::vector<cplx_t> t((size_t)k * k, 0.0);
    for (int c = 0; c < k; ++c)
      for (int m = 0; m < k; ++m) {
        const cplx_t bv = b[idx(m, c)];
        if (bv == 0.0) continue;
        for (int r = 0; r < k; ++r) t[idx(r, c)] += a[idx(r, m)] * bv;
      }
    o.swap(t);
  };
  const int q = 8;
  double cj = 1.0;
  for (int r = 0; r < k; ++r) P[idx(r, r)] = 1.0;  // X^0
  for (int j = 0; j <= q; ++j) {
    if (j > 0) {
      cj *= (double)(q - j + 1) / (double)(j * (2 * q - j + 1));
      mul(P, X, P);
    }
    const double sg = (j % 2) ? -1.0 : 1.0;
    for (size_t i = 0; i < P.size(); ++i) {
      N[i] += cj * P[i];
      D[i] += sg * cj * P[i];
    }
  }
  // Solve D E = N.
  std::vector<int> piv(k);
  for (int c = 0; c < k; ++c) {
    int p = c;
    for (int r = c + 1; r < k; ++r)
      if (std::abs(D[idx(r, c)]) > std::abs(D[idx(p, c)])) p = r;
    piv[c] = p;
    if (p != c) {
      for (int m = 0; m < k; ++m) std::swap(D[idx(c, m)], D[idx(p, m)]);
      for (int m = 0; m < k; ++m) std::swap(N[idx(c, m)], N[idx(p, m)]);
    }
    const cplx_t d = D[idx(c, c)];
    for (int r = c + 1; r < k; ++r) {
      const cplx_t f = D[idx(r, c)] / d;
      if (f == 0.0) continue;
      for (int m = c; m < k; ++m) D[idx(r, m)] -= f * D[idx(c, m)];
      for (int m = 0; m < k; ++m) N[idx(r, m)] -= f * N[idx(c, m)];
    }
  }
  E.assign((size_t)k * k, 0.0);
  for (int m = 0; m < k; ++m)
    for (int r = k - 1; r >= 0; --r) {
      cplx_t s = N[idx(r, m)];
      for (int c = r + 1; c < k; ++c) s -= D[idx(r, c)] * E[idx(c, m)];
      E[idx(r, m)] = s / D[idx(r, r)];
    }
  for (int i = 0; i < sq; ++i) mul(E, E, E);
}

// ---------------------------------------------------------- drivers

struct Krylov {
  mh_handle h;
  mh_heff_op op;                     // dense operator (when bs is null)
  const mh_bs_plan_s *bs = nullptr;  // block-sparse operator plan
  const void *bp1 = nullptr, *bp2 = nullptr;
  bool cplx;
  size_t n, nd;  // scalars, doubles per vector
  int kmax;
  LanLayout lay;
  char *heff_ws;
  double *V, *w;
  void *Mt;
  double *st;
  int G;
  bool arnoldi = false;  // full orthogonalisation, Hessenberg H (non-Hermitian H_eff)
  int nproj = 0;                        // projectors applied after every H_eff x
  const void *const *proj_src = nullptr;
  double *P = nullptr;                  // their copies, stride nd

  Krylov(mh_handle hh, const mh_heff_op &o, int k)
      : h(hh), op(o), cplx(o.dtype == MH_C64), kmax(k), lay(k) {
    n = heff_vec_len(op);
    nd = cplx ? 2 * n : n;
    G = nblocks(nd);
  }
  Krylov(mh_handle hh, const mh_bs_plan_s &p, const void *p1, const void *p2, int k)
      : h(hh), op{}, bs(&p), bp1(p1), bp2(p2), cplx(false), kmax(k), lay(k) {
    n = bs_vec_len(p);
    nd = n;
    G = nblocks(nd);
  }

  int apply(const double *x, double *y) {
    if (bs) MH_TRY(bs_apply_ws(h, *bs, x, bp1, bp2, y, heff_ws));
    else MH_TRY(heff_apply_ws(h, op, x, y, heff_ws, op.kind == MH_HEFF0 ? nullptr : Mt, true));
    return project(y);
  }

  // for l in order: y -= dot(y, p_l) p_l (dmrgexcited.jl:1-6), each with
  // the y left by the previous projector, as the reference loops.
  int project(double *y) {
    double *pp = st + lay.o_ppart, *pc = st + lay.o_pcoef;
    for (int l = 0; l < nproj; ++l) {
      const double *pl = P + (size_t)l * nd;
      if (cplx) hipLaunchKernelGGL(k_mdot_part<true>, dim3(G), dim3(RB), 0, h->stream, pl, n, 1, y, pp);
      else hipLaunchKernelGGL(k_mdot_part<false>, dim3(G), dim3(RB), 0, h->stream, pl, n, 1, y, pp);
      hipLaunchKernelGGL(k_mdot_reduce, dim3(1), dim3(RB), 0, h->stream, pp, G, pc);
      if (cplx) hipLaunchKernelGGL(k_conj_coef, dim3(1), dim3(64), 0, h->stream, pc, 1);
      if (cplx) hipLaunchKernelGGL(k_msub<true>, dim3(G), dim3(RB), 0, h->stream, n, 1, pl, pc, y, nullptr, nullptr);
      else hipLaunchKernelGGL(k_msub<false>, dim3(G), dim3(RB), 0, h->stream, n, 1, pl, pc, y, nullptr, nullptr);
    }
    return nproj ? check_launch(h, "projection") : MH_OK;
  }

  int setup() {
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    size_t hb = bs ? bs_ws_bytes(*bs) : heff_ws_bytes(op);
    // V_0 .. V_{kmax-1} and w, packed at stride nd (the multi-vector kernels
    // index V_j at j * nd), then room for the MPO mode matrix.
    size_t vb = al(nd * sizeof(double) * (kmax + 1));
    size_t mtb = al(sizeof(double) * 2 * 4096 * 4);
    size_t pb = al(nd * sizeof(double) * nproj);
    MH_TRY(ensure_ws(h, hb + vb + mtb + pb));
    heff_ws = (char *)h->ws;
    V = (double *)(heff_ws + hb);
    w = V + (size_t)kmax * nd;
    Mt = (char *)V + vb;
    P = (double *)((char *)Mt + mtb);
    for (int l = 0; l < nproj; ++l)
      MH_HIP(h, hipMemcpyAsync(P + (size_t)l * nd, proj_src[l], nd * sizeof(double),
                               hipMemcpyDeviceToDevice, h->stream));
    MH_TRY(ensure_red(h, lay.total));
    st = h->red;
    // Mixed forms: the environments are packed once for the whole solve.
    if (!bs && op.prec != MH_PREC_F64) MH_TRY(heff_prepare_mixed(h, op, heff_ws));
    if (!bs && op.kind != MH_HEFF0) {
      int kind = op.kind == MH_HEFF2 ? MODE_TWO : MODE_SITE;
      MH_TRY(build_mode_matrix(h, op.dtype, kind, op.wl, op.wm, op.wr, op.d1, op.d2,
                               op.W1, op.W2, Mt));
    }
    hipLaunchKernelGGL(k_fill, dim3((lay.total + 255) / 256), dim3(256), 0, h->stream, st,
                       lay.total, 0.0);
    return check_launch(h, "lanczos init");
  }

  double *vec(int j) { return V + (size_t)j * nd; }

  int launch_nrm(const double *x, double *part) {
    if (cplx) hipLaunchKernelGGL(k_nrm_part<true>, dim3(G), dim3(RB), 0, h->stream, x, n, part);
    else hipLaunchKernelGGL(k_nrm_part<false>, dim3(G), dim3(RB), 0, h->stream, x, n, part);
    return check_launch(h, "nrm");
  }

  // normalised start vector into V_0, norm into st[3] (projected first when
  // the operator carries projectors, as the reference projects its start
  // vector, dmrgexcited.jl:90)
  int start(const void *x0) {
    double *part = st + lay.o_part;
    if (nproj) {
      MH_HIP(h, hipMemcpyAsync(w, x0, nd * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
      MH_TRY(project(w));
      x0 = w;
    }
    if (small()) {
      hipLaunchKernelGGL(k_start_small, dim3(1), dim3(SB_T), 0, h->stream, nd,
                         (const double *)x0, vec(0), st);
      return check_launch(h, "lanczos start");
    }
    MH_TRY(launch_nrm((const double *)x0, part));
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(RB), 0, h->stream, part, G, st + 3, 1);
    unsigned gb = (unsigned)((nd + 255) / 256);
    if (cplx) hipLaunchKernelGGL(k_div<true>, dim3(gb), dim3(256), 0, h->stream, n, (const double *)x0, vec(0), st + 3, nullptr);
    else hipLaunchKernelGGL(k_div<false>, dim3(gb), dim3(256), 0, h->stream, n, (const double *)x0, vec(0), st + 3, nullptr);
    return check_launch(h, "lanczos start");
  }

  // One Arnoldi step i: w = A V_i, two classical Gram-Schmidt passes against
  // V_0..V_i whose coefficients sum to the Hessenberg column H[0..i, i],
  // H[i+1, i] = |w| (kept as beta_i), V_{i+1} = w / |w|.
  int step_arnoldi(int i, double btol, bool write_next) {
    double *part = st + lay.o_part;
    MH_TRY(apply(vec(i), w));
    int slot;
    prof_begin(h, MH_K_VEC, 0.0, &slot);
    double *mpart = st + lay.o_mpart;
    double *coef = st + lay.o_coef;
    for (int pass = 0; pass < 2; ++pass) {
      if (cplx) hipLaunchKernelGGL(k_mdot_part<true>, dim3(G), dim3(RB), 0, h->stream, V, vstride_n(), i + 1, w, mpart);
      else hipLaunchKernelGGL(k_mdot_part<false>, dim3(G), dim3(RB), 0, h->stream, V, vstride_n(), i + 1, w, mpart);
      hipLaunchKernelGGL(k_mdot_reduce, dim3(i + 1), dim3(RB), 0, h->stream, mpart, G, coef);
      hipLaunchKernelGGL(k_hcol, dim3(1), dim3(128), 0, h->stream, st, lay.o_h, kmax + 1, i, i + 1, coef, pass);
      double *pp = (pass == 1) ? part : nullptr;
      if (cplx) hipLaunchKernelGGL(k_msub<true>, dim3(G), dim3(RB), 0, h->stream, vstride_n(), i + 1, V, coef, w, st, pp);
      else hipLaunchKernelGGL(k_msub<false>, dim3(G), dim3(RB), 0, h->stream, vstride_n(), i + 1, V, coef, w, st, pp);
    }
    hipLaunchKernelGGL(k_take_beta, dim3(1), dim3(RB), 0, h->stream, part, G, st, lay.o_beta, i, btol);
    if (write_next) {
      unsigned gb = (unsigned)((nd + 255) / 256);
      if (cplx) hipLaunchKernelGGL(k_div<true>, dim3(gb), dim3(256), 0, h->stream, n, w, vec(i + 1), st + lay.o_beta + i, st);
      else hipLaunchKernelGGL(k_div<false>, dim3(gb), dim3(256), 0, h->stream, n, w, vec(i + 1), st + lay.o_beta + i, st);
    }
    prof_end(h, slot);
    return check_launch(h, "arnoldi step");
  }

  bool small() const { return h->krylov_small && nd <= SMALL_ND && kmax <= SB_VMAX; }

  // Any step on short vectors: the matvec, then one workgroup for the rest.
  int step_small(int i, double btol, int mode, bool write_next) {
    MH_TRY(apply(vec(i), w));
    int slot;
    prof_begin(h, MH_K_VEC, 0.0, &slot);
    double *nx = write_next ? vec(i + 1) : nullptr;
    if (cplx) hipLaunchKernelGGL(k_step_small<true>, dim3(1), dim3(SB_T), 0, h->stream, n, V, w, nx, st, lay.o_alpha, lay.o_beta, lay.o_h, kmax + 1, i, mode, btol);
    else hipLaunchKernelGGL(k_step_small<false>, dim3(1), dim3(SB_T), 0, h->stream, n, V, w, nx, st, lay.o_alpha, lay.o_beta, lay.o_h, kmax + 1, i, mode, btol);
    prof_end(h, slot);
    return check_launch(h, "krylov step");
  }

  // One Lanczos step i: w = H V_i, alpha_i, w -= ..., (reorth), beta_i, V_{i+1}.
  int step(int i, double btol, int reorth, bool write_next) {
    if (small()) return step_small(i, btol, arnoldi ? 2 : (reorth ? 1 : 0), write_next);
    if (arnoldi) return step_arnoldi(i, btol, write_next);
    double *part = st + lay.o_part;
    MH_TRY(apply(vec(i), w));
    int slot;
    prof_begin(h, MH_K_VEC, 0.0, &slot);
    if (cplx) hipLaunchKernelGGL(k_dot_part<true>, dim3(G), dim3(RB), 0, h->stream, vec(i), w, n, part);
    else hipLaunchKernelGGL(k_dot_part<false>, dim3(G), dim3(RB), 0, h->stream, vec(i), w, n, part);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(RB), 0, h->stream, part, G, st + 4, 0);
    hipLaunchKernelGGL(k_take_alpha, dim3(1), dim3(64), 0, h->stream, st + 4, st, lay.o_alpha, i);
    const double *qm = i > 0 ? vec(i - 1) : vec(i);
    if (cplx) hipLaunchKernelGGL(k_lan_update<true>, dim3(G), dim3(RB), 0, h->stream, n, w, vec(i), qm, st, lay.o_alpha, lay.o_beta, i, part);
    else hipLaunchKernelGGL(k_lan_update<false>, dim3(G), dim3(RB), 0, h->stream, n, w, vec(i), qm, st, lay.o_alpha, lay.o_beta, i, part);
    if (reorth) {
      // Two passes of classical Gram-Schmidt against V_0..V_i in FP64.
      double *mpart = st + lay.o_mpart;
      double *coef = st + lay.o_coef;
      for (int pass = 0; pass < 2; ++pass) {
        if (cplx) hipLaunchKernelGGL(k_mdot_part<true>, dim3(G), dim3(RB), 0, h->stream, V, vstride_n(), i + 1, w, mpart);
        else hipLaunchKernelGGL(k_mdot_part<false>, dim3(G), dim3(RB), 0, h->stream, V, vstride_n(), i + 1, w, mpart);
        hipLaunchKernelGGL(k_mdot_reduce, dim3(i + 1), dim3(RB), 0, h->stream, mpart, G, coef);
        double *pp = (pass == 1) ? part : nullptr;
        if (cplx) hipLaunchKernelGGL(k_msub<true>, dim3(G), dim3(RB), 0, h->stream, vstride_n(), i + 1, V, coef, w, st, pp);
        else hipLaunchKernelGGL(k_msub<false>, dim3(G), dim3(RB), 0, h->stream, vstride_n(), i + 1, V, coef, w, st, pp);
      }
    }
    hipLaunchKernelGGL(k_take_beta, dim3(1), dim3(RB), 0, h->stream, part, G, st, lay.o_beta, i, btol);
    if (write_next) {
      unsigned gb = (unsigned)((nd + 255) / 256);
      if (cplx) hipLaunchKernelGGL(k_div<true>, dim3(gb), dim3(256), 0, h->stream, n, w, vec(i + 1), st + lay.o_beta + i, st);
      else hipLaunchKernelGGL(k_div<false>, dim3(gb), dim3(256), 0, h->stream, n, w, vec(i + 1), st + lay.o_beta + i, st);
    }
    prof_end(h, slot);
    return check_launch(h, "lanczos step");
  }

  size_t vstride_n() const { return n; }

  int read_state(std::vector<double> &hs) {
    hs.resize(lay.o_coef);
    MH_HIP(h, hipMemcpyAsync(h->red_host, st, sizeof(double) * lay.o_coef, hipMemcpyDeviceToHost, h->stream));
    MH_HIP(h, hipStreamSynchronize(h->stream));
    std::memcpy(hs.data(), h->red_host, sizeof(double) * lay.o_coef);
    return MH_OK;
  }

  // Arnoldi: the Hessenberg columns, hc[2*(j + (kmax+1)*i) + {0,1}].
  int read_hessenberg(std::vector<double> &hc) {
    const size_t nh = 2 * (size_t)(kmax + 1) * kmax;
    hc.resize(nh);
    MH_HIP(h, hipMemcpyAsync(h->red_host, st + lay.o_h, sizeof(double) * nh, hipMemcpyDeviceToHost, h->stream));
    MH_HIP(h, hipStreamSynchronize(h->stream));
    std::memcpy(hc.data(), h->red_host, sizeof(double) * nh);
    return MH_OK;
  }

  int combine(int k, const double *coef2, void *out) {
    std::memcpy(h->red_host, coef2, sizeof(double) * 2 * k);
    MH_HIP(h, hipMemcpyAsync(st + lay.o_coef, h->red_host, sizeof(double) * 2 * k, hipMemcpyHostToDevice, h->stream));
    unsigned gb = (unsigned)((n + 255) / 256);
    if (cplx) hipLaunchKernelGGL(k_combine<true>, dim3(gb), dim3(256), 0, h->stream, n, k, V, st + lay.o_coef, (double *)out);
    else hipLaunchKernelGGL(k_combine<false>, dim3(gb), dim3(256), 0, h->stream, n, k, V, st + lay.o_coef, (double *)out);
    MH_TRY(check_launch(h, "combine"));
    // The host buffer is reused by the next call: wait for the upload.
    MH_HIP(h, hipStreamSynchronize(h->stream));
    return MH_OK;
  }
};

// Lanczos steps [1, kmax) as one hipGraph.  The steps have no host
// synchronisation (alpha, beta and the early-exit flag live in device
// memory), so they are captured once on the handle's private stream and the
// executable graph is launched on the handle's stream; a driver that repeats
// the same steps (the sub-steps of exponentiate) launches it again.  Step 0
// runs eagerly before, so tuners have fixed their choices (capture skips
// tuning).  Without graphs, or if capture is refused, the steps run eagerly.
struct StepGraph {
  hipGraphExec_t exec = nullptr;
  hipStream_t launched_on = nullptr;
  ~StepGraph() {
    if (exec) {
      if (launched_on) (void)hipStreamSynchronize(launched_on);
      (void)hipGraphExecDestroy(exec);
    }
  }
};

static int run_steps(mh_handle h, Krylov &K, int kmax, double btol, int reorth, StepGraph &g) {
  if (kmax <= 1) return MH_OK;
  if (g.exec) {
    g.launched_on = h->stream;
    MH_HIP(h, hipGraphLaunch(g.exec, h->stream));
    return MH_OK;
  }
  auto eager = [&]() -> int {
    for (int i = 1; i < kmax; ++i) MH_TRY(K.step(i, btol, reorth, i + 1 < kmax));
    return MH_OK;
  };
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  (void)hipStreamIsCapturing(h->stream, &cs);
  // short-vector steps are a few launches each: capture and instantiation
  // cost more than the graph saves unless capture is forced (graphs == 2)
  if (!h->graphs || (h->graphs == 1 && K.small()) || h->profiling ||
      cs != hipStreamCaptureStatusNone)
    return eager();
  if (!h->cap_stream &&
      hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    h->cap_stream = nullptr;
    return eager();
  }
  hipStream_t saved = h->stream;
  h->stream = h->cap_stream;
  if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    h->stream = saved;
    return eager();
  }
  int st = MH_OK;
  for (int i = 1; i < kmax && st == MH_OK; ++i) st = K.step(i, btol, reorth, i + 1 < kmax);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(h->stream, &graph);
  h->stream = saved;
  hipError_t ei = hipErrorUnknown;
  if (st == MH_OK && e == hipSuccess && graph)
    ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    // nothing of the captured sequence ran: do it eagerly
    (void)hipGetLastError();
    g.exec = nullptr;
    h->last_error.clear();
    return eager();
  }
  g.launched_on = h->stream;
  MH_HIP(h, hipGraphLaunch(g.exec, h->stream));
  return MH_OK;
}

// Lanczos eigensolve on a prepared Krylov driver (dense or block-sparse).
static int lanczos_eig_run(mh_handle h, Krylov &K, const void *x0, int kmax, double btol,
                           int reorth, double *eigval, void *xout, int *info, int *niter) {
  MH_TRY(K.setup());
  MH_TRY(K.start(x0));
  MH_TRY(K.step(0, btol, reorth, kmax > 1));
  StepGraph g;
  MH_TRY(run_steps(h, K, kmax, btol, reorth, g));
  std::vector<double> hs;
  MH_TRY(K.read_state(hs));
  if (!(hs[3] > 0.0) || !std::isfinite(hs[3]))
    return fail(h, MH_ERR_ARG, "initial vector has zero or non-finite norm");
  int k = (int)hs[1];
  bool converged = hs[0] != 0.0;
  if (k < 1) k = 1;
  std::vector<double> T((size_t)k * k, 0.0), ev(k), Z((size_t)k * k);
  for (int i = 0; i < k; ++i) {
    T[i + k * i] = hs[K.lay.o_alpha + i];
    if (i + 1 < k) {
      T[i + 1 + k * i] = hs[K.lay.o_beta + i];
      T[i + k * (i + 1)] = hs[K.lay.o_beta + i];
    }
  }
  sym_eig(k, T.data(), ev.data(), Z.data());
  *eigval = ev[0];  // "SR": smallest
  std::vector<double> c2(2 * k, 0.0);
  for (int i = 0; i < k; ++i) c2[2 * i] = Z[i];
  MH_TRY(K.combine(k, c2.data(), xout));
  if (info) *info = converged ? 0 : -1;
  if (niter) *niter = k;
  if (!std::isfinite(*eigval)) return fail(h, MH_ERR_NOCONV, "non-finite Ritz value");
  return MH_OK;
}

// exp(dt H_eff) x0 by Krylov sub-steps: Lanczos (Hermitian H_eff, the
// tridiagonal exponentiated through its eigen-decomposition) or Arnoldi
// (any H_eff, the Hessenberg matrix exponentiated by Pade).  Each sub-step
// takes the largest tau (halving from the remaining time) whose a-posteriori
// error beta0 * h_{k,k-1} * |e_k^T exp(tau H_k) e1| is below tol * |tau|.
static int krylov_expm(mh_handle h, const mh_heff_op *op, const void *x0, double dt_re,
                       double dt_im, int kdim, double tol, bool arnoldi, void *xout,
                       int *info) {
  if (!h || !op) return MH_ERR_ARG;
  MH_TRY(heff_check(h, *op));
  if (!x0 || !xout) return fail(h, MH_ERR_ARG, "null vector");
  if (kdim < 2 || kdim > 64) return fail(h, MH_ERR_ARG, "kdim must be in [2, 64]");
  if (op->dtype == MH_F64 && dt_im != 0.0)
    return fail(h, MH_ERR_DTYPE, "complex time step needs MH_C64 data");
  if (!(tol > 0.0)) return fail(h, MH_ERR_ARG, "tol must be positive");
  const std::complex<double> dt_total(dt_re, dt_im);
  const double adt = std::abs(dt_total);
  if (adt == 0.0) {
    size_t n = heff_vec_len(*op);
    MH_HIP(h, hipMemcpyAsync(xout, x0, n * dtype_bytes(op->dtype), hipMemcpyDeviceToDevice, h->stream));
    if (info) *info = 0;
    return MH_OK;
  }
  Krylov K(h, *op, kdim);
  K.arnoldi = arnoldi;
  MH_TRY(K.setup());
  StepGraph graph;
  const size_t n = heff_vec_len(*op);
  const void *src = x0;
  std::complex<double> done_t(0.0, 0.0);
  int nsub = 0, status = 0;
  while (true) {
    std::complex<double> rem = dt_total - done_t;
    if (std::abs(rem) <= 1e-14 * adt) break;
    if (++nsub > 4096) { status = -1; break; }
    // Rebuild the Krylov state for this sub-step (st is cleared by setup()).
    if (nsub > 1) {
      hipLaunchKernelGGL(k_fill, dim3((K.lay.total + 255) / 256), dim3(256), 0, h->stream, K.st, K.lay.total, 0.0);
    }
    MH_TRY(K.start(src));
    // Happy breakdown threshold relative to the vector norm 1.  Steps
    // 1..kdim-1 are the same launches in every sub-step: one graph.
    MH_TRY(K.step(0, 1e-13, 1, kdim > 1));
    MH_TRY(run_steps(h, K, kdim, 1e-13, 1, graph));
    std::vector<double> hs;
    MH_TRY(K.read_state(hs));
    const double b0 = hs[3];
    if (!(b0 > 0.0)) {  // zero vector stays zero
      MH_HIP(h, hipMemcpyAsync(xout, src, n * dtype_bytes(op->dtype), hipMemcpyDeviceToDevice, h->stream));
      break;
    }
    int k = (int)hs[1];
    if (k < 1) k = 1;
    const bool invariant = hs[0] != 0.0;
    const double bk = hs[K.lay.o_beta + k - 1];
    std::function<void(std::complex<double>, std::vector<std::complex<double>> &)> expT_e1;
    std::vector<double> T, ev, Z, hc;
    std::vector<cplx_t> Hk;
    if (!arnoldi) {
      T.assign((size_t)k * k, 0.0);
      ev.resize(k);
      Z.resize((size_t)k * k);
      for (int i = 0; i < k; ++i) {
        T[i + k * i] = hs[K.lay.o_alpha + i];
        if (i + 1 < k) T[i + 1 + k * i] = T[i + k * (i + 1)] = hs[K.lay.o_beta + i];
      }
      sym_eig(k, T.data(), ev.data(), Z.data());
      expT_e1 = [&](std::complex<double> tau, std::vector<std::complex<double>> &f) {
        f.assign(k, 0.0);
        for (int c = 0; c < k; ++c) {
          std::complex<double> e = std::exp(tau * ev[c]) * Z[0 + k * c];
          for (int r = 0; r < k; ++r) f[r] += Z[r + k * c] * e;
        }
      };
    } else {
      MH_TRY(K.read_hessenberg(hc));
      Hk.assign((size_t)k * k, 0.0);
      for (int i = 0; i < k; ++i) {
        for (int j = 0; j <= i; ++j) {
          const size_t o = 2 * ((size_t)j + (size_t)(kdim + 1) * i);
          Hk[j + (size_t)k * i] = cplx_t(hc[o], op->dtype == MH_C64 ? hc[o + 1] : 0.0);
        }
        if (i + 1 < k) Hk[i + 1 + (size_t)k * i] = hs[K.lay.o_beta + i];
      }
      expT_e1 = [&](std::complex<double> tau, std::vector<std::complex<double>> &f) {
        std::vector<cplx_t> A(Hk.size()), E;
        for (size_t q = 0; q < Hk.size(); ++q) A[q] = tau * Hk[q];
        expm_small(k, A, E);
        f.assign(E.begin(), E.begin() + k);  // first column
      };
    }
    std::complex<double> tau = rem;
    std::vector<std::complex<double>> f;
    for (int tries = 0; tries < 60; ++tries) {
      expT_e1(tau, f);
      double err = invariant ? 0.0 : b0 * bk * std::abs(f[k - 1]);
      if (err <= tol * std::abs(tau)) break;
      tau *= 0.5;
    }
    std::vector<double> c2(2 * k);
    for (int i = 0; i < k; ++i) {
      c2[2 * i] = b0 * f[i].real();
      c2[2 * i + 1] = b0 * f[i].imag();
    }
    if (op->dtype == MH_F64)
      for (int i = 0; i < k; ++i) c2[2 * i + 1] = 0.0;
    MH_TRY(K.combine(k, c2.data(), xout));
    src = xout;
    done_t += tau;
  }
  if (info) *info = status;
  return status == 0 ? MH_OK : fail(h, MH_ERR_NOCONV, "exponentiate: too many sub-steps");
}

}  // namespace mh

using namespace mh;

extern "C" {

int mh_dot(mh_handle h, int dtype, size_t n, const void *x, const void *y, double *out) {
  if (!h || !out) return MH_ERR_ARG;
  if (dtype != MH_F64 && dtype != MH_C64) return fail(h, MH_ERR_DTYPE, "unknown dtype");
  if (n == 0) { out[0] = out[1] = 0.0; return MH_OK; }
  if (!x || !y) return fail(h, MH_ERR_ARG, "null vector");
  MH_TRY(ensure_red(h, 2 * MAXPART + 8));
  size_t nd = dtype == MH_C64 ? 2 * n : n;
  int G = nblocks(nd);
  double *part = h->red + 8;
  if (dtype == MH_C64) hipLaunchKernelGGL(k_dot_part<true>, dim3(G), dim3(RB), 0, h->stream, (const double *)x, (const double *)y, n, part);
  else hipLaunchKernelGGL(k_dot_part<false>, dim3(G), dim3(RB), 0, h->stream, (const double *)x, (const double *)y, n, part);
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(RB), 0, h->stream, part, G, h->red, 0);
  MH_TRY(check_launch(h, "dot"));
  MH_HIP(h, hipMemcpyAsync(h->red_host, h->red, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  MH_HIP(h, hipStreamSynchronize(h->stream));
  out[0] = h->red_host[0];
  out[1] = dtype == MH_C64 ? h->red_host[1] : 0.0;
  return MH_OK;
}

int mh_nrm2(mh_handle h, int dtype, size_t n, const void *x, double *out) {
  if (!h || !out) return MH_ERR_ARG;
  if (dtype != MH_F64 && dtype != MH_C64) return fail(h, MH_ERR_DTYPE, "unknown dtype");
  if (n == 0) { *out = 0.0; return MH_OK; }
  if (!x) return fail(h, MH_ERR_ARG, "null vector");
  MH_TRY(ensure_red(h, 2 * MAXPART + 8));
  size_t nd = dtype == MH_C64 ? 2 * n : n;
  int G = nblocks(nd);
  double *part = h->red + 8;
  if (dtype == MH_C64) hipLaunchKernelGGL(k_nrm_part<true>, dim3(G), dim3(RB), 0, h->stream, (const double *)x, n, part);
  else hipLaunchKernelGGL(k_nrm_part<false>, dim3(G), dim3(RB), 0, h->stream, (const double *)x, n, part);
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(RB), 0, h->stream, part, G, h->red, 1);
  MH_TRY(check_launch(h, "nrm2"));
  MH_HIP(h, hipMemcpyAsync(h->red_host, h->red, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  MH_HIP(h, hipStreamSynchronize(h->stream));
  *out = h->red_host[0];
  return MH_OK;
}

int mh_axpby(mh_handle h, int dtype, size_t n, double a_re, double a_im, const void *x,
             double b_re, double b_im, void *y) {
  if (!h) return MH_ERR_ARG;
  if (dtype != MH_F64 && dtype != MH_C64) return fail(h, MH_ERR_DTYPE, "unknown dtype");
  if (n == 0) return MH_OK;
  if (!x || !y) return fail(h, MH_ERR_ARG, "null vector");
  if (dtype == MH_F64 && (a_im != 0.0 || b_im != 0.0))
    return fail(h, MH_ERR_DTYPE, "complex scalar with F64 data");
  unsigned gb = (unsigned)((n + 255) / 256);
  if (dtype == MH_C64) hipLaunchKernelGGL(k_axpby<true>, dim3(gb), dim3(256), 0, h->stream, n, a_re, a_im, (const double *)x, b_re, b_im, (double *)y);
  else hipLaunchKernelGGL(k_axpby<false>, dim3(gb), dim3(256), 0, h->stream, n, a_re, a_im, (const double *)x, b_re, b_im, (double *)y);
  return check_launch(h, "axpby");
}

int mh_lanczos_eig(mh_handle h, const mh_heff_op *op, const void *x0, int kmax,
                   double btol, int reorth, double *eigval, void *xout, int *info,
                   int *niter) {
  if (!h || !op || !eigval) return MH_ERR_ARG;
  MH_TRY(heff_check(h, *op));
  if (!x0 || !xout) return fail(h, MH_ERR_ARG, "null vector");
  if (kmax < 1 || kmax > 128) return fail(h, MH_ERR_ARG, "kmax must be in [1, 128]");
  Krylov K(h, *op, kmax);
  return lanczos_eig_run(h, K, x0, kmax, btol, reorth, eigval, xout, info, niter);
}

int mh_lanczos_eig_proj(mh_handle h, const mh_heff_op *op, int nproj, const void *const *proj,
                        const void *x0, int kmax, double btol, int reorth, double *eigval,
                        void *xout, int *info, int *niter) {
  if (!h || !op || !eigval) return MH_ERR_ARG;
  MH_TRY(heff_check(h, *op));
  if (!x0 || !xout) return fail(h, MH_ERR_ARG, "null vector");
  if (kmax < 1 || kmax > 128) return fail(h, MH_ERR_ARG, "kmax must be in [1, 128]");
  if (nproj < 0 || nproj > NPMAX) return fail(h, MH_ERR_ARG, "0 to %d projectors", NPMAX);
  if (nproj > 0 && !proj) return fail(h, MH_ERR_ARG, "null projector list");
  for (int l = 0; l < nproj; ++l)
    if (!proj[l]) return fail(h, MH_ERR_ARG, "null projector %d", l);
  Krylov K(h, *op, kmax);
  K.nproj = nproj;
  K.proj_src = proj;
  return lanczos_eig_run(h, K, x0, kmax, btol, reorth, eigval, xout, info, niter);
}

int mh_bs_lanczos_eig(mh_handle h, mh_bs_plan plan, const void *p1, const void *p2,
                      const void *x0, int kmax, double btol, int reorth, double *eigval,
                      void *xout, int *info, int *niter) {
  if (!h || !plan || !eigval) return MH_ERR_ARG;
  if (!x0 || !xout) return fail(h, MH_ERR_ARG, "null vector");
  if (kmax < 1 || kmax > 128) return fail(h, MH_ERR_ARG, "kmax must be in [1, 128]");
  MH_TRY(bs_check_square(h, *plan, p1, p2));
  MH_TRY(bs_prepare(h, *plan));
  Krylov K(h, *plan, p1, p2, kmax);
  return lanczos_eig_run(h, K, x0, kmax, btol, reorth, eigval, xout, info, niter);
}

int mh_lanczos_expm(mh_handle h, const mh_heff_op *op, const void *x0, double dt_re,
                    double dt_im, int kdim, double tol, void *xout, int *info) {
  return krylov_expm(h, op, x0, dt_re, dt_im, kdim, tol, false, xout, info);
}

int mh_arnoldi_expm(mh_handle h, const mh_heff_op *op, const void *x0, double dt_re,
                    double dt_im, int kdim, double tol, void *xout, int *info) {
  return krylov_expm(h, op, x0, dt_re, dt_im, kdim, tol, true, xout, info);
}

}  // extern "C"
