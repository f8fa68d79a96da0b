// One-workgroup SVD of a small matrix on the device: the truncated SVD of the
// dense two-site split (auxiliary/stable_svd.jl:4-18, truncated by
// auxiliary/truncation.jl:27-47) at the bond dimensions where a library
// gesvd is a chain of launches many times longer than its arithmetic (a
// 48 x 48 complex gesvd takes about 4 ms on the device).
//
// Method: one-sided Jacobi (Hestenes; Demmel & Veselic, SIAM J. Matrix Anal.
// Appl. 13 (1992) 1204).  The columns a_p, a_q of A (m x n, m >= n) are
// rotated in pairs until every pair is orthogonal to |a_p^H a_q| <=
// m eps ||a_p|| ||a_q||; the rotations accumulate in V, the column norms are
// the singular values and the normalised columns the left singular vectors.
// For a pair with gamma = a_p^H a_q = |gamma| e (|e| = 1), alpha = |a_p|^2,
// beta = |a_q|^2, d = beta - alpha, the rotation angle theta (|theta| <=
// pi/4, tan 2 theta = 2 |gamma| / d) has
//   r = sqrt(d^2 + 4 |gamma|^2),  c = cos theta = sqrt((1 + |d| / r) / 2),
//   s = sin theta = sign(d) |gamma| / (r c)
// (the same angle as t = tan theta = sign(zeta) / (|zeta| + sqrt(1 +
// zeta^2)), zeta = d / (2 |gamma|), with no cancellation in either form), and
//   a_p <- c a_p - s conj(e) a_q,   a_q <- s e a_p + c a_q
// (the same on the columns of V), which zeroes the new a_p^H a_q.  Pairs run
// in round-robin (tournament) order, one pair per 16-lane row, n/2 disjoint
// pairs per round, both matrices in LDS.  By default the iteration runs on
// R^H of a column-pivoted QR of A (k_svd_small below), which for graded
// spectra takes a third of the sweeps.  One-sided Jacobi resolves small
// singular values to high relative accuracy, so the split's cut is taken on
// values at least as accurate as LAPACK gesdd's.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdlib>

#include "mh_internal.h"

namespace mh {

namespace {

constexpr int SV_THREADS = 1024;
constexpr int SV_WAVES = SV_THREADS / 64;
constexpr int SV_ROW = 16;                     // lanes per column pair (one DPP row)
constexpr int SV_ROWS = SV_THREADS / SV_ROW;   // pairs rotated at once
constexpr int SV_SWEEPS = 60;  // sweeps allowed (convergence is quadratic: ~10)
// entries of LDS slack after the matrices: a row reads whole 16-entry strips
// of a column, at most 16 E - 1 <= m + 30 entries from its start (the
// register counts E below), so past the last column by at most 30
constexpr int SV_SLACK = 32;
// Columns whose squared norm lies below this (A scaled to a largest entry in
// [1/2, 1)) take no rotation: they are zero to far below eps64 ||A||, and the
// bound keeps every quantity of the rotation a normal number.
constexpr double SV_TINY = 1e-120;
constexpr int SV_NMAX = 128;                   // largest n (LDS holds n <= 100)
constexpr double SV_TOL3Z = 1.4901161193847656e-08;  // sqrt(eps64), xLAQP2's switch
// static LDS beside the dynamic matrices (tau, perm, ranks, norms, flags)
constexpr size_t SV_STATIC_LDS = 8192;

__device__ inline double sv_abs2(double a) { return a * a; }
__device__ inline double sv_abs2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ inline double sv_amax(double a) { return fabs(a); }
__device__ inline double sv_amax(double2 a) { return fmax(fabs(a.x), fabs(a.y)); }
// conj(a) b
__device__ inline double2 sv_cmul(double a, double b) { return make_double2(a * b, 0.0); }
__device__ inline double2 sv_cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
// c a - w b and u a + c b for the rotation (w = s conj(e), u = s e)
__device__ inline double sv_rot1(double c, double a, double w, double b) { return c * a - w * b; }
__device__ inline double2 sv_rot1(double c, double2 a, double2 w, double2 b) {
  return make_double2(c * a.x - (w.x * b.x - w.y * b.y), c * a.y - (w.x * b.y + w.y * b.x));
}
__device__ inline double sv_rot2(double u, double a, double c, double b) { return u * a + c * b; }
__device__ inline double2 sv_rot2(double2 u, double2 a, double c, double2 b) {
  return make_double2(u.x * a.x - u.y * a.y + c * b.x, u.x * a.y + u.y * a.x + c * b.y);
}
template <class T> __device__ inline T sv_conj(T a);
template <> __device__ inline double sv_conj<double>(double a) { return a; }
template <> __device__ inline double2 sv_conj<double2>(double2 a) { return make_double2(a.x, -a.y); }
template <class T> __device__ inline T sv_scale(T a, double s);
template <> __device__ inline double sv_scale<double>(double a, double s) { return a * s; }
template <> __device__ inline double2 sv_scale<double2>(double2 a, double s) {
  return make_double2(a.x * s, a.y * s);
}
template <class T> __device__ inline T sv_unit(bool one);
template <> __device__ inline double sv_unit<double>(bool one) { return one ? 1.0 : 0.0; }
template <> __device__ inline double2 sv_unit<double2>(bool one) {
  return make_double2(one ? 1.0 : 0.0, 0.0);
}

// 1/sqrt(x) and 1/x: the hardware estimate and one Newton step (relative
// error near 1e-14; x normal and positive)
__device__ inline double sv_rsq(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  return y * fma(-0.5 * x * y, y, 1.5);
}
__device__ inline double sv_rcp(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  return y * fma(-x, y, 2.0);
}

// Rotation (c, w = s conj(e), u = s e) of a pair from alpha, beta and gamma
// = gr + i gi (g2 = |gamma|^2 > 0), in the r / c / s form above.  The angle
// carries the estimates' error (near 1e-14 relative), which only moves the
// next residual a_p^H a_q by as much; the SVD's accuracy rests on the
// rotation being unitary, so a final Newton step on 1 / sqrt(c^2 + s^2) (and
// on 1 / |e| for complex e) makes it unitary to rounding.
template <class T> struct SvRot;
template <> struct SvRot<double> {
  double c, w, u;
  __device__ SvRot(double al, double be, double gr, double, double g2) {
    const double d = be - al;
    const double q = fma(d, d, 4.0 * g2);
    const double rq = sv_rsq(q);               // 1 / r
    const double y = fma(0.5 * fabs(d), rq, 0.5);
    const double ry = sv_rsq(y);
    double cc = y * ry;                        // sqrt(y)
    const double off = g2 * sv_rsq(g2);       // |gamma|
    double ss = copysign(off * rq * ry, d);    // |gamma| / (r c), signed
    const double f = fma(-0.5, fma(cc, cc, ss * ss), 1.5);
    cc *= f;
    ss *= f;
    c = cc;
    w = u = gr < 0.0 ? -ss : ss;               // s e, e = sign(gamma) exactly
  }
};
template <> struct SvRot<double2> {
  double c;
  double2 w, u;
  __device__ SvRot(double al, double be, double gr, double gi, double g2) {
    const double d = be - al;
    const double q = fma(d, d, 4.0 * g2);
    const double rq = sv_rsq(q);
    const double y = fma(0.5 * fabs(d), rq, 0.5);
    const double ry = sv_rsq(y);
    double cc = y * ry;
    const double rg = sv_rsq(g2);
    const double off = g2 * rg;
    double ss = copysign(off * rq * ry, d);
    const double f = fma(-0.5, fma(cc, cc, ss * ss), 1.5);
    cc *= f;
    ss *= f;
    double er = gr * rg, ei = gi * rg;
    const double fe = fma(-0.5, fma(er, er, ei * ei), 1.5);
    er *= fe;
    ei *= fe;
    c = cc;
    w = make_double2(ss * er, -ss * ei);
    u = make_double2(ss * er, ss * ei);
  }
};

// sum over the 64 lanes, returned to every lane (fixed butterfly order)
__device__ inline double sv_wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline double sv_wmax(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// a double moved between the lanes of each 16-lane row by one DPP move of
// each 32-bit half
template <int CTRL>
__device__ inline double sv_dpp(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sums of four values over each 16-lane row, returned to every lane of the
// row: pairs and quads by quad permutations, then the two quads of each half
// row (half mirror) and the two halves (mirror).  Each step adds two
// partial sums that the partner lane adds in the other order, so every lane
// of a row ends with the same bits (its rotation is the same on all lanes).
__device__ inline void sv_rsum16x4(double &a, double &b, double &c, double &d) {
  a += sv_dpp<0xB1>(a); b += sv_dpp<0xB1>(b); c += sv_dpp<0xB1>(c); d += sv_dpp<0xB1>(d);
  a += sv_dpp<0x4E>(a); b += sv_dpp<0x4E>(b); c += sv_dpp<0x4E>(c); d += sv_dpp<0x4E>(d);
  a += sv_dpp<0x141>(a); b += sv_dpp<0x141>(b); c += sv_dpp<0x141>(c); d += sv_dpp<0x141>(d);
  a += sv_dpp<0x140>(a); b += sv_dpp<0x140>(b); c += sv_dpp<0x140>(c); d += sv_dpp<0x140>(d);
}

// LDS: A (n columns of m), V (n columns of n), SV_SLACK entries, the column
// norms (n); the values a strip reads past a column's end are not used
template <class T>
__host__ __device__ inline size_t sv_lds_bytes(int m, int n) {
  return ((size_t)n * m + (size_t)n * n + SV_SLACK) * sizeof(T) + (size_t)n * sizeof(double);
}

// position -> column of round r of the tournament over N (even) columns:
// position 0 stays, the others rotate (pos - 1 + r < 2 (N - 1))
__device__ inline int sv_pos(int pos, int r, int N) {
  if (pos == 0) return 0;
  int x = pos - 1 + r;
  if (x >= N - 1) x -= N - 1;
  return x + 1;
}

// complex helpers on T (double: the imaginary part is zero)
__device__ inline double sv_re(double a) { return a; }
__device__ inline double sv_re(double2 a) { return a.x; }
__device__ inline double sv_im(double) { return 0.0; }
__device__ inline double sv_im(double2 a) { return a.y; }
template <class T> __device__ inline T sv_cplx(double re, double im);
template <> __device__ inline double sv_cplx<double>(double re, double) { return re; }
template <> __device__ inline double2 sv_cplx<double2>(double re, double im) {
  return make_double2(re, im);
}
__device__ inline double sv_mul(double a, double b) { return a * b; }
__device__ inline double2 sv_mul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double sv_sub(double a, double b) { return a - b; }
__device__ inline double2 sv_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// sums of two values over each 16-lane row (the order of sv_rsum16x4)
__device__ inline void sv_rsum16x2(double &a, double &b) {
  a += sv_dpp<0xB1>(a); b += sv_dpp<0xB1>(b);
  a += sv_dpp<0x4E>(a); b += sv_dpp<0x4E>(b);
  a += sv_dpp<0x141>(a); b += sv_dpp<0x141>(b);
  a += sv_dpp<0x140>(a); b += sv_dpp<0x140>(b);
}

// entry i of the Householder vector stored below the diagonal of column j
// (v_j = 1 implicit), as a value
template <class T>
__device__ __forceinline__ T sv_vref(const T *v, int i, int j) {
  T a = v[i];
  if (i == j) a = sv_unit<T>(true);
  return a;
}

// Cyclic one-sided Jacobi on the n columns (rows entries each, leading
// dimension rows) at sA, rotations accumulated into sV (n x n, set to the
// identity here).  One workgroup, 64 rows of 16 lanes; a row takes one
// column pair of the round.  E > 0: each lane keeps its E entries of a_p and
// a_q (rows l, l + 16, ...; rows <= 16 E) in registers from the products to
// the rotation; E = 0 reads them from LDS twice (any number of rows).
// Returns the sweeps run (the last one rotated nothing), or SV_SWEEPS + 1
// when the limit was reached.
template <class T, int E>
__device__ __forceinline__ int sv_jacobi(T *sA, int rows, T *sV, int n, int *rot) {
  const int tid = threadIdx.x, row = tid >> 4, l16 = tid & 15;
  for (int e = tid; e < n * n; e += SV_THREADS) {
    const int j = e / n, i = e - j * n;
    sV[e] = sv_unit<T>(i == j);
  }
  const int N = n + (n & 1);  // an odd n pairs with an empty column
  const double tol = (double)rows * DBL_EPSILON, tol2 = tol * tol;
  const T zero = sv_unit<T>(false);
  for (int sweep = 0; sweep < SV_SWEEPS; ++sweep) {
    if (tid == 0) *rot = 0;
    __syncthreads();
    for (int r = 0; r < N - 1; ++r) {
      for (int k = row; k < N / 2; k += SV_ROWS) {
        int p = sv_pos(k, r, N), q = sv_pos(N - 1 - k, r, N);
        if (p > q) {
          const int x = p;
          p = q;
          q = x;
        }
        if (q >= n) continue;  // the whole row takes this branch
        T *ap = sA + (size_t)p * rows, *aq = sA + (size_t)q * rows;
        double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
        T ra[E > 0 ? E : 1], rb[E > 0 ? E : 1];
        if constexpr (E > 0) {
#pragma unroll
          for (int j = 0; j < E; ++j) {
            const int i = l16 + SV_ROW * j;
            // whole strips: reads past the column stay inside the allocation
            ra[j] = ap[i];
            rb[j] = aq[i];
            if (i >= rows) {
              ra[j] = zero;
              rb[j] = zero;
            }
            al += sv_abs2(ra[j]);
            be += sv_abs2(rb[j]);
            const double2 g = sv_cmul(ra[j], rb[j]);
            gr += g.x;
            gi += g.y;
          }
        } else {
          for (int i = l16; i < rows; i += SV_ROW) {
            const T a = ap[i], b = aq[i];
            al += sv_abs2(a);
            be += sv_abs2(b);
            const double2 g = sv_cmul(a, b);
            gr += g.x;
            gi += g.y;
          }
        }
        sv_rsum16x4(al, be, gr, gi);
        const double g2 = gr * gr + gi * gi;
        if (!(al > SV_TINY && be > SV_TINY && g2 > tol2 * al * be)) continue;
        const SvRot<T> R(al, be, gr, gi, g2);
        if constexpr (E > 0) {
#pragma unroll
          for (int j = 0; j < E; ++j) {
            const int i = l16 + SV_ROW * j;
            if (i < rows) {
              ap[i] = sv_rot1(R.c, ra[j], R.w, rb[j]);
              aq[i] = sv_rot2(R.u, ra[j], R.c, rb[j]);
            }
          }
        } else {
          for (int i = l16; i < rows; i += SV_ROW) {
            const T a = ap[i], b = aq[i];
            ap[i] = sv_rot1(R.c, a, R.w, b);
            aq[i] = sv_rot2(R.u, a, R.c, b);
          }
        }
        T *vp = sV + (size_t)p * n, *vq = sV + (size_t)q * n;
        for (int i = l16; i < n; i += SV_ROW) {
          const T a = vp[i], b = vq[i];
          vp[i] = sv_rot1(R.c, a, R.w, b);
          vq[i] = sv_rot2(R.u, a, R.c, b);
        }
        if (l16 == 0) *rot = 1;
      }
      __syncthreads();
    }
    const int any = *rot;
    __syncthreads();  // every thread has read rot before it is cleared
    if (!any) return sweep + 1;
  }
  return SV_SWEEPS + 1;
}

// Column norms of the n columns (rows entries, leading dimension rows) into
// sig, then each column's rank in descending order (ties by index) into rk.
template <class T>
__device__ __forceinline__ void sv_rank(const T *sA, int rows, int n, double *sig, int *rk) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int j = wv; j < n; j += SV_WAVES) {
    double s2 = 0.0;
    for (int i = lane; i < rows; i += 64) s2 += sv_abs2(sA[(size_t)j * rows + i]);
    s2 = sv_wsum(s2);
    if (lane == 0) sig[j] = sqrt(s2);
  }
  __syncthreads();
  for (int j = wv; j < n; j += SV_WAVES) {
    const double sj = sig[j];
    double c = 0.0;
    for (int i = lane; i < n; i += 64) {
      const double si = sig[i];
      c += (si > sj || (si == sj && i < j)) ? 1.0 : 0.0;
    }
    c = sv_wsum(c);
    if (lane == 0) rk[j] = (int)c;
  }
  __syncthreads();
}

// Thin SVD of one m x n matrix (m >= n) in one workgroup.
//
// A is scaled by a power of two on entry (its largest entry into [1/2, 1):
// exact), S by its inverse at the end.  With precond = 1 the Jacobi
// iteration runs on X = R^H of the column-pivoted QR A P = Q R (Householder,
// the steps of LAPACK xGEQP3 / xLAQP2 with their norm-downdating safeguard),
// which for graded matrices takes a third of the sweeps (Drmac & Veselic,
// SIAM J. Matrix Anal. Appl. 29 (2008) 1322): X V_x = Y (orthogonal
// columns, norms sigma) gives R = V_x Sigma U_y^H, so A = (Q V_x) Sigma
// (P U_y)^H with U_y = Y Sigma^-1.  The left vectors Q [V_x; 0] are
// orthonormal by construction; the right ones are normalised columns of Y,
// so a column left below SV_TINY (an input rank-deficient to far below eps)
// would have none: then the plain iteration on A runs instead (V
// accumulated, the zero columns' U left zero).  The Householder vectors wait
// in U's storage while the Jacobi iteration has the LDS.
template <class T, int E>
__global__ __launch_bounds__(SV_THREADS) void k_svd_small(int m, int n, const T *__restrict__ Ain,
                                                          int lda, T *__restrict__ U,
                                                          double *__restrict__ S,
                                                          T *__restrict__ V,
                                                          int *__restrict__ sweeps, int precond) {
  extern __shared__ double2 sv_lds[];
  T *sA = (T *)sv_lds;
  T *sV = sA + (size_t)n * m;
  double *sig = (double *)(sV + (size_t)n * n + SV_SLACK);
  __shared__ int rot, flag, piv;
  __shared__ double wmax[SV_WAVES];
  __shared__ T tau[SV_NMAX];
  __shared__ int perm[SV_NMAX], rk[SV_NMAX];
  __shared__ double vn1[SV_NMAX], vn2[SV_NMAX];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int row = tid >> 4, l16 = tid & 15;
  const T zero = sv_unit<T>(false);
  double amax = 0.0;
  for (int e = tid; e < n * m; e += SV_THREADS) {
    const int j = e / m, i = e - j * m;
    const T a = Ain[i + (size_t)lda * j];
    sA[e] = a;
    amax = fmax(amax, sv_amax(a));
  }
  if (tid < SV_SLACK) sV[(size_t)n * n + tid] = zero;
  amax = sv_wmax(amax);
  if (lane == 0) wmax[wv] = amax;
  __syncthreads();
  amax = 0.0;
  for (int w = 0; w < SV_WAVES; ++w) amax = fmax(amax, wmax[w]);
  int ex = 0;
  if (amax > 0.0 && amax <= DBL_MAX) (void)frexp(amax, &ex);
  const double scale = ldexp(1.0, -ex);  // largest entry -> [1/2, 1)
  for (int e = tid; e < n * m; e += SV_THREADS) sA[e] = sv_scale(sA[e], scale);
  __syncthreads();
  int total = 0;
  bool plain = !precond;
  if (precond) {
    // ---- column-pivoted Householder QR, in place (R above the diagonal,
    // the vectors v (v_j = 1 implicit) below it, tau apart) ----
    for (int k = row; k < n; k += SV_ROWS) {
      double s2 = 0.0, z = 0.0;
      for (int i = l16; i < m; i += SV_ROW) s2 += sv_abs2(sA[(size_t)k * m + i]);
      sv_rsum16x2(s2, z);
      if (l16 == 0) {
        vn1[k] = vn2[k] = sqrt(s2);
        perm[k] = k;
      }
    }
    if (tid == 0) flag = 0;
    __syncthreads();
    for (int j = 0; j < n; ++j) {
      if (wv == 0) {
        // pivot: the largest remaining partial norm, first index on ties
        double best = -1.0;
        int bi = n;
        for (int k = j + lane; k < n; k += 64) {
          const double v = vn1[k];
          if (v > best) {
            best = v;
            bi = k;
          }
        }
        for (int o = 32; o > 0; o >>= 1) {
          const double ov = __shfl_xor(best, o, 64);
          const int oi = __shfl_xor(bi, o, 64);
          if (ov > best || (ov == best && oi < bi)) {
            best = ov;
            bi = oi;
          }
        }
        if (best * best < SV_TINY) {
          if (lane == 0) flag = 1;  // rank-deficient far below eps: plain path
        } else {
          if (bi != j) {
            for (int i = lane; i < m; i += 64) {
              const T a = sA[(size_t)j * m + i];
              sA[(size_t)j * m + i] = sA[(size_t)bi * m + i];
              sA[(size_t)bi * m + i] = a;
            }
            if (lane == 0) {
              const double t1 = vn1[j], t2 = vn2[j];
              vn1[j] = vn1[bi];
              vn2[j] = vn2[bi];
              vn1[bi] = t1;
              vn2[bi] = t2;
              const int tp = perm[j];
              perm[j] = perm[bi];
              perm[bi] = tp;
            }
          }
          __threadfence_block();  // the swapped column, read by other lanes below
          __builtin_amdgcn_wave_barrier();
          // Householder vector of column j, rows j .. m-1 (xLARFG without the
          // rescaling loop: the column's norm is at least sqrt(SV_TINY))
          const T alpha = sA[(size_t)j * m + j];
          double xn2 = 0.0;
          for (int i = j + 1 + lane; i < m; i += 64) xn2 += sv_abs2(sA[(size_t)j * m + i]);
          xn2 = sv_wsum(xn2);
          T tj = zero;
          if (xn2 > 0.0) {
            const double beta = -copysign(sqrt(sv_abs2(alpha) + xn2), sv_re(alpha));
            tj = sv_cplx<T>((beta - sv_re(alpha)) / beta, -sv_im(alpha) / beta);
            const T d = sv_cplx<T>(sv_re(alpha) - beta, sv_im(alpha));
            const double d2 = sv_abs2(d);
            const T sc = sv_scale(sv_conj(d), 1.0 / d2);  // 1 / (alpha - beta)
            for (int i = j + 1 + lane; i < m; i += 64)
              sA[(size_t)j * m + i] = sv_mul(sc, sA[(size_t)j * m + i]);
            if (lane == 0) sA[(size_t)j * m + j] = sv_cplx<T>(beta, 0.0);
          }
          if (lane == 0) tau[j] = tj;
        }
      }
      __syncthreads();
      if (flag) break;
      // apply H_j^H = I - conj(tau_j) v v^H to the columns right of j, then
      // downdate their partial norms (xLAQP2)
      const T tj = tau[j];
      const T *v = sA + (size_t)j * m;
      for (int k = j + 1 + row; k < n; k += SV_ROWS) {
        T *ak = sA + (size_t)k * m;
        double ajk2 = 0.0;
        if (sv_abs2(tj) > 0.0) {
          double wr = 0.0, wi = 0.0;
          for (int i = j + l16; i < m; i += SV_ROW) {
            const double2 g = sv_cmul(sv_vref(v, i, j), ak[i]);
            wr += g.x;
            wi += g.y;
          }
          sv_rsum16x2(wr, wi);
          const T f = sv_mul(sv_conj(tj), sv_cplx<T>(wr, wi));
          for (int i = j + l16; i < m; i += SV_ROW) {
            const T a = sv_sub(ak[i], sv_mul(sv_vref(v, i, j), f));
            ak[i] = a;
            if (i == j) ajk2 = sv_abs2(a);
          }
        } else if (l16 == 0) {
          ajk2 = sv_abs2(ak[j]);
        }
        ajk2 = __shfl(ajk2, lane & ~(SV_ROW - 1), 64);  // the row's lane 0 holds it
        const double nv1 = vn1[k], nv2 = vn2[k];
        double nv = nv1;
        bool recompute = false;
        if (nv1 != 0.0) {
          const double t = fmax(0.0, 1.0 - ajk2 / (nv1 * nv1));
          const double r = nv1 / nv2;
          if (t * r * r <= SV_TOL3Z)
            recompute = true;
          else
            nv = nv1 * sqrt(t);
        }
        if (recompute) {
          double s2 = 0.0, z = 0.0;
          for (int i = j + l16; i < m; i += SV_ROW)
            if (i > j) s2 += sv_abs2(ak[i]);
          sv_rsum16x2(s2, z);
          nv = sqrt(s2);
        }
        if (l16 == 0) {
          vn1[k] = nv;
          if (recompute) vn2[k] = nv;
        }
      }
      __syncthreads();
    }
    plain = flag != 0;
  }
  if (!plain) {
    // the Householder vectors and R to U's storage; X = R^H into LDS
    for (int e = tid; e < n * m; e += SV_THREADS) {
      const int j = e / m, i = e - j * m;
      U[i + (size_t)m * j] = sA[e];
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += SV_THREADS) {
      const int k = e / n, i = e - k * n;  // X[i, k] = conj(R[k, i]), i >= k
      sA[e] = i >= k ? sv_conj(U[k + (size_t)m * i]) : zero;
    }
    __syncthreads();
    total = sv_jacobi<T, E>(sA, n, sV, n, &rot);
    sv_rank(sA, n, n, sig, rk);
    if (wv == 0) {
      bool low = false;
      for (int j = lane; j < n; j += 64) low = low || sig[j] * sig[j] < SV_TINY;
      if (__ballot(low) != 0 && lane == 0) flag = 1;
    }
    __syncthreads();
    plain = flag != 0;
    if (!plain) {
      // right vectors: rows of P U_y
      for (int e = tid; e < n * n; e += SV_THREADS) {
        const int j = e / n, i = e - j * n;
        V[perm[i] + (size_t)n * rk[j]] = sv_scale(sA[e], 1.0 / sig[j]);
      }
      __syncthreads();
      // left vectors: Q [V_x; 0], Q = H_0 H_1 ... H_{n-1}
      for (int e = tid; e < n * m; e += SV_THREADS) {
        const int j = e / m, i = e - j * m;
        T z = zero;
        if (i < n) z = sV[i + (size_t)n * j];
        sA[e] = z;
      }
      __syncthreads();
      for (int j = n - 1; j >= 0; --j) {
        const T tj = tau[j];
        if (sv_abs2(tj) > 0.0) {
          const T *v = U + (size_t)m * j;
          for (int c = row; c < n; c += SV_ROWS) {
            T *zc = sA + (size_t)c * m;
            double wr = 0.0, wi = 0.0;
            for (int i = j + l16; i < m; i += SV_ROW) {
              const double2 g = sv_cmul(sv_vref(v, i, j), zc[i]);
              wr += g.x;
              wi += g.y;
            }
            sv_rsum16x2(wr, wi);
            const T f = sv_mul(tj, sv_cplx<T>(wr, wi));
            for (int i = j + l16; i < m; i += SV_ROW) zc[i] = sv_sub(zc[i], sv_mul(sv_vref(v, i, j), f));
          }
          __syncthreads();
        }
      }
      for (int e = tid; e < n * m; e += SV_THREADS) {
        const int j = e / m, i = e - j * m;
        U[i + (size_t)m * rk[j]] = sA[e];
      }
      if (tid < n) S[rk[tid]] = total <= SV_SWEEPS ? ldexp(sig[tid], ex) : NAN;
      if (sweeps && tid == 0) *sweeps = total <= SV_SWEEPS ? total : -1;
      return;
    }
  }
  if (precond) {
    // the plain iteration starts again from A
    __syncthreads();
    for (int e = tid; e < n * m; e += SV_THREADS) {
      const int j = e / m, i = e - j * m;
      sA[e] = sv_scale(Ain[i + (size_t)lda * j], scale);
    }
    __syncthreads();
  }
  const int sw = sv_jacobi<T, E>(sA, m, sV, n, &rot);
  total += sw;
  const bool ok = sw <= SV_SWEEPS;
  sv_rank(sA, m, n, sig, rk);
  for (int e = tid; e < n * m; e += SV_THREADS) {
    const int j = e / m, i = e - j * m;
    const double sj = sig[j];
    U[i + (size_t)m * rk[j]] = sv_scale(sA[e], sj > 0.0 ? 1.0 / sj : 0.0);
  }
  for (int e = tid; e < n * n; e += SV_THREADS) {
    const int j = e / n;
    V[e - j * n + (size_t)n * rk[j]] = sV[e];
  }
  // a sweep limit reached marks the result: every value NaN
  if (tid < n) S[rk[tid]] = ok ? ldexp(sig[tid], ex) : NAN;
  if (sweeps && tid == 0) *sweeps = ok ? total : -1;
}

// the kernel for m rows: entries per lane kept in registers, or none
template <class T>
const void *sv_kernel(int m) {
  if (m <= 16) return (const void *)k_svd_small<T, 1>;
  if (m <= 32) return (const void *)k_svd_small<T, 2>;
  if (m <= 48) return (const void *)k_svd_small<T, 3>;
  if (m <= 64) return (const void *)k_svd_small<T, 4>;
  if (m <= 96) return (const void *)k_svd_small<T, 6>;
  if (m <= 128) return (const void *)k_svd_small<T, 8>;
  return (const void *)k_svd_small<T, 0>;
}

int sv_lds_limit(mh_handle h) {
  int per_block = 0;
  (void)hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device);
  return per_block > 0 ? per_block : 65536;
}

}  // namespace

}  // namespace mh

using namespace mh;

extern "C" int mh_svd_small_fits(mh_handle h, int dtype, int m, int n, int *fits) {
  if (!h || !fits) return MH_ERR_ARG;
  if (dtype != MH_F64 && dtype != MH_C64) return fail(h, MH_ERR_DTYPE, "dtype %d", dtype);
  MH_HIP(h, hipSetDevice(h->device));
  const size_t need = dtype == MH_C64 ? sv_lds_bytes<double2>(m, n) : sv_lds_bytes<double>(m, n);
  *fits = (m >= n && n >= 1 && n <= SV_NMAX && need + SV_STATIC_LDS <= (size_t)sv_lds_limit(h))
              ? 1
              : 0;
  return MH_OK;
}

extern "C" int mh_svd_small(mh_handle h, int dtype, int m, int n, const void *A, int lda, void *U,
                            double *S, void *V, int *sweeps) {
  if (!h) return MH_ERR_ARG;
  if (dtype != MH_F64 && dtype != MH_C64) return fail(h, MH_ERR_DTYPE, "dtype %d", dtype);
  if (!A || !U || !S || !V) return fail(h, MH_ERR_ARG, "null matrix or output buffer");
  if (n < 1 || m < n || lda < m)
    return fail(h, MH_ERR_SHAPE, "needs m >= n >= 1 and lda >= m (m %d, n %d, lda %d)",
                m, n, lda);
  MH_HIP(h, hipSetDevice(h->device));
  const size_t lds = dtype == MH_C64 ? sv_lds_bytes<double2>(m, n) : sv_lds_bytes<double>(m, n);
  if (n > SV_NMAX || lds + SV_STATIC_LDS > (size_t)sv_lds_limit(h))
    return fail(h, MH_ERR_SHAPE, "%d x %d needs %zu bytes of LDS", m, n, lds + SV_STATIC_LDS);
  // MH_SVD_PLAIN=1: Jacobi on A itself, without the pivoted-QR preconditioner
  const char *plain_env = getenv("MH_SVD_PLAIN");
  int precond = (plain_env && atoi(plain_env) == 1) ? 0 : 1;
  const void *fn = dtype == MH_C64 ? sv_kernel<double2>(m) : sv_kernel<double>(m);
  if (lds > 65536)
    MH_HIP(h, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  void *args[] = {&m, &n, &A, &lda, &U, &S, &V, &sweeps, &precond};
  MH_HIP(h, hipLaunchKernel(fn, dim3(1), dim3(SV_THREADS), args, lds, h->stream));
  return check_launch(h, "k_svd_small");
}
