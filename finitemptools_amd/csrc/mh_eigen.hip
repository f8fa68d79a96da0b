// Batched FP64 symmetric eigensolver for the sector density matrices of the
// block-sparse two-site split (the FP64 step that replaces the per-sector
// truncated SVD of auxiliary/stable_svd.jl:4-18 on graded tensors).
//
//   mh_sytrd_batched : Householder tridiagonalisation of every matrix of a
//                      batch in ONE cooperative launch.  Matrix m gets G_m
//                      workgroups that own its columns cyclically and
//                      exchange one vector and one column per step, instead
//                      of the handful of launches a library reduction spends
//                      per column.  Output follows
//                      LAPACK dsytrd (uplo = 'L'): D, E, tau, and reflector
//                      m's vector v_i (v_i[i+1] = 1 implicit) in A[i+2:, i].
//   mh_stedc_batched : eigen-decomposition of the tridiagonals by rocSOLVER's
//                      divide and conquer (dstedc), one call per matrix, the
//                      calls spread over side streams.
//   mh_ormtr_batched : selected eigenvector columns back-transformed,
//                      z <- H_0 H_1 ... H_{n-2} z, one wavefront per column.
//
// Algorithm of the tridiagonalisation (Golub & Van Loan, Alg. 8.3.1; the
// LAPACK dsytd2 recurrence): with the column i of A^(i) split as
// (alpha, x), H_i = I - tau v v^T maps it to (beta, 0); p = tau A v,
// w = p - (tau/2)(p.v) v, A^(i+1) = A^(i) - v w^T - w v^T.  Each workgroup
// applies the rank-2 update of step i-1 to its own columns lazily, in the
// same pass that forms its entries of p_i = tau_i A^(i) v_i (A symmetric, so
// row j . v = column j . v); p_i and the next column travel between the
// workgroups of one matrix through double-buffered global rows.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <thread>
#include <vector>

#include "mh_internal.h"

namespace mh {

namespace {

constexpr int SY_THREADS = 256;
constexpr int SY_WAVES = SY_THREADS / 64;
constexpr int SY_MAXT = 48;          // matrices per launch (kernel-argument table)
constexpr int SY_MAXN = 2048;        // largest order (LDS: 3 n doubles per workgroup)
constexpr int SY_COLS_PER_WG = 16;   // target columns per workgroup
constexpr int SY_MAX_G = 64;         // workgroups per matrix
constexpr unsigned SY_SPIN_LIMIT = 1u << 22;

struct SyTask {
  int64_t a_off;  // element offset of the matrix in A (column-major, ld = n)
  int64_t v_off;  // element offset of D / E / tau and of the exchange rows
  int n;          // order
  int G;          // workgroups of this matrix
  int g0;         // first workgroup of this matrix in the grid
  int pad;
};

struct SyTable {
  int nt;
  int pad;
  int64_t xrow;  // stride between the two halves of the exchange buffers
  SyTask t[SY_MAXT];
};

// All lanes of the block get the sum; the order of additions is fixed, so
// every workgroup of a matrix computes bit-identical values.
__device__ inline double block_sum(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < SY_WAVES; ++i) s += red[i];
  return s;
}

__device__ inline double xload(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void xstore(double *p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Barrier of the G workgroups of one matrix: a monotone counter, target
// G * (number of barriers passed).  The wait is bounded: past the spin limit
// the workgroup raises *err and leaves, and once *err is set every later
// wait returns at once, so the grid always drains.
__device__ inline void group_barrier(unsigned *cnt, unsigned target, int *err) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    atomicAdd(cnt, 1u);
    unsigned spins = 0;
    // relaxed polls (no cache invalidation per poll); one acquire below
    while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      if (++spins > SY_SPIN_LIMIT) {
        atomicOr(err, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __threadfence();
  }
  __syncthreads();
}

__global__ __launch_bounds__(SY_THREADS) void k_sytrd_group(SyTable tab, double *__restrict__ A,
                                                            double *__restrict__ D,
                                                            double *__restrict__ E,
                                                            double *__restrict__ tau,
                                                            double *__restrict__ P,
                                                            double *__restrict__ BC,
                                                            unsigned *__restrict__ cnt,
                                                            int *__restrict__ err) {
  extern __shared__ double sm[];
  __shared__ double red[SY_WAVES];
  int t = 0;
  while (t + 1 < tab.nt && (int)blockIdx.x >= tab.t[t + 1].g0) ++t;
  const SyTask tk = tab.t[t];
  const int n = tk.n, G = tk.G, g = (int)blockIdx.x - tk.g0;
  double *a = A + tk.a_off;
  double *vp = sm, *wp = sm + n, *vc = sm + 2 * n;
  double *const P0 = P + tk.v_off, *const P1 = P + tab.xrow + tk.v_off;
  double *const B0 = BC + tk.v_off, *const B1 = BC + tab.xrow + tk.v_off;
  unsigned *c = cnt + t;
  unsigned nbar = 0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;

  for (int r = threadIdx.x; r < 3 * n; r += SY_THREADS) sm[r] = 0.0;
  if (g == 0)
    for (int r = threadIdx.x; r < n; r += SY_THREADS) xstore(B0 + r, a[r]);
  group_barrier(c, (unsigned)G * ++nbar, err);

  double tau_prev = 0.0;
  for (int i = 0; i < n; ++i) {
    const double *bc = (i & 1) ? B1 : B0;
    // Phase A (every workgroup, redundantly): w_{i-1}, column i of A^(i),
    // and the reflector v_i, tau_i, beta_i.
    if (i > 0) {
      const double *pp = ((i - 1) & 1) ? P1 : P0;
      double s = 0.0;
      for (int r = i + threadIdx.x; r < n; r += SY_THREADS) s += xload(pp + r) * vp[r];
      const double h = 0.5 * tau_prev * block_sum(s, red);
      for (int r = i + threadIdx.x; r < n; r += SY_THREADS) wp[r] = xload(pp + r) - h * vp[r];
      __syncthreads();
    }
    const double vpi = vp[i], wpi = wp[i];
    double s2 = 0.0;
    for (int r = i + threadIdx.x; r < n; r += SY_THREADS) {
      const double x = xload(bc + r) - vp[r] * wpi - wp[r] * vpi;
      vc[r] = x;
      if (r >= i + 2) s2 += x * x;
    }
    const double xn2 = block_sum(s2, red);  // its barriers publish vc
    const double di = vc[i];
    if (i == n - 1) {
      if (g == 0 && threadIdx.x == 0) {
        D[tk.v_off + i] = di;
        E[tk.v_off + i] = 0.0;
        tau[tk.v_off + i] = 0.0;
      }
      if (i % G == g && threadIdx.x == 0) a[i + (int64_t)i * n] = di;
      break;
    }
    const double alpha = vc[i + 1];
    double beta = alpha, tu = 0.0, scale = 1.0;
    if (xn2 > 0.0) {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      tu = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    __syncthreads();
    for (int r = i + 2 + threadIdx.x; r < n; r += SY_THREADS) vc[r] *= scale;
    if (threadIdx.x == 0) {
      vc[i] = 0.0;
      vc[i + 1] = 1.0;
    }
    __syncthreads();
    if (i % G == g) {  // owner of column i stores the reflector (LAPACK layout)
      double *col = a + (int64_t)i * n;
      for (int r = i + 2 + threadIdx.x; r < n; r += SY_THREADS) col[r] = vc[r];
      if (threadIdx.x == 0) {
        col[i] = di;
        col[i + 1] = beta;
      }
    }
    if (g == 0 && threadIdx.x == 0) {
      D[tk.v_off + i] = di;
      E[tk.v_off + i] = beta;
      tau[tk.v_off + i] = tu;
    }
    // Phase B: own columns j >= i+1 (j = g mod G), one wavefront per column:
    // apply step i-1's rank-2 update, form p_i[j]; the owner of column i+1
    // also publishes that column of A^(i) for step i+1.
    double *pc = (i & 1) ? P1 : P0;
    double *bn = ((i + 1) & 1) ? B1 : B0;
    const int j0 = i + 1 + (((g - (i + 1)) % G) + G) % G;
    for (int j = j0 + wv * G; j < n; j += SY_WAVES * G) {
      double *col = a + (int64_t)j * n;
      const double vpj = vp[j], wpj = wp[j];
      const bool pub = (j == i + 1);
      double acc = 0.0;
      for (int r = i + 1 + lane; r < n; r += 64) {
        const double x = col[r] - vp[r] * wpj - wp[r] * vpj;
        col[r] = x;
        if (pub) xstore(bn + r, x);
        acc += x * vc[r];
      }
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
      if (lane == 0) xstore(pc + j, tu * acc);
    }
    double *sw = vp;  // v_i becomes the lazy update's vector of the next step
    vp = vc;
    vc = sw;
    tau_prev = tu;
    group_barrier(c, (unsigned)G * ++nbar, err);
  }
}

// The same reduction with every workgroup's columns resident in LDS for the
// whole launch (orders up to SL_MAXN).  The step is the one above; what
// changes is where the data lives and how the workgroups meet:
//   - a workgroup keeps its columns (cyclic, j = g + q G) in LDS, so the lazy
//     rank-2 update and the dot products of a step touch no global memory;
//   - wavefront 0 forms w_{i-1}, the column i of A^(i) and v_i in registers
//     (lane l holds rows l + 64 k) and hands them to the other waves through
//     LDS; only p and the next column travel through global memory;
//   - every step has exchange slots of its own (p_i and the column i+1 of
//     A^(i+1)), filled with an "empty" bit pattern before the launch; a
//     reader polls the rows it needs until none is empty.  Data arrival is
//     the only synchronisation between workgroups: no barrier, no flag, and
//     a writer never waits for its stores to be acknowledged.
constexpr int SL_KRMAX = 18;              // rows per lane, largest template
constexpr int SL_MAXN = 64 * SL_KRMAX;    // largest order of the LDS form
constexpr int SL_THREADS = 512;           // 8 wavefronts share a workgroup's columns
constexpr int SL_WAVES = SL_THREADS / 64;
// An exchange slot not yet written: a signalling NaN.  Arithmetic only ever
// produces the canonical quiet NaN, and the one raw copy (column 0) is
// canonicalised, so no published value has this pattern.
constexpr long long SL_EMPTY = 0x7FF4DEAD7FF4DEADLL;

struct SlTask {
  int64_t a_off, v_off, x_off;  // matrix, D/E/tau, exchange slots (2 n xs doubles)
  int n, G, g0, ncl;            // order, workgroups, first workgroup, LDS columns each
  int xs, pad;                  // exchange slot length: order padded to a multiple of 64
};

// A double moved between lanes by one DPP move of each 32-bit half.
template <int CTRL, int ROWS>
__device__ inline double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// value of v in lane src (uniform src), read into scalar registers
__device__ inline double lane_value(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes of a wavefront (every lane active), returned to all:
// pairs and quads by quad permutations, rows of 16 by two rotations, then
// row 0 + 1 and row 2 + 3 by the row-15 broadcast and the total into row 3 by
// the row-31 broadcast; lane 63 holds it.  The order of additions is fixed,
// so every workgroup that sums the same data gets the same bits.
__device__ inline double wave_sum64(double v) {
  v += dpp_f64<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v += dpp_f64<0x124, 0xF>(v);  // row_ror:4
  v += dpp_f64<0x128, 0xF>(v);  // row_ror:8
  v += dpp_f64<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_f64<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  return lane_value(v, 63);
}

// wave_sum64 of N values at once (the same additions per value, the N
// reduction chains interleaved)
template <int N>
__device__ inline void wave_sum64_n(double (&v)[N]) {
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] += dpp_f64<0xB1, 0xF>(v[q]);
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] += dpp_f64<0x4E, 0xF>(v[q]);
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] += dpp_f64<0x124, 0xF>(v[q]);
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] += dpp_f64<0x128, 0xF>(v[q]);
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] += dpp_f64<0x142, 0xA>(v[q]);
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] += dpp_f64<0x143, 0xC>(v[q]);
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = lane_value(v[q], 63);
}

// value at row idx of a lane-distributed vector (lane idx % 64, slot idx / 64)
template <int KR>
__device__ inline double row_value(const double (&x)[KR], int idx) {
  const int kk = idx >> 6;
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < KR; ++k)
    if (k == kk) t = x[k];
  return lane_value(t, idx & 63);
}

__device__ inline bool sl_empty(double v) { return __double_as_longlong(v) == SL_EMPTY; }

// A wavefront's read of step i's inputs: its rows (slots wv + 8 t) of
// p_{i-1} (if with_p) and of the column i of A^(i), polled until no needed
// row is empty.  Slots below k0 hold only finished rows and slots from kr on
// lie past the order, so neither is read; every load stays inside the
// matrix's slots.  Rows outside [i, n) come back as zero.  The wait is
// bounded: past the spin limit the error flag is raised, and a raised error
// flag ends every wait.
template <int T>
__device__ inline void sl_poll(const double *pp, const double *bc, bool with_p, int i, int n,
                               int kr, int wv, int lane, double (&wp)[T], double (&bv)[T],
                               int *err, unsigned spin_limit) {
  const int k0 = i >> 6;
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const int sl = wv + SL_WAVES * t;
    wp[t] = bv[t] = 0.0;
    if (sl >= k0 && sl < kr) {
      const int r = lane + 64 * sl;
      if (with_p) wp[t] = xload(pp + r);
      bv[t] = xload(bc + r);
    }
  }
  for (unsigned spins = 0;; ++spins) {
    unsigned miss = 0;
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const int sl = wv + SL_WAVES * t;
      if (sl >= k0 && sl < kr) {
        const int r = lane + 64 * sl;
        const bool need = r >= i && r < n;
        if (need && with_p && sl_empty(wp[t])) miss |= 1u << (2 * t);
        if (need && sl_empty(bv[t])) miss |= 2u << (2 * t);
      }
    }
    if (__ballot(miss != 0) == 0) break;
    const int e = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (e != 0 || spins >= spin_limit) {
      if (e == 0 && lane == 0) atomicOr(err, 1);
      break;
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const int sl = wv + SL_WAVES * t;
      if (sl >= k0 && sl < kr) {
        const int r = lane + 64 * sl;
        if (miss & (1u << (2 * t))) wp[t] = xload(pp + r);
        if (miss & (2u << (2 * t))) bv[t] = xload(bc + r);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const int r = lane + 64 * (wv + SL_WAVES * t);
    const bool in = r >= i && r < n;
    wp[t] = in ? wp[t] : 0.0;
    bv[t] = in ? bv[t] : 0.0;
  }
}

// the element of a row-distributed vector at slot t (uniform t < T)
template <int T>
__device__ inline double slot_value(const double (&x)[T], int t) {
  double v = x[0];
#pragma unroll
  for (int q = 1; q < T; ++q)
    if (q == t) v = x[q];
  return v;
}

// KR: row slots per lane of the launch's largest order; a matrix of order n
// uses its first kr = ceil(n / 64) slots.  A workgroup's ncl columns: the
// last 8 CW (the longest-lived) in registers, wavefront w holding columns
// nlds + w + 8 m (m < CW) with row l + 64 k of column m in lane l, slot k;
// the first nlds = ncl - 8 CW (if any) in LDS at stride 64 kr.
//
// The vector work of a step (forming w_{i-1}, the column i of A^(i), v_i and
// tau_i) is spread over the eight wavefronts by rows: wavefront w holds the
// row slots w + 8 t (lane l: row l + 64 slot), reads its rows of the exchange
// slots itself and keeps its rows of v in registers; the two dot products
// meet in LDS at the workgroup barriers B1 and B2, in a fixed order, so every
// workgroup of the matrix computes the same bits.  After B3 the columns are
// processed one per wavefront: step i-1's rank-2 update and the dot product
// with v_i in one pass.
template <int KR, int CW>
__global__ __launch_bounds__(SL_THREADS) void k_sytrd_lds(const SlTask *__restrict__ tabv, int nt,
                                                          double *__restrict__ A,
                                                          double *__restrict__ D,
                                                          double *__restrict__ E,
                                                          double *__restrict__ tau,
                                                          double *__restrict__ X,
                                                          int *__restrict__ err,
                                                          unsigned spin_limit, int drop_step) {
  constexpr int T = (KR + SL_WAVES - 1) / SL_WAVES;  // row slots per lane of a wavefront
  extern __shared__ double sm[];
  __shared__ double red[2][SL_WAVES];
  __shared__ double sc[4];  // v_{i-1}[i], p_{i-1}[i], x[i], x[i+1]
  int lo = 0, hi = nt - 1;
  while (lo < hi) {
    const int md = (lo + hi + 1) >> 1;
    if (tabv[md].g0 <= (int)blockIdx.x) lo = md; else hi = md - 1;
  }
  const SlTask tk = tabv[lo];
  const int n = tk.n, G = tk.G, g = (int)blockIdx.x - tk.g0, ncl = tk.ncl, xs = tk.xs;
  const int kr = xs >> 6;  // slots in use: rows padded to xs = 64 kr
  const int np = xs;
  const int nlds = std::max(0, ncl - SL_WAVES * CW);  // columns kept in LDS
  double *a = A + tk.a_off;
  // LDS: nlds columns of np rows (rows n..np-1 zero), w_{i-1}, and a ring of
  // two v vectors (v_i at ring slot i & 1)
  double *cols = sm;
  double *sW = sm + (size_t)nlds * np;
  double *sR = sW + np;
  // exchange slots: column i of A^(i) at xb + 2 i xs, p_i at xb + (2 i + 1) xs
  double *const xb = X + tk.x_off;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));

  for (int q = 0; q < nlds; ++q) {
    const int j = g + q * G;
    if (j >= n) break;
    for (int r = threadIdx.x; r < np; r += SL_THREADS)
      cols[(size_t)q * np + r] = (r < n) ? a[r + (int64_t)j * n] : 0.0;
  }
  double colr[CW > 0 ? CW : 1][KR];  // (with CW = 0 unused)
#pragma unroll
  for (int m = 0; m < CW; ++m) {
    const int q = nlds + wv + SL_WAVES * m;
    const int j = g + q * G;
    const bool live = q < ncl && j < n;
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      const int r = lane + 64 * k;
      colr[m][k] = (live && r < n) ? a[r + (int64_t)j * n] : 0.0;
    }
  }
  for (int r = threadIdx.x; r < 3 * np; r += SL_THREADS) sW[r] = 0.0;
  if (g == 0)
    for (int r = threadIdx.x; r < n; r += SL_THREADS)
      xstore(xb + r, __builtin_canonicalize(a[r]));
  __syncthreads();

  double vpr[T];  // own rows of v_{i-1}
#pragma unroll
  for (int t = 0; t < T; ++t) vpr[t] = 0.0;
  double tau_prev = 0.0;
  for (int i = 0; i < n; ++i) {
    const bool last = (i == n - 1);
    const int k0 = i >> 6;  // slots below k0 hold only finished rows
    double *const vprev = sR + ((i + 1) & 1) * np;  // v_{i-1}
    double *const vcur = sR + (i & 1) * np;         // v_i
    const int si = i >> 6, si1 = (i + 1) >> 6;      // slots of rows i and i + 1
    const bool row_i = (wv == (si & 7)) && lane == (i & 63);
    const bool row_i1 = (wv == (si1 & 7)) && lane == ((i + 1) & 63);
    double wpr[T], bcr[T];
    {
      const double *bc = xb + 2 * (int64_t)i * xs;
      sl_poll<T>(bc - xs, bc, i > 0, i, n, kr, wv, lane, wpr, bcr, err, spin_limit);
    }
    {
      double s = 0.0;
#pragma unroll
      for (int t = 0; t < T; ++t) s += wpr[t] * vpr[t];
      s = wave_sum64(s);
      if (lane == 0) red[0][wv] = s;
      if (row_i) {
        sc[0] = slot_value(vpr, si >> 3);
        sc[1] = slot_value(wpr, si >> 3);
      }
    }
    __syncthreads();  // B1
    double xr[T];
    {
      double S = 0.0;
#pragma unroll
      for (int w = 0; w < SL_WAVES; ++w) S += red[0][w];
      const double hh = (i > 0) ? 0.5 * tau_prev * S : 0.0;
      const double vpi = sc[0], wpi = sc[1] - hh * sc[0];
      double s2 = 0.0;
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const int sl = wv + SL_WAVES * t;
        const int r = lane + 64 * sl;
        wpr[t] -= hh * vpr[t];                             // w_{i-1}
        xr[t] = bcr[t] - vpr[t] * wpi - wpr[t] * vpi;      // zero outside [i, n)
        if (r >= i + 2) s2 += xr[t] * xr[t];
        if (sl >= k0 && sl < kr) sW[r] = wpr[t];
      }
      s2 = wave_sum64(s2);
      if (lane == 0) red[1][wv] = s2;
      if (row_i) sc[2] = slot_value(xr, si >> 3);
      if (row_i1) sc[3] = slot_value(xr, si1 >> 3);
    }
    __syncthreads();  // B2
    double tu = 0.0;
    {
      double xn2 = 0.0;
#pragma unroll
      for (int w = 0; w < SL_WAVES; ++w) xn2 += red[1][w];
      const double di = sc[2];
      double beta = 0.0, scale = 1.0;
      if (!last) {
        const double alpha = sc[3];
        beta = alpha;
        if (xn2 > 0.0) {
          beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
          tu = (beta - alpha) / beta;
          scale = 1.0 / (alpha - beta);
        }
      }
      const bool own = (i % G == g);
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const int sl = wv + SL_WAVES * t;
        const int r = lane + 64 * sl;
        const double vv = (r >= i + 2) ? xr[t] * scale : ((r == i + 1) ? 1.0 : 0.0);
        if (sl >= k0 && sl < kr) vcur[r] = vv;
        if (own && r >= i + 2 && r < n) a[r + (int64_t)i * n] = vv;
        vpr[t] = vv;
      }
      if (wv == 0 && lane == 0) {
        if (own) {
          a[i + (int64_t)i * n] = di;
          if (!last) a[i + 1 + (int64_t)i * n] = beta;
        }
        if (g == 0) {
          D[tk.v_off + i] = di;
          E[tk.v_off + i] = beta;
          tau[tk.v_off + i] = tu;
        }
      }
    }
    __syncthreads();  // B3: v_i and w_{i-1} in LDS
    if (last) break;
    tau_prev = tu;
    {
      // step i-1's rank-2 update of the own columns and their p_i entries
      // over the slots from k0 on: v_{i-1} and w_{i-1} vanish above row i and
      // v_i above row i + 1, so the rows of the slots below are final and
      // contribute nothing; the zero padding stays zero
      double vpB[KR], wpB[KR], vcB[KR];
#pragma unroll
      for (int k = 0; k < KR; ++k) {
        vpB[k] = wpB[k] = vcB[k] = 0.0;
        if (k >= k0 && k < kr) {
          const int r = lane + 64 * k;
          vpB[k] = vprev[r];
          wpB[k] = sW[r];
          vcB[k] = vcur[r];
        }
      }
      double *pc = xb + (2 * (int64_t)i + 1) * xs;
      double *bn = pc + xs;
      // failure-path test hook: workgroup 0 withholds its p entries of one
      // step, so the readers' bounded wait must end in the error flag
      const bool drop = (i == drop_step && g == 0);
      // register columns: no LDS traffic; finished slots see zero factors;
      // the dot products of all of them reduced together
      if (CW > 0) {
        double acc[CW > 0 ? CW : 1];
#pragma unroll
        for (int m = 0; m < CW; ++m) {
          const int q = nlds + wv + SL_WAVES * m;
          const int j = g + q * G;
          acc[m] = 0.0;
          if (q < ncl && j > i && j < n) {
            const double vpj = vprev[j], wpj = sW[j];
#pragma unroll
            for (int k = 0; k < KR; ++k) {
              colr[m][k] = colr[m][k] - vpB[k] * wpj - wpB[k] * vpj;
              acc[m] += colr[m][k] * vcB[k];
            }
            if (j == i + 1) {  // the next step's column, for every workgroup
#pragma unroll
              for (int k = 0; k < KR; ++k) {
                const int r = lane + 64 * k;
                if (k >= k0 && k < kr && r >= i + 1 && r < n) xstore(bn + r, colr[m][k]);
              }
            }
          }
        }
        wave_sum64_n(acc);
        if (lane == 0) {
#pragma unroll
          for (int m = 0; m < CW; ++m) {
            const int q = nlds + wv + SL_WAVES * m;
            const int j = g + q * G;
            if (q < ncl && j > i && j < n && !drop) xstore(pc + j, tu * acc[m]);
          }
        }
      }
      const int q0 = std::max(0, (i + 1 - g + G - 1) / G);
      for (int q = q0 + wv; q < nlds; q += SL_WAVES) {
        const int j = g + q * G;
        if (j >= n) break;
        const double vpj = vprev[j], wpj = sW[j];
        double *col = cols + (size_t)q * np;
        // loads, arithmetic and stores in separate passes: the arithmetic
        // has no branches (finished slots compute zeros), so the loads are
        // all in flight before the first wait and no store waits on another
        double xc[KR];
#pragma unroll
        for (int k = 0; k < KR; ++k) {
          xc[k] = 0.0;
          if (k >= k0 && k < kr) xc[k] = col[lane + 64 * k];
        }
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < KR; ++k) {
          xc[k] = xc[k] - vpB[k] * wpj - wpB[k] * vpj;
          acc += xc[k] * vcB[k];
        }
#pragma unroll
        for (int k = 0; k < KR; ++k)
          if (k >= k0 && k < kr) col[lane + 64 * k] = xc[k];
        if (j == i + 1) {  // the next step's column, for every workgroup
#pragma unroll
          for (int k = 0; k < KR; ++k) {
            const int r = lane + 64 * k;
            if (k >= k0 && k < kr && r >= i + 1 && r < n) xstore(bn + r, xc[k]);
          }
        }
        acc = wave_sum64(acc);
        if (lane == 0 && !drop) xstore(pc + j, tu * acc);
      }
    }
  }
}

// fill the exchange slots with the empty pattern
__global__ void k_sl_fill(long long *__restrict__ x, size_t len) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < len;
       e += (size_t)gridDim.x * blockDim.x)
    x[e] = SL_EMPTY;
}

// z <- Q z for selected columns, Q = H_0 H_1 ... H_{n-2}; one wavefront per
// column, the column held in LDS, reflectors read from A in reverse order.
constexpr int OR_THREADS = 256;
constexpr int OR_WAVES = OR_THREADS / 64;

struct OrTask {
  int64_t a_off, v_off, z_off;
  int n, col0, ncol, w0;  // w0: first workgroup of this matrix
};
struct OrTable {
  int nt;
  int pad;
  OrTask t[SY_MAXT];
};

__global__ __launch_bounds__(OR_THREADS) void k_ormtr_cols(OrTable tab, const double *__restrict__ A,
                                                           const double *__restrict__ tau,
                                                           double *__restrict__ Z) {
  extern __shared__ double zs[];
  int t = 0;
  while (t + 1 < tab.nt && (int)blockIdx.x >= tab.t[t + 1].w0) ++t;
  const OrTask tk = tab.t[t];
  const int n = tk.n, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int jc = ((int)blockIdx.x - tk.w0) * OR_WAVES + wv;  // column within the selection
  if (jc >= tk.ncol) return;  // whole wavefront leaves; no block barrier below
  double *z = zs + (int64_t)wv * n;
  double *zg = Z + tk.z_off + (int64_t)(tk.col0 + jc) * n;
  for (int r = lane; r < n; r += 64) z[r] = zg[r];
  const double *a = A + tk.a_off;
  const double *tv = tau + tk.v_off;
  for (int i = n - 2; i >= 0; --i) {
    const double ti = tv[i];
    if (ti == 0.0) continue;  // uniform across the wavefront
    const double *v = a + (int64_t)i * n;
    // v = [0 .. 0, 1 (row i+1), A[i+2:, i]]
    double s = (lane == 0) ? z[i + 1] : 0.0;
    for (int r = i + 2 + lane; r < n; r += 64) s += v[r] * z[r];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    const double f = ti * s;
    if (lane == 0) z[i + 1] -= f;
    for (int r = i + 2 + lane; r < n; r += 64) z[r] -= f * v[r];
  }
  for (int r = lane; r < n; r += 64) zg[r] = z[r];
}

// Blocked back-transformation (compact WY, Schreiber & Van Loan 1989; the
// LAPACK dlarft/dlarfb scheme): the reflectors H_i = I - tau_i v_i v_i^T of
// the tridiagonalisation (v_i[i+1] = 1, v_i[r] = A[r, i] for r >= i+2) are
// grouped in blocks of WY_NB, Q_b = H_{i0} ... H_{i0+NB-1} = I - V T V^T with
// T upper triangular, and Z <- Q Z is applied block by block from the last:
// W = V^T Z, W = T W, Z -= V W.  k_wy_t forms T for every (matrix, block);
// k_wy_apply keeps a tile of WY_CT columns of Z in LDS for all the blocks.
constexpr int WY_NB = 32;
constexpr int WY_CT = 16;

struct WyTask {
  int64_t a_off, v_off, z_off, t_off;  // matrix, tau, Z, first T block
  int n, col0, ncol, w0;               // w0: first workgroup (k_wy_apply) or block (k_wy_t)
};

typedef double wy_d4 __attribute__((ext_vector_type(4)));

// entry r of reflector i (zero for i outside [0, nrefl) or r outside the column)
__device__ inline double vref(const double *a, int n, int nrefl, int i, int r) {
  if (i >= nrefl || r >= n || r <= i) return 0.0;
  return (r == i + 1) ? 1.0 : a[r + (int64_t)i * n];
}


__global__ __launch_bounds__(256) void k_wy_t(const WyTask *__restrict__ tab, int nt,
                                              const double *__restrict__ A,
                                              const double *__restrict__ tau,
                                              double *__restrict__ T) {
  __shared__ double Gp[4][WY_NB][WY_NB + 1];
  __shared__ double Ts[WY_NB][WY_NB + 1];
  int lo = 0, hi = nt - 1;
  while (lo < hi) {
    const int md = (lo + hi + 1) >> 1;
    if (tab[md].w0 <= (int)blockIdx.x) lo = md; else hi = md - 1;
  }
  const WyTask tk = tab[lo];
  const int n = tk.n, b = (int)blockIdx.x - tk.w0;
  const int nrefl = n - 1;  // H_0 .. H_{n-2}
  const int i0 = b * WY_NB, nb = min(WY_NB, nrefl - i0);
  const double *a = A + tk.a_off;
  const double *tv = tau + tk.v_off;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  // G = V^T V on FP64 MFMA: the four waves take alternate 4-row K steps; an
  // operand lane holds v_{i0 + li}[r] for A (V^T) and for B (V) alike
  wy_d4 g00 = {0.0, 0.0, 0.0, 0.0}, g01 = g00, g10 = g00, g11 = g00;
  for (int ks = wv; i0 + 1 + 4 * ks < n; ks += 4) {
    const int r = i0 + 1 + 4 * ks + lk;
    const double a0 = vref(a, n, nrefl, i0 + li, r);
    const double a1 = vref(a, n, nrefl, i0 + 16 + li, r);
    g00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a0, g00, 0, 0, 0);
    g01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a1, g01, 0, 0, 0);
    g10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, a0, g10, 0, 0, 0);
    g11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, a1, g11, 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    Gp[wv][lk + 4 * j][li] = g00[j];
    Gp[wv][lk + 4 * j][16 + li] = g01[j];
    Gp[wv][16 + lk + 4 * j][li] = g10[j];
    Gp[wv][16 + lk + 4 * j][16 + li] = g11[j];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < WY_NB * WY_NB; e += 256) {
    const int l = e / WY_NB, j = e % WY_NB;
    Gp[0][l][j] = Gp[0][l][j] + Gp[1][l][j] + Gp[2][l][j] + Gp[3][l][j];
    Ts[l][j] = 0.0;
  }
  __syncthreads();
  // T[j][j] = tau_j; T[0:j, j] = -tau_j T[0:j, 0:j] G[0:j, j]
  for (int j = 0; j < nb; ++j) {
    const double tj = tv[i0 + j];
    if (threadIdx.x < j) {
      const int l = threadIdx.x;
      double s = 0.0;
      for (int m = l; m < j; ++m) s += Ts[l][m] * Gp[0][m][j];
      Ts[l][j] = -tj * s;
    }
    if (threadIdx.x == 0) Ts[j][j] = tj;
    __syncthreads();
  }
  double *to = T + tk.t_off + (int64_t)b * WY_NB * WY_NB;
  for (int e = threadIdx.x; e < WY_NB * WY_NB; e += 256) to[e] = Ts[e / WY_NB][e % WY_NB];
}

// One workgroup per WY_CT-column tile of Z (all n rows in LDS, row-major),
// on FP64 MFMA (16x16x4: lane l holds A[l%16][l/16], B[l/16][l%16]; the
// accumulator's register j is C[l/16 + 4j][l%16]).  Per block, from the last:
//   W = V^T Z   (32 x WY_CT; the four waves take alternate 4-row K steps and
//                add their partial tiles through LDS),
//   W = T W,
//   Z -= V W    (16-row tiles of Z over the waves, K = the 32 reflectors).
// V is read in MFMA operand order straight from A (the L2 holds it: every
// tile workgroup of a matrix reads the same reflectors).
__global__ __launch_bounds__(256) void k_wy_apply(const WyTask *__restrict__ tab, int nt,
                                                  const double *__restrict__ A,
                                                  const double *__restrict__ T,
                                                  double *__restrict__ Z) {
  extern __shared__ double zt[];  // n rows x WY_CT columns, row-major
  __shared__ double Wp[4][WY_NB][WY_CT];
  __shared__ double W2[WY_NB][WY_CT + 1];
  __shared__ double Tb[WY_NB][WY_NB + 1];
  int lo = 0, hi = nt - 1;
  while (lo < hi) {
    const int md = (lo + hi + 1) >> 1;
    if (tab[md].w0 <= (int)blockIdx.x) lo = md; else hi = md - 1;
  }
  const WyTask tk = tab[lo];
  const int n = tk.n, tile = (int)blockIdx.x - tk.w0;
  const int c0 = tk.col0 + tile * WY_CT, nc = min(WY_CT, tk.col0 + tk.ncol - c0);
  const double *a = A + tk.a_off;
  double *z = Z + tk.z_off;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  for (int e = threadIdx.x; e < WY_CT * n; e += 256) {
    const int c = e / n, r = e % n;
    zt[r * WY_CT + c] = (c < nc) ? z[r + (int64_t)(c0 + c) * n] : 0.0;
  }
  const int nrefl = n - 1;
  const int nblk = (nrefl + WY_NB - 1) / WY_NB;
  for (int b = nblk - 1; b >= 0; --b) {
    const int i0 = b * WY_NB, rstart = i0 + 1;
    const double *tb = T + tk.t_off + (int64_t)b * WY_NB * WY_NB;
    for (int e = threadIdx.x; e < WY_NB * WY_NB; e += 256) Tb[e / WY_NB][e % WY_NB] = tb[e];
    __syncthreads();  // also orders the previous block's Z update before these reads
    // W = V^T Z, partial per wave
    wy_d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    // four K steps per pass: all operand loads first, then the MFMAs
    for (int ks = wv; rstart + 4 * ks < n; ks += 16) {
      double a0[4], a1[4], bz[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = rstart + 4 * (ks + 4 * u) + lk;
        a0[u] = vref(a, n, nrefl, i0 + li, r);
        a1[u] = vref(a, n, nrefl, i0 + 16 + li, r);
        bz[u] = (r < n) ? zt[r * WY_CT + li] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[u], bz[u], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[u], bz[u], acc1, 0, 0, 0);
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      Wp[wv][lk + 4 * j][li] = acc0[j];
      Wp[wv][16 + lk + 4 * j][li] = acc1[j];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < WY_NB * WY_CT; e += 256) {
      const int l = e / WY_CT, c = e % WY_CT;
      Wp[0][l][c] = Wp[0][l][c] + Wp[1][l][c] + Wp[2][l][c] + Wp[3][l][c];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < WY_NB * WY_CT; e += 256) {
      const int l = e / WY_CT, c = e % WY_CT;
      double s2 = 0.0;
      for (int m = l; m < WY_NB; ++m) s2 += Tb[l][m] * Wp[0][m][c];
      W2[l][c] = s2;
    }
    __syncthreads();
    // Z -= V W2
    for (int t = wv; rstart + 16 * t < n; t += 4) {
      const int rb = rstart + 16 * t;
      wy_d4 acc = {0.0, 0.0, 0.0, 0.0};
      double av[WY_NB / 4];
#pragma unroll
      for (int ks = 0; ks < WY_NB / 4; ++ks) av[ks] = vref(a, n, nrefl, i0 + 4 * ks + lk, rb + li);
#pragma unroll
      for (int ks = 0; ks < WY_NB / 4; ++ks)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ks], W2[4 * ks + lk][li], acc, 0, 0, 0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = rb + lk + 4 * j;
        if (r < n) zt[r * WY_CT + li] -= acc[j];
      }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < WY_CT * n; e += 256) {
    const int c = e / n, r = e % n;
    if (c < nc) z[r + (int64_t)(c0 + c) * n] = zt[r * WY_CT + c];
  }
}

struct EigState {
  std::vector<hipStream_t> streams;
  std::vector<rocblas_handle> rb;
  std::vector<hipEvent_t> done;
  hipEvent_t ready = nullptr;
  int *info = nullptr;
  int info_len = 0;
  int coresident = 0;  // sytrd workgroups that fit on the device at once
  int coresident_lds = -1;
  void *wy = nullptr;  // back-transformation: task table and T blocks
  size_t wy_bytes = 0;
};

EigState *state(mh_handle h) {
  if (!h->eig_state) h->eig_state = new EigState();
  return (EigState *)h->eig_state;
}

int ensure_eig_scratch(mh_handle h, size_t bytes) {
  if (h->eig_bytes >= bytes) return MH_OK;
  MH_HIP(h, hipStreamSynchronize(h->stream));
  if (h->eig) MH_HIP(h, hipFree(h->eig));
  h->eig = nullptr;
  h->eig_err = nullptr;
  h->eig_bytes = 0;
  size_t b = (bytes + 255) & ~size_t(255);
  hipError_t e = hipMalloc(&h->eig, b);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    h->eig = nullptr;
    return fail(h, MH_ERR_OOM, "eigensolver scratch of %zu bytes: %s", b, hipGetErrorString(e));
  }
  h->eig_bytes = b;
  return MH_OK;
}

int eig_streams(mh_handle h, int k) {
  EigState *s = state(h);
  while ((int)s->streams.size() < k) {
    hipStream_t st;
    MH_HIP(h, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    s->streams.push_back(st);
    rocblas_handle rb;
    if (rocblas_create_handle(&rb) != rocblas_status_success)
      return fail(h, MH_ERR_HIP, "rocblas_create_handle failed");
    if (rocblas_set_stream(rb, st) != rocblas_status_success)
      return fail(h, MH_ERR_HIP, "rocblas_set_stream failed");
    s->rb.push_back(rb);
    hipEvent_t ev;
    MH_HIP(h, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    s->done.push_back(ev);
  }
  if (!s->ready) MH_HIP(h, hipEventCreateWithFlags(&s->ready, hipEventDisableTiming));
  return MH_OK;
}

}  // namespace

void eig_release(mh_handle h) {
  EigState *s = (EigState *)h->eig_state;
  if (s) {
    for (auto st : s->streams) (void)hipStreamSynchronize(st);
    for (auto rb : s->rb) (void)rocblas_destroy_handle(rb);
    for (auto st : s->streams) (void)hipStreamDestroy(st);
    for (auto ev : s->done) (void)hipEventDestroy(ev);
    if (s->ready) (void)hipEventDestroy(s->ready);
    if (s->info) (void)hipFree(s->info);
    if (s->wy) (void)hipFree(s->wy);
    delete s;
  }
  h->eig_state = nullptr;
  if (h->eig) (void)hipFree(h->eig);
  h->eig = nullptr;
  h->eig_err = nullptr;
  h->eig_bytes = 0;
}

// k_sytrd_lds for KR = 1 .. SL_KRMAX row slots per lane, with CW register
// columns per wavefront: as many as the register file holds beside the three
// broadcast vectors and an LDS column in flight, (CW + 4) KR doubles within
// about 99 per lane; the table keeps every instance free of register spills
constexpr int kSlCw[SL_KRMAX + 1] = {0, 6, 6, 6, 6, 6, 6, 6, 6, 6, 5, 4, 4, 3, 2, 2, 1, 0, 0};
constexpr int sl_cw(int kr) { return kSlCw[kr]; }
template <int KR>
constexpr const void *sl_kernel() { return (const void *)k_sytrd_lds<KR, sl_cw(KR)>; }
const void *const kSlKernels[SL_KRMAX] = {
    sl_kernel<1>(),  sl_kernel<2>(),  sl_kernel<3>(),  sl_kernel<4>(),  sl_kernel<5>(),
    sl_kernel<6>(),  sl_kernel<7>(),  sl_kernel<8>(),  sl_kernel<9>(),  sl_kernel<10>(),
    sl_kernel<11>(), sl_kernel<12>(), sl_kernel<13>(), sl_kernel<14>(), sl_kernel<15>(),
    sl_kernel<16>(), sl_kernel<17>(), sl_kernel<18>()};

// Plan and launch k_sytrd_lds.  Matrix m gets G_m workgroups holding
// ncl = ceil(n / G) columns each: up to 8 CW in registers, the rest in LDS.
// The planner starts from the fewest workgroups the LDS budget allows and
// hands the device's remaining workgroup slots, one at a time, to the matrix
// with the longest estimated run (steps x per-step time, the step priced by
// its fixed exchange and vector work plus a cost per register column and a
// larger one per LDS column), so the largest matrices spread out most.  If
// the batch does not fit at once, the matrices go, largest first, into
// consecutive launches.
int sytrd_lds(mh_handle h, int nmat, const int *n, const int64_t *a_off, const int64_t *v_off,
              double *A, double *D, double *E, double *tau) {
  int ncu = 0, lds_cu = 0;
  MH_HIP(h, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device));
  MH_HIP(h, hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                  h->device));
  if (lds_cu <= 0) lds_cu = 64 * 1024;
  int nmax = 0;
  for (int m = 0; m < nmat; ++m) nmax = std::max(nmax, n[m]);
  const int KR = (nmax + 63) / 64;
  if (KR < 1 || KR > SL_KRMAX) return MH_ERR_UNSUPPORTED;
  const int regc = SL_WAVES * sl_cw(KR);  // register columns per workgroup
  const void *kern = kSlKernels[KR - 1];
  // test hooks of the failure path: the bound of every wait between
  // workgroups (polls) and a step whose p entries workgroup 0 withholds
  unsigned spin_limit = SY_SPIN_LIMIT;
  int drop_step = -1;
  if (const char *e = getenv("MH_SYTRD_SPIN")) spin_limit = (unsigned)std::max(1, atoi(e));
  if (const char *e = getenv("MH_SYTRD_DROP")) drop_step = atoi(e);
  std::vector<int> order(nmat);
  for (int m = 0; m < nmat; ++m) order[m] = m;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return n[x] > n[y]; });
  // per-step cost model (microseconds): fixed part, per register column,
  // per LDS column
  constexpr double kFix = 3.4, kReg = 0.02, kLds = 0.085;
  auto est = [&](int nm, int G) {
    const int ncl = (nm + G - 1) / G;
    return nm * (kFix + kReg * std::min(ncl, regc) + kLds * std::max(0, ncl - regc));
  };
  const int64_t budget = (int64_t)lds_cu - 2048;  // LDS columns + 3 vectors
  std::vector<int> G(nmat);
  std::vector<SlTask> tab(nmat);
  int gmin = 0;
  for (int pi = 0; pi < nmat; ++pi) {
    const int nm = n[order[pi]];
    const int xs = (nm + 63) & ~63;
    const int64_t capc = (budget / 8 - 3 * (int64_t)xs) / xs + regc;
    if (capc < 1) return MH_ERR_UNSUPPORTED;
    G[pi] = (int)((nm + capc - 1) / capc);
    gmin += G[pi];
  }
  auto lds_of = [&](int pi) {
    const int nm = n[order[pi]];
    const int xs = (nm + 63) & ~63;
    const int ncl = (nm + G[pi] - 1) / G[pi];
    return ((size_t)std::max(0, ncl - regc) + 3) * xs * sizeof(double);
  };
  size_t lds = 0;
  for (int pi = 0; pi < nmat; ++pi) lds = std::max(lds, lds_of(pi));
  int per_cu = 0;
  MH_HIP(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, SL_THREADS, lds));
  const int cap = per_cu * ncu;
  if (cap < 1) return MH_ERR_UNSUPPORTED;
  for (int pi = 0; pi < nmat; ++pi)
    if (G[pi] > cap) return MH_ERR_UNSUPPORTED;
  if (gmin <= cap) {
    // spread: extra workgroups to the longest estimated run (only while it
    // shortens that run; the LDS need only falls as G grows)
    int gt = gmin;
    while (gt < cap) {
      int best = -1;
      double tb = 0.0;
      for (int pi = 0; pi < nmat; ++pi) {
        const int nm = n[order[pi]];
        if (G[pi] >= nm) continue;
        const double t = est(nm, G[pi]);
        if (t > tb && est(nm, G[pi] + 1) < t) {
          tb = t;
          best = pi;
        }
      }
      if (best < 0) break;
      ++G[best];
      ++gt;
    }
  }
  int64_t xsum = 0;  // exchange doubles of the batch
  for (int pi = 0; pi < nmat; ++pi) {
    const int m = order[pi];
    const int xs = (n[m] + 63) & ~63;
    tab[pi] = SlTask{a_off[m], v_off[m], xsum, n[m], G[pi], 0, (n[m] + G[pi] - 1) / G[pi], xs, 0};
    xsum += 2 * (int64_t)n[m] * xs;
  }
  // launches: consecutive runs of the plan whose workgroups fit at once;
  // g0 counts from the launch's first workgroup
  std::vector<std::pair<int, int>> launches;  // (first task, task count)
  {
    int first = 0, g = 0;
    for (int pi = 0; pi < nmat; ++pi) {
      if (g + tab[pi].G > cap) {
        launches.push_back({first, pi - first});
        first = pi;
        g = 0;
      }
      tab[pi].g0 = g;
      g += tab[pi].G;
    }
    launches.push_back({first, nmat - first});
  }
  // scratch: the exchange slots (every step of every matrix), the task
  // table, the error flag
  const size_t xbytes = ((size_t)xsum * sizeof(double) + 255) & ~size_t(255);
  const size_t tbytes = ((size_t)nmat * sizeof(SlTask) + 255) & ~size_t(255);
  MH_TRY(ensure_eig_scratch(h, xbytes + tbytes + 256));
  char *base = (char *)h->eig;
  double *X = (double *)base;
  SlTask *dtab = (SlTask *)(base + xbytes);
  int *err = (int *)(base + xbytes + tbytes);
  h->eig_err = err;
  MH_HIP(h, hipMemsetAsync(err, 0, sizeof(int), h->stream));
  MH_HIP(h, hipMemcpyAsync(dtab, tab.data(), (size_t)nmat * sizeof(SlTask), hipMemcpyHostToDevice,
                           h->stream));
  {
    const size_t len = (size_t)xsum;
    const int blocks = (int)std::min<size_t>((len + 255) / 256, 4096);
    hipLaunchKernelGGL(k_sl_fill, dim3(std::max(blocks, 1)), dim3(256), 0, h->stream,
                       (long long *)X, len);
    MH_TRY(check_launch(h, "k_sl_fill"));
  }
  for (const auto &L : launches) {
    SlTask *t0 = dtab + L.first;
    int nt = L.second;
    int g = 0;
    for (int pi = L.first; pi < L.first + L.second; ++pi) g += tab[pi].G;
    void *args[] = {&t0, &nt, &A, &D, &E, &tau, &X, &err, &spin_limit, &drop_step};
    MH_HIP(h, hipLaunchCooperativeKernel(kern, dim3(g), dim3(SL_THREADS), args, (unsigned)lds,
                                         h->stream));
  }
  return MH_OK;
}

// Largest dynamic LDS (the Z tile) k_wy_apply may take: the device's limit
// per workgroup less the kernel's static LDS.
int wy_lds_limit(mh_handle h, size_t *out) {
  int per_block = 0;
  MH_HIP(h, hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                  h->device));
  hipFuncAttributes fa{};
  MH_HIP(h, hipFuncGetAttributes(&fa, (const void *)k_wy_apply));
  *out = (per_block > (int)fa.sharedSizeBytes) ? (size_t)per_block - fa.sharedSizeBytes : 0;
  return MH_OK;
}

// Blocked back-transformation of the selected columns: T blocks for every
// (matrix, reflector block), then one workgroup per WY_CT-column tile.
int ormtr_wy(mh_handle h, int nmat, const int *n, const int64_t *a_off, const int64_t *v_off,
             const double *A, const double *tau, const int64_t *z_off, const int *col0,
             const int *ncol, double *Z) {
  std::vector<WyTask> tt, ta;
  int64_t toff = 0;
  int nb_tot = 0, wg_tot = 0, nmax = 0;
  for (int m = 0; m < nmat; ++m) {
    if (ncol[m] == 0 || n[m] < 2) continue;  // order 1: Q = I
    const int nblk = (n[m] - 1 + WY_NB - 1) / WY_NB;
    WyTask t{a_off[m], v_off[m], z_off[m], toff, n[m], col0[m], ncol[m], nb_tot};
    tt.push_back(t);
    t.w0 = wg_tot;
    ta.push_back(t);
    nb_tot += nblk;
    wg_tot += (ncol[m] + WY_CT - 1) / WY_CT;
    toff += (int64_t)nblk * WY_NB * WY_NB;
    nmax = std::max(nmax, n[m]);
  }
  if (tt.empty()) return MH_OK;
  EigState *st = state(h);
  const size_t tab_bytes = (tt.size() * sizeof(WyTask) + 255) & ~size_t(255);
  const size_t need = 2 * tab_bytes + (size_t)toff * sizeof(double);
  if (st->wy_bytes < need) {
    MH_HIP(h, hipStreamSynchronize(h->stream));
    if (st->wy) MH_HIP(h, hipFree(st->wy));
    st->wy = nullptr;
    st->wy_bytes = 0;
    hipError_t e = hipMalloc(&st->wy, need);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      st->wy = nullptr;
      return fail(h, MH_ERR_OOM, "mh_ormtr_batched: %zu bytes: %s", need, hipGetErrorString(e));
    }
    st->wy_bytes = need;
  }
  char *base = (char *)st->wy;
  WyTask *dtt = (WyTask *)base, *dta = (WyTask *)(base + tab_bytes);
  double *T = (double *)(base + 2 * tab_bytes);
  MH_HIP(h, hipMemcpyAsync(dtt, tt.data(), tt.size() * sizeof(WyTask), hipMemcpyHostToDevice,
                           h->stream));
  MH_HIP(h, hipMemcpyAsync(dta, ta.data(), ta.size() * sizeof(WyTask), hipMemcpyHostToDevice,
                           h->stream));
  const int nt = (int)tt.size();
  hipLaunchKernelGGL(k_wy_t, dim3(nb_tot), dim3(256), 0, h->stream, dtt, nt, A, tau, T);
  MH_TRY(check_launch(h, "k_wy_t"));
  const size_t lds = (size_t)WY_CT * nmax * sizeof(double);
  hipLaunchKernelGGL(k_wy_apply, dim3(wg_tot), dim3(256), lds, h->stream, dta, nt, A, T, Z);
  MH_TRY(check_launch(h, "k_wy_apply"));
  return MH_OK;
}

}  // namespace mh

using namespace mh;

extern "C" int mh_sytrd_batched(mh_handle h, int nmat, const int *n, const int64_t *a_off,
                                const int64_t *v_off, double *A, double *D, double *E,
                                double *tau) {
  if (!h || nmat < 0 || (nmat > 0 && (!n || !a_off || !v_off || !A || !D || !E || !tau)))
    return fail(h, MH_ERR_ARG, "mh_sytrd_batched: null argument");
  if (nmat == 0) return MH_OK;
  int nmax = 0;
  int64_t vend = 0;
  for (int m = 0; m < nmat; ++m) {
    if (n[m] < 1 || n[m] > SY_MAXN)
      return fail(h, MH_ERR_SHAPE, "mh_sytrd_batched: order %d of matrix %d outside [1, %d]",
                  n[m], m, SY_MAXN);
    if (a_off[m] < 0 || v_off[m] < 0)
      return fail(h, MH_ERR_SHAPE, "mh_sytrd_batched: negative offset (matrix %d)", m);
    nmax = std::max(nmax, n[m]);
    vend = std::max(vend, v_off[m] + n[m]);
  }
  MH_HIP(h, hipSetDevice(h->device));
  EigState *st = state(h);
  if (nmax <= SL_MAXN && !getenv("MH_SYTRD_GLOBAL")) {
    int status = sytrd_lds(h, nmat, n, a_off, v_off, A, D, E, tau);
    if (status != MH_ERR_UNSUPPORTED) return status;  // else: does not fit, global form
  }
  const size_t lds = (size_t)3 * nmax * sizeof(double);
  if (st->coresident_lds != (int)lds) {
    int per_cu = 0, ncu = 0;
    MH_HIP(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_sytrd_group,
                                                            SY_THREADS, lds));
    MH_HIP(h, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device));
    st->coresident = per_cu * ncu;
    st->coresident_lds = (int)lds;
  }
  if (st->coresident < 1)
    return fail(h, MH_ERR_UNSUPPORTED, "mh_sytrd_batched: kernel does not fit on the device");
  // scratch: two exchange halves for p and for the broadcast column, the
  // barrier counters (one per matrix of a launch) and the error flag
  const size_t xbytes = (size_t)vend * sizeof(double);
  const size_t need = 4 * xbytes + 256 + (SY_MAXT + 64) * sizeof(unsigned);
  MH_TRY(ensure_eig_scratch(h, need));
  double *P = (double *)h->eig;
  double *BC = P + 2 * vend;
  unsigned *cnt = (unsigned *)((char *)h->eig + ((4 * xbytes + 255) & ~size_t(255)));
  int *err = (int *)(cnt + SY_MAXT);
  h->eig_err = err;
  MH_HIP(h, hipMemsetAsync(err, 0, sizeof(int), h->stream));
  // largest matrices first, in launches of at most SY_MAXT matrices whose
  // workgroups are all resident at once
  std::vector<int> order(nmat);
  for (int m = 0; m < nmat; ++m) order[m] = m;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return n[x] > n[y]; });
  size_t pos = 0;
  while (pos < order.size()) {
    SyTable tab{};
    tab.xrow = vend;
    int g = 0;
    while (pos < order.size() && tab.nt < SY_MAXT) {
      const int m = order[pos];
      int G = std::min(SY_MAX_G, std::max(1, (n[m] + SY_COLS_PER_WG - 1) / SY_COLS_PER_WG));
      G = std::min(G, std::max(1, n[m]));
      if (g + G > st->coresident) {
        if (tab.nt > 0) break;
        G = st->coresident;  // a single matrix larger than the device: fewer groups
      }
      SyTask &tk = tab.t[tab.nt++];
      tk.a_off = a_off[m];
      tk.v_off = v_off[m];
      tk.n = n[m];
      tk.G = G;
      tk.g0 = g;
      g += G;
      ++pos;
    }
    MH_HIP(h, hipMemsetAsync(cnt, 0, SY_MAXT * sizeof(unsigned), h->stream));
    void *args[] = {&tab, &A, &D, &E, &tau, &P, &BC, &cnt, &err};
    MH_HIP(h, hipLaunchCooperativeKernel((const void *)k_sytrd_group, dim3(g), dim3(SY_THREADS),
                                         args, (unsigned)lds, h->stream));
  }
  return MH_OK;
}

extern "C" int mh_stedc_batched(mh_handle h, int nmat, const int *n, const int64_t *v_off,
                                const int64_t *z_off, double *D, double *E, double *Z) {
  if (!h || nmat < 0 || (nmat > 0 && (!n || !v_off || !z_off || !D || !E || !Z)))
    return fail(h, MH_ERR_ARG, "mh_stedc_batched: null argument");
  if (nmat == 0) return MH_OK;
  for (int m = 0; m < nmat; ++m)
    if (n[m] < 1) return fail(h, MH_ERR_SHAPE, "mh_stedc_batched: order %d", n[m]);
  MH_HIP(h, hipSetDevice(h->device));
  int k = 4;
  if (const char *e = getenv("MH_EIG_STREAMS")) k = std::max(1, std::min(16, atoi(e)));
  k = std::min(k, nmat);
  MH_TRY(eig_streams(h, k));
  EigState *st = state(h);
  if (st->info_len < nmat) {
    MH_HIP(h, hipStreamSynchronize(h->stream));
    if (st->info) MH_HIP(h, hipFree(st->info));
    st->info = nullptr;
    st->info_len = 0;
    MH_HIP(h, hipMalloc(&st->info, (size_t)std::max(nmat, 64) * sizeof(int)));
    st->info_len = std::max(nmat, 64);
  }
  MH_HIP(h, hipMemsetAsync(st->info, 0, (size_t)nmat * sizeof(int), h->stream));
  MH_HIP(h, hipEventRecord(st->ready, h->stream));
  for (int q = 0; q < k; ++q) MH_HIP(h, hipStreamWaitEvent(st->streams[q], st->ready, 0));
  // largest first, each to the stream with the least estimated work; one
  // host thread per stream issues that stream's calls (a dstedc call can
  // block its host thread, which would otherwise serialise the streams)
  std::vector<int> order(nmat);
  for (int m = 0; m < nmat; ++m) order[m] = m;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return n[x] > n[y]; });
  std::vector<double> load(k, 0.0);
  std::vector<std::vector<int>> work(k);
  for (int m : order) {
    const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[q] += 0.5 + 0.003 * n[m];  // ms: a fixed latency plus a size term
    work[q].push_back(m);
  }
  std::vector<int> bad(k, -1), bad_status(k, 0);
  auto issue = [&](int q) {
    (void)hipSetDevice(h->device);
    for (int m : work[q]) {
      rocblas_status rs = rocsolver_dstedc(st->rb[q], rocblas_evect_tridiagonal, n[m],
                                           D + v_off[m], E + v_off[m], Z + z_off[m], n[m],
                                           st->info + m);
      if (rs != rocblas_status_success) {
        bad[q] = m;
        bad_status[q] = (int)rs;
        return;
      }
    }
  };
  if (k == 1) {
    issue(0);
  } else {
    std::vector<std::thread> th;
    for (int q = 0; q < k; ++q) th.emplace_back(issue, q);
    for (auto &x : th) x.join();
  }
  for (int q = 0; q < k; ++q)
    if (bad[q] >= 0)
      return fail(h, MH_ERR_HIP, "rocsolver_dstedc (matrix %d, n = %d) failed: status %d", bad[q],
                  n[bad[q]], bad_status[q]);
  for (int q = 0; q < k; ++q) {
    MH_HIP(h, hipEventRecord(st->done[q], st->streams[q]));
    MH_HIP(h, hipStreamWaitEvent(h->stream, st->done[q], 0));
  }
  return MH_OK;
}

extern "C" int mh_ormtr_batched(mh_handle h, int nmat, const int *n, const int64_t *a_off,
                                const int64_t *v_off, const double *A, const double *tau,
                                const int64_t *z_off, const int *col0, const int *ncol, double *Z) {
  if (!h || nmat < 0 ||
      (nmat > 0 && (!n || !a_off || !v_off || !A || !tau || !z_off || !col0 || !ncol || !Z)))
    return fail(h, MH_ERR_ARG, "mh_ormtr_batched: null argument");
  int nmax = 0;
  for (int m = 0; m < nmat; ++m) {
    if (n[m] < 1 || n[m] > SY_MAXN || col0[m] < 0 || ncol[m] < 0 || col0[m] + ncol[m] > n[m])
      return fail(h, MH_ERR_SHAPE, "mh_ormtr_batched: matrix %d: n %d, columns [%d, %d)", m, n[m],
                  col0[m], col0[m] + ncol[m]);
    nmax = std::max(nmax, n[m]);
  }
  MH_HIP(h, hipSetDevice(h->device));
  size_t wy_max = 0;
  MH_TRY(wy_lds_limit(h, &wy_max));
  if ((size_t)WY_CT * nmax * sizeof(double) <= wy_max && !getenv("MH_ORMTR_COLS"))
    return ormtr_wy(h, nmat, n, a_off, v_off, A, tau, z_off, col0, ncol, Z);
  const size_t lds = (size_t)OR_WAVES * nmax * sizeof(double);
  int m = 0;
  while (m < nmat) {
    OrTable tab{};
    int w = 0;
    while (m < nmat && tab.nt < SY_MAXT) {
      if (ncol[m] > 0) {
        OrTask &tk = tab.t[tab.nt++];
        tk.a_off = a_off[m];
        tk.v_off = v_off[m];
        tk.z_off = z_off[m];
        tk.n = n[m];
        tk.col0 = col0[m];
        tk.ncol = ncol[m];
        tk.w0 = w;
        w += (ncol[m] + OR_WAVES - 1) / OR_WAVES;
      }
      ++m;
    }
    if (w == 0) continue;
    hipLaunchKernelGGL(k_ormtr_cols, dim3(w), dim3(OR_THREADS), lds, h->stream, tab, A, tau, Z);
    MH_TRY(check_launch(h, "k_ormtr_cols"));
  }
  return MH_OK;
}

extern "C" int mh_eig_check(mh_handle h, int nmat) {
  if (!h || nmat < 0) return fail(h, MH_ERR_ARG, "mh_eig_check: bad argument");
  MH_HIP(h, hipSetDevice(h->device));
  MH_HIP(h, hipStreamSynchronize(h->stream));
  EigState *st = state(h);
  if (h->eig_err) {
    int err = 0;
    MH_HIP(h, hipMemcpy(&err, h->eig_err, sizeof(int), hipMemcpyDeviceToHost));
    if (err) return fail(h, MH_ERR_HIP, "mh_sytrd_batched: a wait between workgroups timed out");
  }
  if (int *dce = dc_error_flag(h)) {
    int err = 0;
    MH_HIP(h, hipMemcpy(&err, dce, sizeof(int), hipMemcpyDeviceToHost));
    if (err)
      return fail(h, MH_ERR_NOCONV,
                  "mh_stedc_dc_batched: the QL iteration of a leaf did not converge");
  }
  if (nmat > 0 && st->info) {
    std::vector<int> info(std::min(nmat, st->info_len));
    MH_HIP(h, hipMemcpy(info.data(), st->info, info.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (size_t m = 0; m < info.size(); ++m)
      if (info[m] != 0)
        return fail(h, MH_ERR_NOCONV, "rocsolver_dstedc: matrix %zu did not converge (info %d)", m,
                    info[m]);
  }
  return MH_OK;
}
