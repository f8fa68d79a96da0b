// Batched divide-and-conquer eigensolver for symmetric tridiagonal matrices:
// the middle stage of the device eigensolver behind the graded two-site split
// (Householder tridiagonalisation -> this -> back-transformation), replacing
// one library dstedc call per sector by a few launches for all sectors.
//
// Method (Cuppen's divide and conquer with the Gu-Eisenstat eigenvectors, as
// in Demmel, "Applied Numerical Linear Algebra", sec. 5.3.3, and Gu &
// Eisenstat, SIAM J. Matrix Anal. Appl. 16 (1995) 172):
//   T = diag(T1', T2') + |beta| u u^T,   u = e_k + sign(beta) e_{k+1},
// where beta = T[k, k+1] couples the halves and T1', T2' have |beta| taken
// off the touching diagonal entries.  Leaves (order <= DC_LEAF) are solved by
// the implicit QL iteration with Wilkinson shifts (tql2).  A merge of two
// solved halves Q1 L1 Q1^T, Q2 L2 Q2^T works on D + rho z z^T with
// D = diag(L1, L2) sorted, z = [last row of Q1; sign(beta) first row of Q2]/sqrt2,
// rho = 2|beta|:
//   deflation (rho |z_j| <= tol, or a Givens rotation of two close d's that
//   zeroes one z component), tol = 8 eps max(|D|max, rho);
//   secular roots 1/rho + sum_j z_j^2/(d_j - lambda) = 0, each computed as
//   lambda = d_origin + tau from the nearer pole (so d_j - lambda keeps its
//   relative accuracy), by a bracketed iteration on a two-pole osculatory
//   model of the secular function (Li, "Solving secular equations stably and
//   efficiently", 1994);
//   zhat_i^2 = prod_j (lambda_j - d_i) / (rho prod_{j != i} (d_j - d_i)), the
//   z of which the computed roots are the exact eigenvalues, and eigenvectors
//   (zhat_i / (d_i - lambda_j))_i normalised: orthogonal to working precision
//   without reorthogonalisation;
//   new eigenvectors = [Q1 0; 0 Q2] (permuted, rotated) x U: one GEMM.
// All nodes of one tree height across all matrices run in the same launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mh_internal.h"

namespace mh {

namespace {

constexpr int DC_LEAF = 16;       // largest leaf order
constexpr int DC_MAXN = 4096;     // largest matrix order
constexpr int DC_QL_MAXIT = 60;   // QL iterations per eigenvalue of a leaf
constexpr int DC_ROOT_MAXIT = 100;
constexpr double DC_EPS = 2.220446049250313e-16;

struct DcLeaf {
  int64_t z_off;  // Z base of the matrix
  int64_t v_off;  // D / E base of the matrix
  int ld;         // matrix order (leading dimension of Z)
  int off, sz;    // rows/columns [off, off + sz) of the matrix
  int pad;
};

struct DcNode {
  int64_t z_off, v_off;
  int64_t q_off;  // node's work region in the gather / U buffers (n^2 doubles)
  int64_t e_off;  // node's range in the per-element work arrays
  int ld;
  int off, n1, n2;
  int id;         // index into the per-node info arrays
  int pad;
};

struct DcWork {
  double *Q2, *U;                                   // gathered eigenvectors, merge eigenvectors
  double *dk, *zk, *tau, *lamk, *zh, *dval;         // per element
  int *org, *gcol, *dest, *rota, *rotb;             // per element
  double *rotc, *rots;                              // per element (rotations of a node)
  int *nK, *nRot;                                   // per node
  double *nRho;                                     // per node
};

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// ---- leaves: implicit QL (tql2) on an order <= 32 tridiagonal -------------
// One wavefront per leaf.  Lane r holds d[r] and e[r] (e[r] couples r and
// r + 1) in registers; the scalar recurrence reads them with readlane (the
// index is uniform) and writes them back with a lane-select, so the whole
// wavefront runs the same scalar code without touching LDS.  Lane r keeps
// row r of the eigenvector matrix in LDS and applies every plane rotation to
// it.  The leaf also clears its row band of Z (the columns outside its
// block), so that Z is block diagonal for the merges above it.
__device__ inline double rdl(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__global__ __launch_bounds__(64) void k_dc_leaf(const DcLeaf *__restrict__ lv,
                                                double *__restrict__ D,
                                                const double *__restrict__ E,
                                                double *__restrict__ Z, int *__restrict__ err) {
  __shared__ double sz[DC_LEAF][DC_LEAF + 1];
  __shared__ int srank[DC_LEAF];
  const DcLeaf L = lv[blockIdx.x];
  const int n = L.sz, r = threadIdx.x;
  const double *e = E + L.v_off;
  double *d = D + L.v_off;
  double dr = 0.0, er = 0.0;
  if (r < n) {
    dr = d[L.off + r];
    // the tears of the merges above: |beta| off both touching diagonals
    if (r == 0 && L.off > 0) dr -= fabs(e[L.off - 1]);
    if (r == n - 1 && L.off + n < L.ld) dr -= fabs(e[L.off + n - 1]);
    er = (r < n - 1) ? e[L.off + r] : 0.0;
    for (int c = 0; c < n; ++c) sz[r][c] = (r == c) ? 1.0 : 0.0;
  }
  bool bad = false;
  for (int l = 0; l < n && !bad; ++l) {
    int iter = 0, m;
    do {
      // m: first index >= l with a negligible coupling e[m] (n - 1 if none)
      const double dn = __shfl_down(dr, 1);
      const bool small = r >= l && r < n - 1 && fabs(er) <= DC_EPS * (fabs(dr) + fabs(dn));
      const unsigned long long mask = __ballot(small);
      m = mask ? (int)__ffsll((long long)mask) - 1 : n - 1;
      if (m != l) {
        if (iter++ == DC_QL_MAXIT) {
          bad = true;
          break;
        }
        const double dl = rdl(dr, l), el = rdl(er, l);
        double g = (rdl(dr, l + 1) - dl) / (2.0 * el);
        double rr = hypot(g, 1.0);
        g = rdl(dr, m) - dl + el / (g + copysign(rr, g));
        double s = 1.0, c = 1.0, p = 0.0;
        bool underflow = false;
        for (int i = m - 1; i >= l; --i) {
          const double ei = rdl(er, i);
          const double f = s * ei, b = c * ei;
          rr = hypot(f, g);
          er = (r == i + 1) ? rr : er;
          if (rr == 0.0) {
            dr = (r == i + 1) ? dr - p : dr;
            er = (r == m) ? 0.0 : er;
            underflow = true;
            break;
          }
          s = f / rr;
          c = g / rr;
          g = rdl(dr, i + 1) - p;
          rr = (rdl(dr, i) - g) * s + 2.0 * c * b;
          p = s * rr;
          dr = (r == i + 1) ? g + p : dr;
          g = c * rr - b;
          if (r < n) {
            const double zf = sz[r][i + 1], zi = sz[r][i];
            sz[r][i + 1] = s * zi + c * zf;
            sz[r][i] = c * zi - s * zf;
          }
        }
        if (underflow) continue;
        dr = (r == l) ? dr - p : dr;
        er = (r == l) ? g : er;
        er = (r == m) ? 0.0 : er;
      }
    } while (m != l);
  }
  if (bad) {
    if (r == 0) *err = 1;
    return;  // uniform: every lane took the same path
  }
  // ascending order: rank of each eigenvalue, ties broken by index
  int k = 0;
  for (int j = 0; j < n; ++j) {
    const double dj = rdl(dr, j);
    k += (dj < dr) || (dj == dr && j < r);
  }
  if (r < n) {
    srank[r] = k;
    d[L.off + k] = dr;
  }
  __syncthreads();
  if (r < n) {
    double *zr = Z + L.z_off + (L.off + r);
    for (int c = 0; c < L.ld; ++c) zr[(int64_t)c * L.ld] = 0.0;
    for (int c = 0; c < n; ++c) zr[(int64_t)(L.off + srank[c]) * L.ld] = sz[r][c];
  }
}

// ---- merge step 1: sort, deflate ------------------------------------------
// One workgroup per node.  The two children's eigenvalues (ascending each)
// are merged by rank; thread 0 runs the deflation scan (it is sequential:
// a Givens deflation carries its surviving index to the next comparison).
constexpr int PREP_THREADS = 256;

// Exclusive prefix sum of one int per thread over the workgroup (wavefront
// scans, then the wavefront totals); *total gets the sum.
__device__ inline int prep_scan(int v, int *wsum, int *total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < PREP_THREADS / 64; ++q) {
    off += (q < wv) ? wsum[q] : 0;
    tot += wsum[q];
  }
  __syncthreads();  // wsum is reused by the next scan
  *total = tot;
  return off + x - v;
}

__global__ __launch_bounds__(PREP_THREADS) void k_dc_prep(const DcNode *__restrict__ tab,
                                                          const double *__restrict__ D,
                                                          const double *__restrict__ E,
                                                          const double *__restrict__ Z,
                                                          DcWork w) {
  extern __shared__ double smem[];
  __shared__ double red[PREP_THREADS / 64];
  __shared__ int wsum[PREP_THREADS / 64];
  const DcNode nd = tab[blockIdx.x];
  const int n1 = nd.n1, n2 = nd.n2, n = n1 + n2, tid = threadIdx.x;
  double *sD = smem;
  double *sZ = smem + n;
  int *sC = (int *)(smem + 2 * n);
  int *st = sC + n;
  int *sL = st + n;  // positions of the entries with a non-negligible z
  const double *lam = D + nd.v_off + nd.off;
  const double beta = E[nd.v_off + nd.off + n1 - 1];
  const double sgn = beta >= 0.0 ? 1.0 : -1.0;
  const double rho = 2.0 * fabs(beta);
  const int64_t ld = nd.ld;
  const double *z1 = Z + nd.z_off + (nd.off + n1 - 1);
  const double *z2 = Z + nd.z_off + (nd.off + n1);
  const double rs2 = 0.70710678118654752440;
  double mx = 0.0;
  for (int t = tid; t < n; t += PREP_THREADS) {
    const double v = lam[t];
    int pos;
    double zv;
    if (t < n1) {
      int lo = 0, hi = n2;  // count of right values < v
      while (lo < hi) {
        const int md = (lo + hi) >> 1;
        if (lam[n1 + md] < v) lo = md + 1; else hi = md;
      }
      pos = t + lo;
      zv = z1[(nd.off + t) * ld] * rs2;
    } else {
      int lo = 0, hi = n1;  // count of left values <= v
      while (lo < hi) {
        const int md = (lo + hi) >> 1;
        if (lam[md] <= v) lo = md + 1; else hi = md;
      }
      pos = (t - n1) + lo;
      zv = sgn * z2[(nd.off + t) * ld] * rs2;
    }
    sD[pos] = v;
    sZ[pos] = zv;
    sC[pos] = t;
    mx = fmax(mx, fabs(v));
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
  if ((tid & 63) == 0) red[tid >> 6] = mx;
  __syncthreads();
  double dmax = 0.0;
  for (int i = 0; i < PREP_THREADS / 64; ++i) dmax = fmax(dmax, red[i]);
  const double tol = 8.0 * DC_EPS * fmax(dmax, rho);
  const int64_t e0 = nd.e_off;
  // every thread owns a contiguous chunk of positions
  const int per = (n + PREP_THREADS - 1) / PREP_THREADS;
  const int j0 = min(n, tid * per), j1 = min(n, j0 + per);
  // (a) entries with a negligible z deflate at once (no rotation ever
  // touches them first); the others are listed in position order
  int cnt = 0;
  for (int j = j0; j < j1; ++j) {
    st[j] = 0;
    cnt += (rho * fabs(sZ[j]) > tol) ? 1 : 0;
  }
  int m = 0;
  int q = prep_scan(cnt, wsum, &m);
  for (int j = j0; j < j1; ++j)
    if (rho * fabs(sZ[j]) > tol) sL[q++] = j;
  __syncthreads();
  // (b) the Givens deflation of close neighbours: a chain along the listed
  // entries, so one thread walks it (the graded matrices list few entries at
  // the upper merges)
  if (tid == 0) {
    int nrot = 0, pj = -1;
    for (int k = 0; k < m; ++k) {
      const int j = sL[k];
      if (pj < 0) {
        pj = j;
        continue;
      }
      const double zp = sZ[pj], zj = sZ[j];
      const double tz = hypot(zj, zp);
      const double c = zj / tz, s = -zp / tz;
      const double t = sD[j] - sD[pj];
      if (fabs(t * c * s) <= tol) {
        sZ[j] = tz;
        sZ[pj] = 0.0;
        w.rota[e0 + nrot] = sC[pj];
        w.rotb[e0 + nrot] = sC[j];
        w.rotc[e0 + nrot] = c;
        w.rots[e0 + nrot] = s;
        ++nrot;
        const double dp = sD[pj] * c * c + sD[j] * s * s;
        sD[j] = sD[pj] * s * s + sD[j] * c * c;
        sD[pj] = dp;
        st[pj] = 0;
      } else {
        st[pj] = 1;
      }
      pj = j;
    }
    if (pj >= 0) st[pj] = 1;
    w.nRot[nd.id] = nrot;
    w.nRho[nd.id] = rho;
  }
  __syncthreads();
  // (c) Q2 column order: kept (ascending d), then deflated (ascending
  // position)
  int kc = 0;
  for (int j = j0; j < j1; ++j) kc += st[j];
  int K = 0;
  int ck = prep_scan(kc, wsum, &K);
  int cd = K + (j0 - ck);  // deflated entries before this chunk follow the kept ones
  for (int j = j0; j < j1; ++j) {
    const int c = st[j] ? ck++ : cd++;
    w.gcol[e0 + c] = sC[j];
    if (st[j]) {
      w.dk[e0 + c] = sD[j];
      w.zk[e0 + c] = sZ[j];
    } else {
      w.dval[e0 + c] = sD[j];
    }
  }
  if (tid == 0) w.nK[nd.id] = K;
}

// ---- merge step 2: Givens rotations of the deflation on Z's columns -------
__global__ __launch_bounds__(256) void k_dc_rot(const DcNode *__restrict__ tab,
                                                double *__restrict__ Z, DcWork w) {
  const DcNode nd = tab[blockIdx.y];
  const int nrot = w.nRot[nd.id];
  const int n = nd.n1 + nd.n2;
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (nrot == 0 || r >= n) return;
  const int64_t ld = nd.ld;
  double *zr = Z + nd.z_off + (nd.off + r) + (int64_t)nd.off * ld;
  for (int t = 0; t < nrot; ++t) {
    const int64_t a = w.rota[nd.e_off + t], b = w.rotb[nd.e_off + t];
    const double c = w.rotc[nd.e_off + t], s = w.rots[nd.e_off + t];
    const double x = zr[a * ld], y = zr[b * ld];
    zr[a * ld] = c * x + s * y;
    zr[b * ld] = c * y - s * x;
  }
}

// ---- merge step 3: gather the node's columns (kept, then deflated) --------
__global__ __launch_bounds__(256) void k_dc_gather(const DcNode *__restrict__ tab,
                                                   const double *__restrict__ Z, DcWork w) {
  const DcNode nd = tab[blockIdx.y];
  const int n = nd.n1 + nd.n2;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (c >= n) return;
  const int64_t ld = nd.ld;
  const double *src = Z + nd.z_off + nd.off + (int64_t)(nd.off + w.gcol[nd.e_off + c]) * ld;
  double *dst = w.Q2 + nd.q_off + (int64_t)c * n;
  for (int r = lane; r < n; r += 64) dst[r] = src[r];
}

// ---- merge step 4: secular roots, one wavefront per root ------------------
__global__ __launch_bounds__(256) void k_dc_roots(const DcNode *__restrict__ tab, DcWork w) {
  const DcNode nd = tab[blockIdx.y];
  const int K = w.nK[nd.id];
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= K) return;
  const double rho = w.nRho[nd.id];
  const double *dk = w.dk + nd.e_off, *zk = w.zk + nd.e_off;
  int o, a, b;
  double lo, hi;
  if (K == 1) {
    if (lane == 0) {
      w.org[nd.e_off] = 0;
      w.tau[nd.e_off] = rho * zk[0] * zk[0];
      w.lamk[nd.e_off] = dk[0] + rho * zk[0] * zk[0];
    }
    return;
  }
  if (i < K - 1) {
    const double mid = 0.5 * (dk[i + 1] - dk[i]);
    double s = 0.0;
    for (int j = lane; j < K; j += 64) s += zk[j] * zk[j] / ((dk[j] - dk[i]) - mid);
    const double g = 1.0 / rho + wave_sum(s);
    if (g >= 0.0) {
      o = i;
      lo = 0.0;
      hi = mid;
    } else {
      o = i + 1;
      lo = -mid;
      hi = 0.0;
    }
    a = i;
    b = i + 1;
  } else {
    double s = 0.0;
    for (int j = lane; j < K; j += 64) s += zk[j] * zk[j];
    o = K - 1;
    lo = 0.0;
    hi = rho * wave_sum(s);
    a = K - 1;
    b = K;
  }
  const double dorg = dk[o];
  double tau = 0.5 * (lo + hi);
  for (int it = 0; it < DC_ROOT_MAXIT; ++it) {
    double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0;
    for (int j = lane; j < K; j += 64) {
      const double den = (dk[j] - dorg) - tau;
      const double t = zk[j] * zk[j] / den;
      if (j <= a) {
        psi += t;
        dpsi += t / den;
      } else {
        phi += t;
        dphi += t / den;
      }
    }
    psi = wave_sum(psi);
    dpsi = wave_sum(dpsi);
    phi = wave_sum(phi);
    dphi = wave_sum(dphi);
    const double g = 1.0 / rho + psi + phi;
    const double errb = 8.0 * (1.0 / rho + fabs(psi) + phi) + fabs(tau) * (dpsi + dphi);
    if (fabs(g) <= DC_EPS * errb) break;
    if (g < 0.0) lo = tau; else hi = tau;
    // osculatory model: psi ~ s/(da - t) + S, phi ~ p/(db - t) + P
    const double da = dk[a] - dorg;
    const double sa = dpsi * (da - tau) * (da - tau);
    double cc = 1.0 / rho + (psi - sa / (da - tau));
    double nt = 0.5 * (lo + hi);
    bool have = false;
    if (b < K) {
      const double db = dk[b] - dorg;
      const double pb = dphi * (db - tau) * (db - tau);
      cc += phi - pb / (db - tau);
      // cc (da - t)(db - t) + sa (db - t) + pb (da - t) = 0
      const double A = cc, B = -(cc * (da + db) + sa + pb), C = cc * da * db + sa * db + pb * da;
      if (A == 0.0) {
        if (B != 0.0) {
          nt = -C / B;
          have = true;
        }
      } else {
        const double disc = B * B - 4.0 * A * C;
        if (disc >= 0.0) {
          const double q = -0.5 * (B + copysign(sqrt(disc), B));
          const double r1 = q / A, r2 = (q != 0.0) ? C / q : r1;
          const bool ok1 = lo < r1 && r1 < hi, ok2 = lo < r2 && r2 < hi;
          if (ok1 && ok2) {
            nt = (fabs(r1 - tau) <= fabs(r2 - tau)) ? r1 : r2;
            have = true;
          } else if (ok1) {
            nt = r1;
            have = true;
          } else if (ok2) {
            nt = r2;
            have = true;
          }
        }
      }
    } else {
      cc += phi;
      if (cc != 0.0) {
        nt = da + sa / cc;
        have = true;
      }
    }
    if (!have || !(lo < nt && nt < hi)) nt = 0.5 * (lo + hi);
    if (nt == tau || hi - lo <= 2.0 * DC_EPS * fmax(fabs(lo), fabs(hi))) {
      tau = nt;
      break;
    }
    tau = nt;
  }
  if (lane == 0) {
    w.org[nd.e_off + i] = o;
    w.tau[nd.e_off + i] = tau;
    w.lamk[nd.e_off + i] = dorg + tau;
  }
}

// ---- merge step 5: Gu-Eisenstat z, one thread per component ---------------
__global__ __launch_bounds__(256) void k_dc_zhat(const DcNode *__restrict__ tab, DcWork w) {
  const DcNode nd = tab[blockIdx.y];
  const int K = w.nK[nd.id];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= K) return;
  const double rho = w.nRho[nd.id];
  const double *dk = w.dk + nd.e_off;
  const int *org = w.org + nd.e_off;
  const double *tau = w.tau + nd.e_off;
  const double di = dk[i];
  // lambda_j - d_i = (d_org(j) - d_i) + tau_j, paired with d_j - d_i (j < i)
  // or d_{j+1} - d_i (j >= i): every factor is a positive ratio (interlacing)
  double pr = ((dk[org[K - 1]] - di) + tau[K - 1]) / rho;
  for (int j = 0; j < K - 1; ++j) {
    const double lmd = (dk[org[j]] - di) + tau[j];
    pr *= lmd / ((j < i) ? (dk[j] - di) : (dk[j + 1] - di));
  }
  w.zh[nd.e_off + i] = copysign(sqrt(fmax(pr, 0.0)), w.zk[nd.e_off + i]);
}

// ---- merge step 6: eigenvectors of D + rho z z^T, one wavefront per column
__global__ __launch_bounds__(256) void k_dc_vecs(const DcNode *__restrict__ tab, DcWork w) {
  const DcNode nd = tab[blockIdx.y];
  const int K = w.nK[nd.id];
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (j >= K) return;
  const double *dk = w.dk + nd.e_off, *zh = w.zh + nd.e_off;
  const double dorg = dk[w.org[nd.e_off + j]], t = w.tau[nd.e_off + j];
  double s = 0.0;
  for (int i = lane; i < K; i += 64) {
    const double v = zh[i] / ((dk[i] - dorg) - t);
    s += v * v;
  }
  const double inv = 1.0 / sqrt(wave_sum(s));
  double *u = w.U + nd.q_off + (int64_t)j * K;
  for (int i = lane; i < K; i += 64) u[i] = inv * zh[i] / ((dk[i] - dorg) - t);
}

// ---- merge step 7: final order of the node's eigenvalues ------------------
// Bitonic sort of (value, output column) in LDS; D gets the ascending values,
// dest[c] the position of column c.
__global__ __launch_bounds__(256) void k_dc_order(const DcNode *__restrict__ tab,
                                                  double *__restrict__ D, DcWork w) {
  extern __shared__ double smem[];
  const DcNode nd = tab[blockIdx.x];
  const int n = nd.n1 + nd.n2, K = w.nK[nd.id];
  int N2 = 1;
  while (N2 < n) N2 <<= 1;
  double *kv = smem;
  int *ki = (int *)(smem + N2);
  for (int t = threadIdx.x; t < N2; t += 256) {
    kv[t] = (t < n) ? ((t < K) ? w.lamk[nd.e_off + t] : w.dval[nd.e_off + t]) : INFINITY;
    ki[t] = t;
  }
  __syncthreads();
  for (int k = 2; k <= N2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < N2; t += 256) {
        const int p = t ^ j;
        if (p > t) {
          const bool up = (t & k) == 0;
          const double a = kv[t], b = kv[p];
          const int ia = ki[t], ib = ki[p];
          const bool gt = (a > b) || (a == b && ia > ib);
          if (gt == up) {
            kv[t] = b;
            kv[p] = a;
            ki[t] = ib;
            ki[p] = ia;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int t = threadIdx.x; t < n; t += 256) {
    D[nd.v_off + nd.off + t] = kv[t];
    if (ki[t] < n) w.dest[nd.e_off + ki[t]] = t;  // (a NaN value could leave a pad here)
  }
}

// ---- merge step 8: Z(node, kept) = Q2(:, 0:K) U, columns to their places --
// 64 x 64 output tile per workgroup, 4 x 4 per thread, K panels of 16 in LDS.
__global__ __launch_bounds__(256) void k_dc_gemm(const DcNode *__restrict__ tab,
                                                 double *__restrict__ Z, DcWork w) {
  __shared__ double As[16][64 + 1];
  __shared__ double Bs[16][64 + 1];
  const DcNode nd = tab[blockIdx.z];
  const int n = nd.n1 + nd.n2, K = w.nK[nd.id];
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  if (r0 >= n || c0 >= K) return;
  const double *A = w.Q2 + nd.q_off;  // n x K used, ld n
  const double *B = w.U + nd.q_off;   // K x K, ld K
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += 16) {
    for (int t = threadIdx.x; t < 16 * 64; t += 256) {
      const int kk = t >> 6, rr = t & 63;
      const int gk = k0 + kk;
      As[kk][rr] = (gk < K && r0 + rr < n) ? A[(int64_t)gk * n + r0 + rr] : 0.0;
    }
    for (int t = threadIdx.x; t < 16 * 64; t += 256) {
      const int cc = t >> 4, kk = t & 15;
      const int gk = k0 + kk;
      Bs[kk][cc] = (gk < K && c0 + cc < K) ? B[(int64_t)(c0 + cc) * K + gk] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = As[kk][tx + 16 * q];
        b[q] = Bs[kk][ty + 16 * q];
      }
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[p][q] = fma(a[p], b[q], acc[p][q]);
    }
    __syncthreads();
  }
  const int64_t ld = nd.ld;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int c = c0 + ty + 16 * q;
    if (c >= K) continue;
    double *zc = Z + nd.z_off + nd.off + (int64_t)(nd.off + w.dest[nd.e_off + c]) * ld;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int r = r0 + tx + 16 * p;
      if (r < n) zc[r] = acc[p][q];
    }
  }
}

// ---- merge step 9: deflated columns to their places -----------------------
__global__ __launch_bounds__(256) void k_dc_defl(const DcNode *__restrict__ tab,
                                                 double *__restrict__ Z, DcWork w) {
  const DcNode nd = tab[blockIdx.y];
  const int n = nd.n1 + nd.n2, K = w.nK[nd.id];
  const int c = K + blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (c >= n) return;
  const int64_t ld = nd.ld;
  const double *src = w.Q2 + nd.q_off + (int64_t)c * n;
  double *dst = Z + nd.z_off + nd.off + (int64_t)(nd.off + w.dest[nd.e_off + c]) * ld;
  for (int r = lane; r < n; r += 64) dst[r] = src[r];
}

struct DcState {
  void *buf = nullptr;
  size_t bytes = 0;
  int *err = nullptr;  // leaf QL failure flag (device)
};

DcState *dc_state(mh_handle h) {
  if (!h->dc_state) h->dc_state = new DcState();
  return (DcState *)h->dc_state;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

void dc_release(mh_handle h) {
  DcState *s = (DcState *)h->dc_state;
  if (s) {
    if (s->buf) (void)hipFree(s->buf);
    delete s;
  }
  h->dc_state = nullptr;
}

int *dc_error_flag(mh_handle h) {
  DcState *s = (DcState *)h->dc_state;
  return s ? s->err : nullptr;
}

}  // namespace mh

using namespace mh;

extern "C" int mh_stedc_dc_batched(mh_handle h, int nmat, const int *n, const int64_t *v_off,
                                   const int64_t *z_off, double *D, const double *E, double *Z) {
  if (!h || nmat < 0 || (nmat > 0 && (!n || !v_off || !z_off || !D || !E || !Z)))
    return fail(h, MH_ERR_ARG, "mh_stedc_dc_batched: null argument");
  if (nmat == 0) return MH_OK;
  for (int m = 0; m < nmat; ++m) {
    if (n[m] < 1 || n[m] > DC_MAXN)
      return fail(h, MH_ERR_SHAPE, "mh_stedc_dc_batched: order %d of matrix %d outside [1, %d]",
                  n[m], m, DC_MAXN);
    if (v_off[m] < 0 || z_off[m] < 0)
      return fail(h, MH_ERR_SHAPE, "mh_stedc_dc_batched: negative offset (matrix %d)", m);
  }
  MH_HIP(h, hipSetDevice(h->device));
  // the tree of every matrix: halves until the order is at most DC_LEAF;
  // internal nodes grouped by height (leaves 0)
  std::vector<DcLeaf> leaves;
  std::vector<std::vector<DcNode>> byh;
  int64_t qsum = 0, esum = 0;
  int nodes = 0;
  for (int m = 0; m < nmat; ++m) {
    const int nm = n[m];
    const int64_t qb = qsum, eb = esum;
    qsum += (int64_t)nm * nm;
    esum += nm;
    // iterative post-order: (off, sz) -> height
    struct Fr {
      int off, sz, state, hl;
    };
    std::vector<Fr> stk;
    stk.push_back({0, nm, 0, 0});
    int last_h = 0;
    while (!stk.empty()) {
      Fr &f = stk.back();
      if (f.sz <= DC_LEAF) {
        leaves.push_back({z_off[m], v_off[m], nm, f.off, f.sz, 0});
        last_h = 0;
        stk.pop_back();
        continue;
      }
      const int h1 = f.sz / 2;
      if (f.state == 0) {
        f.state = 1;
        stk.push_back({f.off, h1, 0, 0});
      } else if (f.state == 1) {
        f.state = 2;
        f.hl = last_h;
        const int off = f.off, sz = f.sz;
        stk.push_back({off + h1, sz - h1, 0, 0});
      } else {
        const int hgt = 1 + std::max(f.hl, last_h);
        if ((int)byh.size() < hgt) byh.resize(hgt);
        DcNode nd{};
        nd.z_off = z_off[m];
        nd.v_off = v_off[m];
        nd.q_off = qb + (int64_t)f.off * nm;
        nd.e_off = eb + f.off;
        nd.ld = nm;
        nd.off = f.off;
        nd.n1 = h1;
        nd.n2 = f.sz - h1;
        nd.id = nodes++;
        byh[hgt - 1].push_back(nd);
        last_h = hgt;
        stk.pop_back();
      }
    }
  }
  // one table: leaves, then the nodes of each height
  std::vector<int64_t> hstart(byh.size());
  std::vector<int> hmax(byh.size(), 0);
  size_t tab_bytes = align256(leaves.size() * sizeof(DcLeaf));
  for (size_t k = 0; k < byh.size(); ++k) {
    hstart[k] = (int64_t)tab_bytes;
    tab_bytes += align256(byh[k].size() * sizeof(DcNode));
    for (auto &nd : byh[k]) hmax[k] = std::max(hmax[k], nd.n1 + nd.n2);
  }
  const size_t S1 = (size_t)std::max<int64_t>(esum, 1), S2 = (size_t)std::max<int64_t>(qsum, 1);
  const size_t NN = (size_t)std::max(nodes, 1);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  };
  const size_t o_tab = take(tab_bytes), o_err = take(sizeof(int));
  const size_t o_q2 = take(S2 * 8), o_u = take(S2 * 8);
  const size_t o_d[6] = {take(S1 * 8), take(S1 * 8), take(S1 * 8), take(S1 * 8), take(S1 * 8),
                         take(S1 * 8)};
  const size_t o_i[5] = {take(S1 * 4), take(S1 * 4), take(S1 * 4), take(S1 * 4), take(S1 * 4)};
  const size_t o_rc = take(S1 * 8), o_rs = take(S1 * 8);
  const size_t o_nk = take(NN * 4), o_nr = take(NN * 4), o_rho = take(NN * 8);
  DcState *st = dc_state(h);
  if (st->bytes < off) {
    MH_HIP(h, hipStreamSynchronize(h->stream));
    if (st->buf) MH_HIP(h, hipFree(st->buf));
    st->buf = nullptr;
    st->bytes = 0;
    st->err = nullptr;
    hipError_t e = hipMalloc(&st->buf, off);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      st->buf = nullptr;
      return fail(h, MH_ERR_OOM, "mh_stedc_dc_batched: work buffer of %zu bytes: %s", off,
                  hipGetErrorString(e));
    }
    st->bytes = off;
  }
  char *base = (char *)st->buf;
  st->err = (int *)(base + o_err);
  std::vector<char> host(tab_bytes, 0);
  memcpy(host.data(), leaves.data(), leaves.size() * sizeof(DcLeaf));
  for (size_t k = 0; k < byh.size(); ++k)
    memcpy(host.data() + hstart[k], byh[k].data(), byh[k].size() * sizeof(DcNode));
  MH_HIP(h, hipMemcpyAsync(base + o_tab, host.data(), tab_bytes, hipMemcpyHostToDevice, h->stream));
  MH_HIP(h, hipMemsetAsync(st->err, 0, sizeof(int), h->stream));
  DcWork w;
  w.Q2 = (double *)(base + o_q2);
  w.U = (double *)(base + o_u);
  w.dk = (double *)(base + o_d[0]);
  w.zk = (double *)(base + o_d[1]);
  w.tau = (double *)(base + o_d[2]);
  w.lamk = (double *)(base + o_d[3]);
  w.zh = (double *)(base + o_d[4]);
  w.dval = (double *)(base + o_d[5]);
  w.org = (int *)(base + o_i[0]);
  w.gcol = (int *)(base + o_i[1]);
  w.dest = (int *)(base + o_i[2]);
  w.rota = (int *)(base + o_i[3]);
  w.rotb = (int *)(base + o_i[4]);
  w.rotc = (double *)(base + o_rc);
  w.rots = (double *)(base + o_rs);
  w.nK = (int *)(base + o_nk);
  w.nRot = (int *)(base + o_nr);
  w.nRho = (double *)(base + o_rho);
  const DcLeaf *dleaves = (const DcLeaf *)(base + o_tab);
  hipLaunchKernelGGL(k_dc_leaf, dim3((unsigned)leaves.size()), dim3(64), 0, h->stream, dleaves, D,
                     E, Z, st->err);
  MH_TRY(check_launch(h, "k_dc_leaf"));
  for (size_t k = 0; k < byh.size(); ++k) {
    const unsigned nn = (unsigned)byh[k].size();
    if (nn == 0) continue;
    const DcNode *tab = (const DcNode *)(base + o_tab + hstart[k]);
    const int nm = hmax[k];
    const size_t prep_lds = (size_t)nm * (2 * sizeof(double) + 3 * sizeof(int));
    int n2p = 1;
    while (n2p < nm) n2p <<= 1;
    const size_t ord_lds = (size_t)n2p * (sizeof(double) + sizeof(int));
    hipLaunchKernelGGL(k_dc_prep, dim3(nn), dim3(PREP_THREADS), prep_lds, h->stream, tab, D, E, Z,
                       w);
    MH_TRY(check_launch(h, "k_dc_prep"));
    hipLaunchKernelGGL(k_dc_rot, dim3((nm + 255) / 256, nn), dim3(256), 0, h->stream, tab, Z, w);
    MH_TRY(check_launch(h, "k_dc_rot"));
    hipLaunchKernelGGL(k_dc_gather, dim3((nm + 3) / 4, nn), dim3(256), 0, h->stream, tab, Z, w);
    MH_TRY(check_launch(h, "k_dc_gather"));
    hipLaunchKernelGGL(k_dc_roots, dim3((nm + 3) / 4, nn), dim3(256), 0, h->stream, tab, w);
    MH_TRY(check_launch(h, "k_dc_roots"));
    hipLaunchKernelGGL(k_dc_zhat, dim3((nm + 255) / 256, nn), dim3(256), 0, h->stream, tab, w);
    MH_TRY(check_launch(h, "k_dc_zhat"));
    hipLaunchKernelGGL(k_dc_vecs, dim3((nm + 3) / 4, nn), dim3(256), 0, h->stream, tab, w);
    MH_TRY(check_launch(h, "k_dc_vecs"));
    hipLaunchKernelGGL(k_dc_order, dim3(nn), dim3(256), ord_lds, h->stream, tab, D, w);
    MH_TRY(check_launch(h, "k_dc_order"));
    hipLaunchKernelGGL(k_dc_gemm, dim3((nm + 63) / 64, (nm + 63) / 64, nn), dim3(256), 0, h->stream,
                       tab, Z, w);
    MH_TRY(check_launch(h, "k_dc_gemm"));
    hipLaunchKernelGGL(k_dc_defl, dim3((nm + 3) / 4, nn), dim3(256), 0, h->stream, tab, Z, w);
    MH_TRY(check_launch(h, "k_dc_defl"));
  }
  return MH_OK;
}
