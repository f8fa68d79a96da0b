// Schur-form MPO site tensors and vector-valued environments at the boundary.
//
// The reference's MPOHamiltonian site tensor (SchurMPOTensor,
// algorithms/derivatives.jl:7-17,24-49; mpohamiltonian/transfer.jl:48-75) is
// a wl x wr matrix whose entries are either an operator O[s', s] or a
// scalar c (c times the identity), and its environments are Vectors of one
// (D x D) matrix per MPO channel.  The engine's kernels take the dense site
// tensor W[wl, s', wr, s] and one environment array E[a, w, a'] instead:
// mh_schur_fold builds W from the entries (sum over the terms of one site,
// identity entries included, which is the same operator as the reference's
// opkeys + scalkeys sums), and mh_env_gather / mh_env_scatter move between
// the Vector of channel matrices and E on the device.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "mh_internal.h"

using namespace mh;

extern "C" int mh_schur_fold(int dtype, int wl, int wr, int d, int nterms, const int *ti,
                             const int *tj, const int *kind, const double *coef,
                             const void *ops, void *W) {
  if (dtype != MH_F64 && dtype != MH_C64) return MH_ERR_DTYPE;
  if (wl <= 0 || wr <= 0 || d <= 0 || nterms < 0) return MH_ERR_SHAPE;
  if (!W || (nterms > 0 && (!ti || !tj || !kind || !coef))) return MH_ERR_ARG;
  const int ew = dtype == MH_C64 ? 2 : 1;
  double *w = (double *)W;
  const size_t nw = (size_t)wl * d * wr * d * ew;
  std::memset(w, 0, nw * sizeof(double));
  const double *op = (const double *)ops;
  size_t nop = 0;  // operator entries consume consecutive d x d matrices
  for (int t = 0; t < nterms; ++t) {
    const int i = ti[t], j = tj[t];
    if (i < 0 || i >= wl || j < 0 || j >= wr) return MH_ERR_SHAPE;
    if (kind[t] != 0 && kind[t] != 1) return MH_ERR_ARG;
    if (kind[t] == 1 && !op) return MH_ERR_ARG;
    const double cr = coef[2 * t], ci = coef[2 * t + 1];
    if (dtype == MH_F64 && ci != 0.0) return MH_ERR_DTYPE;
    for (int s = 0; s < d; ++s)      // s_in
      for (int sp = 0; sp < d; ++sp) {  // s_out
        double vr, vi = 0.0;
        if (kind[t] == 0) {
          if (sp != s) continue;
          vr = cr;
          vi = ci;
        } else {
          const size_t o = (nop * d * d + sp + (size_t)d * s) * ew;
          const double orr = op[o], oi = dtype == MH_C64 ? op[o + 1] : 0.0;
          vr = cr * orr - ci * oi;
          vi = cr * oi + ci * orr;
        }
        // W[wl, s_out, wr, s_in], column-major
        const size_t e = ((size_t)i + (size_t)wl * (sp + (size_t)d * (j + (size_t)wr * s))) * ew;
        w[e] += vr;
        if (dtype == MH_C64) w[e + 1] += vi;
      }
    if (kind[t] == 1) ++nop;
  }
  return MH_OK;
}

// E[a, w, a'] (extents D1, nw, D2) <-> parts[w] (D1 x D2, column-major)
static int env_copy(mh_handle h, int dtype, int D1, int D2, int nw, const void *const *parts,
                    void *const *outs, void *E, const void *Ein) {
  if (!h) return MH_ERR_ARG;
  if (dtype != MH_F64 && dtype != MH_C64) return fail(h, MH_ERR_DTYPE, "dtype %d", dtype);
  if (D1 <= 0 || D2 <= 0 || nw <= 0) return fail(h, MH_ERR_SHAPE, "non-positive extent");
  MH_HIP(h, hipSetDevice(h->device));
  const size_t eb = dtype_bytes(dtype);
  const size_t col = (size_t)D1 * eb;  // one column of a part
  for (int k = 0; k < nw; ++k) {
    if (parts) {
      if (!parts[k] || !E) return fail(h, MH_ERR_ARG, "null part %d or environment", k);
      // column a' of part k -> E[:, k, a'] (pitch D1*nw elements)
      MH_HIP(h, hipMemcpy2DAsync((char *)E + (size_t)k * col, col * nw, parts[k], col, col, D2,
                                 hipMemcpyDeviceToDevice, h->stream));
    } else {
      if (!outs[k] || !Ein) return fail(h, MH_ERR_ARG, "null part %d or environment", k);
      MH_HIP(h, hipMemcpy2DAsync(outs[k], col, (const char *)Ein + (size_t)k * col, col * nw, col,
                                 D2, hipMemcpyDeviceToDevice, h->stream));
    }
  }
  return MH_OK;
}

extern "C" int mh_env_gather(mh_handle h, int dtype, int D1, int D2, int nw,
                             const void *const *parts, void *E) {
  if (!parts) return h ? fail(h, MH_ERR_ARG, "null part list") : MH_ERR_ARG;
  return env_copy(h, dtype, D1, D2, nw, parts, nullptr, E, nullptr);
}

extern "C" int mh_env_scatter(mh_handle h, int dtype, int D1, int D2, int nw, const void *E,
                              void *const *parts) {
  if (!parts) return h ? fail(h, MH_ERR_ARG, "null part list") : MH_ERR_ARG;
  return env_copy(h, dtype, D1, D2, nw, nullptr, parts, nullptr, E);
}
