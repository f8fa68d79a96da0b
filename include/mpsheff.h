/*
 * mpsheff.h -- C ABI of the MI355X (gfx950) effective-Hamiltonian engine.
 *
 * The engine evaluates the contractions behind H_eff.x and the left/right
 * environment updates of finite-MPS algorithms (DMRG, TDVP) on one GPU.
 * Every tensor is a device pointer to a dense column-major array whose index
 * order is the Julia/TensorKit order of the reference:
 *
 *   MPS site tensor   A[a, s, b]            extents (Dl, d, Dr)
 *   two-site tensor   X[a, s1, s2, b]       extents (Dl, d1, d2, Dr)
 *   bond tensor       C[a, b]               extents (Dl, Dr)
 *   MPO site tensor   W[wl, s_out, wr, s_in] extents (wl, d, wr, d)
 *   environment       E[a_bra, w, a_ket]    extents (D, w, D)
 *
 * The first index is fastest in memory.  dtype MH_C64 means interleaved
 * (re, im) pairs of doubles (Julia ComplexF64).  Every entry point returns an
 * mh_status; on failure mh_last_error(h) describes the cause.  Nothing aborts.
 *
 * Work is enqueued on the handle's stream (mh_set_stream); outputs go into
 * caller-provided device buffers.  Intermediates live in a workspace the
 * handle owns; it grows on first use of a larger shape and is then reused.
 * A handle is not thread-safe: one host thread drives one handle.
 */
#ifndef MPSHEFF_H
#define MPSHEFF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mh_context *mh_handle;

enum mh_status {
  MH_OK = 0,
  MH_ERR_ARG = 1,         /* null pointer, bad enum          (Julia ArgumentError)     */
  MH_ERR_SHAPE = 2,       /* extents inconsistent            (Julia DimensionMismatch) */
  MH_ERR_DTYPE = 3,       /* dtype/precision not supported   (Julia ArgumentError)     */
  MH_ERR_HIP = 4,         /* a HIP runtime call failed                                  */
  MH_ERR_OOM = 5,         /* device allocation failed                                    */
  MH_ERR_NOCONV = 6,      /* Krylov iteration did not reach its tolerance               */
  MH_ERR_UNSUPPORTED = 7  /* feature not built for this configuration                  */
};

enum mh_dtype { MH_F64 = 0, MH_C64 = 1 };

/* Arithmetic used for the O(D^3) contractions.  Storage stays FP64/C64.
 *   MH_PREC_F64    : FP64 MFMA (v_mfma_f64_16x16x4f64), the plain path.
 *   MH_PREC_F32    : operands rounded to FP32, FP32 MFMA, FP32 accumulation.
 *   MH_PREC_BF16X3 : each operand split as hi+lo BF16 pair, three BF16 MFMA
 *                    products (hi*hi + hi*lo + lo*hi), FP32 accumulation.      */
enum mh_prec { MH_PREC_F64 = 0, MH_PREC_F32 = 1, MH_PREC_BF16X3 = 2 };

/* ---- handle lifecycle ---------------------------------------------------- */
int mh_create(int device, mh_handle *out);
int mh_destroy(mh_handle h);
/* stream: a hipStream_t, or NULL for the null stream. */
int mh_set_stream(mh_handle h, void *stream);
int mh_synchronize(mh_handle h);
const char *mh_last_error(mh_handle h);
int mh_version(void);

/* ---- device memory for hosts without their own allocator ----------------- */
int mh_malloc(mh_handle h, size_t bytes, void **ptr);
int mh_free(mh_handle h, void *ptr);
int mh_upload(mh_handle h, void *dst_device, const void *src_host, size_t bytes);
int mh_download(mh_handle h, void *dst_host, const void *src_device, size_t bytes);

/* ---- effective Hamiltonians (algorithms/derivatives.jl) ------------------ */

/* y[a,s',b] = sum L[a,w,a'] x[a',s,b'] W[w,s',w',s] R[b,w',b']
 * replaces ac_prime(x::MPSTensor, m::MPOTensor, hleft, hright),
 * algorithms/derivatives.jl:2-4.  L: (Dl,wl,Dl), R: (Dr,wr,Dr), W: (wl,d,wr,d). */
int mh_ac_prime(mh_handle h, int dtype, int prec, int Dl, int d, int Dr, int wl,
                int wr, const void *x, const void *W, const void *L,
                const void *R, void *y);

/* y[a,s1',s2',b] = sum L[a,w,a'] x[a',s1,s2,b'] W1[w,s1',m,s1] W2[m,s2',w',s2] R[b,w',b']
 * replaces ac2_prime(x::MPOTensor, h1, h2, hleft, hright),
 * algorithms/derivatives.jl:20-22. */
int mh_ac2_prime(mh_handle h, int dtype, int prec, int Dl, int d1, int d2,
                 int Dr, int wl, int wm, int wr, const void *x, const void *W1,
                 const void *W2, const void *L, const void *R, void *y);

/* mh_ac2_prime without any allocation: no reference counterpart.  For
 * callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops, hosts that forbid hipMalloc).  Same arguments and
 * result as mh_ac2_prime; prec must be MH_PREC_F64 (MH_ERR_UNSUPPORTED
 * otherwise).  The intermediates live in a static 64 MiB device array that
 * each GPU holds until the process ends (the handle's workspace is not
 * touched): (wl + wr)*Dl*d1*d2*Dr elements (8 bytes real, 16 complex) plus
 * the folded MPO matrix.  An operator that needs more returns MH_ERR_SHAPE
 * (at w = 5, d = 2: D up to 458 real, 323 complex).  GEMM tiles come
 * from the heuristic (no tuning runs), so results agree with mh_ac2_prime to
 * rounding, not bit for bit.  All calls on one GPU share its array: calls
 * that overlap on one GPU (two streams, two host threads) race. */
int mh_ac2_prime_noalloc(mh_handle h, int dtype, int prec, int Dl, int d1, int d2, int Dr,
                         int wl, int wm, int wr, const void *x, const void *W1,
                         const void *W2, const void *L, const void *R, void *y);

/* out[p, j, q] = sum_i in[p, i, q] Mt[i, j]: a small matrix on the middle
 * index of a 3-index tensor (in: P x ni x Q, out: P x nj x Q, Mt: ni x nj,
 * all column-major on the device; ni, nj <= 4096).  With Mt = G^T it applies
 * a gate G[s', s] to the physical legs of a one-, two- or three-site tensor
 * (bond_evolution, _apply_impl, circuit/apply_gates.jl:9-31,74-79). */
int mh_mode_product(mh_handle h, int dtype, int P, int ni, int nj, int Q, const void *in,
                    const void *Mt, void *out);

/* The right bond left open: P[a,s',w',b'] = sum L[a,w,a'] x[a',s,b'] W[w,s',w',s]
 * (FP64 / C64 arithmetic), i.e. ac_prime before its last contraction with R.
 * P: (Dl, d, wr, Dr) = the CH subspace-expansion term L.A.W with (w', b')
 * fused, w' fastest (expansion/chexpand.jl:23-25); the left-open term of
 * chexpand.jl:34-36 is the same contraction of the bond-transposed tensors. */
int mh_ac_prime_open(mh_handle h, int dtype, int Dl, int d, int Dr, int wl, int wr,
                     const void *x, const void *W, const void *L, void *P);

/* y[a,b] = sum L[a,w,a'] x[a',b'] R[b,w,b']
 * replaces c_prime(x::MPSBondTensor, hleft, hright), algorithms/derivatives.jl:52-54. */
int mh_c_prime(mh_handle h, int dtype, int prec, int Dl, int Dr, int w,
               const void *x, const void *L, const void *R, void *y);

/* ---- environment updates (mpo/transfer.jl) ------------------------------- */

/* hnew[b,w',b'] = sum hold[a,w,a'] psiB[a',s,b'] W[w,s',w',s] conj(psiA[a,s',b])
 * replaces updateleft(hold, psiAj, hj::MPOTensor, psiBj), mpo/transfer.jl:31-34.
 * hold: (Dl,wl,Dl), psiA/psiB: (Dl,d,Dr), hnew: (Dr,wr,Dr). */
int mh_updateleft(mh_handle h, int dtype, int prec, int Dl, int d, int Dr,
                  int wl, int wr, const void *hold, const void *psiA,
                  const void *W, const void *psiB, void *hnew);

/* hnew[a,w,a'] = sum hold[b,w',b'] psiB[a',s,b'] W[w,s',w',s] conj(psiA[a,s',b])
 * replaces updateright(hold, psiAj, hj::MPOTensor, psiBj), mpo/transfer.jl:26-29.
 * hold: (Dr,wr,Dr), hnew: (Dl,wl,Dl). */
int mh_updateright(mh_handle h, int dtype, int prec, int Dl, int d, int Dr,
                   int wl, int wr, const void *hold, const void *psiA,
                   const void *W, const void *psiB, void *hnew);

/* ---- Schur-form MPO tensors and Vector environments ---------------------- */

/* W[wl, s_out, wr, s_in] (host arrays, column-major; dtype MH_F64 or MH_C64)
 * from the entries of a Schur-form site tensor (the reference's
 * MPOHamiltonian: ac_prime(x, ham::AbstractSiteMPOTensor, hleft::Vector,
 * hright::Vector), algorithms/derivatives.jl:7-17,24-49): term t adds
 * coef[t] * O_t at channel pair (ti[t], tj[t]), where O_t is the next d x d
 * matrix of ops (O[s_out, s_in], column-major) when kind[t] = 1, or the
 * identity when kind[t] = 0 (the scalar entries, scalkeys).  coef holds
 * (re, im) pairs; a real W needs im = 0.  Host only: no handle, no device. */
int mh_schur_fold(int dtype, int wl, int wr, int d, int nterms, const int *ti,
                  const int *tj, const int *kind, const double *coef, const void *ops,
                  void *W);
/* The reference's Vector environments (one D1 x D2 device matrix per MPO
 * channel, hleft[w][a, a']) <-> the engine's E[a, w, a'] (extents D1, nw,
 * D2), as strided device copies on the handle's stream. */
int mh_env_gather(mh_handle h, int dtype, int D1, int D2, int nw, const void *const *parts,
                  void *E);
int mh_env_scatter(mh_handle h, int dtype, int D1, int D2, int nw, const void *E,
                   void *const *parts);

/* ---- operator descriptor for the device-resident Krylov drivers ---------- */
enum mh_heff_kind { MH_HEFF0 = 0, MH_HEFF1 = 1, MH_HEFF2 = 2 };

typedef struct mh_heff_op {
  int kind;   /* mh_heff_kind: c_prime / ac_prime / ac2_prime              */
  int dtype;  /* mh_dtype                                                  */
  int prec;   /* mh_prec                                                   */
  int Dl, d1, d2, Dr; /* d1 unused for HEFF0, d2 used only by HEFF2        */
  int wl, wm, wr;     /* wm used only by HEFF2; HEFF0 uses wl (= wr)       */
  const void *W1, *W2, *L, *R;
} mh_heff_op;

/* number of scalar entries of the vector the operator acts on */
int mh_heff_size(const mh_heff_op *op, size_t *n);
/* y = H_eff x for the described operator */
int mh_heff_apply(mh_handle h, const mh_heff_op *op, const void *x, void *y);

/* Lowest eigenpair by three-term Lanczos with the Krylov basis kept on the
 * device; only the tridiagonal coefficients cross to the host.
 * Follows simple_lanczos_solver(A, v, "SR", kmax, btol),
 * auxiliary/simple_lanczos.jl:83-120 (via _eig_solver, algorithms/dmrg.jl:3-6):
 * first step :27-42, further steps :49-68, early exit when beta < btol,
 * host eigen of the k x k tridiagonal :115, Ritz vector :70-78.
 * reorth = 0 reproduces the reference (no re-orthogonalisation); reorth = 1
 * adds FP64 full re-orthogonalisation (used by the mixed-precision forms).
 * info = 0 converged (beta < btol), -1 ran kmax steps (reference semantics).
 * niter = Krylov dimension used.  xout may alias x0.                        */
int mh_lanczos_eig(mh_handle h, const mh_heff_op *op, const void *x0, int kmax,
                   double btol, int reorth, double *eigval, void *xout,
                   int *info, int *niter);

/* mh_lanczos_eig on the projected operator x -> P(H_eff x) with
 * P(y): for l = 0..nproj-1 in order, y -= dot(y, p_l) p_l (dot conjugate-
 * linear in y), applied to the start vector too: the excited-state solve of
 * algorithms/dmrgexcited.jl:1-6,22-25,86-90 (projector vectors p_l are the
 * lower states contracted with the overlap environments, c_proj,
 * derivatives.jl:65-67).  proj: nproj <= 8 device vectors of x's size. */
int mh_lanczos_eig_proj(mh_handle h, const mh_heff_op *op, int nproj, const void *const *proj,
                        const void *x0, int kmax, double btol, int reorth, double *eigval,
                        void *xout, int *info, int *niter);

/* xout = exp(dt * H_eff) x0 by a Lanczos (Hermitian H_eff) Krylov space of
 * dimension <= kdim, repeated in sub-steps until the Krylov error estimate of
 * each sub-step is below tol.  Replaces _exp_evolve -> KrylovKit.exponentiate,
 * algorithms/tdvp.jl:3-6.  dt = dt_re + i dt_im; a non-real dt needs MH_C64. */
int mh_lanczos_expm(mh_handle h, const mh_heff_op *op, const void *x0,
                    double dt_re, double dt_im, int kdim, double tol,
                    void *xout, int *info);

/* The same with an Arnoldi Krylov space (full FP64 classical Gram-Schmidt,
 * two passes; the Hessenberg matrix exponentiated on the host by scaling and
 * squaring of the [8/8] Pade approximant): any H_eff, Hermitian or not.
 * Replaces exponentiate(...; ishermitian=false), KrylovKit's Arnoldi driver,
 * the reference's default in algorithms/tdvp.jl:3-6,16,78,153. */
int mh_arnoldi_expm(mh_handle h, const mh_heff_op *op, const void *x0,
                    double dt_re, double dt_im, int kdim, double tol,
                    void *xout, int *info);

/* enable = 1 (default; environment MH_GRAPHS sets 0/1/2): the Lanczos
 * steps after the first of mh_lanczos_eig / mh_bs_lanczos_eig /
 * mh_lanczos_expm are captured into one hipGraph and launched as a whole
 * (exponentiate re-launches it for every sub-step), except steps on short
 * vectors (the single-workgroup step kernel, a few launches per step),
 * which run eagerly because capturing them costs more than it saves;
 * enable = 2 captures those too; 0 captures nothing.  Results are identical
 * to the eager launches. */
int mh_set_graphs(mh_handle h, int enable);
/* enable = 0: run the first step of the FP64 one-/two-site H_eff as a GEMM
 * followed by a separate mode-product launch instead of the fused kernel
 * (default on; environment MH_FUSION=0). */
int mh_set_fusion(mh_handle h, int enable);
/* enable = 1 (default; environment MH_KRYLOV_SMALL=0 turns it off): a Krylov
 * step on vectors of at most 16384 doubles runs its vector work (alpha, the
 * update, Gram-Schmidt passes, beta, next basis vector) as one launch of one
 * workgroup instead of the multi-block reduction kernels; the same
 * recurrence with a different summation order. */
int mh_set_krylov_small(mh_handle h, int enable);
/* enable = 1 (default 0; environment MH_BS_PLAIN=1): the GEMM stages of a
 * block-sparse plan run as one strided FP64 GEMM launch per sector product
 * (output block and K-segment, segments accumulated), the per-sector loop of
 * a plain port of TensorKit's block contraction, instead of one grouped
 * launch per stage.  FP64 plans only; the baseline the grouped form is
 * measured against. */
int mh_set_bs_plain(mh_handle h, int enable);
/* GEMM configurations (tile shape, K depth, split-K, tile order, waves) the
 * per-shape tuner of each precision chooses from; force_candidate fixes
 * configuration `cand` of the calling precision's list for every GEMM of the
 * handle that admits it (-1: back to tuning).  Used by the tests to check
 * each configuration. */
int mh_gemm_candidates(int prec, int *n);
int mh_gemm_force_candidate(mh_handle h, int cand);

/* ---- Krylov vector primitives (simple_lanczos.jl:28-38,50-63) ------------ */
/* out[0] + i out[1] = <x, y>  (conjugate-linear in x, as Julia dot) */
int mh_dot(mh_handle h, int dtype, size_t n, const void *x, const void *y,
           double *out);
int mh_nrm2(mh_handle h, int dtype, size_t n, const void *x, double *out);
/* y = a x + b y with complex a = (a_re, a_im), b = (b_re, b_im) */
int mh_axpby(mh_handle h, int dtype, size_t n, double a_re, double a_im,
             const void *x, double b_re, double b_im, void *y);

/* ---- plain strided GEMM used by the sweep glue --------------------------- */
/* C[m,n] = alpha * sum_k op(A)[m,k] B[k,n] + beta * C[m,n],
 * element (m,k) of A at A + m*sam + k*sak (op = conj if conjA), likewise for
 * B and C.  Used for the two-site contraction A[s]*A[s+1] (dmrg.jl:98). */
int mh_gemm(mh_handle h, int dtype, int prec, int M, int N, int K,
            const void *A, int64_t sam, int64_t sak, int conjA,
            const void *B, int64_t sbk, int64_t sbn, void *C, int64_t scm,
            int64_t scn, double alpha_re, double alpha_im, double beta_re,
            double beta_im);

/* ---- adjoint (differentiation) ------------------------------------------- */
/* G[w,s',w',s] = sum L[a,w,a'] conj(psiA[a,s',b]) psiB[a',s,b'] R[b,w',b']:
 * the pullback of ac_prime / updateleft / updateright onto the MPO tensor,
 * <psiA| H_eff(W) |psiB> = sum G.W.  The reference has no differentiation
 * (Project.toml:6-11 has no ChainRulesCore); this is the kernel behind the
 * energy gradient dE/dtheta = sum_i <G_i, dW_i/dtheta> (see the callers of
 * ac_prime, derivatives.jl:2-4, and of updateleft, mpo/transfer.jl:31-34).
 * L: (Dl,wl,Dl), psiA/psiB: (Dl,d,Dr), R: (Dr,wr,Dr), G: (wl,d,wr,d).
 * Needs w*d <= 64 on both sides and (w d)^2 <= 1024.  Deterministic.      */
int mh_mpo_grad(mh_handle h, int dtype, int prec, int Dl, int d, int Dr, int wl, int wr,
                const void *L, const void *psiA, const void *psiB, const void *R, void *G);

/* ---- block-sparse (abelian-symmetric) contractions ----------------------- */
/* A tensor with U(1)-type charge conservation is a list of dense column-major
 * blocks in one device buffer (TensorKit GradedSpace tensors; the Hubbard
 * U1xU1 site operators of utilities/fermion_siteops.jl:37-62).  A host
 * planner that knows the charge sectors lowers one contraction to a plan of
 * stages; each stage is ONE launch over all sectors:
 *   grouped GEMM: tasks (one output block C, M x N, ldc) each summing
 *     K-segments opA (M x K) * opB (K x N) addressed by (offset, ld);
 *   block mode product: destination blocks, each a linear combination of
 *     equally shaped source blocks (the MPO entries between the sectors).
 * Plans replace TensorKit's per-sector loop behind
 *   ac2_prime(x, h1, h2, hleft, hright)  algorithms/derivatives.jl:20-22,
 *   ac_prime(x, h, hleft, hright)        algorithms/derivatives.jl:2-4,
 *   updateleft / updateright             mpo/transfer.jl:26-34
 * on graded tensors.  Arithmetic is FP64 storage with FP64 or mixed
 * (MH_PREC_F32 / MH_PREC_BF16X3, FP32 accumulation) matrix-core products.
 * Buffers: the operand table of every stage indexes six buffers.          */
typedef struct mh_bs_plan_s *mh_bs_plan;
enum mh_bs_buffer {
  MH_BS_X = 0,   /* input vector (ket site tensor for environment updates) */
  MH_BS_Y = 1,   /* output                                                 */
  MH_BS_P1 = 2,  /* left environment (H_eff, updateleft) / old right env   */
  MH_BS_P2 = 3,  /* right environment (H_eff) / bra site tensor (updates)  */
  MH_BS_T1 = 4,  /* intermediates, held in the handle's workspace          */
  MH_BS_T2 = 5
};
/* sizes[6]: elements of each buffer; every table entry added later is
 * checked against them, so a plan never addresses outside its buffers.
 * A plan is built on the host: h may be NULL while building (errors then
 * carry no message).  Its first apply / Lanczos on a handle uploads the
 * tables to that handle's device in one allocation (not under stream
 * capture); the plan then stays on that device. */
int mh_bs_plan_create(mh_handle h, int prec, const int64_t *sizes, mh_bs_plan *plan);
/* task_i[5*t] = {M, N, seg0, nseg, ldc}, task_coff[t] = offset of C;
 * seg_off[2*s] = {offset of A, offset of B}, seg_i[3*s] = {K, lda, ldb}.
 * amc: A is m-contiguous (A(m,k) at m + k*lda), else A(m,k) at m*lda + k;
 * bkc: B is k-contiguous (B(k,n) at k + n*ldb), else B(k,n) at k*ldb + n. */
int mh_bs_plan_add_gemm(mh_handle h, mh_bs_plan plan, int a_src, int b_src, int c_src,
                        int amc, int bkc, int ntask, const int32_t *task_i,
                        const int64_t *task_coff, int nseg, const int64_t *seg_off,
                        const int32_t *seg_i);
/* dst_i[5*b] = {rows, cols, ld, c0, nc}: block b at dst_off[b] is
 * sum_{c in [c0, c0+nc)} coef[c] * (rows x cols block at src_off[c], ld src_ld[c]) */
int mh_bs_plan_add_mode(mh_handle h, mh_bs_plan plan, int in_src, int out_src, int ndst,
                        const int64_t *dst_off, const int32_t *dst_i, int ncontrib,
                        const int64_t *src_off, const int32_t *src_ld, const double *coef);
int mh_bs_plan_flops(mh_bs_plan plan, double *flops);  /* sum 2mnk + mode flops */
/* Grouped-GEMM tile variants (tile shape x K depth).  By default every GEMM
 * stage picks its variant on first use by timing all of them, cached per
 * shape class in the handle (MH_AUTOTUNE=0: always variant 0).
 * force_variant fixes one variant for every GEMM stage (-1: back to tuning);
 * used by the tests to check each variant. */
int mh_bs_plan_variants(mh_bs_plan plan, int *n);
int mh_bs_plan_force_variant(mh_bs_plan plan, int variant);
int mh_bs_plan_destroy(mh_handle h, mh_bs_plan plan);
/* run the plan: y = contraction(x; p1, p2) */
int mh_bs_apply(mh_handle h, mh_bs_plan plan, const void *x, const void *p1, const void *p2,
                void *y);
/* ---- native planner: plans straight from the charge sectors -------------
 * A bond is a list of sectors (charge, dimension) in the caller's storage
 * order; the planners lay the blocks out as
 *   MPS-type A[a, s1..sn, b], q(a) + sum q(s) = q(b): blocks (qa, s1..sn) in
 *     bond order of qa, then the physical indices (s1 slowest); the blocks
 *     of one left sector are adjacent: one D_qa x ncols column-major matrix;
 *   environment E[a_bra, w, a_ket], q(a_bra) = q(w) + q(a_ket): blocks
 *     (q, w) in bond order of q, then channel; one sector's blocks form a
 *     D_q x K_q column-major matrix.
 * These are the layouts of the Python planner (blocksparse.py), which lists
 * the sectors sorted by charge; both produce the same tables.
 * charges: nq ints per sector (nq = 1..4 U(1) factors, |q| < 16384).
 * phys: d charges of the physical index (nq ints each).  qwl / qwr: charge
 * of every MPO channel on the left / right MPO bond (nq ints each); MPO
 * tensors W[wl, s_out, wr, s_in] column-major FP64.  Errors: MH_ERR_ARG for
 * an MPO entry that breaks the charge rule or an environment sector that no
 * block reaches, MH_ERR_SHAPE for a non-positive dimension.               */
typedef struct {
  int n;                  /* number of sectors   */
  int nq;                 /* charge components   */
  const int32_t *charges; /* [n * nq]            */
  const int32_t *dims;    /* [n]                 */
} mh_bs_sectors;
/* ac2_prime(x, W1, W2, L, R) (algorithms/derivatives.jl:20-22): x on
 * left | d | d | right; buffers X = x, Y = y, P1 = L (over left, channels
 * qwl), P2 = R (over right, channels qwr); W1: (wl,d,wm,d), W2: (wm,d,wr,d). */
int mh_bs_plan_heff2(mh_handle h, int prec, const mh_bs_sectors *left,
                     const mh_bs_sectors *right, int d, const int32_t *phys, int wl,
                     const int32_t *qwl, int wm, int wr, const int32_t *qwr, const double *W1,
                     const double *W2, mh_bs_plan *plan);
/* updateleft(L, A, W, B) (mpo/transfer.jl:31-34): A, B on left | d | right;
 * X = ket B, P1 = L (over left, qwl), P2 = bra A, Y = L' (over right, qwr). */
int mh_bs_plan_updateleft(mh_handle h, int prec, const mh_bs_sectors *left,
                          const mh_bs_sectors *right, int d, const int32_t *phys, int wl,
                          const int32_t *qwl, int wr, const int32_t *qwr, const double *W,
                          mh_bs_plan *plan);
/* updateright(R, A, W, B) (mpo/transfer.jl:26-29): X = ket B, P1 = R (over
 * right, qwr), P2 = bra A, Y = R' (over left, qwl). */
int mh_bs_plan_updateright(mh_handle h, int prec, const mh_bs_sectors *left,
                           const mh_bs_sectors *right, int d, const int32_t *phys, int wl,
                           const int32_t *qwl, int wr, const int32_t *qwr, const double *W,
                           mh_bs_plan *plan);
/* two-site tensor x[a,s1,s2,b] = A1[a,s1,m] A2[m,s2,b] (dmrg.jl:98), FP64:
 * X = A1 (left | d | mid), P1 = A2 (mid | d | right), Y = x. */
int mh_bs_plan_two_site(mh_handle h, const mh_bs_sectors *left, const mh_bs_sectors *mid,
                        const mh_bs_sectors *right, int d, const int32_t *phys,
                        mh_bs_plan *plan);
/* sizes[6] of a plan's buffers (elements), as mh_bs_plan_create took them. */
int mh_bs_plan_sizes(mh_bs_plan plan, int64_t *sizes);
/* Inspection of a plan's stages (host tables, no device needed):
 * info[8] = {0 (gemm), a_src, b_src, c_src, amc, bkc, ntask, nseg} or
 *           {1 (mode), in_src, out_src, 0, 0, 0, ndst, ncontrib};
 * the table getters fill arrays in the form of _add_gemm / _add_mode (any
 * pointer may be NULL). */
int mh_bs_plan_stage_count(mh_bs_plan plan, int *n);
int mh_bs_plan_stage_info(mh_bs_plan plan, int stage, int32_t *info);
int mh_bs_plan_stage_gemm(mh_bs_plan plan, int stage, int32_t *task_i, int64_t *task_coff,
                          int64_t *seg_off, int32_t *seg_i);
int mh_bs_plan_stage_mode(mh_bs_plan plan, int stage, int64_t *dst_off, int32_t *dst_i,
                          int64_t *src_off, int32_t *src_ld, double *coef);
/* mh_lanczos_eig with a block-sparse H_eff plan (x and y of equal size);
 * the Lanczos vectors are the flat block buffers. */
int mh_bs_lanczos_eig(mh_handle h, mh_bs_plan plan, const void *p1, const void *p2,
                      const void *x0, int kmax, double btol, int reorth, double *eigval,
                      void *xout, int *info, int *niter);

/* ---- TensorKit's coupled-sector block form ------------------------------
 * A graded MPS-type tensor A[a, s1..sn, b] (nphys = n in {1, 2}; layout as
 * the native planner's) seen as a TensorMap with codomain (a, s1..sk) and
 * domain (s_{k+1}..sn, b), k = ncod - 1 (ncod in 1..n+1): TensorKit stores
 * it as one column-major matrix per coupled charge c, blocks(t)[c]
 * (the matrices behind abelian_matrix_from_dense,
 * auxiliary/abelianmatrix.jl:112-124, and the graded states of
 * mps/initializers.jl:34-73): rows are the codomain pieces (qa, s1..sk) with
 * q(a) + sum q(s) = c, columns the domain pieces (s_{k+1}..sn, qb) with
 * q(b) - sum q(s) = c.  Convention here: coupled charges in ascending
 * lexicographic order; row pieces by left sector (caller's order), then
 * s1..sk (s1 slowest); column pieces by right sector, then s_{k+1}..sn; the
 * matrices concatenated in one buffer (element offsets[c]).  A host binding
 * whose fusion-tree order differs permutes row or column pieces of equal
 * charge.  mh_bs_coupled_info: nc coupled sectors; charges [nc * nq], rows
 * [nc], cols [nc], offsets [nc + 1] (any output but nc may be NULL; host
 * only, no device).  mh_bs_to_coupled / mh_bs_from_coupled: device copies
 * between the block buffer (the layout's size) and the coupled buffer
 * (offsets[nc] elements) on the handle's stream; they wait for the stream. */
int mh_bs_coupled_info(const mh_bs_sectors *left, const mh_bs_sectors *right, int d,
                       const int32_t *phys, int nphys, int ncod, int *nc, int32_t *charges,
                       int32_t *rows, int32_t *cols, int64_t *offsets);
int mh_bs_to_coupled(mh_handle h, const mh_bs_sectors *left, const mh_bs_sectors *right, int d,
                     const int32_t *phys, int nphys, int ncod, const double *blocks,
                     double *coupled);
int mh_bs_from_coupled(mh_handle h, const mh_bs_sectors *left, const mh_bs_sectors *right, int d,
                       const int32_t *phys, int nphys, int ncod, const double *coupled,
                       double *blocks);

/* ---- batched symmetric eigensolver (block-sparse two-site split) ---------
 * The FP64 step of the graded two-site split: the truncated SVD of every
 * charge sector (TensorKit tsvd under auxiliary/stable_svd.jl:4-18,
 * truncated by auxiliary/truncation.jl:27-47) is taken from the sector
 * reduced density matrices rho_c = M_c M_c^T, whose eigenpairs give u and
 * s^2.  nmat symmetric matrices: matrix m of order n[m] (1 <= n[m] <= 2048)
 * is column-major with ld = n[m] at element offset a_off[m] of A; D, E and tau
 * hold n[m] entries each from element offset v_off[m]; matrices and their
 * D/E/tau ranges must not overlap.  Offsets and orders are host arrays.     */
/* Householder tridiagonalisation, all matrices in one cooperative launch;
 * output as LAPACK dsytrd with uplo = 'L' (reflector i in A[i+2:, i]). */
int mh_sytrd_batched(mh_handle h, int nmat, const int *n, const int64_t *a_off,
                     const int64_t *v_off, double *A, double *D, double *E, double *tau);
/* Eigenvalues (ascending, into D) and eigenvectors of the tridiagonals
 * (column-major, ld = n[m], at element offset z_off[m] of Z) by rocSOLVER
 * dstedc, one call per matrix on side streams; E is overwritten. */
int mh_stedc_batched(mh_handle h, int nmat, const int *n, const int64_t *v_off,
                     const int64_t *z_off, double *D, double *E, double *Z);
/* The same eigen-decomposition of the tridiagonals (D: diagonal in,
 * eigenvalues ascending out; E: off-diagonal, E[v_off[m] + i] couples rows i
 * and i + 1, read only; Z: eigenvectors, column-major, ld = n[m], at element
 * offset z_off[m]) by a batched divide and conquer on the device: every
 * merge of one tree height, over all matrices, in the same few launches
 * (1 <= n[m] <= 4096).  Replaces the per-matrix dstedc calls of
 * mh_stedc_batched; a leaf that does not converge is reported by
 * mh_eig_check. */
int mh_stedc_dc_batched(mh_handle h, int nmat, const int *n, const int64_t *v_off,
                        const int64_t *z_off, double *D, const double *E, double *Z);
/* Columns [col0[m], col0[m] + ncol[m]) of Z_m <- Q_m Z_m, Q_m the product of
 * the reflectors mh_sytrd_batched left in A and tau: eigenvectors of the
 * original matrices. */
int mh_ormtr_batched(mh_handle h, int nmat, const int *n, const int64_t *a_off,
                     const int64_t *v_off, const double *A, const double *tau,
                     const int64_t *z_off, const int *col0, const int *ncol, double *Z);
/* Batched strided FP64 GEMMs of unrelated shapes in one launch on the
 * handle's stream: C_t = alpha op(A_t) op(B_t) + beta C_t for t < count (the
 * per-sector products of the graded split, which TensorKit runs as one BLAS
 * call per sector block).  desc: count rows of 12 int64 (host array): device
 * pointers A, B, C, extents M, N, K, and element strides sam, sak (A[m, k]
 * at A + m sam + k sak), sbk, sbn, scm, scn.  C is not read when beta = 0;
 * C must not overlap A or B; tasks with M = 0 or N = 0 are skipped. */
int mh_gemm_batched(mh_handle h, int count, const int64_t *desc, double alpha, double beta);

/* Waits for the handle's stream; MH_ERR_NOCONV if a dstedc call of the last
 * mh_stedc_batched (nmat matrices) or a leaf of the last mh_stedc_dc_batched
 * failed to converge, MH_ERR_HIP if a
 * wait between workgroups of the last tridiagonalisation timed out. */
int mh_eig_check(mh_handle h, int nmat);

/* Eigen-decomposition of `batch` Hermitian (MH_C64) or symmetric (MH_F64)
 * n x n matrices, one workgroup per matrix with the matrix in LDS: the small
 * reduced density matrices of the dense two-site split (the FP64 step behind
 * auxiliary/stable_svd.jl:4-18 and truncation.jl:27-47 at small bonds), where
 * a library syevd is launch-latency bound.  A: batch matrices, column-major,
 * contiguous (read only).  W: batch x n eigenvalues, ascending.  Z: batch x
 * n x n, column-major; columns col0 .. n-1 hold the eigenvectors of W[col0..]
 * (the columns below col0 are left unfinished).  info (device, nullable):
 * per matrix 0, or 1 when the QL iteration did not converge (W is then NaN).
 * n above mh_heev_small_max -> MH_ERR_SHAPE. */
int mh_heev_small(mh_handle h, int dtype, int n, int batch, const void *A, double *W, void *Z,
                  int col0, int *info);
/* largest order mh_heev_small takes on the handle's device for dtype */
int mh_heev_small_max(mh_handle h, int dtype, int *nmax);

/* Thin SVD A = U diag(S) V^H of one m x n matrix with m >= n (MH_F64 or
 * MH_C64), one workgroup with A and V in LDS, by column-pivoted QR and
 * one-sided Jacobi on R^H (MH_SVD_PLAIN=1 in the environment: Jacobi on A
 * itself; an input rank-deficient far below eps takes that path too): the
 * truncated SVD of the dense two-site split (auxiliary/stable_svd.jl:4-18,
 * LAPACK gesdd in the reference's tsvd) at the small bonds where a library
 * gesvd is launch-latency bound.  A: column-major, leading dimension lda
 * (read only).  U: m x n, S: n (descending), V: n x n (right singular vectors
 * as columns), all column-major on the device; stream-ordered, no host
 * synchronisation.  sweeps (device, nullable): the Jacobi sweeps run (the
 * last one rotated no pair), or -1 when the sweep limit was reached; every S
 * entry is then NaN.  MH_ERR_SHAPE when m < n or the matrices do not fit the
 * device's LDS (mh_svd_small_fits tells beforehand). */
int mh_svd_small(mh_handle h, int dtype, int m, int n, const void *A, int lda, void *U, double *S,
                 void *V, int *sweeps);
/* *fits = 1 if mh_svd_small takes an m x n matrix of dtype on this device */
int mh_svd_small_fits(mh_handle h, int dtype, int m, int n, int *fits);

/* ---- measurement --------------------------------------------------------- */
enum mh_kernel_class { MH_K_GEMM = 0, MH_K_WSTAGE = 1, MH_K_VEC = 2, MH_K_NCLASS = 3 };
/* enable = 1: bracket every launch with HIP events on the handle's stream */
int mh_profile_enable(mh_handle h, int enable);
/* total device milliseconds, launch count and algorithmic flops recorded for
 * one kernel class since the last reset; waits for the recorded events.   */
int mh_profile_read(mh_handle h, int kclass, double *ms, int64_t *launches,
                    double *flops);
int mh_profile_reset(mh_handle h);
/* FP64 matrix-core throughput this device sustains on a register-only MFMA
 * loop (no memory traffic), in TFLOP/s: the measured roofline ceiling. */
int mh_probe_mfma_f64(mh_handle h, double *tflops);

#ifdef __cplusplus
}
#endif
#endif /* MPSHEFF_H */
