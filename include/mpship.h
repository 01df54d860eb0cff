/*
 * mpship.h -- C ABI of the MI355X (gfx950) MPS time-evolution hot path.
 *
 * Drop-in replacement for the tensor-network hot path of
 * jdnz/MatrixProductStates.jl: environment contractions (src/mps_base.jl),
 * one-site variational compression sweeps (src/compression.jl), RK4 and
 * Arnoldi steps (src/timeevolution.jl) and QR/SVD gauge moves
 * (src/orthonormalisation.jl, compressionSVD! in src/compression.jl).
 *
 * Conventions (identical to the reference, which is column-major Julia):
 *   - complex numbers are interleaved (re, im) doubles (Julia ComplexF64);
 *   - MPS site A[Dl, d, Dr], MPO site O[wl, d_out, wr, d_in], column-major;
 *   - operator environments E[Dket, w, Dbra], LEFT ones stored conjugated
 *     (mps_base.jl:219-241); overlap environments L[Dbra, Dket], R[Dket, Dbra];
 *   - sites and environment slots are 0-based: site j (Julia jj = j+1) has its
 *     left environment in slot j and its right environment in slot j+1;
 *   - every function returns 0 on success, a negative MPS_E_* code otherwise;
 *     mps_last_error() gives a message (thread-local);
 *   - shapes are checked on the host before anything is queued: a chain
 *     length, local dim or bond that does not fit, an environment slot smaller
 *     than the states' bonds need, or (in the H-applying sweeps, RK4 and
 *     Arnoldi, whose states the reference builds from one dims vector) states
 *     of differing bonds return MPS_E_ARG (the reference's DimensionMismatch)
 *     and leave every argument unchanged.
 *
 * All state (MPS sites, MPOs, environments) is device resident and owned by
 * the library.  Calls on one context are ordered on the context's HIP stream;
 * functions that return a host scalar synchronise that stream.  Distinct
 * contexts may be used from distinct host threads.
 */
#ifndef MPSHIP_H
#define MPSHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPS_OK 0
#define MPS_E_ARG (-1)      /* invalid argument / DimensionMismatch        */
#define MPS_E_HIP (-2)      /* HIP runtime error                            */
#define MPS_E_ALLOC (-3)    /* device allocation failed                     */
#define MPS_E_STATE (-4)    /* call not valid in the current state          */
#define MPS_E_NUMERIC (-5)  /* numerical failure (e.g. SVD did not converge) */

/* Both precisions form complex products with three real MFMA (3M: ArBr,
 * AiBi, (Ar+Ai)(Br+Bi)).  The rounding bound is normwise, about
 * 3 K eps |a||b| per complex dot product of length K, not componentwise: a
 * real or imaginary part far below |a||b| has an absolute error of that
 * size (eps = 2^-53 for FP64, 2^-24 for FP32). */
#define MPS_PREC_FP64 0     /* plain path: complex FP64 on MFMA (3M)        */
#define MPS_PREC_FP32 1     /* precision form: operator terms in complex FP32 (3M) */

typedef struct mps_ctx_s* mps_ctx_t;
typedef struct mps_state_s* mps_state_t;  /* Vector{Array{TN,3}} on device    */
typedef struct mps_mpo_s* mps_mpo_t;      /* Vector{Array{TN,4}} on device    */
typedef struct mps_env_s* mps_env_t;      /* build_env result on device       */

/* ---------------------------------------------------------------- context */
const char* mps_last_error(void);
int mps_version(void);
int mps_ctx_create(int device, mps_ctx_t* out);
int mps_ctx_destroy(mps_ctx_t ctx);
/* Lockstep batch (no counterpart in the reference; it serves ensembles of
 * quantum trajectories of one system, spin_model_quantum_trajectory_rk4.jl:
 * 115-209 run for many jump records at once).  A batched context holds nb
 * copies of every state, MPO and environment created on it (copy_bytes of
 * containers per copy, plus a workspace of ws_bytes per copy, reserved at
 * creation), and every sweep, QR and norm call on it runs all copies with
 * the same kernels in the same launches.  Uploads write every copy.  One copy
 * is reached through a view on a companion context (same stream); entry
 * points that return per-call host values (scalar products, local
 * expectation values, SVD truncation) and the raw mps_dev_* calls refuse a
 * batched context. */
int mps_ctx_create_batch(int device, int nb, int64_t copy_bytes, int64_t ws_bytes, mps_ctx_t* out);
/* Workspace per copy (ws_bytes above) for RK4 and sweep calls at bond
 * dimension <= Dmax, local dimension <= d and MPO bond <= wmax. */
int mps_batch_workspace_bytes(int device, int nb, int64_t Dmax, int64_t d, int64_t wmax, int64_t* out);
/* A plain context on the batched context's stream, for views. */
int mps_ctx_create_companion(mps_ctx_t batched, mps_ctx_t* out);
/* Ragged pool (no counterpart in the reference; it serves ensembles of
 * trajectories of chains of different lengths, the uneven-N form of
 * spin_model_quantum_trajectory_rk4.jl:115-209).  Declares that copy b of
 * the batched context holds a chain of n[b] sites (n sorted longest first,
 * n[0] = nmax) in containers of nmax slots: its sites 0..n[b]/2-1 in slots
 * 0..n[b]/2-1 and the rest right-aligned, ending at slot nmax-1.  Build the
 * states with the nmax-site dims rule (both halves of every chain then see
 * their own bonds) and set each copy through views (mps_state_view gives copy
 * b's own n[b]-site chain; mps_state_product refuses the pool).  RK4 and the
 * H-applying half-sweeps then run every copy in the same launches: per site
 * slot, the copies present there; each copy's arithmetic is that of its chain
 * alone in a pool of the same size.  Other sweeps refuse a ragged pool
 * (MPS_E_STATE).  Call once, before the first sweep. */
int mps_ctx_set_ragged(mps_ctx_t ctx, int nmax, const int* n);
/* Use an external HIP stream (hipStream_t passed as void*); NULL = own stream. */
int mps_ctx_set_stream(mps_ctx_t ctx, void* stream);
int mps_ctx_synchronize(mps_ctx_t ctx);
/* Rank-revealing cutoff of the Householder generator: a reflector whose
 * trailing column norm is <= tol*||A||_F is the identity (tau = 0).  0 keeps
 * the exact LAPACK zgeqrf/zungqr semantics of qrq! (orthonormalisation.jl:180-198). */
int mps_ctx_set_qr_rank_tol(mps_ctx_t ctx, double tol);
/* Form of the QR behind qrq!/leftorth!/rightorth! (no counterpart in the
 * reference, which calls LAPACK): 1 = one persistent launch per factorization
 * where it applies (m <= 2560, n <= 256; on a batched context one launch for
 * every copy; panels of m <= 256 only for n > 16 on up to 64 copies), 0 = a
 * panel and a trailing-update launch per 8 or 16 columns, -1 = the process
 * default (1 unless MPS_QR_CHAIN=0 is set).  Both give zgeqr2's reflectors (the same Q to
 * rounding); the setting is for callers that compare or time them. */
int mps_ctx_set_qr_single_launch(mps_ctx_t ctx, int on);
/* MPS_PREC_FP64 (default) or MPS_PREC_FP32 for the operator (MPO) contractions. */
int mps_ctx_set_precision(mps_ctx_t ctx, int prec);
/* Refinement of the precision form (no counterpart in the reference; it is
 * the reference's own compress_sum_var_apply_H_left! / _right!,
 * compression.jl:889-944, 1007-1063, called again with the last RK4 stage's
 * arguments): with n > 0 (even, at most 64), every mps_rk4stp_apply_H_half
 * call on this context follows stage 4 with n half-sweeps (left, right, ...)
 * in complex FP64 that refit stage 4's target -1/3 psi + 1/3 psi1 +
 * 2/3 psi2 + 1/3 H3 psi3, each started from the previous result; the result
 * stays left canonical.  The context keeps one or two states of psi's shape
 * for it (the 10-argument form's copy of psi and the refit's intermediate);
 * on a batched context they come from the slab, so copy_bytes must hold two
 * more states of psi's shape.  Default 0. */
int mps_ctx_set_refine_sweeps(mps_ctx_t ctx, int n);
/* Enable / disable hipGraph capture of whole RK4 steps (default on). */
int mps_ctx_set_graphs(mps_ctx_t ctx, int enable);
/* Measurement extension (no counterpart in the reference): when enabled, every
 * GEMM and QR launch is bracketed by HIP events on the context stream and
 * graphs are not used.  mps_ctx_kernel_stats synchronises, fills
 * out[0..7] = {gemm_ms, gemm_flops, gemm_bytes, gemm_launches, qr_ms,
 * qr_flops, qr_launches, launches} (algorithmic flops/bytes) and resets;
 * launches = kernel launches the library has checked and queued in this
 * process so far (every launch is checked for a non-empty grid and block, the
 * kernel's thread limit, 32-bit grid sizes and the device's LDS limit, and
 * refused with MPS_E_STATE otherwise). */
int mps_ctx_set_profiling(mps_ctx_t ctx, int enable);
int mps_ctx_kernel_stats(mps_ctx_t ctx, double* out);
/* GEMM plan tuning (no counterpart in the reference; it serves the offline
 * autotuner mpsamd/autotune.py, whose table csrc/gemm_tuned.inc is compiled
 * into the library).  Between mps_ctx_gemm_record(ctx, 1) and (ctx, 0) every
 * contraction the context launches eagerly is recorded (captured graphs are
 * dropped first).  mps_ctx_gemm_recorded lists the distinct shapes in order of
 * first launch, 10 int64 per row: M, N, K, batch, copy class (floor(log2 of
 * the lockstep copies the plan is made for)), arithmetic (MPS_PREC_*), the
 * tile and split-K depth in use, the copies, the partials workspace bytes;
 * *rows is the number of shapes (at most max_rows are written).
 * mps_ctx_gemm_time re-launches the recorded GEMM of that row, on its own
 * operands and stream, iters times with the given tile (32 or 64) and split-K
 * depth, and returns the mean microseconds per launch; a depth whose partials
 * exceed the recorded call's workspace returns MPS_E_ARG before anything is
 * queued.  The re-launches overwrite the recorded outputs: time on scratch
 * workloads only. */
int mps_ctx_gemm_record(mps_ctx_t ctx, int enable);
int mps_ctx_gemm_recorded(mps_ctx_t ctx, int64_t* out, int max_rows, int* rows);
int mps_ctx_gemm_time(mps_ctx_t ctx, int row, int tm, int splits, int iters, double* us);

/* ------------------------------------------------------- states (a1, a3) */
/* Vector{Array{TN,3}} with N sites, local dims d[N], bond dims dims[N+1]
 * (mpsdims, MatrixProductStates.jl:41-51). Contents are zero.           */
int mps_state_create(mps_ctx_t ctx, int N, const int64_t* d, const int64_t* dims, mps_state_t* out);
int mps_state_destroy(mps_state_t s);
int mps_state_dims(mps_state_t s, int64_t* dims_out /* N+1 */);
int mps_state_upload_site(mps_state_t s, int site, const double* host, int64_t dl, int64_t d, int64_t dr);
int mps_state_download_site(mps_state_t s, int site, double* host, int64_t dl, int64_t d, int64_t dr);
/* The same with ComplexF32 host buffers (interleaved floats), the
 * TN = Complex{Float32} arrays of the reference's GPU examples
 * (rydberg_density_matrix_gpu.jl:124-125): widened to / narrowed from the
 * library's complex FP64 storage on the device.  Plain contexts only. */
int mps_state_upload_site_f32(mps_state_t s, int site, const float* host, int64_t dl, int64_t d, int64_t dr);
int mps_state_download_site_f32(mps_state_t s, int site, float* host, int64_t dl, int64_t d, int64_t dr);
/* dst .= copy.(src)  (e.g. spin_model_quantum_trajectory_rk4.jl:130) */
int mps_state_copy(mps_state_t dst, mps_state_t src);
/* mpsproductstate (constructors.jl:8-32, 37-61): local state vectors
 * states[n*d + s] (complex interleaved, per site), left canonical. */
int mps_state_product(mps_state_t s, const double* states);

/* ------------------------------------------------------------- MPOs (a1) */
int mps_mpo_create(mps_ctx_t ctx, int N, const int64_t* d, const int64_t* wdims /* N+1 */, mps_mpo_t* out);
int mps_mpo_destroy(mps_mpo_t o);
int mps_mpo_upload_site(mps_mpo_t o, int site, const double* host, int64_t wl, int64_t d, int64_t wr);
/* ComplexF32 MPO site (makempo of a Complex{Float32} TN), widened on the device. */
int mps_mpo_upload_site_f32(mps_mpo_t o, int site, const float* host, int64_t wl, int64_t d, int64_t wr);
/* O[site][wl_idx, :, wr_idx, :] = op (d x d) -- the per-step writes of
 * update_h! (spin_model_quantum_trajectory_rk4.jl:274-284) and
 * update_L_Ryd! (rydberg_density_matrix_gpu.jl:393-403). */
int mps_mpo_set_block(mps_mpo_t o, int site, int64_t wl_idx, int64_t wr_idx, const double* op);

/* -------------------------------------------------------- environments (a2) */
/* build_env(TN, TA, dims1, dims2, mpo_dims) (constructors.jl:347-359) */
int mps_env_create_op(mps_ctx_t ctx, int N, const int64_t* dims1, const int64_t* dims2,
                      const int64_t* wdims, mps_env_t* out);
/* build_env(TN, TA, dims1, dims2) (constructors.jl:363-375) */
int mps_env_create(mps_ctx_t ctx, int N, const int64_t* dims1, const int64_t* dims2, mps_env_t* out);
int mps_env_destroy(mps_env_t e);
int mps_env_download_slot(mps_env_t e, int slot, double* host, int64_t n_elems);

/* ------------------------------------- L1 contractions on raw device memory
 * (a4-a9, a23; src/mps_base.jl).  Pointers are device addresses of
 * interleaved complex-double column-major arrays; shapes as in the reference. */
/* prod_AOR (mps_base.jl:7-19): out[Dl,wl,d,Dr0] from A[Dl,d,Dr], O[wl,d,wr,d], R[Dr,wr,Dr0] */
int mps_dev_prod_AOR(mps_ctx_t ctx, void* out, const void* A, int64_t Dl, int64_t d, int64_t Dr,
                     const void* O, int64_t wl, int64_t wr, const void* R, int64_t Dr0);
/* prod_AOL (mps_base.jl:26-39): out[Dl0,d,Dr,wr] from A[Dl,d,Dr], O, L[Dl,wl,Dl0] */
int mps_dev_prod_AOL(mps_ctx_t ctx, void* out, const void* A, int64_t Dl, int64_t d, int64_t Dr,
                     const void* O, int64_t wl, int64_t wr, const void* L, int64_t Dl0);
/* prod_LR (3x4) (mps_base.jl:45-61): out[l3,r3,r4] = L[r1*r2,l3]^H * R[r1*r2, r3*r4] */
int mps_dev_prod_LR_34(mps_ctx_t ctx, void* out, const void* L, const void* R,
                       int64_t r1, int64_t r2, int64_t l3, int64_t r3, int64_t r4);
/* prod_LR (4x3) (mps_base.jl:65-81): out[l1,l2,r3] = L[l1*l2, l3*l4] * R[l3*l4, r3] */
int mps_dev_prod_LR_43(mps_ctx_t ctx, void* out, const void* L, const void* R,
                       int64_t l1, int64_t l2, int64_t l3, int64_t l4, int64_t r3);
/* prod_LR (2x3) (mps_base.jl:85-108): out = L[l1,r1] * R[r1, r2*r3]; co: L^T (92-118) */
int mps_dev_prod_LR_23(mps_ctx_t ctx, void* out, const void* L, const void* R,
                       int64_t l1, int64_t r1, int64_t r2, int64_t r3, int transpose_L);
/* prod_LR (3x2) (mps_base.jl:122-137): out = L[l1*l2, l3] * R[l3, r2] */
int mps_dev_prod_LR_32(mps_ctx_t ctx, void* out, const void* L, const void* R,
                       int64_t l1, int64_t l2, int64_t l3, int64_t r2);
/* update_re (mps_base.jl:142-160): out[r1,r2,a1] = AOR[r1*r2, r3*r4] * A0[a1, r3*r4]^H */
int mps_dev_update_re(mps_ctx_t ctx, void* out, const void* A0, const void* AOR,
                      int64_t r1, int64_t r2, int64_t r3, int64_t r4, int64_t a1);
/* update_le (mps_base.jl:223-241): out[l3,l4,a3] = AOL[l1*l2, l3*l4]^H * A0[l1*l2, a3] */
int mps_dev_update_le(mps_ctx_t ctx, void* out, const void* A0, const void* AOL,
                      int64_t l1, int64_t l2, int64_t l3, int64_t l4, int64_t a3);
/* update_renv (mps_base.jl:185-199): out[ar1, a1] = AR[ar1, d*a2] * A0[a1, d*a2]^H */
int mps_dev_update_renv(mps_ctx_t ctx, void* out, const void* A0, const void* AR,
                        int64_t a1, int64_t d, int64_t a2, int64_t ar1);
/* update_lenv (mps_base.jl:267-281): out[a3, al3] = A0[a1*d, a3]^H * AL[a1*d, al3]
 * mode 0: update_lenv; 1: update_lenv_co (283-289, AL^T conj(A0) -> [al3, a3]);
 * 2: update_lenv_nc (319-325, A0^T AL). */
int mps_dev_update_lenv(mps_ctx_t ctx, void* out, const void* A0, const void* AL,
                        int64_t a1, int64_t d, int64_t a3, int64_t al3, int mode);
/* local_exp (mps_base.jl:341-349): trace(O * A_p * A0_p^H) -> out (host, 2 doubles) */
int mps_dev_local_exp(mps_ctx_t ctx, double* out, const void* A0, const void* A,
                      int64_t a1, int64_t d, int64_t a3, const double* O_host);

/* ---------------------------------------------- L2 gauge moves (a10-a12) */
/* qrq! (orthonormalisation.jl:180-198): in place, the m x n matrix at dev
 * (element (r,c) at r*rs + c*cs) is replaced by the Q of its Householder QR
 * (zgeqrf + zungqr semantics, m >= n).  R (n x n, column-major, upper) is
 * written to R_out when non-NULL (Julia qr, used by leftorth/rightorth). */
int mps_dev_qrq(mps_ctx_t ctx, void* dev, int64_t m, int64_t n, int64_t rs, int64_t cs, void* R_out);
/* leftorth! / rightorth! (orthonormalisation.jl:164-171, 227-234) on one site in place. */
int mps_leftorth_site(mps_state_t s, int site);
int mps_rightorth_site(mps_state_t s, int site);
/* leftorthogonalizeQR! / leftorthonormalizationQR! (orthonormalisation.jl:29-56) */
int mps_leftorthogonalize_qr(mps_state_t s, double* norm2_out);
int mps_leftorthonormalize_qr(mps_state_t s, double* norm2_out);
/* rightorthogonalizeQR! / rightorthonormalizationQR! (orthonormalisation.jl:64-90) */
int mps_rightorthogonalize_qr(mps_state_t s, double* norm2_out);
int mps_rightorthonormalize_qr(mps_state_t s, double* norm2_out);
/* normalize_lo! (orthonormalisation.jl:97-103) and norm_lo (:117-121) */
int mps_normalize_lo(mps_state_t s, double* norm2_out);
int mps_norm_lo(mps_state_t s, double* norm2_out);

/* ------------------------------------------------ L3 sweeps (a13-a19, a21) */
/* compress_var_apply_H_left! (compression.jl:321-347) */
int mps_compress_var_apply_H_left(mps_state_t A0, mps_env_t envop, mps_state_t Ain,
                                  mps_state_t A, mps_mpo_t H);
/* compress_sum_var_apply_H_left! (compression.jl:889-944); coef: numv complex
 * (interleaved); A: numv states; env: numv-1 overlap environments. */
int mps_compress_sum_var_apply_H_left(mps_state_t A0, const mps_env_t* env, mps_env_t envop,
                                      mps_state_t Ain, int numv, const double* coef,
                                      const mps_state_t* A, mps_mpo_t H);
/* compress_sum_var_apply_H_right! (compression.jl:1007-1063).  A0 may be A[0]
 * (the aliasing of timeevolution.jl:178-186). */
int mps_compress_sum_var_apply_H_right(mps_state_t A0, const mps_env_t* env, mps_env_t envop,
                                       mps_state_t Ain, int numv, const double* coef,
                                       const mps_state_t* A, mps_mpo_t H);
/* compress_var_apply_H! (compression.jl:217-258) */
int mps_compress_var_apply_H(mps_state_t A0, mps_env_t envop, mps_state_t Ain, mps_state_t A,
                             mps_mpo_t H, int maxsweeps);
/* compress_var! (compression.jl:65-103) */
int mps_compress_var(mps_state_t A0, mps_state_t A, mps_env_t env, int maxsweeps);
/* compress_sum_var (compression.jl:496-565): result written to A0out (same
 * dims as Ain); Ain is the initial guess and may alias A0out. */
int mps_compress_sum_var(mps_state_t A0out, mps_state_t Ain, int numv, const double* coef,
                         const mps_state_t* A, int maxsweeps);
/* compress_sum_var_apply_H! (compression.jl:805-887), by intent (the
 * reference's environment initialisation at :816-817 throws). */
int mps_compress_sum_var_apply_H(mps_state_t A0, mps_env_t envop, int numv, const double* coef,
                                 const mps_state_t* A, mps_mpo_t H, int maxsweeps);
/* compressionSVD! (compression.jl:1280-1312): err_trim / ent_ent have N entries. */
int mps_compression_svd(mps_state_t s, int64_t dmax, double tol, double* err_glob,
                        double* err_trim, double* ent_ent);

/* ------------------------------------------------ L3 time steppers (a16, a20) */
/* Views of copy b of a batched container, on a companion context: they share
 * the container's memory and must be destroyed before it. */
int mps_state_view(mps_state_t s, int b, mps_ctx_t companion, mps_state_t* out);
int mps_env_view(mps_env_t e, int b, mps_ctx_t companion, mps_env_t* out);
int mps_mpo_view(mps_mpo_t m, int b, mps_ctx_t companion, mps_mpo_t* out);
/* normalize_lo! (orthonormalisation.jl:97-103) of every copy of a batched
 * state: norm2[b] (nb doubles) for copy b. */
int mps_normalize_lo_batch(mps_state_t s, double* norm2);

/* RK4stp_apply_H_half! (timeevolution.jl:178-196).  psi0 == NULL selects the
 * 10-argument form (result overwrites psi); otherwise the 11-argument form.
 * Returns once the step is queued on the context's stream; later calls on the
 * same context, downloads and scalar-returning calls order after it.  Small
 * sites run, from the second call with a given set of handles on, as one
 * captured graph, when every handle of the call was created on this context
 * (otherwise eagerly); sites of 32768 elements or more run eagerly, their overlap
 * terms on a second stream of the context, and the environment builds of
 * stages 2-4 on a third, behind the previous stage's sweep, into envop and a
 * second operator environment the context keeps; both streams are joined
 * back into the context's stream before the call returns.  Such a call also
 * builds, behind its last sweep, the stage-1 environments of a following
 * call with psi = this call's result and the same H1 and envop; that call
 * uses them only if no entry point that writes a state, MPO or environment
 * ran in between.  envop is scratch: its content after the call is not part
 * of the result. */
int mps_rk4stp_apply_H_half(double dt, mps_mpo_t H1, mps_mpo_t H2, mps_mpo_t H3,
                            mps_state_t psi0, mps_state_t psi, mps_state_t psi1,
                            mps_state_t psi2, mps_state_t psi3, mps_env_t env1,
                            mps_env_t env2, mps_env_t env3, mps_env_t envop);
/* Arnoldi_expm! (timeevolution.jl:88-119); coef_out (numv complex) optional. */
int mps_arnoldi_expm(mps_mpo_t A, int numv, const mps_state_t* rhov, mps_env_t env,
                     mps_env_t envop, double* coef_out);

/* Two-site variational compression with SVD truncation:
 * compress_var_two_site (compression.jl:110-164), compress_var_apply_H_two_site
 * {,_left,_right} (:355-486), compress_sum_var_two_site (:574-647) and
 * compress_sum_var_apply_H_two_site{,_left,_right} (:1072-1273).  The
 * reference's versions call an svdecon it never defines; here svdecon is the
 * thin SVD and the sweeps are built by intent.  Target: sum_l coef[l] A[l]
 * over the overlap terms (numv of them with H == NULL, numv-1 otherwise) plus
 * coef[numv-1] H A[numv-1].  mode 0: maxsweeps full sweeps (left, then right)
 * after left environments from Ain; 1: one left half-sweep (the _left
 * variants, result right canonical with U S on site 1); 2: one right
 * half-sweep after right environments from Ain (the _right variants, Ain
 * right canonical; result left canonical with S V^H on the last site).  Every
 * bond keeps min(#{k : cumsum(s^2)/sum(s^2) < 1 - svd_tol} + 1, dlim) values;
 * A0's bonds change accordingly (read them with mps_state_dims) within its
 * capacity, which must cover Ain's bonds.  env (overlap terms) and envop must
 * be sized for A0's capacity against each target. */
int mps_compress_two_site(mps_state_t A0, const mps_env_t* env, mps_env_t envop, mps_state_t Ain, int numv,
                          const double* coef, const mps_state_t* A, mps_mpo_t H, int mode, int maxsweeps,
                          double svd_tol, int64_t dlim);
/* steady_state_search! (steadystate.jl:6-86): sweeps of local
 * smallest-eigenvalue problems of <At| L^dag L |At> over the sites of At
 * (left canonical on entry, not necessarily normalised; overwritten with the
 * result).  W is the MPO of Lconj * L that the reference contracts through
 * its two-operator environments: W = applyMPOtoMPO(L, Lconj)
 * (MatrixProductStates.jl:67-99), bond w_L * w_Lconj; envop is an operator
 * environment for W's bonds.  Each local problem (the reference's
 * eigs(...; nev = 1, which = :SR, v0 = At[jj], tol = eigs_tol)) is solved
 * densely on the device for local spaces of at most 1536 elements (one-sided
 * Jacobi SVD of the positive semidefinite local operator), otherwise by a
 * Lanczos iteration from the current site to a residual below
 * eigs_tol * max(|theta|, the largest Ritz value).  Stops as the reference
 * does, after maxsweeps or when |ae| > converge_rat |ae at the sweep's first
 * site| or |ae| < 1e-10.  evalue_out: the last local eigenvalue;
 * sweeps_out: sweeps run. */
int mps_steady_state_search(mps_state_t At, mps_mpo_t W, mps_env_t envop, int maxsweeps, double converge_rat,
                            double eigs_tol, double* evalue_out, int* sweeps_out);

/* ------------------------------------ scalar products and observables (a22-a24) */
/* scal_prod (MatrixProductStates.jl:345-358): <A|B> */
int mps_scal_prod(mps_state_t A, mps_state_t B, double* out);
/* scal_op_prod (MatrixProductStates.jl:327-340): conj(<A|O|B>) */
int mps_scal_op_prod(mps_state_t A, mps_mpo_t O, mps_state_t B, double* out);
/* scal_prod_no_conj (MatrixProductStates.jl:366-379): bilinear A . B */
int mps_scal_prod_no_conj(mps_state_t A, mps_state_t B, double* out);
/* scal_prod_no_conj without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same result as mps_scal_prod_no_conj (its GEMMs do not split
 * K, so the last bits may differ).  It works in a fixed 24 MiB device buffer
 * of the library, one per GPU, held until the process ends: the two running
 * environments and the per-site product need
 * 32 (max_j dims_A[j] dims_B[j]) + 16 max_j (Dl_A(j) d_j Dr_B(j)) bytes
 * (D = 256, d = 9: about 11.5 MiB); larger inputs return MPS_E_ARG.  Calls
 * on one GPU are serialised by a lock and synchronise before returning, so
 * they never overlap on the buffer; nothing in it must start at zero.
 * MPS_E_STATE on a batched context or with kernel timing on. */
int mps_scal_prod_no_conj_noalloc(mps_state_t A, mps_state_t B, double* out);
/* measure_excitations! (spin_model_quantum_trajectory_rk4.jl:220-231):
 * out[n] = real(local_exp(A[n] R, A[n], op)) for a left-canonical, normalised state. */
int mps_measure_local(mps_state_t A, const double* op /* d x d complex */, double* out /* N */);
/* apply_site_operator! (MatrixProductStates.jl:135-143) */
int mps_apply_site_operator(mps_state_t A, int site, const double* op /* d x d complex */);

#ifdef __cplusplus
}
#endif
#endif /* MPSHIP_H */
