/*
 * kpm_hip.h -- C ABI of libkpm_hip.so, the MI355X (gfx950) Chebyshev moment
 * engine that replaces KPM.jl's CUDA package extension (ext/KPMCUDAExt.jl).
 *
 * Conventions
 *   - Every entry point returns an int status (KPM_OK = 0, negative on error);
 *     kpm_last_error() returns a thread-local message for the last failure.
 *   - Host blocks use the Julia caller's layout: probe blocks are NH x NR
 *     column-major, interleaved complex double (ComplexF64); mu_all is NR x NC
 *     column-major; the 2-D table mu is NC x NC column-major with
 *     mu[n + NC*m] = mu[n,m]; the 3-D table is NC x NC x NC column-major.
 *   - Sparse inputs are Julia SparseMatrixCSC arrays (colptr, rowval, nzval),
 *     int64 indices with index_base 0 or 1, interleaved complex double values.
 *   - One HIP stream per context; calls on one context must be serialised by
 *     the caller; different contexts are independent.
 *   - The library owns all device memory it allocates; handles are released
 *     with the matching *_destroy.
 */
#ifndef KPM_HIP_H
#define KPM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------
 * KPM_EINVAL   <-> ArgumentError / AssertionError in the reference
 *                  (src/moment.jl:431,434,438,512,515-519,767-807,913-973)
 * KPM_EUNIMPL  <-> throw("unimplemented") (src/moment.jl:734,882)
 * KPM_EUNSTABLE<-> error() of the recurrence stability guard
 *                  (src/utils/chebyshev_iteration.jl:100-132)             */
#define KPM_OK 0
#define KPM_EINVAL (-1)
#define KPM_ENOMEM (-2)
#define KPM_EHIP (-3)
#define KPM_EUNSTABLE (-4)
#define KPM_EUNIMPL (-5)

/* recurrence-vector storage: FP64 = ComplexF64 throughout (the reference's
 * dt_cplx, src/KPM.jl:5-7); FP32 = ComplexF32 vectors, FP64 arithmetic and
 * FP64 moment accumulation. */
#define KPM_VEC_F64 0
#define KPM_VEC_F32 1
/* kpm_bench_contract only: slots held as split-BF16 hi/lo planes */
#define KPM_VEC_PLANES 2

/* kpm_2d contraction arithmetic: FP64 MFMA; FP32 MFMA (FP64 promotion every
 * 256 reals of K); 3-pass split BF16 (hi*hi + hi*lo + lo*hi on BF16 MFMA,
 * FP32 accumulation promoted to FP64 every 256 reals) */
#define KPM_GEMM_F64 0
#define KPM_GEMM_F32 1
#define KPM_GEMM_BF16X3 2

/* moment_parity of kpm_2d! (src/moment.jl:913-915,1067-1069) */
#define KPM_PARITY_NONE 0
#define KPM_PARITY_ODD 1
#define KPM_PARITY_EVEN 2

/* indices of kpm_ctx_counters */
#define KPM_COUNT_STEP 0    /* fused recurrence passes (every mode) */
#define KPM_COUNT_DUAL 1    /* of which right-chain + Jb dual passes */
#define KPM_COUNT_DOT 2     /* dot-only passes (re-reads of two blocks) */
#define KPM_COUNT_COMBINE 3 /* separate moment-reduction launches */
#define KPM_COUNT_SYNC 4    /* host synchronisations of the moment paths */
#define KPM_COUNT_N 8

typedef struct kpm_ctx kpm_ctx;     /* device + stream + cached workspaces */
typedef struct kpm_csr kpm_csr;     /* device CSR operator (H, J_alpha, ...) */
typedef struct kpm_block kpm_block; /* device NH x NR recurrence block */

/* ---- errors / version ---------------------------------------------------- */
const char* kpm_last_error(void);
int kpm_abi_version(void);

/* ---- device layer: src/device.jl (ACTIVE_DEVICE, whichcore, device_zeros,
 *      device_free_memory) and ext/KPMCUDAExt.jl:13-20,43-45 ------------- */
/* number of visible GPUs (0 -> the caller stays on the CPU device,
 * the equivalent of CUDA.functional() == false at ext/KPMCUDAExt.jl:16) */
int kpm_device_count(int* count);
int kpm_ctx_create(int device, kpm_ctx** out);
int kpm_ctx_destroy(kpm_ctx* ctx);
int kpm_ctx_synchronize(kpm_ctx* ctx);
/* No counterpart in the reference (a switch of this engine).  enable = 1
 * (default): kpm_2d and kpm_3d compute each right-chain step and the Jb
 * product of the previous order in one pass over H and Jb merged onto the
 * union of their patterns (operators with sorted rows, two or more lanes per
 * row).  enable = 0: two separate passes.  The tables are the same either
 * way; the switch exists to show that. */
int kpm_ctx_set_dual_pass(kpm_ctx* ctx, int enable);
/* No counterpart in the reference (a switch of this engine).  enable = 1
 * (default): in the split-BF16 contraction mode, when every contraction of a
 * kpm_2d call has 128 or more slots a side, the recurrence passes write the
 * slots as hi/lo BF16 planes (8 bytes per complex value) and the contraction
 * reads them without a split.  enable = 0: FP64 / FP32 slots, split on every
 * read.  The tables are bitwise the same either way. */
int kpm_ctx_set_bf16_planes(kpm_ctx* ctx, int enable);
/* No counterpart in the reference (a limit of this engine).  The largest
 * merged operator, in bytes, the dual pass may allocate; a larger one makes
 * kpm_2d / kpm_3d run the two passes separately, as when the allocation
 * fails.  bytes < 0 (default): no limit. */
int kpm_ctx_set_merged_limit(kpm_ctx* ctx, int64_t bytes);
/* No counterpart in the reference (instrumentation).  Copies the first
 * min(n, KPM_COUNT_N) launch counters (KPM_COUNT_* indices) of the context;
 * reset != 0 then zeroes them. */
int kpm_ctx_counters(kpm_ctx* ctx, int64_t* counts, int n, int reset);
/* device_free_memory(::Dev) -- ext/KPMCUDAExt.jl:45, src/device.jl:66-67 */
int kpm_device_free_memory(kpm_ctx* ctx, size_t* free_bytes, size_t* total_bytes);
/* name of the device, e.g. for logs */
int kpm_device_name(kpm_ctx* ctx, char* buf, size_t len);

/* ---- operators: to_device(::Dev, ::SparseMatrixCSC, T) -- ext:26-27 ------
 * Builds the device CSR of A (not of its transpose) from Julia CSC arrays.
 * Values are stored complex FP64, or real FP64 when every imaginary part is
 * zero (the moments are unchanged; it halves the operator's bytes). */
int kpm_csr_from_csc(kpm_ctx* ctx, int64_t n, const int64_t* colptr, const int64_t* rowval,
                     const double* nzval, int index_base, kpm_csr** out);
/* Device CSR of adjoint(A) from the CSC arrays of A: the operator of
 * mul!(Jα_psi, Jα', ψ) at src/moment.jl:680. */
int kpm_csr_adjoint_from_csc(kpm_ctx* ctx, int64_t n, const int64_t* colptr,
                             const int64_t* rowval, const double* nzval, int index_base,
                             kpm_csr** out);
/* Device CSR from CSR arrays (row pointers / column indices), e.g. when the
 * caller already holds H in row-major sparse form. */
int kpm_csr_from_csr(kpm_ctx* ctx, int64_t n, const int64_t* rowptr, const int64_t* colind,
                     const double* vals, int index_base, kpm_csr** out);
int kpm_csr_destroy(kpm_csr* A);
int kpm_csr_info(const kpm_csr* A, int64_t* n, int64_t* nnz, int* is_real);

/* ---- recurrence blocks: maybe_on_device_zeros(dt_cplx, NH, NR) and
 *      maybe_to_device/maybe_to_host of probe blocks -- ext:28,36,43 ------- */
int kpm_block_create(kpm_ctx* ctx, int64_t NH, int64_t NR, int vec_prec, kpm_block** out);
int kpm_block_destroy(kpm_block* b);
/* host NH x NR column-major ComplexF64 -> device block (rounded for FP32) */
int kpm_block_upload(kpm_block* b, const double* host);
/* device block -> host NH x NR column-major ComplexF64 */
int kpm_block_download(const kpm_block* b, double* host);
int kpm_block_copy(kpm_block* dst, const kpm_block* src);
int kpm_block_zero(kpm_block* b);
/* random_phase_vectors + normalize_by_col on the device (src/frontend.jl:
 * 212-216, src/utils/vectors.jl:14-20; kpm_1d! draws its random probes on the
 * device, src/moment.jl:493-497): b = exp(2 pi i U) with unit-norm columns.
 * U is the lane-parallel xoshiro256++ stream of kpmhip.rng (lane j seeded by
 * splitmix64 outputs 4j..4j+3 of `seed`, value i of the column-major fill
 * from lane i % lanes), not Julia's Xoshiro. */
int kpm_block_random_phase(kpm_block* b, uint64_t seed, int64_t lanes);

/* ---- recurrence primitives: ext/KPMCUDAExt.jl:84-150 ---------------------
 * chebyshev_iter_single(H, V_pp, V_p[, V_out]) (chebyshev_iteration.jl:39-91,
 * 135-167): V_out <- 2 H V_p - V_pp.  V_out may be V_pp (in place).
 * The pass also reduces <V_p|V_p> and <V_p|V_out> (FP64, per probe column),
 * the two dots kpm_1d! asks for next (src/moment.jl:463-484); they are kept
 * for the columnwise_dot! calls that follow (kpm_accum_dot,
 * kpm_colwise_dot) until any call writes a block of this context. */
int kpm_cheb_step(kpm_ctx* ctx, const kpm_csr* H, const kpm_block* Vpp, const kpm_block* Vp,
                  kpm_block* Vout);
/* mul!(Y, A, X, alpha, beta) with real alpha, beta: Y <- alpha A X + beta Y;
 * like kpm_cheb_step it keeps <X|X> and <X|Y> for the next dots
 * (src/moment.jl:442-449: mul! then columnwise_dot! of mu0 and mu1). */
int kpm_spmm(kpm_ctx* ctx, const kpm_csr* A, double alpha, const kpm_block* X, double beta,
             kpm_block* Y);
/* columnwise_dot!(target, n, A, B; alpha, beta_col, beta_scale) with a HOST
 * target (src/utils/vectors.jl:103-117): target[i,n] = alpha <A_i|B_i> +
 * beta_scale target[i,beta_col]; the target is a host NR x NC column-major
 * ComplexF64 matrix, n and beta_col are 1-based (beta_col == 0: no beta
 * term).  FP64 accumulation; synchronises with the host. */
int kpm_colwise_dot(kpm_ctx* ctx, double* target, int64_t NC, int64_t n, const kpm_block* A,
                    const kpm_block* B, double alpha, int64_t beta_col, double beta_scale);

/* ---- device moment accumulator: moment_accumulator / columnwise_dot! /
 *      copy_moment_accumulator! on device arrays (ext/KPMCUDAExt.jl:125-150,
 *      src/utils/vectors.jl:98-122) ---------------------------------------
 * A zeroed NR x NC ComplexF64 matrix in device memory.  kpm_accum_dot is
 * columnwise_dot! into it: when <A|B> is a dot the last kpm_cheb_step /
 * kpm_spmm of this context reduced (A = its V_p, B = V_p or V_out) and no
 * block was written since, the call only records a combine of the cached
 * partials, which runs inside the next primitive step's launch (or in one
 * small launch at the latest at kpm_accum_download); otherwise it runs one
 * dot pass and one combine.  It never synchronises with the host.
 * kpm_accum_download copies the matrix to the host once. */
typedef struct kpm_accum kpm_accum;
int kpm_accum_create(kpm_ctx* ctx, int64_t NR, int64_t NC, kpm_accum** out);
int kpm_accum_destroy(kpm_accum* acc);
int kpm_accum_dot(kpm_ctx* ctx, kpm_accum* acc, int64_t n, const kpm_block* A, const kpm_block* B,
                  double alpha, int64_t beta_col, double beta_scale);
int kpm_accum_download(kpm_ctx* ctx, const kpm_accum* acc, double* host);

/* ---- moment engine: src/moment.jl ----------------------------------------
 * kpm_1d! equal-vector doubling form (src/moment.jl:420-492): NC even,
 * NC/2 fused SpMM+dot passes; mu_all NR x NC column-major ComplexF64 out. */
int kpm_1d(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t NR, int64_t NH,
           const double* psi, double* mu_all, int vec_prec);
/* kpm_1d! bra/ket form (src/moment.jl:498-556): NC >= 2, NC-1 passes. */
int kpm_1d_lr(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t NR, int64_t NH,
              const double* psi_l, const double* psi_r, double* mu_all, int vec_prec);
/* kpm_1d_current! (src/moment.jl:656-714): Jadj is the device adjoint of
 * J_alpha (kpm_csr_adjoint_from_csc). */
int kpm_1d_current(kpm_ctx* ctx, const kpm_csr* H, const kpm_csr* Jadj, int64_t NC, int64_t NR,
                   int64_t NH, const double* psi, double* mu_all, int vec_prec);
/* Same engines on probe blocks already resident in device memory (the
 * block is read, not modified). */
int kpm_1d_block(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, const kpm_block* psi,
                 double* mu_all);
int kpm_1d_lr_block(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, const kpm_block* psi_l,
                    const kpm_block* psi_r, double* mu_all);
/* kpm_1d with avg_output = true on equal vectors (src/moment.jl:339-418 with
 * :410-416): mu_avg (NC ComplexF64) = sum(mu_all, dims = 1) ./ NR, the moments
 * dos_moments stores (src/frontend.jl:254-256).  The doubling passes sum the
 * moment dots over the probe columns as they go (a few bytes of partials per
 * wave instead of NR columns per workgroup), so mu_all is never formed.
 * Same arguments and errors as kpm_1d / kpm_1d_block otherwise. */
int kpm_1d_avg(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t NR, int64_t NH,
               const double* psi, double* mu_avg, int vec_prec);
int kpm_1d_block_avg(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, const kpm_block* psi,
                     double* mu_avg);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory while they run (code under HIP graph capture, serving loops).
 * kpm_1d_block with every buffer supplied: psi (read), the two recurrence
 * workspaces w0, w1 (overwritten; the alpha_all of src/moment.jl:428-429),
 * same shape and precision as psi, and mu, a device accumulator of at least
 * NR x NC (kpm_accum_create) that receives the moments.  The dot partials
 * live in the library's static 8 MiB scratch array on each device, so NRp
 * may be at most 16,384 (a clear KPM_EINVAL beyond).  Calls that overlap on
 * one device share that array and race.  Launches only: no allocation, no
 * host synchronisation, no memset or memcpy. */
int kpm_1d_block_noalloc(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, const kpm_block* psi,
                         kpm_block* w0, kpm_block* w1, kpm_accum* mu);

/* kpm_2d_blocking (src/moment.jl:759-810).  budget < 0 means "auto" (50% of
 * free device memory, src/moment.jl:741); arr_size/right_block <= 0 mean :auto. */
int kpm_2d_blocking(kpm_ctx* ctx, int64_t NH, int64_t NR, int64_t NC, int64_t budget,
                    int64_t arr_size, int64_t right_block, int64_t* K_left, int64_t* K_right,
                    int64_t* bytes);
/* kpm_2d! / _kpm_2d_blocked! (src/moment.jl:887-1088):
 * mu[n,m] = (1/NR) sum_r <T_{m-1} psi_l_r | Jb T_{n-1} Ja psi_r_r>,
 * blocked as K_left left slots x K_right right slots, contraction on the
 * MFMA cores.  mu is NC x NC column-major ComplexF64 (overwritten). */
int kpm_2d(kpm_ctx* ctx, const kpm_csr* H, const kpm_csr* Ja, const kpm_csr* Jb, int64_t NC,
           int64_t NR, int64_t NH, const double* psi_l, const double* psi_r, int64_t K_left,
           int64_t K_right, int parity, int mn_sym, int vec_prec, int gemm_mode, double* mu);
/* kpm_3d! (src/moment.jl:1106-1227): mu NC x NC x NC column-major, one
 * kpm_2d table per outer order n1, mu[:,:,n1] with Jgamma as the inner Ja. */
int kpm_3d(kpm_ctx* ctx, const kpm_csr* H, const kpm_csr* Ja, const kpm_csr* Jb,
           const kpm_csr* Jg, int64_t NC, int64_t NR, int64_t NH, const double* psi_l,
           const double* psi_r, int64_t K_left, int64_t K_right, int vec_prec, int gemm_mode,
           double* mu);

/* ---- matrix-function action: src/utils/chebyshev_action.jl ------------------
 * chebyshev_action! (src/utils/chebyshev_action.jl:47-125) with the fused
 * accumulation broadcast_assign! (src/utils/chebyshev_action.jl:163-178;
 * ext/KPMCUDAExt.jl:273-299) and the relative stability guard
 * _check_chebyshev_columns (src/utils/chebyshev_iteration.jl:119-132):
 *   out[:, :, k] = sum_{n=1..NC} C[n, k] T_{n-1}(H) V,   k = 1..K.
 * V: NH x NR column-major complex; C: NC x K column-major complex, used
 * verbatim; out: NH x NR x K column-major complex, overwritten.  The
 * recurrence runs in vec_prec, the accumulation in FP64.  check_every = 0
 * disables the guard; a tripped guard returns KPM_EUNSTABLE with the
 * reference's message (iteration and growth) and leaves out unspecified. */
int kpm_cheb_action(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t K, int64_t NR, int64_t NH,
                    const double* V, const double* C, double* out, int vec_prec,
                    int64_t check_every);
/* Device-resident form (the reference's CuArray method with out resident
 * with Hn): V a recurrence block (its precision is the recurrence's), out an
 * FP64 block of (K*NH) x NR whose rows [k*NH, (k+1)*NH) receive out[:, :, k]. */
int kpm_cheb_action_block(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t K, const kpm_block* V,
                          const double* C, kpm_block* out, int64_t check_every);

/* ---- spectral edges for normalizeH (src/utils/external.jl:41-76) -------------
 * The reference estimates the extreme eigenvalues with Arpack eigs on the
 * host (tol 1e-3, maxiter 300); this runs an FP64 Lanczos recurrence on the
 * device from a seeded random start vector and returns the extreme Ritz
 * values once both move by less than tol * max(|emin|, |emax|) over 10 steps
 * (or at an invariant subspace, or after max_iter steps).  iters may be NULL. */
int kpm_spectral_edges(kpm_ctx* ctx, const kpm_csr* H, int64_t max_iter, double tol, uint64_t seed,
                       double* emin, double* emax, int64_t* iters);

/* ---- reconstruction heads (FP64, device) ------------------------------------
 * chebyshev_lin_trans (src/utils/chebyshev_lintrans.jl:12-27; the device
 * method ext/KPMCUDAExt.jl:207-236) as used by dos (src/applications/dos.jl:
 * 128-133): y[i] = sum_j mu_tilde[j] cos(n_grid[j] acos(x[i])), x in [-1, 1];
 * mu_tilde and y complex (interleaved), host arrays. */
int kpm_cheb_lin_trans(kpm_ctx* ctx, int64_t Nx, const double* x, int64_t Nn, const int64_t* n_grid,
                       const double* mu_tilde, double* y);
/* The quadrature-node contraction of kubo_bastin_cond
 * (src/applications/conductivity.jl:133-155): for the kernel-weighted table
 * mu_tilde (NC x NC complex column-major) and K nodes x_k in (-1, 1) with
 * weights phi_k, acc = sum_k phi_k sum_nm mu_tilde[n,m] Gamma_nm(x_k), with
 * Gamma_nm(x) = (x - i m s) e^{i m th} T_n(x) + (x + i n s) e^{-i n th} T_m(x).
 * acc: one complex value.  The caller applies -2 NH / (area a^2) Re(acc). */
int kpm_kubo_bastin_nodes(kpm_ctx* ctx, int64_t NC, const double* mu_tilde, int64_t K,
                          const double* x, const double* phi, double* acc);

/* A kernel-weighted kpm_3d table resident on the device for the node
 * contractions of cpge / d_cpge (src/applications/cpge.jl:66-120, where the
 * reference holds mu_matrix = reshape(mu_tilde, NC, NC^2) for every node).
 * mu_tilde: NC^3 complex, n fastest, then m, then p (Julia column-major
 * mu[n,m,p]); copied during the call.  Destroy a table before its context. */
typedef struct kpm_table3 kpm_table3;
int kpm_table3_create(kpm_ctx* ctx, int64_t NC, const double* mu_tilde, kpm_table3** out);
int kpm_table3_destroy(kpm_table3* t);
/* out[q] = sum_{n,m,p} mu_tilde[n,m,p] a[q,n] b[q,m] c[q,p] for q < Q: the
 * _cpge_contract! of src/applications/cpge.jl:59-64 (collapsed = mu^T slot1,
 * right = reshape(collapsed) slot3, sum(slot2 .* right); plain transposes)
 * with a = slot1, b = slot2, c = slot3.  a, b, c: Q x NC complex, row q
 * contiguous; out: Q complex.  Q <= 2^20. */
int kpm_trilinear_forms(kpm_table3* t, int64_t Q, const double* a, const double* b, const double* c,
                        double* out);
/* The same forms with coefficient rows generated on the device from one
 * 5-double descriptor per row (desc_x: Q x 5, row-major):
 *   (0, c, -, -, -)       delta row: d[0] = 1/pi, d[1] = c/pi,
 *                         d[n] = 2c d[n-1] - d[n-2]  (chebyshev_delta_theta!,
 *                         src/applications/spectral_coefficients.jl:22-33)
 *   (1, g0r, g0i, qr, qi) resolvent row: g[0] = g0, g[n] = g[n-1] q
 *                         (green_coefficients!, spectral_coefficients.jl:84-86)
 *   (2, -, -, -, -)       zero row (the delta-annulus exclusion,
 *                         src/applications/cpge.jl:30-36)
 * so a quadrature round ships O(Q) scalars instead of O(Q NC) values. */
int kpm_trilinear_forms_rec(kpm_table3* t, int64_t Q, const double* desc_a, const double* desc_b,
                            const double* desc_c, double* out);
/* kpm_table3_create from the raw table mu (NC^3, n fastest) and the per-axis
 * weight k[n] = kernel(n, NC) h_n: the device applies
 * mu~ = ((mu k[n]) k[m]) k[p], the axis order of muND_apply_kernel_and_h
 * (src/kernels/kernels.jl:92-103; cpge.jl:122-125). */
int kpm_table3_create_weighted(kpm_ctx* ctx, int64_t NC, const double* mu, const double* kvec,
                               kpm_table3** out);
/* Measurement helper (no reference counterpart): total ms of `reps`
 * kpm_trilinear_forms launches (contraction + fixed-order reduce, no
 * transfers) of Q forms with device-generated unit-modulus coefficients,
 * after one untimed launch. */
int kpm_bench_trilinear(kpm_table3* t, int64_t Q, int64_t reps, float* ms);

/* ---- reverse-mode gradient ---------------------------------------------------
 * No counterpart in the reference.  For the equal-vector moments
 * mu_n = (1/NR) sum_r <psi_r|T_n(H)|psi_r> (n < NC; the moments of
 * kpm_1d!, src/moment.jl:420-492) and g[n] = dL/dmu_n (NC reals), writes
 * grad[k] = dL/dH_k for every stored entry k of H, in the order of the CSC
 * arrays H was built from (interleaved complex; dL = sum_k grad[k] dH_k).
 * Reverse Chebyshev (Clenshaw) recurrence.  ckpt_stride 1 keeps every
 * forward vector; K > 1 keeps checkpoint pairs every K steps and recomputes
 * one segment at a time; 0 keeps every vector when they fit in 60% of the
 * free device memory (plus this context's cached adjoint workspace) and
 * uses K = ceil(sqrt(NC)) otherwise.  The workspace stays cached in the
 * context (grow-only, released by kpm_ctx_destroy), so repeated gradient
 * steps allocate it once.  H must come from kpm_csr_from_csc. */
int kpm_1d_adjoint(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t NR, int64_t NH,
                   const double* psi, const double* g, double* grad, int vec_prec,
                   int64_t ckpt_stride);
/* No counterpart in the reference.  Reverse mode through the kpm_2d table
 * mu[n,m] = (1/NR) sum_r <T_m(H) psi_l_r | Jb T_n(H) Ja psi_r_r> (n, m < NC;
 * the table of _kpm_2d_blocked!, src/moment.jl:987-1088): for the cotangent
 * gbar (NC x NC complex column-major, dL = Re sum_nm conj(gbar[n,m])
 * dmu[n,m]) writes the derivatives on the stored entries of H, Ja and Jb in
 * the order of the CSC arrays each was built from (interleaved complex,
 * dL = Re sum_k grad[k] dA_k).  Reverse Clenshaw over both chains; keeps all
 * 2 NC chain vectors and two mixes of them (4 NC blocks, FP64), cached in
 * the context.  H, Ja, Jb from kpm_csr_from_csc; Jb_adj = adjoint(Jb) from
 * kpm_csr_adjoint_from_csc.  psi_l, psi_r: NH x NR column-major complex. */
int kpm_2d_adjoint(kpm_ctx* ctx, const kpm_csr* H, const kpm_csr* Ja, const kpm_csr* Jb,
                   const kpm_csr* Jb_adj, int64_t NC, int64_t NR, int64_t NH, const double* psi_l,
                   const double* psi_r, const double* gbar, double* grad_H, double* grad_Ja,
                   double* grad_Jb);
/* The two halves of kpm_1d_adjoint, so a loss can be formed from the
 * moments of the same forward pass (no reference counterpart).
 * kpm_1d_adjoint_forward runs the forward recurrence once, keeps its
 * vectors (or checkpoints, as ckpt_stride says) in the context and writes
 * the moments mu_all[r, n] = <psi_r|T_n(H)|psi_r> (NR x NC column-major
 * complex, the equal-vector moments of kpm_1d! as bra/ket dots of every
 * order).  kpm_1d_adjoint_reverse then takes g = dL/dmu_n (NC reals,
 * probe-averaged moments) and writes grad as kpm_1d_adjoint does; it needs
 * the same operator, and any later forward pass replaces the kept one.
 * Gradients are deterministic: rows wider than 64 words sum their row
 * chunks in a fixed order. */
int kpm_1d_adjoint_forward(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t NR, int64_t NH,
                           const double* psi, double* mu_all, int vec_prec, int64_t ckpt_stride);
int kpm_1d_adjoint_reverse(kpm_ctx* ctx, const kpm_csr* H, const double* g, double* grad);
/* Host bookkeeping for per-bond gradients (no reference counterpart): for a
 * CSC pattern with sorted rows holding both (i, j) and (j, i), and grad on that
 * pattern (the output of kpm_1d_adjoint), write every pair i < j with
 * dL/dt_ij = Re(u M_ij + conj(u) M_ji) / a, u = H_ij / |H_ij|.  bi, bj, val
 * need room for nnz entries; nbonds receives the count (0-based indices). */
int kpm_bond_gradient(int64_t n, const int64_t* colptr, const int64_t* rowval, const double* h,
                      const double* grad, double a, int64_t* bi, int64_t* bj, double* val,
                      int64_t* nbonds);

/* ---- batched realisations ---------------------------------------------------
 * No single-call counterpart in the reference, which runs one disorder
 * realisation per process and calls kpm_1d! on each
 * (scripts/anderson_seebeck_ensemble.jl:1-20, src/moment.jl:420-492).
 * kpm_batch_create packs nreal independent operators, given as ONE
 * block-diagonal Julia CSC matrix with realisation r on rows/columns
 * [row_offsets[r], row_offsets[r+1]), into a padded device CSR. */
typedef struct kpm_batch kpm_batch;
int kpm_batch_create(kpm_ctx* ctx, int nreal, const int64_t* row_offsets, const int64_t* colptr,
                     const int64_t* rowval, const double* nzval, int index_base, kpm_batch** out);
int kpm_batch_destroy(kpm_batch* B);
/* kpm_1d! equal-vector doubling moments of every realisation in one fused
 * launch per step.  psi: the stacked probes (row_offsets[nreal] x NR,
 * column-major ComplexF64); mu: nreal blocks of NR x NC column-major
 * ComplexF64.  use_graph != 0 records the NC/2 steps once into a HIP graph
 * and replays it on later calls of the same shape. */
int kpm_batch_1d(kpm_ctx* ctx, kpm_batch* B, int64_t NC, int64_t NR, const double* psi, double* mu,
                 int vec_prec, int use_graph);
/* The two halves of kpm_batch_1d: stage the stacked probes into a resident
 * padded device block, then run the recurrence from it (the probes stay
 * resident for repeated runs). */
int kpm_batch_upload(kpm_ctx* ctx, kpm_batch* B, int64_t NR, const double* psi, int vec_prec);
int kpm_batch_run(kpm_ctx* ctx, kpm_batch* B, int64_t NC, double* mu, int use_graph);
/* Launch-shape tuner of the batched pass (no reference counterpart): times
 * every candidate on the uploaded probes and records the fastest. */
int kpm_batch_autotune(kpm_ctx* ctx, kpm_batch* B, int reps, int* best_unroll, int* best_grid,
                       float* best_ms_per_step);
int kpm_batch_info(const kpm_batch* B, int64_t* padded_rows, int64_t* nnz, int* nseg);
/* Instrumentation (no reference counterpart): how many times kpm_batch_run
 * has recorded its step graph.  The graph is recorded again whenever the
 * shape or a device buffer it uses changes (probe upload at another width,
 * a new launch schedule). */
int kpm_batch_graph_builds(const kpm_batch* B, int64_t* builds);

/* ---- measurement ---------------------------------------------------------
 * Launch the fused recurrence+doubling-dot pass `nsteps` times on the
 * context stream over blocks X (v_n) and Z (v_{n-1}, overwritten), bracketed
 * by HIP events on that stream; *ms is the elapsed time of all passes.
 * The output is not meaningful (timing only). */
int kpm_bench_step(kpm_ctx* ctx, const kpm_csr* H, kpm_block* X, kpm_block* Z, int64_t nsteps,
                   float* ms);
/* kpm_bench_step for either doubling pass: mode 0 the per-probe pass of
 * kpm_1d (MODE_DBL), mode 1 the probe-summed pass of kpm_1d_avg. */
int kpm_bench_step_mode(kpm_ctx* ctx, const kpm_csr* H, kpm_block* X, kpm_block* Z,
                        int64_t nsteps, int mode, float* ms);
/* Streaming copy src -> dst (same shape) `reps` times on the context
 * stream, bracketed by HIP events: the plain copy kernel of known size that
 * calibrates the HBM byte counters and measures the bandwidth ceiling. */
int kpm_bench_copy(kpm_ctx* ctx, kpm_block* dst, const kpm_block* src, int64_t reps, float* ms);
/* Measurement helper (no reference counterpart): total ms of `reps`
 * back-to-back kpm_2d contractions of nL x nR slots of K complex values
 * (K a multiple of 16; the _kpm_2d_blocked! GEMM, src/moment.jl:1055-1061),
 * after one untimed launch, on scratch slots the call allocates, fills with
 * pseudo-random values in [-0.5, 0.5) (the MFMA kernels run measurably
 * faster on all-zero operands) and frees.  pitch: complex values between
 * consecutive slots (>= K, a multiple of 8); 0 selects kpm_2d's own pitch, K
 * plus 256 bytes.  vec_prec KPM_VEC_PLANES (with gemm_mode KPM_GEMM_BF16X3,
 * nL and nR >= 128): slots stored as split-BF16 planes, as kpm_2d writes them
 * in that mode (kpm_ctx_set_bf16_planes). */
int kpm_bench_contract(kpm_ctx* ctx, int64_t nL, int64_t nR, int64_t K, int64_t pitch, int vec_prec,
                       int gemm_mode, int64_t reps, float* ms);
/* HIP-event timer on the context stream: record into slot 0..15, read the
 * elapsed milliseconds between two recorded slots. */
int kpm_event_record(kpm_ctx* ctx, int slot);
int kpm_event_elapsed(kpm_ctx* ctx, int slot_a, int slot_b, float* ms);

/* ---- launch-shape tuner ---------------------------------------------------
 * No counterpart in the reference (it has no tuning knob).  Times every
 * candidate geometry of the fused pass (rows in flight per lane x either a
 * persistent grid of 2..16 workgroups per CU or a one-shot grid of 2..64
 * row tiles per workgroup) on the caller's operator and blocks X, Z (Z is
 * overwritten), records the fastest for this shape and returns it
 * (best_grid: the workgroups it launches).  The probe-summed doubling pass
 * and the plain pass without dots are timed over the same candidates and
 * recorded on their own (kpm_tune_get_mode).  X and Z should hold data like the engine's (random
 * phases, say): the pass's speed depends on the values, and all-zero blocks
 * rank the schedules differently. */
int kpm_autotune_step(kpm_ctx* ctx, const kpm_csr* H, kpm_block* X, kpm_block* Z, int reps,
                      int* best_unroll, int* best_grid, float* best_ms_per_step);
/* Set the table entry for a shape directly (unroll 1 or 2; grid 0 = auto). */
int kpm_tune_set(kpm_ctx* ctx, const kpm_csr* H, int64_t NR, int vec_prec, int unroll, int grid);
/* Set a one-shot schedule for a shape: one workgroup per `chunk` consecutive
 * row tiles, chunks in dispatch order (xcd_grouped = 0), each XCD sweeping
 * its own contiguous eighth (xcd_grouped = 1), or in dispatch order with each
 * XCD taking runs of m = xcd_grouped consecutive chunks (2 .. 64), which keeps
 * one front and puts a row's near neighbours in the same XCD's L2.  The engines fall back to the
 * persistent schedule when the grid would exceed their partials buffers. */
int kpm_tune_set_chunked(kpm_ctx* ctx, const kpm_csr* H, int64_t NR, int vec_prec, int unroll,
                         int chunk, int xcd_grouped);
/* The launch shape the engines use for a shape (chunk 0: persistent grid). */
int kpm_tune_get(kpm_ctx* ctx, const kpm_csr* H, int64_t NR, int vec_prec, int* unroll, int* grid,
                 int* chunk, int* xcd_grouped);
/* The same for each pass the tuner times separately: mode 0 the per-probe
 * passes with dots (kpm_tune_get), mode 1 the probe-summed doubling pass of
 * kpm_1d_avg / kpm_1d_block_avg, mode 2 the plain passes without dots
 * (kpm_2d / kpm_3d chains, the matrix-function action, the adjoint's source
 * steps).  kpm_tune_set and kpm_tune_set_chunked set all three. */
int kpm_tune_get_mode(kpm_ctx* ctx, const kpm_csr* H, int64_t NR, int vec_prec, int mode,
                      int* unroll, int* grid, int* chunk, int* xcd_grouped);

#ifdef __cplusplus
}
#endif
#endif /* KPM_HIP_H */
