/*
 * kpm_oracle.c -- CPU restatement of KPM.jl's Chebyshev moment engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the CPU
 * baseline; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product path (libkpm_hip.so) never links or calls it.
 *
 * What it restates (reference files, read-only, cited file:line):
 *   - SparseArrays' 5-argument CSC mul! y <- a*H*x + b*y, per probe column,
 *     threaded over columns above 2^14 elements
 *       src/utils/chebyshev_iteration.jl:51-75
 *   - columnwise_dot!: target[i,n] = alpha*<A_i|B_i> + beta_scale*target[i,beta_col]
 *       src/utils/vectors.jl:103-117
 *   - kpm_1d! equal-vector doubling       src/moment.jl:420-492
 *   - kpm_1d! bra/ket                     src/moment.jl:498-556
 *   - kpm_1d_current!                     src/moment.jl:656-714
 *   - kpm_2d_blocking (integer planner)   src/moment.jl:759-810
 *   - _kpm_2d_blocked!                    src/moment.jl:987-1088
 *   - kpm_3d!                             src/moment.jl:1106-1227
 *   - chebyshev_action! + broadcast_assign! + the relative stability guard
 *       src/utils/chebyshev_action.jl:47-125,163-178
 *       src/utils/chebyshev_iteration.jl:119-132
 *
 * Layout conventions follow the Julia caller: probe blocks are NH x NR
 * column-major interleaved complex doubles; mu_all is NR x NC column-major;
 * the 2-D table is NC x NC column-major with mu[n + NC*m] = mu[n,m] (0-based);
 * the 3-D table is NC^3 column-major.  H, J are CSC with 0-based int64
 * colptr/rowval and interleaved complex nzval (Julia's SparseMatrixCSC).
 *
 * Every function returns 0 on success and a negative code on invalid input.
 * Parity pins (closed forms, dense references, ED) live in tests/.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define OR_EINVAL (-1)
#define OR_ENOMEM (-2)

/* chebyshev_iteration.jl:53 -- below this many elements run serially. */
#define CHEB_SERIAL_CUTOFF (1 << 14)

static int g_threads = 0; /* 0: OpenMP default */

void kpm_oracle_set_threads(int n) { g_threads = n; }

int kpm_oracle_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

/* SparseArrays CSC 5-arg mul!: y <- beta*y (unless beta == 1), then for each
 * column j: ax = x[j]*alpha; for k in nzrange(j): y[row[k]] += nz[k]*ax. */
static void csc_mul5(int64_t n, const int64_t* colptr, const int64_t* rowval,
                     const cplx* nz, const cplx* x, cplx* y, cplx alpha, cplx beta) {
    if (beta != 1.0) {
        if (beta == 0.0) {
            for (int64_t i = 0; i < n; ++i) y[i] = 0.0;
        } else {
            for (int64_t i = 0; i < n; ++i) y[i] *= beta;
        }
    }
    for (int64_t j = 0; j < n; ++j) {
        const cplx ax = x[j] * alpha;
        for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k) y[rowval[k]] += nz[k] * ax;
    }
}

/* mul!(y, adjoint(A), x): y[j] = sum_k conj(nz[k]) * x[row[k]] per column j. */
static void csc_adjoint_mul(int64_t n, const int64_t* colptr, const int64_t* rowval,
                            const cplx* nz, const cplx* x, cplx* y) {
    for (int64_t j = 0; j < n; ++j) {
        cplx acc = 0.0;
        for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k) acc += conj(nz[k]) * x[rowval[k]];
        y[j] = acc;
    }
}

typedef struct {
    int64_t n;
    const int64_t* colptr;
    const int64_t* rowval;
    const cplx* nz;
} csc_t;

/* Block SpMM over NR columns: V_out[:,c] <- alpha*H*V_in[:,c] + beta*V_out[:,c].
 * chebyshev_iteration.jl:59-75 threads over the columns above the cutoff. */
static void block_mul5(const csc_t* H, int64_t NH, int64_t NR, const cplx* Vin, cplx* Vout,
                       cplx alpha, cplx beta) {
    const int serial = kpm_oracle_get_threads() == 1 || NH * NR <= CHEB_SERIAL_CUTOFF;
    if (serial) {
        for (int64_t c = 0; c < NR; ++c)
            csc_mul5(H->n, H->colptr, H->rowval, H->nz, Vin + c * NH, Vout + c * NH, alpha, beta);
    } else {
#pragma omp parallel for schedule(static) num_threads(kpm_oracle_get_threads())
        for (int64_t c = 0; c < NR; ++c)
            csc_mul5(H->n, H->colptr, H->rowval, H->nz, Vin + c * NH, Vout + c * NH, alpha, beta);
    }
}

/* chebyshev_iter_single(H, V_pp, V_p): V_pp <- 2 H V_p - V_pp. */
static void cheb_single(const csc_t* H, int64_t NH, int64_t NR, cplx* Vpp, const cplx* Vp) {
    block_mul5(H, NH, NR, Vp, Vpp, 2.0, -1.0);
}

/* 3-address form (chebyshev_iteration.jl:157-167): V_out <- V_pp (unless the
 * same buffer), then the in-place step on V_out. */
static void cheb_single3(const csc_t* H, int64_t NH, int64_t NR, const cplx* Vpp, const cplx* Vp,
                         cplx* Vout) {
    if (Vout != Vpp) memcpy(Vout, Vpp, sizeof(cplx) * (size_t)(NH * NR));
    cheb_single(H, NH, NR, Vout, Vp);
}

/* LinearAlgebra.dot(a, b) = sum conj(a_i) b_i.  The reference reaches BLAS
 * zdotc, whose summation order is the library's own (not reproducible
 * here); this sums blocks of 64 terms in order and adds the block sums
 * pairwise, so the rounding error grows as log n rather than n (a strictly
 * sequential sum drifts by ~1e-12 on 1e5-element probe columns). */
static cplx vdot_seq(int64_t n, const cplx* a, const cplx* b) {
    cplx acc = 0.0;
    for (int64_t i = 0; i < n; ++i) acc += conj(a[i]) * b[i];
    return acc;
}
static cplx vdot(int64_t n, const cplx* a, const cplx* b) {
    if (n <= 64) return vdot_seq(n, a, b);
    const int64_t h = (n / 2 + 63) / 64 * 64;  /* split on a block boundary */
    return vdot(h, a, b) + vdot(n - h, a + h, b + h);
}

/* columnwise_dot! (vectors.jl:103-117), target NR x NC column-major, n and
 * beta_col are 1-based like the reference (beta_col == 0: no beta term). */
static void colwise_dot(cplx* target, int64_t NR, int64_t n, const cplx* A, const cplx* B,
                        int64_t NH, cplx alpha, int64_t beta_col, cplx beta_scale) {
    for (int64_t i = 0; i < NR; ++i) {
        cplx beta = beta_col == 0 ? 0.0 : beta_scale * target[i + NR * (beta_col - 1)];
        target[i + NR * (n - 1)] = alpha * vdot(NH, A + i * NH, B + i * NH) + beta;
    }
}

/* kpm_1d! equal-vector doubling form, moment.jl:420-492. */
int kpm_oracle_1d(int64_t NH, const int64_t* colptr, const int64_t* rowval, const double* nzval,
                  int64_t NC, int64_t NR, const double* psi, double* mu_all) {
    if (NH <= 0 || NR <= 0 || NC < 2 || (NC % 2) != 0) return OR_EINVAL; /* moment.jl:434 */
    csc_t H = {NH, colptr, rowval, (const cplx*)nzval};
    cplx* mu = (cplx*)mu_all;
    const size_t blk = (size_t)(NH * NR);
    cplx* a_all = (cplx*)calloc(2 * blk, sizeof(cplx));
    if (!a_all) return OR_ENOMEM;
    cplx* views[2] = {a_all, a_all + blk};
    memcpy(views[0], psi, sizeof(cplx) * blk);                 /* :440 */
    block_mul5(&H, NH, NR, views[0], views[1], 1.0, 0.0);       /* :442 mul! */
    colwise_dot(mu, NR, 1, views[0], views[0], NH, 1.0, 0, 0.0); /* :448 */
    colwise_dot(mu, NR, 2, views[0], views[1], NH, 1.0, 0, 0.0); /* :449 */
    int ip = 1, ipp = 0;                                        /* 0-based ip=2, ipp=1 */
    for (int64_t n = 2; n <= NC / 2; ++n) {                     /* :460 */
        cheb_single(&H, NH, NR, views[ipp], views[ip]);
        colwise_dot(mu, NR, 2 * n - 1, views[ip], views[ip], NH, 2.0, 1, -1.0);
        colwise_dot(mu, NR, 2 * n, views[ip], views[ipp], NH, 2.0, 2, -1.0);
        ip = 1 - ip;
        ipp = 1 - ipp;
    }
    free(a_all);
    return 0;
}

/* kpm_1d! bra/ket form, moment.jl:498-556: NC-1 SpMMs, one dot per step. */
int kpm_oracle_1d_lr(int64_t NH, const int64_t* colptr, const int64_t* rowval, const double* nzval,
                     int64_t NC, int64_t NR, const double* psi_l, const double* psi_r,
                     double* mu_all) {
    if (NH <= 0 || NR <= 0 || NC < 2) return OR_EINVAL; /* :512 */
    csc_t H = {NH, colptr, rowval, (const cplx*)nzval};
    cplx* mu = (cplx*)mu_all;
    const size_t blk = (size_t)(NH * NR);
    cplx* a_all = (cplx*)calloc(3 * blk, sizeof(cplx));
    if (!a_all) return OR_ENOMEM;
    cplx* views[2] = {a_all, a_all + blk};
    cplx* l = a_all + 2 * blk;
    memcpy(l, psi_l, sizeof(cplx) * blk);        /* :527 */
    memcpy(views[0], psi_r, sizeof(cplx) * blk); /* :528 */
    block_mul5(&H, NH, NR, views[0], views[1], 1.0, 0.0);
    colwise_dot(mu, NR, 1, l, views[0], NH, 1.0, 0, 0.0);
    colwise_dot(mu, NR, 2, l, views[1], NH, 1.0, 0, 0.0);
    int ip = 1, ipp = 0;
    for (int64_t n = 3; n <= NC; ++n) { /* :546 */
        cheb_single(&H, NH, NR, views[ipp], views[ip]);
        colwise_dot(mu, NR, n, l, views[ipp], NH, 1.0, 0, 0.0);
        ip = 1 - ip;
        ipp = 1 - ipp;
    }
    free(a_all);
    return 0;
}

/* kpm_1d_current!, moment.jl:656-714: Ja^dagger applied once (:680). */
int kpm_oracle_1d_current(int64_t NH, const int64_t* colptr, const int64_t* rowval,
                          const double* nzval, const int64_t* jcolptr, const int64_t* jrowval,
                          const double* jnzval, int64_t NC, int64_t NR, const double* psi,
                          double* mu_all) {
    if (NH <= 0 || NR <= 0 || NC < 2) return OR_EINVAL;
    csc_t H = {NH, colptr, rowval, (const cplx*)nzval};
    cplx* mu = (cplx*)mu_all;
    const size_t blk = (size_t)(NH * NR);
    cplx* a_all = (cplx*)calloc(3 * blk, sizeof(cplx));
    if (!a_all) return OR_ENOMEM;
    cplx* views[2] = {a_all, a_all + blk};
    cplx* jpsi = a_all + 2 * blk;
    memcpy(views[0], psi, sizeof(cplx) * blk);
    for (int64_t c = 0; c < NR; ++c)
        csc_adjoint_mul(NH, jcolptr, jrowval, (const cplx*)jnzval, views[0] + c * NH, jpsi + c * NH);
    colwise_dot(mu, NR, 1, jpsi, views[0], NH, 1.0, 0, 0.0);
    block_mul5(&H, NH, NR, views[0], views[1], 1.0, 0.0);
    colwise_dot(mu, NR, 2, jpsi, views[1], NH, 1.0, 0, 0.0);
    int ip = 1, ipp = 0;
    for (int64_t n = 3; n <= NC; ++n) {
        cheb_single(&H, NH, NR, views[ipp], views[ip]);
        colwise_dot(mu, NR, n, jpsi, views[ipp], NH, 1.0, 0, 0.0);
        ip = 1 - ip;
        ipp = 1 - ipp;
    }
    free(a_all);
    return 0;
}

static int64_t fld(int64_t a, int64_t b) { /* floor division, b > 0 */
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
    return q;
}

/* kpm_2d_blocking, moment.jl:759-810.  arr_size/right_block <= 0 mean :auto.
 * Returns OR_EINVAL where the reference throws ArgumentError. */
int kpm_oracle_2d_blocking(int64_t NH, int64_t NR, int64_t NC, int64_t budget, int64_t arr_size,
                           int64_t right_block, int64_t* out_K_left, int64_t* out_K_right,
                           int64_t* out_bytes) {
    if (NH <= 0 || NR <= 0 || NC < 2) return OR_EINVAL;
    if (budget < 0) return OR_EINVAL;
    const int64_t block_bytes = NH * NR * 16;
    const int64_t K_right = right_block <= 0 ? (NC < 16 ? NC : 16) : right_block;
    const int64_t base = (K_right + 7) * block_bytes;
    const int64_t per_left = block_bytes + K_right * 16;
    int64_t K_left;
    if (arr_size <= 0) {
        int64_t kmax = fld(budget - base, per_left);
        if (kmax < 2) return OR_EINVAL;
        K_left = kmax < NC ? kmax : NC;
    } else {
        if (arr_size < 2) return OR_EINVAL;
        K_left = arr_size;
    }
    *out_K_left = K_left;
    *out_K_right = K_right;
    *out_bytes = base + K_left * per_left;
    return 0;
}

/* The GEMM of _kpm_2d_blocked! (moment.jl:1055-1061, ZGEMM 'C','N'):
 * out[jn + ldo*jm] = scale * sum_k conj(R_jn[k]) L_jm[k] for jn < nR, jm < nL,
 * slots of K complex values back to back.  Cache-blocked: a task holds
 * kGroup left slots and walks K in chunks of kChunk, so each left chunk is
 * read once per task and each right chunk once per chunk of a task.  Every
 * entry sums its chunks in increasing k and each chunk in increasing k, so
 * the result does not depend on the thread count. */
enum { kChunk = 1024, kGroup = 16 };
static void contract_slots(const cplx* R, int64_t nR, const cplx* L, int64_t nL, size_t K,
                           cplx* out, int64_t ldo, double scale) {
    const int64_t ntask = (nL + kGroup - 1) / kGroup;
#pragma omp parallel for schedule(dynamic, 1) num_threads(kpm_oracle_get_threads())
    for (int64_t t = 0; t < ntask; ++t) {
        const int64_t m0 = t * kGroup;
        const int64_t mc = nL - m0 < kGroup ? nL - m0 : kGroup;
        cplx* acc = (cplx*)calloc((size_t)(nR * mc), sizeof(cplx));
        for (size_t c0 = 0; c0 < K; c0 += kChunk) {
            const size_t c1 = c0 + kChunk < K ? c0 + kChunk : K;
            for (int64_t jn = 0; jn < nR; ++jn) {
                const cplx* r = R + (size_t)jn * K;
                for (int64_t jm = 0; jm < mc; ++jm) {
                    const cplx* l = L + (size_t)(m0 + jm) * K;
                    double sr = 0.0, si = 0.0;
                    for (size_t k = c0; k < c1; ++k) {
                        const double ar = creal(r[k]), ai = cimag(r[k]);
                        const double br = creal(l[k]), bi = cimag(l[k]);
                        sr += ar * br + ai * bi;
                        si += ar * bi - ai * br;
                    }
                    acc[jn + nR * jm] += CMPLX(sr, si);
                }
            }
        }
        for (int64_t jm = 0; jm < mc; ++jm)
            for (int64_t jn = 0; jn < nR; ++jn)
                out[jn + ldo * (m0 + jm)] = acc[jn + nR * jm] * scale;
        free(acc);
    }
}

/* _kpm_2d_blocked!, moment.jl:987-1088.  parity: 0 NONE, 1 ODD, 2 EVEN.
 * mu is NC x NC column-major, overwritten (kpm_2d! does fill!(mu, 0), :916). */
int kpm_oracle_2d(int64_t NH, const int64_t* hcp, const int64_t* hrv, const double* hnz,
                  const int64_t* acp, const int64_t* arv, const double* anz, const int64_t* bcp,
                  const int64_t* brv, const double* bnz, int64_t NC, int64_t NR,
                  const double* psi_l, const double* psi_r, int64_t K_left, int64_t K_right,
                  int parity, int mn_sym, double* mu_out) {
    if (NH <= 0 || NR <= 0 || NC < 1 || K_left < 2 || K_right < 1) return OR_EINVAL;
    if (parity < 0 || parity > 2) return OR_EINVAL;
    csc_t H = {NH, hcp, hrv, (const cplx*)hnz};
    csc_t Ja = {NH, acp, arv, (const cplx*)anz};
    csc_t Jb = {NH, bcp, brv, (const cplx*)bnz};
    cplx* mu = (cplx*)mu_out;
    const size_t blk = (size_t)(NH * NR);
    const size_t nblocks = 6 + (size_t)K_left + (size_t)K_right;
    cplx* ws = (cplx*)calloc(nblocks * blk, sizeof(cplx));
    cplx* mub = (cplx*)calloc((size_t)(K_left * K_right), sizeof(cplx));
    if (!ws || !mub) {
        free(ws);
        free(mub);
        return OR_ENOMEM;
    }
    for (int64_t i = 0; i < NC * NC; ++i) mu[i] = 0.0;
    cplx* psi0l = ws;
    cplx* jpsi0r = ws + blk;
    cplx* psiw = ws + 2 * blk;
    cplx* rr[3] = {ws + 3 * blk, ws + 4 * blk, ws + 5 * blk};
    cplx* L = ws + 6 * blk;                     /* K_left slots */
    cplx* R = ws + (6 + (size_t)K_left) * blk;  /* K_right slots */
    memcpy(psi0l, psi_l, sizeof(cplx) * blk);
    /* left starter (:1013-1018) */
    memcpy(L, psi0l, sizeof(cplx) * blk);
    block_mul5(&H, NH, NR, psi0l, psiw, 1.0, 0.0);
    memcpy(L + blk, psiw, sizeof(cplx) * blk);
    /* right starter (:1021) */
    block_mul5(&Ja, NH, NR, (const cplx*)psi_r, jpsi0r, 1.0, 0.0);

    const int64_t reps = (NC + K_left - 1) / K_left;
    const double invNR = 1.0 / (double)NR;
    for (int64_t rep = 1; rep <= reps; ++rep) {
        const int64_t m1 = (rep - 1) * K_left + 1;
        const int64_t m2 = rep * K_left < NC ? rep * K_left : NC;
        const int64_t rep_size = m2 - m1 + 1;
        for (int64_t i = 3; i <= rep_size; ++i) /* chebyshev_iter(H, views, rep_size) */
            cheb_single3(&H, NH, NR, L + (size_t)(i - 3) * blk, L + (size_t)(i - 2) * blk,
                         L + (size_t)(i - 1) * blk);
        const int64_t nmax = mn_sym ? m2 : NC;
        for (int64_t n1 = 1; n1 <= nmax; n1 += K_right) {
            const int64_t n2 = n1 + K_right - 1 < nmax ? n1 + K_right - 1 : nmax;
            const int64_t n_count = n2 - n1 + 1;
            for (int64_t j = 1; j <= n_count; ++j) {
                const int64_t n = n1 + j - 1;
                /* r_i(n) = mod(n-1,3), r_ip = mod(n-2,3), r_ipp = mod(n-3,3) (vectors.jl:125-127) */
                const int ri = (int)((n - 1) % 3), rip = (int)((n + 1) % 3), ripp = (int)(n % 3);
                if (n == 1) {
                    memcpy(rr[0], jpsi0r, sizeof(cplx) * blk);
                } else if (n == 2) {
                    block_mul5(&H, NH, NR, jpsi0r, rr[1], 1.0, 0.0);
                } else {
                    cheb_single3(&H, NH, NR, rr[ripp], rr[rip], rr[ri]);
                }
                block_mul5(&Jb, NH, NR, rr[ri], R + (size_t)(j - 1) * blk, 1.0, 0.0);
            }
            /* mub = R^H L / NR  (ZGEMM 'C','N', :1055-1061) */
            contract_slots(R, n_count, L, rep_size, blk, mub, K_right, invNR);
            for (int64_t jm = 1; jm <= rep_size; ++jm) {
                const int64_t m = m1 + jm - 1;
                for (int64_t jn = 1; jn <= n_count; ++jn) {
                    const int64_t n = n1 + jn - 1;
                    const int keep = parity == 0 || (parity == 1 && ((m + n) % 2 == 1)) ||
                                     (parity == 2 && ((m + n) % 2 == 0));
                    if (keep) mu[(n - 1) + NC * (m - 1)] += conj(mub[(jn - 1) + K_right * (jm - 1)]);
                }
            }
        }
        if (rep < reps) { /* chebyshev_iter_wrap(H, views, K_left) */
            cheb_single3(&H, NH, NR, L + (size_t)(K_left - 2) * blk, L + (size_t)(K_left - 1) * blk, L);
            cheb_single3(&H, NH, NR, L + (size_t)(K_left - 1) * blk, L, L + blk);
        }
    }
    if (mn_sym) { /* :1078-1086 */
        for (int64_t m = 1; m <= NC; ++m)
            for (int64_t n = m + 1; n <= NC; ++n) {
                mu[(m - 1) + NC * (n - 1)] = creal(mu[(m - 1) + NC * (n - 1)]);
                mu[(n - 1) + NC * (m - 1)] = mu[(m - 1) + NC * (n - 1)];
            }
    }
    free(ws);
    free(mub);
    return 0;
}

/* kpm_3d!, moment.jl:1106-1227: for each n1, kpm_2d!(H, Jg, Jb, ..., psi_l,
 * T_{n1-1} Ja psi_r) into mu[:,:,n1]. */
int kpm_oracle_3d(int64_t NH, const int64_t* hcp, const int64_t* hrv, const double* hnz,
                  const int64_t* acp, const int64_t* arv, const double* anz, const int64_t* bcp,
                  const int64_t* brv, const double* bnz, const int64_t* gcp, const int64_t* grv,
                  const double* gnz, int64_t NC, int64_t NR, const double* psi_l,
                  const double* psi_r, int64_t K_left, int64_t K_right, double* mu_out) {
    if (NH <= 0 || NR <= 0 || NC < 1) return OR_EINVAL;
    csc_t H = {NH, hcp, hrv, (const cplx*)hnz};
    csc_t Ja = {NH, acp, arv, (const cplx*)anz};
    const size_t blk = (size_t)(NH * NR);
    cplx* rr = (cplx*)calloc(3 * blk, sizeof(cplx));
    if (!rr) return OR_ENOMEM;
    cplx* mu = (cplx*)mu_out;
    const size_t slab = (size_t)(NC * NC);
    int rc = 0;
    for (int64_t n1 = 1; n1 <= NC && rc == 0; ++n1) {
        const int ri = (int)((n1 - 1) % 3), rip = (int)((n1 + 1) % 3), ripp = (int)(n1 % 3);
        if (n1 == 1) {
            block_mul5(&Ja, NH, NR, (const cplx*)psi_r, rr, 1.0, 0.0);
        } else if (n1 == 2) {
            block_mul5(&H, NH, NR, rr + (size_t)rip * blk, rr + blk, 1.0, 0.0);
        } else {
            cheb_single3(&H, NH, NR, rr + (size_t)ripp * blk, rr + (size_t)rip * blk,
                         rr + (size_t)ri * blk);
        }
        rc = kpm_oracle_2d(NH, hcp, hrv, hnz, gcp, grv, gnz, bcp, brv, bnz, NC, NR, psi_l,
                           (const double*)(rr + (size_t)ri * blk), K_left, K_right, 0, 0,
                           (double*)(mu + slab * (size_t)(n1 - 1)));
    }
    free(rr);
    return rc;
}

/* chebyshev_iter(H, views, n) on an NH x NR x n workspace (chebyshev_iteration.jl:16-20):
 * slots 3..n filled from slots 1, 2 with the 3-address step. */
int kpm_oracle_cheb_iter(int64_t NH, const int64_t* colptr, const int64_t* rowval,
                         const double* nzval, int64_t NR, int64_t nslots, double* slots) {
    if (NH <= 0 || NR <= 0 || nslots < 2) return OR_EINVAL;
    csc_t H = {NH, colptr, rowval, (const cplx*)nzval};
    const size_t blk = (size_t)(NH * NR);
    cplx* s = (cplx*)slots;
    for (int64_t i = 3; i <= nslots; ++i)
        cheb_single3(&H, NH, NR, s + (size_t)(i - 3) * blk, s + (size_t)(i - 2) * blk,
                     s + (size_t)(i - 1) * blk);
    return 0;
}

/* A bounded run of the doubling loop for the CPU baseline: `steps` in-place
 * recurrence steps plus the two fused-by-reference dots per step, exactly the
 * per-step work of kpm_1d! (moment.jl:460-487).  Writes 2*steps moments per
 * column into mu_all (NR x 2*steps). */
int kpm_oracle_1d_steps(int64_t NH, const int64_t* colptr, const int64_t* rowval,
                        const double* nzval, int64_t NR, int64_t steps, double* v0, double* v1,
                        double* mu_all) {
    if (NH <= 0 || NR <= 0 || steps < 1) return OR_EINVAL;
    csc_t H = {NH, colptr, rowval, (const cplx*)nzval};
    cplx* views[2] = {(cplx*)v0, (cplx*)v1};
    cplx* mu = (cplx*)mu_all;
    int ip = 1, ipp = 0;
    for (int64_t n = 1; n <= steps; ++n) {
        cheb_single(&H, NH, NR, views[ipp], views[ip]);
        colwise_dot(mu, NR, 2 * n - 1, views[ip], views[ip], NH, 2.0, 0, 0.0);
        colwise_dot(mu, NR, 2 * n, views[ip], views[ipp], NH, 2.0, 0, 0.0);
        ip = 1 - ip;
        ipp = 1 - ipp;
    }
    return 0;
}

/* Forward-mode tangent of the direct recurrence (the derivative of kpm_1d's
 * moments along a perturbation dH with H's sparsity pattern): alpha_0 = psi,
 * alpha_1 = H psi, alpha_{k+1} = 2 H alpha_k - alpha_{k-1};
 * dalpha_0 = 0, dalpha_1 = dH psi, dalpha_{k+1} = 2 dH alpha_k + 2 H dalpha_k - dalpha_{k-1};
 * mu[:,n] = <psi|alpha_n>, dmu[:,n] = <psi|dalpha_n> (NR x NC each).
 * dnzval holds dH's values in H's CSC order. */
int kpm_oracle_1d_tangent(int64_t NH, const int64_t* colptr, const int64_t* rowval,
                          const double* nzval, const double* dnzval, int64_t NC, int64_t NR,
                          const double* psi, double* mu_all, double* dmu_all) {
    if (NH <= 0 || NR <= 0 || NC < 2) return OR_EINVAL;
    const cplx* nz = (const cplx*)nzval;
    const cplx* dnz = (const cplx*)dnzval;
    const cplx* p = (const cplx*)psi;
    cplx* mu = (cplx*)mu_all;
    cplx* dmu = (cplx*)dmu_all;
    cplx* w = (cplx*)calloc((size_t)(5 * NH), sizeof(cplx));
    if (!w) return OR_ENOMEM;
    for (int64_t r = 0; r < NR; ++r) {
        const cplx* v = p + r * NH;
        cplx *a0 = w, *a1 = w + NH, *d0 = w + 2 * NH, *d1 = w + 3 * NH, *tmp = w + 4 * NH;
        memcpy(a0, v, sizeof(cplx) * NH);
        for (int64_t i = 0; i < NH; ++i) d0[i] = 0.0;
        csc_mul5(NH, colptr, rowval, nz, a0, a1, 1.0, 0.0);
        csc_mul5(NH, colptr, rowval, dnz, a0, d1, 1.0, 0.0);
        mu[r] = vdot(NH, v, a0);
        dmu[r] = 0.0;
        mu[r + NR] = vdot(NH, v, a1);
        dmu[r + NR] = vdot(NH, v, d1);
        for (int64_t n = 2; n < NC; ++n) {
            /* d_{n} = 2 dH a_{n-1} + 2 H d_{n-1} - d_{n-2}, into d0 (holds d_{n-2}) */
            csc_mul5(NH, colptr, rowval, dnz, a1, tmp, 2.0, 0.0);
            csc_mul5(NH, colptr, rowval, nz, d1, d0, 2.0, -1.0);
            for (int64_t i = 0; i < NH; ++i) d0[i] += tmp[i];
            csc_mul5(NH, colptr, rowval, nz, a1, a0, 2.0, -1.0);
            mu[r + NR * n] = vdot(NH, v, a0);
            dmu[r + NR * n] = vdot(NH, v, d0);
            cplx* t = a0; a0 = a1; a1 = t;
            t = d0; d0 = d1; d1 = t;
        }
    }
    free(w);
    return 0;
}

/* chebyshev_action! (chebyshev_action.jl:47-125): two-slot recurrence,
 * out[:, :, k] += C[n, k] * T_{n-1} V after every step (broadcast_assign!,
 * chebyshev_action.jl:163-178), out zeroed first.  V NH x NR, C NC x K and
 * out NH x NR x K, all column-major complex.  The relative guard
 * (chebyshev_iteration.jl:119-132) runs at iteration 1 when NC == 2 and at
 * every check_every-th iteration and the last one otherwise; on a trip the
 * function returns -4 and stores the iteration and sqrt(max growth) in
 * guard[0], guard[1].  check_every == 0 disables it. */
int kpm_oracle_cheb_action(int64_t NH, const int64_t* colptr, const int64_t* rowval,
                           const double* nzval, int64_t NC, int64_t K, int64_t NR,
                           const double* V, const double* Cc, double* out_, int64_t check_every,
                           double* guard) {
    if (NH <= 0 || NR <= 0 || NC < 1 || K < 1 || check_every < 0) return OR_EINVAL;
    csc_t H = {NH, colptr, rowval, (const cplx*)nzval};
    const cplx* C = (const cplx*)Cc;
    cplx* out = (cplx*)out_;
    const size_t blk = (size_t)(NH * NR);
    cplx* s1 = (cplx*)malloc(sizeof(cplx) * blk);
    cplx* s2 = (cplx*)malloc(sizeof(cplx) * blk);
    double* ref = (double*)malloc(sizeof(double) * (size_t)NR);
    if (!s1 || !s2 || !ref) {
        free(s1);
        free(s2);
        free(ref);
        return OR_ENOMEM;
    }
    memcpy(s1, V, sizeof(cplx) * blk);
    for (int64_t c = 0; c < NR; ++c) ref[c] = creal(vdot(NH, s1 + c * NH, s1 + c * NH));
    memset(out, 0, sizeof(cplx) * blk * (size_t)K);
    cplx* slots[2] = {s1, s2};
    int rc = 0;
    /* broadcast_assign!: out_k .+= x .* c_k */
#define ACCUM(x, n)                                                        \
    for (int64_t k = 0; k < K; ++k) {                                     \
        const cplx ck = C[(n) + NC * k];                                   \
        cplx* o = out + (size_t)k * blk;                                   \
        for (size_t e = 0; e < blk; ++e) o[e] += (x)[e] * ck;              \
    }
    /* _check_chebyshev_columns(slot, iteration, ref_sq) */
#define GUARD(x, it)                                                       \
    do {                                                                   \
        double g2 = -INFINITY;                                             \
        int nan = 0;                                                       \
        for (int64_t c = 0; c < NR; ++c) {                                 \
            const double d = ref[c] == 0.0 ? 1.0 : ref[c];                 \
            const double r = creal(vdot(NH, (x) + c * NH, (x) + c * NH)) / d; \
            if (r != r) nan = 1;                                           \
            if (r > g2) g2 = r;                                            \
        }                                                                  \
        if (nan || !(g2 <= 2.25)) {                                        \
            if (guard) {                                                   \
                guard[0] = (double)(it);                                   \
                guard[1] = nan ? NAN : sqrt(g2);                           \
            }                                                              \
            rc = -4;                                                       \
        }                                                                  \
    } while (0)
    ACCUM(slots[0], 0);
    if (NC >= 2) {
        block_mul5(&H, NH, NR, slots[0], slots[1], 1.0, 0.0);
        ACCUM(slots[1], 1);
        if (NC == 2 && check_every > 0) GUARD(slots[1], 1);
        int ip = 1, ipp = 0;
        for (int64_t n = 3; n <= NC && rc == 0; ++n) {
            cheb_single(&H, NH, NR, slots[ipp], slots[ip]);
            ACCUM(slots[ipp], n - 1);
            const int64_t it = n - 1;
            if (check_every > 0 && (it % check_every == 0 || n == NC)) GUARD(slots[ipp], it);
            ip = 1 - ip;
            ipp = 1 - ipp;
        }
    }
#undef ACCUM
#undef GUARD
    free(s1);
    free(s2);
    free(ref);
    return rc;
}
