// engine2d.hip -- kpm_2d (src/moment.jl:887-1088) and kpm_3d
// (src/moment.jl:1106-1227) on the device.
//
// Blocking follows _kpm_2d_blocked! exactly: a K_left-slot left array
// L = T_m psi_l filled by chebyshev_iter and continued by chebyshev_iter_wrap
// (src/utils/chebyshev_iteration.jl:10-36), a 3-slot right ring
// T_n Ja psi_r restarted every rep (:1039-1050), K_right slots
// R = Jb T_n Ja psi_r (:1051) and one contraction per (rep, right block)
// accumulated on the device (no per-block D2H sync).  The 3-address steps
// (V_out <- 2 H V_p - V_pp) are single fused launches: V_pp is read and
// V_out written by the same pass, no copy.
//
// Slots are stored back to back with a stride of NH*NRp complex values
// rounded up to 16 (zero tail), so the contraction sees dense vectors.
//
// Right chain and Jb in one pass: the order-n step of the right ring already
// gathers T_{n-1} Ja psi_r, which is what Jb needs for order n-1.  H and Jb
// merged onto the union of their patterns (each entry carrying both values,
// zero where one operator has no entry) give both products from one set of
// gathers (MODE_DUAL), so each order's R is emitted one step late and every
// right block ends with one standalone Jb pass.  Each product sums its own
// nonzeros in its own row order, so the tables are those of separate passes.
#include <algorithm>
#include <climits>
#include <vector>

#include "bf16_planes.h"
#include "kpm_internal.h"

namespace kpm {

StepConfig tuned_config(kpm_ctx* ctx, const kpm_csr* A, int64_t NR, int prec, int mode);

static int spmm(kpm_ctx* ctx, const kpm_csr* A, const kpm_block* x, const kpm_block* z,
                kpm_block* out, double alpha, double beta) {
  StepLaunch s{A, x, z, out, nullptr, alpha, beta, MODE_NONE};
  return launch_step(ctx, s, tuned_config(ctx, A, x->NR, x->prec, MODE_NONE), nullptr);
}

// ---- merged operator of the dual pass -----------------------------------------

// The union's row pointers come from a three-kernel scan: k_union_count
// writes each row's entry count and each 256-row block's total,
// k_scan_blocks turns the block totals into block offsets, and k_union_fill
// scans its block's counts in LDS, adds the block offset and writes the
// entries.

constexpr int kMergeBlock = 256;

// inclusive scan of one value per thread over a 256-thread workgroup
__device__ inline int block_scan(int v, int* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = 1; off < kMergeBlock; off <<= 1) {
    const int u = tid >= off ? sh[tid - off] : 0;
    __syncthreads();
    sh[tid] += u;
    __syncthreads();
  }
  return sh[tid];
}

// entries of row i of the union of two operators with strictly increasing
// columns per row
__device__ inline int union_count(const int* __restrict__ rpa, const int* __restrict__ cia,
                                  const int* __restrict__ rpb, const int* __restrict__ cib, int i) {
  int a = rpa[i], b = rpb[i], c = 0;
  const int ae = rpa[i + 1], be = rpb[i + 1];
  while (a < ae || b < be) {
    const int ca = a < ae ? cia[a] : INT_MAX;
    const int cb = b < be ? cib[b] : INT_MAX;
    a += ca <= cb ? 1 : 0;
    b += cb <= ca ? 1 : 0;
    ++c;
  }
  return c;
}

// cnt[i] <- entries of union row i, bsum[block] <- the block's total,
// maxrow <- the largest row
__global__ void __launch_bounds__(kMergeBlock)
    k_union_count(const int* __restrict__ rpa, const int* __restrict__ cia,
                  const int* __restrict__ rpb, const int* __restrict__ cib, int n,
                  int* __restrict__ cnt, int* __restrict__ bsum, int* __restrict__ maxrow) {
  __shared__ int sh[kMergeBlock];
  const int i = blockIdx.x * kMergeBlock + threadIdx.x;
  const int c = i < n ? union_count(rpa, cia, rpb, cib, i) : 0;
  if (i < n) {
    cnt[i] = c;
    atomicMax(maxrow, c);
  }
  const int tot = block_scan(c, sh);
  if (threadIdx.x == kMergeBlock - 1) bsum[blockIdx.x] = tot;
}

// bsum[0..nb) totals -> exclusive offsets, in place; one workgroup, each
// thread a contiguous chunk
__global__ void __launch_bounds__(kMergeBlock) k_scan_blocks(int* __restrict__ bsum, int nb) {
  __shared__ int sh[kMergeBlock];
  const int chunk = (nb + kMergeBlock - 1) / kMergeBlock;
  const int lo = min((int)threadIdx.x * chunk, nb), hi = min(lo + chunk, nb);
  int s = 0;
  for (int k = lo; k < hi; ++k) s += bsum[k];
  int run = block_scan(s, sh) - s;
  for (int k = lo; k < hi; ++k) {
    const int v = bsum[k];
    bsum[k] = run;
    run += v;
  }
}

__device__ inline double2 csr_value(const void* v, int real, int k) {
  return real ? make_double2(reinterpret_cast<const double*>(v)[k], 0.0)
              : reinterpret_cast<const double2*>(v)[k];
}

// union row pointers (rpu[0] = 0, rpu[i + 1] = end of row i) and entries in
// column order: columns, A's values, B's values
__global__ void __launch_bounds__(kMergeBlock)
    k_union_fill(const int* __restrict__ rpa, const int* __restrict__ cia,
                 const void* __restrict__ va, int a_real, const int* __restrict__ rpb,
                 const int* __restrict__ cib, const void* __restrict__ vb, int b_real, int n,
                 const int* __restrict__ cnt, const int* __restrict__ boff, int* __restrict__ rpu,
                 int* __restrict__ ciu, double2* __restrict__ vua, double2* __restrict__ vub) {
  __shared__ int sh[kMergeBlock];
  const int i = blockIdx.x * kMergeBlock + threadIdx.x;
  const int c = i < n ? cnt[i] : 0;
  const int end = boff[blockIdx.x] + block_scan(c, sh);
  if (i == 0) rpu[0] = 0;
  if (i >= n) return;
  rpu[i + 1] = end;
  int a = rpa[i], b = rpb[i], o = end - c;
  const int ae = rpa[i + 1], be = rpb[i + 1];
  while (a < ae || b < be) {
    const int ca = a < ae ? cia[a] : INT_MAX;
    const int cb = b < be ? cib[b] : INT_MAX;
    const bool ta = ca <= cb, tb = cb <= ca;
    ciu[o] = ta ? ca : cb;
    vua[o] = ta ? csr_value(va, a_real, a) : make_double2(0.0, 0.0);
    vub[o] = tb ? csr_value(vb, b_real, b) : make_double2(0.0, 0.0);
    a += ta ? 1 : 0;
    b += tb ? 1 : 0;
    ++o;
  }
}

// A and B merged (M: A's values on the union pattern, *vals2: B's), built
// on the device and kept in the context for the same pair of operators.
// *ok stays false when the dual pass does not apply: unsorted rows, rows of
// one lane at this block width (no spread metadata), no device memory for
// the merged operator, or the dual pass switched off.
static int dual_operator(kpm_ctx* ctx, const kpm_csr* A, const kpm_csr* B, int64_t NR, int prec,
                         kpm_csr* M, const void** vals2, bool* ok) {
  *ok = false;
  if (!ctx->dual_pass || !A->sorted || !B->sorted || A->n != B->n || A->n < 1) return KPM_OK;
  const int64_t n = A->n, cap = A->nnz + B->nnz + kCsrPad;  // union <= sum, plus the tail
  if (cap >= INT32_MAX) return KPM_OK;
  const size_t brp = ((size_t)(n + 1) * 4 + 255) / 256 * 256;
  const size_t bci = ((size_t)cap * 4 + 255) / 256 * 256;
  const size_t bv = (size_t)cap * sizeof(double2);
  const int64_t nb = (n + kMergeBlock - 1) / kMergeBlock;
  // scratch after the operator: row counts (n), block offsets (nb), max row
  const size_t bcnt = ((size_t)n * 4 + 255) / 256 * 256;
  const size_t bblk = ((size_t)nb * 4 + 255) / 256 * 256;
  char* base = nullptr;
  if (ctx->merged_key[0] != A->serial || ctx->merged_key[1] != B->serial) {
    ctx->merged_key[0] = ctx->merged_key[1] = 0;
    // no room for the merged operator (over the context's limit, or the
    // allocation fails; ensure clears HIP's latched status): the two passes
    // run separately
    const size_t need = brp + bci + 2 * bv + bcnt + bblk + 256;
    if (ctx->merged_limit >= 0 && need > (size_t)ctx->merged_limit) return KPM_OK;
    if (ensure(ctx, ctx->merged, need) != KPM_OK) {
      set_error("");
      return KPM_OK;
    }
    base = (char*)ctx->merged.ptr;
    int* rpu = (int*)base;
    int* cnt = (int*)(base + brp + bci + 2 * bv);
    int* bsum = (int*)(base + brp + bci + 2 * bv + bcnt);
    int* dmax = (int*)(base + brp + bci + 2 * bv + bcnt + bblk);
    KPM_HIP_TRY(hipMemsetAsync(dmax, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(k_union_count, dim3((unsigned)nb), dim3(kMergeBlock), 0, ctx->stream,
                       A->rowptr, A->colind, B->rowptr, B->colind, (int)n, cnt, bsum, dmax);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kMergeBlock), 0, ctx->stream, bsum, (int)nb);
    hipLaunchKernelGGL(k_union_fill, dim3((unsigned)nb), dim3(kMergeBlock), 0, ctx->stream,
                       A->rowptr, A->colind, A->vals, A->is_real, B->rowptr, B->colind, B->vals,
                       B->is_real, (int)n, cnt, bsum, rpu, (int*)(base + brp),
                       (double2*)(base + brp + bci), (double2*)(base + brp + bci + bv));
    KPM_HIP_TRY(hipGetLastError());
    int hm[2] = {0, 0};
    KPM_HIP_TRY(hipMemcpyAsync(&hm[0], dmax, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    KPM_HIP_TRY(hipMemcpyAsync(&hm[1], rpu + n, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->merged_max_row = hm[0];
    ctx->merged_nnz = hm[1];
    ctx->merged_key[0] = A->serial;
    ctx->merged_key[1] = B->serial;
  }
  base = (char*)ctx->merged.ptr;
  M->ctx = ctx;
  M->n = n;
  M->nnz = ctx->merged_nnz;
  M->is_real = 0;
  M->max_row_nnz = ctx->merged_max_row;
  M->sorted = 1;
  M->serial = 0;
  M->rowptr = (int*)base;
  M->colind = (int*)(base + brp);
  M->vals = base + brp + bci;
  *vals2 = base + brp + bci + bv;
  if (step_geom(M, NR, prec, StepConfig{}).lpr < 2) return KPM_OK;
  *ok = true;
  return KPM_OK;
}

// out = alpha*H*x + beta*z and out2 = Jb*x in one pass over the merged
// operator; with out2_cols > 0, out2 is split into blocks of that many
// complex columns, out2_stride complex values apart
static int spmm_dual(kpm_ctx* ctx, const kpm_csr* H, const kpm_csr* M, const void* vals2,
                     const kpm_block* x, const kpm_block* z, kpm_block* out, kpm_block* out2,
                     double alpha, double beta, int64_t out2_cols = 0, int64_t out2_stride = 0) {
  StepLaunch s{M, x, z, out, nullptr, alpha, beta, MODE_DUAL};
  s.gamma = 1.0;
  s.vals2 = vals2;
  s.out2 = out2;
  s.out2_cols = out2_cols;
  s.out2_stride = out2_stride;
  return launch_step(ctx, s, tuned_config(ctx, H, x->NR, x->prec, MODE_NONE), nullptr);
}

static void slot_view(void* base, int64_t stride_bytes, int64_t idx, int64_t NH, int64_t NR,
                      int prec, kpm_ctx* ctx, kpm_block* out) {
  out->ctx = ctx;
  out->NH = NH;
  out->NR = NR;
  out->NRp = padded_nr(NR, prec);
  out->prec = prec;
  out->d = (char*)base + stride_bytes * idx;
  out->owned = false;
}

struct Plan2d {
  int64_t NC, NR, NH, K_left, K_right;
  int parity, mn_sym, prec, gemm_mode;
};

// ---- split-BF16 mode on slots stored as planes ---------------------------------
//
// In the split-BF16 contraction mode the contraction reads each slot only
// through its hi/lo BF16 split.  So the passes that produce the slots write
// that split (bf16_planes.h, 8 bytes per complex value) instead of the FP64
// or FP32 words the contraction would split on every read: the right slots
// Jb T_n Ja psi_r are written only as planes (the dual pass's second product,
// and the Jb pass), and the left chain runs on a ring of three blocks whose
// step also writes the planes of its output into the left slot.  The
// contraction then moves half the operand bytes per stage and does no split
// (k_contract_planes_t128); its table is bitwise the one from the split of
// the stored slots.  Taken when every contraction of the call has 128 or
// more slots a side (the 128 x 128 tile kernel).

// planes of a block's first n complex values
template <typename E>
__global__ void __launch_bounds__(256)
    k_block_planes(const E* __restrict__ src, int64_t n, uint32_t* __restrict__ planes) {
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256)
    store_plane_value(planes, k, (double)src[k].x, (double)src[k].y);
}

static int block_planes(kpm_ctx* ctx, const kpm_block* b, void* planes) {
  const int64_t n = b->NH * b->NRp;
  const unsigned grid = (unsigned)std::max<int64_t>(
      1, std::min<int64_t>((n + 255) / 256, (int64_t)std::max(ctx->num_cus, 8) * 16));
  if (b->prec == KPM_VEC_F32)
    hipLaunchKernelGGL(k_block_planes<float2>, dim3(grid), dim3(256), 0, ctx->stream,
                       (const float2*)b->d, n, (uint32_t*)planes);
  else
    hipLaunchKernelGGL(k_block_planes<double2>, dim3(grid), dim3(256), 0, ctx->stream,
                       (const double2*)b->d, n, (uint32_t*)planes);
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

// out = alpha*A*x + beta*z, and the planes of out at `planes` (only the
// planes when planes_only)
static int spmm_planes(kpm_ctx* ctx, const kpm_csr* A, const kpm_block* x, const kpm_block* z,
                       kpm_block* out, double alpha, double beta, void* planes, bool planes_only) {
  StepLaunch s{A, x, z, out, nullptr, alpha, beta, MODE_NONE};
  s.planes = planes;
  s.planes_only = planes_only;
  return launch_step(ctx, s, tuned_config(ctx, A, x->NR, x->prec, MODE_NONE), nullptr);
}

// out = alpha*H*x + beta*z, and the planes of Jb*x at `planes` (dual pass)
static int spmm_dual_planes(kpm_ctx* ctx, const kpm_csr* H, const kpm_csr* M, const void* vals2,
                            const kpm_block* x, const kpm_block* z, kpm_block* out, void* planes,
                            double alpha, double beta) {
  StepLaunch s{M, x, z, out, nullptr, alpha, beta, MODE_DUAL};
  s.gamma = 1.0;
  s.vals2 = vals2;
  s.planes = planes;
  return launch_step(ctx, s, tuned_config(ctx, H, x->NR, x->prec, MODE_NONE), nullptr);
}

// every contraction run_2d would launch has 128 or more slots a side
static bool plane_tables(const Plan2d& p) {
  if (p.gemm_mode != KPM_GEMM_BF16X3 || p.K_left < 128 || p.K_right < 128) return false;
  const int64_t reps = (p.NC + p.K_left - 1) / p.K_left;
  for (int64_t rep = 1; rep <= reps; ++rep) {
    const int64_t m1 = (rep - 1) * p.K_left + 1, m2 = std::min(rep * p.K_left, p.NC);
    if (m2 - m1 + 1 < 128) return false;
    const int64_t nmax = p.mn_sym ? m2 : p.NC;
    for (int64_t n1 = 1; n1 <= nmax; n1 += p.K_right)
      if (std::min(n1 + p.K_right - 1, nmax) - n1 + 1 < 128) return false;
  }
  return true;
}

static int run_2d_planes(kpm_ctx* ctx, const kpm_csr* H, const kpm_csr* Ja, const kpm_csr* Jb,
                         const Plan2d& p, const kpm_block* psi_l, const kpm_block* psi_r, void* mu_dev) {
  const int64_t NH = p.NH, NR = p.NR, NC = p.NC;
  const int prec = p.prec;
  const int64_t NRp = padded_nr(NR, prec);
  const int64_t used = NH * NRp;
  const int64_t Kp = (used + 15) / 16 * 16;
  const int64_t eb = elem_bytes(prec);
  const int64_t pitch = Kp + 256 / eb;   // complex values between slots, as in run_2d
  const int64_t stride = pitch * 8;      // plane bytes per slot
  const int64_t KL = p.K_left, KR = p.K_right;
  KPM_TRY(ensure(ctx, ctx->slots_l, (size_t)(stride * KL)));
  KPM_TRY(ensure(ctx, ctx->slots_r, (size_t)(stride * KR)));
  if (Kp > used) {  // the zero tail of each slot (the partial block's written words come later)
    const int64_t t0 = used / 16 * 128;
    KPM_HIP_TRY(hipMemset2DAsync((char*)ctx->slots_l.ptr + t0, (size_t)stride, 0, (size_t)(Kp * 8 - t0),
                                 (size_t)KL, ctx->stream));
    KPM_HIP_TRY(hipMemset2DAsync((char*)ctx->slots_r.ptr + t0, (size_t)stride, 0, (size_t)(Kp * 8 - t0),
                                 (size_t)KR, ctx->stream));
  }
  auto lp = [&](int64_t i) { return (void*)((char*)ctx->slots_l.ptr + stride * i); };
  auto rp = [&](int64_t i) { return (void*)((char*)ctx->slots_r.ptr + stride * i); };
  kpm_block rr[3], jpsi0r, lr[3], sink;
  for (int i = 0; i < 3; ++i) KPM_TRY(work_block(ctx, ctx->ring[i], NH, NR, prec, &rr[i]));
  KPM_TRY(work_block(ctx, ctx->ring[3], NH, NR, prec, &jpsi0r));
  KPM_TRY(work_block(ctx, ctx->ring[4], NH, NR, prec, &sink));  // out of planes-only passes (not written)
  for (int i = 0; i < 3; ++i) KPM_TRY(work_block(ctx, ctx->ring[5 + i], NH, NR, prec, &lr[i]));

  kpm_csr M;
  const void* v2 = nullptr;
  bool dual = false;
  KPM_TRY(dual_operator(ctx, H, Jb, NR, prec, &M, &v2, &dual));
  KPM_HIP_TRY(hipMemsetAsync(mu_dev, 0, (size_t)(NC * NC) * sizeof(double2), ctx->stream));
  KPM_TRY(spmm(ctx, Ja, psi_r, nullptr, &jpsi0r, 1.0, 0.0));  // right starter (:1021)

  const int64_t reps = (NC + KL - 1) / KL;
  const double scale = 1.0 / (double)NR;
  for (int64_t rep = 1; rep <= reps; ++rep) {
    const int64_t m1 = (rep - 1) * KL + 1;
    const int64_t m2 = std::min(rep * KL, NC);
    const int64_t rep_size = m2 - m1 + 1;
    // left orders m1-1 .. m2-1 (0-based): the starters (:1013-1018), then
    // chebyshev_iter and, from the second rep, chebyshev_iter_wrap, which
    // continue the same recurrence on the ring
    for (int64_t o = m1 - 1; o < m2; ++o) {
      void* pl = lp(o - (m1 - 1));
      if (o == 0) {
        KPM_TRY(kpm_block_copy(&lr[0], psi_l));
        KPM_TRY(block_planes(ctx, psi_l, pl));
      } else if (o == 1) {
        KPM_TRY(spmm_planes(ctx, H, psi_l, nullptr, &lr[1], 1.0, 0.0, pl, false));
      } else {
        KPM_TRY(spmm_planes(ctx, H, &lr[(o - 1) % 3], &lr[(o - 2) % 3], &lr[o % 3], 2.0, -1.0, pl, false));
      }
    }
    const int64_t nmax = p.mn_sym ? m2 : NC;
    for (int64_t n1 = 1; n1 <= nmax; n1 += KR) {
      const int64_t n2 = std::min(n1 + KR - 1, nmax);
      const int64_t n_count = n2 - n1 + 1;
      for (int64_t j = 1; j <= n_count; ++j) {
        const int64_t n = n1 + j - 1;
        const int ri = (int)((n - 1) % 3), rip = (int)((n + 1) % 3), ripp = (int)(n % 3);
        const bool fuse = dual && j >= 2;  // this step also emits the planes of R order n - 1
        if (n == 1) {
          KPM_TRY(kpm_block_copy(&rr[0], &jpsi0r));
        } else if (n == 2) {
          if (fuse) KPM_TRY(spmm_dual_planes(ctx, H, &M, v2, &jpsi0r, nullptr, &rr[1], rp(j - 2), 1.0, 0.0));
          else KPM_TRY(spmm(ctx, H, &jpsi0r, nullptr, &rr[1], 1.0, 0.0));
        } else {
          if (fuse)
            KPM_TRY(spmm_dual_planes(ctx, H, &M, v2, &rr[rip], &rr[ripp], &rr[ri], rp(j - 2), 2.0, -1.0));
          else KPM_TRY(spmm(ctx, H, &rr[rip], &rr[ripp], &rr[ri], 2.0, -1.0));
        }
        if (!dual) KPM_TRY(spmm_planes(ctx, Jb, &rr[ri], nullptr, &sink, 1.0, 0.0, rp(j - 1), true));
      }
      if (dual)  // the block's last order
        KPM_TRY(spmm_planes(ctx, Jb, &rr[(int)((n2 - 1) % 3)], nullptr, &sink, 1.0, 0.0, rp(n_count - 1), true));
      KPM_TRY(launch_contract_planes(ctx, ctx->slots_l.ptr, (int)rep_size, ctx->slots_r.ptr, (int)n_count,
                                     Kp, pitch, scale, mu_dev, NC, n1 - 1, m1 - 1, p.parity));
    }
  }
  return KPM_OK;
}

// mu_dev: NC x NC column-major double2, overwritten
static int run_2d(kpm_ctx* ctx, const kpm_csr* H, const kpm_csr* Ja, const kpm_csr* Jb,
                  const Plan2d& p, const kpm_block* psi_l, const kpm_block* psi_r, void* mu_dev) {
  if (ctx->bf16_planes && plane_tables(p)) return run_2d_planes(ctx, H, Ja, Jb, p, psi_l, psi_r, mu_dev);
  const int64_t NH = p.NH, NR = p.NR, NC = p.NC;
  const int prec = p.prec;
  const int64_t NRp = padded_nr(NR, prec);
  const int64_t used = NH * NRp;                 // complex values per block
  const int64_t Kp = (used + 15) / 16 * 16;      // contracted length in complex values
  const int64_t eb = elem_bytes(prec);
  // Slot pitch: Kp plus 256 bytes, so the slots of one contraction tile do not
  // sit a power of two apart (which maps them all to one L2 set and channel).
  const int64_t pitch = Kp + 256 / eb;
  const int64_t stride = pitch * eb;
  const int64_t KL = p.K_left, KR = p.K_right;
  KPM_TRY(ensure(ctx, ctx->slots_l, (size_t)(stride * KL)));
  KPM_TRY(ensure(ctx, ctx->slots_r, (size_t)(stride * KR)));
  if (Kp > used) {
    KPM_HIP_TRY(hipMemset2DAsync((char*)ctx->slots_l.ptr + used * eb, (size_t)stride, 0,
                                 (size_t)((Kp - used) * eb), (size_t)KL, ctx->stream));
    KPM_HIP_TRY(hipMemset2DAsync((char*)ctx->slots_r.ptr + used * eb, (size_t)stride, 0,
                                 (size_t)((Kp - used) * eb), (size_t)KR, ctx->stream));
  }
  kpm_block rr[3], jpsi0r;
  for (int i = 0; i < 3; ++i) KPM_TRY(work_block(ctx, ctx->ring[i], NH, NR, prec, &rr[i]));
  KPM_TRY(work_block(ctx, ctx->ring[3], NH, NR, prec, &jpsi0r));
  std::vector<kpm_block> L(KL), R(KR);
  for (int64_t i = 0; i < KL; ++i) slot_view(ctx->slots_l.ptr, stride, i, NH, NR, prec, ctx, &L[i]);
  for (int64_t i = 0; i < KR; ++i) slot_view(ctx->slots_r.ptr, stride, i, NH, NR, prec, ctx, &R[i]);

  kpm_csr M;
  const void* v2 = nullptr;
  bool dual = false;
  KPM_TRY(dual_operator(ctx, H, Jb, NR, prec, &M, &v2, &dual));
  KPM_HIP_TRY(hipMemsetAsync(mu_dev, 0, (size_t)(NC * NC) * sizeof(double2), ctx->stream));
  // left starter (:1013-1018) and right starter (:1021)
  KPM_TRY(kpm_block_copy(&L[0], psi_l));
  KPM_TRY(spmm(ctx, H, psi_l, nullptr, &L[1], 1.0, 0.0));
  KPM_TRY(spmm(ctx, Ja, psi_r, nullptr, &jpsi0r, 1.0, 0.0));

  const int64_t reps = (NC + KL - 1) / KL;
  const double scale = 1.0 / (double)NR;
  for (int64_t rep = 1; rep <= reps; ++rep) {
    const int64_t m1 = (rep - 1) * KL + 1;
    const int64_t m2 = std::min(rep * KL, NC);
    const int64_t rep_size = m2 - m1 + 1;
    for (int64_t i = 3; i <= rep_size; ++i)  // chebyshev_iter(H, L, rep_size)
      KPM_TRY(spmm(ctx, H, &L[i - 2], &L[i - 3], &L[i - 1], 2.0, -1.0));
    const int64_t nmax = p.mn_sym ? m2 : NC;
    for (int64_t n1 = 1; n1 <= nmax; n1 += KR) {
      const int64_t n2 = std::min(n1 + KR - 1, nmax);
      const int64_t n_count = n2 - n1 + 1;
      for (int64_t j = 1; j <= n_count; ++j) {
        const int64_t n = n1 + j - 1;
        const int ri = (int)((n - 1) % 3), rip = (int)((n + 1) % 3), ripp = (int)(n % 3);
        // dual: this step also emits R of order n - 1 (rr[rip]) when that
        // order belongs to this block
        const bool fuse = dual && j >= 2;
        if (n == 1) {
          KPM_TRY(kpm_block_copy(&rr[0], &jpsi0r));
        } else if (n == 2) {
          if (fuse) KPM_TRY(spmm_dual(ctx, H, &M, v2, &jpsi0r, nullptr, &rr[1], &R[j - 2], 1.0, 0.0));
          else KPM_TRY(spmm(ctx, H, &jpsi0r, nullptr, &rr[1], 1.0, 0.0));
        } else {
          if (fuse) KPM_TRY(spmm_dual(ctx, H, &M, v2, &rr[rip], &rr[ripp], &rr[ri], &R[j - 2], 2.0, -1.0));
          else KPM_TRY(spmm(ctx, H, &rr[rip], &rr[ripp], &rr[ri], 2.0, -1.0));
        }
        if (!dual) KPM_TRY(spmm(ctx, Jb, &rr[ri], nullptr, &R[j - 1], 1.0, 0.0));
      }
      if (dual)  // the block's last order
        KPM_TRY(spmm(ctx, Jb, &rr[(int)((n2 - 1) % 3)], nullptr, &R[n_count - 1], 1.0, 0.0));
      KPM_TRY(launch_contract(ctx, ctx->slots_l.ptr, (int)rep_size, ctx->slots_r.ptr, (int)n_count,
                              Kp, pitch, prec, p.gemm_mode, scale, mu_dev, NC, n1 - 1, m1 - 1,
                              p.parity));
    }
    if (rep < reps) {  // chebyshev_iter_wrap(H, L, K_left)
      KPM_TRY(spmm(ctx, H, &L[KL - 1], &L[KL - 2], &L[0], 2.0, -1.0));
      KPM_TRY(spmm(ctx, H, &L[0], &L[KL - 1], &L[1], 2.0, -1.0));
    }
  }
  return KPM_OK;
}

// ---- kpm_3d with the outer orders batched as probe columns ------------------
//
// kpm_3d's NC nested kpm_2d calls share the left chain T_m psi_l and differ
// only in the right starter Jg T_{p-1} Ja psi_r.  With every left order
// resident, the left chain is built once and P outer orders run as one right
// chain over a wide block of P * NRp columns: column block q holds outer
// order b + q.  Every column sees exactly the arithmetic of the narrow chain,
// so every contraction input is bitwise that of the nested form.  After each
// order's Jb pass the wide block is scattered into P dense per-order slot
// arrays (the contraction's layout, one pitch), and the whole batch is one
// contraction of the NC left slots against its P * NC right slots.

// column block q of a wide block (NRw per row) <- a narrow block (NRp per row)
template <typename E>
__global__ void __launch_bounds__(256)
    k_to_wide(const E* __restrict__ src, E* __restrict__ dst, int64_t NH, int NRp, int NRw, int q) {
  const int64_t n = NH * NRp;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const int64_t row = i / NRp;
    const int c = (int)(i - row * NRp);
    dst[row * NRw + (int64_t)q * NRp + c] = src[i];
  }
}

// wide block -> P narrow slots: slot q at dst + q * qstride (elements)
template <typename E>
__global__ void __launch_bounds__(256)
    k_from_wide(const E* __restrict__ src, E* __restrict__ dst, int64_t NH, int NRp, int P,
                int64_t qstride) {
  const int NRw = NRp * P;
  const int64_t n = NH * NRw;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const int64_t row = i / NRw;
    const int c = (int)(i - row * NRw);
    const int q = c / NRp;
    const int r = c - q * NRp;
    dst[(int64_t)q * qstride + row * NRp + r] = src[i];
  }
}

static unsigned copy_grid(kpm_ctx* ctx, int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)std::max(ctx->num_cus, 8) * 16));
}

static int to_wide(kpm_ctx* ctx, const kpm_block* src, void* dst, int NRw, int q) {
  const int64_t n = src->NH * src->NRp;
  if (src->prec == KPM_VEC_F32)
    hipLaunchKernelGGL(k_to_wide<float2>, dim3(copy_grid(ctx, n)), dim3(256), 0, ctx->stream,
                       (const float2*)src->d, (float2*)dst, src->NH, (int)src->NRp, NRw, q);
  else
    hipLaunchKernelGGL(k_to_wide<double2>, dim3(copy_grid(ctx, n)), dim3(256), 0, ctx->stream,
                       (const double2*)src->d, (double2*)dst, src->NH, (int)src->NRp, NRw, q);
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

static int from_wide(kpm_ctx* ctx, const kpm_block* src, int64_t NRp, int P, void* dst,
                     int64_t qstride) {
  const int64_t n = src->NH * NRp * P;
  if (src->prec == KPM_VEC_F32)
    hipLaunchKernelGGL(k_from_wide<float2>, dim3(copy_grid(ctx, n)), dim3(256), 0, ctx->stream,
                       (const float2*)src->d, (float2*)dst, src->NH, (int)NRp, P, qstride);
  else
    hipLaunchKernelGGL(k_from_wide<double2>, dim3(copy_grid(ctx, n)), dim3(256), 0, ctx->stream,
                       (const double2*)src->d, (double2*)dst, src->NH, (int)NRp, P, qstride);
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

// Outer orders per batch: the largest power of two P <= NC, <= 64, such that
// the NC left slots, P * NC right slots and four wide blocks fit in half of
// the free memory (plus what this context already holds for them); 0 when
// fewer than two fit.
static int64_t batch_3d(kpm_ctx* ctx, const Plan2d& p) {
  const int64_t NRp = padded_nr(p.NR, p.prec);
  const int64_t eb = elem_bytes(p.prec);
  const int64_t Kp = (p.NH * NRp + 15) / 16 * 16;
  const int64_t stride = (Kp + 256 / eb) * eb;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
  size_t held = ctx->slots_l.bytes + ctx->slots_r.bytes + ctx->mu.bytes;
  for (const auto& b : ctx->ring) held += b.bytes;
  const int64_t budget = (int64_t)((fr + held) / 2);
  const int64_t block = p.NH * NRp * eb;
  int64_t P = 1;
  while (P * 2 <= std::min<int64_t>(p.NC, 64)) {
    const int64_t Q = P * 2;
    const int64_t need = stride * p.NC * (1 + Q) + 4 * block * Q + Q * p.NC * p.NC * 16;
    if (need > budget) break;
    P = Q;
  }
  return P >= 2 ? P : 0;
}

static int run_3d_batched(kpm_ctx* ctx, const kpm_csr* H, const kpm_csr* Ja, const kpm_csr* Jb,
                          const kpm_csr* Jg, const Plan2d& p, int64_t P, const kpm_block* l,
                          const kpm_block* r, kpm_block* outer, double* mu) {
  const int64_t NH = p.NH, NR = p.NR, NC = p.NC;
  const int prec = p.prec;
  const int64_t NRp = padded_nr(NR, prec);
  const int64_t NRw = NRp * P;
  if (padded_nr(NRw, prec) != NRw) return fail(KPM_EINVAL, "kpm_3d batch width is not a row width");
  const int64_t used = NH * NRp;
  const int64_t Kp = (used + 15) / 16 * 16;
  const int64_t eb = elem_bytes(prec);
  const int64_t pitch = Kp + 256 / eb;
  const int64_t stride = pitch * eb;
  KPM_TRY(ensure(ctx, ctx->slots_l, (size_t)(stride * NC)));
  KPM_TRY(ensure(ctx, ctx->slots_r, (size_t)(stride * NC * P)));
  if (Kp > used) {
    KPM_HIP_TRY(hipMemset2DAsync((char*)ctx->slots_l.ptr + used * eb, (size_t)stride, 0,
                                 (size_t)((Kp - used) * eb), (size_t)NC, ctx->stream));
    KPM_HIP_TRY(hipMemset2DAsync((char*)ctx->slots_r.ptr + used * eb, (size_t)stride, 0,
                                 (size_t)((Kp - used) * eb), (size_t)(NC * P), ctx->stream));
  }
  const size_t slab = (size_t)(NC * NC);
  KPM_TRY(ensure(ctx, ctx->mu, slab * (size_t)P * sizeof(double2)));
  kpm_block W[3], Rw, tmp;
  for (int i = 0; i < 3; ++i) KPM_TRY(work_block(ctx, ctx->ring[i], NH, NRw, prec, &W[i]));
  KPM_TRY(work_block(ctx, ctx->ring[3], NH, NRw, prec, &Rw));
  KPM_TRY(work_block(ctx, ctx->ring[4], NH, NR, prec, &tmp));
  std::vector<kpm_block> L(NC);
  for (int64_t i = 0; i < NC; ++i) slot_view(ctx->slots_l.ptr, stride, i, NH, NR, prec, ctx, &L[i]);
  kpm_csr M;
  const void* v2 = nullptr;
  bool dual = false;
  KPM_TRY(dual_operator(ctx, H, Jb, NRw, prec, &M, &v2, &dual));

  // the left chain, once (:1013-1018 and chebyshev_iter over all NC orders)
  KPM_TRY(kpm_block_copy(&L[0], l));
  if (NC > 1) KPM_TRY(spmm(ctx, H, l, nullptr, &L[1], 1.0, 0.0));
  for (int64_t i = 3; i <= NC; ++i) KPM_TRY(spmm(ctx, H, &L[i - 2], &L[i - 3], &L[i - 1], 2.0, -1.0));

  const double scale = 1.0 / (double)NR;
  const int64_t qstride = stride / eb;  // one right slot, in elements
  // each batch table comes back through page-locked staging
  KPM_TRY(ensure_pinned(ctx, slab * (size_t)P * sizeof(double2)));
  const double* host = (const double*)ctx->pinned.ptr;
  for (int64_t b = 1; b <= NC; b += P) {
    const int64_t Pb = std::min<int64_t>(P, NC - b + 1);
    // wide starter: column block q = Jg T_{b+q-1} Ja psi_r (unused blocks stay zero)
    KPM_HIP_TRY(hipMemsetAsync(W[0].d, 0, (size_t)(NH * NRw * eb), ctx->stream));
    for (int64_t q = 0; q < Pb; ++q) {
      const int64_t n1 = b + q;
      const int ri = (int)((n1 - 1) % 3), rip = (int)((n1 + 1) % 3), ripp = (int)(n1 % 3);
      if (n1 == 1) {
        KPM_TRY(spmm(ctx, Ja, r, nullptr, &outer[0], 1.0, 0.0));
      } else if (n1 == 2) {
        KPM_TRY(spmm(ctx, H, &outer[rip], nullptr, &outer[1], 1.0, 0.0));
      } else {
        KPM_TRY(spmm(ctx, H, &outer[rip], &outer[ripp], &outer[ri], 2.0, -1.0));
      }
      KPM_TRY(spmm(ctx, Jg, &outer[ri], nullptr, &tmp, 1.0, 0.0));
      KPM_TRY(to_wide(ctx, &tmp, W[0].d, (int)NRw, (int)q));
    }
    // the right chain over all outer orders of the batch (:1039-1051); with
    // the dual pass, the step of order n writes Jb W of order n - 1 straight
    // into that order's P slots (column block q -> slot q * NC + n - 2)
    for (int64_t n = 1; n <= NC; ++n) {
      const int ri = (int)((n - 1) % 3), rip = (int)((n + 1) % 3), ripp = (int)(n % 3);
      kpm_block slot;
      if (dual && n >= 2) slot_view(ctx->slots_r.ptr, stride, n - 2, NH, NRw, prec, ctx, &slot);
      if (n == 2) {
        if (dual) KPM_TRY(spmm_dual(ctx, H, &M, v2, &W[0], nullptr, &W[1], &slot, 1.0, 0.0, NRp, qstride * NC));
        else KPM_TRY(spmm(ctx, H, &W[0], nullptr, &W[1], 1.0, 0.0));
      } else if (n >= 3) {
        if (dual) KPM_TRY(spmm_dual(ctx, H, &M, v2, &W[rip], &W[ripp], &W[ri], &slot, 2.0, -1.0, NRp, qstride * NC));
        else KPM_TRY(spmm(ctx, H, &W[rip], &W[ripp], &W[ri], 2.0, -1.0));
      }
      if (!dual) {
        KPM_TRY(spmm(ctx, Jb, &W[ri], nullptr, &Rw, 1.0, 0.0));
        KPM_TRY(from_wide(ctx, &Rw, NRp, (int)P, (char*)ctx->slots_r.ptr + stride * (n - 1),
                          qstride * NC));
      }
    }
    if (dual) {  // the last order
      KPM_TRY(spmm(ctx, Jb, &W[(int)((NC - 1) % 3)], nullptr, &Rw, 1.0, 0.0));
      KPM_TRY(from_wide(ctx, &Rw, NRp, (int)P, (char*)ctx->slots_r.ptr + stride * (NC - 1),
                        qstride * NC));
    }
    // one contraction of the NC left slots against all Pb * NC right slots of
    // the batch (one pitch), into the stacked table
    // mu_b[(q NC + n) + Pb NC m], then reordered into the Pb outer slabs
    const int64_t ld = Pb * NC;
    KPM_HIP_TRY(hipMemsetAsync(ctx->mu.ptr, 0, slab * (size_t)Pb * sizeof(double2), ctx->stream));
    KPM_TRY(launch_contract(ctx, ctx->slots_l.ptr, (int)NC, ctx->slots_r.ptr, (int)ld, Kp, pitch,
                            prec, p.gemm_mode, scale, ctx->mu.ptr, ld, 0, 0, KPM_PARITY_NONE));
    KPM_HIP_TRY(hipMemcpyAsync(ctx->pinned.ptr, ctx->mu.ptr, slab * (size_t)Pb * sizeof(double2),
                               hipMemcpyDeviceToHost, ctx->stream));
    KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int64_t q = 0; q < Pb; ++q) {
      double* dst = mu + 2 * slab * (size_t)(b - 1 + q);
      for (int64_t m = 0; m < NC; ++m)
        std::copy(host + 2 * (q * NC + ld * m), host + 2 * (q * NC + ld * m + NC), dst + 2 * NC * m);
    }
  }
  return KPM_OK;
}

static void apply_mn_sym(std::vector<double>& mu, int64_t NC) {  // :1078-1086
  for (int64_t m = 0; m < NC; ++m)
    for (int64_t n = m + 1; n < NC; ++n) {
      const int64_t a = m + NC * n, b = n + NC * m;
      mu[2 * a + 1] = 0.0;
      mu[2 * b] = mu[2 * a];
      mu[2 * b + 1] = 0.0;
    }
}

static int validate_2d(const kpm_csr* H, const kpm_csr* Ja, const kpm_csr* Jb, const Plan2d& p) {
  if (!H || !Ja || !Jb) return fail(KPM_EINVAL, "null operator");
  if (p.parity < 0 || p.parity > 2) return fail(KPM_EINVAL, "moment_parity not understood.");
  if (p.NH != H->n || Ja->n != p.NH || Jb->n != p.NH)
    return fail(KPM_EINVAL, "operator dimensions do not match NH");
  if (p.NR <= 0 || p.NC < 1) return fail(KPM_EINVAL, "NR and NC must be positive");
  if (p.K_left < 2) return fail(KPM_EINVAL, "ψall_l must have at least two slots");
  if (p.K_right < 1) return fail(KPM_EINVAL, "JTnHJψr must have at least one slot");
  if (p.K_left > INT32_MAX || p.K_right > INT32_MAX) return fail(KPM_EINVAL, "block too large");
  if (p.prec != KPM_VEC_F64 && p.prec != KPM_VEC_F32) return fail(KPM_EINVAL, "bad vec_prec");
  if (p.gemm_mode != KPM_GEMM_F64 && p.gemm_mode != KPM_GEMM_F32 && p.gemm_mode != KPM_GEMM_BF16X3)
    return fail(KPM_EINVAL, "bad gemm_mode");
  return KPM_OK;
}

static int64_t fld(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}

}  // namespace kpm

using namespace kpm;

extern "C" {

int kpm_2d_blocking(kpm_ctx* ctx, int64_t NH, int64_t NR, int64_t NC, int64_t budget,
                    int64_t arr_size, int64_t right_block, int64_t* K_left, int64_t* K_right,
                    int64_t* bytes) {
  if (!K_left || !K_right || !bytes) return fail(KPM_EINVAL, "null output");
  if (NH <= 0) return fail(KPM_EINVAL, "NH must be positive (got " + std::to_string(NH) + ")");
  if (NR <= 0) return fail(KPM_EINVAL, "NR must be positive (got " + std::to_string(NR) + ")");
  if (NC < 2) return fail(KPM_EINVAL, "NC must be at least 2 (got " + std::to_string(NC) + ")");
  const int64_t block_bytes = NH * NR * 16;
  if (budget < 0) {  // :auto -> 50% of free device memory (:741)
    if (!ctx) return fail(KPM_EINVAL, "auto budget needs a context");
    size_t f = 0, t = 0;
    KPM_HIP_TRY(hipSetDevice(ctx->device));
    KPM_HIP_TRY(hipMemGetInfo(&f, &t));
    budget = (int64_t)(f / 2);
  }
  const int64_t KR = right_block <= 0 ? std::min<int64_t>(NC, 16) : right_block;
  const int64_t base = (KR + 7) * block_bytes;
  const int64_t per_left = block_bytes + KR * 16;
  int64_t KL;
  if (arr_size <= 0) {
    const int64_t kmax = fld(budget - base, per_left);
    if (kmax < 2)
      return fail(KPM_EINVAL, "workspace_bytes=" + std::to_string(budget) +
                                  " cannot fit the minimum kpm_2d workspace (" +
                                  std::to_string(base + 2 * per_left) + " bytes at right_block=" +
                                  std::to_string(KR) + ")");
    KL = std::min(kmax, NC);
  } else {
    if (arr_size < 2) return fail(KPM_EINVAL, "arr_size must be :auto or an integer at least 2");
    KL = arr_size;
  }
  *K_left = KL;
  *K_right = KR;
  *bytes = base + KL * per_left;
  return KPM_OK;
}

int kpm_2d(kpm_ctx* ctx, const kpm_csr* H, const kpm_csr* Ja, const kpm_csr* Jb, int64_t NC,
           int64_t NR, int64_t NH, const double* psi_l, const double* psi_r, int64_t K_left,
           int64_t K_right, int parity, int mn_sym, int vec_prec, int gemm_mode, double* mu) {
  if (!ctx || !psi_l || !psi_r || !mu) return fail(KPM_EINVAL, "null argument");
  Plan2d p{NC, NR, NH, K_left, K_right, parity, mn_sym, vec_prec, gemm_mode};
  KPM_TRY(validate_2d(H, Ja, Jb, p));
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  kpm_block l, r;
  KPM_TRY(work_block(ctx, ctx->lwork, NH, NR, vec_prec, &l));
  KPM_TRY(work_block(ctx, ctx->work[0], NH, NR, vec_prec, &r));
  KPM_TRY(ensure(ctx, ctx->staging, (size_t)(NH * NR) * 16));
  KPM_HIP_TRY(hipMemcpyAsync(ctx->staging.ptr, psi_l, (size_t)(NH * NR) * 16, hipMemcpyHostToDevice, ctx->stream));
  KPM_TRY(launch_upload(ctx, ctx->staging.ptr, &l));
  KPM_HIP_TRY(hipMemcpyAsync(ctx->staging.ptr, psi_r, (size_t)(NH * NR) * 16, hipMemcpyHostToDevice, ctx->stream));
  KPM_TRY(launch_upload(ctx, ctx->staging.ptr, &r));
  KPM_TRY(ensure(ctx, ctx->mu, (size_t)(NC * NC) * sizeof(double2)));
  KPM_TRY(run_2d(ctx, H, Ja, Jb, p, &l, &r, ctx->mu.ptr));
  KPM_HIP_TRY(hipMemcpyAsync(mu, ctx->mu.ptr, (size_t)(NC * NC) * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
  KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (mn_sym) {
    std::vector<double> m(mu, mu + 2 * NC * NC);
    apply_mn_sym(m, NC);
    std::copy(m.begin(), m.end(), mu);
  }
  return KPM_OK;
}

int kpm_3d(kpm_ctx* ctx, const kpm_csr* H, const kpm_csr* Ja, const kpm_csr* Jb, const kpm_csr* Jg,
           int64_t NC, int64_t NR, int64_t NH, const double* psi_l, const double* psi_r,
           int64_t K_left, int64_t K_right, int vec_prec, int gemm_mode, double* mu) {
  if (!ctx || !psi_l || !psi_r || !mu || !Jg) return fail(KPM_EINVAL, "null argument");
  Plan2d p{NC, NR, NH, K_left, K_right, KPM_PARITY_NONE, 0, vec_prec, gemm_mode};
  KPM_TRY(validate_2d(H, Ja, Jb, p));
  if (Jg->n != NH) return fail(KPM_EINVAL, "operator dimensions do not match NH");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  kpm_block l, r, outer[3];
  KPM_TRY(work_block(ctx, ctx->lwork, NH, NR, vec_prec, &l));
  KPM_TRY(work_block(ctx, ctx->work[0], NH, NR, vec_prec, &r));
  for (int i = 0; i < 3; ++i) KPM_TRY(work_block(ctx, ctx->work[1 + i], NH, NR, vec_prec, &outer[i]));
  KPM_TRY(ensure(ctx, ctx->staging, (size_t)(NH * NR) * 16));
  KPM_HIP_TRY(hipMemcpyAsync(ctx->staging.ptr, psi_l, (size_t)(NH * NR) * 16, hipMemcpyHostToDevice, ctx->stream));
  KPM_TRY(launch_upload(ctx, ctx->staging.ptr, &l));
  KPM_HIP_TRY(hipMemcpyAsync(ctx->staging.ptr, psi_r, (size_t)(NH * NR) * 16, hipMemcpyHostToDevice, ctx->stream));
  KPM_TRY(launch_upload(ctx, ctx->staging.ptr, &r));
  // every left order resident: the batched form (left chain once, P outer
  // orders per right chain) when at least two outer orders fit
  if (K_left >= NC) {
    const int64_t P = batch_3d(ctx, p);
    if (P >= 2) return run_3d_batched(ctx, H, Ja, Jb, Jg, p, P, &l, &r, outer, mu);
  }
  KPM_TRY(ensure(ctx, ctx->mu, (size_t)(NC * NC) * sizeof(double2)));
  const size_t slab = (size_t)(NC * NC);
  for (int64_t n1 = 1; n1 <= NC; ++n1) {
    const int ri = (int)((n1 - 1) % 3), rip = (int)((n1 + 1) % 3), ripp = (int)(n1 % 3);
    if (n1 == 1) {
      KPM_TRY(spmm(ctx, Ja, &r, nullptr, &outer[0], 1.0, 0.0));
    } else if (n1 == 2) {
      KPM_TRY(spmm(ctx, H, &outer[rip], nullptr, &outer[1], 1.0, 0.0));
    } else {
      KPM_TRY(spmm(ctx, H, &outer[rip], &outer[ripp], &outer[ri], 2.0, -1.0));
    }
    // kpm_2d!(H, Jg, Jb, ..., psi_l, T_{n1-1} Ja psi_r) into mu[:,:,n1]
    KPM_TRY(run_2d(ctx, H, Jg, Jb, p, &l, &outer[ri], ctx->mu.ptr));
    KPM_HIP_TRY(hipMemcpyAsync(mu + 2 * slab * (size_t)(n1 - 1), ctx->mu.ptr, slab * sizeof(double2),
                               hipMemcpyDeviceToHost, ctx->stream));
  }
  KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return KPM_OK;
}

}  // extern "C"
