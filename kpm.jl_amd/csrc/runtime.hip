// runtime.hip -- contexts, operators, blocks, errors, timing.
//
// This is the device layer of the replacement for ext/KPMCUDAExt.jl:
//   CUDADevice + __init__ activation          ext/KPMCUDAExt.jl:13-20
//   to_device(::SparseMatrixCSC) -> CSR         ext/KPMCUDAExt.jl:26-27
//   to_device / maybe_to_host of dense blocks   ext/KPMCUDAExt.jl:28,36-39
//   device_zeros / device_free_memory           ext/KPMCUDAExt.jl:43-45
#include <algorithm>
#include <atomic>
#include <cstring>
#include <string>
#include <vector>

#include "kpm_internal.h"

namespace kpm {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int ensure(kpm_ctx* ctx, Buffer& b, size_t bytes) {
  if (b.bytes >= bytes && b.ptr) return KPM_OK;
  if (b.ptr) {
    KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    KPM_HIP_TRY(hipFree(b.ptr));
    b.ptr = nullptr;
    b.bytes = 0;
  }
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(&b.ptr, bytes);
  if (e != hipSuccess) {
    b.ptr = nullptr;
    // HIP keeps a failed call's status until hipGetLastError reads it; a
    // caller that tolerates the failure (and falls back) must not see it
    // again at its next launch check
    (void)hipGetLastError();
    return fail(KPM_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                hipGetErrorString(e));
  }
  b.bytes = bytes;
  return KPM_OK;
}

int work_block(kpm_ctx* ctx, Buffer& buf, int64_t NH, int64_t NR, int prec, kpm_block* out) {
  const int64_t NRp = padded_nr(NR, prec);
  KPM_TRY(ensure(ctx, buf, (size_t)(NH * NRp) * elem_bytes(prec)));
  out->ctx = ctx;
  out->NH = NH;
  out->NR = NR;
  out->NRp = NRp;
  out->prec = prec;
  out->d = buf.ptr;
  out->owned = false;
  return KPM_OK;
}

// identity and row order of a new operator: the serial keys the merged
// operator of the dual pass (a freed operator's address can be reused, its
// serial cannot), a merge of two operators needs strictly increasing column
// indices in every row, and the fused pass reuses the diagonal's gather when
// every row stores its diagonal within the pass's gather slots
static void stamp_csr(kpm_csr* A, const int* rp, const int* ci) {
  static std::atomic<uint64_t> next{1};
  A->serial = next.fetch_add(1);
  bool sorted = true;
  for (int64_t i = 0; i < A->n && sorted; ++i)
    for (int k = rp[i] + 1; k < rp[i + 1]; ++k)
      if (ci[k] <= ci[k - 1]) {
        sorted = false;
        break;
      }
  A->sorted = sorted ? 1 : 0;
  int slots = 0;
  for (int64_t i = 0; i < A->n && slots < INT32_MAX; ++i) {
    int pos = INT32_MAX;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (ci[k] == (int)i) {
        pos = k - rp[i] + 1;
        break;
      }
    slots = std::max(slots, pos);
  }
  A->diag_slots = slots;
}

static void free_buffer(Buffer& b) {
  if (b.ptr) (void)hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
}

__global__ void k_round_f32(const double* __restrict__ src, float* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];  // round to nearest even, as the FP32 pass's conversion on load
}

int attach_vals32(kpm_ctx* ctx, kpm_csr* A) {
  const int64_t nreal = (A->nnz + kCsrPad) * (A->is_real ? 1 : 2);
  hipError_t e = hipMalloc(&A->vals32, (size_t)nreal * sizeof(float));
  if (e != hipSuccess) {
    A->vals32 = nullptr;
    (void)hipGetLastError();  // clear the latched status (see ensure)
    return fail(KPM_ENOMEM, std::string("hipMalloc for FP32 operator values failed: ") +
                                hipGetErrorString(e));
  }
  KPM_HIP_TRY(hipMemsetAsync(A->vals32, 0, (size_t)nreal * sizeof(float), ctx->stream));  // the zero tail
  if (A->nnz > 0) {
    const int64_t n = A->nnz * (A->is_real ? 1 : 2);
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_round_f32, dim3(blocks), dim3(256), 0, ctx->stream, (const double*)A->vals,
                       (float*)A->vals32, n);
    KPM_HIP_TRY(hipGetLastError());
  }
  return KPM_OK;
}

// Build device CSR arrays from host (rowptr, colind, vals) already in CSR
// form with 0-based int64 indices.
int upload_csr(kpm_ctx* ctx, int64_t n, const std::vector<int64_t>& rp,
                      const std::vector<int64_t>& ci, const std::vector<double>& v_cplx,
                      kpm_csr** out) {
  const int64_t nnz = rp[n];
  if (n >= INT32_MAX || nnz >= INT32_MAX)
    return fail(KPM_EINVAL, "operator too large for int32 device indices (n or nnz >= 2^31)");
  bool is_real = true;
  for (int64_t k = 0; k < nnz; ++k)
    if (v_cplx[2 * k + 1] != 0.0) {
      is_real = false;
      break;
    }
  std::vector<int> rp32(n + 1), ci32(std::max<int64_t>(nnz, 1));
  int maxrow = 0;
  for (int64_t i = 0; i <= n; ++i) rp32[i] = (int)rp[i];
  for (int64_t i = 0; i < n; ++i) maxrow = std::max(maxrow, (int)(rp[i + 1] - rp[i]));
  for (int64_t k = 0; k < nnz; ++k) ci32[k] = (int)ci[k];
  kpm_csr* A = new kpm_csr();
  A->ctx = ctx;
  A->n = n;
  A->nnz = nnz;
  A->is_real = is_real ? 1 : 0;
  A->max_row_nnz = maxrow;
  stamp_csr(A, rp32.data(), ci32.data());
  const size_t vbytes = (size_t)(nnz + kCsrPad) * (is_real ? sizeof(double) : 2 * sizeof(double));
  const size_t cbytes = sizeof(int) * (size_t)(nnz + kCsrPad);
  hipError_t e = hipMalloc(&A->rowptr, sizeof(int) * (n + 1));
  if (e == hipSuccess) e = hipMalloc(&A->colind, cbytes);
  if (e == hipSuccess) e = hipMalloc(&A->vals, vbytes);
  if (e == hipSuccess) e = hipMemset(A->colind, 0, cbytes);  // the zero tail
  if (e == hipSuccess) e = hipMemset(A->vals, 0, vbytes);
  if (e != hipSuccess) {
    if (A->rowptr) (void)hipFree(A->rowptr);
    if (A->colind) (void)hipFree(A->colind);
    if (A->vals) (void)hipFree(A->vals);
    delete A;
    (void)hipGetLastError();  // clear the latched status (see ensure)
    return fail(KPM_ENOMEM, std::string("hipMalloc for CSR failed: ") + hipGetErrorString(e));
  }
  std::vector<double> vr;
  const void* vsrc = v_cplx.data();
  if (is_real) {
    vr.resize(std::max<int64_t>(nnz, 1));
    for (int64_t k = 0; k < nnz; ++k) vr[k] = v_cplx[2 * k];
    vsrc = vr.data();
  }
  e = hipMemcpy(A->rowptr, rp32.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess && nnz > 0)
    e = hipMemcpy(A->colind, ci32.data(), sizeof(int) * nnz, hipMemcpyHostToDevice);
  if (e == hipSuccess && nnz > 0) e = hipMemcpy(A->vals, vsrc, nnz * (is_real ? 8 : 16), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(A->rowptr);
    (void)hipFree(A->colind);
    (void)hipFree(A->vals);
    delete A;
    return fail(KPM_EHIP, std::string("hipMemcpy for CSR failed: ") + hipGetErrorString(e));
  }
  const int rc = attach_vals32(ctx, A);
  if (rc != KPM_OK) {
    kpm_csr_destroy(A);
    return rc;
  }
  *out = A;
  return KPM_OK;
}

int ensure_pinned(kpm_ctx* ctx, size_t bytes) {
  Buffer& b = ctx->pinned;
  if (b.bytes >= bytes && b.ptr) return KPM_OK;
  if (b.ptr) {
    KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    KPM_HIP_TRY(hipHostFree(b.ptr));
    b.ptr = nullptr;
    b.bytes = 0;
  }
  hipError_t e = hipHostMalloc(&b.ptr, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    b.ptr = nullptr;
    return fail(KPM_ENOMEM, "hipHostMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                hipGetErrorString(e));
  }
  b.bytes = bytes;
  return KPM_OK;
}

int csr_from_csc_staged(kpm_ctx* ctx, int64_t n, const int64_t* colptr, const int64_t* rowval,
                        const double* nzval, int base, bool adjoint, std::vector<int>* perm,
                        kpm_csr** out) {
  const int64_t nnz = colptr[n] - base;
  if (n >= INT32_MAX || nnz >= INT32_MAX)
    return fail(KPM_EINVAL, "operator too large for int32 device indices (n or nnz >= 2^31)");
  bool is_real = true;
  for (int64_t k = 0; k < nnz && is_real; ++k) is_real = nzval[2 * k + 1] == 0.0;
  const int64_t nz1 = std::max<int64_t>(nnz, 1);
  const size_t brp = (size_t)(n + 1) * sizeof(int), bci = (size_t)nz1 * sizeof(int);
  const size_t bv = (size_t)nz1 * (is_real ? sizeof(double) : 2 * sizeof(double));
  // device arrays: the staged entries plus the zero tail
  const size_t bci_dev = (size_t)(nnz + kCsrPad) * sizeof(int);
  const size_t bv_dev = (size_t)(nnz + kCsrPad) * (is_real ? sizeof(double) : 2 * sizeof(double));
  const size_t off_ci = (brp + 255) / 256 * 256, off_v = off_ci + (bci + 255) / 256 * 256;
  KPM_TRY(ensure_pinned(ctx, off_v + bv));
  char* base_p = (char*)ctx->pinned.ptr;
  int* rp = (int*)base_p;
  int* ci = (int*)(base_p + off_ci);
  double* vr = (double*)(base_p + off_v);
  const double sgn = adjoint ? -1.0 : 1.0;
  if (adjoint) {
    // row j of adjoint(A) is column j of A, conjugated: no scatter
    for (int64_t j = 0; j <= n; ++j) rp[j] = (int)(colptr[j] - base);
    for (int64_t k = 0; k < nnz; ++k) {
      ci[k] = (int)(rowval[k] - base);
      if (is_real) {
        vr[k] = nzval[2 * k];
      } else {
        vr[2 * k] = nzval[2 * k];
        vr[2 * k + 1] = sgn * nzval[2 * k + 1];
      }
    }
  } else {
    // CSR(A): count entries per row, then scatter columns in increasing order
    std::fill(rp, rp + n + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) rp[rowval[k] - base + 1]++;
    for (int64_t i = 0; i < n; ++i) rp[i + 1] += rp[i];
    std::vector<int> next(rp, rp + n);
    if (perm) perm->resize((size_t)nnz);
    for (int64_t j = 0; j < n; ++j)
      for (int64_t k = colptr[j] - base; k < colptr[j + 1] - base; ++k) {
        const int dst = next[rowval[k] - base]++;
        ci[dst] = (int)j;
        if (is_real) {
          vr[dst] = nzval[2 * k];
        } else {
          vr[2 * dst] = nzval[2 * k];
          vr[2 * dst + 1] = nzval[2 * k + 1];
        }
        if (perm) (*perm)[k] = dst;
      }
  }
  int maxrow = 0;
  for (int64_t i = 0; i < n; ++i) maxrow = std::max(maxrow, rp[i + 1] - rp[i]);
  kpm_csr* A = new kpm_csr();
  A->ctx = ctx;
  A->n = n;
  A->nnz = nnz;
  A->is_real = is_real ? 1 : 0;
  A->max_row_nnz = maxrow;
  stamp_csr(A, rp, ci);
  hipError_t e = hipMalloc(&A->rowptr, brp);
  if (e == hipSuccess) e = hipMalloc(&A->colind, bci_dev);
  if (e == hipSuccess) e = hipMalloc(&A->vals, bv_dev);
  if (e == hipSuccess) e = hipMemsetAsync(A->colind, 0, bci_dev, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(A->vals, 0, bv_dev, ctx->stream);
  if (e != hipSuccess) {
    if (A->rowptr) (void)hipFree(A->rowptr);
    if (A->colind) (void)hipFree(A->colind);
    if (A->vals) (void)hipFree(A->vals);
    delete A;
    (void)hipGetLastError();  // clear the latched status (see ensure)
    return fail(KPM_ENOMEM, std::string("hipMalloc for CSR failed: ") + hipGetErrorString(e));
  }
  e = hipMemcpyAsync(A->rowptr, rp, brp, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(A->colind, ci, bci, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(A->vals, vr, bv, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // staging is reused by the next call
  if (e != hipSuccess) {
    (void)hipFree(A->rowptr);
    (void)hipFree(A->colind);
    (void)hipFree(A->vals);
    delete A;
    return fail(KPM_EHIP, std::string("CSR upload failed: ") + hipGetErrorString(e));
  }
  const int rc = attach_vals32(ctx, A);
  if (rc != KPM_OK) {
    kpm_csr_destroy(A);
    return rc;
  }
  *out = A;
  return KPM_OK;
}

int check_csc(int64_t n, const int64_t* colptr, const int64_t* rowval, int base) {
  if (n <= 0) return fail(KPM_EINVAL, "operator dimension must be positive");
  if (base != 0 && base != 1) return fail(KPM_EINVAL, "index_base must be 0 or 1");
  if (!colptr) return fail(KPM_EINVAL, "null colptr");
  if (colptr[0] - base != 0) return fail(KPM_EINVAL, "colptr must start at index_base");
  const int64_t nnz = colptr[n] - base;
  if (nnz < 0) return fail(KPM_EINVAL, "negative nnz");
  for (int64_t j = 0; j < n; ++j)
    if (colptr[j + 1] < colptr[j]) return fail(KPM_EINVAL, "colptr must be nondecreasing");
  for (int64_t k = 0; k < nnz; ++k) {
    const int64_t r = rowval[k] - base;
    if (r < 0 || r >= n) return fail(KPM_EINVAL, "row index out of range");
  }
  return KPM_OK;
}

}  // namespace kpm

using namespace kpm;

extern "C" {

const char* kpm_last_error(void) { return g_last_error.c_str(); }

int kpm_abi_version(void) { return 1; }

int kpm_device_count(int* count) {
  if (!count) return fail(KPM_EINVAL, "null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return KPM_OK;
}

int kpm_ctx_create(int device, kpm_ctx** out) {
  if (!out) return fail(KPM_EINVAL, "null output");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(KPM_EHIP, "no HIP device is available");
  if (device < 0 || device >= n) return fail(KPM_EINVAL, "device index out of range");
  KPM_HIP_TRY(hipSetDevice(device));
  kpm_ctx* ctx = new kpm_ctx();
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
  if (ctx->num_cus <= 0) ctx->num_cus = 256;
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  for (int i = 0; e == hipSuccess && i < 16; ++i) e = hipEventCreate(&ctx->events[i]);
  if (e != hipSuccess) {
    delete ctx;
    return fail(KPM_EHIP, std::string("context creation failed: ") + hipGetErrorString(e));
  }
  // the allocation-free entry point's scratch address on this device; without
  // it only that entry point is unavailable
  if (noalloc_scratch_init(ctx) != KPM_OK) {
    (void)hipGetLastError();
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
  }
  *out = ctx;
  return KPM_OK;
}

int kpm_ctx_destroy(kpm_ctx* ctx) {
  if (!ctx) return KPM_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  free_buffer(ctx->partials);
  free_buffer(ctx->mu);
  free_buffer(ctx->staging);
  free_buffer(ctx->gemm_part);
  for (auto& b : ctx->work) free_buffer(b);
  free_buffer(ctx->lwork);
  free_buffer(ctx->slots_l);
  free_buffer(ctx->slots_r);
  for (auto& b : ctx->ring) free_buffer(b);
  free_buffer(ctx->adjoint);
  free_buffer(ctx->adj_grad);
  free_buffer(ctx->adj2);
  free_buffer(ctx->adj2_aux);
  free_buffer(ctx->merged);
  free_buffer(ctx->prim.part);
  if (ctx->pinned.ptr) (void)hipHostFree(ctx->pinned.ptr);
  for (auto& ev : ctx->events)
    if (ev) (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return KPM_OK;
}

int kpm_ctx_synchronize(kpm_ctx* ctx) {
  if (!ctx) return fail(KPM_EINVAL, "null context");
  KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return KPM_OK;
}

int kpm_ctx_set_dual_pass(kpm_ctx* ctx, int enable) {
  if (!ctx) return fail(KPM_EINVAL, "null context");
  ctx->dual_pass = enable ? 1 : 0;
  return KPM_OK;
}

int kpm_ctx_set_bf16_planes(kpm_ctx* ctx, int enable) {
  if (!ctx) return fail(KPM_EINVAL, "null context");
  ctx->bf16_planes = enable ? 1 : 0;
  return KPM_OK;
}

int kpm_ctx_set_merged_limit(kpm_ctx* ctx, int64_t bytes) {
  if (!ctx) return fail(KPM_EINVAL, "null context");
  ctx->merged_limit = bytes;
  ctx->merged_key[0] = ctx->merged_key[1] = 0;  // re-plan the merged operator
  return KPM_OK;
}

int kpm_ctx_counters(kpm_ctx* ctx, int64_t* counts, int n, int reset) {
  if (!ctx || (n > 0 && !counts)) return fail(KPM_EINVAL, "bad arguments");
  for (int i = 0; i < n && i < KPM_COUNT_N; ++i) counts[i] = ctx->counters[i];
  if (reset)
    for (int i = 0; i < KPM_COUNT_N; ++i) ctx->counters[i] = 0;
  return KPM_OK;
}

int kpm_device_free_memory(kpm_ctx* ctx, size_t* free_bytes, size_t* total_bytes) {
  if (!ctx) return fail(KPM_EINVAL, "null context");
  size_t f = 0, t = 0;
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  KPM_HIP_TRY(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return KPM_OK;
}

int kpm_device_name(kpm_ctx* ctx, char* buf, size_t len) {
  if (!ctx || !buf || len == 0) return fail(KPM_EINVAL, "bad arguments");
  hipDeviceProp_t prop;
  KPM_HIP_TRY(hipGetDeviceProperties(&prop, ctx->device));
  // the marketing name can be empty (it is on some driver stacks): then the
  // device index and its CU count name the part
  if (prop.name[0])
    std::snprintf(buf, len, "%s (%s)", prop.name, prop.gcnArchName);
  else
    std::snprintf(buf, len, "device %d, %d CUs (%s)", ctx->device, prop.multiProcessorCount,
                  prop.gcnArchName);
  return KPM_OK;
}

int kpm_csr_from_csc(kpm_ctx* ctx, int64_t n, const int64_t* colptr, const int64_t* rowval,
                     const double* nzval, int index_base, kpm_csr** out) {
  if (!ctx || !out) return fail(KPM_EINVAL, "null argument");
  KPM_TRY(check_csc(n, colptr, rowval, index_base));
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  std::vector<int> perm;
  KPM_TRY(csr_from_csc_staged(ctx, n, colptr, rowval, nzval, index_base, false, &perm, out));
  (*out)->csc_to_csr = std::move(perm);
  return KPM_OK;
}

int kpm_csr_adjoint_from_csc(kpm_ctx* ctx, int64_t n, const int64_t* colptr, const int64_t* rowval,
                             const double* nzval, int index_base, kpm_csr** out) {
  if (!ctx || !out) return fail(KPM_EINVAL, "null argument");
  KPM_TRY(check_csc(n, colptr, rowval, index_base));
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  return csr_from_csc_staged(ctx, n, colptr, rowval, nzval, index_base, true, nullptr, out);
}

int kpm_csr_from_csr(kpm_ctx* ctx, int64_t n, const int64_t* rowptr, const int64_t* colind,
                     const double* vals, int index_base, kpm_csr** out) {
  if (!ctx || !out) return fail(KPM_EINVAL, "null argument");
  // the CSR arrays of A are the CSC arrays of A^T: validate with the same rules
  KPM_TRY(check_csc(n, rowptr, colind, index_base));
  const int64_t nnz = rowptr[n] - index_base;
  std::vector<int64_t> rp(n + 1), ci(std::max<int64_t>(nnz, 1));
  std::vector<double> v(2 * std::max<int64_t>(nnz, 1));
  for (int64_t i = 0; i <= n; ++i) rp[i] = rowptr[i] - index_base;
  for (int64_t k = 0; k < nnz; ++k) {
    ci[k] = colind[k] - index_base;
    v[2 * k] = vals[2 * k];
    v[2 * k + 1] = vals[2 * k + 1];
  }
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  return upload_csr(ctx, n, rp, ci, v, out);
}

int kpm_csr_destroy(kpm_csr* A) {
  if (!A) return KPM_OK;
  (void)hipSetDevice(A->ctx->device);
  (void)hipStreamSynchronize(A->ctx->stream);
  (void)hipFree(A->rowptr);
  (void)hipFree(A->colind);
  (void)hipFree(A->vals);
  if (A->vals32) (void)hipFree(A->vals32);
  if (A->csc_to_csr_dev) (void)hipFree(A->csc_to_csr_dev);
  delete A;
  return KPM_OK;
}

int kpm_csr_info(const kpm_csr* A, int64_t* n, int64_t* nnz, int* is_real) {
  if (!A) return fail(KPM_EINVAL, "null operator");
  if (n) *n = A->n;
  if (nnz) *nnz = A->nnz;
  if (is_real) *is_real = A->is_real;
  return KPM_OK;
}

int kpm_block_create(kpm_ctx* ctx, int64_t NH, int64_t NR, int vec_prec, kpm_block** out) {
  if (!ctx || !out) return fail(KPM_EINVAL, "null argument");
  if (NH <= 0 || NR <= 0) return fail(KPM_EINVAL, "block dimensions must be positive");
  if (vec_prec != KPM_VEC_F64 && vec_prec != KPM_VEC_F32) return fail(KPM_EINVAL, "bad vec_prec");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  kpm_block* b = new kpm_block();
  b->ctx = ctx;
  b->NH = NH;
  b->NR = NR;
  b->NRp = padded_nr(NR, vec_prec);
  b->prec = vec_prec;
  const size_t bytes = (size_t)(NH * b->NRp) * elem_bytes(vec_prec);
  hipError_t e = hipMalloc(&b->d, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // clear the latched status (see ensure)
    delete b;
    return fail(KPM_ENOMEM, "hipMalloc of block failed: " + std::string(hipGetErrorString(e)));
  }
  e = hipMemsetAsync(b->d, 0, bytes, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(b->d);
    delete b;
    return fail(KPM_EHIP, std::string("block clear failed: ") + hipGetErrorString(e));
  }
  *out = b;
  return KPM_OK;
}

int kpm_block_destroy(kpm_block* b) {
  if (!b) return KPM_OK;
  if (b->ctx) ++b->ctx->epoch;  // its address may come back for another block
  if (b->owned && b->d) {
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    (void)hipFree(b->d);
  }
  delete b;
  return KPM_OK;
}

int kpm_block_upload(kpm_block* b, const double* host) {
  if (!b || !host) return fail(KPM_EINVAL, "null argument");
  kpm_ctx* ctx = b->ctx;
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  ++ctx->epoch;
  const size_t hbytes = (size_t)(b->NH * b->NR) * 16;
  KPM_TRY(ensure(ctx, ctx->staging, hbytes));
  KPM_HIP_TRY(hipMemcpyAsync(ctx->staging.ptr, host, hbytes, hipMemcpyHostToDevice, ctx->stream));
  KPM_TRY(launch_upload(ctx, ctx->staging.ptr, b));
  KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return KPM_OK;
}

int kpm_block_download(const kpm_block* b, double* host) {
  if (!b || !host) return fail(KPM_EINVAL, "null argument");
  kpm_ctx* ctx = b->ctx;
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const size_t hbytes = (size_t)(b->NH * b->NR) * 16;
  KPM_TRY(ensure(ctx, ctx->staging, hbytes));
  KPM_TRY(launch_download(ctx, b, ctx->staging.ptr));
  KPM_HIP_TRY(hipMemcpyAsync(host, ctx->staging.ptr, hbytes, hipMemcpyDeviceToHost, ctx->stream));
  KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return KPM_OK;
}

int kpm_block_copy(kpm_block* dst, const kpm_block* src) {
  if (!dst || !src) return fail(KPM_EINVAL, "null argument");
  if (dst->NH != src->NH || dst->NR != src->NR || dst->prec != src->prec)
    return fail(KPM_EINVAL, "block shapes or precisions differ");
  if (dst == src || dst->d == src->d) return KPM_OK;
  kpm_ctx* ctx = dst->ctx;
  ++ctx->epoch;
  KPM_HIP_TRY(hipMemcpyAsync(dst->d, src->d, (size_t)(dst->NH * dst->NRp) * elem_bytes(dst->prec),
                             hipMemcpyDeviceToDevice, ctx->stream));
  return KPM_OK;
}

int kpm_block_zero(kpm_block* b) {
  if (!b) return fail(KPM_EINVAL, "null argument");
  ++b->ctx->epoch;
  KPM_HIP_TRY(hipMemsetAsync(b->d, 0, (size_t)(b->NH * b->NRp) * elem_bytes(b->prec), b->ctx->stream));
  return KPM_OK;
}

int kpm_event_record(kpm_ctx* ctx, int slot) {
  if (!ctx || slot < 0 || slot >= 16) return fail(KPM_EINVAL, "bad event slot");
  KPM_HIP_TRY(hipEventRecord(ctx->events[slot], ctx->stream));
  return KPM_OK;
}

int kpm_event_elapsed(kpm_ctx* ctx, int slot_a, int slot_b, float* ms) {
  if (!ctx || !ms || slot_a < 0 || slot_a >= 16 || slot_b < 0 || slot_b >= 16)
    return fail(KPM_EINVAL, "bad event slot");
  KPM_HIP_TRY(hipEventSynchronize(ctx->events[slot_b]));
  KPM_HIP_TRY(hipEventElapsedTime(ms, ctx->events[slot_a], ctx->events[slot_b]));
  return KPM_OK;
}

}  // extern "C"
