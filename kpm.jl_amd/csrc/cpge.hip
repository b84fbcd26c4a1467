// cpge.hip -- trilinear forms of the kpm_3d moment table on the MFMA cores.
//
// The CPGE reconstruction evaluates, at every quadrature node theta, three
// trilinear forms of the kernel-weighted table mu~[n,m,p] (NC^3 complex, n
// fastest, then m, then p):
//   F = sum_{n,m,p} mu~[n,m,p] a[n] b[m] c[p]
// (_cpge_contract!, src/applications/cpge.jl:59-64: collapsed = mu~^T a,
// right = reshape(collapsed) c, F = sum(b .* right); plain transposes, no
// conjugation; the node function that builds a, b, c is :66-120).
//
// A call evaluates Q forms (three per node, a batch of nodes) in one pass
// over the table.  With columns j = m + NC p of the NC x NC^2 matrix mu~,
//   X[q, j] = sum_n a_q[n] mu~[n, j]
// is a (Q x NC) x (NC x NC^2) complex GEMM, run as two real GEMMs over the
// interleaved reals k' = 2n + {0, 1} of a column:
//   Re X = sum_k' a'[q, k'] mu'[k', j]     a' = (ar, -ai) per complex
//   Im X = sum_k' a''[q, k'] mu'[k', j]    a'' = (ai, ar)
// on v_mfma_f64_16x16x4f64 (a' and a'' are formed in registers from the
// pair partner).  The epilogue multiplies X[q, j] by b_q[m] c_q[p] and sums
// over the workgroup's columns in a fixed order (16-lane butterfly, then the
// four waves); a second kernel adds the per-workgroup partials in column
// order.  Results do not depend on timing or the launch shape.
//
// The table stays resident on the device (kpm_table3) across the many
// calls of one adaptive quadrature.
#include "kpm_internal.h"

#include <algorithm>

namespace kpm {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kTQ = 64;  // forms per workgroup: four 16-row MFMA blocks
constexpr int kTJ = 64;  // table columns per workgroup: 16 per wave
constexpr int kStage = 32;  // reals of a column per stage (8 per lane group)
constexpr int kRB = kTQ / 16;
constexpr int kLdaA = kStage + 2;  // LDS row stride of the form rows (16-byte pad)

// Each lane (g = lane / 16, i = lane % 16) owns column j0 + i of its wave's
// 16 columns and, in each 32-real stage, reals k0 + 8g .. k0 + 8g + 7 of that
// column and of its four form rows q0 + 16 rb + i; MFMA t of the stage
// contracts element t of every lane group (a fixed permutation of k', the
// same for both operands).  The form rows are common to the four waves: each
// stage of them is loaded once per workgroup (coalesced, double-buffered in
// LDS, the 16-byte row pad puts the 16 rows a lane group reads at one offset
// in 16 different bank quads); the table columns stream from global memory
// one stage ahead.
__global__ void __launch_bounds__(256)
    k_trilinear(const double* __restrict__ mu, int NC, const double* __restrict__ A,
                const double2* __restrict__ B, const double2* __restrict__ C, int Q, int ntx,
                int nqy, double2* __restrict__ part) {
  __shared__ double2 red[4][kTQ];
  __shared__ double As[2][kTQ * kLdaA];
  // Workgroup b runs on XCD b % 8.  XCD x takes the column tiles x, x + 8, ...
  // and, for each, all form blocks in consecutive dispatch slots, so a tile
  // is fetched into that XCD's L2 once and reused by every form block; the
  // coefficient rows (Q x NC) stay L2-resident.  The grid holds ntx rounded
  // up to a multiple of 8 tiles; the padding tiles exit at once.
  const int b = blockIdx.x, local = b >> 3;
  const int tile = (local / nqy) * 8 + (b & 7), qblk = local % nqy;
  if (tile >= ntx) return;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, i = lane & 15;
  const int64_t NC2 = (int64_t)NC * NC;
  const int64_t j = (int64_t)tile * kTJ + wave * 16 + i;
  const bool jok = j < NC2;
  const int64_t jc = jok ? j : NC2 - 1;
  const int q0 = qblk * kTQ;
  const int K2 = 2 * NC;
  const double* mcol = mu + (int64_t)K2 * jc;
  // this thread's share of a stage of form rows: kTQ rows x 16 double2
  // words, four words per thread (rows beyond Q repeat row Q - 1)
  constexpr int kAW = kTQ * (kStage / 2) / 256;
  const double* asrc[kAW];
  int adst[kAW], awi[kAW];
#pragma unroll
  for (int w = 0; w < kAW; ++w) {
    const int word = tid + 256 * w;
    const int row = word / (kStage / 2), wi = word % (kStage / 2);
    const int q = q0 + row;
    asrc[w] = A + (int64_t)K2 * (q < Q ? q : Q - 1);
    adst[w] = row * kLdaA + 2 * wi;
    awi[w] = 2 * wi;
  }
  // reals beyond 2NC (the last stage) read as zero; the clamped addresses
  // stay inside the column / row
  auto load_b = [&](int st, double* bb) {
    const int kb = st * kStage + 8 * g;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = kb + 2 * t;
      const bool ok = k < K2;
      const double2 v = *reinterpret_cast<const double2*>(mcol + (ok ? k : 0));
      bb[2 * t] = ok ? v.x : 0.0;
      bb[2 * t + 1] = ok ? v.y : 0.0;
    }
  };
  auto load_a = [&](int st, double2* ra) {
#pragma unroll
    for (int w = 0; w < kAW; ++w) {
      const int k = st * kStage + awi[w];
      const bool ok = k < K2;
      const double2 v = *reinterpret_cast<const double2*>(asrc[w] + (ok ? k : 0));
      ra[w] = ok ? v : make_double2(0.0, 0.0);
    }
  };
  auto store_a = [&](int buf, const double2* ra) {
#pragma unroll
    for (int w = 0; w < kAW; ++w) *reinterpret_cast<double2*>(&As[buf][adst[w]]) = ra[w];
  };

  d4 are[kRB], aim[kRB];
#pragma unroll
  for (int rb = 0; rb < kRB; ++rb) are[rb] = aim[rb] = d4{0.0, 0.0, 0.0, 0.0};
  const int nst = (K2 + kStage - 1) / kStage;
  double bc[8];
  double2 ra[kAW];
  load_b(0, bc);
  load_a(0, ra);
  store_a(0, ra);
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const int buf = st & 1;
    const bool more = st + 1 < nst;  // uniform
    double bn[8];
    if (more) {  // next stage in flight during the MFMAs
      load_b(st + 1, bn);
      load_a(st + 1, ra);
    }
    // all four row blocks' fragments first, then the MFMAs with the eight
    // accumulators interleaved (each accumulator still takes t in order), so
    // a dependent MFMA never waits on the one just issued
    double ac[kRB][8];
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) {
      const double* pa = &As[buf][(rb * 16 + i) * kLdaA + 8 * g];
#pragma unroll
      for (int t = 0; t < 8; t += 2) {
        const double2 v = *reinterpret_cast<const double2*>(pa + t);
        ac[rb][t] = v.x;
        ac[rb][t + 1] = v.y;
      }
    }
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb) {
        const double ar = (t & 1) ? -ac[rb][t] : ac[rb][t];
        const double ai = (t & 1) ? ac[rb][t - 1] : ac[rb][t + 1];
        are[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bc[t], are[rb], 0, 0, 0);
        aim[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bc[t], aim[rb], 0, 0, 0);
      }
    if (more) {
      store_a(buf ^ 1, ra);
#pragma unroll
      for (int t = 0; t < 8; ++t) bc[t] = bn[t];
    }
    __syncthreads();
  }

  // epilogue: accumulator element v of lane (g, i) is X[row g + 4v][column i]
  const int m = (int)(jc % NC), p = (int)(jc / NC);
#pragma unroll
  for (int rb = 0; rb < kRB; ++rb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = rb * 16 + g + 4 * v;
      const int q = q0 + r;
      double sx = 0.0, sy = 0.0;
      if (jok && q < Q) {
        const double2 bq = B[(int64_t)q * NC + m], cq = C[(int64_t)q * NC + p];
        const double wr = bq.x * cq.x - bq.y * cq.y, wi = bq.x * cq.y + bq.y * cq.x;
        const double xr = are[rb][v], xi = aim[rb][v];
        sx = xr * wr - xi * wi;
        sy = xr * wi + xi * wr;
      }
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) {
        sx += __shfl_xor(sx, off, 16);
        sy += __shfl_xor(sy, off, 16);
      }
      if (i == 0) red[wave][r] = make_double2(sx, sy);
    }
  __syncthreads();
  if (threadIdx.x < kTQ) {
    const int q = q0 + threadIdx.x;
    if (q < Q) {
      double sx = 0.0, sy = 0.0;
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        sx += red[w][threadIdx.x].x;
        sy += red[w][threadIdx.x].y;
      }
      part[(int64_t)tile * Q + q] = make_double2(sx, sy);
    }
  }
}

// out[q] = sum over column tiles, in tile order
__global__ void k_trilinear_reduce(const double2* __restrict__ part, int ntiles, int Q,
                                   double2* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Q) return;
  double sx = 0.0, sy = 0.0;
  // loads in batches of 16 (in flight together), added in tile order
  constexpr int U = 16;
  int t = 0;
  for (; t + U <= ntiles; t += U) {
    double2 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = part[(int64_t)(t + u) * Q + q];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      sx += v[u].x;
      sy += v[u].y;
    }
  }
  for (; t < ntiles; ++t) {
    const double2 v = part[(int64_t)t * Q + q];
    sx += v.x;
    sy += v.y;
  }
  out[q] = make_double2(sx, sy);
}

// mu~[n,m,p] = ((mu[n,m,p] k[n]) k[m]) k[p]: the damping kernel and h_n
// along each axis, applied axis by axis in the order of
// muND_apply_kernel_and_h (src/kernels/kernels.jl:92-103)
__global__ void k_weight3(double2* __restrict__ mu, const double* __restrict__ k, int NC) {
  const int64_t total = (int64_t)NC * NC * NC;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int n = (int)(idx % NC);
    const int m = (int)((idx / NC) % NC);
    const int p = (int)(idx / ((int64_t)NC * NC));
    double2 v = mu[idx];
    v.x *= k[n];
    v.y *= k[n];
    v.x *= k[m];
    v.y *= k[m];
    v.x *= k[p];
    v.y *= k[p];
    mu[idx] = v;
  }
}

// Coefficient rows from per-row recurrence descriptors (5 doubles each):
//   kind 0, c:        delta[0] = 1/pi, delta[1] = c/pi,
//                     delta[n] = 2c delta[n-1] - delta[n-2]
//                     (chebyshev_delta_theta!, spectral_coefficients.jl:22-33)
//   kind 1, g0, q:    g[0] = g0, g[n] = g[n-1] q
//                     (green_coefficients!, spectral_coefficients.jl:84-86)
//   kind 2:           zeros (the delta-annulus exclusion, cpge.jl:30-36)
// One thread per row, n in order, so each row follows the reference's
// sequential recurrence.
__global__ void k_expand_rows(const double* __restrict__ desc, int64_t rows, int NC,
                              double2* __restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const double* d = desc + 5 * r;
  double2* o = out + r * NC;
  const int kind = (int)d[0];
  if (kind == 0) {
    const double c = d[1];
    const double inv_pi = 0.31830988618379067154;
    double dm2 = inv_pi, dm1 = c / 3.14159265358979323846;
    o[0] = make_double2(dm2, 0.0);
    if (NC > 1) o[1] = make_double2(dm1, 0.0);
    for (int n = 2; n < NC; ++n) {
      const double dn = 2.0 * c * dm1 - dm2;
      o[n] = make_double2(dn, 0.0);
      dm2 = dm1;
      dm1 = dn;
    }
  } else if (kind == 1) {
    double gr = d[1], gi = d[2];
    const double qr = d[3], qi = d[4];
    o[0] = make_double2(gr, gi);
    for (int n = 1; n < NC; ++n) {
      const double nr = gr * qr - gi * qi, ni = gr * qi + gi * qr;
      gr = nr;
      gi = ni;
      o[n] = make_double2(gr, gi);
    }
  } else {
    for (int n = 0; n < NC; ++n) o[n] = make_double2(0.0, 0.0);
  }
}

// deterministic unit-modulus coefficient rows for the timing helper
__global__ void k_fill_phase(double2* __restrict__ v, int64_t n, double step) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s, c;
  sincos(step * (double)(i % 4099), &s, &c);
  v[i] = make_double2(c, s);
}

}  // namespace kpm

struct kpm_table3 {
  kpm_ctx* ctx = nullptr;
  int64_t NC = 0;
  double2* mu = nullptr;  // NC^3, n fastest
  kpm::Buffer work;       // coefficient rows, partials and results of a call
};

using namespace kpm;

namespace {

// device coefficient rows da (real view), db, dc -> host out (Q complex)
int run_forms(kpm_table3* t, int64_t Q, const double* da, const double2* db, const double2* dc,
              double2* dpart, double2* dout, double* out) {
  kpm_ctx* ctx = t->ctx;
  const int64_t NC = t->NC, NC2 = NC * NC;
  const int64_t ntx = (NC2 + kTJ - 1) / kTJ, nqy = (Q + kTQ - 1) / kTQ;
  const unsigned grid = (unsigned)((ntx + 7) / 8 * 8 * nqy);
  hipLaunchKernelGGL(k_trilinear, dim3(grid), dim3(256), 0, ctx->stream, (const double*)t->mu, (int)NC,
                     da, db, dc, (int)Q, (int)ntx, (int)nqy, dpart);
  KPM_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_trilinear_reduce, dim3((unsigned)((Q + 255) / 256)), dim3(256), 0, ctx->stream,
                     (const double2*)dpart, (int)ntx, (int)Q, dout);
  KPM_HIP_TRY(hipGetLastError());
  KPM_HIP_TRY(hipMemcpyAsync(out, dout, (size_t)Q * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
  KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return KPM_OK;
}

}  // namespace

extern "C" {

int kpm_table3_create(kpm_ctx* ctx, int64_t NC, const double* mu_tilde, kpm_table3** out) {
  if (!ctx || !mu_tilde || !out) return fail(KPM_EINVAL, "null argument");
  *out = nullptr;
  if (NC < 1 || NC > 2048) return fail(KPM_EINVAL, "NC must lie in [1, 2048]");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const size_t bytes = (size_t)(NC * NC * NC) * sizeof(double2);
  kpm_table3* t = new kpm_table3;
  t->ctx = ctx;
  t->NC = NC;
  if (hipMalloc(&t->mu, bytes) != hipSuccess) {
    delete t;
    (void)hipGetLastError();
    return fail(KPM_ENOMEM, "device allocation of the moment table failed");
  }
  if (hipMemcpyAsync(t->mu, mu_tilde, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    (void)hipFree(t->mu);
    delete t;
    return fail(KPM_EHIP, "moment table upload failed");
  }
  *out = t;
  return KPM_OK;
}

int kpm_table3_destroy(kpm_table3* t) {
  if (!t) return KPM_OK;
  (void)hipSetDevice(t->ctx->device);
  (void)hipStreamSynchronize(t->ctx->stream);
  if (t->mu) (void)hipFree(t->mu);
  if (t->work.ptr) (void)hipFree(t->work.ptr);
  delete t;
  return KPM_OK;
}

int kpm_trilinear_forms(kpm_table3* t, int64_t Q, const double* a, const double* b, const double* c,
                        double* out) {
  if (!t || (Q > 0 && (!a || !b || !c || !out))) return fail(KPM_EINVAL, "null argument");
  if (Q < 0 || Q > (1 << 20)) return fail(KPM_EINVAL, "Q must lie in [0, 2^20]");
  if (Q == 0) return KPM_OK;
  kpm_ctx* ctx = t->ctx;
  const int64_t NC = t->NC, NC2 = NC * NC;
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const int64_t ntx = (NC2 + kTJ - 1) / kTJ, nqy = (Q + kTQ - 1) / kTQ;
  if ((ntx + 7) / 8 * 8 * nqy > INT32_MAX) return fail(KPM_EINVAL, "too many forms in one call");
  const size_t bcoef = (size_t)(Q * NC) * sizeof(double2), bpart = (size_t)(ntx * Q) * sizeof(double2),
               bout = (size_t)Q * sizeof(double2);
  KPM_TRY(ensure(ctx, t->work, 3 * bcoef + bpart + bout));
  char* base = (char*)t->work.ptr;
  double* da = (double*)base;
  double2* db = (double2*)(base + bcoef);
  double2* dc = (double2*)(base + 2 * bcoef);
  double2* dpart = (double2*)(base + 3 * bcoef);
  double2* dout = (double2*)(base + 3 * bcoef + bpart);
  KPM_HIP_TRY(hipMemcpyAsync(da, a, bcoef, hipMemcpyHostToDevice, ctx->stream));
  KPM_HIP_TRY(hipMemcpyAsync(db, b, bcoef, hipMemcpyHostToDevice, ctx->stream));
  KPM_HIP_TRY(hipMemcpyAsync(dc, c, bcoef, hipMemcpyHostToDevice, ctx->stream));
  return run_forms(t, Q, da, db, dc, dpart, dout, out);
}

int kpm_trilinear_forms_rec(kpm_table3* t, int64_t Q, const double* desc_a, const double* desc_b,
                            const double* desc_c, double* out) {
  if (!t || (Q > 0 && (!desc_a || !desc_b || !desc_c || !out))) return fail(KPM_EINVAL, "null argument");
  if (Q < 0 || Q > (1 << 20)) return fail(KPM_EINVAL, "Q must lie in [0, 2^20]");
  if (Q == 0) return KPM_OK;
  for (const double* d : {desc_a, desc_b, desc_c})
    for (int64_t r = 0; r < Q; ++r) {
      const double k = d[5 * r];
      if (!(k == 0.0 || k == 1.0 || k == 2.0)) return fail(KPM_EINVAL, "row kind must be 0, 1 or 2");
    }
  kpm_ctx* ctx = t->ctx;
  const int64_t NC = t->NC, NC2 = NC * NC;
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const int64_t ntx = (NC2 + kTJ - 1) / kTJ, nqy = (Q + kTQ - 1) / kTQ;
  if ((ntx + 7) / 8 * 8 * nqy > INT32_MAX) return fail(KPM_EINVAL, "too many forms in one call");
  const size_t bcoef = (size_t)(Q * NC) * sizeof(double2), bpart = (size_t)(ntx * Q) * sizeof(double2),
               bout = (size_t)Q * sizeof(double2), bdesc = (size_t)(15 * Q) * sizeof(double);
  KPM_TRY(ensure(ctx, t->work, 3 * bcoef + bpart + bout + bdesc));
  char* base = (char*)t->work.ptr;
  double2* rows = (double2*)base;  // a | b | c, each Q x NC
  double2* dpart = (double2*)(base + 3 * bcoef);
  double2* dout = (double2*)(base + 3 * bcoef + bpart);
  double* ddesc = (double*)(base + 3 * bcoef + bpart + bout);
  const size_t b1 = (size_t)(5 * Q) * sizeof(double);
  KPM_HIP_TRY(hipMemcpyAsync(ddesc, desc_a, b1, hipMemcpyHostToDevice, ctx->stream));
  KPM_HIP_TRY(hipMemcpyAsync(ddesc + 5 * Q, desc_b, b1, hipMemcpyHostToDevice, ctx->stream));
  KPM_HIP_TRY(hipMemcpyAsync(ddesc + 10 * Q, desc_c, b1, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_expand_rows, dim3((unsigned)((3 * Q + 63) / 64)), dim3(64), 0, ctx->stream,
                     (const double*)ddesc, 3 * Q, (int)NC, rows);
  KPM_HIP_TRY(hipGetLastError());
  return run_forms(t, Q, (const double*)rows, rows + Q * NC, rows + 2 * Q * NC, dpart, dout, out);
}

int kpm_table3_create_weighted(kpm_ctx* ctx, int64_t NC, const double* mu, const double* kvec,
                               kpm_table3** out) {
  if (!kvec) return fail(KPM_EINVAL, "null argument");
  KPM_TRY(kpm_table3_create(ctx, NC, mu, out));
  kpm_table3* t = *out;
  double* dk = nullptr;
  if (hipMalloc(&dk, (size_t)NC * sizeof(double)) != hipSuccess) {
    kpm_table3_destroy(t);
    *out = nullptr;
    (void)hipGetLastError();
    return fail(KPM_ENOMEM, "device allocation failed");
  }
  const int64_t total = NC * NC * NC;
  const unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
  int rc = KPM_OK;
  if (hipMemcpyAsync(dk, kvec, (size_t)NC * sizeof(double), hipMemcpyHostToDevice, ctx->stream) !=
      hipSuccess)
    rc = fail(KPM_EHIP, "kernel vector upload failed");
  if (rc == KPM_OK) {
    hipLaunchKernelGGL(k_weight3, dim3(g), dim3(256), 0, ctx->stream, t->mu, (const double*)dk, (int)NC);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = fail(KPM_EHIP, "kernel weighting failed");
  }
  (void)hipFree(dk);
  if (rc != KPM_OK) {
    kpm_table3_destroy(t);
    *out = nullptr;
  }
  return rc;
}

int kpm_bench_trilinear(kpm_table3* t, int64_t Q, int64_t reps, float* ms) {
  if (!t || !ms || Q < 1 || Q > (1 << 20) || reps < 1) return fail(KPM_EINVAL, "bad arguments");
  kpm_ctx* ctx = t->ctx;
  const int64_t NC = t->NC, NC2 = NC * NC;
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const int64_t ntx = (NC2 + kTJ - 1) / kTJ, nqy = (Q + kTQ - 1) / kTQ;
  if ((ntx + 7) / 8 * 8 * nqy > INT32_MAX) return fail(KPM_EINVAL, "too many forms in one call");
  const size_t bcoef = (size_t)(Q * NC) * sizeof(double2), bpart = (size_t)(ntx * Q) * sizeof(double2),
               bout = (size_t)Q * sizeof(double2);
  KPM_TRY(ensure(ctx, t->work, 3 * bcoef + bpart + bout));
  char* base = (char*)t->work.ptr;
  double2* dc = (double2*)base;
  double2* dpart = (double2*)(base + 3 * bcoef);
  double2* dout = (double2*)(base + 3 * bcoef + bpart);
  const int64_t n = 3 * Q * NC;
  hipLaunchKernelGGL(k_fill_phase, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, dc, n,
                     0.618);
  KPM_HIP_TRY(hipGetLastError());
  auto launch = [&]() {
    hipLaunchKernelGGL(k_trilinear, dim3((unsigned)((ntx + 7) / 8 * 8 * nqy)), dim3(256), 0,
                       ctx->stream, (const double*)t->mu, (int)NC, (const double*)dc,
                       (const double2*)(dc + Q * NC), (const double2*)(dc + 2 * Q * NC), (int)Q,
                       (int)ntx, (int)nqy, dpart);
    hipLaunchKernelGGL(k_trilinear_reduce, dim3((unsigned)((Q + 255) / 256)), dim3(256), 0,
                       ctx->stream, (const double2*)dpart, (int)ntx, (int)Q, dout);
  };
  launch();  // warm-up
  KPM_HIP_TRY(hipEventRecord(ctx->events[14], ctx->stream));
  for (int64_t r = 0; r < reps; ++r) launch();
  KPM_HIP_TRY(hipEventRecord(ctx->events[15], ctx->stream));
  KPM_HIP_TRY(hipGetLastError());
  KPM_HIP_TRY(hipEventSynchronize(ctx->events[15]));
  KPM_HIP_TRY(hipEventElapsedTime(ms, ctx->events[14], ctx->events[15]));
  return KPM_OK;
}

}  // extern "C"
