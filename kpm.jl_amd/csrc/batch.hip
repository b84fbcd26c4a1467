// batch.hip -- many independent realisations in one launch per recurrence
// step (the disorder-ensemble workload).
//
// The reference runs one realisation per process and calls kpm_1d! on each
// (scripts/anderson_seebeck_ensemble.jl:1-20; src/moment.jl:420-492).  Here
// all realisations are packed into one block-diagonal CSR whose diagonal
// blocks are padded with empty rows to a multiple of kPad rows, so no row
// tile straddles two realisations.  Each step is one segmented fused pass
// (k_step<..., SEG = true>) plus one per-realisation fixed-order reduction;
// the NC/2 steps can be recorded once into a HIP graph and replayed.
#include <algorithm>
#include <tuple>
#include <vector>

#include "kpm_internal.h"

namespace kpm {

constexpr int64_t kPad = 512;  // >= the largest row tile (4 waves x 64 rows x 2)

struct BatchSched {
  int64_t TR = 0, ntiles = 0;
  int G = 0, nseg = 0;
  Buffer tile_lo, treal, seg0, sfirst, scount;
};

}  // namespace kpm

struct kpm_batch {
  kpm_ctx* ctx = nullptr;
  int nreal = 0;
  std::vector<int64_t> off;   // unpadded row offsets (nreal+1)
  std::vector<int64_t> poff;  // padded row offsets (nreal+1)
  kpm_csr* A = nullptr;       // padded block-diagonal operator
  kpm::BatchSched sched;
  kpm::Buffer partials;
  kpm::Buffer probes;  // resident padded probe block (kpm_batch_upload)
  int64_t probe_NR = 0;
  int probe_prec = -1;
  // recorded step sequence, keyed on the shape and on every device address
  // the captured nodes use; dropped whenever one of those buffers may move
  hipGraphExec_t graph = nullptr;
  std::tuple<int64_t, int64_t, int, int64_t, int, const void*, const void*, const void*, const void*,
             const void*, const void*, const void*, const void*, const void*, const void*>
      gkey{};
  int64_t graph_builds = 0;  // instantiations so far (kpm_batch_graph_builds)
};

namespace kpm {

// Per-realisation moments of one step from the segment partials of the
// segmented pass.  One workgroup per realisation: its partials are contiguous
// ([segment][d][NRp] double2 over segments s0 .. s0 + ns - 1, W = nd * NRp
// columns), so each column gets SL = 256 / W segment lanes (columns taken 256
// at a time when W is larger).  Lane j of a column adds segments j, j + SL, ...
// in order and the column's thread then adds the SL lane sums in lane order.
// The order depends on W only, so eager runs and graph replays give the same
// bits.  (One thread per output, adding every segment itself, took 15.4 us per
// step on the 256-realisation cubic ensemble: a handful of workgroups, each
// thread waiting on its loads in turn.)
constexpr int kCombineThreads = 256;

__global__ void __launch_bounds__(kCombineThreads)
    k_batch_combine(const double2* __restrict__ partials, const int* __restrict__ sfirst,
                    const int* __restrict__ scount, int nd, int NRp, int NR, int64_t NC,
                    double2* __restrict__ mu, Combine cb) {
  __shared__ double2 lane_sum[kCombineThreads];
  const int r = blockIdx.x, t = threadIdx.x;
  const int W = nd * NRp;
  const int CW = min(W, kCombineThreads);  // columns per round
  const int SL = kCombineThreads / CW;     // segment lanes per column
  const int j = t / CW;
  const int s0 = sfirst[r], ns = scount[r];
  const double2* P = partials + (size_t)s0 * W;
  double2* m = mu + (size_t)r * NR * NC;
  for (int c0 = 0; c0 < W; c0 += CW) {
    const int c = c0 + t % CW;
    double sr = 0.0, si = 0.0;
    if (j < SL && c < W) {
#pragma unroll 4
      for (int s = j; s < ns; s += SL) {
        const double2 v = P[(size_t)s * W + c];
        sr += v.x;
        si += v.y;
      }
    }
    lane_sum[t] = make_double2(sr, si);
    __syncthreads();
    if (t < CW && c < W) {
      double tr = 0.0, ti = 0.0;
      for (int q = 0; q < SL; ++q) {
        const double2 a = lane_sum[q * CW + t];
        tr += a.x;
        ti += a.y;
      }
      const int d = c / NRp, col = c % NRp;
      if (col < NR) {
        double rr = cb.alpha[d] * tr, ri = cb.alpha[d] * ti;
        if (cb.bcol[d] >= 0) {
          const double2 bv = m[col + (size_t)NR * cb.bcol[d]];
          rr = fma(cb.bscale[d], bv.x, rr);
          ri = fma(cb.bscale[d], bv.y, ri);
        }
        m[col + (size_t)NR * cb.tcol[d]] = make_double2(rr, ri);
      }
    }
    __syncthreads();
  }
}

static int upload_ints(kpm_ctx* ctx, Buffer& b, const std::vector<int>& v) {
  KPM_TRY(ensure(ctx, b, std::max<size_t>(v.size(), 1) * sizeof(int)));
  KPM_HIP_TRY(hipMemcpy(b.ptr, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
  return KPM_OK;
}

static void drop_graph(kpm_batch* B);

// tile ranges, tile realisations and segment slots for row tiles of TR rows
// on G workgroups.  A rebuild may reallocate the schedule buffers that a
// recorded graph reads, so it drops the graph first.
static int build_sched(kpm_batch* B, int64_t TR, int G) {
  BatchSched& S = B->sched;
  if (S.TR == TR && S.G == G) return KPM_OK;
  drop_graph(B);
  kpm_ctx* ctx = B->ctx;
  const int64_t ntiles = B->poff[B->nreal] / TR;
  std::vector<int> treal(ntiles), tile_lo(G + 1), seg0(G), sfirst(B->nreal, 0), scount(B->nreal, 0);
  for (int r = 0; r < B->nreal; ++r)
    for (int64_t t = B->poff[r] / TR; t < B->poff[r + 1] / TR; ++t) treal[t] = r;
  for (int b = 0; b <= G; ++b) tile_lo[b] = (int)((int64_t)b * ntiles / G);
  int nseg = 0;
  std::vector<int> segreal;
  for (int b = 0; b < G; ++b) {
    seg0[b] = nseg;
    for (int t = tile_lo[b]; t < tile_lo[b + 1]; ++t)
      if (t == tile_lo[b] || treal[t] != treal[t - 1]) {
        segreal.push_back(treal[t]);
        ++nseg;
      }
  }
  // segments of one realisation are consecutive (blocks and tiles are ordered)
  for (int s = nseg - 1; s >= 0; --s) sfirst[segreal[s]] = s;
  for (int s = 0; s < nseg; ++s) scount[segreal[s]]++;
  KPM_TRY(upload_ints(ctx, S.treal, treal));
  KPM_TRY(upload_ints(ctx, S.tile_lo, tile_lo));
  KPM_TRY(upload_ints(ctx, S.seg0, seg0));
  KPM_TRY(upload_ints(ctx, S.sfirst, sfirst));
  KPM_TRY(upload_ints(ctx, S.scount, scount));
  S.TR = TR;
  S.G = G;
  S.ntiles = ntiles;
  S.nseg = nseg;
  return KPM_OK;
}

static void drop_graph(kpm_batch* B) {
  if (B->graph) {
    (void)hipStreamSynchronize(B->ctx->stream);
    (void)hipGraphExecDestroy(B->graph);
    B->graph = nullptr;
  }
}

static void free_buf(Buffer& b) {
  if (b.ptr) (void)hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
}

// the NC/2 doubling steps of every realisation (kpm_1d!, src/moment.jl:440-487)
static int batch_steps(kpm_ctx* ctx, kpm_batch* B, int64_t NC, const StepGeom& g, kpm_block* w0,
                       kpm_block* w1, void* mu) {
  const BatchSched& S = B->sched;
  const int nd = 2;
  const int64_t NR = w0->NR;
  auto combine = [&](const Combine& c) -> int {
    hipLaunchKernelGGL(k_batch_combine, dim3((unsigned)B->nreal), dim3(kCombineThreads), 0, ctx->stream,
                       (const double2*)B->partials.ptr, (const int*)S.sfirst.ptr,
                       (const int*)S.scount.ptr, nd, (int)w0->NRp, (int)NR, NC, (double2*)mu, c);
    KPM_HIP_TRY(hipGetLastError());
    return KPM_OK;
  };
  kpm_block* v[2] = {w0, w1};
  StepLaunch s0{B->A, v[0], nullptr, v[1], nullptr, 1.0, 0.0, MODE_DBL_SEG};
  KPM_TRY(launch_step_seg(ctx, s0, g, S.G, S.ntiles, (const int*)S.tile_lo.ptr,
                          (const int*)S.treal.ptr, (const int*)S.seg0.ptr,
                          (double2*)B->partials.ptr));
  KPM_TRY(combine(Combine{2, {0, 1}, {1.0, 1.0}, {-1, -1}, {0.0, 0.0}}));
  int ip = 1, ipp = 0;
  for (int64_t n = 2; n <= NC / 2; ++n) {
    StepLaunch s{B->A, v[ip], v[ipp], v[ipp], nullptr, 2.0, -1.0, MODE_DBL_SEG};
    KPM_TRY(launch_step_seg(ctx, s, g, S.G, S.ntiles, (const int*)S.tile_lo.ptr,
                            (const int*)S.treal.ptr, (const int*)S.seg0.ptr,
                            (double2*)B->partials.ptr));
    KPM_TRY(combine(Combine{2, {(int)(2 * n - 2), (int)(2 * n - 1)}, {2.0, 2.0}, {0, 1}, {-1.0, -1.0}}));
    ip = 1 - ip;
    ipp = 1 - ipp;
  }
  return KPM_OK;
}

}  // namespace kpm

using namespace kpm;

extern "C" {

int kpm_batch_create(kpm_ctx* ctx, int nreal, const int64_t* row_offsets, const int64_t* colptr,
                     const int64_t* rowval, const double* nzval, int index_base, kpm_batch** out) {
  if (!ctx || !row_offsets || !out || nreal <= 0) return fail(KPM_EINVAL, "bad arguments");
  if (row_offsets[0] != 0) return fail(KPM_EINVAL, "row_offsets must start at 0");
  for (int r = 0; r < nreal; ++r)
    if (row_offsets[r + 1] <= row_offsets[r]) return fail(KPM_EINVAL, "empty or unordered realisation");
  const int64_t n = row_offsets[nreal];
  KPM_TRY(check_csc(n, colptr, rowval, index_base));
  std::vector<int64_t> poff(nreal + 1, 0);
  for (int r = 0; r < nreal; ++r)
    poff[r + 1] = poff[r] + (row_offsets[r + 1] - row_offsets[r] + kPad - 1) / kPad * kPad;
  const int64_t P = poff[nreal];
  if (P >= INT32_MAX) return fail(KPM_EINVAL, "packed operator too large for int32 indices");
  // realisation of each packed row, and the padded index map
  std::vector<int> real_of(n);
  for (int r = 0; r < nreal; ++r)
    for (int64_t i = row_offsets[r]; i < row_offsets[r + 1]; ++i) real_of[i] = r;
  auto pmap = [&](int64_t i) { const int r = real_of[i]; return poff[r] + (i - row_offsets[r]); };
  const int64_t nnz = colptr[n] - index_base;
  std::vector<int64_t> rp(P + 1, 0), ci(std::max<int64_t>(nnz, 1));
  std::vector<double> v(2 * std::max<int64_t>(nnz, 1));
  for (int64_t j = 0; j < n; ++j)
    for (int64_t k = colptr[j] - index_base; k < colptr[j + 1] - index_base; ++k) {
      const int64_t i = rowval[k] - index_base;
      if (real_of[i] != real_of[j])
        return fail(KPM_EINVAL, "operator is not block diagonal over the given row offsets");
      rp[pmap(i) + 1]++;
    }
  for (int64_t i = 0; i < P; ++i) rp[i + 1] += rp[i];
  std::vector<int64_t> next(rp.begin(), rp.end() - 1);
  for (int64_t j = 0; j < n; ++j)
    for (int64_t k = colptr[j] - index_base; k < colptr[j + 1] - index_base; ++k) {
      const int64_t i = pmap(rowval[k] - index_base);
      const int64_t dst = next[i]++;
      ci[dst] = pmap(j);
      v[2 * dst] = nzval[2 * k];
      v[2 * dst + 1] = nzval[2 * k + 1];
    }
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  kpm_batch* B = new kpm_batch();
  B->ctx = ctx;
  B->nreal = nreal;
  B->off.assign(row_offsets, row_offsets + nreal + 1);
  B->poff = poff;
  int rc = upload_csr(ctx, P, rp, ci, v, &B->A);
  if (rc != KPM_OK) {
    delete B;
    return rc;
  }
  *out = B;
  return KPM_OK;
}

int kpm_batch_destroy(kpm_batch* B) {
  if (!B) return KPM_OK;
  (void)hipSetDevice(B->ctx->device);
  drop_graph(B);
  (void)hipStreamSynchronize(B->ctx->stream);
  free_buf(B->sched.tile_lo);
  free_buf(B->sched.treal);
  free_buf(B->sched.seg0);
  free_buf(B->sched.sfirst);
  free_buf(B->sched.scount);
  free_buf(B->partials);
  kpm_csr_destroy(B->A);
  delete B;
  return KPM_OK;
}

int kpm_batch_upload(kpm_ctx* ctx, kpm_batch* B, int64_t NR, const double* psi, int vec_prec) {
  if (!ctx || !B || !psi) return fail(KPM_EINVAL, "null argument");
  if (NR <= 0) return fail(KPM_EINVAL, "NR must be positive");
  if (vec_prec != KPM_VEC_F64 && vec_prec != KPM_VEC_F32) return fail(KPM_EINVAL, "bad vec_prec");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const int64_t P = B->poff[B->nreal], n = B->off[B->nreal];
  // the recorded graph copies from the probe block; it may move here
  const void* old_probes = B->probes.ptr;
  kpm_block pb;
  KPM_TRY(work_block(ctx, B->probes, P, NR, vec_prec, &pb));
  if (B->probes.ptr != old_probes) drop_graph(B);
  // stacked (n x NR, column-major) -> padded rows, then the device layout
  std::vector<double> padded((size_t)P * NR * 2, 0.0);
  for (int64_t c = 0; c < NR; ++c)
    for (int r = 0; r < B->nreal; ++r) {
      const int64_t len = B->off[r + 1] - B->off[r];
      std::copy(psi + 2 * (c * n + B->off[r]), psi + 2 * (c * n + B->off[r] + len),
                padded.data() + 2 * (c * P + B->poff[r]));
    }
  KPM_TRY(ensure(ctx, ctx->staging, padded.size() * sizeof(double)));
  KPM_HIP_TRY(hipMemcpyAsync(ctx->staging.ptr, padded.data(), padded.size() * sizeof(double),
                             hipMemcpyHostToDevice, ctx->stream));
  KPM_TRY(launch_upload(ctx, ctx->staging.ptr, &pb));
  KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  B->probe_NR = NR;
  B->probe_prec = vec_prec;
  return KPM_OK;
}

int kpm_batch_run(kpm_ctx* ctx, kpm_batch* B, int64_t NC, double* mu, int use_graph) {
  if (!ctx || !B || !mu) return fail(KPM_EINVAL, "null argument");
  if (B->probe_prec < 0) return fail(KPM_EINVAL, "no probes uploaded (kpm_batch_upload)");
  if (NC % 2 != 0 || NC < 2) return fail(KPM_EINVAL, "Invalid NC: NC should be even.");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const int64_t NR = B->probe_NR;
  const int vec_prec = B->probe_prec;
  const int64_t P = B->poff[B->nreal];
  const StepConfig cfg = tuned_config(ctx, B->A, NR, vec_prec, MODE_DBL);
  const StepGeom g = step_geom(B->A, NR, vec_prec, cfg);
  const int64_t ntiles = P / g.TR;
  const int G = cfg.grid > 0 ? cfg.grid : default_grid(ctx, ntiles);
  KPM_TRY(build_sched(B, g.TR, G));
  const int64_t NRp = padded_nr(NR, vec_prec);
  KPM_TRY(ensure(ctx, B->partials, (size_t)std::max(B->sched.nseg, 1) * 2 * NRp * sizeof(double2)));
  kpm_block w0, w1, pb;
  KPM_TRY(work_block(ctx, ctx->work[0], P, NR, vec_prec, &w0));
  KPM_TRY(work_block(ctx, ctx->work[1], P, NR, vec_prec, &w1));
  KPM_TRY(work_block(ctx, B->probes, P, NR, vec_prec, &pb));
  KPM_TRY(ensure(ctx, ctx->mu, (size_t)B->nreal * NR * NC * sizeof(double2)));
  auto run = [&]() -> int {
    KPM_HIP_TRY(hipMemcpyAsync(w0.d, pb.d, (size_t)(P * NRp) * elem_bytes(vec_prec),
                               hipMemcpyDeviceToDevice, ctx->stream));
    return batch_steps(ctx, B, NC, g, &w0, &w1, ctx->mu.ptr);
  };
  if (use_graph) {
    const BatchSched& S = B->sched;
    auto key = std::make_tuple(NC, NR, vec_prec, g.TR, G, (const void*)w0.d, (const void*)w1.d,
                               (const void*)ctx->mu.ptr, (const void*)B->partials.ptr,
                               (const void*)S.treal.ptr, (const void*)pb.d, (const void*)S.tile_lo.ptr,
                               (const void*)S.seg0.ptr, (const void*)S.sfirst.ptr,
                               (const void*)S.scount.ptr);
    if (!B->graph || B->gkey != key) {
      drop_graph(B);
      hipGraph_t graph;
      KPM_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      int rc = run();
      hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
      if (rc != KPM_OK) return rc;
      if (e != hipSuccess) return fail(KPM_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
      e = hipGraphInstantiate(&B->graph, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return fail(KPM_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
      B->gkey = key;
      ++B->graph_builds;
    }
    KPM_HIP_TRY(hipGraphLaunch(B->graph, ctx->stream));
  } else {
    KPM_TRY(run());
  }
  KPM_HIP_TRY(hipMemcpyAsync(mu, ctx->mu.ptr, (size_t)B->nreal * NR * NC * sizeof(double2),
                             hipMemcpyDeviceToHost, ctx->stream));
  KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return KPM_OK;
}

int kpm_batch_1d(kpm_ctx* ctx, kpm_batch* B, int64_t NC, int64_t NR, const double* psi, double* mu,
                 int vec_prec, int use_graph) {
  if (NC % 2 != 0 || NC < 2) return fail(KPM_EINVAL, "Invalid NC: NC should be even.");
  KPM_TRY(kpm_batch_upload(ctx, B, NR, psi, vec_prec));
  return kpm_batch_run(ctx, B, NC, mu, use_graph);
}

// Time every candidate geometry of the segmented pass on the uploaded
// probes (work blocks are overwritten) and record the fastest for this
// shape; returns it.
int kpm_batch_autotune(kpm_ctx* ctx, kpm_batch* B, int reps, int* best_unroll, int* best_grid,
                       float* best_ms_per_step) {
  if (!ctx || !B || reps < 1) return fail(KPM_EINVAL, "bad arguments");
  if (B->probe_prec < 0) return fail(KPM_EINVAL, "no probes uploaded (kpm_batch_upload)");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const int64_t NR = B->probe_NR;
  const int prec = B->probe_prec;
  const int64_t P = B->poff[B->nreal];
  const int64_t NRp = padded_nr(NR, prec);
  kpm_block w0, w1, pb;
  KPM_TRY(work_block(ctx, ctx->work[0], P, NR, prec, &w0));
  KPM_TRY(work_block(ctx, ctx->work[1], P, NR, prec, &w1));
  KPM_TRY(work_block(ctx, B->probes, P, NR, prec, &pb));
  KPM_HIP_TRY(hipMemcpyAsync(w0.d, pb.d, (size_t)(P * NRp) * elem_bytes(prec), hipMemcpyDeviceToDevice,
                             ctx->stream));
  KPM_HIP_TRY(hipMemcpyAsync(w1.d, pb.d, (size_t)(P * NRp) * elem_bytes(prec), hipMemcpyDeviceToDevice,
                             ctx->stream));
  KPM_TRY(ensure(ctx, ctx->mu, (size_t)B->nreal * NR * 2 * sizeof(double2)));  // the timed combines' target
  float best = 1e30f;
  StepConfig bestc;
  // candidates: 2..16 workgroups per CU (each sweeping a long contiguous
  // tile range), and short ranges of 2..32 tiles per workgroup (a one-shot
  // grid handed out in tile order)
  std::vector<StepConfig> cand;
  for (int u : {1, 2}) {
    for (int pc : {2, 4, 8, 16}) {
      StepConfig c;
      c.unroll = u;
      c.grid = std::max(ctx->num_cus, 8) * pc;
      cand.push_back(c);
    }
    const StepGeom g0 = step_geom(B->A, NR, prec, StepConfig{u, 0, 0, 0});
    const int64_t ntiles = P / g0.TR;
    for (int ch : {2, 4, 8, 16, 32}) {
      const int64_t G = (ntiles + ch - 1) / ch;
      if (G <= (int64_t)std::max(ctx->num_cus, 8) * 16 || G > (int64_t)1 << 24) continue;
      StepConfig c;
      c.unroll = u;
      c.grid = (int)G;
      cand.push_back(c);
    }
  }
  for (const StepConfig& c : cand) {
    {
      const StepGeom g = step_geom(B->A, NR, prec, c);
      KPM_TRY(build_sched(B, g.TR, c.grid));
      KPM_TRY(ensure(ctx, B->partials, (size_t)std::max(B->sched.nseg, 1) * 2 * NRp * sizeof(double2)));
      StepLaunch s{B->A, &w0, &w1, &w1, nullptr, 2.0, -1.0, MODE_DBL_SEG};
      // a step as batch_steps runs it: the fused pass and its per-realisation
      // combine (whose cost grows with the segments a schedule makes)
      auto launch = [&]() -> int {
        KPM_TRY(launch_step_seg(ctx, s, g, B->sched.G, B->sched.ntiles, (const int*)B->sched.tile_lo.ptr,
                                (const int*)B->sched.treal.ptr, (const int*)B->sched.seg0.ptr,
                                (double2*)B->partials.ptr));
        hipLaunchKernelGGL(k_batch_combine, dim3((unsigned)B->nreal), dim3(kCombineThreads), 0, ctx->stream,
                           (const double2*)B->partials.ptr, (const int*)B->sched.sfirst.ptr,
                           (const int*)B->sched.scount.ptr, 2, (int)NRp, (int)NR, (int64_t)2,
                           (double2*)ctx->mu.ptr, Combine{2, {0, 1}, {2.0, 2.0}, {-1, -1}, {0.0, 0.0}});
        KPM_HIP_TRY(hipGetLastError());
        return KPM_OK;
      };
      KPM_TRY(launch());
      KPM_HIP_TRY(hipEventRecord(ctx->events[12], ctx->stream));
      for (int r = 0; r < reps; ++r) KPM_TRY(launch());
      KPM_HIP_TRY(hipEventRecord(ctx->events[13], ctx->stream));
      KPM_HIP_TRY(hipEventSynchronize(ctx->events[13]));
      float ms = 0.f;
      KPM_HIP_TRY(hipEventElapsedTime(&ms, ctx->events[12], ctx->events[13]));
      ms /= (float)reps;
      if (ms < best) {
        best = ms;
        bestc = c;
      }
    }
  }
  tune_record(ctx, B->A, NR, prec, bestc);
  if (best_unroll) *best_unroll = bestc.unroll;
  if (best_grid) *best_grid = bestc.grid;
  if (best_ms_per_step) *best_ms_per_step = best;
  return KPM_OK;
}

int kpm_batch_graph_builds(const kpm_batch* B, int64_t* builds) {
  if (!B || !builds) return fail(KPM_EINVAL, "null argument");
  *builds = B->graph_builds;
  return KPM_OK;
}

int kpm_batch_info(const kpm_batch* B, int64_t* padded_rows, int64_t* nnz, int* nseg) {
  if (!B) return fail(KPM_EINVAL, "null batch");
  if (padded_rows) *padded_rows = B->poff[B->nreal];
  if (nnz) *nnz = B->A->nnz;
  if (nseg) *nseg = B->sched.nseg;
  return KPM_OK;
}

}  // extern "C"
