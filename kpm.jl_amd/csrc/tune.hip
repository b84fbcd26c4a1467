// tune.hip -- launch-shape table for the fused recurrence pass.
//
// A shape key is (rows, nnz, NR, precision, mode, operator value type).  The
// tuner (kpm_autotune_step) times every candidate (rows per lane in flight x
// workgroup count) on the caller's own operator and blocks and records the
// fastest; engines look the key up on every launch and fall back to the
// default geometry.  Moments do not depend on the choice beyond the order of
// the FP64 partial sums.
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "kpm_internal.h"

namespace kpm {

using Key = std::tuple<int, int64_t, int64_t, int64_t, int, int, int>;

static std::mutex g_mu;
static std::map<Key, StepConfig> g_table;

static Key make_key(kpm_ctx* ctx, const kpm_csr* A, int64_t NR, int prec, int mode) {
  return Key{ctx->device, A->n, A->nnz, NR, prec, mode, A->is_real};
}

StepConfig tuned_config(kpm_ctx* ctx, const kpm_csr* A, int64_t NR, int prec, int mode) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_table.find(make_key(ctx, A, NR, prec, mode));
  if (it != g_table.end()) return it->second;
  StepConfig c;
  c.unroll = 2;
  c.grid = 0;
  return c;
}

void tune_record(kpm_ctx* ctx, const kpm_csr* A, int64_t NR, int prec, const StepConfig& c) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (int mode : {MODE_NONE, MODE_DBL, MODE_LR}) g_table[make_key(ctx, A, NR, prec, mode)] = c;
}

}  // namespace kpm

using namespace kpm;

extern "C" {

// Time every candidate geometry of the MODE_DBL pass (the kpm_1d hot loop)
// over `reps` launches each on blocks X, Z (Z is overwritten), record the
// fastest for this shape (it then applies to all modes of the shape), and
// return it.  Candidates: rows in flight per lane {1, 2} x workgroups per CU
// {2, 4, 8, 16}.
int kpm_autotune_step(kpm_ctx* ctx, const kpm_csr* H, kpm_block* X, kpm_block* Z, int reps,
                      int* best_unroll, int* best_grid, float* best_ms_per_step) {
  if (!ctx || !H || !X || !Z || reps < 1) return fail(KPM_EINVAL, "bad arguments");
  if (X->NH != H->n || Z->NH != H->n || X->NR != Z->NR || X->prec != Z->prec)
    return fail(KPM_EINVAL, "block shapes do not match");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  ++ctx->epoch;
  const int unrolls[2] = {1, 2};
  const int per_cu[4] = {2, 4, 8, 16};
  float best = 1e30f;
  StepConfig bestc;
  for (int u : unrolls)
    for (int pc : per_cu) {
      StepConfig c;
      c.unroll = u;
      c.grid = std::max(ctx->num_cus, 8) * pc;
      StepLaunch s{H, X, Z, Z, nullptr, 2.0, -1.0, MODE_DBL};
      KPM_TRY(launch_step(ctx, s, c, nullptr));  // warm-up
      KPM_HIP_TRY(hipEventRecord(ctx->events[12], ctx->stream));
      for (int r = 0; r < reps; ++r) KPM_TRY(launch_step(ctx, s, c, nullptr));
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
  tune_record(ctx, H, X->NR, X->prec, bestc);
  if (best_unroll) *best_unroll = bestc.unroll;
  if (best_grid) *best_grid = bestc.grid;
  if (best_ms_per_step) *best_ms_per_step = best;
  return KPM_OK;
}

// Set a table entry directly (e.g. from a saved tuning file).
int kpm_tune_set(kpm_ctx* ctx, const kpm_csr* H, int64_t NR, int vec_prec, int unroll, int grid) {
  if (!ctx || !H || NR <= 0) return fail(KPM_EINVAL, "bad arguments");
  if (unroll != 1 && unroll != 2) return fail(KPM_EINVAL, "unroll must be 1 or 2");
  if (grid < 0) return fail(KPM_EINVAL, "grid must be >= 0");
  StepConfig c;
  c.unroll = unroll;
  c.grid = grid;
  tune_record(ctx, H, NR, vec_prec, c);
  return KPM_OK;
}

}  // extern "C"
