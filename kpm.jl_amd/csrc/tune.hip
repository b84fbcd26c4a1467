// tune.hip -- launch-shape table for the fused recurrence pass.
//
// A shape key is (rows, nnz, NR, precision, mode, operator value type).  The
// tuner (kpm_autotune_step) times every candidate (rows per lane in flight x
// workgroup count) on the caller's own operator and blocks and records the
// fastest; engines look the key up on every launch and fall back to the
// default geometry.  Moments do not depend on the choice beyond the order of
// the FP64 partial sums.
#include <algorithm>
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

void tune_record_avg(kpm_ctx* ctx, const kpm_csr* A, int64_t NR, int prec, const StepConfig& c) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_table[make_key(ctx, A, NR, prec, MODE_DBLA)] = c;
}

// the passes without dots (MODE_NONE's entry: kpm_2d / kpm_3d chains and
// their dual passes, the matrix-function action, the adjoint's source steps)
static void tune_record_plain(kpm_ctx* ctx, const kpm_csr* A, int64_t NR, int prec, const StepConfig& c) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_table[make_key(ctx, A, NR, prec, MODE_NONE)] = c;
}

// reps plain passes Z <- 2 H X - Z in place, no dots
static int time_plain_steps(kpm_ctx* ctx, const kpm_csr* H, kpm_block* X, kpm_block* Z,
                            const StepConfig& c, int reps, float* ms) {
  StepLaunch s{H, X, Z, Z, nullptr, 2.0, -1.0, MODE_NONE};
  for (int r = 0; r < 2; ++r) KPM_TRY(launch_step(ctx, s, c, nullptr));
  KPM_HIP_TRY(hipEventRecord(ctx->events[12], ctx->stream));
  for (int r = 0; r < reps; ++r) KPM_TRY(launch_step(ctx, s, c, nullptr));
  KPM_HIP_TRY(hipEventRecord(ctx->events[13], ctx->stream));
  KPM_HIP_TRY(hipEventSynchronize(ctx->events[13]));
  KPM_HIP_TRY(hipEventElapsedTime(ms, ctx->events[12], ctx->events[13]));
  return KPM_OK;
}

}  // namespace kpm

using namespace kpm;

extern "C" {

// Time every candidate geometry of the MODE_DBL pass (the kpm_1d hot loop)
// over `reps` launches each on blocks X, Z (Z is overwritten), record the
// fastest for this shape (it then applies to the passes with dots), and
// return it; the probe-summed and the plain passes get their own entries.  Candidates: rows in flight per lane {1, 2} x workgroups per CU
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
  // candidates: persistent grids of 2..16 workgroups per CU, and one-shot
  // grids of 2..64 tiles per workgroup (in-order or per-XCD chunk order) of
  // at most kMaxChunkGridPerCU workgroups per CU
  std::vector<StepConfig> cand;
  for (int u : unrolls) {
    for (int pc : per_cu) {
      StepConfig c;
      c.unroll = u;
      c.grid = std::max(ctx->num_cus, 8) * pc;
      cand.push_back(c);
    }
    const StepGeom g = step_geom(H, X->NR, X->prec, StepConfig{u, 0, 0, 0});
    const int64_t ntiles = (X->NH + g.TR - 1) / g.TR;
    const int64_t cap = (int64_t)kMaxChunkGridPerCU * std::max(ctx->num_cus, 1);
    for (int ch : {2, 4, 8, 16, 32, 64}) {
      const int64_t gc = ((ntiles + ch - 1) / ch + 7) / 8 * 8;
      if (gc > cap || gc < 2LL * std::max(ctx->num_cus, 8)) continue;  // too many, or too few to fill the chip
      for (int xm : {0, 1, 2, 4, 8}) {
        StepConfig c;
        c.unroll = u;
        c.chunk = ch;
        c.xmap = xm;
        cand.push_back(c);
      }
    }
  }
  // each candidate runs as the engine runs it: fused passes with the
  // deferred moment reductions their grid implies.  The per-probe pass
  // (MODE_DBL, whose shape the passes with dots share), the probe-summed pass
  // (MODE_DBLA: a few bytes of partials per wave instead of NRp per
  // workgroup, so it favours shorter chunks) and the plain pass without dots
  // (MODE_NONE) are tuned separately.  (On the Haldane 256 x 256, NR 16
  // shape the per-probe pass's pick, a persistent grid, made kpm_2d's plain
  // and dual passes 10 ms slower per call than one-shot chunks of 2 tiles.)
  auto best_of = [&](int mode, StepConfig* bc, float* bms) -> int {
    *bms = 1e30f;
    for (const StepConfig& c : cand) {
      // the better of two timings: candidates often lie within a few percent
      // of one another, and one timing alone let a transient pick the shape
      float ms = 0.f, ms2 = 0.f;
      if (mode == MODE_NONE) {
        KPM_TRY(time_plain_steps(ctx, H, X, Z, c, reps, &ms));
        KPM_TRY(time_plain_steps(ctx, H, X, Z, c, reps, &ms2));
      } else {
        KPM_TRY(time_engine_steps(ctx, H, X, Z, c, reps, &ms, mode));
        KPM_TRY(time_engine_steps(ctx, H, X, Z, c, reps, &ms2, mode));
      }
      ms = std::min(ms, ms2) / (float)reps;
      if (ms < *bms) {
        *bms = ms;
        *bc = c;
      }
    }
    return KPM_OK;
  };
  float best = 0.f, best_avg = 0.f, best_plain = 0.f;
  StepConfig bestc, bestc_avg, bestc_plain;
  KPM_TRY(best_of(MODE_DBL, &bestc, &best));
  KPM_TRY(best_of(MODE_DBLA, &bestc_avg, &best_avg));
  KPM_TRY(best_of(MODE_NONE, &bestc_plain, &best_plain));
  tune_record(ctx, H, X->NR, X->prec, bestc);
  tune_record_avg(ctx, H, X->NR, X->prec, bestc_avg);
  tune_record_plain(ctx, H, X->NR, X->prec, bestc_plain);
  if (best_unroll) *best_unroll = bestc.unroll;
  if (best_grid) {  // the workgroup count the winner launches with
    const StepGeom g = step_geom(H, X->NR, X->prec, bestc);
    const int64_t ntiles = (X->NH + g.TR - 1) / g.TR;
    *best_grid = bestc.chunk > 0 ? (int)(((ntiles + bestc.chunk - 1) / bestc.chunk + 7) / 8 * 8)
                                 : (bestc.grid > 0 ? bestc.grid : default_grid(ctx, ntiles));
  }
  if (best_ms_per_step) *best_ms_per_step = best;
  return KPM_OK;
}

// Set a one-shot (chunked) table entry directly.
int kpm_tune_set_chunked(kpm_ctx* ctx, const kpm_csr* H, int64_t NR, int vec_prec, int unroll,
                         int chunk, int xcd_grouped) {
  if (!ctx || !H || NR <= 0) return fail(KPM_EINVAL, "bad arguments");
  if (unroll != 1 && unroll != 2) return fail(KPM_EINVAL, "unroll must be 1 or 2");
  if (chunk < 1) return fail(KPM_EINVAL, "chunk must be >= 1 tile");
  if (xcd_grouped < 0 || xcd_grouped > 64) return fail(KPM_EINVAL, "xcd_grouped must be 0 .. 64");
  StepConfig c;
  c.unroll = unroll;
  c.chunk = chunk;
  c.xmap = xcd_grouped;
  tune_record(ctx, H, NR, vec_prec, c);
  tune_record_avg(ctx, H, NR, vec_prec, c);
  return KPM_OK;
}

// The table entry the engines use for a shape (defaults when untuned):
// mode 0 the per-probe passes, mode 1 the probe-summed doubling pass, mode 2
// the plain passes without dots.
int kpm_tune_get_mode(kpm_ctx* ctx, const kpm_csr* H, int64_t NR, int vec_prec, int mode,
                      int* unroll, int* grid, int* chunk, int* xcd_grouped) {
  if (!ctx || !H || NR <= 0 || mode < 0 || mode > 2) return fail(KPM_EINVAL, "bad arguments");
  const StepConfig c = tuned_config(ctx, H, NR, vec_prec, mode == 2 ? MODE_NONE : mode ? MODE_DBLA : MODE_DBL);
  if (unroll) *unroll = c.unroll;
  if (grid) *grid = c.grid;
  if (chunk) *chunk = c.chunk;
  if (xcd_grouped) *xcd_grouped = c.xmap;
  return KPM_OK;
}

int kpm_tune_get(kpm_ctx* ctx, const kpm_csr* H, int64_t NR, int vec_prec, int* unroll, int* grid,
                 int* chunk, int* xcd_grouped) {
  if (!ctx || !H || NR <= 0) return fail(KPM_EINVAL, "bad arguments");
  const StepConfig c = tuned_config(ctx, H, NR, vec_prec, MODE_DBL);
  if (unroll) *unroll = c.unroll;
  if (grid) *grid = c.grid;
  if (chunk) *chunk = c.chunk;
  if (xcd_grouped) *xcd_grouped = c.xmap;
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
  tune_record_avg(ctx, H, NR, vec_prec, c);
  return KPM_OK;
}

}  // extern "C"
