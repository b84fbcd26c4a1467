// dot_steps.h -- a sequence of fused recurrence passes whose moment dots
// are reduced one or two launches later (shared by the moment engines and
// the adjoint's moment-keeping forward pass).
#pragma once

#include <algorithm>

#include "kpm_internal.h"

namespace kpm {

// Dot steps alternate between two halves of ctx->partials; each launch also
// finishes the previous step's moment columns (PendingCombine), and only the
// last step's columns take a separate combine launch.
struct DotSteps {
  kpm_ctx* ctx;
  const kpm_csr* H;
  int64_t NR, NRp;
  double2* half[2];
  int k = 0;
  int max_grid = 0;          // partial rows each half holds (0: kMaxGridPerCU * CUs)
  double2* l2buf[2] = {nullptr, nullptr};  // slice sums of the two-level reduction
  double2* mu_dev = nullptr;  // null: ctx->mu
  // reductions still owed: a final one (into the moment columns) and a
  // first level (partials -> slice sums) whose final follows a launch later
  struct Fin {
    bool on = false;
    const double2* part = nullptr;
    int G = 0;
    Combine cb{0, {0, 0}, {0.0, 0.0}, {-1, -1}, {0.0, 0.0}};
  } fin;
  struct L1 {
    bool on = false;
    const double2* part = nullptr;
    double2* out = nullptr;
    int G = 0;
    Combine cb{0, {0, 0}, {0.0, 0.0}, {-1, -1}, {0.0, 0.0}};
  } l1;

  const StepConfig* cfg = nullptr;  // launch shape (null: the tuned one)

  double2* mu() const { return mu_dev ? mu_dev : (double2*)ctx->mu.ptr; }
  // halves (and slice-sum buffers) in ctx->partials, sized for the grid of
  // the first block's shape
  int init(int nd, const kpm_block* x, int mode) {
    max_grid = cfg ? partials_grid_cfg(ctx, H, x->NH, x->NR, x->prec, *cfg)
                   : partials_grid(ctx, H, x->NH, x->NR, x->prec, mode);
    // MODE_DBLA writes one (single-column) partial row per wave and row chunk
    if (mode == MODE_DBLA) max_grid *= kStepWaves * step_geom(H, x->NR, x->prec, StepConfig{}).chunks;
    const size_t per = (size_t)max_grid * nd * NRp;
    const size_t l2per = (size_t)((max_grid + kSlice - 1) / kSlice) * nd * NRp;
    KPM_TRY(ensure(ctx, ctx->partials, (2 * per + 2 * l2per) * sizeof(double2)));
    half[0] = (double2*)ctx->partials.ptr;
    half[1] = half[0] + per;
    l2buf[0] = half[1] + per;
    l2buf[1] = l2buf[0] + l2per;
    return KPM_OK;
  }
  // caller-owned halves of max_grid (<= kTwoLevelMin) workgroups each
  void init_external(double2* h0, double2* h1, int grid_cap) {
    half[0] = h0;
    half[1] = h1;
    max_grid = std::min(grid_cap, kTwoLevelMin);
  }
  // one fused pass whose moment columns are described by c
  int step(const kpm_block* x, const kpm_block* z, kpm_block* out, const kpm_block* l, double alpha,
           double beta, int mode, const Combine& c) {
    StepLaunch s{H, x, z, out, l, alpha, beta, mode};
    s.part_out = half[k & 1];
    s.max_grid = max_grid;
    PendingCombine pc;
    pc.mu = mu();
    pc.NRp = (int)NRp;
    pc.NR = (int)NR;
    if (fin.on) {
      pc.part = fin.part;
      pc.G = fin.G;
      pc.cb = fin.cb;
    }
    if (l1.on) {
      pc.l1_part = l1.part;
      pc.l1_out = l1.out;
      pc.l1_G = l1.G;
      pc.l1_nd = l1.cb.nd;
    }
    s.pend = (fin.on || l1.on) ? &pc : nullptr;
    int G = 0;
    KPM_TRY(launch_step(ctx, s, cfg ? *cfg : tuned_config(ctx, H, x->NR, x->prec, mode), &G));
    fin.on = false;
    if (l1.on) {  // its first level ran in this launch; the final is next
      fin.on = true;
      fin.part = l1.out;
      fin.G = (l1.G + kSlice - 1) / kSlice;
      fin.cb = l1.cb;
      l1.on = false;
    }
    if (G > kTwoLevelMin && l2buf[0]) {
      l1.on = true;
      l1.part = half[k & 1];
      l1.out = l2buf[k & 1];
      l1.G = G;
      l1.cb = c;
    } else {
      if (fin.on) {  // both would be final next launch: finish the older now
        KPM_TRY(launch_combine_from(ctx, fin.part, fin.G, NR, NRp, fin.cb, mu()));
      }
      fin.on = true;
      fin.part = half[k & 1];
      fin.G = G;
      fin.cb = c;
    }
    ++k;
    return KPM_OK;
  }
  // the owed reductions as launches of their own, older step first
  int finish() {
    if (fin.on) {
      fin.on = false;
      KPM_TRY(launch_combine_from(ctx, fin.part, fin.G, NR, NRp, fin.cb, mu()));
    }
    if (l1.on) {
      l1.on = false;
      KPM_TRY(launch_reduce1(ctx, l1.part, l1.G, l1.cb.nd, NR, NRp, l1.out));
      KPM_TRY(launch_combine_from(ctx, l1.out, (l1.G + kSlice - 1) / kSlice, NR, NRp, l1.cb, mu()));
    }
    return KPM_OK;
  }
};

}  // namespace kpm
