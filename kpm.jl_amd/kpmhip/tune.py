"""Launch-shape autotuner of the fused recurrence pass (no reference
counterpart; the reference has no tuning knob).

``autotune(H, NR, precision)`` times every candidate geometry (rows in
flight per lane x a persistent grid of 2..16 workgroups per CU or a one-shot
grid of 2..64 row tiles per workgroup) of three passes: the kpm_1d hot-loop
pass, its probe-summed form, and the plain pass without dots.  It runs them
on the caller's operator and a probe-sized block of random-phase probes, and
records the fastest of each for that shape in the library's table; every
later engine call on the same shape uses them.  The blocks hold probes because the pass's
speed depends on the values it moves: on all-zero blocks the schedules rank
differently (graphene 2M sites, NR 64, FP32: per-XCD runs of 8 chunks time
fastest on zeros, 561 us against 568-570 us for runs of 2 or none, but slowest
on random phases, 587 us against 569-575 us).
"""
from __future__ import annotations

import ctypes as C

from ._lib import check
from .device import DeviceBlock, get_context, prec_code, to_device


def autotune(H, NR: int, precision="f64", reps: int = 8) -> dict:
    ctx = get_context()
    Hd = to_device(H, ctx)
    n = Hd.shape[0]
    x = DeviceBlock(n, NR, precision, ctx).random_phase(0x5EED)
    z = DeviceBlock(n, NR, precision, ctx).random_phase(0x5EED + 1)
    unroll, grid, ms = C.c_int(0), C.c_int(0), C.c_float(0.0)
    check(ctx.lib.kpm_autotune_step(ctx.handle, Hd.handle, x.handle, z.handle, int(reps),
                                    C.byref(unroll), C.byref(grid), C.byref(ms)), "autotune")
    out = {"unroll": unroll.value, "grid": grid.value, "ms_per_pass": ms.value,
           "precision": prec_code(precision)}
    out.update({k: v for k, v in tune_get(Hd, NR, precision).items() if k in ("chunk", "xmap")})
    out["avg"] = tune_get(Hd, NR, precision, avg=True)
    out["plain"] = tune_get(Hd, NR, precision, plain=True)
    return out


def tune_get(H, NR: int, precision="f64", avg: bool = False, plain: bool = False) -> dict:
    """The launch shape the engines use for (H, NR, precision); chunk 0 is
    the persistent grid.  ``avg``: the probe-summed doubling pass of the
    averaged kpm_1d instead of the per-probe passes; ``plain``: the passes
    without dots (kpm_2d / kpm_3d chains, the action, the adjoint's source
    steps)."""
    ctx = get_context()
    Hd = to_device(H, ctx)
    u, g, ch, xm = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
    mode = 2 if plain else 1 if avg else 0
    check(ctx.lib.kpm_tune_get_mode(ctx.handle, Hd.handle, int(NR), prec_code(precision), mode,
                                    C.byref(u), C.byref(g), C.byref(ch), C.byref(xm)), "tune_get")
    return {"unroll": u.value, "grid": g.value, "chunk": ch.value, "xmap": xm.value}


def tune_set_chunked(H, NR: int, precision="f64", unroll: int = 1, chunk: int = 64,
                     xcd_grouped: int = 0) -> None:
    """Record the one-shot schedule (chunk row tiles per workgroup) for a
    shape.  xcd_grouped: 0 chunks in dispatch order, 1 one contiguous eighth
    per XCD, m >= 2 dispatch order with runs of m chunks per XCD."""
    ctx = get_context()
    Hd = to_device(H, ctx)
    check(ctx.lib.kpm_tune_set_chunked(ctx.handle, Hd.handle, int(NR), prec_code(precision),
                                       int(unroll), int(chunk), int(xcd_grouped)),
          "tune_set_chunked")


def tune_set(H, NR: int, precision="f64", unroll: int = 2, grid: int = 0) -> None:
    ctx = get_context()
    Hd = to_device(H, ctx)
    check(ctx.lib.kpm_tune_set(ctx.handle, Hd.handle, int(NR), prec_code(precision), int(unroll),
                               int(grid)), "tune_set")
