"""Launch-shape autotuner of the fused recurrence pass (no reference
counterpart; the reference has no tuning knob).

``autotune(H, NR, precision)`` times every candidate geometry (rows in flight
per lane x workgroups per CU) of the kpm_1d hot-loop pass on the caller's
operator and a probe-sized block, and records the fastest for that shape in
the library's table; every later engine call on the same shape uses it.
"""
from __future__ import annotations

import ctypes as C

from ._lib import check
from .device import DeviceBlock, get_context, prec_code, to_device


def autotune(H, NR: int, precision="f64", reps: int = 5) -> dict:
    ctx = get_context()
    Hd = to_device(H, ctx)
    n = Hd.shape[0]
    x = DeviceBlock(n, NR, precision, ctx)
    z = DeviceBlock(n, NR, precision, ctx)
    unroll, grid, ms = C.c_int(0), C.c_int(0), C.c_float(0.0)
    check(ctx.lib.kpm_autotune_step(ctx.handle, Hd.handle, x.handle, z.handle, int(reps),
                                    C.byref(unroll), C.byref(grid), C.byref(ms)), "autotune")
    return {"unroll": unroll.value, "grid": grid.value, "ms_per_pass": ms.value,
            "precision": prec_code(precision)}


def tune_set(H, NR: int, precision="f64", unroll: int = 2, grid: int = 0) -> None:
    ctx = get_context()
    Hd = to_device(H, ctx)
    check(ctx.lib.kpm_tune_set(ctx.handle, Hd.handle, int(NR), prec_code(precision), int(unroll),
                               int(grid)), "tune_set")
