"""kpm_2d's split-BF16 contraction on slots written as BF16 planes.

In the split-BF16 contraction mode (gemm="bf16x3") the recurrence passes of
kpm_2d (src/moment.jl:987-1088) can write each slot as its hi/lo BF16 split
instead of FP64 / FP32 words.  The left chain then runs on a three-block ring
and also writes planes, the right slots Jb T_n Ja psi_r are written only as
planes (by the dual pass or the Jb pass), and the contraction reads them
without a split.  The split is the one the contraction applies to stored
slots, and the MFMA order is the same, so the table must be bitwise the
FP64-slot table of the same blocking.  This is checked over blockings with
wraps and several right blocks, mn_sym, moment parity, the separate Jb pass,
FP32 vectors and unequal probes.  The plane table is also held to the split
mode's accuracy against the FP64 contraction, and a blocking with fewer than
128 slots a side is checked to keep the FP64-slot path.
"""
import numpy as np
import pytest

import kpmhip as K
from kpmhip import models
from kpmhip.rng import Xoshiro256pp

pytestmark = pytest.mark.gpu


def haldane(L):
    H, Jx, Jy, _ = models.haldane_model(L, L, t2=0.1)
    return (H / 3.5).tocsc(), Jx.tocsc(), Jy.tocsc()


def table(H, Jx, Jy, NC, NR, psi, planes, **kw):
    ctx = K.get_context()
    ctx.set_bf16_planes(planes)
    try:
        return K.kpm_2d(H, Jx, Jy, NC, NR, psi_in=psi, gemm="bf16x3", **kw)
    finally:
        ctx.set_bf16_planes(True)


CASES = [
    # (L, NR, NC, kwargs)
    (12, 4, 256, dict(arr_size=256, right_block=256)),
    (12, 4, 384, dict(arr_size=128, right_block=128)),            # 3 reps (wraps) x 3 right blocks
    (12, 4, 256, dict(arr_size=128, right_block=128, mn_sym=True)),
    (12, 4, 256, dict(arr_size=256, right_block=128, moment_parity="ODD")),
    (10, 6, 256, dict(arr_size=256, right_block=256, precision="f32")),
    (10, 16, 128, dict(arr_size=128, right_block=128)),
    # 192 slots a side: the second 128-slot tile of each side is partial (its
    # rows past the last slot are clamped on load and masked on store)
    (12, 4, 384, dict(arr_size=192, right_block=192)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_planes_bitwise_equal_to_slots(case):
    L, NR, NC, kw = CASES[case]
    H, Jx, Jy = haldane(L)
    psi = K.random_phase_vectors(Xoshiro256pp(5 + case), H.shape[0], NR)
    print(f"bf16 planes case {case}: L={L} NR={NR} NC={NC} {kw}", flush=True)
    on = table(H, Jx, Jy, NC, NR, psi, True, **kw)
    off = table(H, Jx, Jy, NC, NR, psi, False, **kw)
    assert np.array_equal(on, off)
    assert np.abs(on).max() > 0
    ref = K.kpm_2d(H, Jx, Jy, NC, NR, psi_in=psi, gemm="f64", **kw)
    err = np.max(np.abs(on - ref)) / np.max(np.abs(ref))
    assert 1e-9 < err < 1e-4  # the 3-pass split's level: the BF16 path ran


def test_separate_jb_pass_and_unequal_probes():
    """The dual pass off: the right slots come from the Jb pass's planes only
    (its FP64 output is not written); distinct left and right probes."""
    H, Jx, Jy = haldane(12)
    NR, NC = 4, 256
    pl = K.random_phase_vectors(Xoshiro256pp(41), H.shape[0], NR)
    pr = K.random_phase_vectors(Xoshiro256pp(42), H.shape[0], NR)
    ctx = K.get_context()
    ctx.set_dual_pass(False)
    try:
        print("bf16 planes, dual pass off", flush=True)
        outs = []
        for planes in (True, False):
            ctx.set_bf16_planes(planes)
            outs.append(K.kpm_2d(H, Jx, Jy, NC, NR, psi_in_l=pl, psi_in_r=pr, gemm="bf16x3",
                                 arr_size=256, right_block=256))
    finally:
        ctx.set_dual_pass(True)
        ctx.set_bf16_planes(True)
    assert np.array_equal(outs[0], outs[1])
    # and the dual pass's planes give the same table as the separate Jb pass
    fused = K.kpm_2d(H, Jx, Jy, NC, NR, psi_in_l=pl, psi_in_r=pr, gemm="bf16x3", arr_size=256,
                     right_block=256)
    assert np.array_equal(fused, outs[0])


def test_small_blocks_keep_slot_path():
    """A last rep of fewer than 128 left slots: the whole call keeps the
    FP64-slot contraction, so the switch changes nothing."""
    H, Jx, Jy = haldane(10)
    NR, NC = 4, 200
    psi = K.random_phase_vectors(Xoshiro256pp(7), H.shape[0], NR)
    kw = dict(arr_size=128, right_block=128)
    on = table(H, Jx, Jy, NC, NR, psi, True, **kw)
    off = table(H, Jx, Jy, NC, NR, psi, False, **kw)
    assert np.array_equal(on, off)
    ref = K.kpm_2d(H, Jx, Jy, NC, NR, psi_in=psi, gemm="f64", **kw)
    err = np.max(np.abs(on - ref)) / np.max(np.abs(ref))
    assert 1e-9 < err < 1e-4


def test_kpm_3d_nested_form():
    """kpm_3d's nested form (src/moment.jl:1106-1227) runs one kpm_2d per outer
    order; with blocks of 128 slots (fewer left slots than NC, so not the
    batched form) each of them takes the plane path: bitwise the FP64-slot
    table."""
    H, Jx, Jy = haldane(4)
    NR, NC = 2, 256
    psi = K.random_phase_vectors(Xoshiro256pp(77), H.shape[0], NR)
    ctx = K.get_context()
    outs = []
    try:
        for planes in (True, False):
            print(f"bf16 planes kpm_3d nested planes={planes}", flush=True)
            ctx.set_bf16_planes(planes)
            outs.append(K.kpm_3d(H, Jx, Jy, Jx, NC, NR, psi_in=psi, gemm="bf16x3", arr_size=128,
                                 right_block=128))
    finally:
        ctx.set_bf16_planes(True)
    assert np.array_equal(outs[0], outs[1])
    assert np.abs(outs[0]).max() > 0
