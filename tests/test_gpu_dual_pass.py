"""The fused right-chain pass of kpm_2d and kpm_3d against the same engine
with the passes run separately.

In the fused pass, each right-chain step T_n = 2 H T_{n-1} - T_{n-2} also
emits Jb T_{n-1}, the R slot of the previous order, from the same gathers.
H and Jb are merged onto the union of their patterns for it. Each product
still sums its own nonzeros in its own row order, so the tables must be
bitwise equal to those of two separate passes (src/moment.jl:1039-1051 is the
right chain they restate), and equal to the oracle.

Cases:
- Haldane currents (Jb's pattern inside H's), FP64 and FP32 blocks, every
  blocking of the reference's blocked path, parities and mn_sym;
- operators whose patterns differ, with union rows beyond 16 nonzeros (the
  spill-over loop of the pass), and a real H merged with a complex Jb;
- one-lane rows (NR = 1 at FP64), where the fused pass does not apply;
- kpm_3d with the outer orders batched as column blocks, and nested;
- an empty Jb or Ja (no stored entries);
- a Jb with a duplicated stored entry, whose repeated column switches the
  fused pass off;
- a merged operator over the context's size limit, which falls back to the
  separate passes as when its allocation fails.
The context's dual-pass counter shows which form actually ran.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
import oracle
from kpmhip import models
from kpmhip.rng import Xoshiro256pp
from matrices import ragged_hermitian

pytestmark = pytest.mark.gpu

TOL = dict(rtol=1e-12, atol=1e-12)


def phases(seed, n, nr):
    return K.random_phase_vectors(Xoshiro256pp(seed), n, nr)


LAST_DUAL = {}


def fused_and_separate(fn, expect_fused=True):
    """Run fn with the dual pass off, then on; check with the context's
    dual-pass counter that the separate run used none and that the fused run
    used it exactly when expected."""
    ctx = K.get_context()
    ctx.set_dual_pass(False)
    try:
        ctx.counters(reset=True)
        sep = fn()
        assert ctx.counters()["dual"] == 0
    finally:
        ctx.set_dual_pass(True)
    ctx.counters(reset=True)
    got = fn()
    dual = ctx.counters()["dual"]
    LAST_DUAL["n"] = dual
    if expect_fused:
        assert dual > 0, "the fused pass did not run"
    else:
        assert dual == 0, f"the fused pass ran {dual} times where it must not apply"
    return got, sep


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_kpm2d_haldane_fused_equals_separate(prec):
    H, Jx, Jy, area = models.haldane_model(6, 6, t2=0.2)
    a, b, Hn = K.normalizeH(H, center=True)
    NR, NC = 5, 24
    pl, pr = phases(3, H.shape[0], NR), phases(4, H.shape[0], NR)
    ref = oracle.kpm_2d(Hn, Jx, Jy, NC, pl, pr, NC, NC)
    for KL, KR in ((NC, NC), (5, 3), (2, 1), (7, 24), (NC, 2)):
        print(f"haldane kpm_2d {prec} KL={KL} KR={KR}", flush=True)
        # a right block of one order has no previous order to fuse with
        got, sep = fused_and_separate(lambda: K.kpm_2d(Hn, Jx, Jy, NC, NR, psi_in_l=pl, psi_in_r=pr,
                                                       arr_size=KL, right_block=KR, precision=prec),
                                      expect_fused=KR >= 2)
        np.testing.assert_array_equal(got, sep)
        if prec == "f64":
            np.testing.assert_allclose(got, ref, **TOL)
        else:
            assert np.max(np.abs(got - ref)) < 1e-5
    got, sep = fused_and_separate(lambda: K.kpm_2d(Hn, Jx, Jy, NC, NR, psi_in=pl, arr_size=5,
                                                   right_block=3, mn_sym=True, precision=prec,
                                                   moment_parity="EVEN"))
    np.testing.assert_array_equal(got, sep)


def test_kpm2d_differing_patterns_and_long_rows():
    n, NR, NC = 97, 4, 20
    H = ragged_hermitian(n, 3, seed=5)
    Ja = ragged_hermitian(n, 5, seed=6)
    Jb = ragged_hermitian(n, 20, seed=7)  # union rows up to 41 nonzeros
    pl, pr = phases(8, n, NR), phases(9, n, NR)
    ref = oracle.kpm_2d(H, Ja, Jb, NC, pl, pr, NC, NC)
    for KL, KR in ((NC, NC), (6, 4)):
        print(f"ragged kpm_2d KL={KL} KR={KR}", flush=True)
        got, sep = fused_and_separate(lambda: K.kpm_2d(H, Ja, Jb, NC, NR, psi_in_l=pl, psi_in_r=pr,
                                                       arr_size=KL, right_block=KR))
        np.testing.assert_array_equal(got, sep)
        np.testing.assert_allclose(got, ref, **TOL)
    # real H merged with a complex Jb
    Hr = sp.csc_matrix(H.real)
    ref = oracle.kpm_2d(Hr, Ja, Jb, NC, pl, pr, NC, NC)
    got, sep = fused_and_separate(lambda: K.kpm_2d(Hr, Ja, Jb, NC, NR, psi_in_l=pl, psi_in_r=pr))
    np.testing.assert_array_equal(got, sep)
    np.testing.assert_allclose(got, ref, **TOL)


def test_kpm2d_one_lane_rows():
    H, Jx, Jy, area = models.haldane_model(4, 4, t2=0.2)
    Hn = (H / 4.0).tocsc()
    NC = 12
    pl, pr = phases(11, H.shape[0], 1), phases(12, H.shape[0], 1)
    got, sep = fused_and_separate(lambda: K.kpm_2d(Hn, Jx, Jy, NC, 1, psi_in_l=pl, psi_in_r=pr,
                                                   arr_size=5, right_block=3), expect_fused=False)
    np.testing.assert_array_equal(got, sep)
    np.testing.assert_allclose(got, oracle.kpm_2d(Hn, Jx, Jy, NC, pl, pr, 5, 3), **TOL)


@pytest.mark.parametrize("NR", [3, 8])
def test_kpm3d_fused_equals_separate(NR):
    H, Jx, Jy, area = models.haldane_model(5, 5, t2=0.2)
    Hn = (H / 4.0).tocsc()
    Jz = ((Jx + 2.0 * Jy) / 3.0).tocsc()
    NC = 11
    pl, pr = phases(41 + NR, H.shape[0], NR), phases(42 + NR, H.shape[0], NR)
    ref = oracle.kpm_3d(Hn, Jx, Jy, Jz, NC, pl, pr, 3, 3)
    print(f"kpm_3d batched NR={NR}", flush=True)
    got, sep = fused_and_separate(lambda: K.kpm_3d(Hn, Jx, Jy, Jz, NC, NR, psi_in_l=pl, psi_in_r=pr))
    np.testing.assert_array_equal(got, sep)
    np.testing.assert_allclose(got, ref, **TOL)
    print(f"kpm_3d nested NR={NR}", flush=True)
    got, sep = fused_and_separate(lambda: K.kpm_3d(Hn, Jx, Jy, Jz, NC, NR, psi_in_l=pl, psi_in_r=pr,
                                                   arr_size=NC - 1, right_block=4))
    np.testing.assert_array_equal(got, sep)
    np.testing.assert_allclose(got, ref, **TOL)


def test_kpm2d_empty_current_operator():
    """An empty Jb (no stored entries): the merged operator is H's pattern with
    zero second values, and every table entry is zero; an empty Ja likewise."""
    n, NR, NC = 50, 4, 10
    H = ragged_hermitian(n, 3, seed=21)
    J = ragged_hermitian(n, 2, seed=22)
    Z = sp.csc_matrix((n, n), dtype=np.complex128)
    pl, pr = phases(23, n, NR), phases(24, n, NR)
    for Ja, Jb in ((J, Z), (Z, J)):
        got, sep = fused_and_separate(lambda: K.kpm_2d(H, Ja, Jb, NC, NR, psi_in_l=pl, psi_in_r=pr))
        np.testing.assert_array_equal(got, sep)
        assert np.all(got == 0)


def test_kpm2d_duplicate_entries_take_separate_passes():
    """A Jb stored with a duplicated entry (its value split in two) has a row
    whose column indices repeat, so the merge does not apply and the passes
    run separately; the table still equals the oracle's on the summed
    matrix."""
    n, NR, NC = 60, 4, 12
    H = ragged_hermitian(n, 3, seed=31)
    Jb = ragged_hermitian(n, 2, seed=32).tocsc()
    Ja = ragged_hermitian(n, 2, seed=33)
    k = Jb.indptr[5]  # first stored entry of column 5, stored twice at half value
    data = np.insert(Jb.data, k, Jb.data[k] / 2)
    data[k + 1] /= 2
    indices = np.insert(Jb.indices, k, Jb.indices[k])
    indptr = Jb.indptr.copy()
    indptr[6:] += 1
    Jdup = sp.csc_matrix((data, indices, indptr), shape=(n, n))
    assert Jdup.nnz == Jb.nnz + 1 and not Jdup.has_canonical_format
    pl, pr = phases(34, n, NR), phases(35, n, NR)
    ref = oracle.kpm_2d(H, Ja, Jb, NC, pl, pr, NC, NC)
    got, sep = fused_and_separate(lambda: K.kpm_2d(H, Ja, Jdup, NC, NR, psi_in_l=pl, psi_in_r=pr),
                                  expect_fused=False)
    np.testing.assert_array_equal(got, sep)
    np.testing.assert_allclose(got, ref, **TOL)


def test_kpm2d_merged_operator_over_limit_falls_back():
    """No room for the merged operator (here: a context limit of 0 bytes, the
    same path as a failed allocation): the passes run separately, the table
    still equals the oracle's, and lifting the limit brings the fused pass
    back on the same operators."""
    H, Jx, Jy, area = models.haldane_model(5, 5, t2=0.2)
    Hn = (H / 4.0).tocsc()
    NR, NC = 4, 16
    pl, pr = phases(51, H.shape[0], NR), phases(52, H.shape[0], NR)
    ref = oracle.kpm_2d(Hn, Jx, Jy, NC, pl, pr, NC, NC)
    ctx = K.get_context()
    ctx.set_merged_limit(0)
    try:
        ctx.counters(reset=True)
        got = K.kpm_2d(Hn, Jx, Jy, NC, NR, psi_in_l=pl, psi_in_r=pr)
        assert ctx.counters()["dual"] == 0
    finally:
        ctx.set_merged_limit(-1)
    np.testing.assert_allclose(got, ref, **TOL)
    ctx.counters(reset=True)
    again = K.kpm_2d(Hn, Jx, Jy, NC, NR, psi_in_l=pl, psi_in_r=pr)
    assert ctx.counters()["dual"] > 0
    np.testing.assert_array_equal(again, got)


def test_failed_allocation_does_not_poison_the_next_call():
    """A device allocation that fails (a block far beyond HBM) raises
    MemoryError and leaves no latched HIP status behind: the next engine call
    on the same context runs and matches the oracle."""
    with pytest.raises(MemoryError):
        K.DeviceBlock(1 << 40, 64)
    H = ragged_hermitian(40, 3, seed=61)
    psi = phases(62, 40, 3)
    got = K.kpm_1d(H, 16, 3, psi_in=psi, avg_output=False)
    np.testing.assert_allclose(got, oracle.kpm_1d(H, 16, psi), **TOL)
