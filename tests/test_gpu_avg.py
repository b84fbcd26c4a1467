"""The probe-averaged kpm_1d (avg_output=true, src/moment.jl:339-418 with
:410-416) on the engine's probe-summed doubling pass (kpm_1d_avg /
kpm_1d_block_avg): the moment dots are summed over the probe columns inside
the fused pass, so only a few bytes of partials per wave leave it.

Checked against the oracle's per-probe table averaged over probes (FP64 to
1e-12), against the engine's own per-probe path (its probe mean, to
roundoff), over padded and multi-chunk probe widths, complex operators,
every launch shape the tuner may record, and for bitwise repeatability.
"""
import ctypes

import numpy as np
import pytest

import kpmhip as K
import oracle
from kpmhip import models
from kpmhip.rng import Xoshiro256pp

pytestmark = pytest.mark.gpu


def phases(seed, n, nr):
    return K.random_phase_vectors(Xoshiro256pp(seed), n, nr)


def square(L=64):
    return (models.square_anderson(L, 2.0, seed=42) / 5.5).tocsc()


def test_c1_average_vs_oracle():
    """c1 (square L=64 Anderson W=2, NC=512, NR=8): the averaged moments equal
    the oracle's per-probe table averaged over probes, to 1e-12."""
    H = square(64)
    psi = phases(42, H.shape[0], 8)
    print("avg c1 NH=4096 NC=512 NR=8", flush=True)
    got = K.kpm_1d(H, 512, 8, psi_in=psi)
    ref = oracle.kpm_1d(H, 512, psi).mean(axis=0)
    assert got.dtype == np.float64 and got.shape == (512,)
    np.testing.assert_allclose(got, ref.real, rtol=0, atol=1e-12)
    assert abs(got[0] - 1.0) < 1e-13  # unit probes: mu_0 = 1


@pytest.mark.parametrize("NR", [1, 2, 3, 5, 16, 33, 64, 65, 130, 256])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_probe_widths(NR, prec):
    """Padded widths (NRp > NR: zero columns add nothing) and rows of several
    64-word chunks (NR 130 FP64, NR 256 FP32: one partial per wave and row
    chunk)."""
    H = (models.graphene(12, 10, disorder=0.5, seed=3) / 3.6).tocsc()
    psi = phases(11 + NR, H.shape[0], NR)
    NC = 40
    print(f"avg widths NR={NR} {prec}", flush=True)
    ref = oracle.kpm_1d(H, NC, psi).mean(axis=0).real
    got = K.kpm_1d(H, NC, NR, psi_in=psi, precision=prec)
    tol = 1e-12 if prec == "f64" else 2e-5
    np.testing.assert_allclose(got, ref, rtol=0, atol=tol)
    # the same engine's per-probe table, averaged on the host
    per = K.kpm_1d(H, NC, NR, psi_in=psi, avg_output=False, precision=prec).mean(axis=0).real
    np.testing.assert_allclose(got, per, rtol=0, atol=1e-13)


def test_complex_operator_and_block_entry():
    """Haldane (complex H, 10 entries per row) through kpm_1d_block_avg on a
    device probe block: the complex probe sums divided by NR."""
    H, _, _, _ = models.haldane_model(10, 9, t2=0.1)
    Hn = (H / 3.5).tocsc()
    NR, NC = 7, 64
    psi = phases(9, Hn.shape[0], NR)
    ref = oracle.kpm_1d(Hn, NC, psi).mean(axis=0)
    ctx = K.get_context()
    Hd = K.to_device(Hn)
    blk = K.DeviceBlock.from_host(psi, "f64")
    out = np.zeros(NC, dtype=np.complex128)
    K._lib.check(ctx.lib.kpm_1d_block_avg(ctx.handle, Hd.handle, NC, blk.handle,
                                          K._lib.dptr(out)), "kpm_1d_block_avg")
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-12)
    # imaginary parts are roundoff for a Hermitian operator and equal vectors
    assert np.max(np.abs(out.imag)) < 1e-13


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_launch_shapes_and_repeatability(prec):
    """Every persistent and one-shot shape gives the same averages to
    roundoff (only the FP64 grouping of the sums moves); one shape run twice
    is bitwise the same; the tuner's pick is one of the shapes."""
    H = (models.graphene(40, 40, disorder=0.5, seed=9) / 3.6).tocsc()
    NR, NC = 16, 64
    psi = phases(19, H.shape[0], NR)
    K.tune_set(H, NR, prec, unroll=2, grid=0)
    base = K.kpm_1d(H, NC, NR, psi_in=psi, precision=prec)
    again = K.kpm_1d(H, NC, NR, psi_in=psi, precision=prec)
    assert np.array_equal(base, again)
    for unroll in (1, 2):
        for grid in (8, 64, 2048):
            print(f"avg shape {prec} unroll={unroll} grid={grid}", flush=True)
            K.tune_set(H, NR, prec, unroll=unroll, grid=grid)
            got = K.kpm_1d(H, NC, NR, psi_in=psi, precision=prec)
            np.testing.assert_allclose(got, base, rtol=0, atol=1e-13)
        for chunk, xm in ((1, 0), (2, 0), (4, 0), (16, 0), (2, 1), (1, 4), (4, 2), (2, 8)):
            print(f"avg shape {prec} unroll={unroll} chunk={chunk} xmap={xm}", flush=True)
            K.tune_set_chunked(H, NR, prec, unroll=unroll, chunk=chunk, xcd_grouped=xm)
            got = K.kpm_1d(H, NC, NR, psi_in=psi, precision=prec)
            np.testing.assert_allclose(got, base, rtol=0, atol=1e-13)
            per = K.kpm_1d(H, NC, NR, psi_in=psi, avg_output=False, precision=prec)
            np.testing.assert_allclose(per.mean(axis=0).real, base, rtol=0, atol=1e-13)
    picked = K.autotune(H, NR, prec)
    assert picked["avg"]["unroll"] in (1, 2)
    got = K.kpm_1d(H, NC, NR, psi_in=psi, precision=prec)
    np.testing.assert_allclose(got, base, rtol=0, atol=1e-13)
    K.tune_set(H, NR, prec, unroll=2, grid=0)


def test_two_level_reduction_large_grid():
    """A one-shot grid of one tile per workgroup on 2^17 rows writes 2^15+
    wave partials, so the sums go through the two-level reduction."""
    H = (models.graphene(256, 256, disorder=0.5, seed=4) / 3.6).tocsc()
    NR, NC = 8, 24
    psi = phases(23, H.shape[0], NR)
    ref = oracle.kpm_1d(H, NC, psi).mean(axis=0).real
    for chunk in (1, 64):
        print(f"avg two-level chunk={chunk}", flush=True)
        K.tune_set_chunked(H, NR, "f64", unroll=2, chunk=chunk)
        got = K.kpm_1d(H, NC, NR, psi_in=psi)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)
    K.tune_set(H, NR, "f64", unroll=2, grid=0)


def test_errors_and_reference_semantics():
    H = square(8)
    psi = phases(3, H.shape[0], 4)
    with pytest.raises(AssertionError):
        K.kpm_1d(H, 7, 4, psi_in=psi)                      # odd NC (moment.jl:434)
    with pytest.raises(AssertionError):
        K.kpm_1d(H, 8, 4, psi_in=psi[:-1])                 # probe shape (:438)
    ctx = K.get_context()
    Hd = K.to_device(H)
    out = np.zeros(8, dtype=np.complex128)
    rc = ctx.lib.kpm_1d_avg(ctx.handle, Hd.handle, 9, 4, H.shape[0],
                            K._lib.dptr(np.asfortranarray(psi)), K._lib.dptr(out), 0)
    assert rc == K._lib.KPM_EINVAL
    # force_norm still normalises the caller's probes in place (dos_test.jl:117-118)
    p = 3.0 * phases(4, H.shape[0], 4)
    mu = K.kpm_1d(H, 8, 4, psi_in=p, force_norm=True)
    assert abs(mu[0] - 1.0) < 1e-13
    np.testing.assert_allclose(np.linalg.norm(p, axis=0), 1.0, rtol=1e-14)
    # random probes: unit columns, so mu_0 = 1
    mu = K.kpm_1d(H, 8, 5)
    assert abs(mu[0] - 1.0) < 1e-13
