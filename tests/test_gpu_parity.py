"""Parity of the HIP path (libkpm_hip.so through its C ABI) with the oracle
and the golden fixtures.  FP64 results must match to 1e-12 (the reference's
kpm_1d parity target); FP32-vector results are bounded against the FP64 path.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
import oracle
from kpmhip import models
from kpmhip.rng import Xoshiro256pp

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = dict(rtol=1e-12, atol=1e-12)


def load(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


def phases(seed, n, nr):
    return K.random_phase_vectors(Xoshiro256pp(seed), n, nr)


def c1_model():
    """square L=64 Anderson W=2 with a fixed rescaling (a, b) shared by both sides."""
    H = models.square_anderson(64, 2.0, seed=42)
    a, b = 5.5, 0.0
    return (H / a).tocsc()


# ---- golden fixtures --------------------------------------------------------------------

def test_kpm1d_golden_fixtures():
    g = load("kpm1d_dense.npz")
    NC = int(g["NC"])
    H = sp.csc_matrix(g["H"])
    mu = K.kpm_1d(H, NC, g["psi"].shape[1], psi_in=g["psi"], avg_output=False)
    np.testing.assert_allclose(mu, g["mu_eq"], **TOL)
    mu = K.kpm_1d(H, NC, 3, psi_in_l=g["psi_l"], psi_in_r=g["psi_r"], avg_output=False)
    np.testing.assert_allclose(mu, g["mu_lr"], **TOL)
    mu = K.kpm_1d_current(H, sp.csc_matrix(g["J"]), NC, 3, psi_in=g["psi"], avg_output=False)
    np.testing.assert_allclose(mu, g["mu_cur"], **TOL)


def test_kpm2d_golden_fixtures_blockings():
    g = load("kpm2d_dense.npz")
    NC = int(g["NC"])
    args = [sp.csc_matrix(g[k]) for k in ("H", "Ja", "Jb")]
    for KL, KR in ((2, 1), (3, 3), (5, 16), (NC, NC), (NC + 7, 2)):
        mu = K.kpm_2d(*args, NC, 2, psi_in_l=g["psi_l"], psi_in_r=g["psi_r"], arr_size=KL,
                      right_block=KR)
        np.testing.assert_allclose(mu, g["mu"], **TOL)


def test_kpm3d_golden_fixture():
    g = load("kpm3d_dense.npz")
    NC = int(g["NC"])
    mu = K.kpm_3d(*[sp.csc_matrix(g[k]) for k in ("H", "Ja", "Jb", "Jg")], NC, 2,
                  psi_in_l=g["psi_l"], psi_in_r=g["psi_r"], arr_size=3, right_block=3)
    np.testing.assert_allclose(mu, g["mu"], **TOL)


# ---- c1: square L=64 Anderson, NC=512, NR=8 vs the oracle at 1e-12 ------------------------

def test_c1_kpm1d_vs_oracle():
    H = c1_model()
    psi = phases(42, H.shape[0], 8)
    print("c1 kpm_1d NH=4096 NC=512 NR=8", flush=True)
    got = K.kpm_1d(H, 512, 8, psi_in=psi, avg_output=False)
    ref = oracle.kpm_1d(H, 512, psi)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)
    assert np.max(np.abs(got[:, 0] - 1.0)) < 1e-13  # unit probes: mu_0 = 1


def test_determinism_bitwise():
    H = c1_model()
    psi = phases(5, H.shape[0], 8)
    a = K.kpm_1d(H, 64, 8, psi_in=psi, avg_output=False)
    b = K.kpm_1d(H, 64, 8, psi_in=psi, avg_output=False)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_autotuner_choice_invariance(prec):
    """Every launch shape the tuner may record (rows per lane x workgroups)
    gives the same moments; only the partial-sum grouping moves, so they agree
    to roundoff, and the tuner's own pick is one of them."""
    H = (models.graphene(40, 40, disorder=0.5, seed=9) / 3.6).tocsc()
    NR, NC = 16, 64
    psi = phases(19, H.shape[0], NR)
    K.tune_set(H, NR, prec, unroll=2, grid=0)
    base = K.kpm_1d(H, NC, NR, psi_in=psi, avg_output=False, precision=prec)
    tol = 1e-13  # the stored vectors are identical; only FP64 dot grouping moves
    for unroll in (1, 2):
        for grid in (8, 64, 512, 2048):
            print(f"tune prec={prec} unroll={unroll} grid={grid}", flush=True)
            K.tune_set(H, NR, prec, unroll=unroll, grid=grid)
            got = K.kpm_1d(H, NC, NR, psi_in=psi, avg_output=False, precision=prec)
            np.testing.assert_allclose(got, base, rtol=0, atol=tol)
    picked = K.autotune(H, NR, prec)
    assert picked["unroll"] in (1, 2) and picked["grid"] % 8 == 0
    got = K.kpm_1d(H, NC, NR, psi_in=psi, avg_output=False, precision=prec)
    np.testing.assert_allclose(got, base, rtol=0, atol=tol)
    K.tune_set(H, NR, prec, unroll=2, grid=0)


@pytest.mark.parametrize("NR", [1, 2, 3, 5, 16, 33, 64, 65, 130])
def test_kpm1d_probe_widths(NR):
    """padded probe widths (NRp > NR) and multi-chunk rows (NR > 64)."""
    H = (models.graphene(12, 10, disorder=0.5, seed=3) / 3.6).tocsc()
    psi = phases(11 + NR, H.shape[0], NR)
    NC = 40
    ref = oracle.kpm_1d(H, NC, psi)
    np.testing.assert_allclose(K.kpm_1d(H, NC, NR, psi_in=psi, avg_output=False), ref, **TOL)
    pl = phases(100 + NR, H.shape[0], NR)
    ref_lr = oracle.kpm_1d_lr(H, 17, pl, psi)
    got_lr = K.kpm_1d(H, 17, NR, psi_in_l=pl, psi_in_r=psi, avg_output=False)
    np.testing.assert_allclose(got_lr, ref_lr, **TOL)


def test_kpm1d_haldane_complex_and_current():
    H, Jx, Jy, area = models.haldane_model(8, 8, t2=0.1)
    Hn = (H / 3.5).tocsc()
    psi = phases(9, Hn.shape[0], 4)
    np.testing.assert_allclose(K.kpm_1d(Hn, 64, 4, psi_in=psi, avg_output=False),
                               oracle.kpm_1d(Hn, 64, psi), **TOL)
    np.testing.assert_allclose(K.kpm_1d_current(Hn, Jx, 31, 4, psi_in=psi, avg_output=False),
                               oracle.kpm_1d_current(Hn, Jx, 31, psi), **TOL)


def test_empty_rows_and_nc2():
    n = 50
    H = sp.lil_matrix((n, n), dtype=complex)
    for i in range(0, n - 1, 3):  # rows 2, 5, ... stay empty in one direction
        H[i, i + 1] = 0.3 + 0.1j
        H[i + 1, i] = 0.3 - 0.1j
    H = H.tocsc()
    psi = phases(1, n, 3)
    np.testing.assert_allclose(K.kpm_1d(H, 2, 3, psi_in=psi, avg_output=False),
                               oracle.kpm_1d(H, 2, psi), **TOL)
    np.testing.assert_allclose(K.kpm_1d(H, 30, 3, psi_in=psi, avg_output=False),
                               oracle.kpm_1d(H, 30, psi), **TOL)


def test_argument_errors():
    H = c1_model()
    psi = phases(1, H.shape[0], 2)
    with pytest.raises(AssertionError):
        K.kpm_1d(H, 33, 2, psi_in=psi)  # odd NC (src/moment.jl:434)
    with pytest.raises(AssertionError):
        K.kpm_1d(H, 32, 3, psi_in=psi)  # probe shape (:438)
    with pytest.raises(K.ArgumentError):
        K.kpm_2d(H, H, H, 8, 2, psi_in=psi, moment_parity="BOTH")


def test_stochastic_trace_unbiased():
    """test/dos_test.jl:66-78: H = 0.8 |u><u| with u the uniform state; probes
    that were mean-centred would be orthogonal to u and give mu_1 = 0 instead
    of 0.8 / D.  Checked with explicit host probes and with the device's own
    random-phase probes."""
    D = 8
    u = np.full(D, 1 / np.sqrt(D))
    H = sp.csc_matrix(0.8 * np.outer(u, u))
    NR = 512
    psi = phases(21, D, NR)
    mu = K.kpm_1d(H, 4, NR, psi_in=psi)
    assert abs(mu[1] - 0.8 / D) <= 0.3 * 0.8 / D
    mu_dev = K.kpm_1d(H, 4, NR)
    assert abs(mu_dev[1] - 0.8 / D) <= 0.3 * 0.8 / D


def test_force_norm_and_nr_parallel():
    """test/dos_test.jl:46-64,95-141"""
    H = c1_model()
    rng = Xoshiro256pp(19)
    raw = K.random_phase_vectors(rng, H.shape[0], 4) * np.array([1.0, 5.0, 0.2, 12.0])
    pf = raw.copy(order="F")
    mf = K.kpm_1d(H, 16, 4, psi_in=pf, force_norm=True)
    pp = raw.copy(order="F")
    K.normalize_by_col(pp)
    mp = K.kpm_1d(H, 16, 4, psi_in=pp)
    np.testing.assert_allclose(mf, mp, rtol=1e-12)
    np.testing.assert_allclose(pf, pp, rtol=1e-12)  # normalised in place
    ms = K.kpm_1d(H, 16, 4, psi_in=pp, NR_parallel=False)
    np.testing.assert_allclose(ms, mp, atol=1e-12)


# ---- primitives (method table) ---------------------------------------------------------------

def test_chebyshev_iter_and_colwise_dot():
    rng = Xoshiro256pp(1234)
    NH, NR, n = 300, 3, 6
    H = (models.graphene(15, 10, disorder=0.3) / 3.5).tocsc()
    s0 = phases(1, NH, NR)
    s1 = phases(2, NH, NR)
    slots = [K.DeviceBlock(NH, NR) for _ in range(n)]
    slots[0].upload(s0)
    slots[1].upload(s1)
    K.chebyshev_iter(H, slots, n)
    ref = np.zeros((NH, NR, n), dtype=complex, order="F")
    ref[:, :, 0], ref[:, :, 1] = s0, s1
    ref = oracle.cheb_iter(H, ref)
    for i in range(n):
        np.testing.assert_allclose(slots[i].download(), ref[:, :, i], **TOL)
    # 2-slot wrap in place (V_out === V_pp)
    K.chebyshev_iter_wrap(H, slots[:2], 2)
    Hd = H.toarray()
    v0 = 2 * Hd @ s1 - s0
    v1 = 2 * Hd @ v0 - s1
    np.testing.assert_allclose(slots[0].download(), v0, **TOL)
    np.testing.assert_allclose(slots[1].download(), v1, **TOL)
    tgt = np.zeros((NR, 4), dtype=np.complex128, order="F")
    K.columnwise_dot(tgt, 1, slots[2], slots[3])
    K.columnwise_dot(tgt, 2, slots[2], slots[3], alpha=2.0, beta_col=1, beta_scale=-1.0)
    d = np.einsum("ir,ir->r", ref[:, :, 2].conj(), ref[:, :, 3])
    np.testing.assert_allclose(tgt[:, 0], d, **TOL)
    np.testing.assert_allclose(tgt[:, 1], 2 * d - d, **TOL)


# ---- kpm_2d / kpm_3d vs the oracle -------------------------------------------------------------

def test_kpm2d_haldane_vs_oracle_blockings():
    H, Jx, Jy, area = models.haldane_model(6, 6, t2=0.2)
    a, b, Hn = K.normalizeH(H, center=True)
    NR, NC = 5, 24
    pl, pr = phases(3, H.shape[0], NR), phases(4, H.shape[0], NR)
    ref = oracle.kpm_2d(Hn, Jx, Jy, NC, pl, pr, NC, NC)
    for KL, KR in ((NC, NC), (5, 3), (2, 1), (7, 24)):
        got = K.kpm_2d(Hn, Jx, Jy, NC, NR, psi_in_l=pl, psi_in_r=pr, arr_size=KL, right_block=KR)
        np.testing.assert_allclose(got, ref, **TOL)
    for par, code in (("ODD", 1), ("EVEN", 2)):
        got = K.kpm_2d(Hn, Jx, Jy, NC, NR, psi_in_l=pl, psi_in_r=pr, arr_size=5, right_block=3,
                       moment_parity=par)
        np.testing.assert_allclose(got, oracle.kpm_2d(Hn, Jx, Jy, NC, pl, pr, 5, 3, parity=code),
                                   **TOL)
    got = K.kpm_2d(Hn, Jx, Jy, NC, NR, psi_in=pl, arr_size=5, right_block=3, mn_sym=True)
    np.testing.assert_allclose(got, oracle.kpm_2d(Hn, Jx, Jy, NC, pl, pl, 5, 3, mn_sym=True), **TOL)


def test_kpm2d_sigma_xy_identity_probes():
    """kubo_bastin_test.jl:19-56 on GPU moments: σxy quantised, equal to the
    oracle-moment σxy to 1e-9."""
    H, Jx, Jy, area = models.haldane_model(6, 6, t2=0.2)
    a, b, Hn = K.normalizeH(H, center=True)
    D = H.shape[0]
    psi = np.eye(D, dtype=complex)
    NC = 96
    mu = K.kpm_2d(Hn, Jx, Jy, NC, D, psi_in=psi)
    ref = oracle.kpm_2d(Hn, Jx, Jy, NC, psi, psi, NC, NC)
    np.testing.assert_allclose(mu, ref, rtol=1e-11, atol=1e-12)
    s = K.kubo_bastin_cond_raw(mu, a, 0.0, b=b, NH=D, area=area)
    s_ref = K.kubo_bastin_cond_raw(ref, a, 0.0, b=b, NH=D, area=area)
    assert abs(s - s_ref) < 1e-9
    assert abs(abs(s) - 1) < 0.05


def test_kpm3d_vs_oracle():
    rng = Xoshiro256pp(804)
    H, Jx, Jy, area = models.haldane_model(3, 3, t2=0.2)
    Hn = (H / 4.0).tocsc()
    NR, NC = 2, 6
    pl, pr = phases(5, H.shape[0], NR), phases(6, H.shape[0], NR)
    got = K.kpm_3d(Hn, Jx, Jy, Jx, NC, NR, psi_in_l=pl, psi_in_r=pr, arr_size=3, right_block=3)
    ref = oracle.kpm_3d(Hn, Jx, Jy, Jx, NC, pl, pr, 3, 3)
    np.testing.assert_allclose(got, ref, **TOL)
    # the engine's own slot plan (all slots resident) gives the same table
    auto = K.kpm_3d(Hn, Jx, Jy, Jx, NC, NR, psi_in_l=pl, psi_in_r=pr)
    np.testing.assert_allclose(auto, ref, **TOL)


@pytest.mark.parametrize("NR", [1, 3, 8, 33])
def test_kpm3d_batched_outer_orders(NR):
    """The engine's kpm_3d with every left order resident runs the outer
    orders as column blocks of one wide right chain (powers of two, so NC = 11
    gives a partial last batch), with three distinct current operators, against
    the oracle's nested form and the explicit nested blocking; the FP32 form is
    bounded against FP64."""
    H, Jx, Jy, area = models.haldane_model(5, 5, t2=0.2)
    Hn = (H / 4.0).tocsc()
    Jz = ((Jx + 2.0 * Jy) / 3.0).tocsc()  # a third, distinct current
    NC = 11
    pl, pr = phases(31 + NR, H.shape[0], NR), phases(32 + NR, H.shape[0], NR)
    print(f"kpm_3d batched NR={NR} NC={NC}", flush=True)
    ref = oracle.kpm_3d(Hn, Jx, Jy, Jz, NC, pl, pr, 3, 3)
    got = K.kpm_3d(Hn, Jx, Jy, Jz, NC, NR, psi_in_l=pl, psi_in_r=pr)
    np.testing.assert_allclose(got, ref, **TOL)
    nested = K.kpm_3d(Hn, Jx, Jy, Jz, NC, NR, psi_in_l=pl, psi_in_r=pr, arr_size=NC - 1, right_block=4)
    np.testing.assert_allclose(nested, ref, **TOL)
    g32 = K.kpm_3d(Hn, Jx, Jy, Jz, NC, NR, psi_in_l=pl, psi_in_r=pr, precision="f32")
    assert np.max(np.abs(g32 - got)) < 1e-5


@pytest.mark.parametrize("prec,NR", [("f64", 16), ("f32", 32), ("f64", 64), ("f64", 13)])
def test_kpm2d_wide_blocks_vs_oracle(prec, NR):
    """kpm_2d at wider probe blocks (NRp = 16, 32, 64 complex per row; NR = 13
    pads to 16) in every contraction mode, against the oracle and against an
    explicit blocking (arr_size = NC - 1, so the left ring wraps), plus the
    EVEN parity mask."""
    H, Jx, Jy, area = models.haldane_model(8, 8, t2=0.2)
    Hn = (H / 4.0).tocsc()
    NC = 24
    pl, pr = phases(21, H.shape[0], NR), phases(22, H.shape[0], NR)
    print(f"wide-block kpm_2d prec={prec} NR={NR}", flush=True)
    got = K.kpm_2d(Hn, Jx, Jy, NC, NR, psi_in_l=pl, psi_in_r=pr, precision=prec)
    sep = K.kpm_2d(Hn, Jx, Jy, NC, NR, psi_in_l=pl, psi_in_r=pr, precision=prec, arr_size=NC - 1,
                   right_block=NC)
    ref = oracle.kpm_2d(Hn, Jx, Jy, NC, pl, pr, NC, NC)
    tol = TOL if prec == "f64" else dict(rtol=0, atol=2e-5 * np.abs(ref).max())
    np.testing.assert_allclose(got, ref, **tol)
    np.testing.assert_allclose(got, sep, rtol=1e-12, atol=1e-13 * np.abs(ref).max())
    for gemm in ("f32", "bf16x3"):
        g2 = K.kpm_2d(Hn, Jx, Jy, NC, NR, psi_in_l=pl, psi_in_r=pr, precision=prec, gemm=gemm)
        assert np.abs(g2 - ref).max() <= 2e-4 * np.abs(ref).max()
    par = K.kpm_2d(Hn, Jx, Jy, NC, NR, psi_in_l=pl, psi_in_r=pr, precision=prec, moment_parity="EVEN")
    n, m = np.indices(par.shape)
    assert np.all(par[(n + m) % 2 == 1] == 0)
    np.testing.assert_allclose(par[(n + m) % 2 == 0], got[(n + m) % 2 == 0], rtol=1e-12, atol=1e-14)


# ---- FP32 recurrence -------------------------------------------------------------------------

def test_fp32_vectors_error_bound():
    H = (models.graphene(64, 64, disorder=0.5, seed=8) / 3.6).tocsc()
    psi = phases(8, H.shape[0], 16)
    m64 = K.kpm_1d(H, 512, 16, psi_in=psi, avg_output=False)
    m32 = K.kpm_1d(H, 512, 16, psi_in=psi, avg_output=False, precision="f32")
    err = np.max(np.abs(m32.mean(axis=0) - m64.mean(axis=0)))
    print(f"fp32 max |dmu| (avg over probes) = {err:.3e}", flush=True)
    assert err < 1e-6


def test_fp32_kpm2d_gemm_modes():
    H, Jx, Jy, area = models.haldane_model(6, 6, t2=0.2)
    a, b, Hn = K.normalizeH(H, center=True)
    NR, NC = 4, 32
    pl = phases(3, H.shape[0], NR)
    ref = oracle.kpm_2d(Hn, Jx, Jy, NC, pl, pl, NC, NC)
    for prec, gemm, tol in (("f64", "f32", 2e-5), ("f32", "f32", 2e-5), ("f32", "f64", 2e-5),
                            ("f64", "bf16x3", 2e-4), ("f32", "bf16x3", 2e-4)):
        got = K.kpm_2d(Hn, Jx, Jy, NC, NR, psi_in=pl, precision=prec, gemm=gemm)
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print(f"kpm_2d prec={prec} gemm={gemm} rel err={err:.2e}", flush=True)
        assert err < tol


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("NC", [70, 200])
def test_bf16x3_contraction_long_k(prec, NC):
    """The split-BF16 contraction over many K stages and split-K chunks
    (Haldane 32 x 32, NR 16: 32,768 complex per slot) against the FP64 MFMA
    contraction of the same slots: error at the level of the 3-pass split
    (above FP32 rounding alone, so the BF16 path is the one that ran).
    NC = 70 runs the 64 x 64 kernel (tiles of 64 plus a 6-wide remainder),
    NC = 200 the 128 x 128 kernel with FP32 chunk sums (128 + 72)."""
    H, Jx, Jy, area = models.haldane_model(32, 32, t2=0.2)
    Hn = (H / 4.0).tocsc()
    NR = 16
    pl = phases(11, H.shape[0], NR)
    ref = K.kpm_2d(Hn, Jx, Jy, NC, NR, psi_in=pl, precision=prec, gemm="f64")
    got = K.kpm_2d(Hn, Jx, Jy, NC, NR, psi_in=pl, precision=prec, gemm="bf16x3")
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"bf16x3 vs f64 contraction ({prec} slots): rel err {err:.2e}", flush=True)
    assert 1e-9 < err < 1e-4
