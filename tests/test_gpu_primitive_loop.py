"""The reference's device loop of kpm_1d! (src/moment.jl:440-489), replayed
call for call through the C ABI that the ROCm extension's method table binds
(ext/KPMCUDAExt.jl:84-150):

    α_all[:, :, 1] = psi                      kpm_block_upload
    mul!(α2, H, α1)                           kpm_spmm
    mu_acc = moment_accumulator(α_all, mu)    kpm_accum_create
    columnwise_dot!(mu_acc, 1, α1, α1)        kpm_accum_dot
    columnwise_dot!(mu_acc, 2, α1, α2)        kpm_accum_dot
    for n in 2:NC/2
        chebyshev_iter_single(H, α[ipp], α[ip])          kpm_cheb_step
        columnwise_dot!(mu_acc, 2n-1, α[ip], α[ip]; 2, 1, -1)   kpm_accum_dot
        columnwise_dot!(mu_acc, 2n,   α[ip], α[ipp]; 2, 2, -1)  kpm_accum_dot
    copy_moment_accumulator!(mu, mu_acc)      kpm_accum_download

The step reduces both dots in its own pass, so the loop runs one recurrence
launch per step, no dot pass and one host synchronisation in all, and its
moments are bitwise those of the fused engine kpm_1d (same launches, same
fixed-order reductions) and equal to the oracle.
"""
import ctypes as C
import time

import numpy as np
import pytest

import kpmhip as K
import oracle
from kpmhip import _lib, models
from kpmhip.rng import Xoshiro256pp

pytestmark = pytest.mark.gpu


def phases(seed, n, nr):
    return K.random_phase_vectors(Xoshiro256pp(seed), n, nr)


def replay_kpm_1d(ctx, Hd, a1, a2, NC, NR, psi=None):
    """moment.jl:440-489 through the C ABI; a1, a2 are the two workspace
    blocks (α_all[:, :, 1:2]).  Returns mu_all (NR x NC)."""
    lib, h = ctx.lib, ctx.handle
    if psi is not None:
        a1.upload(psi)
    _lib.check(lib.kpm_spmm(h, Hd.handle, 1.0, a1.handle, 0.0, a2.handle), "mul!")
    acc = C.c_void_p()
    _lib.check(lib.kpm_accum_create(h, NR, NC, C.byref(acc)), "moment_accumulator")
    try:
        _lib.check(lib.kpm_accum_dot(h, acc, 1, a1.handle, a1.handle, 1.0, 0, 0.0), "dot")
        _lib.check(lib.kpm_accum_dot(h, acc, 2, a1.handle, a2.handle, 1.0, 0, 0.0), "dot")
        views = [a1, a2]
        ip, ipp = 1, 0
        for n in range(2, NC // 2 + 1):
            _lib.check(lib.kpm_cheb_step(h, Hd.handle, views[ipp].handle, views[ip].handle,
                                         views[ipp].handle), "chebyshev_iter_single")
            _lib.check(lib.kpm_accum_dot(h, acc, 2 * n - 1, views[ip].handle, views[ip].handle,
                                         2.0, 1, -1.0), "dot")
            _lib.check(lib.kpm_accum_dot(h, acc, 2 * n, views[ip].handle, views[ipp].handle,
                                         2.0, 2, -1.0), "dot")
            ip, ipp = 1 - ip, 1 - ipp
        mu = np.zeros((NR, NC), dtype=np.complex128, order="F")
        _lib.check(lib.kpm_accum_download(h, acc, _lib.dptr(mu)), "copy_moment_accumulator!")
    finally:
        lib.kpm_accum_destroy(acc)
    return mu


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_replayed_loop_equals_engine_and_oracle(prec):
    H = (models.graphene(16, 12, disorder=0.5, seed=3) / 3.6).tocsc()
    NH, NR, NC = H.shape[0], 6, 64
    psi = phases(5, NH, NR)
    ctx = K.get_context()
    Hd = K.to_device(H)
    a1, a2 = K.DeviceBlock(NH, NR, prec), K.DeviceBlock(NH, NR, prec)
    ctx.counters(reset=True)
    mu = replay_kpm_1d(ctx, Hd, a1, a2, NC, NR, psi)
    cnt = ctx.counters()
    print(f"replay {prec}: counters {cnt}", flush=True)
    assert cnt["step"] == NC // 2  # one fused pass per step: mul! + NC/2 - 1 steps
    assert cnt["dot"] == 0          # no dot re-reads a block
    assert cnt["combine"] == 1      # only the last step's columns take a launch of their own
    assert cnt["sync"] == 1         # copy_moment_accumulator!
    eng = K.kpm_1d(Hd, NC, NR, psi_in=psi, avg_output=False, precision=prec)
    np.testing.assert_array_equal(mu, eng)
    ref = oracle.kpm_1d(H, NC, psi)
    if prec == "f64":
        np.testing.assert_allclose(mu, ref, rtol=1e-12, atol=1e-12)
    else:
        assert np.max(np.abs(mu - ref)) < 1e-5


def test_dots_of_other_blocks_and_stale_cache():
    """columnwise_dot! of blocks that are not the last step's pair takes a
    dot pass; a block written after the step (here an upload) makes the
    cached dots stale, so the dot is recomputed from the new contents."""
    H = (models.graphene(10, 10, disorder=0.3, seed=4) / 3.5).tocsc()
    NH, NR = H.shape[0], 3
    ctx = K.get_context()
    lib, h = ctx.lib, ctx.handle
    Hd = K.to_device(H)
    s0, s1, s2 = phases(11, NH, NR), phases(12, NH, NR), phases(13, NH, NR)
    b0, b1, b2 = (K.DeviceBlock.from_host(s) for s in (s0, s1, s2))
    acc = K.DeviceAccumulator(NR, 4)
    ctx.counters(reset=True)
    K.chebyshev_iter_single(Hd, b0, b1)            # b0 <- 2 H b1 - b0, dots <b1|b1>, <b1|b0>
    v0 = 2 * (H @ s1) - s0
    acc.dot(1, b1, b0)                              # cached
    acc.dot(2, b0, b2)                              # not a dot of the step: dot pass
    assert ctx.counters()["dot"] == 1
    b1.upload(s2)                                   # the step's V_p changes
    acc.dot(3, b1, b1, alpha=2.0, beta_col=1, beta_scale=-1.0)  # must not use the cache
    assert ctx.counters()["dot"] == 2
    got = acc.download()
    c = lambda u, v: np.einsum("ir,ir->r", u.conj(), v)
    np.testing.assert_allclose(got[:, 0], c(s1, v0), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(got[:, 1], c(v0, s2), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(got[:, 2], 2 * c(s2, s2) - c(s1, v0), rtol=1e-12, atol=1e-12)
    assert np.all(got[:, 3] == 0)
    del lib, h


def test_host_target_columnwise_dot_uses_the_step():
    """columnwise_dot! into a host target after a step: the same values as
    the device accumulator, with no dot pass."""
    H = (models.graphene(8, 8, disorder=0.3, seed=6) / 3.5).tocsc()
    NH, NR = H.shape[0], 4
    ctx = K.get_context()
    s0, s1 = phases(21, NH, NR), phases(22, NH, NR)
    b0, b1 = K.DeviceBlock.from_host(s0), K.DeviceBlock.from_host(s1)
    ctx.counters(reset=True)
    K.chebyshev_iter_single(H, b0, b1)
    tgt = np.zeros((NR, 2), dtype=np.complex128, order="F")
    K.columnwise_dot(tgt, 1, b1, b1)
    K.columnwise_dot(tgt, 2, b1, b0, alpha=2.0, beta_col=1, beta_scale=-1.0)
    assert ctx.counters()["dot"] == 0
    v0 = 2 * (H @ s1) - s0
    d = np.einsum("ir,ir->r", s1.conj(), s1)
    np.testing.assert_allclose(tgt[:, 0], d, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(tgt[:, 1], 2 * np.einsum("ir,ir->r", s1.conj(), v0) - d,
                               rtol=1e-12, atol=1e-12)
    with pytest.raises(K.ArgumentError):
        K.columnwise_dot(tgt, 1, b1, b1, alpha=1.0 + 1.0j)


def test_replayed_loop_time_within_five_percent_of_engine():
    """Graphene 512 x 512 x 2 (NH = 524,288), NR = 32, FP64, NC = 256: the
    primitive loop, driven from Python one call at a time, runs within 5% of
    the fused engine on the same device-resident probes (best of 3 each)."""
    H = (models.graphene(512, 512, disorder=0.5, seed=7) / 3.6).tocsc()
    NH, NR, NC = H.shape[0], 32, 256
    ctx = K.get_context()
    Hd = K.to_device(H)
    psi = K.DeviceBlock(NH, NR).random_phase(9)
    a1, a2 = K.DeviceBlock(NH, NR), K.DeviceBlock(NH, NR)
    mu_e = np.zeros((NR, NC), dtype=np.complex128, order="F")

    def engine():
        _lib.check(ctx.lib.kpm_1d_block(ctx.handle, Hd.handle, NC, psi.handle, _lib.dptr(mu_e)))
        return mu_e

    def loop():
        a1.copy_from(psi)
        return replay_kpm_1d(ctx, Hd, a1, a2, NC, NR)

    def best(fn):
        fn()
        ts = []
        for _ in range(3):
            ctx.synchronize()
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    te, tl = best(engine), best(loop)
    print(f"kpm_1d engine {te * 1e3:.2f} ms, replayed loop {tl * 1e3:.2f} ms "
          f"({tl / te:.3f}x)", flush=True)
    np.testing.assert_array_equal(loop(), engine())
    assert tl <= 1.05 * te
