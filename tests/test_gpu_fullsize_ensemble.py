"""The batched disorder ensemble at full size: 256 cubic Anderson
realisations with L drawn from {16, 24, 32, 40, 48} (11.7M rows in all),
W uniform in [2, 20], each rescaled by its Gershgorin bound, NC = 1024,
NR = 4, one fused launch per step for all realisations, replayed from a HIP
graph (the reference runs kpm_1d! once per realisation,
scripts/anderson_seebeck_ensemble.jl, src/moment.jl:420-492).

- FP32 recurrence, batched and graph-replayed, against the unbatched FP32
  kpm_1d of every realisation.  The recurrence vectors of the two paths are
  the same values (each row's arithmetic is the same); only the grouping of
  the FP64 dot partials differs, so the moments agree to 1e-13.
- FP64 batched run: the first realisation of each of L = 16, 32, 48 against
  the oracle at 1e-12.
- FP32 against FP64, probe-averaged: 1e-6 (the FP32-form bound).
"""
import numpy as np
import pytest

import kpmhip as K
import oracle
from kpmhip import models
from kpmhip.rng import Xoshiro256pp

pytestmark = pytest.mark.gpu

NREAL, NC, NR = 256, 1024, 4


@pytest.fixture(scope="module")
def ensemble():
    rng = Xoshiro256pp(2025, lanes=4)
    Ls = [(16, 24, 32, 40, 48)[int(5 * rng.rand())] for _ in range(NREAL)]
    Ws = [2.0 + 18.0 * rng.rand() for _ in range(NREAL)]
    print(f"building {NREAL} cubic realisations ...", flush=True)
    Hs = []
    for i, (L, W) in enumerate(zip(Ls, Ws)):
        H, *_ = models.cubic_model(L, W=W, seed=1000 + i)
        Hs.append((H / (6.0 + W / 2 + 0.1)).tocsc())
        if i % 64 == 63:
            print(f"  {i + 1} built", flush=True)
    # probes drawn on the device (the lane-parallel xoshiro256++ stream), one
    # host copy per realisation feeds every path
    psis = [K.DeviceBlock(H.shape[0], NR).random_phase(5000 + i).download()
            for i, H in enumerate(Hs)]
    return Ls, Hs, psis


def test_ensemble_batched_fp32_equals_unbatched(ensemble):
    Ls, Hs, psis = ensemble
    B = K.BatchOperator(Hs)
    B.upload(psis, NR, "f32")
    print(f"batched FP32, graph: {B.info()}", flush=True)
    eager = B.run(NC)
    g1 = B.run(NC, use_graph=True)
    g2 = B.run(NC, use_graph=True)
    np.testing.assert_array_equal(g1, eager)
    np.testing.assert_array_equal(g2, eager)
    worst = 0.0
    for i, (H, p) in enumerate(zip(Hs, psis)):
        un = K.kpm_1d(H, NC, NR, psi_in=p, avg_output=False, precision="f32")
        worst = max(worst, float(np.max(np.abs(un - g1[i]))))
    print(f"  max|batched - unbatched| over {NREAL} realisations = {worst:.3e}", flush=True)
    assert worst <= 1e-13


def test_ensemble_fp64_vs_oracle_and_fp32(ensemble):
    Ls, Hs, psis = ensemble
    B = K.BatchOperator(Hs)
    m64 = K.kpm_batch_1d(B, NC, NR, psis, avg_output=False, use_graph=True)
    m32 = K.kpm_batch_1d(B, NC, NR, psis, precision="f32", avg_output=False, use_graph=True)
    d = np.max(np.abs(m32.mean(axis=1) - m64.mean(axis=1)))
    print(f"ensemble FP32 vs FP64 probe-averaged max|dmu| = {d:.3e}", flush=True)
    assert d <= 1e-6
    for L in (16, 32, 48):
        i = Ls.index(L)
        print(f"realisation {i} (L={L}) vs oracle", flush=True)
        ref = oracle.kpm_1d(Hs[i], NC, psis[i])
        np.testing.assert_allclose(m64[i], ref, rtol=0, atol=1e-12, err_msg=f"L={L}")
