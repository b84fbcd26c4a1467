"""Random-phase probes drawn on the device (kpm_block_random_phase) against
the host generator: the lane-parallel xoshiro256++ integer stream is shared
bit for bit, so the phases agree to the last ulp of sin/cos and the
normalisation to rounding (src/frontend.jl:212-216, src/moment.jl:493-497)."""
import numpy as np
import pytest

import kpmhip as K
from kpmhip import models
from kpmhip.rng import Xoshiro256pp

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("NH,NR,lanes", [(1000, 5, 1024), (777, 3, 7), (300, 130, 1024), (1, 1, 4)])
def test_device_probes_match_host_stream(NH, NR, lanes):
    print(f"probes NH={NH} NR={NR} lanes={lanes}", flush=True)
    seed = 1234 + NH
    host = K.random_phase_vectors(Xoshiro256pp(seed, lanes=lanes), NH, NR)
    dev = K.DeviceBlock(NH, NR, "f64").random_phase(seed, lanes).download()
    np.testing.assert_allclose(dev, host, rtol=0, atol=4e-16)
    np.testing.assert_allclose(np.linalg.norm(dev, axis=0), 1.0, rtol=0, atol=1e-14)
    d32 = K.DeviceBlock(NH, NR, "f32").random_phase(seed, lanes).download()
    np.testing.assert_allclose(d32, host, rtol=0, atol=2e-7)


def test_kpm1d_random_probes_on_device():
    H = (models.graphene(16, 16, disorder=0.5, seed=3) / 3.6).tocsc()
    mu = K.kpm_1d(H, 64, 8)
    assert abs(mu[0] - 1.0) < 1e-13          # unit probes: mu_0 = 1
    assert np.all(np.abs(mu) <= 1.0 + 1e-12)  # |T_n| <= 1 on the rescaled spectrum
    mu2 = K.kpm_1d(H, 64, 8)
    assert np.max(np.abs(mu - mu2)) > 0       # fresh probes per call
