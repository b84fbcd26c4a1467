"""The reference's own GPU-vs-CPU checks, with the oracle as the CPU side:
test/gpu_smoke.jl:50-90 (kpm_1d on a 4096-site chain with explicit
random-phase probes, equal-vector and left/right with odd NC = 257) and the
device branch of test/optional_cuda_test.jl:31-37 (a real diagonal 2 x 2
operator through the complex workspace: mu_0 = 1).  The reference asserts
atol 1e-8 against its CPU path; the oracle comparison here is 1e-12."""
import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
import oracle
from kpmhip.rng import Xoshiro256pp

pytestmark = pytest.mark.gpu


def chain(NH):
    off = np.full(NH - 1, -0.45 + 0j)
    return sp.diags([off, off], [1, -1], shape=(NH, NH), format="csc")


def probes(seed, NH, NR):
    r = Xoshiro256pp(seed)
    psi = np.exp(2j * np.pi * r.rand(NH * NR).reshape(NH, NR, order="F"))
    return psi / np.linalg.norm(psi, axis=0)


def test_device_active_and_trivial_real_operator():
    assert K.activate(0) is True and K.whichcore() is True
    H = sp.csc_matrix(([0.3, -0.3], ([0, 1], [0, 1])), shape=(2, 2))
    mu = K.kpm_1d(H, 8, 2)
    assert abs(mu[0] - 1.0) <= 1e-6


def test_kpm1d_chain_gpu_equals_cpu():
    NH, NC, NR = 4096, 256, 8
    H = chain(NH)
    psi = probes(42, NH, NR)
    print("gpu_smoke chain kpm_1d NH=4096 NC=256 NR=8", flush=True)
    got = K.kpm_1d(H, NC, NR, psi_in=psi.copy())
    ref = oracle.kpm_1d(H, NC, psi).mean(axis=0).real
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)


def test_kpm1d_left_right_odd_nc_gpu_equals_cpu():
    NH, NC, NR = 4096, 257, 4
    H = chain(NH)
    pl, pr = probes(43, NH, NR), probes(44, NH, NR)
    print("gpu_smoke chain left/right NC=257 NR=4", flush=True)
    mu = np.zeros((NR, NC), dtype=np.complex128)
    K.kpm_1d_(H, NC, NR, NH, mu, pl, pr)
    ref = oracle.kpm_1d_lr(H, NC, pl, pr)
    np.testing.assert_allclose(mu, ref, rtol=0, atol=1e-12)
