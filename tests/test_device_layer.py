"""The device layer's host-side contract, restating test/optional_cuda_test.jl
for a machine without a usable device: whichcore() is false, maybe_to_host
returns host arrays and host sparse matrices as the same objects, and the
engines refuse to run (no CPU fallback) instead of computing on the host."""
import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
from kpmhip import device as D


def test_whichcore_follows_the_active_device():
    saved = D.ACTIVE_DEVICE[0]
    try:
        D.ACTIVE_DEVICE[0] = D.CPUDevice()
        assert K.whichcore() is False
        D.ACTIVE_DEVICE[0] = D.HIPDevice(0)
        assert K.whichcore() is True
    finally:
        D.ACTIVE_DEVICE[0] = saved


def test_maybe_to_host_passes_host_objects_through():
    a = np.array([1.0, 2.0])
    assert K.maybe_to_host(a) is a
    s = sp.csc_matrix(([1.0, 2.0], ([0, 1], [0, 1])), shape=(2, 2))
    assert K.maybe_to_host(s) is s


def test_no_device_no_host_fallback():
    try:
        n = D.device_count()
    except K.KPMHipMissing:
        n = 0
    if n > 0:
        pytest.skip("a device is visible; the GPU tests cover this path")
    assert K.activate(0) is False
    H = sp.csc_matrix(([0.3, -0.3], ([0, 1], [0, 1])), shape=(2, 2))
    with pytest.raises(K.KPMHipMissing, match="no CPU fallback"):
        K.kpm_1d(H, 8, 2)
