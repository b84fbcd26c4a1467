"""Kubo–Bastin conductivity from device moment tables against ED, restating
test/kubo_bastin_test.jl:58-152 on the 6 x 6 Haldane torus (t2 = 0.2, D = 72)
with identity probes (an exact trace, so only kernel broadening and
quadrature remain), NC = 96:

- a rigidly shifted spectrum (H + 2.5) keeps the Hall plateau at E_F = 2.5:
  the centred rescale recovers b within 0.05 and sigma_xy matches the
  unshifted T = 0 ED value within 0.03;
- the trivial phase (m = 1.6) reproduces its non-quantised ED value within
  0.03;
- sigma_xx vanishes in the gap (< 1e-4) and, in the band (E_F = -2), matches
  the Lorentzian-broadened ED at eta = a lambda sqrt(1 - x_F^2) / NC within
  15% for Lorentz kernels lambda = 3, 4.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
from kpmhip import models
import oracle as O

pytestmark = pytest.mark.gpu

NC = 96
H, JX, JY, AREA = models.haldane_model(6, 6, t=1.0, t2=0.2)
D = H.shape[0]
PSI = np.eye(D, dtype=np.complex128)


def test_shifted_spectrum_keeps_the_plateau():
    shift = 2.5
    Hs = (H + shift * sp.identity(D)).tocsc()
    a_s, b_s, Hs_n = K.normalizeH(Hs, center=True)
    assert abs(b_s - shift) < 0.05
    mu = K.kpm_2d(Hs_n, JX, JY, NC, D, D, psi_in=PSI)
    sig = K.kubo_bastin_cond_raw(mu, a_s, shift, b=b_s, NH=D, area=AREA)
    ed = O.ed_hall_conductivity_T0(H, JX, JY, AREA, Ef=0.0)
    print(f"shifted plateau {sig} vs ED {ed}", flush=True)
    assert abs(sig - ed) < 0.03


def test_trivial_phase_tracks_ed():
    Ht, Jxt, Jyt, at = models.haldane_model(6, 6, t=1.0, t2=0.2, m=1.6)
    ed = O.ed_hall_conductivity_T0(Ht, Jxt, Jyt, at, Ef=0.0)
    assert abs(ed) < 0.2
    a_t, b_t, Ht_n = K.normalizeH(Ht, center=True)
    mu = K.kpm_2d(Ht_n, Jxt, Jyt, NC, D, D, psi_in=PSI)
    sig = K.kubo_bastin_cond_raw(mu, a_t, 0.0, b=b_t, NH=D, area=at)
    print(f"trivial phase {sig} vs ED {ed}", flush=True)
    assert abs(sig - ed) < 0.03


def test_sigma_xx_gap_and_band():
    a, b, Hn = K.normalizeH(H, center=True)
    mu = K.kpm_2d(Hn, JX, JX, NC, D, D, psi_in=PSI)
    assert abs(K.kubo_bastin_cond_raw(mu, a, 0.0, b=b, NH=D, area=AREA)) < 1e-4
    Ef = -2.0
    xf = (Ef - b) / a
    for lam in (3.0, 4.0):
        eta = a * lam * np.sqrt(1 - xf ** 2) / NC
        sig = K.kubo_bastin_cond_raw(mu, a, Ef, b=b, NH=D, area=AREA, kernel=K.LorentzKernels(lam))
        ed = O.ed_kubo_bastin_broadened(H, JX, JX, AREA, Ef=Ef, eta=eta)
        print(f"sigma_xx lambda={lam}: KPM {sig} vs broadened ED {ed}", flush=True)
        assert abs(sig - ed) <= 0.15 * abs(ed)
