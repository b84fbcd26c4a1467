"""Bond (d-wave) pairing through the device stiffness driver, restating
test/bond_stiffness_test.jl on the 3 x 3 periodic square lattice.

The pairing matrix follows an even d-wave channel: D_ij = D_ji = Δ w_ij on
every nearest-neighbour bond, with w = +1 on x bonds and -1 on y bonds
(pairing_matrix, src/applications/bdg.jl:493-549). The operator is
BdGOperator(h; mu, U, n, D, hole_convention = conjugate), and the tables come
from the engine with identity probes at NC = 256.

Checks:
- Zero gap: the complete and the paramagnetic stiffness vanish below 1e-10.
- Rigid d-wave: the complete stiffness equals (2 g_J / V)(F''_SC - F''_N) of
  the ED free energy under a Peierls vector potential, within 2e-2. The
  curvature is step-converged to 1e-3, and the diamagnetic term matters
  (> 1e-4).
- Paramagnetic response: Pi_SC against the dense Lehmann sum on the Nambu
  matrix [xi D; D^H -conj(xi)], within 5%.
- U(1) rigidity: D -> e^{0.4 i} D changes neither stiffness beyond 1e-10.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
from kpmhip import models
import oracle as O

pytestmark = pytest.mark.gpu

H, POS, DISP = models.square_model(3, 3)
N = 9
QY = np.array([0.0, 2 * math.pi / 3])
BETA, VOLUME, G_J, MU = 10.0, 9.0, 1.0, -0.7
DENSITY = np.full(N, 0.4)
U = np.full(N, 2.0)
PSI = np.eye(2 * N, dtype=np.complex128)


def dwave_bonds():
    Hc = sp.coo_matrix(H)
    bonds, weights = [], []
    for i, j in zip(Hc.row, Hc.col):
        if i >= j:
            continue
        d = np.asarray(DISP(i, j) if callable(DISP) else DISP[(i, j)])
        isx, isy = abs(d[0]) == 1, abs(d[1]) == 1
        assert isx != isy
        bonds.append((i, j))
        weights.append(1.0 if isx else -1.0)
    return bonds, weights


def pairing_matrix(amplitude):
    bonds, weights = dwave_bonds()
    I, J, V = [], [], []
    for (i, j), w in zip(bonds, weights):
        I += [i, j]
        J += [j, i]
        V += [amplitude * w, amplitude * w]  # even channel
    return sp.csc_matrix((np.asarray(V, dtype=np.complex128), (I, J)), shape=(N, N))


def response(D):
    op = K.BdGOperator(H, mu=MU, U=U, n=DENSITY, D=D, hole_convention="conjugate")
    return K.superfluid_stiffness(op, POS, QY, beta=BETA, eta=0.0, NC=256, psi_in=PSI,
                                  volume=VOLUME, g_J=G_J, disp=DISP, include_diamagnetic=True)


def test_zero_gap_cancellation_with_structural_bond_zeros():
    r0 = response(pairing_matrix(0.0))
    assert abs(r0.Ds_over_pi_complete) < 1e-10
    assert abs(r0.Ds_over_pi) < 1e-10


def test_rigid_dwave_stiffness_vs_free_energy_curvature():
    D = pairing_matrix(0.45)
    r = response(D)
    Dd, Dz = D.toarray(), np.zeros((N, N), dtype=np.complex128)

    def F(Dm, A):
        return O.ed_bdg_free_energy(H, POS, DISP, MU, U, DENSITY, Dm, A, q=QY, dir=1, beta=BETA)

    def curv(Dm, d):
        return (F(Dm, d) - 2 * F(Dm, 0.0) + F(Dm, -d)) / d ** 2

    anchor = (2 * G_J / VOLUME) * (curv(Dd, 1e-3) - curv(Dz, 1e-3))
    half = (2 * G_J / VOLUME) * (curv(Dd, 5e-4) - curv(Dz, 5e-4))
    print(f"d-wave complete stiffness {r.Ds_over_pi_complete} vs curvature {anchor} "
          f"(half step {half}); paramagnetic {r.Ds_over_pi}", flush=True)
    assert abs(r.Ds_over_pi_complete - anchor) <= 2e-2 * abs(anchor)
    assert abs(anchor - half) <= 1e-3 * abs(anchor)
    assert abs(r.Ds_over_pi_complete - r.Ds_over_pi) > 1e-4


def test_paramagnetic_response_vs_dense_ed():
    D = pairing_matrix(0.45)
    r = response(D)
    xi = H.toarray() - MU * np.eye(N) - np.diag(U * DENSITY / 2)
    Dd = D.toarray()
    Hn = np.block([[xi, Dd], [Dd.conj().T, -np.conj(xi)]])
    Jq = K.nambu_current_q(H, POS, QY, dir=1, disp=DISP, hole_convention="conjugate")
    Jmq = K.nambu_current_q(H, POS, -QY, dir=1, disp=DISP, hole_convention="conjugate")
    ed = G_J * O.ed_two_energy_response(Hn, Jq, Jmq, beta=BETA, eta=0.0) / VOLUME
    print(f"d-wave Pi_SC {r.Pi_SC} vs ED {ed}", flush=True)
    assert abs(np.real(r.Pi_SC) - ed.real) <= 0.05 * abs(ed.real)


def test_u1_rigidity():
    D = pairing_matrix(0.45)
    r = response(D)
    rp = response(np.exp(0.4j) * D)
    assert abs(rp.Ds_over_pi - r.Ds_over_pi) <= 1e-10
    assert abs(rp.Ds_over_pi_complete - r.Ds_over_pi_complete) <= 1e-10
