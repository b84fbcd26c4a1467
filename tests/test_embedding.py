"""The lattice embedding decides the current operator and the stiffness,
restating test/ssh_invariance_test.jl:50-105 on a 6-cell SSH ring (t1 = 1,
t2 = 0.6, circumference 6).  In the split embedding B sits half a cell after
A; in the co-located one A and B share a position.  The Nambu current picks up
the bond displacement (0.5 vs 0 inside a cell, 0.5 vs 1 across cells), and
the stiffness of the same BdG operator with the same probes is reproducible
for identical geometry and different between embeddings."""
import math
import warnings

import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
from kpmhip.rng import Xoshiro256pp

NCELL, T1, T2, L = 6, 1.0, 0.6, 6.0
N = 2 * NCELL


def A(n):
    return 2 * n


def B(n):
    return 2 * n + 1


def ssh_ring():
    h = sp.lil_matrix((N, N))
    for n in range(NCELL):
        h[A(n), B(n)] = h[B(n), A(n)] = -T1
        m = (n + 1) % NCELL
        h[B(n), A(m)] = h[A(m), B(n)] = -T2
    return h.tocsc()


def positions(split):
    pos = np.zeros((N, 1))
    for n in range(NCELL):
        pos[A(n), 0] = n
        pos[B(n), 0] = n + (0.5 if split else 0.0)
    return pos


def ring_disp(pos):
    def disp(i, j):
        d = pos[i, 0] - pos[j, 0]
        return np.array([d - L * round(d / L)])
    return disp


HS = ssh_ring()
POS_A, POS_B = positions(True), positions(False)
DISP_A, DISP_B = ring_disp(POS_A), ring_disp(POS_B)


def test_embedding_decides_the_current_operator():
    Ja = K.nambu_current_q(HS, POS_A, [0.0], disp=DISP_A).toarray()
    Jb = K.nambu_current_q(HS, POS_B, [0.0], disp=DISP_B).toarray()
    for n in range(NCELL):
        a, b = A(n), B(n)
        assert Jb[a, b] == 0 and Jb[a + N, b + N] == 0
        assert abs(Ja[a, b] - 0.5 * T1) <= 1e-14 and abs(Ja[a + N, b + N] - 0.5 * T1) <= 1e-14
    b1, a2 = B(0), A(1)  # split embedding d = -0.5, co-located d = -1
    assert abs(Ja[b1, a2] - 0.5 * T2) <= 1e-14 and abs(Jb[b1, a2] - T2) <= 1e-14
    assert abs(Ja[b1 + N, a2 + N] - 0.5 * T2) <= 1e-14 and abs(Jb[b1 + N, a2 + N] - T2) <= 1e-14
    assert abs(Ja[b1, a2] / Jb[b1, a2] - 0.5) <= 1e-14


@pytest.mark.gpu
def test_identical_geometry_identical_response_different_embedding_different():
    op = K.BdGOperator(HS, mu=0.2, U=0.0, n=np.zeros(N), Delta=np.full(N, 0.3),
                       hole_convention="conjugate")
    q = [2 * math.pi / L]

    def stiffness(pos, disp):
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            r = K.superfluid_stiffness(op, pos, q, beta=8.0, eta=0.15, dir=1, disp=disp, NC=256,
                                       NR=6, rng=Xoshiro256pp(7), volume=L, include_diamagnetic=True)
        assert any("longitudinal" in str(x.message) for x in w)
        return r

    r1 = stiffness(POS_A, DISP_A)
    r1b = stiffness(POS_A, DISP_A)
    r2 = stiffness(POS_B, DISP_B)
    print(f"SSH stiffness split {r1.Ds_over_pi_complete} co-located {r2.Ds_over_pi_complete}; "
          f"Dia_SC {r1.Dia_SC} {r2.Dia_SC}", flush=True)
    assert abs(r1.Ds_over_pi_complete - r1b.Ds_over_pi_complete) <= 1e-12 * abs(r1.Ds_over_pi_complete)
    assert abs(r1.Ds_over_pi_complete - r2.Ds_over_pi_complete) > 1e-3
    assert abs(r1.Dia_SC - r2.Dia_SC) > 1e-3
