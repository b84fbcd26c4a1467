"""The device Lanczos spectral-edge estimate of normalizeH against exact
extremes: dense eigh at small sizes and closed-form band edges (the 1e-3
tolerance of the reference's Arpack call, src/utils/external.jl:63-72)."""
import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
from kpmhip import models
from kpmhip.rescale import spectral_edges

pytestmark = pytest.mark.gpu


def dense_edges(H):
    ev = np.linalg.eigvalsh(H.toarray() if sp.issparse(H) else H)
    return ev.min(), ev.max()


@pytest.mark.parametrize("name", ["anderson", "haldane", "random", "cubic"])
def test_edges_vs_dense(name):
    rng = np.random.default_rng(4)
    if name == "anderson":
        H = models.square_anderson(32, 2.0, seed=5)
    elif name == "haldane":
        H, *_ = models.haldane_model(10, 10, t2=0.2)
    elif name == "cubic":
        H, *_ = models.cubic_model(8, W=6.0, seed=3)
    else:
        A = rng.standard_normal((200, 200)) + 1j * rng.standard_normal((200, 200))
        H = sp.csc_matrix((A + A.conj().T) / 2 + 3.0 * np.eye(200))
    lo, hi = spectral_edges(H)
    elo, ehi = dense_edges(H)
    rho = max(abs(elo), abs(ehi))
    print(f"{name}: lanczos [{lo:.6f}, {hi:.6f}] dense [{elo:.6f}, {ehi:.6f}]", flush=True)
    assert abs(lo - elo) <= 1e-3 * rho and abs(hi - ehi) <= 1e-3 * rho
    # Ritz values never leave the spectrum
    assert lo >= elo - 1e-9 * rho and hi <= ehi + 1e-9 * rho


def test_edges_closed_forms_and_tiny():
    # periodic chain with on-site shift: band [shift - 2t, shift + 2t] (N even)
    H = models.chain(4000, t=1.0, shift=0.7)
    lo, hi = spectral_edges(H)
    assert abs(lo - (0.7 - 2.0)) < 3e-3 and abs(hi - (0.7 + 2.0)) < 3e-3
    # clean graphene torus: band edges exactly -3t, +3t (the Gamma point)
    G = models.graphene(64, 64, t=1.0)
    lo, hi = spectral_edges(G)
    assert abs(lo + 3.0) < 3e-3 and abs(hi - 3.0) < 3e-3
    # 1x1 and 3x3: Krylov space exhausted, exact
    lo, hi = spectral_edges(sp.csc_matrix(np.array([[0.37 + 0j]])))
    assert abs(lo - 0.37) < 1e-14 and abs(hi - 0.37) < 1e-14
    M = np.array([[1.0, 0.5, 0], [0.5, -2.0, 0.25j], [0, -0.25j, 0.3]])
    lo, hi = spectral_edges(sp.csc_matrix(M))
    ev = np.linalg.eigvalsh(M)
    assert abs(lo - ev[0]) < 1e-12 and abs(hi - ev[-1]) < 1e-12


def test_normalizeH_lanczos_vs_arpack():
    H = models.square_anderson(40, 3.0, seed=9) + 0.4 * sp.identity(1600)
    a1, b1, Hn1 = K.normalizeH(H, center=True)
    a2, b2, Hn2 = K.normalizeH(H, center=True, method="arpack")
    assert abs(a1 - a2) <= 1e-3 * a2 and abs(b1 - b2) <= 1e-3 * a2
    ev = np.linalg.eigvalsh(Hn1.toarray())
    assert ev.min() > -1 and ev.max() < 1
    a3, Hn3 = K.normalizeH(H)
    a4, _ = K.normalizeH(H, method="arpack")
    assert abs(a3 - a4) <= 1e-3 * a4
    with pytest.raises(K.ArgumentError):
        K.normalizeH(H, method="power")
