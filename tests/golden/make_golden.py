"""Writes the golden fixtures of tests/golden/*.npz.

The reference is Julia and cannot run here, and its own tests hold no
numeric fixtures for this path (they recompute dense references in Julia).
These fixtures are therefore made from the same independent dense references
the reference's tests use -- explicit T_n(H) matrices
(test/kpm2d_blocking_test.jl:25-58) and exact diagonalisation
(test/dos_test.jl:39-43) -- on small seeded inputs, in extended-precision-free
plain numpy FP64.  Inputs come from the build's xoshiro256++ generator.
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "kpm.jl_amd"))

from kpmhip.rng import Xoshiro256pp  # noqa: E402


def herm(rng, n, scale):
    a = (rng.rand(n, n) - 0.5) + 1j * (rng.rand(n, n) - 0.5)
    h = (a + a.conj().T) / 2
    return h / (scale * np.max(np.abs(np.linalg.eigvalsh(h))))


def phases(rng, n, nr):
    p = np.exp(2j * np.pi * rng.rand(n, nr))
    return p / np.linalg.norm(p, axis=0)


def cheb(H, NC):
    T = [np.eye(H.shape[0], dtype=complex), H.astype(complex)]
    while len(T) < NC:
        T.append(2 * H @ T[-1] - T[-2])
    return T[:NC]


def sparsify(rng, H, keep=0.35):
    """zero out a symmetric random pattern (keeps Hermiticity)."""
    n = H.shape[0]
    mask = rng.rand(n, n) < keep
    mask = mask | mask.T | np.eye(n, dtype=bool)
    return np.where(mask, H, 0)


def main():
    rng = Xoshiro256pp(2024, lanes=8)
    # g1: kpm_1d equal / bra-ket / current
    NH, NR, NC = 24, 3, 16
    H = sparsify(rng, herm(rng, NH, 1.2))
    J = (rng.rand(NH, NH) - 0.5) + 1j * (rng.rand(NH, NH) - 0.5)
    J = sparsify(rng, J, 0.3)
    psi, pl, pr = phases(rng, NH, NR), phases(rng, NH, NR), phases(rng, NH, NR)
    T = cheb(H, NC)
    mu_eq = np.array([[np.vdot(psi[:, i], T[n] @ psi[:, i]) for n in range(NC)] for i in range(NR)])
    mu_lr = np.array([[np.vdot(pl[:, i], T[n] @ pr[:, i]) for n in range(NC)] for i in range(NR)])
    mu_cur = np.array([[np.vdot(psi[:, i], J @ (T[n] @ psi[:, i])) for n in range(NC)]
                       for i in range(NR)])
    lam, V = np.linalg.eigh(H)
    w = np.abs(V.conj().T @ psi) ** 2
    mu_ed = (np.cos(np.arange(NC)[:, None] * np.arccos(np.clip(lam, -1, 1))[None, :]) @ w).T
    np.savez(os.path.join(HERE, "kpm1d_dense.npz"), H=H, J=J, psi=psi, psi_l=pl, psi_r=pr, NC=NC,
             mu_eq=mu_eq, mu_lr=mu_lr, mu_cur=mu_cur, mu_ed=mu_ed)
    # g2: kpm_2d
    NH, NR, NC = 16, 2, 9
    H = sparsify(rng, herm(rng, NH, 1.2))
    Ja = sparsify(rng, herm(rng, NH, 1.0), 0.4)
    Jb = sparsify(rng, herm(rng, NH, 1.0), 0.4)
    pl, pr = phases(rng, NH, NR), phases(rng, NH, NR)
    T = cheb(H, NC)
    mu2 = np.zeros((NC, NC), dtype=complex)
    for m in range(NC):
        for n in range(NC):
            for i in range(NR):
                mu2[n, m] += np.vdot(T[m] @ pl[:, i], Jb @ T[n] @ Ja @ pr[:, i]) / NR
    np.savez(os.path.join(HERE, "kpm2d_dense.npz"), H=H, Ja=Ja, Jb=Jb, psi_l=pl, psi_r=pr, NC=NC,
             mu=mu2)
    # g3: kpm_3d
    NH, NR, NC = 8, 2, 4
    H = herm(rng, NH, 1.2)
    Ja, Jb, Jg = herm(rng, NH, 1.0), herm(rng, NH, 1.0), herm(rng, NH, 1.0)
    pl, pr = phases(rng, NH, NR), phases(rng, NH, NR)
    T = cheb(H, NC)
    mu3 = np.zeros((NC, NC, NC), dtype=complex)
    for n1 in range(NC):
        for n3 in range(NC):
            for n2 in range(NC):
                for i in range(NR):
                    mu3[n2, n3, n1] += np.vdot(T[n3] @ pl[:, i],
                                               Jb @ T[n2] @ Jg @ T[n1] @ Ja @ pr[:, i]) / NR
    np.savez(os.path.join(HERE, "kpm3d_dense.npz"), H=H, Ja=Ja, Jb=Jb, Jg=Jg, psi_l=pl, psi_r=pr,
             NC=NC, mu=mu3)
    # g4: chebyshev_action (test/chebyshev_action_test.jl:20-41 reference form)
    NH, NR, NC, K = 20, 2, 12, 3
    H = sparsify(rng, herm(rng, NH, 1.05))
    V = (rng.rand(NH, NR) - 0.5) + 1j * (rng.rand(NH, NR) - 0.5)
    Cc = (rng.rand(NC, K) - 0.5) + 1j * (rng.rand(NC, K) - 0.5)
    T = cheb(H, NC)
    out = np.zeros((NH, NR, K), dtype=complex)
    for k in range(K):
        for n in range(NC):
            out[:, :, k] += Cc[n, k] * (T[n] @ V)
    np.savez(os.path.join(HERE, "action_dense.npz"), H=H, V=V, C=Cc, out=out)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
