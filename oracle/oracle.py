"""The oracle -- TEST INFRASTRUCTURE ONLY.

Python access to the C restatement of KPM.jl's CPU moment engine
(oracle/kpm_oracle.c -> oracle/liboracle.so) plus independent numpy dense
references used to pin it.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module; the product path
(kpmhip + libkpm_hip.so) never does.

It also restates the reference's host reconstruction heads in numpy FP64
(``dos`` after src/applications/dos.jl:76-149 and src/utils/chebyshev_lintrans.jl,
``kubo_bastin_cond`` after src/applications/conductivity.jl:101-157), against
which the device reconstruction kernels are checked.

Dense references (numpy, tiny sizes) restate the reference's own test
references: _chebyshev_matrices / _kpm2d_dense_reference /
_kpm3d_dense_reference (test/kpm2d_blocking_test.jl:25-58) and the ED moment
identity mu_n = sum_i |<v_i|psi>|^2 T_n(lambda_i) (test/dos_test.jl:39-43).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int64)
_i = C.c_int64

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        sigs = {
            "kpm_oracle_set_threads": (None, [C.c_int]),
            "kpm_oracle_get_threads": (C.c_int, []),
            "kpm_oracle_1d": (C.c_int, [_i, _pi, _pi, _pd, _i, _i, _pd, _pd]),
            "kpm_oracle_1d_lr": (C.c_int, [_i, _pi, _pi, _pd, _i, _i, _pd, _pd, _pd]),
            "kpm_oracle_1d_current": (C.c_int, [_i, _pi, _pi, _pd, _pi, _pi, _pd, _i, _i, _pd, _pd]),
            "kpm_oracle_2d_blocking": (C.c_int, [_i, _i, _i, _i, _i, _i, _pi, _pi, _pi]),
            "kpm_oracle_2d": (C.c_int, [_i, _pi, _pi, _pd, _pi, _pi, _pd, _pi, _pi, _pd, _i, _i,
                                        _pd, _pd, _i, _i, C.c_int, C.c_int, _pd]),
            "kpm_oracle_3d": (C.c_int, [_i, _pi, _pi, _pd, _pi, _pi, _pd, _pi, _pi, _pd, _pi, _pi,
                                        _pd, _i, _i, _pd, _pd, _i, _i, _pd]),
            "kpm_oracle_cheb_iter": (C.c_int, [_i, _pi, _pi, _pd, _i, _i, _pd]),
            "kpm_oracle_1d_steps": (C.c_int, [_i, _pi, _pi, _pd, _i, _i, _pd, _pd, _pd]),
            "kpm_oracle_1d_tangent": (C.c_int, [_i, _pi, _pi, _pd, _pd, _i, _i, _pd, _pd, _pd]),
            "kpm_oracle_cheb_action": (C.c_int, [_i, _pi, _pi, _pd, _i, _i, _i, _pd, _pd, _pd, _i,
                                                 _pd]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def set_threads(n: int):
    lib().kpm_oracle_set_threads(int(n))


def get_threads() -> int:
    return int(lib().kpm_oracle_get_threads())


class CSC:
    """Julia SparseMatrixCSC arrays (0-based) of a square matrix."""

    def __init__(self, A):
        A = sp.csc_matrix(A, dtype=np.complex128)
        A.sort_indices()
        self.n = A.shape[0]
        self.colptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
        self.rowval = np.ascontiguousarray(A.indices, dtype=np.int64)
        self.nz = np.ascontiguousarray(A.data, dtype=np.complex128)

    def args(self):
        return (self.n, self.colptr.ctypes.data_as(_pi), self.rowval.ctypes.data_as(_pi),
                self.nz.ctypes.data_as(_pd))

    def args3(self):
        return self.args()[1:]


def _dp(a):
    assert a.flags.f_contiguous or a.flags.c_contiguous
    return a.ctypes.data_as(_pd)


def _check(rc, what):
    if rc != 0:
        raise ValueError(f"oracle {what} rejected its arguments (status {rc})")


def _blk(psi):
    a = np.asarray(psi, dtype=np.complex128)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    return np.asfortranarray(a)


def kpm_1d(H, NC: int, psi) -> np.ndarray:
    """kpm_1d! doubling (src/moment.jl:420-492): NR x NC complex."""
    A = CSC(H)
    p = _blk(psi)
    NR = p.shape[1]
    mu = np.zeros((NR, NC), dtype=np.complex128, order="F")
    _check(lib().kpm_oracle_1d(*A.args(), NC, NR, _dp(p), _dp(mu)), "kpm_1d")
    return mu


def kpm_1d_lr(H, NC: int, psi_l, psi_r) -> np.ndarray:
    A = CSC(H)
    pl, pr = _blk(psi_l), _blk(psi_r)
    NR = pl.shape[1]
    mu = np.zeros((NR, NC), dtype=np.complex128, order="F")
    _check(lib().kpm_oracle_1d_lr(*A.args(), NC, NR, _dp(pl), _dp(pr), _dp(mu)), "kpm_1d_lr")
    return mu


def kpm_1d_current(H, J, NC: int, psi) -> np.ndarray:
    A = CSC(H)
    Jc = CSC(J)
    p = _blk(psi)
    NR = p.shape[1]
    mu = np.zeros((NR, NC), dtype=np.complex128, order="F")
    _check(lib().kpm_oracle_1d_current(*A.args(), *Jc.args3(), NC, NR, _dp(p), _dp(mu)),
           "kpm_1d_current")
    return mu


def kpm_1d_tangent(H, dH_data, NC: int, psi):
    """Moments of the direct recurrence and their derivative along dH, whose
    values dH_data follow H's (sorted) CSC order: (mu, dmu), NR x NC each."""
    A = CSC(H)
    d = np.ascontiguousarray(dH_data, dtype=np.complex128)
    if d.shape != A.nz.shape:
        raise ValueError("dH_data must match H's stored entries")
    p = _blk(psi)
    NR = p.shape[1]
    mu = np.zeros((NR, NC), dtype=np.complex128, order="F")
    dmu = np.zeros((NR, NC), dtype=np.complex128, order="F")
    _check(lib().kpm_oracle_1d_tangent(*A.args(), d.ctypes.data_as(_pd), NC, NR, _dp(p), _dp(mu),
                                       _dp(dmu)), "1d_tangent")
    return mu, dmu


def kpm_2d_blocking(NH, NR, NC, budget, arr_size=0, right_block=0):
    KL, KR, by = C.c_int64(), C.c_int64(), C.c_int64()
    rc = lib().kpm_oracle_2d_blocking(NH, NR, NC, budget, arr_size, right_block, C.byref(KL),
                                      C.byref(KR), C.byref(by))
    if rc != 0:
        raise ValueError("ArgumentError")
    return {"arr_size": KL.value, "right_block": KR.value, "bytes": by.value}


def kpm_2d(H, Ja, Jb, NC: int, psi_l, psi_r, K_left: int, K_right: int, parity: int = 0,
           mn_sym: bool = False) -> np.ndarray:
    A, Ac, Bc = CSC(H), CSC(Ja), CSC(Jb)
    pl, pr = _blk(psi_l), _blk(psi_r)
    NR = pl.shape[1]
    mu = np.zeros((NC, NC), dtype=np.complex128, order="F")
    _check(lib().kpm_oracle_2d(*A.args(), *Ac.args3(), *Bc.args3(), NC, NR, _dp(pl), _dp(pr),
                               K_left, K_right, parity, 1 if mn_sym else 0, _dp(mu)), "kpm_2d")
    return mu


def kpm_2d_tangent(H, Ja, Jb, dH, dJa, dJb, NC: int, psi_l, psi_r):
    """Forward-mode tangent of the kpm_2d table (numpy, small sizes):
    mu[n,m] = (1/NR) sum_r <l_m|Jb r_n> with the left chain l_m = T_m(H) psi_l
    and the right chain r_n = T_n(H) Ja psi_r of _kpm_2d_blocked!
    (src/moment.jl:1013-1061; conj placement of :1055-1070), differentiated
    along (dH, dJa, dJb) through the recurrences
    dl_1 = dH l_0, dl_k = 2 dH l_{k-1} + 2 H dl_{k-1} - dl_{k-2} (and the same
    for r with r_0 = Ja psi_r, dr_0 = dJa psi_r).  Returns (mu, dmu), NC x NC."""
    H, Ja, Jb = (sp.csr_matrix(X, dtype=np.complex128) for X in (H, Ja, Jb))
    dH, dJa, dJb = (sp.csr_matrix(X, dtype=np.complex128) for X in (dH, dJa, dJb))
    pl = np.asarray(psi_l, dtype=np.complex128).reshape(H.shape[0], -1)
    pr = np.asarray(psi_r, dtype=np.complex128).reshape(H.shape[0], -1)
    NR = pl.shape[1]

    def chain(v, dv):
        X, dX = [v], [dv]
        if NC > 1:
            X.append(H @ v)
            dX.append(dH @ v + H @ dv)
        for _ in range(2, NC):
            X.append(2 * (H @ X[-1]) - X[-2])
            dX.append(2 * (dH @ X[-2]) + 2 * (H @ dX[-1]) - dX[-2])
        return np.array(X), np.array(dX)

    L, dL = chain(pl, np.zeros_like(pl))
    R, dR = chain(Ja @ pr, dJa @ pr)
    JR = np.array([Jb @ r for r in R])
    dJR = np.array([dJb @ r + Jb @ d for r, d in zip(R, dR)])
    mu = np.einsum("mir,nir->nm", L.conj(), JR) / NR
    dmu = (np.einsum("mir,nir->nm", dL.conj(), JR) + np.einsum("mir,nir->nm", L.conj(), dJR)) / NR
    return mu, dmu


def kpm_3d(H, Ja, Jb, Jg, NC: int, psi_l, psi_r, K_left: int = 3, K_right: int | None = None):
    K_right = min(NC, 16) if K_right is None else K_right
    A, Ac, Bc, Gc = CSC(H), CSC(Ja), CSC(Jb), CSC(Jg)
    pl, pr = _blk(psi_l), _blk(psi_r)
    NR = pl.shape[1]
    mu = np.zeros((NC, NC, NC), dtype=np.complex128, order="F")
    _check(lib().kpm_oracle_3d(*A.args(), *Ac.args3(), *Bc.args3(), *Gc.args3(), NC, NR, _dp(pl),
                               _dp(pr), K_left, K_right, _dp(mu)), "kpm_3d")
    return mu


class OracleUnstable(RuntimeError):
    """The oracle's restatement of the recurrence stability guard tripped;
    .iteration and .growth are the reference's error() payload."""

    def __init__(self, iteration: int, growth: float):
        super().__init__(f"Chebyshev recurrence is unstable at iteration {iteration} "
                         f"(maximum column growth {growth} > 1.5)")
        self.iteration = iteration
        self.growth = growth


def cheb_action(H, V, C, check_every: int = 16) -> np.ndarray:
    """chebyshev_action! (src/utils/chebyshev_action.jl:47-125):
    out[:, :, k] = sum_n C[n, k] T_{n-1}(H) V, NH x NR x K complex."""
    A = CSC(H)
    v = _blk(V)
    c = np.asarray(C, dtype=np.complex128)
    if c.ndim == 1:
        c = c.reshape(-1, 1)
    c = np.asfortranarray(c)
    NC, K = c.shape
    NH, NR = v.shape
    out = np.zeros((NH, NR, K), dtype=np.complex128, order="F")
    g = np.zeros(2)
    rc = lib().kpm_oracle_cheb_action(*A.args(), NC, K, NR, _dp(v), _dp(c), _dp(out),
                                      int(check_every), _dp(g))
    if rc == -4:
        raise OracleUnstable(int(g[0]), float(g[1]))
    _check(rc, "cheb_action")
    return out


def cheb_iter(H, slots: np.ndarray) -> np.ndarray:
    """chebyshev_iter over an NH x NR x n workspace (slots 1, 2 given)."""
    A = CSC(H)
    s = np.asfortranarray(np.array(slots, dtype=np.complex128))
    NH, NR, n = s.shape
    _check(lib().kpm_oracle_cheb_iter(*A.args(), NR, n, _dp(s)), "cheb_iter")
    return s


class StepRunner:
    """Bounded run of the doubling loop for CPU-baseline timing."""

    def __init__(self, H, psi):
        self.A = CSC(H)
        self.v0 = _blk(psi).copy(order="F")
        self.v1 = self.v0.copy(order="F")
        self.NR = self.v0.shape[1]

    def run(self, steps: int) -> np.ndarray:
        mu = np.zeros((self.NR, 2 * steps), dtype=np.complex128, order="F")
        _check(lib().kpm_oracle_1d_steps(*self.A.args(), self.NR, steps, _dp(self.v0), _dp(self.v1),
                                         _dp(mu)), "1d_steps")
        return mu


# ---- independent dense references -----------------------------------------------


def chebyshev_matrices(H, NC: int):
    """T_0..T_{NC-1}(H) as dense matrices (test/kpm2d_blocking_test.jl:25-35)."""
    Hd = H.toarray() if sp.issparse(H) else np.asarray(H, dtype=np.complex128)
    N = Hd.shape[0]
    T = [np.eye(N, dtype=np.complex128)]
    if NC == 1:
        return T
    T.append(Hd.astype(np.complex128))
    for _ in range(2, NC):
        T.append(2 * Hd @ T[-1] - T[-2])
    return T


def action_dense(H, V, C):
    """Column-by-column three-term recurrence reference of the action
    (test/chebyshev_action_test.jl:20-41): out[:, j, k] = sum_n C[n,k] T_n V[:, j]."""
    Hd = H.toarray() if sp.issparse(H) else np.asarray(H, dtype=np.complex128)
    v = _blk(V)
    c = np.asarray(C, dtype=np.complex128).reshape(np.shape(C)[0], -1)
    NC, K = c.shape
    out = np.zeros(v.shape + (K,), dtype=np.complex128)
    for j in range(v.shape[1]):
        prev = v[:, j].copy()
        cur = Hd @ prev
        for k in range(K):
            out[:, j, k] += c[0, k] * prev
            if NC >= 2:
                out[:, j, k] += c[1, k] * cur
        for n in range(2, NC):
            prev, cur = cur, 2 * (Hd @ cur) - prev
            for k in range(K):
                out[:, j, k] += c[n, k] * cur
    return out


def action_ed(H, V, C):
    """ED form of the action: U diag(sum_n C[n,k] cos(n theta)) U^H V
    (test/chebyshev_action_test.jl:53-59)."""
    Hd = H.toarray() if sp.issparse(H) else np.asarray(H, dtype=np.complex128)
    lam, U = np.linalg.eigh(Hd)
    th = np.arccos(np.clip(lam, -1, 1))
    v = _blk(V)
    c = np.asarray(C, dtype=np.complex128).reshape(np.shape(C)[0], -1)
    f = np.cos(np.arange(c.shape[0])[:, None] * th[None, :]).T @ c      # (N, K)
    proj = U.conj().T @ v
    return np.stack([U @ (f[:, k:k + 1] * proj) for k in range(c.shape[1])], axis=2)


def kpm1d_dense(H, psi, NC):
    """[<psi_i|T_n psi_i>] NR x NC (test/kpm2d_blocking_test.jl:249)."""
    T = chebyshev_matrices(H, NC)
    p = _blk(psi)
    return np.array([[np.vdot(p[:, i], T[n] @ p[:, i]) for n in range(NC)]
                     for i in range(p.shape[1])])


def kpm1d_lr_dense(H, psi_l, psi_r, NC):
    T = chebyshev_matrices(H, NC)
    pl, pr = _blk(psi_l), _blk(psi_r)
    return np.array([[np.vdot(pl[:, i], T[n] @ pr[:, i]) for n in range(NC)]
                     for i in range(pl.shape[1])])


def current_dense(H, J, psi, NC):
    """[<psi_i|J T_n psi_i>] (test/kpm2d_blocking_test.jl:266)."""
    T = chebyshev_matrices(H, NC)
    Jd = J.toarray() if sp.issparse(J) else np.asarray(J)
    p = _blk(psi)
    return np.array([[np.vdot(p[:, i], Jd @ (T[n] @ p[:, i])) for n in range(NC)]
                     for i in range(p.shape[1])])


def kpm2d_dense(H, Ja, Jb, psi_l, psi_r, NC):
    """mu[n,m] = sum_i <T_m psi_l_i | Jb T_n Ja psi_r_i>/NR
    (test/kpm2d_blocking_test.jl:37-45)."""
    T = chebyshev_matrices(H, NC)
    Ja = Ja.toarray() if sp.issparse(Ja) else np.asarray(Ja)
    Jb = Jb.toarray() if sp.issparse(Jb) else np.asarray(Jb)
    pl, pr = _blk(psi_l), _blk(psi_r)
    NR = pl.shape[1]
    Lm = np.stack([T[m] @ pl for m in range(NC)])            # m, N, NR
    Rn = np.stack([Jb @ (T[n] @ (Ja @ pr)) for n in range(NC)])  # n, N, NR
    return np.einsum("mir,nir->nm", Lm.conj(), Rn) / NR


def kpm3d_dense(H, Ja, Jb, Jg, psi_l, psi_r, NC):
    """mu[n2,n3,n1] (test/kpm2d_blocking_test.jl:47-58)."""
    T = chebyshev_matrices(H, NC)
    dense = lambda X: X.toarray() if sp.issparse(X) else np.asarray(X)  # noqa: E731
    Ja, Jb, Jg = dense(Ja), dense(Jb), dense(Jg)
    pl, pr = _blk(psi_l), _blk(psi_r)
    NR = pl.shape[1]
    mu = np.zeros((NC, NC, NC), dtype=np.complex128)
    Lm = np.stack([T[m] @ pl for m in range(NC)])
    for n1 in range(NC):
        v = T[n1] @ (Ja @ pr)
        Rn = np.stack([Jb @ (T[n2] @ (Jg @ v)) for n2 in range(NC)])
        mu[:, :, n1] = np.einsum("mir,nir->nm", Lm.conj(), Rn) / NR
    return mu


def ed_moments(H, psi, NC):
    """mu_n = sum_i |<v_i|psi>|^2 T_n(lambda_i) (test/dos_test.jl:39-43), per column."""
    Hd = H.toarray() if sp.issparse(H) else np.asarray(H)
    lam, V = np.linalg.eigh(Hd)
    p = _blk(psi)
    w = np.abs(V.conj().T @ p) ** 2
    th = np.arccos(np.clip(lam, -1, 1))
    Tn = np.cos(np.arange(NC)[:, None] * th[None, :])
    return (Tn @ w).T


def ed_evolve(H, psi, t):
    """exp(-iHt) psi = U e^{-i lam t} U^H psi by dense eigendecomposition
    (test/ed_reference.jl:50-59), broadcast over the columns of psi."""
    Hd = H.toarray() if sp.issparse(H) else np.asarray(H)
    lam, U = np.linalg.eigh(Hd)
    p = np.asarray(psi, dtype=np.complex128)
    ph = np.exp(-1j * lam * t)
    proj = U.conj().T @ p
    return U @ (ph[:, None] * proj if p.ndim == 2 else ph * proj)


def _dense(A):
    return np.asarray(A.toarray() if sp.issparse(A) else A, dtype=np.complex128)


def ed_hall_conductivity_T0(H, Ja, Jb, area, *, Ef):
    """T = 0 Hall conductivity in e^2/h from the eigenbasis (TKNN form),
    2 pi (-2/A) sum_{occupied m, empty n} Im[va_mn vb_nm] / (E_m - E_n)^2 with
    v = J / i (test/ed_reference.jl:517-528)."""
    E, U = np.linalg.eigh(_dense(H))
    Va = U.conj().T @ (_dense(Ja) / 1j) @ U
    Vb = U.conj().T @ (_dense(Jb) / 1j) @ U
    occ, emp = E <= Ef, E > Ef
    num = np.imag(Va[np.ix_(occ, emp)] * Vb[np.ix_(emp, occ)].T)
    den = (E[occ][:, None] - E[emp][None, :]) ** 2
    return float(2 * np.pi * (-2.0 / area) * np.sum(num / den))


def ed_kubo_bastin_broadened(H, Ja, Jb, area, *, Ef, eta, beta=np.inf, grid_N=4000):
    """Kubo-Bastin conductivity in e^2/h with both delta(e - H) and G(e)
    broadened into width-eta Lorentzians and the energy integral done on a
    uniform grid over the occupied window padded by 20 eta
    (test/ed_reference.jl:349-385)."""
    E, U = np.linalg.eigh(_dense(H))
    Va = U.conj().T @ (_dense(Ja) / 1j) @ U
    Vb = U.conj().T @ (_dense(Jb) / 1j) @ U
    W = Va * Vb.T  # W_mn = va_mn vb_nm

    def f(e):
        if np.isinf(beta):
            return 1.0 if e <= Ef else 0.0
        return 1.0 / (np.exp(beta * (e - Ef)) + 1.0)

    lo = E.min() - 20 * eta
    hi = Ef if np.isinf(beta) else E.max() + 20 * eta
    grid_N = max(grid_N, int(math.ceil(20 * (hi - lo) / eta)))
    es = np.linspace(lo, hi, grid_N)
    de = es[1] - es[0]
    acc = 0j
    for e in es:
        Lv = (eta / np.pi) / ((e - E) ** 2 + eta ** 2)
        Gp = 1.0 / (e - E + 1j * eta) ** 2
        Gm = 1.0 / (e - E - 1j * eta) ** 2
        t1 = -(Gp @ (W @ Lv))
        t2 = -(Lv @ (W @ Gm))
        acc += f(e) * (t1 - t2) * de
    return float(2 * np.pi * np.real(1j * acc / area))


def ed_chern_marker(H, x, y, Ef, beta=np.inf):
    """m_i = -4 pi Im <i|P X Q Y P|i> with the exact projector P = U f(e - Ef) U^H,
    midpoint occupation at Ef for beta = inf (test/ed_reference.jl:222-247)."""
    Hd = H.toarray() if sp.issparse(H) else np.asarray(H)
    ev, U = np.linalg.eigh(Hd)
    if np.isinf(beta):
        occ = ((ev < Ef).astype(float) + (ev <= Ef).astype(float)) / 2
    else:
        occ = 1.0 / (np.exp(beta * (ev - Ef)) + 1.0)
    P = (U * occ[None, :]) @ U.conj().T
    X = np.diag(np.asarray(x, dtype=float))
    Y = np.diag(np.asarray(y, dtype=float))
    M = P @ X @ (np.eye(len(ev)) - P) @ Y @ P
    return -4 * np.pi * np.imag(np.diag(M))


# ---- host reconstruction heads (restated from the reference) -------------------------


def jackson(n, N):
    """Jackson kernel g_n (src/kernels/jackson_kernel.jl:7-8), zero for n >= N."""
    n = np.asarray(n, dtype=np.float64)
    q = np.pi / (N + 1.0)
    return ((N + 1.0 - n) * np.cos(q * n) + np.sin(q * n) / np.tan(q)) / (N + 1.0) * (n < N)


def hn(n):
    """h_0 = 1, h_n = 2 (src/kernels/kernels.jl:2)."""
    return np.where(np.asarray(n) == 0, 1.0, 2.0)


def lin_trans(x, n_grid, mu_tilde):
    """sum_j mu~_j cos(n_j acos x) (src/utils/chebyshev_lintrans.jl:12-27)."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    return np.cos(np.asarray(n_grid)[None, :] * np.arccos(x)[:, None]) @ np.asarray(mu_tilde)


def dos(mu, a, *, b=0.0, E_grid=None, N_tilde=0, E_range=None, NC=0, kernel=jackson):
    """(E_grid, rho), rho = sum_n h_n g_n mu_n T_n(x) / (a pi sqrt(1-x^2)) for
    |E - b| < a, zero elsewhere (src/applications/dos.jl:76-149, dE_order = 0)."""
    mu = np.asarray(mu)
    NC = len(mu) if NC == 0 else min(NC, len(mu))
    if E_grid is None:
        N_tilde = 2 * NC if N_tilde == 0 else N_tilde
        lo, hi = (b - 0.99 * a, b + 0.99 * a) if E_range is None else E_range
        E_grid = np.arange(N_tilde + 1) * (hi - lo) / N_tilde + lo
    E_grid = np.asarray(E_grid, dtype=np.float64)
    rho = np.zeros(len(E_grid))
    idx = np.abs(E_grid - b) < a
    x = (E_grid[idx] - b) / a
    ns = np.arange(NC)
    mt = mu[:NC] * np.asarray(kernel(ns, NC)) * hn(ns)
    rho[idx] = np.real(lin_trans(x, ns, mt) / (a * np.pi * np.sqrt(1 - x * x)))
    return E_grid, rho


def gauss_chebyshev(n):
    """First-kind Gauss-Chebyshev nodes cos((2k-1)pi/(2n)) (ascending), weights pi/n."""
    k = np.arange(1, n + 1)
    return np.sort(np.cos((2 * k - 1) * np.pi / (2 * n))), np.full(n, np.pi / n)


def kubo_bastin_cond(mu2D, a, Ef, *, b=0.0, NH, area, beta=np.inf, kernel=jackson, NC=None,
                     quad_N=None, edge_cutoff=1e-3):
    """sigma_ab(Ef) (src/applications/conductivity.jl:101-157): Gauss-Chebyshev
    nodes inside the edge cutoff with nonzero Fermi weight, and
    sum_k phi_k sum_nm mu~[n,m] Gamma_nm(x_k) through the factorised GEMMs."""
    mu2D = np.asarray(mu2D)
    NC = mu2D.shape[0] if NC is None else min(NC, *mu2D.shape)
    quad_N = 8 * NC if quad_N is None else quad_N
    w = np.asarray(kernel(np.arange(NC), NC)) * hn(np.arange(NC))
    mt = mu2D[:NC, :NC] * w[:, None] * w[None, :]
    nodes, weights = gauss_chebyshev(quad_N)
    E = a * nodes + b
    if np.isinf(beta):
        fv = ((E < Ef).astype(float) + (E <= Ef).astype(float)) / 2
    else:
        with np.errstate(over="ignore"):
            fv = 1.0 / (np.exp((E - Ef) * beta) + 1.0)
    keep = (np.abs(nodes) < 1 - edge_cutoff) & (fv != 0)
    if not np.any(keep):
        return 0.0
    x = nodes[keep]
    phi = weights[keep] * fv[keep] / (1 - x ** 2) ** 1.5
    ns = np.arange(NC)[:, None]
    th = np.arccos(x)[None, :]
    s = np.sqrt(1 - x ** 2)[None, :]
    C = np.cos(ns * th)
    Em = np.exp(1j * ns * th)
    WM = (x[None, :] - 1j * ns * s) * Em
    WN = (x[None, :] + 1j * ns * s) * np.conj(Em)
    acc = np.sum(phi[None, :] * ((mt.T @ C) * WM)) + np.sum(phi[None, :] * ((mt @ C) * WN))
    return float(-2 * NH / (area * a ** 2) * acc.real)


def kubo_gamma_direct(mt, x):
    """sum_nm mu~[n,m] Gamma_nm(x) with the unfactorised kernel of Garcia et al.
    (PRL 114, 116602): Gamma_nm = (x - i m s) e^{i m th} T_n + (x + i n s) e^{-i n th} T_m."""
    NC = mt.shape[0]
    th = np.arccos(x)
    s = np.sqrt(1 - x * x)
    tot = 0j
    for n in range(NC):
        for m in range(NC):
            g = ((x - 1j * m * s) * np.exp(1j * m * th) * np.cos(n * th) +
                 (x + 1j * n * s) * np.exp(-1j * n * th) * np.cos(m * th))
            tot += mt[n, m] * g
    return tot


# ---- three-current (CPGE) response (src/applications/cpge.jl) ---------------
# cpge_forms restates the node contraction _cpge_contract!
# (src/applications/cpge.jl:59-64).  The remaining helpers restate the
# reference's independent test references (test/cpge_test.jl:37-316): the
# closed-form resolvent coefficients, the eigenbasis integrand and an
# adaptive integral over theta that shares no code with the product's
# quadrature (scipy's QUADPACK instead of the product's Gauss-Kronrod
# restatement).

def cpge_forms(mu_tilde, a, b, c):
    """For each row q: collapsed = mu~^T a_q over the NC x NC^2 view,
    right = reshape(collapsed, NC, NC) c_q, F_q = sum(b_q .* right)."""
    mu_tilde = np.asarray(mu_tilde)
    NC = mu_tilde.shape[0]
    M = np.asfortranarray(mu_tilde).reshape(NC, NC * NC, order="F")
    out = np.empty(a.shape[0], dtype=np.complex128)
    for q in range(a.shape[0]):
        collapsed = M.T @ a[q]
        right = collapsed.reshape(NC, NC, order="F") @ c[q]
        out[q] = np.sum(b[q] * right)
    return out


def _cheb_vec(x, NC):
    T = np.zeros(NC, dtype=np.complex128 if np.iscomplexobj(x) else np.float64)
    if NC == 0:
        return T
    T[0] = 1
    if NC > 1:
        T[1] = x
    for n in range(2, NC):
        T[n] = 2 * x * T[n - 1] - T[n - 2]
    return T


def cpge_ref_delta(x, NC, times_sin=False):
    """T_n(x) / (pi sqrt(1 - x^2)) (test/cpge_test.jl:67); with times_sin,
    that value times sin(theta) = sqrt(1 - x^2), i.e. T_n(x) / pi."""
    if times_sin:
        return _cheb_vec(x, NC) / np.pi
    return _cheb_vec(x, NC) / (np.pi * np.sqrt(1 - x * x))


def cpge_ref_green(z, NC, lam, branch, delta=0.0):
    """Closed-form resolvent coefficients (test/cpge_test.jl:69-96)."""
    n = np.arange(NC)
    if delta > 0 and 1 - delta < abs(z) < 1 + delta:
        return np.zeros(NC, dtype=np.complex128)
    if lam == 0:
        if abs(z) < 1:
            phi = np.arccos(z)
            pre = -1j / np.sin(phi) if branch == "R" else 1j / np.sin(phi)
            sign = -1 if branch == "R" else 1
            return pre * np.exp(sign * 1j * n * phi)
        if z > 1:
            eta = np.arccosh(z)
            return (np.exp(-n * eta) / np.sqrt(z * z - 1)).astype(np.complex128)
        if z < -1:
            eta = np.arccosh(-z)
            return (-((-1.0) ** n) * np.exp(-n * eta) / np.sqrt(z * z - 1)).astype(np.complex128)
        raise ValueError("z = +/-1 is singular")
    w = complex(z, lam) if branch == "R" else complex(z, -lam)
    if branch == "R":
        return -1j * np.exp(-1j * n * np.arccos(w)) / np.sqrt(1 - w * w)
    return 1j * np.exp(1j * n * np.arccos(w)) / np.sqrt(1 - w * w)


def cpge_ref_fermi(x, xF, beta):
    if np.isinf(beta):
        return (float(x < xF) + float(x <= xF)) / 2
    y = beta * (x - xF)
    return np.exp(-y) / (1 + np.exp(-y)) if y > 0 else 1 / (1 + np.exp(y))


def cpge_ref_integral(fun, shifts, xF, beta, lam, *, delta=0.0, rtol=2e-10, atol=2e-12):
    """Adaptive integral of sin(theta) f(cos theta) fun(cos theta) over the
    occupied part of (0, pi), split at the shifted edges (and the delta
    annulus), with the u^2 endpoint map at singular edges when lambda =
    delta = 0 (test/cpge_test.jl:112-198); fun may be array-valued."""
    from scipy.integrate import quad_vec

    lo = np.arccos(np.clip(xF, -1, 1)) if np.isinf(beta) else 0.0
    points, singular = [lo, np.pi], []
    for shift in shifts:
        for edge in (-1.0, 1.0):
            xs = edge - shift
            if -1 < xs < 1:
                th = np.arccos(xs)
                if lo < th < np.pi:
                    points.append(th)
                    if lam == 0 and delta == 0:
                        singular.append(th)
    if delta > 0:
        for shift in shifts:
            for radius in (1 - delta, 1 + delta):
                if radius < 0:
                    continue
                for edge in (-radius, radius):
                    xs = edge - shift
                    if -1 < xs < 1:
                        th = np.arccos(xs)
                        if lo < th < np.pi:
                            points.append(th)
    if not np.isinf(beta) and beta > 50 and -1 < xF < 1:
        points.append(np.arccos(xF))

    def uniq(p, tol=1e-13):
        out = []
        for v in sorted(p):
            if not out or abs(v - out[-1]) > tol:
                out.append(v)
        return out

    points, singular = uniq(points), uniq(singular)

    def isedge(th):
        return any(abs(th - s) < 1e-12 for s in singular)

    # fun(x, s) returns sin(theta) * bracket(x) with s = sin(theta): every
    # bracket term has exactly one delta factor, whose 1/sqrt(1 - x^2) cancels
    # sin(theta) analytically (no 0 * inf where cos(theta) rounds to +/-1)
    def node(th):
        return cpge_ref_fermi(np.cos(th), xF, beta) * np.asarray(fun(np.cos(th), np.sin(th)))

    def quad(f, a, b):
        return quad_vec(f, a, b, epsabs=atol, epsrel=rtol, limit=2000)[0]

    def piece(a, b, le, re):
        if le and re:
            mid = (a + b) / 2
            return piece(a, mid, True, False) + piece(mid, b, False, True)
        w = b - a
        if le:
            return quad(lambda u: node(a + w * u * u) * (2 * w * u), 0.0, 1.0)
        if re:
            return quad(lambda u: node(b - w * u * u) * (2 * w * u), 0.0, 1.0)
        return quad(node, a, b)

    total = 0
    for k in range(len(points) - 1):
        total = total + piece(points[k], points[k + 1], isedge(points[k]), isedge(points[k + 1]))
    return total


def cpge_eigen_setup(Hn, Ja, Jb, Jg, NC):
    E, V = np.linalg.eigh(np.asarray(Hn.toarray() if sp.issparse(Hn) else Hn))
    dense = [np.asarray(J.toarray() if sp.issparse(J) else J) for J in (Ja, Jb, Jg)]
    A, B, C = (V.conj().T @ J @ V for J in dense)
    T = np.stack([_cheb_vec(e, NC) for e in E], axis=1)  # NC x D
    kh = jackson(np.arange(NC), NC) * hn(np.arange(NC))
    return E, A, B, C, T, kh


def cpge_eigen_integrand(setup, NC, omega1, omega2, x, lam, delta=0.0, times_sin=False):
    """The eigenbasis bracket RR + RA + AA at energy x
    (test/cpge_test.jl:200-227): Tr[Ja T_n Jb T_m Jg T_p] puts the (n, m, p)
    coefficient sums at eigenvalues (E_b, E_c, E_a).  times_sin: the bracket
    times sin(theta) (x = cos(theta)), exact at the interval ends."""
    E, A, B, C, T, kh = setup
    Om = omega1 + omega2
    dK = (kh * cpge_ref_delta(x, NC, times_sin)) @ T
    g = lambda z, br: (kh * cpge_ref_green(z, NC, lam, br, delta)) @ T
    rO, r2, r1 = g(x + Om, "R"), g(x + omega2, "R"), g(x + omega1, "R")
    a2, a1, aO = g(x - omega2, "A"), g(x - omega1, "A"), g(x - Om, "A")
    ABC = np.einsum("ab,bc,ca->abc", A, B, C)
    acc = (np.einsum("abc,b,c,a->", ABC, rO, r2, dK) + np.einsum("abc,b,c,a->", ABC, r1, dK, a2) +
           np.einsum("abc,b,c,a->", ABC, dK, a1, aO))
    return acc / len(E)


def cpge_eigen_reference(Hn, Ja, Jb, Jg, NC, omega1, omega2, *, xF, beta, lam, delta=0.0,
                         rtol=2e-10, atol=2e-12):
    """i * the integral of the eigenbasis bracket (test/cpge_test.jl:229-285)."""
    setup = cpge_eigen_setup(Hn, Ja, Jb, Jg, NC)
    shifts = (omega1 + omega2, omega2, omega1, -omega2, -omega1, -omega1 - omega2)
    f = lambda x, s: cpge_eigen_integrand(setup, NC, omega1, omega2, x, lam, delta, times_sin=True)
    return 1j * cpge_ref_integral(f, shifts, xF, beta, lam, delta=delta, rtol=rtol, atol=atol)


def cpge_lambda_table(NC, omega1, omega2, *, xF, beta, lam):
    """Integral of the coefficient tensor rO[n] r2[m] D[p] + r1[n] D[m] a2[p]
    + D[n] a1[m] aO[p] (test/cpge_test.jl:287-309)."""
    Om = omega1 + omega2
    shifts = (Om, omega2, omega1, -omega2, -omega1, -Om)

    def f(x, s):
        D = cpge_ref_delta(x, NC, times_sin=True)
        rO, r2 = cpge_ref_green(x + Om, NC, lam, "R"), cpge_ref_green(x + omega2, NC, lam, "R")
        r1 = cpge_ref_green(x + omega1, NC, lam, "R")
        a2, a1 = cpge_ref_green(x - omega2, NC, lam, "A"), cpge_ref_green(x - omega1, NC, lam, "A")
        aO = cpge_ref_green(x - Om, NC, lam, "A")
        return (np.einsum("n,m,p->nmp", rO, r2, D) + np.einsum("n,m,p->nmp", r1, D, a2) +
                np.einsum("n,m,p->nmp", D, a1, aO))

    return cpge_ref_integral(f, shifts, xF, beta, lam)


# ---- BdG and stiffness ED references (test/ed_reference.jl:576-735) ----------

def bdg_matrix(h, mu, U, n, Delta, convention="intervalley"):
    """Dense reduced Nambu matrix [xi D; D^* hole] with xi = h - mu - U n/2
    (bdg_matrix / bdg_matrix_singlet, test/ed_reference.jl:576-598): hole =
    -xi (intervalley) or -conj(xi) (singlet)."""
    hd = h.toarray() if sp.issparse(h) else np.asarray(h)
    N = hd.shape[0]
    xi = hd - mu * np.eye(N) - np.diag(np.asarray(U) * np.asarray(n) / 2)
    D = np.diag(np.asarray(Delta, dtype=np.complex128))
    hole = -xi if convention == "intervalley" else -np.conj(xi)
    return np.block([[xi, D], [np.conj(D), hole]]).astype(np.complex128)


def ed_two_energy_response(H, Jl, Jr, *, beta, eta, omega=0.0, Ef=0.0):
    """Dense Lehmann sum of the two-energy response
    (test/ed_reference.jl:600-632)."""
    E, V = np.linalg.eigh(np.asarray(H.toarray() if sp.issparse(H) else H))
    Jl = V.conj().T @ np.asarray(Jl.toarray() if sp.issparse(Jl) else Jl) @ V
    Jr = V.conj().T @ np.asarray(Jr.toarray() if sp.issparse(Jr) else Jr) @ V

    def fermi(e):
        if np.isinf(beta):
            return (float(e < Ef) + float(e <= Ef)) / 2
        return 1 / (np.exp(beta * (e - Ef)) + 1)

    occ = np.array([fermi(e) for e in E])
    acc = 0j
    for p in range(len(E)):
        for q in range(len(E)):
            dE = E[p] - E[q]
            if eta == 0:
                if abs(dE) < 1e-8:
                    fm = fermi((E[p] + E[q]) / 2)
                    dd = -beta * fm * (1 - fm)
                else:
                    dd = (occ[p] - occ[q]) / dE
            else:
                dd = (occ[p] - occ[q]) / (omega + dE + 1j * eta)
            acc += Jl[p, q] * Jr[q, p] * dd
    return complex(acc)


def ed_diamagnetic(H, Dhat, *, beta):
    """sum_n f(E_n) <n|Dhat|n> at Fermi level zero
    (test/ed_reference.jl:634-651)."""
    E, V = np.linalg.eigh(np.asarray(H.toarray() if sp.issparse(H) else H))
    Dd = V.conj().T @ np.asarray(Dhat.toarray() if sp.issparse(Dhat) else Dhat) @ V
    f = 1 / (np.exp(beta * E) + 1)
    return float(np.real(np.sum(f * np.diag(Dd))))


def bdg_peierls_matrix(h, pos, disp, A, *, q=None, dir=1, convention="singlet"):
    """Kinetic Nambu part of the Peierls-coupled BdG matrix with a real
    modulation cos(q.m) (test/ed_reference.jl:692-733)."""
    hc = sp.coo_matrix(h)
    N = h.shape[0]
    hp = np.zeros((N, N), dtype=np.complex128)
    hh = np.zeros((N, N), dtype=np.complex128)
    for i, j, v in zip(hc.row, hc.col, hc.data):
        d = disp(i, j)
        m = pos[j, :] + d / 2
        u = 1.0 if q is None else np.cos(np.dot(q, m))
        hp[i, j] = v * np.exp(1j * A * d[dir - 1] * u)
        hole = np.conj(v) if convention == "singlet" else v
        hh[i, j] = -hole * np.exp(-1j * A * d[dir - 1] * u)
    Z = np.zeros((N, N), dtype=np.complex128)
    return np.block([[hp, Z], [Z, hh]])


def ed_bdg_free_energy(h, pos, disp, mu, U, n, Delta, A, *, q=None, dir=1, beta, convention="singlet"):
    """F(A) = -(1/beta) sum_n log(1 + e^{-beta E_n(A)}) of the Peierls-coupled
    kinetic matrix plus the local mu, Hartree and pairing terms (rigid Delta)
    (ed_bdg_free_energy, test/ed_reference.jl:735-810)."""
    N = h.shape[0]
    Dm = np.diag(np.asarray(Delta, dtype=np.complex128)) if np.ndim(Delta) == 1 else \
        np.asarray(Delta, dtype=np.complex128)
    xi0 = -mu * np.eye(N) - np.diag(np.asarray(U) * np.asarray(n) / 2)
    hole0 = -np.conj(xi0) if convention == "singlet" else -xi0
    local = np.block([[xi0, Dm], [Dm.conj().T, hole0]])
    H = bdg_peierls_matrix(h, pos, disp, A, q=q, dir=dir, convention=convention) + local
    E = np.linalg.eigvalsh(H)
    x = -beta * E
    sp_ = np.where(x > 0, x + np.log1p(np.exp(-np.abs(x))), np.log1p(np.exp(np.minimum(x, 0))))
    return float(-np.sum(sp_) / beta)
