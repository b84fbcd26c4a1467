"""The moment engine -- host mirror of src/moment.jl whose compute runs in
libkpm_hip.so on the GPU.

Names, argument meaning and error behaviour follow the reference:
``kpm_1d`` (src/moment.jl:339-418), ``kpm_1d_`` = ``kpm_1d!`` (:420-556),
``kpm_1d_current`` / ``kpm_1d_current_`` (:558-735), ``kpm_2d_blocking``
(:759-810), ``kpm_2d`` / ``kpm_2d_`` (:812-1103), ``kpm_3d`` / ``kpm_3d_``
(:1106-1280).  Julia's ``!`` suffix is spelled with a trailing underscore.

Additions of this engine (keyword-only, defaults reproduce the reference):
``precision`` ('f64' ComplexF64 vectors, 'f32' ComplexF32 vectors with FP64
arithmetic and FP64 moment accumulation) and, for kpm_2d/kpm_3d, ``gemm``
('f64' or 'f32' MFMA contraction).
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib
from ._lib import ArgumentError, check, dptr
from .device import (DeviceBlock, get_context, prec_code, to_device, to_device_adjoint)
from .primitives import normalize_by_col
from .rng import default_rng

dt_cplx = np.complex128
dt_real = np.float64


def _nh(H) -> int:
    return int(H.shape[0])


def _probe(psi, NH: int, NR: int, name: str) -> np.ndarray:
    a = np.asarray(psi)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    if a.shape != (NH, NR):
        raise AssertionError(f"Invalid `{name}` with size {a.shape}. Expecting {NH}, {NR}")
    return np.asfortranarray(a, dtype=np.complex128)


def random_probes(NH: int, NR: int, rng=None) -> np.ndarray:
    """exp(2πi U) with unit-norm columns (src/moment.jl:494-495); U from the
    build's seeded xoshiro256++ generator."""
    rng = rng if rng is not None else default_rng(np.random.SeedSequence().entropy & 0xFFFFFFFF)
    psi = np.exp(rng.rand(NH, NR) * (2j * np.pi))
    normalize_by_col(psi, NR)
    return np.asfortranarray(psi)


# --------------------------------------------------------------------------
# kpm_1d


def kpm_1d_(H, NC: int, NR: int, NH: int, mu_all: np.ndarray, psi_in=None, psi_in_r=None, *,
            verbose: int = 0, precision="f64"):
    """kpm_1d!(H, NC, NR, NH, mu_all[, psi_in[, psi_in_r]]).

    - ``kpm_1d_(H, NC, NR, NH, mu)``: random probes (src/moment.jl:493-497).
    - ``kpm_1d_(H, NC, NR, NH, mu, psi)``: equal-vector doubling (:420-492).
    - ``kpm_1d_(H, NC, NR, NH, mu, psi_l, psi_r)``: bra/ket (:498-556).
    """
    if mu_all.shape != (NR, NC):
        raise AssertionError("size(mu_all) == (NR, NC)")
    ctx = get_context()
    Hd = to_device(H, ctx)
    prec = prec_code(precision)
    out = np.zeros((NR, NC), dtype=np.complex128, order="F")
    if psi_in is None:
        # random probes drawn on the device (src/moment.jl:493-497)
        if NC % 2 != 0:
            raise AssertionError("Invalid NC: NC should be even.")
        if Hd.shape[0] != NH:
            raise AssertionError("H and psi_in disagree on NH")
        blk = DeviceBlock(NH, NR, precision, ctx)
        blk.random_phase(int(np.random.SeedSequence().entropy) & 0xFFFFFFFFFFFFFFFF)
        check(ctx.lib.kpm_1d_block(ctx.handle, Hd.handle, NC, blk.handle, dptr(out)), "kpm_1d!")
        mu_all[...] = out
        return None
    if psi_in_r is None:
        if NC % 2 != 0:
            raise AssertionError("Invalid NC: NC should be even.")
        psi = _probe(psi_in, NH, NR, "psi_in")
        if Hd.shape[0] != NH:
            raise AssertionError("H and psi_in disagree on NH")
        check(ctx.lib.kpm_1d(ctx.handle, Hd.handle, NC, NR, NH, dptr(psi),
                             dptr(out), prec), "kpm_1d!")
    else:
        if NC < 2:
            raise AssertionError("Invalid NC: NC should be at least 2.")
        psl = _probe(psi_in, NH, NR, "psi_in_l")
        psr = _probe(psi_in_r, NH, NR, "psi_in_r")
        check(ctx.lib.kpm_1d_lr(ctx.handle, Hd.handle, NC, NR, NH, dptr(psl),
                                dptr(psr), dptr(out), prec),
              "kpm_1d!")
    mu_all[...] = out
    return None


def _kpm_1d_avg(Hd, NC: int, NR: int, NH: int, psi_in, precision) -> np.ndarray:
    """Probe-averaged equal-vector moments, sum(mu_all, dims=1) ./ NR
    (src/moment.jl:410-416), from the engine's probe-summed doubling pass:
    the NR x NC per-probe table is never formed."""
    if NC % 2 != 0:
        raise AssertionError("Invalid NC: NC should be even.")
    if Hd.shape[0] != NH:
        raise AssertionError("H and psi_in disagree on NH")
    ctx = get_context()
    out = np.zeros(NC, dtype=np.complex128)
    if psi_in is None:
        # random probes drawn on the device (src/moment.jl:493-497)
        blk = DeviceBlock(NH, NR, precision, ctx)
        blk.random_phase(int(np.random.SeedSequence().entropy) & 0xFFFFFFFFFFFFFFFF)
        check(ctx.lib.kpm_1d_block_avg(ctx.handle, Hd.handle, NC, blk.handle, dptr(out)), "kpm_1d!")
    else:
        psi = _probe(psi_in, NH, NR, "psi_in")
        check(ctx.lib.kpm_1d_avg(ctx.handle, Hd.handle, NC, NR, NH, dptr(psi), dptr(out),
                                 prec_code(precision)), "kpm_1d!")
    return out


def kpm_1d(H, NC: int, NR: int, NH: int | None = None, *, psi_in=None, psi_in_l=None,
           psi_in_r=None, force_norm: bool = False, verbose: int = 0, avg_output: bool = True,
           NR_parallel: bool = True, precision="f64"):
    """kpm_1d (src/moment.jl:339-418): NC moments averaged over NR probes
    (real for equal vectors, complex for bra/ket), or the NR x NC matrix with
    ``avg_output=False``.  Averaged equal-vector calls run the probe-summed
    doubling pass (``_kpm_1d_avg``)."""
    NC = int(NC)
    NR = int(NR)
    NH = _nh(H) if NH is None else int(NH)
    mu_all = np.zeros((NR, NC), dtype=np.complex128, order="F")
    lr_mode = False
    Hd = to_device(H)
    if psi_in is None:
        if psi_in_l is not None or psi_in_r is not None:
            if psi_in_l is None or psi_in_r is None:
                raise AssertionError("must set both `psi_in_l` and `psi_in_r` or neither.")
            lr_mode = True
            if force_norm:
                normalize_by_col(psi_in_l, NR)
                normalize_by_col(psi_in_r, NR)
            if NR_parallel:
                kpm_1d_(Hd, NC, NR, NH, mu_all, psi_in_l, psi_in_r, verbose=verbose,
                        precision=precision)
            else:
                for i in range(NR):
                    row = np.zeros((1, NC), dtype=np.complex128, order="F")
                    kpm_1d_(Hd, NC, 1, NH, row, np.asarray(psi_in_l)[:, i:i + 1],
                            np.asarray(psi_in_r)[:, i:i + 1], verbose=verbose, precision=precision)
                    mu_all[i, :] = row[0]
        else:
            if NR_parallel and avg_output:
                return _kpm_1d_avg(Hd, NC, NR, NH, None, precision).real.copy()
            if NR_parallel:
                kpm_1d_(Hd, NC, NR, NH, mu_all, verbose=verbose, precision=precision)
            else:
                for i in range(NR):
                    row = np.zeros((1, NC), dtype=np.complex128, order="F")
                    kpm_1d_(Hd, NC, 1, NH, row, verbose=verbose, precision=precision)
                    mu_all[i, :] = row[0]
    else:
        if psi_in_l is not None or psi_in_r is not None:
            raise AssertionError(
                "must either set `psi_in` or set `psi_in_l` and `psi_in_r`, but not both.")
        if force_norm:
            normalize_by_col(psi_in, NR)
        if NR_parallel and avg_output:
            return _kpm_1d_avg(Hd, NC, NR, NH, psi_in, precision).real.copy()
        if NR_parallel:
            kpm_1d_(Hd, NC, NR, NH, mu_all, psi_in, verbose=verbose, precision=precision)
        else:
            for i in range(NR):
                row = np.zeros((1, NC), dtype=np.complex128, order="F")
                kpm_1d_(Hd, NC, 1, NH, row, np.asarray(psi_in)[:, i:i + 1], verbose=verbose,
                        precision=precision)
                mu_all[i, :] = row[0]
    if avg_output:
        avg = mu_all.sum(axis=0) / NR
        return avg if lr_mode else avg.real.copy()
    return mu_all


# --------------------------------------------------------------------------
# kpm_1d_current


def kpm_1d_current_(H, Ja, NC: int, NR: int, NH: int, mu_all: np.ndarray, psi_in=None,
                    psi_in_r=None, *, verbose: int = 0, precision="f64"):
    """kpm_1d_current!(H, Jα, NC, NR, NH, mu[, psi_in]) (src/moment.jl:656-735).
    The bra/ket form is unimplemented in the reference (:722-735)."""
    if psi_in_r is not None:
        raise NotImplementedError("unimplemented.")
    if mu_all.shape != (NR, NC):
        raise AssertionError("size(mu_all) == (NR, NC)")
    ctx = get_context()
    Hd = to_device(H, ctx)
    Jadj = to_device_adjoint(Ja, ctx)
    if psi_in is None:
        psi_in = random_probes(NH, NR)
    psi = _probe(psi_in, NH, NR, "psi_in")
    out = np.zeros((NR, NC), dtype=np.complex128, order="F")
    check(ctx.lib.kpm_1d_current(ctx.handle, Hd.handle, Jadj.handle, NC, NR, NH,
                                 dptr(psi), dptr(out),
                                 prec_code(precision)), "kpm_1d_current!")
    mu_all[...] = out


def kpm_1d_current(H, Ja, NC: int, NR: int, NH: int | None = None, *, psi_in=None, psi_in_l=None,
                   psi_in_r=None, force_norm: bool = False, verbose: int = 0,
                   avg_output: bool = True, NR_parallel: bool = True, precision="f64"):
    """kpm_1d_current (src/moment.jl:558-654): Γ_n = <ψ|Jα T_n|ψ>, complex."""
    NC = int(NC)
    NR = int(NR)
    NH = _nh(H) if NH is None else int(NH)
    mu_all = np.zeros((NR, NC), dtype=np.complex128, order="F")
    Hd = to_device(H)
    if psi_in is None:
        if psi_in_l is not None or psi_in_r is not None:
            if psi_in_l is None or psi_in_r is None:
                raise AssertionError("must set both `psi_in_l` and `psi_in_r` or neither.")
            if force_norm:
                normalize_by_col(psi_in_l, NR)
                normalize_by_col(psi_in_r, NR)
            kpm_1d_current_(Hd, Ja, NC, NR, NH, mu_all, psi_in_l, psi_in_r)
        else:
            if NR_parallel:
                kpm_1d_current_(Hd, Ja, NC, NR, NH, mu_all, precision=precision)
            else:
                for i in range(NR):
                    row = np.zeros((1, NC), dtype=np.complex128, order="F")
                    kpm_1d_current_(Hd, Ja, NC, 1, NH, row, precision=precision)
                    mu_all[i, :] = row[0]
    else:
        if psi_in_l is not None or psi_in_r is not None:
            raise AssertionError(
                "must either set `psi_in` or set `psi_in_l` and `psi_in_r`, but not both.")
        if force_norm:
            normalize_by_col(psi_in, NR)
        if NR_parallel:
            kpm_1d_current_(Hd, Ja, NC, NR, NH, mu_all, psi_in, precision=precision)
        else:
            for i in range(NR):
                row = np.zeros((1, NC), dtype=np.complex128, order="F")
                kpm_1d_current_(Hd, Ja, NC, 1, NH, row, np.asarray(psi_in)[:, i:i + 1],
                                precision=precision)
                mu_all[i, :] = row[0]
    if avg_output:
        return mu_all.sum(axis=0) / NR
    return mu_all


# --------------------------------------------------------------------------
# kpm_2d


def _auto_budget() -> int:
    """_kpm2d_auto_budget on the active device: 50% of free device memory
    (src/moment.jl:741)."""
    return get_context().free_memory() // 2


def kpm_2d_blocking(NH: int, NR: int, NC: int, *, workspace_bytes="auto", arr_size="auto",
                    right_block="auto"):
    """kpm_2d_blocking (src/moment.jl:759-810): returns a dict with keys
    ``arr_size`` (K_left), ``right_block`` (K_right) and ``bytes``."""
    NH, NR, NC = int(NH), int(NR), int(NC)
    if NH <= 0:
        raise ArgumentError(f"NH must be positive (got {NH})")
    if NR <= 0:
        raise ArgumentError(f"NR must be positive (got {NR})")
    if NC < 2:
        raise ArgumentError(f"NC must be at least 2 (got {NC})")
    block_bytes = NH * NR * 16
    if isinstance(workspace_bytes, str) and workspace_bytes == "auto":
        env = os.environ.get("KPM_WORKSPACE_BYTES", "")
        budget = _auto_budget() if env == "" else int(env)
    elif isinstance(workspace_bytes, (int, np.integer)) and not isinstance(workspace_bytes, bool):
        budget = int(workspace_bytes)
    else:
        raise ArgumentError("workspace_bytes must be :auto or an integer")
    if budget < 0:
        raise ArgumentError("workspace_bytes must be nonnegative")
    if isinstance(right_block, str) and right_block == "auto":
        K_right = min(NC, 16)
    elif isinstance(right_block, (int, np.integer)) and right_block >= 1:
        K_right = int(right_block)
    else:
        raise ArgumentError("right_block must be :auto or a positive integer")
    base = (K_right + 7) * block_bytes
    per_left = block_bytes + K_right * 16
    if isinstance(arr_size, str) and arr_size == "auto":
        kmax = (budget - base) // per_left
        if kmax < 2:
            raise ArgumentError(
                f"workspace_bytes={budget} cannot fit the minimum kpm_2d workspace "
                f"({base + 2 * per_left} bytes at right_block={K_right})")
        K_left = min(kmax, NC)
    elif isinstance(arr_size, (int, np.integer)) and arr_size >= 2:
        K_left = int(arr_size)
    else:
        raise ArgumentError("arr_size must be :auto or an integer at least 2")
    return {"arr_size": K_left, "right_block": K_right, "bytes": base + K_left * per_left}


def engine_blocking(NH: int, NR: int, NC: int, *, precision="f64", workspace_bytes="auto",
                    arr_size="auto", right_block="auto"):
    """The device engine's plan for kpm_2d.  Explicit sizes are kept as given
    (any blocking gives the same moments to rounding,
    test/kpm2d_blocking_test.jl:97-117).  For ':auto' sizes the engine keeps
    every left AND right slot resident when the budget allows (K_left =
    K_right = NC), so the contraction is one MFMA GEMM; otherwise it splits
    the budget between the two slot arrays."""
    prec = prec_code(precision)
    elem = 8 if prec == _lib.VEC_F32 else 16
    nrp = _padded_nr(NR, prec)
    slot = ((NH * nrp + 15) // 16 * 16) * elem
    if isinstance(workspace_bytes, str) and workspace_bytes == "auto":
        env = os.environ.get("KPM_WORKSPACE_BYTES", "")
        budget = _auto_budget() if env == "" else int(env)
    else:
        budget = int(workspace_bytes)
    fixed = 7 * NH * nrp * elem  # probe, ring and starter blocks
    avail = max(budget - fixed, 0) // slot
    KL = None if arr_size == "auto" else int(arr_size)
    KR = None if right_block == "auto" else int(right_block)
    if KL is None and KR is None:
        if avail >= 2 * NC:
            KL, KR = NC, NC
        else:
            KL = max(2, min(NC, avail // 2))
            KR = max(1, min(NC, avail - KL))
    elif KL is None:
        KL = max(2, min(NC, avail - KR))
    elif KR is None:
        KR = max(1, min(NC, avail - KL))
    if KL < 2:
        raise ArgumentError("ψall_l must have at least two slots")
    if KR < 1:
        raise ArgumentError("JTnHJψr must have at least one slot")
    return {"arr_size": int(KL), "right_block": int(KR), "bytes": int((KL + KR) * slot + fixed)}


def _padded_nr(NR: int, prec: int) -> int:
    cpw = 2 if prec == _lib.VEC_F32 else 1
    w = -(-NR // cpw)
    if w <= 64:
        p = 1
        while p < w:
            p <<= 1
        return p * cpw
    return -(-w // 64) * 64 * cpw


_PARITY = {"NONE": 0, "ODD": 1, "EVEN": 2}
# kpm_2d contraction arithmetic: FP64 MFMA, FP32 MFMA, or 3-pass split BF16
_GEMM_MODES = {"f64": _lib.GEMM_F64, "f32": _lib.GEMM_F32, "bf16x3": _lib.GEMM_BF16X3}


def _parity_code(moment_parity) -> int:
    key = str(moment_parity).lstrip(":").upper()
    if key not in _PARITY:
        raise ArgumentError(f"moment_parity={moment_parity} not understood.")
    return _PARITY[key]


def kpm_2d_(H, Ja, Jb, NC: int, NR: int, NH: int, mu: np.ndarray, psi_in_l=None, psi_in_r=None,
            *, arr_size="auto", right_block="auto", workspace_bytes="auto", verbose: int = 0,
            mn_sym: bool = False, moment_parity="NONE", precision="f64", gemm="f64"):
    """kpm_2d!(H, Jα, Jβ, NC, NR, NH, μ[, psi_in[, psi_in_r]]) (src/moment.jl:887-1103):
    μ[n,m] = (1/NR) Σ_r <T_{m-1} ψl_r | Jβ T_{n-1} Jα ψr_r>."""
    parity = _parity_code(moment_parity)
    mu[...] = 0
    if psi_in_l is None:
        psi_in_l = random_probes(NH, NR)
    if psi_in_r is None:
        psi_in_r = psi_in_l
    plan = engine_blocking(NH, NR, NC, precision=precision, workspace_bytes=workspace_bytes,
                           arr_size=arr_size, right_block=right_block)
    K_left, K_right = plan["arr_size"], plan["right_block"]
    if verbose >= 1:
        print(f"kpm_2d blocking: arr_size={K_left}, right_block={K_right}, "
              f"workspace={plan['bytes']} bytes")
    psl = _probe(psi_in_l, NH, NR, "psi_in_l")
    psr = _probe(psi_in_r, NH, NR, "psi_in_r")
    ctx = get_context()
    Hd, Jad, Jbd = to_device(H, ctx), to_device(Ja, ctx), to_device(Jb, ctx)
    out = np.zeros((NC, NC), dtype=np.complex128, order="F")
    gm = _GEMM_MODES.get(str(gemm).lower())
    if gm is None:
        raise ArgumentError(f"gemm must be 'f64', 'f32' or 'bf16x3' (got {gemm!r})")
    check(ctx.lib.kpm_2d(ctx.handle, Hd.handle, Jad.handle, Jbd.handle, NC, NR, NH,
                         dptr(psl), dptr(psr), K_left, K_right,
                         parity, 1 if mn_sym else 0, prec_code(precision), gm,
                         dptr(out)), "kpm_2d!")
    mu[...] = out
    return None


def kpm_2d(H, Ja, Jb, NC: int, NR: int, NH: int | None = None, *, psi_in=None, psi_in_l=None,
           psi_in_r=None, arr_size="auto", right_block="auto", workspace_bytes="auto",
           moment_parity="NONE", verbose: int = 0, mn_sym: bool = False, precision="f64",
           gemm="f64"):
    """kpm_2d (src/moment.jl:812-885): the NC x NC complex moment table."""
    NC, NR = int(NC), int(NR)
    NH = _nh(H) if NH is None else int(NH)
    mu = np.zeros((NC, NC), dtype=np.complex128, order="F")
    kw = dict(arr_size=arr_size, right_block=right_block, workspace_bytes=workspace_bytes,
              verbose=verbose, moment_parity=moment_parity, mn_sym=mn_sym, precision=precision,
              gemm=gemm)
    if psi_in is None and psi_in_l is None and psi_in_r is None:
        kpm_2d_(H, Ja, Jb, NC, NR, NH, mu, **kw)
    elif psi_in is not None and psi_in_l is None and psi_in_r is None:
        kpm_2d_(H, Ja, Jb, NC, NR, NH, mu, psi_in, psi_in, **kw)
    elif psi_in is None and psi_in_l is not None and psi_in_r is not None:
        kpm_2d_(H, Ja, Jb, NC, NR, NH, mu, psi_in_l, psi_in_r, **kw)
    else:
        raise NotImplementedError("unimplemented")
    return mu


# --------------------------------------------------------------------------
# kpm_3d


def kpm_3d_(H, Ja, Jb, Jg, NC: int, NR: int, NH: int, mu: np.ndarray, psi_in_l, psi_in_r, *,
            arr_size="auto", right_block="auto", workspace_bytes="auto", verbose: int = 0,
            precision="f64", gemm="f64"):
    """kpm_3d! (src/moment.jl:1106-1227): μ[n,m,p] = (1/NR) Σ_r
    <T_{m-1} ψl | Jβ T_{n-1} Jγ T_{p-1} Jα ψr>.

    Each of the NC kpm_2d passes uses the engine's slot plan
    (engine_blocking): with ':auto' sizes every left and right slot stays
    resident when the workspace allows, so each pass is one MFMA
    contraction. The reference's default (arr_size = 3, right_block =
    min(NC, 16)) recomputes the right chain NC/3 times per pass; explicit
    sizes are kept as given, and any blocking gives the same table to
    rounding (test/kpm2d_blocking_test.jl:97-117)."""
    plan = engine_blocking(NH, NR, NC, precision=precision, workspace_bytes=workspace_bytes,
                           arr_size=arr_size, right_block=right_block)
    arr_size, right_block = plan["arr_size"], plan["right_block"]
    psl = _probe(psi_in_l, NH, NR, "psi_in_l")
    psr = _probe(psi_in_r, NH, NR, "psi_in_r")
    if arr_size < 2:
        raise ArgumentError("ψall_l must have at least two slots")
    ctx = get_context()
    Hd, Jad, Jbd, Jgd = (to_device(X, ctx) for X in (H, Ja, Jb, Jg))
    # the engine writes every element: fill the caller's array in place when
    # it is a column-major ComplexF64 NC^3 array, else go through a temporary
    direct = (isinstance(mu, np.ndarray) and mu.dtype == np.complex128
              and mu.shape == (NC, NC, NC) and mu.flags.f_contiguous)
    out = mu if direct else np.empty((NC, NC, NC), dtype=np.complex128, order="F")
    gm = _GEMM_MODES[str(gemm).lower()]
    check(ctx.lib.kpm_3d(ctx.handle, Hd.handle, Jad.handle, Jbd.handle, Jgd.handle, NC, NR, NH,
                         dptr(psl), dptr(psr), int(arr_size),
                         right_block, prec_code(precision), gm, dptr(out)),
          "kpm_3d!")
    if not direct:
        mu[...] = out


def kpm_3d(H, Ja, Jb, Jg, NC: int, NR: int, NH: int | None = None, *, arr_size="auto",
           right_block="auto", workspace_bytes="auto", verbose: int = 0, psi_in_l=None,
           psi_in_r=None, psi_in=None, precision="f64", gemm="f64"):
    """kpm_3d (src/moment.jl:1229-1280): the NC x NC x NC table."""
    import warnings
    NC, NR = int(NC), int(NR)
    NH = _nh(H) if NH is None else int(NH)
    mu = np.zeros((NC, NC, NC), dtype=np.complex128, order="F")
    if psi_in is not None:
        if psi_in_l is not None or psi_in_r is not None:
            warnings.warn("`psi_in_l`, `psi_in_r` and `psi_in` are simoutaneously set. "
                          "Taking `psi_in` and discarding the others")
        psi_in_l = psi_in_r = psi_in
    elif psi_in_l is not None or psi_in_r is not None:
        if psi_in_l is None or psi_in_r is None:
            warnings.warn("only one of `psi_in_l` and `psi_in_r` is set. Setting them as the same.")
            psi_in_l = psi_in_l if psi_in_l is not None else psi_in_r
            psi_in_r = psi_in_l
    else:
        psi_in_l = random_probes(NH, NR)
        psi_in_r = psi_in_l
    kpm_3d_(H, Ja, Jb, Jg, NC, NR, NH, mu, psi_in_l, psi_in_r, arr_size=arr_size,
            right_block=right_block, workspace_bytes=workspace_bytes, verbose=verbose,
            precision=precision, gemm=gemm)
    return mu
