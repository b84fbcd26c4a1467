"""libkpm_hip.so loads on a machine without a GPU and exports every entry
point include/kpm_hip.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

import kpmhip
from kpmhip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kpm_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kpm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header():
    lib = _lib.load_library()
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # every declared symbol has a ctypes signature in the binding
    assert set(syms) <= set(_lib.EXPORTED), sorted(set(syms) - set(_lib.EXPORTED))


def test_abi_version_and_error_string():
    lib = _lib.load_library()
    assert lib.kpm_abi_version() == 1
    assert isinstance(_lib.last_error(), str)


def test_invalid_arguments_are_rejected_without_gpu():
    lib = _lib.load_library()
    # null outputs are rejected before any device work
    assert lib.kpm_ctx_create(0, None) == _lib.KPM_EINVAL
    K = ctypes.c_int64()
    rc = lib.kpm_2d_blocking(None, 40, 3, 50, 10, 0, 0, ctypes.byref(K), ctypes.byref(K),
                             ctypes.byref(K))
    assert rc == _lib.KPM_EINVAL  # budget too small -> ArgumentError
    with pytest.raises(kpmhip.ArgumentError):
        _lib.check(rc)


def test_blocking_through_abi_matches_python():
    lib = _lib.load_library()
    NH, NR, NC = 40, 3, 50
    one = NH * NR * 16
    KL, KR, B = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    assert lib.kpm_2d_blocking(None, NH, NR, NC, 30 * one, 0, 4, ctypes.byref(KL), ctypes.byref(KR),
                               ctypes.byref(B)) == 0
    py = kpmhip.kpm_2d_blocking(NH, NR, NC, workspace_bytes=30 * one, right_block=4)
    assert (KL.value, KR.value, B.value) == (py["arr_size"], py["right_block"], py["bytes"])
