"""The C-ABI libraries load without a GPU and export every function their
headers in include/ declare (no compute calls)."""
import ctypes as C
import os
import re

import pytest

from tetfit_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    s = open(os.path.join(ROOT, "include", header)).read()
    s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
    s = re.sub(r"//[^\n]*", "", s)
    s = re.sub(r"#define[^\n]*", "", s)
    return sorted(set(re.findall(r"\b(tf_[a-z0-9_]+)\s*\(", s)))


@pytest.mark.parametrize("header,lib", [("tetfit_hip.h", _ffi.HIP_LIB), ("tetfit_host.h", _ffi.HOST_LIB)])
def test_exports_match_header(header, lib):
    names = declared(header)
    assert len(names) >= 10
    L = C.CDLL(lib)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_device_count_without_gpu_is_callable():
    n = _ffi.hip_lib().tf_device_count()
    assert n >= 0
