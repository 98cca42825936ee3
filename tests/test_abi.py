"""The C-ABI libraries load and export every entry point their headers declare (no compute
calls: this runs without a GPU)."""
from __future__ import annotations

import ctypes as C
import os
import re

import pytest

from sayakaamr_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sayaka_dev_\w+|sy_tree_\w+|sy_mac_\w+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("sayaka_amd.h", N.DEV_LIB), ("sayaka_host.h", N.HOST_LIB),
                                        ("sayaka_amd.h", N.DEV_PHASES_LIB), ("sayaka_phases.h", N.DEV_PHASES_LIB)])
def test_exports(header, lib):
    names = declared(header)
    assert names, header
    so = C.CDLL(lib)
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_device_lib_targets_gfx950():
    """The HIP library carries a gfx950 code object (offload bundle)."""
    blob = open(N.DEV_LIB, "rb").read()
    assert b"gfx950" in blob


def test_last_error_without_gpu_call():
    so = C.CDLL(N.DEV_LIB)
    so.sayaka_dev_last_error.restype = C.c_char_p
    assert isinstance(so.sayaka_dev_last_error(), bytes)


def test_missing_library_raises(monkeypatch):
    from sayakaamr_amd import device as D

    monkeypatch.setattr(N, "DEV_LIB", "/nonexistent/libsayaka_dev.so")
    monkeypatch.setattr(D, "_dev", None)
    with pytest.raises(N.NativeLibraryMissing):
        D.dev()
