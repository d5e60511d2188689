"""The C-ABI libraries load and export every symbol their headers declare
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:fdtd|wop|fdtdb|wdump|nf2ff)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("wipems_fdtd.h", "wipems_amd/libwipems_fdtd.so"),
                                        ("wipems_operator.h", "wipems_amd/libwipems_operator.so"),
                                        ("wipems_dump.h", "wipems_amd/libwipems_dump.so"),
                                        ("wipems_nf2ff.h", "wipems_amd/libwipems_fdtd.so")])
def test_exports(header, lib):
    path = os.path.join(ROOT, lib)
    assert os.path.exists(path), "%s not built" % lib
    L = ctypes.CDLL(path)
    names = declared(header)
    assert names, "no declarations parsed from %s" % header
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_string():
    L = ctypes.CDLL(os.path.join(ROOT, "wipems_amd/libwipems_fdtd.so"))
    L.fdtd_version.restype = ctypes.c_char_p
    assert b"gfx950" in L.fdtd_version()


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from wipems_amd import _native
    monkeypatch.setattr(_native, "PKG", str(tmp_path))
    monkeypatch.setattr(_native, "_libs", {})
    with pytest.raises(_native.FdtdError):
        _native.engine_lib()
