"""The C-ABI libraries load and export every function their headers declare."""
import ctypes
import os
import re

import pytest

from phonomc_amd import runtime as rt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(pmcs?_\w+)\s*\(", text, flags=re.M)))


@pytest.mark.parametrize("header,lib", [("pmc_capi.h", "libpmc_device.so"), ("pmc_solver.h", "libpmc_host.so"),
                                        ("pmc_surrogate.h", "libpmc_device.so")])
def test_every_declared_symbol_is_exported(header, lib):
    names = declared(header)
    assert len(names) > 8
    handle = ctypes.CDLL(os.path.join(rt.BUILD, lib))
    missing = [n for n in names if not hasattr(handle, n)]
    assert not missing, missing


def test_runtime_refuses_missing_library(monkeypatch, tmp_path):
    monkeypatch.setattr(rt, "BUILD", str(tmp_path))
    monkeypatch.setattr(rt, "_lib", None)
    with pytest.raises(RuntimeError, match="missing"):
        rt.lib()
