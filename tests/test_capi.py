"""The C-ABI library loads and exports every symbol include/rca_das.h and
include/rca_svd.h declare; argument validation returns error codes before any
GPU work (CPU-only)."""
import ctypes
import re
from pathlib import Path

import pytest

from simu_rca_ab_amd import _lib

INCLUDE = Path(__file__).resolve().parents[1] / "include"
HEADERS = (INCLUDE / "rca_das.h", INCLUDE / "rca_svd.h")


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_header_declares_the_bound_symbols():
    assert set(declared_functions()) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_version():
    assert "gfx950" in _lib.version()


def test_validation_errors_without_gpu():
    lib = _lib.load()
    rc = lib.rca_forward(None, None, 0, None, 1, None, None, 0, None)
    assert rc == _lib.RCA_ERR_INVALID
    assert b"desc" in lib.rca_last_error()
    d = _lib.RcaDasDesc()
    d.n_pix, d.n_channel, d.n_tx, d.n_sample = 10, 4, 2, 64
    rc = lib.rca_forward(ctypes.byref(d), None, 0, None, 1, None, None, 0, None)
    assert rc == _lib.RCA_ERR_INVALID   # null grid pointers
    with pytest.raises(_lib.RcaError):
        _lib.set_tuning(3, 0, 0)
    _lib.set_tuning(0, 0, 0)
    rc = lib.rca_adjoint(None, None, 1, 1, None, 0, None)
    assert rc == _lib.RCA_ERR_INVALID


def test_svd_validation_without_gpu():
    lib = _lib.load()
    assert lib.rca_svd_workspace_bytes(0, 4) == 0
    assert lib.rca_svd_workspace_bytes(1000, 10) >= 16 * 1000 * 10
    rc = lib.rca_svd_filter(None, None, 1000, 10, None, None, None, None, 0, None)
    assert rc == _lib.RCA_ERR_INVALID
    assert b"NULL" in lib.rca_last_error()
