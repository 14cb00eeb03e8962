"""The C-ABI library loads and exports every symbol include/rca_das.h declares;
argument validation returns error codes before any GPU work (CPU-only)."""
import ctypes
import re
from pathlib import Path

import pytest

from simu_rca_ab_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "rca_das.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)))


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
