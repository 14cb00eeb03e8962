"""The C-ABI library loads and exports every symbol include/rca_das.h,
include/rca_svd.h and include/rca_sim.h declare; argument validation returns error codes before any
GPU work (CPU-only)."""
import ctypes
import re
from pathlib import Path

import pytest

from simu_rca_ab_amd import _lib

INCLUDE = Path(__file__).resolve().parents[1] / "include"
HEADERS = (INCLUDE / "rca_das.h", INCLUDE / "rca_svd.h", INCLUDE / "rca_sim.h")


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


def header_struct_fields(name, header="rca_das.h"):
    text = re.sub(r"/\*.*?\*/", "", (INCLUDE / header).read_text(), flags=re.S)
    body = re.search(r"typedef struct \{(.*?)\}\s*" + name + ";", text, flags=re.S).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"(const\s+)?(\w+)\s*(\*?)\s*(.*)", decl)
        kind = "ptr" if m.group(3) else m.group(2)
        for n in m.group(4).split(","):
            fields.append((n.strip().lstrip("*").strip(), kind))
    return fields


@pytest.mark.parametrize("cls,name,header", [(_lib.RcaDasDesc, "rca_das_desc", "rca_das.h"),
                                             (_lib.RcaSimReceiveArgs, "rca_sim_receive_args",
                                              "rca_sim.h")])
def test_struct_mirrors_match_headers(cls, name, header):
    kinds = {ctypes.c_void_p: "ptr", ctypes.c_int: "int", ctypes.c_float: "float"}
    mirror = [(n, kinds[t]) for n, t in cls._fields_]
    assert mirror == header_struct_fields(name, header)


def test_sim_validation_without_gpu():
    lib = _lib.load()
    rc = lib.rca_sim_txfield(*([None] * 8), 1.0, 1.0, 1e4, 0.0, 0.0, 1540.0, 1, 1, 1, 1, 1, None, None)
    assert rc == _lib.RCA_ERR_INVALID and b"NULL" in lib.rca_last_error()
    rc = lib.rca_sim_receive(None, None)
    assert rc == _lib.RCA_ERR_INVALID
    rc = lib.rca_sim_spectra_to_raw(None, 1, 64, 32, 4, None, None)
    assert rc == _lib.RCA_ERR_INVALID


def _desc_with_dummy_pointers(mode):
    d = _lib.RcaDasDesc()
    for n, t in _lib.RcaDasDesc._fields_:
        if t is ctypes.c_void_p and not n.startswith("ab_law"):
            setattr(d, n, 256)   # never dereferenced: validation fails first
    d.n_pix, d.n_channel, d.n_tx, d.n_sample = 10, 4, 2, 64
    d.speed_of_sound, d.sampling_frequency, d.fnumber = 1540.0, 20e6, 1.0
    d.ab_law_mode = mode
    return d


@pytest.mark.parametrize("mode,msg", [(1, b"ab_law_rx is NULL"), (2, b"ab_law_rx is NULL"),
                                      (3, b"both NULL"), (4, b"unknown aberration-law mode"),
                                      (-1, b"unknown aberration-law mode")])
def test_ablaw_validation_without_gpu(mode, msg):
    lib = _lib.load()
    d = _desc_with_dummy_pointers(mode)
    rc = lib.rca_forward(ctypes.byref(d), None, 0, None, 1, None, None, 0, None)
    assert rc == _lib.RCA_ERR_INVALID
    assert msg in lib.rca_last_error()
    rc = lib.rca_adjoint(ctypes.byref(d), None, 1, 1, None, 0, None)
    assert rc == _lib.RCA_ERR_INVALID
    assert msg in lib.rca_last_error()
