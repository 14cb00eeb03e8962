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
    d.n_sample = (1 << 27) + 1          # sample offsets are 32-bit with a 2^28 clamp
    rc = lib.rca_forward(ctypes.byref(d), None, 0, None, 1, None, None, 0, None)
    assert rc == _lib.RCA_ERR_UNSUPPORTED and b"nSample" in lib.rca_last_error()
    d.n_sample = 64
    with pytest.raises(_lib.RcaError):
        _lib.set_tuning(3, 0, 0)
    _lib.set_tuning(0, 0, 0)
    rc = lib.rca_adjoint(None, None, 1, 1, None, 0, None)
    assert rc == _lib.RCA_ERR_INVALID


def test_deterministic_adjoint_validation_without_gpu():
    """rca_adjoint_det: workspace size (256 B header + 16 B per complex gradient
    sample) and the argument checks that return before any GPU work."""
    lib = _lib.load()
    d = _lib.RcaDasDesc()
    d.n_pix, d.n_channel, d.n_tx, d.n_sample = 10, 4, 3, 64
    assert lib.rca_adjoint_det_workspace_bytes(ctypes.byref(d), 5) == 256 + 16 * 5 * 3 * 4 * 64
    assert lib.rca_adjoint_det_workspace_bytes(ctypes.byref(d), -1) == 0
    assert lib.rca_adjoint_det_workspace_bytes(None, 1) == 0
    need = lib.rca_adjoint_det_workspace_bytes(ctypes.byref(d), 2)
    rc = lib.rca_adjoint_det(ctypes.byref(d), None, 1, 2, None, 0, None, need, None)
    assert rc == _lib.RCA_ERR_INVALID and b"workspace" in lib.rca_last_error()
    rc = lib.rca_adjoint_det(ctypes.byref(d), None, 1, 2, None, 0, ctypes.c_void_p(0x1008), need, None)
    assert rc == _lib.RCA_ERR_INVALID and b"aligned" in lib.rca_last_error()
    rc = lib.rca_adjoint_det(ctypes.byref(d), None, 1, 2, None, 0, ctypes.c_void_p(0x1000), need - 1, None)
    assert rc == _lib.RCA_ERR_INVALID and b"needed" in lib.rca_last_error()
    rc = lib.rca_adjoint_det(ctypes.byref(d), None, 1, 2, None, 0, ctypes.c_void_p(0x1000), need, None)
    assert rc == _lib.RCA_ERR_INVALID   # null grid pointers (check_desc), before any launch


def test_svd_validation_without_gpu():
    lib = _lib.load()
    assert lib.rca_svd_workspace_bytes(0, 4) == 0
    assert lib.rca_svd_workspace_bytes(1000, 10) >= 16 * 1000 * 10
    rc = lib.rca_svd_filter(None, None, 1000, 10, None, None, None, None, 0, None)
    assert rc == _lib.RCA_ERR_INVALID
    assert b"NULL" in lib.rca_last_error()


def header_struct_fields(name, header="rca_das.h"):
    text = re.sub(r"/\*.*?\*/", "", (INCLUDE / header).read_text(), flags=re.S)
    body = re.search(r"typedef struct \{((?:(?!typedef struct).)*?)\}\s*" + name + ";", text,
                     flags=re.S).group(1)
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
                                             (_lib.RcaPlanParams, "rca_plan_params", "rca_das.h"),
                                             (_lib.RcaSimReceiveArgs, "rca_sim_receive_args",
                                              "rca_sim.h")])
def test_struct_mirrors_match_headers(cls, name, header):
    kinds = {ctypes.c_void_p: "ptr", ctypes.c_int: "int", ctypes.c_float: "float",
             ctypes.c_double: "double"}
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


def _plan_params(G, T, cfg, keep):
    """rca_plan_params of a Geometry/Transmit pair (host FP64 arrays)."""
    import numpy as np
    p = _lib.RcaPlanParams()

    def arr(x):
        a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel(order="F"))
        keep.append(a)
        return a.ctypes.data

    p.grid_x, p.grid_y, p.grid_z = (arr(G[k]) for k in ("gridBfX", "gridBfY", "gridBfZ"))
    p.ele_x, p.ele_y, p.ele_z = (arr(G[k]) for k in ("posEleX", "posEleY", "posEleZ"))
    if "thetaX" in T:
        p.theta_x, p.theta_y = arr(T["thetaX"]), arr(T["thetaY"])
        p.n_tx = np.size(T["thetaX"])
    else:
        p.vs_x, p.vs_y, p.vs_z = (arr(T[k]) for k in ("virtualSourceX", "virtualSourceY", "virtualSourceZ"))
        p.n_tx = np.size(T["virtualSourceX"])
    p.n_pix = np.size(G["gridBfX"])
    p.n_channel = np.size(G["posEleX"])
    p.n_sample = cfg.n_sample
    p.receive_x = int(bool(T.get("flagReceiveX", True)))
    if "apertureX" in G:
        p.has_aperture = 1
        p.aperture_x, p.aperture_y, p.aperture_z = G["apertureX"], G["apertureY"], G["apertureZ"]
    p.sampling_frequency, p.demodulation_frequency = cfg.fs, cfg.fd
    p.speed_of_sound, p.start_time, p.fnumber = cfg.c, cfg.t0, cfg.fnumber
    return p


@pytest.mark.parametrize("half_index,sources,apertures", [(0, False, True), (1, False, False),
                                                          (0, True, False)])
def test_prepare_geometry_matches_host_prep(half_index, sources, apertures):
    """The C host prep equals the Python drop-in's prep (RCA.m:491-558)."""
    import numpy as np
    from simu_rca_ab_amd import configs as C
    from simu_rca_ab_amd.plan import prepare_geometry
    from tests import helpers as H
    cfg = C.scaled(C.get("cfg1"), nx=4, ny=3, nz=8)
    G, T, P, Op = H.case(cfg, half_index)
    G = dict(G, posEleX=np.asarray(G["posEleX"]) + 3e-4, posEleZ=np.asarray(G["posEleZ"]) + 1e-5)
    if not apertures:
        G = {k: v for k, v in G.items() if not k.startswith("aperture")}
    if sources:
        pr = prepare_geometry(G, T)
        T = {"virtualSourceX": pr["vs"][0] * 1.0, "virtualSourceY": pr["vs"][1] * 1.0,
             "virtualSourceZ": pr["vs"][2] * 1.0, "flagReceiveX": T["flagReceiveX"]}
    want = prepare_geometry(G, T)
    keep = []
    p = _plan_params(G, T, cfg, keep)
    nc, nt = p.n_channel, p.n_tx
    ele = np.zeros(3 * nc, np.float32)
    vs = np.zeros(3 * nt, np.float32)
    d0 = np.zeros(nt, np.float32)
    ap = np.zeros(3, np.float32)
    lib = _lib.load()
    rc = lib.rca_prepare_geometry(ctypes.byref(p), ele.ctypes.data, vs.ctypes.data, d0.ctypes.data,
                                  ap.ctypes.data)
    assert rc == 0, lib.rca_last_error()
    # equal to the last bit, or one FP32 ulp where the FP64 sums round differently
    np.testing.assert_allclose(ele.reshape(3, nc), np.stack(want["ele"]), rtol=2e-7, atol=1e-12)
    np.testing.assert_allclose(vs.reshape(3, nt), np.stack(want["vs"]), rtol=2e-7, atol=1e-12)
    np.testing.assert_allclose(d0, want["d0"], rtol=2e-7)
    np.testing.assert_allclose(ap, want["aperture"], rtol=2e-7)


def test_plan_validation_without_gpu():
    lib = _lib.load()
    plan = ctypes.c_void_p()
    assert lib.rca_plan_create(None, ctypes.byref(plan)) == _lib.RCA_ERR_INVALID
    p = _lib.RcaPlanParams()
    p.n_pix, p.n_channel, p.n_tx, p.n_sample = 10, 4, 2, 64
    assert lib.rca_plan_create(ctypes.byref(p), ctypes.byref(plan)) == _lib.RCA_ERR_INVALID
    assert b"element" in lib.rca_last_error()
    import numpy as np
    e = np.zeros(4)
    th = np.zeros(2)
    p.ele_x = p.ele_y = p.ele_z = e.ctypes.data
    assert lib.rca_plan_create(ctypes.byref(p), ctypes.byref(plan)) == _lib.RCA_ERR_INVALID
    assert b"either" in lib.rca_last_error()     # neither tilts nor sources
    p.theta_x = p.theta_y = p.vs_x = p.vs_y = p.vs_z = th.ctypes.data
    assert lib.rca_plan_create(ctypes.byref(p), ctypes.byref(plan)) == _lib.RCA_ERR_INVALID
    assert b"either" in lib.rca_last_error()     # both
    assert lib.rca_plan_destroy(None) == 0
