"""Build, load and drive the host-side window-cover check (window_check.hip).

``check(...)`` walks the workgroups of a forward or adjoint launch for one
RCA half with the kernels' own tile / window / pass code and returns the
counts: ``violations`` (in-range updates whose taps are not inside the rows of
the pass that takes them, or that no pass or two passes take) and
``skipped_in_aperture`` (lines skipped although a voxel of the tile lies in
their aperture) must both be zero.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

from simu_rca_ab_amd import configs as C
from simu_rca_ab_amd.plan import prepare_geometry

HERE = Path(__file__).resolve().parent
SRC = HERE / "window_check.hip"
DEPS = [SRC, HERE.parent / "simu_rca_ab_amd" / "csrc" / "rca_tile.h"]
OUT = HERE / "_build" / "libwindow_check.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

_F = ctypes.POINTER(ctypes.c_float)


class WinCase(ctypes.Structure):
    _fields_ = [(n, _F) for n in ("gx", "gy", "gz", "ex", "ey", "ez", "vx", "vy", "vz", "d0", "ab", "abt")] + \
        [(n, ctypes.c_int) for n in ("ab_mode", "n_pix", "n_ch", "n_tx", "n_sample", "depth_run",
                                     "receive_x")] + \
        [(n, ctypes.c_float) for n in ("fs", "fd", "c", "t0", "fnumber")] + \
        [(n, ctypes.c_int) for n in ("lo", "hi", "kernel", "sbytes", "fb", "v", "compound",
                                     "lds_budget", "shrink")] + \
        [("blocks", ctypes.POINTER(ctypes.c_longlong)), ("n_blocks", ctypes.c_longlong),
         ("n_threads", ctypes.c_int)]


class WinStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_longlong) for n in ("blocks", "lines", "lines_skipped", "updates_in_range",
                                                 "violations", "skipped_in_aperture",
                                                 "multi_pass_lines", "max_passes")] + \
        [(n, ctypes.c_int) for n in ("tc", "wcap", "step", "e")] + \
        [("first_bx", ctypes.c_longlong)] + \
        [(n, ctypes.c_int) for n in ("first_c", "first_t", "first_base", "first_cb", "first_end",
                                     "first_kind")]


_lib = None
TAPS = {"quadratic": (1, 1), "lanczos": (2, 3), "lanczos3": (2, 3), "lanczos4": (3, 4)}


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not OUT.exists() or OUT.stat().st_mtime < max(p.stat().st_mtime for p in DEPS):
        OUT.parent.mkdir(exist_ok=True)
        tmp = OUT.with_suffix(f".{os.getpid()}.tmp")
        subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
                        "-pthread", "-o", str(tmp), str(SRC)], check=True)
        os.replace(tmp, OUT)
    lib = ctypes.CDLL(str(OUT))
    lib.rca_window_check.argtypes = [ctypes.POINTER(WinCase), ctypes.POINTER(WinStats)]
    lib.rca_window_check.restype = ctypes.c_int
    _lib = lib
    return lib


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def check(cfg: C.RcaConfig, half_index=0, *, kernel="forward", sbytes=8, fb=1, v=1, compound=True,
          lds_kib=0, grid=None, depth_run=None, ab=None, abt=None, ab_mode=0, blocks=None, shrink=0,
          threads=None, interp=None):
    """Window-cover counts of one launch of ``cfg``'s half ``half_index``.
    ``grid``: (X, Y, Z) flattened voxel coordinates (default the config's
    meshgrid volume, depth-fastest); ``depth_run`` defaults to Nz."""
    lib = load()
    h = cfg.halves[half_index]
    G = C.geometry(cfg, h)
    prep = prepare_geometry(G, C.transmit(h))
    if grid is None:
        grid = tuple(np.asarray(G[k]).ravel(order="F") for k in ("gridBfX", "gridBfY", "gridBfZ"))
        if depth_run is None:
            depth_run = cfg.nz
    keep = [_f32(g) for g in grid] + [_f32(a) for a in prep["ele"]] + [_f32(a) for a in prep["vs"]] + \
        [_f32(prep["d0"])]
    arr = keep + [_f32(ab) if ab is not None else None, _f32(abt) if abt is not None else None]
    wc = WinCase()
    for name, a in zip(("gx", "gy", "gz", "ex", "ey", "ez", "vx", "vy", "vz", "d0", "ab", "abt"), arr):
        setattr(wc, name, a.ctypes.data_as(_F) if a is not None else None)
    wc.ab_mode = ab_mode
    wc.n_pix, wc.n_ch, wc.n_tx, wc.n_sample = keep[0].size, cfg.n_elem, len(h.theta), cfg.n_sample
    wc.depth_run = depth_run or 0
    wc.receive_x = int(h.receive_x)
    wc.fs, wc.fd, wc.c, wc.t0, wc.fnumber = cfg.fs, cfg.fd, cfg.c, cfg.t0, cfg.fnumber
    wc.lo, wc.hi = TAPS[interp or cfg.interpolation]
    wc.kernel = 0 if kernel == "forward" else 1
    wc.sbytes, wc.fb, wc.v, wc.compound = sbytes, fb, v, int(compound)
    wc.lds_budget = lds_kib * 1024
    wc.shrink = shrink
    if blocks is not None:
        bl = np.ascontiguousarray(np.asarray(blocks, dtype=np.int64))
        arr.append(bl)
        wc.blocks = bl.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
        wc.n_blocks = bl.size
    wc.n_threads = threads or max(1, min(16, len(os.sched_getaffinity(0))))
    st = WinStats()
    rc = lib.rca_window_check(ctypes.byref(wc), ctypes.byref(st))
    assert rc == 0
    return {f: getattr(st, f) for f, _ in WinStats._fields_}
