"""ORACLE — test infrastructure only.

Only tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker (or the timed CPU
reference path).  The product never routes through it.

Python side of the CPU restatement of the RCA delay-and-sum beamformer:
``rca_oracle.cpp`` holds the kernel body (``bfVirtualSourceGeneric_RCA.cu``)
and the geometry prep (``bfVirtualSourceGeneric_RCA.m:491-558``); this file
restates the wrapper's data handling and output shaping of
``MATLAB/lib/beamformer/RCA/bfVirtualSourceGeneric_RCA.m``:

* data conversion             RCA.m:468-485 (real: IQ(n) = RF(2n) - j RF(2n+1))
* reshape to [Ns, nCh, nTx, nFrame]   RCA.m:576-578
* per-tilt kernel, then tilt sum (compounding) and mask sum   RCA.m:654-711
* output reshape [s1, s2, s3, nTx|1, nFrame]   RCA.m:715-722

Frame chunking (RCA.m:587-594) only splits launches; the result of one call
does not depend on it, so the oracle runs all frames at once.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "build" / "liboracle_rca.so"

INTERP_QUADRATIC = 0
INTERP_LANCZOS_REF = 1
INTERP_LANCZOS = 2

_lib = None


def load():
    """Load the prebuilt oracle library (``make -C oracle`` builds it)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise FileNotFoundError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
    lib = ctypes.CDLL(str(LIB_PATH))
    fp = ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    I, F = ctypes.c_int, ctypes.c_float
    geo = [fp] * 10 + [I, F, F, F, F, F, F, F, F, I, I, I, I, I, I, I, I, fp, I, fp]
    lib.oracle_rca_forward_f64.argtypes = [dp, dp, dp] + geo + [I, I]
    lib.oracle_rca_forward_f32.argtypes = [fp, fp, fp] + geo + [I, I]
    lib.oracle_rca_adjoint_f64.argtypes = [dp, dp, I] + geo + [I]
    lib.oracle_rca_prepare.argtypes = [dp, dp, dp, I, dp, I, dp, dp, dp, dp, dp, I] + [fp] * 8
    lib.oracle_interp_weights.argtypes = [I, I, ctypes.c_double, dp]
    for f in ("oracle_rca_forward_f64", "oracle_rca_forward_f32", "oracle_rca_adjoint_f64",
              "oracle_rca_prepare", "oracle_max_threads", "oracle_interp_weights"):
        getattr(lib, f).restype = ctypes.c_int
    _lib = lib
    return lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel(order="F"))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))


def prepare(Geometry: dict, Transmit: dict) -> dict:
    """Element centring, apertures, virtual sources and minDist (RCA.m:491-558)."""
    lib = load()
    px, py, pz = (_f64(Geometry[k]) for k in ("posEleX", "posEleY", "posEleZ"))
    nch = px.size
    ap = np.array([float(Geometry.get(k, np.nan)) for k in ("apertureX", "apertureY", "apertureZ")])
    use_theta = "thetaX" in Transmit
    if use_theta:
        tx, ty = _f64(Transmit["thetaX"]), _f64(Transmit["thetaY"])
        ntx = tx.size
        vsx = vsy = vsz = np.zeros(1)
    else:
        vsx, vsy, vsz = (_f64(Transmit[k]) for k in ("virtualSourceX", "virtualSourceY",
                                                      "virtualSourceZ"))
        ntx = vsx.size
        tx = ty = np.zeros(1)
    out = {k: np.zeros(nch, np.float32) for k in ("ex", "ey", "ez")}
    out.update({k: np.zeros(max(ntx, 1), np.float32) for k in ("vx", "vy", "vz", "d0")})
    out["ap"] = np.zeros(3, np.float32)
    fp = ctypes.c_float
    rc = lib.oracle_rca_prepare(_p(px, ctypes.c_double), _p(py, ctypes.c_double),
                                _p(pz, ctypes.c_double), nch, _p(ap, ctypes.c_double),
                                int(use_theta), _p(tx, ctypes.c_double), _p(ty, ctypes.c_double),
                                _p(vsx, ctypes.c_double), _p(vsy, ctypes.c_double),
                                _p(vsz, ctypes.c_double), ntx,
                                *(_p(out[k], fp) for k in ("ex", "ey", "ez", "ap", "vx", "vy",
                                                           "vz", "d0")))
    if rc:
        raise ValueError("oracle_rca_prepare failed")
    for k in ("vx", "vy", "vz", "d0"):
        out[k] = out[k][:ntx]
    return out


def _geo_args(grid, prep, params, n_frame, n_sample, interp, lanczos_a, mask_tx, ablaw,
              ablaw_mode, ablaw_tx=None):
    fp = ctypes.c_float
    gx, gy, gz = grid
    ablaw_arr = None if ablaw is None else _f32(ablaw)
    ablaw_tx_arr = None if ablaw_tx is None else _f32(ablaw_tx)
    keep = [gx, gy, gz, ablaw_arr, ablaw_tx_arr]
    args = [_p(gx, fp), _p(gy, fp), _p(gz, fp)]
    args += [_p(prep[k], fp) for k in ("ex", "ey", "ez", "vx", "vy", "vz", "d0")]
    args += [int(params["receive_x"]), float(prep["ap"][0]), float(prep["ap"][1]),
             float(prep["ap"][2]), float(params["fs"]), float(params["fd"]), float(params["c"]),
             float(params["t0"]), float(params["fnumber"]), gx.size, prep["ex"].size,
             prep["vx"].size, int(n_frame), int(n_sample), int(interp), int(lanczos_a),
             int(mask_tx), None if ablaw_arr is None else _p(ablaw_arr, fp), int(ablaw_mode),
             None if ablaw_tx_arr is None else _p(ablaw_tx_arr, fp)]
    return args, keep


def kernel_forward(data_il, grid, prep, params, interp=INTERP_QUADRATIC, lanczos_a=3,
                   mask_tx=False, precision=64, nosum=False, ablaw=None, ablaw_mode=0,
                   nthreads=0, ablaw_tx=None):
    """Per-tilt forward of the reference kernel (RCA.cu:103-391).

    data_il: interleaved real array [nFrame][nTx][nCh][2*Ns] (any shape with
    that C-order size); params["n_sample"] is Ns.  grid: (gx, gy, gz) float32
    flat arrays.  Returns
    (dataBf complex [nFrame, nTx, nPix] (x nCh first if nosum), mask [nTx, nPix]).
    """
    lib = load()
    ntx, nch, npix = prep["vx"].size, prep["ex"].size, grid[0].size
    dt = np.float64 if precision == 64 else np.float32
    ct = ctypes.c_double if precision == 64 else ctypes.c_float
    data = np.ascontiguousarray(np.asarray(data_il, dtype=dt).ravel())
    n_sample = int(params["n_sample"])
    n_frame = data.size // (2 * n_sample * nch * ntx)
    nout = 2 * npix * ntx * n_frame * (nch if nosum else 1)
    out = np.zeros(nout, dt)
    mask = np.zeros(npix * ntx, dt)
    args, keep = _geo_args(grid, prep, params, n_frame, n_sample, interp, lanczos_a, mask_tx,
                           ablaw, ablaw_mode, ablaw_tx)
    fn = lib.oracle_rca_forward_f64 if precision == 64 else lib.oracle_rca_forward_f32
    rc = fn(_p(out, ct), _p(mask, ct), _p(data, ct), *args, int(nosum), int(nthreads))
    del keep
    if rc:
        raise ValueError("oracle forward rejected its arguments")
    out = out.view(np.complex128 if precision == 64 else np.complex64)
    shape = (nch, n_frame, ntx, npix) if nosum else (n_frame, ntx, npix)
    return out.reshape(shape), mask.reshape(ntx, npix)


def kernel_adjoint(g, grid, prep, params, n_frame, interp=INTERP_QUADRATIC, lanczos_a=3,
                   mask_tx=False, ablaw=None, ablaw_mode=0, nthreads=0, ablaw_tx=None):
    """FP64 adjoint of the per-tilt forward.  g: complex [nFrame, 1|nTx, nPix].
    Returns grad complex [nFrame, nTx, nCh, Ns]."""
    lib = load()
    ntx, nch, npix = prep["vx"].size, prep["ex"].size, grid[0].size
    ns = int(params["n_sample"])
    g = np.ascontiguousarray(np.asarray(g, dtype=np.complex128))
    gtx = g.shape[1]
    grad = np.zeros(2 * n_frame * ntx * nch * ns, np.float64)
    gr = g.view(np.float64).ravel()
    args, keep = _geo_args(grid, prep, params, n_frame, ns, interp, lanczos_a, mask_tx, ablaw,
                           ablaw_mode, ablaw_tx)
    dp = ctypes.c_double
    rc = lib.oracle_rca_adjoint_f64(_p(grad, dp), _p(gr, dp), int(gtx), *args, int(nthreads))
    del keep
    if rc:
        raise ValueError("oracle adjoint rejected its arguments")
    return grad.view(np.complex128).reshape(n_frame, ntx, nch, ns)


def interp_weights(interp, lanczos_a, d):
    """Oracle interpolation weights at fractional position d (taps -lo..hi)."""
    q = np.zeros(16, np.float64)
    n = load().oracle_interp_weights(int(interp), int(lanczos_a), float(d),
                                     q.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if n < 0:
        raise ValueError("bad interpolation")
    return q[:n].copy()


def _interp_from_options(Options):
    kind = Options.get("interpolation")
    if kind is None:
        return (INTERP_LANCZOS_REF, 3) if Options.get("flagInterpolationLanczos", False) else (
            INTERP_QUADRATIC, 3)
    kind = str(kind).lower()
    if kind == "quadratic":
        return INTERP_QUADRATIC, 3
    if kind == "lanczos":
        return INTERP_LANCZOS_REF, 3
    if kind.startswith("lanczos"):
        return INTERP_LANCZOS, int(kind[len("lanczos"):])
    raise ValueError(kind)


def bf(data, Geometry, Transmit, Param, Options=None, precision=64, nthreads=0):
    """Oracle of ``[dataBf, mask] = bfVirtualSourceGeneric_RCA(data, Geometry,
    Transmit, Param, Options)`` (RCA.m:1-724).  Returns numpy arrays in the
    reference's shapes, F-ordered: dataBf [s1, s2, s3, nTx|1, nFrame] and mask
    [s1, s2, s3, nTx|1]."""
    Options = dict(Options or {})
    data = np.asarray(data)
    shp = list(data.shape) + [1] * max(0, 4 - data.ndim)
    ns, ncx, ncy = shp[0], shp[1], shp[2]
    nch = ncx * ncy
    flat = data.ravel(order="F")
    if np.iscomplexobj(data):  # RCA.m:481-484
        il = np.empty(2 * flat.size, np.float64)
        il[0::2] = flat.real
        il[1::2] = flat.imag
        ns_iq = ns
    else:  # RCA.m:475-477
        il = flat.astype(np.float64).copy()
        il[1::2] = -il[1::2]
        ns_iq = ns // 2
    # single cast (RCA.m:599-636 casts everything the kernel sees to FP32)
    il = il.astype(np.float32).astype(np.float64 if precision == 64 else np.float32)
    prep = prepare(Geometry, Transmit)
    ntx = prep["vx"].size
    size_grid = list(np.shape(Geometry["gridBfX"])) + [1, 1, 1]
    size_grid = size_grid[:3]
    grid = tuple(_f32(Geometry[k]) for k in ("gridBfX", "gridBfY", "gridBfZ"))
    params = dict(receive_x=bool(Transmit.get("flagReceiveX", True)),
                  fs=np.float32(Param["samplingFrequency"]),
                  fd=np.float32(Param["demodulationFrequency"]),
                  c=np.float32(Param.get("speedOfSound", 1540)),
                  t0=np.float32(Param.get("startTime", 0)),
                  fnumber=np.float32(Param.get("fnumber", 1)), n_sample=ns_iq)
    interp, a = _interp_from_options(Options)
    mask_tx = bool(Options.get("flagMaskTransmit", True))
    if il.size % (2 * ns_iq * nch * ntx):
        raise ValueError("data size is not a multiple of nSample*nChannel*nTx")
    out, mask = kernel_forward(il, grid, prep, params, interp, a, mask_tx, precision,
                               nthreads=nthreads)
    nframe = out.shape[0]
    if Options.get("flagCompounding", True):  # RCA.m:690-697
        out = out.sum(axis=1, keepdims=True)
        mask = mask.sum(axis=0, keepdims=True)
    # [nFrame][T][nPix] C-order == [nPix, T, nFrame] F-order (RCA.m:715-722)
    dataBf = out.transpose(2, 1, 0).reshape(size_grid + [out.shape[1], nframe], order="F")
    maskr = mask.transpose(1, 0).reshape(size_grid + [mask.shape[0]], order="F")
    return dataBf, maskr
