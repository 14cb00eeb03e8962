"""Drop-in RCA beamformer with the reference's call signature.

``bfVirtualSourceGeneric_RCA(data, Geometry, Transmit, Param[, Options])``
mirrors MATLAB/lib/beamformer/RCA/bfVirtualSourceGeneric_RCA.m:1-724 with the
MATLAB structs passed as dicts: same fields, defaults, warnings (identifier
``bfVirtualSourceGeneric_RCA:<tag>``), errors, real/complex data handling and
output shapes.  The beamforming itself runs in the gfx950 HIP kernels (through
:class:`simu_rca_ab_amd.plan.RcaPlan`); there is no CPU path.

Inputs may be numpy arrays (MATLAB index order, any memory order; results come
back as F-ordered numpy arrays, like ``gather``) or torch tensors (results
stay on the GPU as tensors with the same MATLAB shapes).
"""
from __future__ import annotations

import numbers
import warnings

import numpy as np
import torch

from ..plan import RcaPlan, interpolation_name

__all__ = ["bfVirtualSourceGeneric_RCA", "bfVirtualSourceGeneric_RCA_Nosum",
           "bfVirtualSourceGeneric_RCA_ab_1", "bfVirtualSourceGeneric_RCA_ab_2",
           "bfVirtualSourceGeneric_RCA_Nosum_ab", "bfVirtualSourceGeneric_RCA_TxRx",
           "RcaInputError", "RcaWarning"]

_ID = "bfVirtualSourceGeneric_RCA"


class RcaInputError(ValueError):
    """Input validation error; ``identifier`` follows the reference's error ids."""

    def __init__(self, tag: str, message: str):
        super().__init__(f"{_ID}:{tag}: {message}")
        self.identifier = f"{_ID}:{tag}"


class RcaWarning(UserWarning):
    pass


def _warn(tag, message):
    warnings.warn(f"{_ID}:{tag}: {message}", RcaWarning, stacklevel=3)


def _is_num(a):
    if isinstance(a, torch.Tensor):
        return not a.dtype == torch.bool
    if isinstance(a, (numbers.Number, np.number)) and not isinstance(a, bool):
        return True
    if isinstance(a, np.ndarray):
        return a.dtype.kind in "iufc"
    if isinstance(a, (list, tuple)):
        try:
            return np.asarray(a).dtype.kind in "iufc"
        except Exception:
            return False
    return False


def _numel(a):
    return a.numel() if isinstance(a, torch.Tensor) else int(np.size(a))


def _shape(a):
    return tuple(a.shape) if isinstance(a, (torch.Tensor, np.ndarray)) else np.shape(a)


def _is_scalar_num(a):
    return _is_num(a) and _numel(a) == 1


def _validate(data, Geometry, Transmit, Param, Options):
    """Checks of RCA.m:136-455, in the same order and with the same ids."""
    if not _is_num(data):
        raise RcaInputError("dataNotNumeric", "Input data ('data') must be a numeric array!")
    is_complex = data.is_complex() if isinstance(data, torch.Tensor) else np.iscomplexobj(data)
    if not is_complex:
        _warn("dataReal", "Input data ('data') are real. I am assuming that the I and Q "
                          "channels are interleaved.")
    if Options is not None and not isinstance(Options, dict):
        raise RcaInputError("OptionsNotStruct", "Options ('Options') must be a structure!")
    Options = dict(Options or {})
    Options.setdefault("flagInterpolationLanczos", False)
    if not isinstance(Options["flagInterpolationLanczos"], (bool, np.bool_)):
        raise RcaInputError("flagInterpolationLanczosInvalid",
                            "Interpolation method ('Options.flagInterpolationLanczos') is invalid; "
                            "it must be a boolean (true = Lanczos 5-lobes, false = quadratic)!")
    Options.setdefault("flagCompounding", True)
    if not isinstance(Options["flagCompounding"], (bool, np.bool_)):
        raise RcaInputError("flagCompoundingInvalid",
                            "Compounding flag ('Options.flagCompounding') must be a boolean!")
    Options.setdefault("flagMaskTransmit", True)
    if not isinstance(Options["flagMaskTransmit"], (bool, np.bool_)):
        raise RcaInputError("flagMaskTransmitInvalid",
                            "Transmit mask flag ('Options.flagMaskTransmit') must be a boolean!")
    Options.setdefault("decreaseFactorGPU", 1)
    dfg = Options["decreaseFactorGPU"]
    if not _is_scalar_num(dfg) or float(dfg) != np.floor(float(dfg)) or float(dfg) < 0:
        raise RcaInputError("decreaseFactorGPUInvalid",
                            "GPU decrease factor ('Options.decreaseFactorGPU') must be a positive integer!")
    if "interpolation" in Options:
        try:
            interpolation_name(Options)
        except ValueError as e:
            raise RcaInputError("interpolationInvalid", str(e)) from None

    if not isinstance(Geometry, dict):
        raise RcaInputError("gridBfMissing", "Beamforming grids ('Geometry.gridBfX', "
                            "'Geometry.gridBfY', and 'Geometry.gridBfX')are required!")
    if not all(k in Geometry for k in ("gridBfX", "gridBfY", "gridBfZ")):
        raise RcaInputError("gridBfMissing", "Beamforming grids ('Geometry.gridBfX', "
                            "'Geometry.gridBfY', and 'Geometry.gridBfX')are required!")
    g = [Geometry[k] for k in ("gridBfX", "gridBfY", "gridBfZ")]
    if (not all(_is_num(a) for a in g) or len({_numel(a) for a in g}) != 1 or _numel(g[0]) < 1):
        raise RcaInputError("gridBfInvalid", "Beamforming grids ('Geometry.gridBfX', "
                            "'Geometry.gridBfY', and 'Geometry.gridBfX') must must be 3D numeric "
                            "meshgrids and have the same size!")
    if not all(k in Geometry for k in ("posEleX", "posEleY", "posEleZ")):
        raise RcaInputError("posEleMissing", "Element positions ('Geometry.posEleX', "
                            "'Geometry.posEleY', and 'Geometry.posEleZ')are required!")
    e = [Geometry[k] for k in ("posEleX", "posEleY", "posEleZ")]
    if not all(_is_num(a) for a in e) or len({_numel(a) for a in e}) != 1 or _numel(e[0]) < 1:
        raise RcaInputError("posEleInvalid", "Pposition of the element ('Geometry.posEleX', "
                            "'Geometry.posEleY', and 'Geometry.posEleZ') must be numeric arrays "
                            "and have the same size!")
    for ax in "XYZ":
        k = "aperture" + ax
        if k not in Geometry:
            _warn(f"aperture{ax}Default",
                  f"Aperture along the {ax.lower()}-axis ('Param.aperture{ax}') is missing, the "
                  "position of the elements will be used to estimate it. If this is an "
                  "out-of-plane axis it is recommended to set that value manually based on the "
                  "size of the elements!")
        elif not _is_scalar_num(Geometry[k]):
            raise RcaInputError(f"aperture{ax}Invalid", f"Aperture along the {ax.lower()}-axis "
                                f"('Geometry.aperture{ax}') must be a single numeric value!")

    if not isinstance(Transmit, dict):
        raise RcaInputError("TransmitNotStruct", "Transmit ('Transmit') must be a structure array!")
    has_theta = "thetaX" in Transmit or "thetaY" in Transmit
    has_vs = any(k in Transmit for k in ("virtualSourceX", "virtualSourceY", "virtualSourceZ"))
    if has_theta == has_vs:
        raise RcaInputError("ThetaORVirtualSourceRequired",
                            "Origin of the insonification(s) ('thetaX' and 'thetaY' OR "
                            "'virtualSourceX', 'virtualSourceY', and 'virtualSourceZ') is required "
                            "and must be unique (use either angle(s) OR virtual source(s))!")
    if "thetaX" in Transmit and "thetaY" in Transmit:
        _warn("thetaToVirtualSource", "Tilt angles ('Transmit.thetaX' and 'Transmit.thetaY') "
                                      "will be converted to virtual source(s)")
        if _numel(Transmit["thetaX"]) != _numel(Transmit["thetaY"]):
            raise RcaInputError("tiltsNotSameLength",
                                "Tilts ('Param.thetaX' and 'Param.thetaY')  must have the same length!")
    elif all(k in Transmit for k in ("virtualSourceX", "virtualSourceY", "virtualSourceZ")):
        n = {_numel(Transmit[k]) for k in ("virtualSourceX", "virtualSourceY", "virtualSourceZ")}
        if len(n) != 1:
            raise RcaInputError("virtualSourcesNotSameLength",
                                "Virtual sources ('Param.virtualSourceX', 'Param.virtualSourceY', "
                                "and 'Param.virtualSourceZ') must have the same length!")
    else:
        raise RcaInputError("ThetaORVirtualSourceInvalid",
                            "Origin of the insonification(s) ('thetaX' and 'thetaY' OR "
                            "'virtualSourceX', 'virtualSourceY', and 'virtualSourceZ') is incorrect. "
                            "Verify that all axes/positions are well defined and that you are using "
                            "either angle(s) OR virtual source(s))!")
    Transmit = dict(Transmit)
    if "flagReceiveX" not in Transmit:
        # RCA.m:370-373 builds this warning text but never emits it.
        Transmit["flagReceiveX"] = True
    elif not isinstance(Transmit["flagReceiveX"], (bool, np.bool_)):
        raise RcaInputError("flagReceiveXInvalid",
                            "Transmit receive channel flag ('Transmit.flagReceiveX') must be a boolean!")

    if not isinstance(Param, dict):
        raise RcaInputError("ParamNotStruct", "Parameters ('Param') must be a structure array!")
    Param = dict(Param)
    for k, name in (("samplingFrequency", "Sampling frequency"),
                    ("demodulationFrequency", "Demodulation frequency")):
        if k not in Param:
            raise RcaInputError(f"{k}Missing", f"{name} ('Param.{k}') is required!")
        if not _is_scalar_num(Param[k]):
            raise RcaInputError(f"{k}Invalid", f"{name} ('Param.{k}') must be a single numeric value!")
    if "speedOfSound" not in Param:
        Param["speedOfSound"] = 1540
        _warn("speedOfSoundDefault", "Speed of sound ('Param.speedOfSound') is missing, default = 1540 m/s.")
    elif not _is_scalar_num(Param["speedOfSound"]):
        raise RcaInputError("speedOfSoundInvalid",
                            "Speed of sound  ('Param.speedOfSound') must be a single numeric value!")
    if "startTime" not in Param:
        Param["startTime"] = 0
        _warn("startTimeDefault", "Start time ('Param.startTime') is missing, default = 0.")
    elif not _is_scalar_num(Param["startTime"]):
        raise RcaInputError("startTimeInvalid",
                            "Start time ('Param.startTime') must be a single numeric value!")
    if "fnumber" not in Param:
        Param["fnumber"] = 1
        _warn("fnumberDefault", "F-number ('Param.fnumber') is missing, default = 1.")
    elif not _is_scalar_num(Param["fnumber"]):
        # The reference reports an invalid f-number under the startTime id (RCA.m:449-454).
        raise RcaInputError("startTimeInvalid",
                            "F-number  ('Param.startTime') must be a single numeric value!")
    return Transmit, Param, Options


def _scalar(a):
    return float(a.item() if isinstance(a, torch.Tensor) else np.asarray(a).ravel()[0])


def _to_iq(data, n_ch, n_tx, device):
    """RCA.m:468-485 + 576-578: data [nSample, nColX, nColY, nEvent...] ->
    complex64 [nFrame, nTx, nChannel, nSampleIQ] on the device."""
    if isinstance(data, torch.Tensor):
        shp = list(data.shape)
        t = data.to(device)
        flat = t.permute(*reversed(range(t.dim()))).reshape(-1) if t.dim() > 1 else t.reshape(-1)
    else:
        a = np.asarray(data)
        shp = list(a.shape)
        flat = torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to(device)
    shp = shp + [1] * max(0, 4 - len(shp))
    ns, ncx, ncy = int(shp[0]), int(shp[1]), int(shp[2])
    if ncx * ncy != n_ch:
        raise RcaInputError("dataChannelMismatch",
                            f"data has {ncx}x{ncy} channels but Geometry has {n_ch} elements")
    if flat.is_complex():
        iq = flat.to(torch.complex64)
        ns_iq = ns
    else:  # IQ(n) = RF(2n) - j RF(2n+1): negate every second sample (RCA.m:476)
        if ns % 2:
            raise RcaInputError("dataRealOddSamples",
                                "interleaved real data needs an even number of samples")
        r = flat.to(torch.float32).reshape(-1, 2)
        iq = torch.complex(r[:, 0], -r[:, 1])
        ns_iq = ns // 2
    per = ns_iq * n_ch * n_tx
    if iq.numel() % per:
        raise RcaInputError("dataSizeInvalid", f"data holds {iq.numel()} samples, not a multiple "
                                               f"of nSample*nChannel*nTx = {per}")
    return iq.reshape(-1, n_tx, n_ch, ns_iq).contiguous(), ns_iq


def bfVirtualSourceGeneric_RCA(*args, device=None):
    """``[dataBf, mask] = bfVirtualSourceGeneric_RCA(data, Geometry, Transmit, Param[, Options])``.

    Returns ``dataBf`` complex64 [s1, s2, s3, nTx|1, nFrame] and ``mask``
    float32 [s1, s2, s3, nTx|1] where (s1, s2, s3) is the grid shape
    (RCA.m:715-722).  With no argument it builds the HIP library (the
    reference compiles its PTX in that case, RCA.m:112-115).
    """
    return _run(args, device, nosum=False, ablaw=None)


def bfVirtualSourceGeneric_RCA_Nosum(*args, device=None):
    """Per-receive-line variant (MATLAB/lib/beamformer/NoSum/bfVirtualSourceGeneric_RCA_Nosum.m):
    no channel sum and no compounding; ``dataBf`` is complex64
    [s1, s2, s3, nTx, nFrame, nChannel] (Nosum.m:653-718), ``mask`` per tilt."""
    return _run(args, device, nosum=True, ablaw=None)


def bfVirtualSourceGeneric_RCA_ab_1(*args, device=None):
    """Receive aberration law per line, ``Param.AbLawRx`` [nChannel] (s) added to
    the receive delay (Hyp1_save/bfVirtualSourceGeneric_RCA_ab_1.cu:306-311)."""
    return _run(args, device, nosum=False, ablaw="channel")


def bfVirtualSourceGeneric_RCA_ab_2(*args, device=None):
    """Receive aberration law per (tilt, line), ``Param.AbLawRx`` with
    AbLawRx(:)[tx*nChannel + ch] (Hyp2/bfVirtualSourceGeneric_RCA_ab_2.cu:306-311)."""
    return _run(args, device, nosum=False, ablaw="tx_channel")


def bfVirtualSourceGeneric_RCA_Nosum_ab(*args, device=None):
    """NoSum with a per-line receive aberration law
    (NoSum_ab/bfVirtualSourceGeneric_RCA_Nosum_ab.cu:308-313)."""
    return _run(args, device, nosum=True, ablaw="channel")


def bfVirtualSourceGeneric_RCA_TxRx(*args, device=None):
    """Per-(voxel, line) aberration laws on the RCA geometry: ``Param.AbLawRx``
    joins the receive delay and ``Param.AbLawTx`` the total delay of every
    update, both indexed AbLaw(:)[idxPix + idxChannel*nPix], i.e. MATLAB arrays
    [nPix, nChannel] or [s1, s2, s3, nChannel] (the laws of
    exemple_paul/bfVirtualSourceGenericTxRx.cu:268-275, 292-294, .m:612-613).
    An empty law counts as zero."""
    return _run(args, device, nosum=False, ablaw="pixel_channel")


def _run(args, device, nosum, ablaw):
    if len(args) == 0:
        from .. import build
        build.build_library()
        return None
    if not 4 <= len(args) <= 5:
        raise TypeError(f"{_ID} takes 4 or 5 arguments ({len(args)} given)")
    data, Geometry, Transmit, Param = args[:4]
    Options = args[4] if len(args) == 5 else None
    Transmit, Param, Options = _validate(data, Geometry, Transmit, Param, Options)
    if ablaw is not None and "AbLawRx" not in Param:
        raise RcaInputError("AbLawRxMissing", "Receive aberration law ('Param.AbLawRx') is required!")
    if ablaw == "pixel_channel" and "AbLawTx" not in Param:
        raise RcaInputError("AbLawTxMissing", "Transmit aberration law ('Param.AbLawTx') is required!")

    as_torch = isinstance(data, torch.Tensor)
    if device is None:
        device = data.device if as_torch and data.is_cuda else torch.device(
            "cuda", torch.cuda.current_device())
    n_ch = _numel(Geometry["posEleX"])
    n_tx = _numel(Transmit["thetaX"] if "thetaX" in Transmit else Transmit["virtualSourceX"])
    iq, ns_iq = _to_iq(data, n_ch, n_tx, device)

    grid_shape = list(_shape(Geometry["gridBfX"]))
    if len(grid_shape) > 3:
        if int(np.prod(grid_shape[3:])) != 1:
            raise RcaInputError("gridBfInvalid", "beamforming grids must be at most 3-D")
        grid_shape = grid_shape[:3]
    grid_shape = [int(s) for s in grid_shape] + [1] * (3 - len(grid_shape))

    def flat_grid(a):
        if isinstance(a, torch.Tensor):
            return a.permute(*reversed(range(a.dim()))).reshape(-1) if a.dim() > 1 else a.reshape(-1)
        return np.asarray(a, dtype=np.float32).ravel(order="F")

    Geo = dict(Geometry)
    for k in ("gridBfX", "gridBfY", "gridBfZ"):
        Geo[k] = flat_grid(Geometry[k])
    Par = {k: _scalar(Param[k]) for k in ("samplingFrequency", "demodulationFrequency",
                                          "speedOfSound", "startTime", "fnumber")}
    if ablaw is not None:
        for key in ("AbLawRx", "AbLawTx") if ablaw == "pixel_channel" else ("AbLawRx",):
            law = Param[key]
            law = (law.detach().cpu().numpy() if isinstance(law, torch.Tensor) else np.asarray(law))
            if law.size:                       # single(Param.AbLaw..), column-major
                Par[key] = law.astype(np.float32).ravel(order="F")
        if "AbLawRx" not in Par and "AbLawTx" not in Par:
            ablaw = None                       # both laws empty: no delay at all
        Par["AbLawMode"] = ablaw
    plan = RcaPlan.from_reference(Geo, dict(Transmit), Par, Options, n_sample=ns_iq, device=device,
                                  depth_run=grid_shape[0])

    compound = bool(Options["flagCompounding"]) and not nosum
    out, mask = plan.forward(iq, compound=compound, nosum=nosum)
    s1, s2, s3 = grid_shape
    if nosum:
        nc, nf, nt = out.shape[0], out.shape[1], out.shape[2]
        if as_torch:
            dataBf = out.reshape(nc, nf, nt, s3, s2, s1).permute(5, 4, 3, 2, 1, 0)
        else:
            dataBf = out.cpu().numpy().transpose(3, 2, 1, 0).reshape(s1, s2, s3, nt, nf, nc,
                                                                     order="F")
    else:
        nf, nt = out.shape[0], out.shape[1]
        if as_torch:
            dataBf = out.reshape(nf, nt, s3, s2, s1).permute(4, 3, 2, 1, 0)
        else:
            dataBf = out.cpu().numpy().transpose(2, 1, 0).reshape(s1, s2, s3, nt, nf, order="F")
    nt = mask.shape[0]
    if as_torch:
        maskr = mask.reshape(nt, s3, s2, s1).permute(3, 2, 1, 0)
        if not data.is_cuda:
            dataBf, maskr = dataBf.cpu(), maskr.cpu()
        return dataBf, maskr
    maskr = mask.cpu().numpy().transpose(1, 0).reshape(s1, s2, s3, nt, order="F")
    return dataBf, maskr
