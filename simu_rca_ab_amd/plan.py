"""Device-resident beamforming plan of one RCA half, over the C ABI.

``RcaPlan`` keeps the voxel grid, centred element positions, virtual sources
and minimal distances on the GPU (the reference re-uploads them on every
``feval``, bfVirtualSourceGeneric_RCA.m:670-686) and launches the HIP kernels
on torch's current stream.  torch provides device memory and streams only; all
beamforming arithmetic is in ``csrc/rca_das.hip``.

Tensor layouts (C-order):
  iq     complex64 [nFrame, nTx, nChannel, nSample], or the FP16-I/Q form
         (float16 [nFrame, nTx, nChannel, nSample, 2], float32 scale
         [nFrame, nTx, nChannel]) made by :func:`quantize_fp16`
  out    complex64 [nFrame, 1 | nTx, nPix]
  mask   float32   [1 | nTx, nPix]
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib


def _np64(a):
    """Column-major flatten to float64 numpy (host prep runs on the CPU)."""
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu()
        a = a.permute(*reversed(range(a.dim()))).reshape(-1) if a.dim() > 1 else a.reshape(-1)
        return a.to(torch.float64).numpy()
    return np.asarray(a, dtype=np.float64).ravel(order="F")


def prepare_geometry(Geometry: dict, Transmit: dict) -> dict:
    """Host prep of bfVirtualSourceGeneric_RCA.m:491-558 in float64, cast to
    float32 as the wrapper does before the launch (RCA.m:601-617).

    * element positions are centred on their mean (RCA.m:500-502);
    * missing apertures default to the span of the centred positions
      (RCA.m:505-521);
    * tilts become virtual sources at r = 10 m (RCA.m:527-541);
    * minDist_t = min_c |(ey_c - vx_t, ex_c - vy_t, ez_c - vz_t)| — the X/Y
      swap is the reference's (RCA.m:551-558).
    """
    pos = [_np64(Geometry[k]) for k in ("posEleX", "posEleY", "posEleZ")]
    cen = [p - p.mean() for p in pos]
    ap = []
    for k, c in zip(("apertureX", "apertureY", "apertureZ"), cen):
        ap.append(float(Geometry[k]) if k in Geometry else float(c.max() - c.min()))
    if "thetaX" in Transmit:
        tx = np.tan(_np64(Transmit["thetaX"]))
        ty = np.tan(_np64(Transmit["thetaY"]))
        vz = -10.0 / np.sqrt(1.0 + tx ** 2 + ty ** 2)
        vx, vy = vz * tx, vz * ty
    else:
        vx, vy, vz = (_np64(Transmit[k])
                      for k in ("virtualSourceX", "virtualSourceY", "virtualSourceZ"))
    ex, ey, ez = cen
    d0 = np.sqrt((ey[None, :] - vx[:, None]) ** 2 + (ex[None, :] - vy[:, None]) ** 2 +
                 (ez[None, :] - vz[:, None]) ** 2).min(axis=1)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return {"ele": [f32(c) for c in cen], "vs": [f32(vx), f32(vy), f32(vz)], "d0": f32(d0),
            "aperture": [float(np.float32(a)) for a in ap]}


def _depth_run(g) -> int:
    """Length of a grid array's first axis (its fastest in column-major order),
    0 for a flat voxel list."""
    shape = tuple(g.shape) if hasattr(g, "shape") else np.shape(g)
    return int(shape[0]) if len(shape) >= 2 else 0


def _law_mode(rx, tx, n_ch, n_tx, n_pix):
    """Aberration-law layout from its size: [nChannel] per line, [nTx][nChannel]
    per (tilt, line), [nChannel][nPix] per (line, voxel).  A transmit-side law
    only exists per (line, voxel)."""
    if tx is not None:
        return "pixel_channel"
    n = rx.numel()
    fits = [m for m, k in (("channel", n_ch), ("tx_channel", n_tx * n_ch),
                           ("pixel_channel", n_ch * n_pix)) if n == k]
    if not fits:
        raise ValueError(f"aberration law of {n} delays matches no layout (nChannel={n_ch}, "
                         f"nTx={n_tx}, nPix={n_pix})")
    if len(fits) > 1:
        raise ValueError(f"aberration law of {n} delays is ambiguous ({', '.join(fits)}): "
                         "give its mode")
    return fits[0]


def _dev_f32(a, device):
    if isinstance(a, torch.Tensor):
        t = a.detach().to(device=device, dtype=torch.float32)
        return t.reshape(-1) if t.is_contiguous() else t.contiguous().reshape(-1)
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel(order="F"))).to(device)


def quantize_fp16(iq: torch.Tensor):
    """FP16-I/Q storage on the GPU (``rca_quantize_fp16``): half2 samples
    scaled by a per-trace FP32 factor, scale[f, t, c] = max over the trace of
    max(|re|, |im|) (1 for an all-zero trace), samples (re, im) / scale
    rounded to float16.  iq: complex64 [F, T, C, Ns] on a GPU.  Returns
    (float16 [F, T, C, Ns, 2], float32 [F, T, C])."""
    if iq.dtype != torch.complex64 or not iq.is_cuda or iq.dim() < 1:
        raise ValueError("quantize_fp16 takes complex64 I/Q on a GPU")
    iq = iq.contiguous()
    ns = iq.shape[-1]
    h = torch.empty(tuple(iq.shape) + (2,), dtype=torch.float16, device=iq.device)
    scale = torch.empty(tuple(iq.shape[:-1]), dtype=torch.float32, device=iq.device)
    lib = _lib.load()
    _lib.check(lib.rca_quantize_fp16(ctypes.c_void_p(iq.data_ptr()), iq.numel() // max(ns, 1), ns,
                                     ctypes.c_void_p(h.data_ptr()), ctypes.c_void_p(scale.data_ptr()),
                                     ctypes.c_void_p(torch.cuda.current_stream(iq.device).cuda_stream)))
    return h, scale


class RcaPlan:
    """Geometry of one RCA half resident on a GPU.

    Arguments follow the reference kernel's (bfVirtualSourceGeneric_RCA.cu:103-133):
    grids are flat voxel lists (column-major flatten of the meshgrid), element
    positions are already centred, virtual sources and minimal distances are
    per tilt.  ``interp`` is 'quadratic', 'lanczos' (the reference's 6-tap
    kernel), 'lanczos3' or 'lanczos4'.
    """

    def __init__(self, *, grid, ele, vs, min_dist, receive_x, aperture, fs, fd, c, t0, fnumber,
                 n_sample, interp="quadratic", mask_transmit=False, ab_law=None,
                 ab_law_mode=None, depth_run=0, device=None, ab_law_tx=None):
        self.lib = _lib.load()
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        if self.device.type != "cuda":
            raise ValueError("RcaPlan needs a GPU device")
        dev = self.device
        self.grid = [_dev_f32(g, dev) for g in grid]
        self.ele = [_dev_f32(e, dev) for e in ele]
        self.vs = [_dev_f32(v, dev) for v in vs]
        self.min_dist = _dev_f32(min_dist, dev)
        n_pix = self.grid[0].numel()
        n_ch = self.ele[0].numel()
        n_tx = self.vs[0].numel()
        if any(g.numel() != n_pix for g in self.grid):
            raise ValueError("gridBfX/Y/Z must have the same number of voxels")
        if any(e.numel() != n_ch for e in self.ele):
            raise ValueError("posEleX/Y/Z must have the same length")
        if any(v.numel() != n_tx for v in self.vs) or self.min_dist.numel() != n_tx:
            raise ValueError("virtual sources and minDist must have one entry per tilt")
        if interp not in _lib.INTERP:
            raise ValueError(f"unknown interpolation {interp!r}")
        self.ab_law = self.ab_law_tx = None
        mode = 0
        if ab_law is not None or ab_law_tx is not None:
            self.ab_law = _dev_f32(ab_law, dev) if ab_law is not None else None
            self.ab_law_tx = _dev_f32(ab_law_tx, dev) if ab_law_tx is not None else None
            mode = _lib.ABLAW[ab_law_mode if ab_law_mode is not None else
                              _law_mode(self.ab_law, self.ab_law_tx, n_ch, n_tx, n_pix)]
            need = {1: n_ch, 2: n_tx * n_ch, 3: n_ch * n_pix}[mode]
            for name, law in (("AbLawRx", self.ab_law), ("AbLawTx", self.ab_law_tx)):
                if law is not None and law.numel() != need:
                    raise ValueError(f"{name} needs {need} delays, got {law.numel()}")
            if mode != 3 and self.ab_law_tx is not None:
                raise ValueError("a transmit-side law (AbLawTx) needs the per-(pixel, channel) mode")
            if mode != 3 and self.ab_law is None:
                raise ValueError("this aberration-law mode needs AbLawRx")
        self.n_pix, self.n_channel, self.n_tx, self.n_sample = n_pix, n_ch, n_tx, int(n_sample)
        self.interp = interp
        self.receive_x = bool(receive_x)
        d = _lib.RcaDasDesc()
        d.grid_x, d.grid_y, d.grid_z = (g.data_ptr() for g in self.grid)
        d.ele_x, d.ele_y, d.ele_z = (e.data_ptr() for e in self.ele)
        d.vs_x, d.vs_y, d.vs_z = (v.data_ptr() for v in self.vs)
        d.min_dist = self.min_dist.data_ptr()
        d.ab_law_rx = self.ab_law.data_ptr() if self.ab_law is not None else None
        d.ab_law_tx = self.ab_law_tx.data_ptr() if self.ab_law_tx is not None else None
        d.ab_law_mode = mode
        d.n_pix, d.n_channel, d.n_tx, d.n_sample = n_pix, n_ch, n_tx, self.n_sample
        d.receive_x = int(self.receive_x)
        d.mask_transmit = int(bool(mask_transmit))
        d.interp = _lib.INTERP[interp]
        d.aperture_x, d.aperture_y, d.aperture_z = (float(a) for a in aperture)
        d.sampling_frequency = float(fs)
        d.demodulation_frequency = float(fd)
        d.speed_of_sound = float(c)
        d.start_time = float(t0)
        d.fnumber = float(fnumber)
        # grid stored depth-fastest with this many depths: workgroups take depth
        # tiles (a speed hint only; the results do not depend on it)
        self.depth_run = int(depth_run) if depth_run and n_pix % int(depth_run) == 0 else 0
        d.depth_run = self.depth_run
        self.desc = d

    # ------------------------------------------------------------------
    @classmethod
    def from_reference(cls, Geometry, Transmit, Param, Options=None, n_sample=None, device=None,
                       depth_run=None):
        """Plan from the reference's Geometry/Transmit/Param/Options dicts.
        ``depth_run`` is the length of the grid's first (fastest) axis, Nz for
        the reference's [Nz, Nx, Ny] meshgrids; by default it is read from the
        shape of ``Geometry['gridBfZ']`` when that is at least 2-D."""
        Options = dict(Options or {})
        prep = prepare_geometry(Geometry, Transmit)
        grid = [np.asarray(Geometry[k], dtype=np.float32) if not isinstance(Geometry[k], torch.Tensor)
                else Geometry[k] for k in ("gridBfX", "gridBfY", "gridBfZ")]
        return cls(grid=grid, ele=prep["ele"], vs=prep["vs"], min_dist=prep["d0"],
                   receive_x=bool(Transmit.get("flagReceiveX", True)), aperture=prep["aperture"],
                   fs=Param["samplingFrequency"], fd=Param["demodulationFrequency"],
                   c=Param.get("speedOfSound", 1540.0), t0=Param.get("startTime", 0.0),
                   fnumber=Param.get("fnumber", 1.0), n_sample=n_sample,
                   interp=interpolation_name(Options),
                   mask_transmit=bool(Options.get("flagMaskTransmit", True)),
                   ab_law=Param.get("AbLawRx"), ab_law_mode=Param.get("AbLawMode"),
                   ab_law_tx=Param.get("AbLawTx"),
                   depth_run=_depth_run(Geometry["gridBfZ"]) if depth_run is None else depth_run,
                   device=device)

    # ------------------------------------------------------------------
    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _check_iq(self, iq, scale):
        if iq.device != self.device:
            raise ValueError(f"iq is on {iq.device}, plan on {self.device}")
        if not iq.is_contiguous():
            raise ValueError("iq must be contiguous")
        if iq.dtype == torch.complex64:
            if iq.dim() != 4:
                raise ValueError("complex iq must be [nFrame, nTx, nChannel, nSample]")
            dtype = _lib.IQ_C64
        elif iq.dtype == torch.float16:
            if iq.dim() != 5 or iq.shape[-1] != 2:
                raise ValueError("FP16 iq must be [nFrame, nTx, nChannel, nSample, 2]")
            if scale is None or scale.dtype != torch.float32 or not scale.is_contiguous() \
                    or tuple(scale.shape) != tuple(iq.shape[:3]) or scale.device != self.device:
                raise ValueError("FP16 iq needs a contiguous float32 scale [nFrame, nTx, nChannel]")
            dtype = _lib.IQ_F16X2
        else:
            raise ValueError(f"unsupported iq dtype {iq.dtype}")
        f, t, c, n = iq.shape[:4]
        if (t, c, n) != (self.n_tx, self.n_channel, self.n_sample):
            raise ValueError(f"iq shape {tuple(iq.shape)} does not match the plan "
                             f"(nTx={self.n_tx}, nChannel={self.n_channel}, nSample={self.n_sample})")
        return dtype, f

    def forward(self, iq, scale=None, *, compound=True, out=None, mask=None, want_mask=True,
                accumulate=False, nosum=False):
        """Beamform ``iq``; returns (out, mask).

        out is complex64 [F, 1|T, P] (compounded or per tilt), or
        [nChannel, F, T, P] with ``nosum`` (one output per receive line, the
        reference's NoSum variant); mask is float32 [1|T, P] or None."""
        dtype, nf = self._check_iq(iq, scale)
        if nosum and compound:
            raise ValueError("nosum outputs are per tilt and per line: use compound=False")
        nt = 1 if compound else self.n_tx
        shape = (self.n_channel, nf, nt, self.n_pix) if nosum else (nf, nt, self.n_pix)
        if out is None:
            if accumulate:
                raise ValueError("accumulate needs an output tensor")
            out = torch.empty(shape, dtype=torch.complex64, device=self.device)
        elif (out.dtype != torch.complex64 or not out.is_contiguous() or out.device != self.device
              or tuple(out.shape) != shape):
            raise ValueError(f"out must be contiguous complex64 {shape} on {self.device}")
        if want_mask and mask is None:
            if accumulate:
                raise ValueError("accumulate with a mask needs the mask tensor")
            mask = torch.empty((nt, self.n_pix), dtype=torch.float32, device=self.device)
        if mask is not None and (mask.dtype != torch.float32 or not mask.is_contiguous()
                                 or tuple(mask.shape) != (nt, self.n_pix) or mask.device != self.device):
            raise ValueError(f"mask must be contiguous float32 {(nt, self.n_pix)}")
        flags = ((_lib.FLAG_COMPOUND if compound else 0) | (_lib.FLAG_ACCUMULATE if accumulate else 0)
                 | (_lib.FLAG_NOSUM if nosum else 0))
        rc = self.lib.rca_forward(ctypes.byref(self.desc), ctypes.c_void_p(iq.data_ptr()), dtype,
                                  ctypes.c_void_p(scale.data_ptr() if scale is not None else 0),
                                  nf, ctypes.c_void_p(out.data_ptr()),
                                  ctypes.c_void_p(mask.data_ptr() if mask is not None else 0),
                                  flags, self._stream())
        _lib.check(rc)
        return out, mask

    def adjoint(self, grad_out, *, grad_iq=None, accumulate=False, deterministic=False):
        """grad_iq = A^H grad_out for grad_out complex64 [F, 1|T, P].

        deterministic: sum the tiles' contributions as 64-bit fixed-point
        integers in a workspace (``rca_adjoint_det``), so repeated calls give
        bitwise identical gradients; otherwise float atomics (faster, with
        run-to-run differences at the FP32 rounding level)."""
        if (grad_out.dtype != torch.complex64 or grad_out.dim() != 3 or not grad_out.is_contiguous()
                or grad_out.shape[2] != self.n_pix or grad_out.shape[1] not in (1, self.n_tx)
                or grad_out.device != self.device):
            raise ValueError("grad_out must be contiguous complex64 [nFrame, 1|nTx, nPix] on the plan device")
        nf, gt = grad_out.shape[0], grad_out.shape[1]
        shape = (nf, self.n_tx, self.n_channel, self.n_sample)
        if grad_iq is None:
            if accumulate:
                raise ValueError("accumulate needs grad_iq")
            grad_iq = torch.empty(shape, dtype=torch.complex64, device=self.device)
        elif (grad_iq.dtype != torch.complex64 or tuple(grad_iq.shape) != shape
              or not grad_iq.is_contiguous() or grad_iq.device != self.device):
            raise ValueError(f"grad_iq must be contiguous complex64 {shape}")
        flags = _lib.FLAG_ACCUMULATE if accumulate else 0
        if deterministic:
            nb = int(self.lib.rca_adjoint_det_workspace_bytes(ctypes.byref(self.desc), nf))
            ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
            rc = self.lib.rca_adjoint_det(ctypes.byref(self.desc), ctypes.c_void_p(grad_out.data_ptr()),
                                          gt, nf, ctypes.c_void_p(grad_iq.data_ptr()), flags,
                                          ctypes.c_void_p(ws.data_ptr()), nb, self._stream())
        else:
            rc = self.lib.rca_adjoint(ctypes.byref(self.desc), ctypes.c_void_p(grad_out.data_ptr()), gt,
                                      nf, ctypes.c_void_p(grad_iq.data_ptr()), flags, self._stream())
        _lib.check(rc)
        return grad_iq

    def updates(self, n_frame: int = 1) -> int:
        """Dense voxel*tilt*element updates of ``n_frame`` frames."""
        return self.n_pix * self.n_tx * self.n_channel * n_frame


def interpolation_name(Options: dict) -> str:
    """Options -> kernel interpolation.  ``flagInterpolationLanczos`` keeps the
    reference meaning (RCA.m:159-169); ``interpolation`` selects
    'quadratic' | 'lanczos' | 'lanczos3' | 'lanczos4' explicitly."""
    kind = Options.get("interpolation")
    if kind is None:
        return "lanczos" if Options.get("flagInterpolationLanczos", False) else "quadratic"
    kind = str(kind).lower()
    if kind not in _lib.INTERP:
        raise ValueError(f"unknown interpolation {kind!r}")
    return kind
