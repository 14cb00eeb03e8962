"""SVD clutter filter of beamformed ensembles on the GPU (``rca_svd_filter``).

Drop-in for the reference's ``Python/clutterfilter/SVDClutterFilter.py``:

* ``SVDClutterFilter(IQ, mask, wcut)`` keeps its signature and behaviour
  (``SVDClutterFilter.py:10-23``): ``IQ`` is ``[..., Nt]``, the pixels are the
  F-ordered leading axes (``af.moddims``), the kept eigenvectors are
  ``int(wcut[0]*Nt) .. int(wcut[1]*Nt) - 1`` in descending eigenvalue order,
  ``mask`` is accepted and not used, and the result is ``(IQf, s)`` with
  ``s`` the eigenvalues of the Gram matrix (its singular values), descending.
* ``SVDClutterFilterMatlab(IQ, Mask, w, type='band')`` follows
  ``MATLAB/lib/clutterFilter/SVDClutterFilter.m:1-45``: ``w`` in [0, 1],
  ``Ncut = round(w*Nt)``, data divided by ``sqrt(Mask)``, 1-based inclusive
  bands 'low' / 'high' / 'stop' / 'band'.

numpy in gives numpy out; a torch tensor in gives a torch tensor on its
device.  ``svd_filter(iq, keep, mask)`` is the device-level call for a
frame-major ``[nT, nPix]`` complex64 tensor such as the beamformer's output
``out[:, 0]``.  The arithmetic (FP64 Gram, eigen-decomposition and
projection) runs in the HIP library; torch only holds memory.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib


def svd_filter(iq: torch.Tensor, keep, mask: torch.Tensor | None = None):
    """Filter a frame-major complex64 ensemble ``iq [nT, nPix]`` on its GPU.

    ``keep[k]`` keeps the eigenvector of the k-th largest eigenvalue of
    Y^H Y, Y = iq^T / sqrt(mask).  Returns ``(out [nT, nPix] complex64,
    eigenvalues [nT] float64, descending)``."""
    if iq.dtype != torch.complex64 or iq.dim() != 2 or not iq.is_cuda or not iq.is_contiguous():
        raise ValueError("iq must be a contiguous complex64 [nT, nPix] GPU tensor")
    nt, npix = iq.shape
    keep = np.ascontiguousarray(np.asarray(keep, dtype=bool).astype(np.uint8))
    if keep.shape != (nt,):
        raise ValueError(f"keep must have nT={nt} entries")
    if mask is not None:
        mask = mask.reshape(-1)
        if mask.dtype != torch.float32 or mask.numel() != npix or mask.device != iq.device:
            raise ValueError("mask must be float32 with nPix entries on the data's device")
        mask = mask.contiguous()
    lib = _lib.load()
    nbytes = lib.rca_svd_workspace_bytes(npix, nt)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=iq.device)
    out = torch.empty_like(iq)
    ev = torch.empty(nt, dtype=torch.float64, device=iq.device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(iq.device).cuda_stream)
    with torch.cuda.device(iq.device):
        rc = lib.rca_svd_filter(ctypes.c_void_p(iq.data_ptr()),
                                ctypes.c_void_p(mask.data_ptr() if mask is not None else 0),
                                npix, nt, keep.ctypes.data_as(ctypes.c_void_p),
                                ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ev.data_ptr()),
                                ctypes.c_void_p(ws.data_ptr()), nbytes, stream)
    _lib.check(rc)
    return out, ev


def _to_frames(IQ):
    """[..., Nt] (numpy or torch) -> frame-major complex64 [Nt, Npix] on the
    GPU, pixels in F order; plus a function mapping the result back."""
    if isinstance(IQ, torch.Tensor):
        dev = IQ.device if IQ.is_cuda else torch.device("cuda", torch.cuda.current_device())
        x = IQ.to(dev)
        shape = tuple(x.shape)
        k = x.dim() - 1
        # frame t's pixels in F order = C order of the reversed pixel axes
        frames = x.permute(k, *reversed(range(k))).reshape(shape[-1], -1)
        frames = frames.to(torch.complex64).contiguous()

        def back(out):
            y = out.reshape(shape[-1], *reversed(shape[:-1]))
            return y.permute(*reversed(range(1, k + 1)), 0)

        return frames, back
    IQ = np.asarray(IQ)
    shape = IQ.shape
    frames = np.ascontiguousarray(IQ.reshape(-1, shape[-1], order="F").T.astype(np.complex64))
    dev = torch.device("cuda", torch.cuda.current_device())

    def back(out):
        return out.cpu().numpy().T.reshape(shape, order="F")

    return torch.from_numpy(frames).to(dev), back


def _as_out(ev, like_numpy: bool):
    return ev.cpu().numpy() if like_numpy else ev


def keep_python(wcut, nt: int) -> np.ndarray:
    """Eigenvectors SVDClutterFilter.py keeps: ``u[:, int(ncut[0]):int(ncut[1])]``,
    ``ncut = wcut * Nt`` (.py:12, :17), as a mask over descending eigenvalues."""
    ncut = np.asarray(wcut, dtype=np.float64) * nt
    keep = np.zeros(nt, bool)
    keep[int(ncut[0]):int(ncut[1])] = True
    return keep


def _matlab_round(x):
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def keep_matlab(w, nt: int, kind: str = "band") -> np.ndarray:
    """Eigenvectors SVDClutterFilter.m keeps (:13, :26-44): ``Ncut =
    round(w*Nt)`` and the 1-based inclusive bands 'low' 1:Ncut(1), 'high'
    Ncut(1):end, 'stop' [1:Ncut(1), Ncut(2):end], 'band' Ncut(1):Ncut(2).
    Raises ValueError for w outside [0, 1] (:6-8) and IndexError where MATLAB
    would index column 0."""
    w = np.atleast_1d(np.asarray(w, dtype=np.float64))
    if w.min() < 0 or w.max() > 1:
        raise ValueError("w must be is the normalized cutoff frequency [0 1]")
    n = _matlab_round(w * nt).astype(int)

    def start(i):
        if n[i] < 1:
            raise IndexError("Array indices must be positive integers")
        return n[i] - 1

    keep = np.zeros(nt, bool)
    if kind == "low":
        keep[:n[0]] = True
    elif kind == "high":
        keep[start(0):] = True
    elif kind == "stop":
        keep[:n[0]] = True
        keep[start(1):] = True
    else:
        keep[start(0):n[1]] = True
    return keep


def SVDClutterFilter(IQ, mask, wcut):
    """Python/clutterfilter/SVDClutterFilter.py:10-23 (``mask`` is not used,
    as in the reference).  Returns ``(IQf, s)``."""
    nt = (IQ.shape if hasattr(IQ, "shape") else np.shape(IQ))[-1]
    keep = keep_python(wcut, nt)
    frames, back = _to_frames(IQ)
    out, ev = svd_filter(frames, keep)
    return back(out), _as_out(ev, not isinstance(IQ, torch.Tensor))


def SVDClutterFilterMatlab(IQ, Mask, w, type="band"):   # noqa: A002 (the reference's name)
    """MATLAB/lib/clutterFilter/SVDClutterFilter.m:1-45.  Returns
    ``(IQf, eig_val)``; IQf stays divided by sqrt(Mask), as in the reference."""
    shape = IQ.shape if hasattr(IQ, "shape") else np.shape(IQ)
    keep = keep_matlab(w, shape[-1], type)
    frames, back = _to_frames(IQ)
    mask = None
    if Mask is not None and (Mask.numel() if isinstance(Mask, torch.Tensor) else np.size(Mask)):
        if isinstance(Mask, torch.Tensor):
            k = Mask.dim()
            mask = Mask.permute(*reversed(range(k))).reshape(-1)
        else:
            mask = torch.from_numpy(np.asarray(Mask, dtype=np.float32).ravel(order="F"))
        mask = mask.to(device=frames.device, dtype=torch.float32)
    out, ev = svd_filter(frames, keep, mask)
    return back(out), _as_out(ev, not isinstance(IQ, torch.Tensor))


__all__ = ["svd_filter", "keep_python", "keep_matlab", "SVDClutterFilter",
           "SVDClutterFilterMatlab"]
