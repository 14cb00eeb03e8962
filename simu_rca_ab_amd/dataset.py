"""Realigned per-channel dataset of the simulation pipeline (the NoSum caller).

``MATLAB/save_simus_sh.m`` turns each simulated buffer into the network's
training data: the RF of both RCA halves is scaled to a 16-bit range
(:323-345), and for every frame the per-receive-line (NoSum) beamformer is
run on a 17-voxel depth column centred on that frame's tracked bubble
(:229-249, :351-394), with and without the aberration law; the two halves'
lines are concatenated into ``dataBf_al`` [Nz=17, nTx, 2 x nCh, nFrame]
(:401-403).  Here the frames of a half are beamformed in one GPU launch: the
union of the frames' columns is the voxel list and each frame keeps its own
column (per-voxel results do not depend on the other voxels, so this equals
one call per frame).
"""
from __future__ import annotations

import numpy as np

from .beamformer import bfVirtualSourceGeneric_RCA_Nosum, bfVirtualSourceGeneric_RCA_Nosum_ab


def normalize_rf(rf):
    """``RF ./ max(abs(RF(:))) .* 32768`` (save_simus_sh.m:323-345)."""
    m = np.max(np.abs(rf))
    return rf / m * 32768 if m > 0 else rf


def realignment_grids(pos_bubble, wavelength, sub=4, n_lambda=4):
    """Per-frame depth columns (save_simus_sh.m:229-249): Nz = sub*n_lambda + 1
    depths of wavelength/sub from bubble_z - n_lambda*wavelength/2 (1-based
    ``(1:Nz).*dz + Zmin``), one (x, y) at bubble + one pixel.  pos_bubble is
    [3, nFrame]; returns X, Y, Z of shape [Nz, nFrame]."""
    pos = np.asarray(pos_bubble, dtype=np.float64)
    d = wavelength / sub
    nz = sub * n_lambda + 1
    k = np.arange(1, nz + 1)[:, None]
    X = np.broadcast_to(pos[0] + d, (nz, pos.shape[1]))
    Y = np.broadcast_to(pos[1] + d, (nz, pos.shape[1]))
    Z = k * d + (pos[2] - n_lambda * wavelength / 2)[None, :]
    return np.ascontiguousarray(X), np.ascontiguousarray(Y), Z


def realigned_half(rf, Geometry, Transmit, Param, Options, grids, ab_law=None, device=None):
    """One RCA half: rf is the real RF [2*ns, nCh, nTx, nFrame] of the frames
    (``reshape(RF_X(:,:,:,iFrame), size(RF_X,1), nCh, 1, [])`` per frame);
    returns complex64 [Nz, nTx, nCh, nFrame] - every frame beamformed on its
    own column by the NoSum kernel (with the per-line receive law when
    ``ab_law`` is given, ``bfVirtualSourceGeneric_RCA_Nosum_ab``)."""
    X, Y, Z = grids
    nz, nfr = Z.shape
    rf = np.asarray(rf)
    ns2, nch, ntx = rf.shape[0], rf.shape[1], rf.shape[2]
    if rf.shape[3] != nfr:
        raise ValueError("one realignment column per frame")
    G = dict(Geometry, gridBfX=X[:, :, None], gridBfY=Y[:, :, None], gridBfZ=Z[:, :, None])
    data = rf.reshape(ns2, nch, 1, ntx * nfr, order="F")
    Op = dict(Options or {}, flagCompounding=False)
    if ab_law is None:
        out, _ = bfVirtualSourceGeneric_RCA_Nosum(data, G, Transmit, Param, Op, device=device)
    else:
        out, _ = bfVirtualSourceGeneric_RCA_Nosum_ab(data, G, Transmit, dict(Param, AbLawRx=ab_law), Op,
                                                      device=device)
    # out [Nz, nFr(columns), 1, nTx, nFrame, nCh]: frame f keeps column f
    f = np.arange(nfr)
    sel = out[:, f, 0, :, f, :]                      # [nFrame, Nz, nTx, nCh]
    return np.ascontiguousarray(np.moveaxis(sel, 0, 3)).astype(np.complex64)


def realigned_dataset(RF_X, RF_Y, GeometryX, GeometryY, TransmitX, TransmitY, Param, Options,
                      pos_bubble, wavelength, AbLawX=None, AbLawY=None, device=None):
    """``dataBf_al`` [Nz, nTx, 2*nCh, nFrame] = cat(3, half X, half Y) of the
    realigned NoSum outputs (save_simus_sh.m:351-362); with the laws also the
    corrected ``dataBf_al_ab_c`` (:364-376, the per-line law in the receive
    delay).  RF_X/RF_Y: real [2*ns, nCh, nTx, nFrame] (already normalised)."""
    grids = realignment_grids(pos_bubble, wavelength)
    al = np.concatenate([realigned_half(RF_X, GeometryX, TransmitX, Param, Options, grids, device=device),
                         realigned_half(RF_Y, GeometryY, TransmitY, Param, Options, grids, device=device)],
                        axis=2)
    if AbLawX is None:
        return al
    al_c = np.concatenate([
        realigned_half(RF_X, GeometryX, TransmitX, Param, Options, grids, AbLawX, device),
        realigned_half(RF_Y, GeometryY, TransmitY, Param, Options, grids, AbLawY, device)], axis=2)
    return al, al_c


__all__ = ["normalize_rf", "realignment_grids", "realigned_half", "realigned_dataset"]
