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


# ---------------------------------------------------------------------------
# On-disk form of one simulation (save_simus_sh.m:396-407) and the per-job
# arrays the HDF5 conversion builds from it (Python/Convert_job.py:27-84).
# ---------------------------------------------------------------------------
MAT_NAMES = ("dataBf_al", "dataBf_al_ab", "dataBf_al_ab_c", "DelaysAb", "Pos", "Pos_bubble")
POS_MAX_SCATTERERS = 500   # Convert_job.py:42-44: "Pos" rows are padded to 500


def delays_ab(delay_x, delay_y, n_az):
    """``DelaysAb = cat(3, repmat(AbLawX.Delay, 128, 1, naAz),
    repmat(AbLawY.Delay.', 1, 128, naAz))`` (save_simus_sh.m:307-310): the
    row-transmit law varies along the second axis, the column-transmit law
    along the first; [n, n, 2 naAz] float64."""
    dx = np.asarray(delay_x, dtype=np.float64).ravel()
    dy = np.asarray(delay_y, dtype=np.float64).ravel()
    n = dx.size
    if dy.size != n:
        raise ValueError("both laws need one delay per element")
    X = np.broadcast_to(dx[None, :, None], (n, n, n_az))
    Y = np.broadcast_to(dy[:, None, None], (n, n, n_az))
    return np.concatenate([X, Y], axis=2)


def realigned_set(RF_X, RF_Y, RF_X_ab, RF_Y_ab, GeometryX, GeometryY, TransmitX, TransmitY, Param,
                  Options, pos_bubble, wavelength, AbLawX, AbLawY, device=None):
    """The three beamformed arrays of one simulation (save_simus_sh.m:349-394):
    ``dataBf_al`` from the clean RF, ``dataBf_al_ab_c`` from the aberrated RF
    with the receive law applied, ``dataBf_al_ab`` from the aberrated RF
    without it; each complex64 [Nz, nTx, 2 nCh, nFrame]."""
    al = realigned_dataset(RF_X, RF_Y, GeometryX, GeometryY, TransmitX, TransmitY, Param, Options,
                           pos_bubble, wavelength, device=device)
    al_ab, al_ab_c = realigned_dataset(RF_X_ab, RF_Y_ab, GeometryX, GeometryY, TransmitX, TransmitY,
                                       Param, Options, pos_bubble, wavelength, AbLawX, AbLawY,
                                       device=device)
    return {"dataBf_al": al, "dataBf_al_ab": al_ab, "dataBf_al_ab_c": al_ab_c}


def save_simulation(path, arrays: dict):
    """Write the per-simulation set of save_simus_sh.m:396-407 into folder
    ``path``: one MAT-file per variable, named after it, uncompressed
    (``save(..., '-nocompression')``), readable by ``scipy.io.loadmat`` as
    Convert_job.py reads it.  ``arrays`` holds the six names of MAT_NAMES
    (the beamformed arrays complex64, the others float64)."""
    import os
    from scipy.io import savemat
    missing = [n for n in MAT_NAMES if n not in arrays]
    if missing:
        raise ValueError(f"missing variables {missing}")
    os.makedirs(path, exist_ok=True)
    for name in MAT_NAMES:
        a = np.asarray(arrays[name])
        if name.startswith("data"):
            a = a.astype(np.complex64, copy=False)
        else:
            a = a.astype(np.float64, copy=False)
        savemat(os.path.join(path, name + ".mat"), {name: a}, do_compression=False,
                oned_as="column")


def load_simulation(path) -> dict:
    """Read a simulation folder as Convert_job.py:13-14, 31 does: variable
    names are the file names without '.mat', each loaded by name."""
    import os
    from scipy.io import loadmat
    names = sorted(f[:-4] for f in os.listdir(path) if f.endswith(".mat"))
    return {n: loadmat(os.path.join(path, n + ".mat"))[n] for n in names}


def job_arrays(sim_folders) -> dict:
    """The datasets Convert_job.py:27-84 appends to a job's HDF5 file, as
    numpy arrays: for every ``data*`` variable float64 ``<name>_r`` and
    ``<name>_i`` of shape (nSimu,) + shape; ``Pos`` float64
    (nSimu, 500, 4, 672) with each simulation's scatterers in the first rows
    and zeros after; the others float64 (nSimu,) + shape.  Writing them needs
    h5py (not a dependency here)."""
    sims = [load_simulation(f) for f in sim_folders]
    out = {}
    for name in sims[0]:
        if name.startswith("data"):
            out[name + "_r"] = np.stack([np.real(s[name]) for s in sims]).astype(np.float64)
            out[name + "_i"] = np.stack([np.imag(s[name]) for s in sims]).astype(np.float64)
        elif name == "Pos":
            shape = (len(sims), POS_MAX_SCATTERERS) + sims[0][name].shape[1:]
            pos = np.zeros(shape, np.float64)
            for i, s in enumerate(sims):
                pos[i, :s[name].shape[0]] = s[name]
            out[name] = pos
        else:
            out[name] = np.stack([s[name] for s in sims]).astype(np.float64)
    return out


__all__ = ["normalize_rf", "realignment_grids", "realigned_half", "realigned_dataset",
           "realigned_set", "delays_ab", "save_simulation", "load_simulation", "job_arrays",
           "MAT_NAMES"]
