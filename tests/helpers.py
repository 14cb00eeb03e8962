"""Shared builders and error metrics for the tests."""
from __future__ import annotations

import numpy as np

from simu_rca_ab_amd import configs as C, synth


def rel_l2(got, want):
    got = np.asarray(got).ravel()
    want = np.asarray(want).ravel()
    return float(np.linalg.norm(got - want) / np.linalg.norm(want))


def envelope_db_err(got, want, floor_db=-40.0):
    """max |20 log10|got| - 20 log10|want|| over voxels within floor_db of the max."""
    g = np.abs(np.asarray(got).ravel())
    w = np.abs(np.asarray(want).ravel())
    keep = w >= w.max() * 10 ** (floor_db / 20)
    return float(np.max(np.abs(20 * np.log10(np.maximum(g[keep], 1e-30)) -
                               20 * np.log10(w[keep]))))


def case(cfg: C.RcaConfig, half_index: int = 0, compounding: bool = True):
    """(Geometry, Transmit, Param, Options) of one half, grids as meshgrid volumes."""
    h = cfg.halves[half_index]
    return C.geometry(cfg, h), C.transmit(h), C.param(cfg), C.options(cfg, compounding)


def flat_geometry(G):
    """Geometry with the grids flattened column-major (the kernel's voxel list)."""
    Gf = dict(G)
    for k in ("gridBfX", "gridBfY", "gridBfZ"):
        Gf[k] = np.asarray(G[k]).ravel(order="F")
    return Gf


def oracle_params(cfg: C.RcaConfig, half: C.Half):
    return dict(receive_x=half.receive_x, fs=np.float32(cfg.fs), fd=np.float32(cfg.fd),
                c=np.float32(cfg.c), t0=np.float32(cfg.t0), fnumber=np.float32(cfg.fnumber),
                n_sample=cfg.n_sample)


def grid_f32(G):
    return tuple(np.ascontiguousarray(np.asarray(G[k], dtype=np.float32).ravel(order="F"))
                 for k in ("gridBfX", "gridBfY", "gridBfZ"))


def iq_frame(cfg, half_index=0, n_frame=1, seed_shift=0):
    """Seeded analytic I/Q complex64 [F, T, C, Ns] (torch, CPU) for one half."""
    import torch
    outs = []
    for f in range(n_frame):
        c2 = C.scaled(cfg, seeds=(cfg.seeds[0] + 7 * (f + seed_shift), cfg.seeds[1] + 7 * (f + seed_shift),
                                  cfg.seeds[2]))
        pos, amp = synth.scatterers(c2)
        iq = synth.echoes(c2, cfg.halves[half_index], pos, amp)
        if cfg.snr_db is not None:
            iq = synth.add_noise(iq, cfg.snr_db, cfg.seeds[2] + f)
        outs.append(iq)
    return torch.stack(outs).contiguous()


def tile_voxels(n_pix, depth_run, bx, kind="forward", tile_cols=16):
    """Voxel indices of workgroup ``bx`` of a depth-tiled launch (valid ones,
    lane order), as the kernels' voxel_of assigns them: forward tiles are 16
    depths x 16 neighbouring columns (4 waves stacked in depth), adjoint tiles
    64 depths x 4 columns (one column per wave)."""
    cw, dw = (tile_cols, 4) if kind == "forward" else (1, 1)
    wd, td, wc = 64 // cw, dw * (64 // cw), 4 // dw
    ndc = -(-depth_run // td)
    dc, cg = bx % ndc, bx // ndc
    tid = np.arange(256)
    l, w = tid & 63, tid >> 6
    col = cg * (cw * wc) + (w // dw) * cw + (l % cw)
    dep = dc * td + (w % dw) * wd + l // cw
    p = dep + depth_run * col
    return p[(dep < depth_run) & (p < n_pix)]


def n_tiles(n_pix, depth_run, kind="forward", tile_cols=16):
    """(number of workgroups, depth tiles per column group) of a tiled launch."""
    cw, dw = (tile_cols, 4) if kind == "forward" else (1, 1)
    td, tc = dw * (64 // cw), cw * (4 // dw)
    ncol = -(-n_pix // depth_run)
    ndc = -(-depth_run // td)
    return -(-ncol // tc) * ndc, ndc


def edge_tiles(nx, ny, nz, n_random=8, seed=0, kind="forward"):
    """Workgroups of a [Nz, Nx, Ny] grid at the lateral corners and edges and
    the centre, each at the first, middle and last depth tile, plus a few
    seeded random ones; returns the union of their voxel indices (sorted) and
    the workgroup list."""
    n_pix = nx * ny * nz
    nblk, ndc = n_tiles(n_pix, nz, kind)
    tc = 16 if kind == "forward" else 4
    ngroups = nblk // ndc
    def grp(ix, iy):
        return min(ngroups - 1, (iy * nx + ix) // tc)
    gs = {grp(0, 0), grp(nx - 1, 0), grp(0, ny - 1), grp(nx - 1, ny - 1), grp(nx // 2, 0),
          grp(0, ny // 2), grp(nx - 1, ny // 2), grp(nx // 2, ny - 1), grp(nx // 2, ny // 2)}
    blocks = sorted({g * ndc + d for g in gs for d in (0, ndc // 2, ndc - 1)})
    rng = np.random.default_rng(seed)
    blocks = sorted(set(blocks) | set(rng.choice(nblk, size=min(n_random, nblk), replace=False).tolist()))
    vox = np.unique(np.concatenate([tile_voxels(n_pix, nz, b, kind) for b in blocks]))
    return vox, blocks


def rel_l2_t(got, want):
    """rel-L2 of two complex torch tensors, computed where they are."""
    import torch
    d = (got - want).to(torch.complex128)
    return float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(want.to(torch.complex128)))


def envelope_db_err_t(got, want, floor_db=-40.0):
    """envelope_db_err on torch tensors, computed where they are."""
    import torch
    g = got.abs().reshape(-1).to(torch.float64)
    w = want.abs().reshape(-1).to(torch.float64)
    keep = w >= w.max() * 10 ** (floor_db / 20)
    return float((20 * torch.log10(g[keep].clamp_min(1e-30)) - 20 * torch.log10(w[keep])).abs().max())
