"""Seeded analytic I/Q synthesis for the RCA workloads.

Point scatterers insonified by the same virtual-source transmit model the
beamformer assumes (RCA.m:527-558; RCA.cu:256-269) and received by line
elements (RCA.cu:284-309).  A scatterer s with amplitude A_s and round-trip
time tau_s(t, c) contributes to trace (t, c) the demodulated echo

    IQ[n] += A_s * exp(-(n/fs - t0 - tau_s)^2 / (2 sigma^2)) * exp(-j 2 pi f0 tau_s)

with a Gaussian envelope whose FWHM is two cycles of f0.  With this
convention the beamformer's phase rotation exp(+j 2 pi fd tau) re-aligns the
echo exactly at the scatterer (RCA.cu:371-380).

Layout of every returned array: complex64 [nFrame, nTilt, nChannel, nSample]
(C-order), which is the reference kernel's interleaved
[2*nSample, nChannel, nTx, nFrame] column-major buffer (RCA.m:576-578).

The computation runs in torch (float64) on CPU or GPU; each echo touches
only +-6 sigma around its delay, so the cost is O(tilts*channels*scatterers).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import configs as C


def scatterers(cfg: C.RcaConfig):
    """Seeded positions [S, 3] (m) and real Gaussian amplitudes [S]."""
    hx, hy, z0, z1 = cfg.scat_box
    rp = np.random.Generator(np.random.PCG64(cfg.seeds[0]))
    ra = np.random.Generator(np.random.PCG64(cfg.seeds[1]))
    pos = np.stack([rp.uniform(-hx, hx, cfg.n_scat), rp.uniform(-hy, hy, cfg.n_scat),
                    rp.uniform(z0, z1, cfg.n_scat)], axis=1)
    amp = ra.standard_normal(cfg.n_scat)
    return pos, amp


def echoes(cfg: C.RcaConfig, half: C.Half, pos, amp, device="cpu", out=None, geom=None,
           extra_delay=None):
    """Analytic I/Q of one frame of one half: complex64 [nTilt, nCh, Ns].

    ``geom`` optionally replaces the FP64 transmit/receive geometry with
    given arrays (keys vx, vy, vz, d0 per tilt and e per element), e.g. the
    FP32-rounded values a beamformer receives.  ``extra_delay`` (seconds,
    [nCh] or [nTilt, nCh]) delays every echo on that element, as an
    aberrating layer in front of the probe would."""
    dev = torch.device(device)
    f64 = torch.float64
    if geom is None:
        geom = dict(zip(("vx", "vy", "vz", "d0"), C.virtual_sources(half, cfg.n_elem)))
        geom["e"] = C.element_positions(cfg.n_elem)
    vx, vy, vz, d0, e = (torch.as_tensor(np.asarray(geom[k], dtype=np.float64), device=dev)
                         for k in ("vx", "vy", "vz", "d0", "e"))
    p = torch.as_tensor(pos, dtype=f64, device=dev)
    a = torch.as_tensor(amp, dtype=f64, device=dev)
    nt, nc, ns = len(half.theta), cfg.n_elem, cfg.n_sample
    sig = C.sigma_pulse(cfg.fd)
    half_w = int(math.ceil(6 * sig * cfg.fs)) + 1
    win = torch.arange(-half_w, half_w + 1, device=dev, dtype=torch.int64)
    if out is None:
        out = torch.zeros(nt, nc, ns, dtype=torch.complex64, device=dev)
    flat = torch.view_as_real(out).view(nt, nc * ns, 2)
    xs, ys, zs = p[:, 0], p[:, 1], p[:, 2]
    if half.receive_x:
        rx = torch.sqrt((xs[None, :] - e[:, None]) ** 2 + zs[None, :] ** 2)        # [C, S]
    else:
        rx = torch.sqrt((ys[None, :] - e[:, None]) ** 2 + zs[None, :] ** 2)
    chan = torch.arange(nc, device=dev)[:, None, None] * ns
    if extra_delay is not None:
        xd = torch.as_tensor(np.asarray(extra_delay, dtype=np.float64), device=dev)
        xd = xd.reshape(1, nc).expand(nt, nc) if xd.numel() == nc else xd.reshape(nt, nc)
    for t in range(nt):
        if half.receive_x:
            tx = torch.sqrt((ys - vy[t]) ** 2 + (zs - vz[t]) ** 2) - d0[t]          # [S]
        else:
            tx = torch.sqrt((xs - vx[t]) ** 2 + (zs - vz[t]) ** 2) - d0[t]
        tau = (tx[None, :] + rx) / cfg.c                                             # [C, S]
        if extra_delay is not None:
            tau = tau + xd[t][:, None]
        ncen = torch.floor((tau + cfg.t0) * cfg.fs).to(torch.int64)
        n = ncen[..., None] + win                                                    # [C, S, W]
        tn = n.to(f64) / cfg.fs - cfg.t0
        env = torch.exp(-(tn - tau[..., None]) ** 2 / (2 * sig * sig)) * a[None, :, None]
        ph = -2 * math.pi * cfg.fd * tau
        re = env * torch.cos(ph)[..., None]
        im = env * torch.sin(ph)[..., None]
        ok = (n >= 0) & (n < ns)
        idx = (n.clamp(0, ns - 1) + chan)[ok]
        flat[t, :, 0].index_add_(0, idx, re[ok].to(torch.float32))
        flat[t, :, 1].index_add_(0, idx, im[ok].to(torch.float32))
    return out


def add_noise(iq: torch.Tensor, snr_db: float, seed: int):
    """Complex Gaussian noise at ``snr_db`` below each trace's mean power."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pw = (iq.abs() ** 2).mean(dim=-1, keepdim=True).to(torch.float64)
    std = torch.sqrt(pw / (10 ** (snr_db / 10)) / 2).to(torch.float32)
    nz = torch.from_numpy(rng.standard_normal(tuple(iq.shape) + (2,)).astype(np.float32))
    nz = torch.view_as_complex(nz.to(iq.device))
    return iq + nz * std


def frame(cfg: C.RcaConfig, device="cpu"):
    """One frame of every half: list of complex64 [1, nTilt, nCh, Ns]."""
    pos, amp = scatterers(cfg)
    out = []
    for hi, h in enumerate(cfg.halves):
        iq = echoes(cfg, h, pos, amp, device)
        if cfg.snr_db is not None:
            iq = add_noise(iq, cfg.snr_db, cfg.seeds[2] + 17 * hi)
        out.append(iq[None])
    return out


def ensemble(cfg: C.RcaConfig, n_frame: int | None = None, device="cpu"):
    """Doppler ensemble: scatterers moving at seeded velocities |v| <= vmax
    plus a static tissue slab ``tissue_db`` brighter.  Returns one complex64
    [nFrame, nTilt, nCh, Ns] array per half."""
    nf = n_frame or cfg.n_frame
    ex = cfg.extra
    pos, amp = scatterers(cfg)
    rv = np.random.Generator(np.random.PCG64(cfg.seeds[2] + 101))
    d = rv.standard_normal((cfg.n_scat, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    vel = d * rv.uniform(0, ex.get("vmax", 5e-3), (cfg.n_scat, 1))
    nt = int(ex.get("n_tissue", 0))
    if nt:
        rt = np.random.Generator(np.random.PCG64(cfg.seeds[2] + 202))
        hx, hy = cfg.scat_box[0], cfg.scat_box[1]
        tz0, tz1 = ex["tissue_z"]
        tpos = np.stack([rt.uniform(-hx, hx, nt), rt.uniform(-hy, hy, nt),
                         rt.uniform(tz0, tz1, nt)], axis=1)
        tamp = rt.standard_normal(nt) * 10 ** (ex.get("tissue_db", 40.0) / 20)
    outs = []
    for h in cfg.halves:
        arr = torch.zeros(nf, len(h.theta), cfg.n_elem, cfg.n_sample, dtype=torch.complex64,
                          device=device)
        # the tissue slab is static: its echoes once, then each frame adds the
        # moving scatterers' echoes to a copy
        base = echoes(cfg, h, tpos, tamp, device) if nt else None
        for f in range(nf):
            p = pos + vel * (f * ex.get("frame_dt", 1e-3))
            if base is not None:
                arr[f].copy_(base)
            echoes(cfg, h, p, amp, device, out=arr[f])
        outs.append(arr)
    return outs


def to_matlab_layout(iq: torch.Tensor | np.ndarray):
    """[nFrame, nTilt, nCh, Ns] -> the reference's data argument
    [nSample, nColX=nCh, nColY=1, nEvent=nTilt*nFrame] (a column-major view of
    the same memory; script_All.m:371-373)."""
    nf, nt, nc, ns = iq.shape
    if isinstance(iq, np.ndarray):
        return iq.reshape(nf * nt, nc, ns).transpose(2, 1, 0)[:, :, None, :]
    return iq.reshape(nf * nt, nc, ns).permute(2, 1, 0).unsqueeze(2)
