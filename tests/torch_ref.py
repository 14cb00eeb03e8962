"""Plain-PyTorch reference of the per-tilt DAS forward (float64 geometry,
gather + weights as torch ops) for autograd cross-checks on small cases."""
from __future__ import annotations

import math

import numpy as np
import torch


def das_coefficients(grid, prep, params, lo, hi, W, device):
    """Per (tilt, voxel, channel, tap): sample index and complex coefficient
    w * q_k * exp(+j 2 pi fd (tau - 2 z / c)); zero outside aperture/range."""
    f64 = torch.float64
    gx, gy, gz = (torch.as_tensor(g.astype(np.float64), device=device) for g in grid)
    ex, ey, ez = (torch.as_tensor(prep[k].astype(np.float64), device=device) for k in ("ex", "ey", "ez"))
    vx, vy, vz, d0 = (torch.as_tensor(prep[k].astype(np.float64), device=device)
                      for k in ("vx", "vy", "vz", "d0"))
    fs, fd, c, t0, F = (float(np.float64(params[k])) for k in ("fs", "fd", "c", "t0", "fnumber"))
    ns = int(params["n_sample"])
    if params["receive_x"]:
        A = (gx[:, None] - ex[None, :]).abs()
        tx = torch.sqrt((gy[None] - vy[:, None]) ** 2 + (gz[None] - vz[:, None]) ** 2)
    else:
        A = (gy[:, None] - ey[None, :]).abs()
        tx = torch.sqrt((gx[None] - vx[:, None]) ** 2 + (gz[None] - vz[:, None]) ** 2)
    tx = (tx - d0[:, None]) / c
    apF = gz[:, None] / F
    ok_ap = A < apF
    w = 0.5 * (1 + torch.cos(math.pi * A / apF))
    rx = torch.sqrt(A ** 2 + (gz[:, None] - ez[None]) ** 2) / c
    u = (tx[:, :, None] + rx[None] + t0) * fs                  # [T, P, C]
    m = torch.floor(u)
    ok = ok_ap[None] & (m >= lo) & (m < ns - 1 - hi)
    d = u - m
    taps = torch.arange(-lo, hi + 1, device=device, dtype=f64)
    x = d[..., None] - taps
    if W == 0:
        dd = d[..., None]
        q = torch.cat([dd * (dd - 1) / 2, -(dd - 1) * (dd + 1), dd * (dd + 1) / 2], dim=-1)
    else:
        q = torch.sinc(x) * torch.sinc(x / W)
    ph = 2 * math.pi * fd * (u / fs - 2 * gz[None, :, None] / c)
    coef = q * (w[None] * ok)[..., None] * torch.exp(1j * ph)[..., None]
    idx = (m.to(torch.int64)[..., None] + torch.arange(-lo, hi + 1, device=device)).clamp(0, ns - 1)
    return idx, coef.to(torch.complex128)


def das_torch(iq, idx, coef):
    """iq complex [F, T, C, Ns] -> per-tilt output [F, T, P] (autograd-able)."""
    nf, nt, nc, ns = iq.shape
    out = []
    for t in range(nt):
        flat = iq[:, t].reshape(nf, nc * ns)                       # [F, C*Ns]
        lin = idx[t] + (torch.arange(nc, device=iq.device) * ns)[None, :, None]   # [P, C, K]
        g = flat[:, lin.reshape(-1)].reshape(nf, *lin.shape)       # [F, P, C, K]
        out.append((g * coef[t][None]).sum(dim=(-1, -2)))
    return torch.stack(out, dim=1)


def quantize_fp16(iq):
    """Torch statement of the FP16-I/Q storage: per-trace scale =
    max(|re|, |im|) (1 for a zero trace), samples (re, im) / scale rounded to
    float16.  Returns (float16 [..., Ns, 2], float32 [...])."""
    import torch
    r = torch.view_as_real(iq)
    scale = r.abs().amax(dim=(-1, -2)).to(torch.float32)
    scale = torch.where(scale > 0, scale, torch.ones_like(scale))
    return (r / scale[..., None, None]).to(torch.float16).contiguous(), scale.contiguous()


def dequantize_fp16(h, scale):
    """The complex64 values FP16-I/Q storage represents."""
    import torch
    return torch.view_as_complex((h.to(torch.float32) * scale[..., None, None]).contiguous())
