"""Differentiable beamforming workload: DAS forward -> complex 3-D CNN -> MSE.

The complex layers follow the (real, imag) tensor-pair idiom of the
reference's complex CNNs (Python/models/ComplexTorch/ComplexConv.py:48-64,
Complex_blocks.py CReLU): a complex convolution is two real convolutions,
(W_r + j W_i) * (x_r + j x_i) = (W_r*x_r - W_i*x_i) + j (W_r*x_i + W_i*x_r),
and CReLU applies ReLU to the real and imaginary parts separately.  The
beamformer is :func:`simu_rca_ab_amd.autograd.das`, so the backward pass runs
the hand-written adjoint kernel to get the gradient with respect to the I/Q.
"""
from __future__ import annotations

import time

import torch
import torch.nn as nn

from .autograd import das


class ComplexConv3d(nn.Module):
    def __init__(self, cin, cout, k=3):
        super().__init__()
        self.re = nn.Conv3d(cin, cout, k, padding=k // 2)
        self.im = nn.Conv3d(cin, cout, k, padding=k // 2)

    def forward(self, xr, xi):
        return self.re(xr) - self.im(xi), self.re(xi) + self.im(xr)


class ComplexNet(nn.Module):
    """Four complex conv layers (1 -> ch -> ch -> ch -> 1), CReLU between."""

    def __init__(self, ch=32):
        super().__init__()
        self.layers = nn.ModuleList([ComplexConv3d(1, ch), ComplexConv3d(ch, ch),
                                     ComplexConv3d(ch, ch), ComplexConv3d(ch, 1)])

    def forward(self, xr, xi):
        for i, layer in enumerate(self.layers):
            xr, xi = layer(xr, xi)
            if i < len(self.layers) - 1:
                xr, xi = torch.relu(xr), torch.relu(xi)
        return xr, xi


def volume(out: torch.Tensor, nx: int, ny: int, nz: int):
    """[F, 1, P] complex (voxel p = iz + nz*(ix + nx*iy)) -> real/imag
    [F, 1, ny, nx, nz] tensors."""
    v = out[:, 0].reshape(out.shape[0], 1, ny, nx, nz)
    return v.real, v.imag


def train_step(iq, plan, net, target, opt, grid_shape, cnn_dtype=torch.bfloat16):
    """One step: FP32 beamforming (HIP forward, adjoint in the backward), the
    CNN under autocast to ``cnn_dtype`` (bf16 convolutions run on the matrix
    cores; None keeps FP32), FP32 loss."""
    nx, ny, nz = grid_shape
    opt.zero_grad(set_to_none=True)
    y = das(iq, plan, True)
    vr, vi = volume(y, nx, ny, nz)
    # scale the beamformed volume to O(1) before the network
    s = 1.0 / (y.detach().abs().amax() + 1e-12)
    with torch.autocast("cuda", dtype=cnn_dtype or torch.float32, enabled=cnn_dtype is not None):
        pr, pi = net(vr * s, vi * s)
    pr, pi = pr.float(), pi.float()
    loss = torch.mean((pr - target[0]) ** 2) + torch.mean((pi - target[1]) ** 2)
    loss.backward()
    opt.step()
    return loss.detach()


def time_cfg4(plan, iq, grid_shape, steps=3, ch=32, seed=4004, log=None, cnn_dtype=torch.bfloat16):
    """Times forward+backward training steps and the adjoint (backward of the
    beamformer) alone.  Returns a dict of milliseconds."""
    dev = iq.device
    torch.manual_seed(seed)
    net = ComplexNet(ch).to(dev)
    g = torch.Generator(device=dev).manual_seed(seed)
    nx, ny, nz = grid_shape
    tshape = (iq.shape[0], 1, ny, nx, nz)
    target = (torch.randn(tshape, device=dev, generator=g) * 0.1,
              torch.randn(tshape, device=dev, generator=g) * 0.1)
    opt = torch.optim.SGD(net.parameters(), lr=1e-3)
    x = iq.clone().requires_grad_(True)
    say = log or (lambda *a: None)
    say("[dbf] warm-up step (first use compiles the convolution kernels)")
    losses = [train_step(x, plan, net, target, opt, grid_shape, cnn_dtype).item()]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        losses.append(train_step(x, plan, net, target, opt, grid_shape, cnn_dtype).item())
        say(f"[dbf] step {i + 1}/{steps} loss {losses[-1]:.5g}")
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t0) * 1e3 / steps
    gout = torch.randn((iq.shape[0], 1, plan.n_pix), dtype=torch.complex64, device=dev, generator=g)
    plan.adjoint(gout)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.adjoint(gout)
    torch.cuda.synchronize()
    adj_ms = (time.perf_counter() - t0) * 1e3 / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.forward(iq, want_mask=False)
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t0) * 1e3 / steps
    return {"train_step_ms": step_ms, "das_forward_ms": fwd_ms, "das_adjoint_ms": adj_ms,
            "cnn_dtype": str(cnn_dtype).replace("torch.", "") if cnn_dtype else "float32",
            "losses": losses, "updates_per_pass": plan.updates(iq.shape[0])}
