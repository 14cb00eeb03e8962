"""Deterministic adjoint (rca_adjoint_det): the tiles' partial sums are added
as 64-bit fixed-point integers in a workspace and converted to FP32 at the
end, so repeated calls give bitwise identical gradients.  Checked against the
FP64 oracle adjoint (1e-5), against the float-atomic adjoint (rca_adjoint,
1e-6: the two differ only in FP32 summation order and the 2^-S quanta), for
bitwise repeatability, RCA_FLAG_ACCUMULATE, per-tilt gradients, a zero and a
non-finite grad_out, through torch.autograd, and at config 4's full size
(repeatability, the dot-product identity)."""
import numpy as np
import pytest
import torch

from oracle import rca_oracle as O
from simu_rca_ab_amd import configs as C
from simu_rca_ab_amd.autograd import das
from tests import helpers as H
from tests.test_gpu_forward import plan_for, say

pytestmark = pytest.mark.gpu


def _small(dev, nx=7, ny=6, nz=64, tiled=True):
    """Depth-tiled (both halves' kernels: the receive-X half's tiles go to the
    line-outer kernel) or flat geometry (the general kernel, untiled)."""
    from tests.test_gpu_fullsize import tiled_plan
    cfg = C.scaled(C.get("cfg0"), nx=nx, ny=ny, nz=nz)
    plan, G, T = tiled_plan(cfg, 0, dev) if tiled else plan_for(cfg, 0, dev, "lanczos3")
    return cfg, plan, G, T


@pytest.mark.parametrize("per_tilt,tiled", [(False, True), (True, True), (False, False)])
def test_det_adjoint_matches_oracle_and_float_adjoint(dev, per_tilt, tiled):
    cfg, plan, G, T = _small(dev, tiled=tiled)
    nf = 3
    gt = plan.n_tx if per_tilt else 1
    say(f"case deterministic adjoint {cfg.nx}x{cfg.ny}x{cfg.nz}, {nf} frames, grad_tx={gt}, tiled={tiled}")
    g = torch.Generator(device=dev).manual_seed(71)
    y = torch.randn((nf, gt, plan.n_pix), dtype=torch.complex64, device=dev, generator=g)
    a1 = plan.adjoint(y, deterministic=True)
    a2 = plan.adjoint(y, deterministic=True)
    af = plan.adjoint(y)
    torch.cuda.synchronize()
    assert torch.equal(torch.view_as_real(a1), torch.view_as_real(a2)), "two calls differ"
    r = H.rel_l2(a1.cpu().numpy(), af.cpu().numpy())
    say(f"det vs float-atomic adjoint rel-L2 {r:.2e}")
    assert r <= 1e-6
    prep, grid = O.prepare(G, T), H.grid_f32(G)
    prm = H.oracle_params(cfg, cfg.halves[0])
    want = O.kernel_adjoint(y.cpu().numpy(), grid, prep, prm, nf, O.INTERP_LANCZOS, 3)
    r = H.rel_l2(a1.cpu().numpy(), want)
    say(f"det adjoint vs FP64 oracle adjoint rel-L2 {r:.2e}")
    assert r <= 1e-5


def test_det_adjoint_accumulate_zero_and_nonfinite(dev):
    cfg, plan, G, T = _small(dev, nx=5, ny=4)
    say("case deterministic adjoint: accumulate, zero grad_out, NaN and Inf grad_out")
    g = torch.Generator(device=dev).manual_seed(72)
    y = torch.randn((2, 1, plan.n_pix), dtype=torch.complex64, device=dev, generator=g)
    base = torch.randn((2, plan.n_tx, plan.n_channel, plan.n_sample), dtype=torch.complex64,
                       device=dev, generator=g)
    acc = base.clone()
    plan.adjoint(y, grad_iq=acc, accumulate=True, deterministic=True)
    one = plan.adjoint(y, deterministic=True)
    torch.cuda.synchronize()
    r = H.rel_l2(acc.cpu().numpy(), (base + one).cpu().numpy())
    say(f"accumulate: rel-L2 vs base + A^H y {r:.2e}")
    assert r <= 1e-7
    z = plan.adjoint(torch.zeros_like(y), deterministic=True)
    assert torch.count_nonzero(torch.view_as_real(z)) == 0
    yn = y.clone()
    yn[1, 0, 3] = complex(float("nan"), 0.0)
    n = plan.adjoint(yn, deterministic=True)
    torch.cuda.synchronize()
    assert torch.isnan(torch.view_as_real(n)).all()
    yi = y.clone()
    yi[0, 0, 5] = complex(0.0, float("inf"))
    n = plan.adjoint(yi, deterministic=True)
    torch.cuda.synchronize()
    assert torch.isnan(torch.view_as_real(n)).all()


def test_det_adjoint_through_autograd(dev):
    cfg, plan, G, T = _small(dev, nx=6, ny=5)
    say("case autograd with the deterministic backward: two backward passes bitwise equal")
    x = H.iq_frame(cfg, 0, 2).to(dev)
    tgt = torch.randn(2, 1, plan.n_pix, dtype=torch.complex64, device=dev,
                      generator=torch.Generator(device=dev).manual_seed(3))
    grads = []
    for _ in range(2):
        xg = x.clone().requires_grad_()
        (das(xg, plan, True, deterministic=True) - tgt).abs().pow(2).sum().backward()
        grads.append(xg.grad)
    xf = x.clone().requires_grad_()
    (das(xf, plan, True) - tgt).abs().pow(2).sum().backward()
    assert torch.equal(torch.view_as_real(grads[0]), torch.view_as_real(grads[1]))
    r = H.rel_l2(grads[0].cpu().numpy(), xf.grad.cpu().numpy())
    say(f"det vs float-atomic gradient rel-L2 {r:.2e}")
    assert r <= 1e-6


def test_det_adjoint_cfg4_full(dev):
    from tests.test_gpu_fullsize import tiled_plan
    cfg = C.get("cfg4")
    nf = cfg.n_frame
    say("case deterministic adjoint at cfg4 full size: 96x96x192, 16 tilts, 96 lines, 8 frames")
    plan, G, T = tiled_plan(cfg, 0, dev)
    x = H.iq_frame(cfg, 0, n_frame=nf).to(dev)
    y_fwd, _ = plan.forward(x, want_mask=False)
    g = torch.Generator(device=dev).manual_seed(4045)
    yv = torch.randn((nf, 1, plan.n_pix), dtype=torch.complex64, device=dev, generator=g)
    a1 = plan.adjoint(yv, deterministic=True)
    a2 = plan.adjoint(yv, deterministic=True)
    af = plan.adjoint(yv)
    torch.cuda.synchronize()
    assert torch.equal(torch.view_as_real(a1), torch.view_as_real(a2)), "two calls differ"
    r = H.rel_l2(a1.cpu().numpy(), af.cpu().numpy())
    say(f"det vs float-atomic adjoint rel-L2 {r:.2e}")
    assert r <= 1e-6
    lhs = torch.vdot(yv.flatten().to(torch.complex128), y_fwd.flatten().to(torch.complex128))
    rhs = torch.vdot(a1.flatten().to(torch.complex128), x.flatten().to(torch.complex128))
    r = abs(complex(lhs - rhs)) / abs(complex(lhs))
    say(f"dot-product identity with the deterministic adjoint: mismatch {r:.3e}")
    assert r <= 1e-5
