"""torch.autograd through the HIP forward + adjoint kernels vs autograd of the
plain-PyTorch reference (float64) on the same inputs: gradients to 1e-5."""
import numpy as np
import pytest
import torch

from oracle import rca_oracle as O
from simu_rca_ab_amd import configs as C
from simu_rca_ab_amd.autograd import das
from tests import helpers as H, torch_ref as R
from tests.test_gpu_forward import plan_for, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("compound", [True, False])
def test_autograd_matches_torch_reference(dev, compound):
    cfg = C.scaled(C.get("cfg0"), nx=6, ny=5, nz=16)
    say(f"case autograd through DAS, compound={compound}")
    plan, G, T = plan_for(cfg, 0, dev, "lanczos3")
    prep, grid = O.prepare(G, T), H.grid_f32(G)
    idx, coef = R.das_coefficients(grid, prep, H.oracle_params(cfg, cfg.halves[0]), 2, 3, 3.0, dev)
    x = H.iq_frame(cfg, 0, 2).to(dev)
    nt = 1 if compound else 11
    tgt = torch.randn(2, nt, plan.n_pix, dtype=torch.complex64, device=dev,
                      generator=torch.Generator(device=dev).manual_seed(1))
    xg = x.clone().requires_grad_()
    loss = (das(xg, plan, compound) - tgt).abs().pow(2).sum()
    loss.backward()
    xr = x.to(torch.complex128).requires_grad_()
    yr = R.das_torch(xr, idx, coef)
    if compound:
        yr = yr.sum(dim=1, keepdim=True)
    lr = (yr - tgt.to(torch.complex128)).abs().pow(2).sum()
    lr.backward()
    r = H.rel_l2(xg.grad.cpu().numpy(), xr.grad.cpu().numpy())
    say(f"grad rel-L2 {r:.3e}, loss rel {abs(loss.item() - lr.item()) / lr.item():.2e}")
    assert r <= 1e-5


def test_differentiable_beamforming_training_runs(dev):
    """Small config-4-like run: DAS -> complex CNN -> MSE; gradients reach the
    I/Q through the adjoint kernel and SGD lowers the loss."""
    from simu_rca_ab_amd import dbf
    cfg = C.scaled(C.get("cfg4"), nx=12, ny=10, nz=24, n_frame=2)
    say("case differentiable beamforming: 12x10x24 voxels, 2 frames, 8-channel CNN")
    plan, G, T = plan_for(cfg, 0, dev, "lanczos3")
    iq = H.iq_frame(cfg, 0, n_frame=2).to(dev)
    r = dbf.time_cfg4(plan, iq, (cfg.nx, cfg.ny, cfg.nz), steps=4, ch=8)
    say(f"losses {r['losses']}")
    assert all(np.isfinite(r["losses"]))
    assert r["losses"][-1] < r["losses"][0]
    x = iq.clone().requires_grad_(True)
    y = das(x, plan, True)
    y.abs().pow(2).sum().backward()
    assert x.grad is not None and torch.isfinite(torch.view_as_real(x.grad)).all()
    assert x.grad.abs().max() > 0


def test_adjoint_central_finite_differences(dev):
    """The loss L(x) = sum |DAS(x) - t|^2 is quadratic in the I/Q, so the
    central difference (L(x + h v) - L(x - h v)) / 2h with h = 1e-3 equals the
    directional derivative Re<grad, v> up to FP32 rounding of the forward
    (relative 1e-3 here); grad comes from the adjoint kernel through
    torch.autograd (conjugate-Wirtinger convention: dL = Re sum conj(grad) v)."""
    cfg = C.scaled(C.get("cfg0"), nx=6, ny=5, nz=16)
    say("case adjoint vs central finite differences, h=1e-3, 3 random directions")
    plan, G, T = plan_for(cfg, 0, dev, "lanczos3")
    x = H.iq_frame(cfg, 0, 2).to(dev)
    gen = torch.Generator(device=dev).manual_seed(11)
    tgt = torch.randn(2, 1, plan.n_pix, dtype=torch.complex64, device=dev, generator=gen)

    def loss(z):
        return (das(z, plan, True) - tgt).to(torch.complex128).abs().pow(2).sum()

    xg = x.clone().requires_grad_()
    loss(xg).backward()
    g = xg.grad.to(torch.complex128)
    h = 1e-3
    scale = x.abs().max()
    for k in range(3):
        v = torch.randn(x.shape, dtype=torch.complex64, device=dev, generator=gen) * scale
        with torch.no_grad():
            fd = (loss(x + h * v) - loss(x - h * v)).item() / (2 * h)
        an = (g.conj() * v.to(torch.complex128)).real.sum().item()
        r = abs(fd - an) / abs(an)
        say(f"direction {k}: finite difference {fd:.6e}, adjoint {an:.6e}, rel {r:.2e}")
        assert r <= 1e-3
