"""The plain-PyTorch reference DAS (used for autograd checks) agrees with the
FP64 oracle, and its autograd gradient agrees with the oracle adjoint (CPU)."""
import numpy as np
import pytest
import torch

from oracle import rca_oracle as O
from simu_rca_ab_amd import configs as C
from tests import helpers as H, torch_ref as R

CFG = C.scaled(C.get("cfg0"), nx=5, ny=4, nz=10)


@pytest.mark.parametrize("kind,lo,hi,W", [(O.INTERP_LANCZOS, 2, 3, 3.0), (O.INTERP_QUADRATIC, 1, 1, 0)])
def test_torch_reference_matches_oracle(kind, lo, hi, W):
    G, T, P, Op = H.case(CFG)
    prep, grid = O.prepare(G, T), H.grid_f32(G)
    params = H.oracle_params(CFG, CFG.halves[0])
    idx, coef = R.das_coefficients(grid, prep, params, lo, hi, W, "cpu")
    x = H.iq_frame(CFG, 0, 2)
    ref, _ = O.kernel_forward(x.numpy().view(np.float32), grid, prep, params, kind, 3)
    assert H.rel_l2(R.das_torch(x.to(torch.complex128), idx, coef).numpy(), ref) < 1e-12
    # autograd of the torch reference == oracle adjoint
    xg = x.to(torch.complex128).requires_grad_()
    g = torch.randn(2, 1, grid[0].size, dtype=torch.complex128, generator=torch.Generator().manual_seed(0))
    y = R.das_torch(xg, idx, coef).sum(dim=1, keepdim=True)
    (y * g.conj()).real.sum().backward()
    adj = O.kernel_adjoint(g.numpy(), grid, prep, params, 2, kind, 3)
    assert H.rel_l2(xg.grad.numpy(), adj) < 1e-12
