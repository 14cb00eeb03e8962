"""GPU parity of the FP16-I/Q forward and of the adjoint kernel.

FP16-I/Q bar: relative L2 <= 3e-3 against the FP64 oracle on the original
complex64 data, with envelope error <= 0.1 dB over voxels within -30 dB of the
maximum and <= 0.2 dB within -40 dB; <= 1e-5 (and 0.1 dB at -40 dB) against
the FP64 oracle on the dequantised data, i.e. the kernel's FP32 arithmetic
adds nothing to the storage error.  The -40 dB bound is looser than the FP32
one because 11-bit FP16 mantissas leak ~2.4e-4 of each bright echo into the
dim voxels: the FP64 oracle run on the dequantised data shows the same
0.10-0.14 dB worst voxel at the -40 dB floor.
Adjoint bar: relative L2 <= 1e-5 against the FP64 oracle adjoint; the
dot-product identity <A x, y> = <x, A^H y> to 1e-5.
"""
import numpy as np
import pytest
import torch

from oracle import rca_oracle as O
from simu_rca_ab_amd import _lib, configs as C
from simu_rca_ab_amd.plan import quantize_fp16
from tests.torch_ref import dequantize_fp16, quantize_fp16 as quantize_fp16_ref
from tests import helpers as H
from tests.test_gpu_forward import check, oracle_fwd, plan_for, say

pytestmark = pytest.mark.gpu

CFG = C.scaled(C.get("cfg1"), nx=6, ny=6, nz=48, n_scat=300)


@pytest.mark.parametrize("half_index", [0, 1])
def test_fp16_iq(dev, half_index):
    say(f"case FP16-I/Q half={half_index}, 3 frames, lanczos4")
    iq = H.iq_frame(CFG, half_index, n_frame=3)
    plan, G, T = plan_for(CFG, half_index, dev)
    h, s = quantize_fp16(iq.to(dev))
    out16, m16 = plan.forward(h, s)
    out32, _ = plan.forward(iq.to(dev))
    torch.cuda.synchronize()
    ref, _ = oracle_fwd(CFG, half_index, iq, "lanczos4")
    ref = ref.sum(axis=1, keepdims=True)
    check(out16.cpu().numpy(), ref, "fp16 vs FP64 oracle (original data)", rel=3e-3, env=0.2)
    e30 = H.envelope_db_err(out16.cpu().numpy(), ref, floor_db=-30.0)
    say(f"fp16 envelope error within -30 dB: {e30:.4f} dB")
    assert e30 <= 0.1
    check(out16.cpu().numpy(), out32.cpu().numpy(), "fp16 vs FP32 kernel", rel=3e-3, env=0.2)
    deq = dequantize_fp16(h, s).cpu()
    refq, _ = oracle_fwd(CFG, half_index, deq, "lanczos4")
    check(out16.cpu().numpy(), refq.sum(axis=1, keepdims=True), "fp16 vs FP64 oracle (dequantised)")


def test_quantize_kernel_matches_torch_statement(dev):
    """rca_quantize_fp16 against the torch statement of the storage format,
    bit for bit: traces of every magnitude, an all-zero trace, a trace whose
    largest component is imaginary, a length that is not a multiple of the
    workgroup."""
    say("case FP16-I/Q quantisation kernel vs torch statement")
    g = torch.Generator().manual_seed(5)
    x = torch.complex(torch.randn((3, 4, 5, 1000), generator=g), torch.randn((3, 4, 5, 1000), generator=g))
    x = x * torch.logspace(-8, 8, 60).reshape(3, 4, 5, 1)
    x[1, 2, 3] = 0
    x[2, 0, 1, 17] = 1j * 1e3
    x = x.to(torch.complex64).to(dev)
    h, s = quantize_fp16(x)
    hr, sr = quantize_fp16_ref(x)
    torch.cuda.synchronize()
    assert h.shape == hr.shape and s.shape == sr.shape
    assert torch.equal(s, sr) and torch.equal(h, hr)
    assert s[1, 2, 3].item() == 1.0 and not h[1, 2, 3].any()


ADJ = C.scaled(C.get("cfg0"), nx=8, ny=6, nz=24)


@pytest.mark.parametrize("interp", ["lanczos3", "quadratic"])
@pytest.mark.parametrize("gtx", [1, 11])
def test_adjoint_vs_oracle(dev, interp, gtx):
    say(f"case adjoint interp={interp} grad_tx={gtx}")
    plan, G, T = plan_for(ADJ, 0, dev, interp)
    rng = np.random.default_rng(1)
    g = (rng.standard_normal((2, gtx, plan.n_pix)) + 1j * rng.standard_normal((2, gtx, plan.n_pix))
         ).astype(np.complex64)
    got = plan.adjoint(torch.from_numpy(g).to(dev))
    torch.cuda.synchronize()
    prep = O.prepare(G, T)
    kind, a = {"lanczos3": (O.INTERP_LANCZOS, 3), "quadratic": (O.INTERP_QUADRATIC, 3)}[interp]
    ref = O.kernel_adjoint(g, H.grid_f32(G), prep, H.oracle_params(ADJ, ADJ.halves[0]), 2, kind, a)
    r = H.rel_l2(got.cpu().numpy(), ref)
    say(f"adjoint rel-L2 {r:.3e}")
    assert r <= 1e-5


def test_adjoint_frame_blocks_and_accumulate(dev):
    say("case adjoint frame blocks 1 and 2 (3 frames), accumulate")
    plan, G, T = plan_for(ADJ, 0, dev, "lanczos3")
    rng = np.random.default_rng(2)
    g = torch.from_numpy((rng.standard_normal((3, 1, plan.n_pix)) +
                          1j * rng.standard_normal((3, 1, plan.n_pix))).astype(np.complex64)).to(dev)
    try:
        _lib.set_tuning(0, 1, 0)
        a = plan.adjoint(g)
        _lib.set_tuning(0, 2, 0)
        b = plan.adjoint(g)
        c = b.clone()
        plan.adjoint(g, grad_iq=c, accumulate=True)
    finally:
        _lib.set_tuning(0, 0, 0)
    torch.cuda.synchronize()
    assert H.rel_l2(b.cpu().numpy(), a.cpu().numpy()) <= 1e-6
    assert H.rel_l2(c.cpu().numpy(), 2 * a.cpu().numpy()) <= 1e-6


def test_dot_product_identity(dev):
    say("case <A x, y> = <x, A^H y> on the GPU")
    plan, G, T = plan_for(ADJ, 0, dev, "lanczos3")
    x = H.iq_frame(ADJ, 0, n_frame=2).to(dev)
    rng = np.random.default_rng(4)
    y = torch.from_numpy((rng.standard_normal((2, 1, plan.n_pix)) +
                          1j * rng.standard_normal((2, 1, plan.n_pix))).astype(np.complex64)).to(dev)
    Ax, _ = plan.forward(x)
    AHy = plan.adjoint(y)
    lhs = torch.vdot(y.flatten().to(torch.complex128), Ax.flatten().to(torch.complex128))
    rhs = torch.vdot(AHy.flatten().to(torch.complex128), x.flatten().to(torch.complex128))
    r = abs(complex(lhs - rhs)) / abs(complex(lhs))
    say(f"dot-product mismatch {r:.3e}")
    assert r <= 1e-5
