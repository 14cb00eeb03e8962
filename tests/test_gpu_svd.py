"""GPU parity of the SVD clutter filter (rca_svd_filter) against the FP64
oracle (oracle/svd_oracle.py).

The library computes the Gram matrix, its eigen-decomposition and the
projection in FP64 and rounds the result to complex64 once, so the bar is
rel-L2 <= 1e-6 on the filtered ensemble (complex64 rounding is 6e-8) and
1e-9 on the eigenvalues.  Cases: tissue + blood ensembles with high-pass,
band and stop bands, both reference conventions (Python and MATLAB, with a
noise mask), a multi-chunk projection (> 2^20 pixels), torch and numpy
inputs, and the beamformer's own FP16 ensemble output filtered in place.
Both projection paths run: the low-rank kernel (kept or removed set of at
most 16 vectors: high3 over 64 and 300 frames, low2, none, all for its
4-vector instantiation, high6 for 8, stop for 16) and the ZGEMM by the full
projector (band, and the multi-chunk case).
"""
import numpy as np
import pytest
import torch

from oracle import svd_oracle as S
from simu_rca_ab_amd import clutterfilter as CF
from tests import helpers as H
from tests.test_gpu_forward import say

pytestmark = pytest.mark.gpu

REL = 1e-6


def frames_of(Y):
    """[Npix, Nt] -> frame-major complex64 [Nt, Npix] on the GPU."""
    return torch.from_numpy(np.ascontiguousarray(Y.T.astype(np.complex64))).cuda()


@pytest.mark.parametrize("npix,nt,band", [(20000, 64, "high3"), (12000, 200, "band"),
                                          (5000, 32, "stop"), (8000, 48, "low2"),
                                          (3000, 24, "none"), (3000, 24, "all"),
                                          (4000, 300, "high3"), (6000, 40, "high6")])
def test_filter_vs_oracle(dev, npix, nt, band):
    say(f"case svd filter npix={npix} nt={nt} keep={band}")
    Y, _, _ = S.tissue_blood_ensemble(npix, nt, rank=3, tissue_db=40.0, seed=npix)
    Y = Y.astype(np.complex64).astype(np.complex128)   # the GPU sees complex64
    keep = np.ones(nt, bool)
    if band.startswith("high"):
        keep[:int(band[4:])] = False
    elif band == "band":
        keep[:] = False
        keep[int(0.1 * nt):int(0.9 * nt)] = True
    elif band == "stop":
        keep[5:nt // 2] = False
    elif band == "low2":
        keep[:] = False
        keep[:2] = True
    elif band == "none":
        keep[:] = False
    out, ev = CF.svd_filter(frames_of(Y), keep)
    torch.cuda.synchronize()
    ref, w = S.project(Y, keep)
    got = out.cpu().numpy().T
    e = np.max(np.abs(ev.cpu().numpy() - w) / w[0])
    if band == "none":
        say(f"max |out| {np.abs(got).max():.3e}; eigenvalues max rel err {e:.3e}")
        assert not np.any(got)
    else:
        r = H.rel_l2(got, ref)
        say(f"rel-L2 {r:.3e}; eigenvalues max rel err {e:.3e}")
        assert r <= REL
    assert e <= 1e-9


def test_multi_chunk_projection(dev):
    """20 kept and 20 removed vectors: the ZGEMM path, in two pixel chunks."""
    npix, nt = (1 << 20) + 4321, 40
    say(f"case svd filter over {npix} pixels (two projection chunks), nt={nt}, keep 10..29")
    Y, _, _ = S.tissue_blood_ensemble(npix, nt, rank=2, tissue_db=30.0, seed=3)
    Y = Y.astype(np.complex64).astype(np.complex128)
    keep = np.zeros(nt, bool)
    keep[10:30] = True
    out, ev = CF.svd_filter(frames_of(Y), keep)
    torch.cuda.synchronize()
    ref, w = S.project(Y, keep)
    r = H.rel_l2(out.cpu().numpy().T, ref)
    say(f"rel-L2 {r:.3e}")
    assert r <= REL


def test_reference_signatures(dev):
    say("case SVDClutterFilter (python) and SVDClutterFilterMatlab with a mask, numpy and torch")
    rng = np.random.default_rng(9)
    Y, _, _ = S.tissue_blood_ensemble(40 * 30, 50, rank=2, seed=9)
    IQ = Y.reshape(40, 30, 50, order="F").astype(np.complex64)
    got, s = CF.SVDClutterFilter(IQ, None, [0.04, 1.0])
    ref, w = S.svd_filter_python(IQ.astype(np.complex128), [0.04, 1.0])
    assert got.shape == IQ.shape and isinstance(got, np.ndarray)
    assert H.rel_l2(got, ref) <= REL
    np.testing.assert_allclose(s, w, rtol=1e-9, atol=1e-9 * w[0])
    gt, st = CF.SVDClutterFilter(torch.from_numpy(IQ).cuda(), None, [0.04, 1.0])
    assert isinstance(gt, torch.Tensor) and tuple(gt.shape) == IQ.shape
    assert H.rel_l2(gt.cpu().numpy(), ref) <= REL
    mask = (1.0 + rng.random((40, 30))).astype(np.float32)
    for kind, wv in (("band", [0.06, 0.9]), ("high", [0.06]), ("low", [0.5]), ("stop", [0.1, 0.6])):
        got, ev = CF.SVDClutterFilterMatlab(IQ, mask, wv, kind)
        ref, w = S.svd_filter_matlab(IQ.astype(np.complex128), mask.astype(np.float64), wv, kind)
        r = H.rel_l2(got, ref)
        say(f"matlab {kind}: rel-L2 {r:.3e}")
        assert r <= REL
        np.testing.assert_allclose(ev, w, rtol=1e-9, atol=1e-9 * w[0])
    with pytest.raises(ValueError):
        CF.SVDClutterFilterMatlab(IQ, None, [-0.1, 0.5])


def test_beamformed_fp16_ensemble(dev):
    from simu_rca_ab_amd.plan import quantize_fp16
    from tests.test_gpu_tiled import TILE, plans
    say("case beamformed FP16 ensemble (12 frames) -> svd filter")
    flat, tiled, G = plans(TILE, 0, dev, "lanczos4")
    base = H.iq_frame(TILE, 0, n_frame=1)
    ph = torch.exp(1j * torch.linspace(0, 3, 12)).to(torch.complex64)
    iq = (base * ph[:, None, None, None]).contiguous().to(dev)
    iq = iq + 0.01 * torch.randn(iq.shape, dtype=torch.complex64, device=dev)
    h, s = quantize_fp16(iq)
    vol, _ = tiled.forward(h, s, want_mask=False)
    frames = vol[:, 0].contiguous()
    keep = np.ones(12, bool)
    keep[:1] = False
    out, ev = CF.svd_filter(frames, keep)
    torch.cuda.synchronize()
    ref, w = S.project(frames.cpu().numpy().T.astype(np.complex128), keep)
    r = H.rel_l2(out.cpu().numpy().T, ref)
    say(f"rel-L2 {r:.3e}; first eigenvalue share {w[0] / w.sum():.4f}")
    assert r <= REL
