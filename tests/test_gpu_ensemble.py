"""Config 2, the precision mode on a Doppler ensemble: config-1 geometry (128 +
128 elements, 2 x 16 tilts, 128x128x256 voxels, Lanczos a=4), scatterers
moving at seeded velocities up to 5 mm/s with 1 ms between frames, and a
static tissue slab (z 10-11 mm) 40 dB brighter.  The frame subset is the
first 8 of the 200 frames; I/Q stored as FP16 with per-trace FP32 scales.

Bars: FP16 against the FP32 kernel on all 8 frames: relative L2 <= 3e-3 and
envelope error <= 0.1 dB over voxels within -40 dB of the maximum (measured
0.03 dB); FP16 against the FP64 oracle on frame 0 (3000 sampled voxels):
3e-3 and 0.1 dB; the FP32 kernel against the oracle: 1e-5 and 0.1 dB.  The
launch is the depth-tiled one the drop-in uses (depth hint Nz).  All 200
frames run in tests/test_gpu_fullsize.py.
"""
import numpy as np
import pytest
import torch

from oracle import rca_oracle as O
from simu_rca_ab_amd import configs as C, synth
from simu_rca_ab_amd.plan import quantize_fp16
from tests import helpers as H
from tests.test_gpu_forward import check, plan_for, say

pytestmark = pytest.mark.gpu

N_FRAME = 8


def test_cfg2_doppler_ensemble_fp16(dev):
    cfg = C.get("cfg2")
    say(f"case cfg2 Doppler ensemble, first {N_FRAME} of 200 frames, 40 dB tissue slab, FP16 vs FP32 vs oracle")
    halves = synth.ensemble(cfg, n_frame=N_FRAME, device=dev)
    out32 = out16 = mask = None
    for hi in range(2):
        plan, _, _ = plan_for(cfg, hi, dev, depth_run=cfg.nz)
        h, s = quantize_fp16(halves[hi])
        if out32 is None:
            out32, mask = plan.forward(halves[hi])
            out16, _ = plan.forward(h, s, want_mask=False)
        else:
            plan.forward(halves[hi], out=out32, mask=mask, accumulate=True)
            plan.forward(h, s, out=out16, want_mask=False, accumulate=True)
        del plan, h, s
    torch.cuda.synchronize()
    assert torch.isfinite(out32).all() and torch.isfinite(out16).all()
    a16, a32 = out16[:, 0].cpu().numpy(), out32[:, 0].cpu().numpy()
    check(a16, a32, f"fp16 vs FP32 kernel, {N_FRAME} frames", rel=3e-3, env=0.1)
    # the frames differ (moving scatterers), so the ensemble is not one frame repeated
    assert H.rel_l2(a32[N_FRAME - 1], a32[0]) > 1e-3
    rng = np.random.default_rng(321)
    sel = np.sort(rng.choice(cfg.n_vox, 3000, replace=False))
    ref = 0
    for hi in range(2):
        G, T, P, Op = H.case(cfg, hi)
        Gs = dict(G)
        for k in ("gridBfX", "gridBfY", "gridBfZ"):
            Gs[k] = np.asarray(G[k]).ravel(order="F")[sel]
        prep = O.prepare(G, T)
        iq0 = halves[hi][:1].cpu().contiguous()
        r, _ = O.kernel_forward(iq0.numpy().view(np.float32), H.grid_f32(Gs), prep,
                                H.oracle_params(cfg, cfg.halves[hi]), O.INTERP_LANCZOS, 4)
        ref = ref + r[0].sum(axis=0)
    check(a32[0][sel], ref, "FP32 vs FP64 oracle, frame 0 (3000 sampled voxels)")
    check(a16[0][sel], ref, "fp16 vs FP64 oracle, frame 0 (3000 sampled voxels)", rel=3e-3, env=0.1)
