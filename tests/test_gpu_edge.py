"""GPU edge cases of the forward and adjoint kernels against the FP64 oracle
(bar as the other forward tests: rel-L2 <= 1e-5, masks to 1e-5):

* traces of as many samples as taps: the range test of
  bfVirtualSourceGeneric_RCA.cu:323 leaves no update, so the output is zero
  while the mask still sums the Hann weights (.cu:301-304); the adjoint is
  zero (shorter traces too);
* 24-sample traces whose start time is moved to the shallowest depth, so a
  few depths reach the trace: the depth-tiled and untiled launches against
  the oracle, FP32 and FP16-I/Q, and the adjoint against the oracle adjoint;
* a one-voxel grid;
* no voxels or no frames: the calls return at once and write nothing.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import rca_oracle as O
from simu_rca_ab_amd import _lib, configs as C
from simu_rca_ab_amd.plan import quantize_fp16
from tests import helpers as H
from tests.test_gpu_forward import OKIND, check, oracle_fwd, plan_for, say

pytestmark = pytest.mark.gpu

TAPS = {"quadratic": 3, "lanczos3": 6, "lanczos4": 8}


def _oracle_adj(cfg, half_index, g, interp, n_frame):
    G, T, P, Op = H.case(cfg, half_index)
    kind, a = OKIND[interp]
    return O.kernel_adjoint(g, H.grid_f32(G), O.prepare(G, T),
                            H.oracle_params(cfg, cfg.halves[half_index]), n_frame, kind, a)


@pytest.mark.parametrize("interp", list(TAPS))
@pytest.mark.parametrize("depth_run", [0, 64])
def test_traces_as_short_as_the_taps(dev, interp, depth_run):
    cfg = C.scaled(C.get("cfg1"), nx=5, ny=4, nz=64, n_sample=TAPS[interp])
    say(f"case {cfg.n_sample}-sample traces, interp={interp}, depth_run={depth_run}")
    g = torch.Generator().manual_seed(5)
    for half in (0, 1):
        plan, G, T = plan_for(cfg, half, dev, interp, depth_run=depth_run)
        iq = torch.randn((2, plan.n_tx, plan.n_channel, cfg.n_sample), dtype=torch.complex64,
                         generator=g)
        out, mask = plan.forward(iq.to(dev), compound=False)
        torch.cuda.synchronize()
        ref, rmask = oracle_fwd(cfg, half, iq, interp)
        assert not np.any(ref), "the oracle leaves every update out"
        assert torch.count_nonzero(torch.view_as_real(out)) == 0
        assert rmask.max() > 0
        assert H.rel_l2(mask.cpu().numpy(), rmask) <= 1e-5
        y = torch.randn((2, 1, plan.n_pix), dtype=torch.complex64, generator=g).to(dev)
        gi = plan.adjoint(y)
        torch.cuda.synchronize()
        assert torch.count_nonzero(torch.view_as_real(gi)) == 0
    # a trace shorter than the taps: the adjoint is zero as well
    short = C.scaled(cfg, n_sample=TAPS[interp] - 1)
    plan, _, _ = plan_for(short, 0, dev, interp, depth_run=depth_run)
    y = torch.randn((1, 1, plan.n_pix), dtype=torch.complex64, generator=g).to(dev)
    gi = torch.full((1, plan.n_tx, plan.n_channel, short.n_sample), 7 + 7j, dtype=torch.complex64,
                    device=dev)
    plan.adjoint(y, grad_iq=gi)
    torch.cuda.synchronize()
    assert torch.count_nonzero(torch.view_as_real(gi)) == 0


@pytest.mark.parametrize("interp", list(TAPS))
def test_short_traces_reached_by_a_few_depths(dev, interp):
    cfg = C.scaled(C.get("cfg1"), nx=5, ny=4, nz=64, n_sample=24, t0=-6.4e-6, n_scat=200,
                   scat_box=(0.5e-3, 0.4e-3, 5.0e-3, 5.8e-3))
    for half in (0, 1):
        say(f"case 24-sample traces from 6.4 us, interp={interp}, half={half}")
        iq = H.iq_frame(cfg, half, n_frame=2)
        ref, rmask = oracle_fwd(cfg, half, iq, interp)
        assert np.count_nonzero(ref) > 0 and np.count_nonzero(ref) < ref.size
        outs = {}
        for dr in (0, 64):
            plan, G, T = plan_for(cfg, half, dev, interp, depth_run=dr)
            out, mask = plan.forward(iq.to(dev), compound=False)
            torch.cuda.synchronize()
            outs[dr] = out.cpu().numpy()
            check(outs[dr], ref, f"{interp} half {half} depth_run {dr}")
            assert H.rel_l2(mask.cpu().numpy(), rmask) <= 1e-5
        hq, sc = quantize_fp16(iq.to(dev))
        o16, _ = plan.forward(hq, sc, compound=False)
        torch.cuda.synchronize()
        # 0.2 dB: on cases this small the -40 dB floor admits voxels whose FP16
        # storage error alone reaches 0.10-0.14 dB (the oracle on dequantised
        # I/Q shows the same); the 0.1 dB bar is asserted on config 2
        check(o16.cpu().numpy(), outs[64], f"{interp} half {half} FP16-I/Q vs FP32", rel=3e-3, env=0.2)
        gen = torch.Generator().manual_seed(11)
        y = torch.randn((2, 1, plan.n_pix), dtype=torch.complex64, generator=gen)
        gi = plan.adjoint(y.to(dev))
        torch.cuda.synchronize()
        gref = _oracle_adj(cfg, half, y.numpy(), interp, 2)
        r = H.rel_l2(gi.cpu().numpy(), gref)
        say(f"adjoint vs oracle adjoint: rel-L2 {r:.3e}")
        assert r <= 1e-5


def test_one_voxel(dev):
    cfg = C.scaled(C.get("cfg1"), nx=1, ny=1, nz=1, n_scat=50,
                   scat_box=(0.3e-3, 0.3e-3, 4.8e-3, 5.2e-3))
    for half in (0, 1):
        say(f"case one voxel, half={half}")
        iq = H.iq_frame(cfg, half, n_frame=3)
        plan, G, T = plan_for(cfg, half, dev, "lanczos4")
        for compound in (True, False):
            out, mask = plan.forward(iq.to(dev), compound=compound)
            torch.cuda.synchronize()
            ref, rmask = oracle_fwd(cfg, half, iq, "lanczos4")
            if compound:
                ref, rmask = ref.sum(axis=1, keepdims=True), rmask.sum(axis=0, keepdims=True)
            check(out.cpu().numpy(), ref, f"one voxel half {half} compound {compound}")
            assert H.rel_l2(mask.cpu().numpy(), rmask) <= 1e-5
        y = torch.randn((3, 1, 1), dtype=torch.complex64, generator=torch.Generator().manual_seed(3))
        gi = plan.adjoint(y.to(dev))
        torch.cuda.synchronize()
        r = H.rel_l2(gi.cpu().numpy(), _oracle_adj(cfg, half, y.numpy(), "lanczos4", 3))
        say(f"one voxel adjoint vs oracle adjoint: rel-L2 {r:.3e}")
        assert r <= 1e-5


def test_no_voxels_or_no_frames(dev):
    cfg = C.scaled(C.get("cfg1"), nx=4, ny=3, nz=16)
    say("case nPix = 0 and nFrame = 0: nothing written")
    plan, G, T = plan_for(cfg, 0, dev, "lanczos4")
    lib = _lib.load()
    iq = H.iq_frame(cfg, 0, n_frame=1).to(dev)
    out = torch.full((1, 1, plan.n_pix), 3 - 2j, dtype=torch.complex64, device=dev)
    mask = torch.full((1, plan.n_pix), 5.0, dtype=torch.float32, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    empty = _lib.RcaDasDesc.from_buffer_copy(plan.desc)
    empty.n_pix = 0
    for d, nf in ((empty, 1), (plan.desc, 0)):
        rc = lib.rca_forward(ctypes.byref(d), ctypes.c_void_p(iq.data_ptr()), _lib.IQ_C64, None, nf,
                             ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(mask.data_ptr()),
                             _lib.FLAG_COMPOUND, stream)
        assert rc == _lib.RCA_OK
    gi = torch.full((1, plan.n_tx, plan.n_channel, plan.n_sample), 1 + 1j, dtype=torch.complex64,
                    device=dev)
    rc = lib.rca_adjoint(ctypes.byref(plan.desc), ctypes.c_void_p(out.data_ptr()), 1, 0,
                         ctypes.c_void_p(gi.data_ptr()), 0, stream)
    assert rc == _lib.RCA_OK
    torch.cuda.synchronize()
    assert torch.all(out == (3 - 2j)) and torch.all(mask == 5.0)
    assert torch.all(gi == (1 + 1j))
    # no voxels: A is the zero map, so A^H y = 0 over the whole gradient
    rc = lib.rca_adjoint(ctypes.byref(empty), ctypes.c_void_p(out.data_ptr()), 1, 1,
                         ctypes.c_void_p(gi.data_ptr()), 0, stream)
    assert rc == _lib.RCA_OK
    torch.cuda.synchronize()
    assert torch.count_nonzero(torch.view_as_real(gi)) == 0
