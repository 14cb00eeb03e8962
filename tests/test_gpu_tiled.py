"""GPU parity of the depth-tiled forward launch (grid hint ``depth_run``).

With the hint, a forward workgroup owns 16 depths x 16V neighbouring columns
(the adjoint 64 depths x 4V columns) and stages only the sample window that its
voxels' taps can reach on each (tilt chunk, line).  Its
results must match the untiled launch and the FP64 oracle (1e-5, 0.1 dB).
Tiled and untiled launches add the same per-voxel terms, but the tilts per
chunk (set by the LDS budget) fix the order in which tilts are added, so
compounded sums agree to FP32 rounding: TOL = 4e-6 relative L2 (per-tilt
outputs, whose order does not change, agree to 1e-6).  Cases: every
interpolation, both halves, per-tilt and compounded output, frame counts that
leave partial frame blocks, column counts that leave partial column groups, a
hint that does not divide the voxel count, a depth-coarse grid whose windows
are longer than the LDS rows (staged in several passes), several LDS budgets
(tilts per chunk), FP16-I/Q, NoSum, and the adjoint (tiled vs untiled, FP64
oracle adjoint, frame blocks, per-tilt gradients, accumulation).
"""
import numpy as np
import pytest
import torch

from oracle import rca_oracle as O
from simu_rca_ab_amd import _lib, configs as C
from simu_rca_ab_amd.plan import RcaPlan, quantize_fp16
from tests import helpers as H
from tests.test_gpu_forward import OKIND, check, oracle_fwd, say

pytestmark = pytest.mark.gpu

TOL = 4e-6

# 36 columns: two whole 16-column forward groups and a partial third one that
# spans two grid rows (9 x 4 columns), 8 forward / 2 adjoint depth tiles
TILE = C.scaled(C.get("cfg1"), nx=9, ny=4, nz=128, n_scat=300)


def plans(cfg, half_index, dev, interp=None, G=None, depth_run=None):
    G0, T, P, Op = H.case(cfg, half_index)
    G = G or G0
    if interp:
        Op = dict(Op, interpolation=interp)
    Gf = H.flat_geometry(G)
    kw = dict(n_sample=cfg.n_sample, device=dev)
    flat = RcaPlan.from_reference(Gf, T, P, Op, depth_run=0, **kw)
    tiled = RcaPlan.from_reference(Gf, T, P, Op, depth_run=depth_run or cfg.nz, **kw)
    assert flat.depth_run == 0 and tiled.depth_run == (depth_run or cfg.nz)
    return flat, tiled, G


@pytest.mark.parametrize("interp", ["quadratic", "lanczos", "lanczos3", "lanczos4"])
@pytest.mark.parametrize("half_index", [0, 1])
def test_tiled_matches_untiled_and_oracle(dev, interp, half_index):
    say(f"case tiled 9x4x128 interp={interp} half={half_index}, 3 frames, per tilt + compound")
    iq = H.iq_frame(TILE, half_index, n_frame=3)
    flat, tiled, G = plans(TILE, half_index, dev, interp)
    x = iq.to(dev)
    for compound in (False, True):
        o_f, m_f = flat.forward(x, compound=compound)
        o_t, m_t = tiled.forward(x, compound=compound)
        torch.cuda.synchronize()
        r = H.rel_l2(o_t.cpu().numpy(), o_f.cpu().numpy())
        say(f"tiled vs untiled compound={compound}: rel-L2 {r:.3e}")
        assert r <= (TOL if compound else 1e-6)
        assert H.rel_l2(m_t.cpu().numpy(), m_f.cpu().numpy()) <= TOL
    ref, rmask = oracle_fwd(TILE, half_index, iq, interp)
    check(o_t.cpu().numpy(), ref.sum(axis=1, keepdims=True), f"tiled {interp}/{half_index} vs oracle")
    assert H.rel_l2(m_t.cpu().numpy(), rmask.sum(axis=0, keepdims=True)) <= 1e-5


def test_tiled_launch_shapes(dev):
    say("case tiled V x FB shapes, 5 frames")
    iq = H.iq_frame(TILE, 0, n_frame=5).to(dev)
    flat, tiled, G = plans(TILE, 0, dev, "lanczos4")
    base = flat.forward(iq)[0].cpu().numpy()
    try:
        for v, fb, kib in ((1, 1, 0), (2, 1, 0), (4, 1, 0), (1, 2, 0), (2, 2, 0), (1, 4, 0),
                           (2, 4, 0), (1, 8, 0), (1, 1, 16), (2, 1, 16), (1, 1, 64), (1, 8, 64)):
            _lib.set_tuning(v, fb, 0, kib)     # kib: LDS budget, i.e. tilts per chunk
            o = tiled.forward(iq)[0].cpu().numpy()
            r = H.rel_l2(o, base)
            say(f"V={v} FB={fb} LDS={kib} KiB: rel-L2 vs untiled {r:.3e}")
            assert r <= TOL, (v, fb, kib)
    finally:
        _lib.set_tuning(0, 0, 0)


def test_hint_not_dividing_voxels(dev):
    say("case depth_run=100 on a 128-deep grid (partial last column, mixed tiles)")
    iq = H.iq_frame(TILE, 1, n_frame=2).to(dev)
    flat, tiled, G = plans(TILE, 1, dev, "lanczos3")
    tiled.desc.depth_run = 100     # the plan keeps only dividing hints; the C ABI takes any
    o_f, m_f = flat.forward(iq, compound=False)
    o_t, m_t = tiled.forward(iq, compound=False)
    torch.cuda.synchronize()
    assert H.rel_l2(o_t.cpu().numpy(), o_f.cpu().numpy()) <= 1e-6
    assert H.rel_l2(m_t.cpu().numpy(), m_f.cpu().numpy()) <= TOL


def test_windows_longer_than_rows(dev):
    # depths 0.6 mm apart: a 64-depth tile spans 38 mm, about 750 samples per
    # window, while depth-tile rows hold 256 samples (three passes)
    cfg = C.scaled(TILE, nx=4, ny=3)
    x, y, _ = C.grid_axes(cfg)
    z = 5e-3 + np.arange(cfg.nz) * 0.6e-3
    G0, T, P, Op = H.case(cfg, 0)
    G = C.geometry(cfg, cfg.halves[0], C.meshgrid_zxy(x, y, z))
    say("case depth-coarse grid (windows staged in several chunks), 8 frames, FB=8")
    iq = H.iq_frame(cfg, 0, n_frame=8)
    flat, tiled, _ = plans(cfg, 0, dev, "lanczos4", G=G)
    x8 = iq.to(dev)
    base = flat.forward(x8)[0].cpu().numpy()
    try:
        for v, fb in ((1, 8), (2, 4), (2, 1)):
            _lib.set_tuning(v, fb, 0)
            o = tiled.forward(x8)[0].cpu().numpy()
            r = H.rel_l2(o, base)
            say(f"V={v} FB={fb}: rel-L2 vs untiled {r:.3e}")
            assert r <= TOL, (v, fb)
    finally:
        _lib.set_tuning(0, 0, 0)
    ref, _ = oracle_fwd(cfg, 0, iq, "lanczos4", G=G)
    check(base, ref.sum(axis=1, keepdims=True), "depth-coarse vs oracle")


def test_tiled_fp16_and_nosum(dev):
    say("case tiled FP16-I/Q (4 frames) and NoSum")
    iq = H.iq_frame(TILE, 0, n_frame=4).to(dev)
    flat, tiled, G = plans(TILE, 0, dev, "lanczos4")
    h, s = quantize_fp16(iq)
    o_f = flat.forward(h, s)[0]
    o_t = tiled.forward(h, s)[0]
    torch.cuda.synchronize()
    r = H.rel_l2(o_t.cpu().numpy(), o_f.cpu().numpy())
    say(f"fp16 tiled vs untiled: rel-L2 {r:.3e}")
    assert r <= TOL
    n_f = flat.forward(iq[:1], compound=False, nosum=True, want_mask=False)[0]
    n_t = tiled.forward(iq[:1], compound=False, nosum=True, want_mask=False)[0]
    torch.cuda.synchronize()
    r = H.rel_l2(n_t.cpu().numpy(), n_f.cpu().numpy())
    say(f"nosum tiled vs untiled: rel-L2 {r:.3e}")
    assert r <= 1e-6


@pytest.mark.parametrize("nf", [6, 7])
@pytest.mark.parametrize("half_index", [0, 1])
def test_adjoint_frame_split(dev, nf, half_index):
    """Without a tuned shape the adjoint runs the frames in launches of
    4-, 2- and 1-frame blocks (6 = 4 + 2; 7 = 4 + 2 + 1, or one launch of
    4-frame blocks on the receive-x half with depth tiles): every frame of
    the split equals that frame's own one-frame adjoint, for the float-atomic
    and the fixed-point sums, and with accumulation."""
    say(f"case adjoint frame split nf={nf} half={half_index}")
    _, tiled, G = plans(ROWS, half_index, dev, "lanczos4")
    rng = np.random.default_rng(61 + nf + half_index)
    g = torch.from_numpy((rng.standard_normal((nf, 1, tiled.n_pix)) +
                          1j * rng.standard_normal((nf, 1, tiled.n_pix))).astype(np.complex64)).to(dev)
    one = torch.cat([tiled.adjoint(g[f:f + 1].contiguous()) for f in range(nf)])
    for det in (False, True):
        r = H.rel_l2_t(tiled.adjoint(g, deterministic=det), one)
        say(f"deterministic={det}: {nf} frames vs one frame at a time: rel-L2 {r:.3e}")
        assert r <= 1e-6
    base = torch.from_numpy((rng.standard_normal(tuple(one.shape)) +
                             1j * rng.standard_normal(tuple(one.shape))).astype(np.complex64)).to(dev)
    acc = base.clone()
    tiled.adjoint(g, grad_iq=acc, accumulate=True)
    r = H.rel_l2_t(acc, base + one)
    say(f"accumulated: rel-L2 vs base + one-frame adjoints {r:.3e}")
    assert r <= 1e-6


@pytest.mark.parametrize("interp", ["quadratic", "lanczos4"])
@pytest.mark.parametrize("half_index", [0, 1])
def test_tiled_adjoint(dev, interp, half_index):
    say(f"case tiled adjoint interp={interp} half={half_index}: vs untiled, oracle, frame blocks")
    flat, tiled, G = plans(TILE, half_index, dev, interp)
    rng = np.random.default_rng(11 + half_index)
    nf = 3
    g = torch.from_numpy((rng.standard_normal((nf, 1, flat.n_pix)) +
                          1j * rng.standard_normal((nf, 1, flat.n_pix))).astype(np.complex64)).to(dev)
    a_f = flat.adjoint(g).cpu().numpy()
    outs = {}
    try:
        for fb, kib in ((0, 0), (1, 0), (2, 0), (4, 0), (8, 0), (1, 16), (4, 64)):
            _lib.set_tuning(0, fb, 0, kib)
            outs[(fb, kib)] = tiled.adjoint(g).cpu().numpy()
    finally:
        _lib.set_tuning(0, 0, 0, 0)
    for k, o in outs.items():
        r = H.rel_l2(o, a_f)
        say(f"FB={k[0]} LDS={k[1]} KiB: rel-L2 vs untiled {r:.3e}")
        assert r <= TOL, k
    T = H.case(TILE, half_index)[1]
    prep = O.prepare(G, T)
    kind, a = OKIND[interp]
    ref = O.kernel_adjoint(g.cpu().numpy(), H.grid_f32(G), prep,
                           H.oracle_params(TILE, TILE.halves[half_index]), nf, kind, a)
    r = H.rel_l2(outs[(0, 0)], ref)
    say(f"tiled adjoint vs FP64 oracle: rel-L2 {r:.3e}")
    assert r <= 1e-5


# 72 columns in rows of 24: of the adjoint's 16-column tiles (4 voxels per
# lane) those along one grid row (columns 0-15, 32-47, 48-63 and the partial
# 64-71) share the transmit terms and go to the line-outer kernel on the
# receive-x half (half 0); the tile spanning two rows (16-31), and every tile
# of the receive-y half, to das_adj
ROWS = C.scaled(C.get("cfg1"), nx=24, ny=3, nz=128, n_scat=300)


@pytest.mark.parametrize("interp", ["quadratic", "lanczos4"])
@pytest.mark.parametrize("half_index", [0, 1])
def test_adjoint_tiles_along_rows(dev, interp, half_index):
    say(f"case adjoint, tiles along grid rows (line-outer + general kernels), interp={interp} "
        f"half={half_index}: vs untiled, oracle, frame blocks, per-tilt gradients")
    flat, tiled, G = plans(ROWS, half_index, dev, interp)
    rng = np.random.default_rng(31 + half_index)
    nf = 5   # a frame block of 4 and a partial one
    g = torch.from_numpy((rng.standard_normal((nf, 1, flat.n_pix)) +
                          1j * rng.standard_normal((nf, 1, flat.n_pix))).astype(np.complex64)).to(dev)
    a_f = flat.adjoint(g).cpu().numpy()
    outs = {}
    try:
        for v, fb in ((0, 0), (4, 1), (4, 2), (4, 4), (2, 4)):
            _lib.set_tuning(v, fb, 0, 0)
            outs[(v, fb)] = tiled.adjoint(g).cpu().numpy()
    finally:
        _lib.set_tuning(0, 0, 0, 0)
    for k, o in outs.items():
        r = H.rel_l2(o, a_f)
        say(f"V={k[0]} FB={k[1]}: rel-L2 vs untiled {r:.3e}")
        assert r <= TOL, k
    T = H.case(ROWS, half_index)[1]
    kind, a = OKIND[interp]
    ref = O.kernel_adjoint(g.cpu().numpy(), H.grid_f32(G), O.prepare(G, T),
                           H.oracle_params(ROWS, ROWS.halves[half_index]), nf, kind, a)
    r = H.rel_l2(outs[(0, 0)], ref)
    say(f"tiled adjoint vs FP64 oracle: rel-L2 {r:.3e}")
    assert r <= 1e-5
    gt = torch.from_numpy((rng.standard_normal((2, flat.n_tx, flat.n_pix)) +
                           1j * rng.standard_normal((2, flat.n_tx, flat.n_pix))).astype(np.complex64)).to(dev)
    r = H.rel_l2(tiled.adjoint(gt).cpu().numpy(), flat.adjoint(gt).cpu().numpy())
    say(f"per-tilt gradients, tiled vs untiled: rel-L2 {r:.3e}")
    assert r <= TOL


@pytest.mark.parametrize("kind", ["channel", "tx_channel"])
def test_adjoint_tiles_along_rows_laws(dev, kind):
    from tests.test_gpu_ablaw import laws, plans as law_plans
    say(f"case adjoint, tiles along grid rows, law {kind}: tiled vs untiled vs FP64 oracle adjoint")
    pkw, okw = laws(ROWS, kind, seed=23)
    flat, tiled, G, T = law_plans(ROWS, 0, dev, "lanczos4", pkw)
    rng = np.random.default_rng(29)
    nf = 4
    g = torch.from_numpy((rng.standard_normal((nf, 1, flat.n_pix)) +
                          1j * rng.standard_normal((nf, 1, flat.n_pix))).astype(np.complex64)).to(dev)
    a_f = flat.adjoint(g).cpu().numpy()
    a_t = tiled.adjoint(g).cpu().numpy()
    r = H.rel_l2(a_t, a_f)
    say(f"tiled vs untiled: rel-L2 {r:.3e}")
    assert r <= TOL
    kind_o, a = OKIND["lanczos4"]
    ref = O.kernel_adjoint(g.cpu().numpy(), H.grid_f32(G), O.prepare(G, T),
                           H.oracle_params(ROWS, ROWS.halves[0]), nf, kind_o, a, **okw)
    r = H.rel_l2(a_t, ref)
    say(f"tiled adjoint vs FP64 oracle: rel-L2 {r:.3e}")
    assert r <= 1e-5


def test_tiled_adjoint_per_tilt_and_accumulate(dev):
    say("case tiled adjoint of per-tilt outputs (grad_tx = nTx), accumulate")
    flat, tiled, G = plans(TILE, 0, dev, "lanczos3")
    rng = np.random.default_rng(21)
    g = torch.from_numpy((rng.standard_normal((2, flat.n_tx, flat.n_pix)) +
                          1j * rng.standard_normal((2, flat.n_tx, flat.n_pix))
                          ).astype(np.complex64)).to(dev)
    a_f = flat.adjoint(g)
    a_t = tiled.adjoint(g)
    c = a_t.clone()
    tiled.adjoint(g, grad_iq=c, accumulate=True)
    torch.cuda.synchronize()
    r = H.rel_l2(a_t.cpu().numpy(), a_f.cpu().numpy())
    say(f"per-tilt tiled vs untiled: rel-L2 {r:.3e}")
    assert r <= TOL
    assert H.rel_l2(c.cpu().numpy(), 2 * a_t.cpu().numpy()) <= 1e-6


def test_nosum_tilt_chunks_frames_mask_fp16_accumulate(dev):
    """NoSum takes tilt chunks and stores each update into a cleared output:
    5 frames (frame blocks of 4 + 1), the transmit mask on (zeros and mask 0
    for tilts that do not insonify a voxel), FP16-I/Q, accumulation."""
    say("case NoSum tiled: 5 frames, transmit mask, FP16, accumulate vs FP64 oracle")
    G0, T, P, Op = H.case(TILE, 1)
    # a 4 mm square aperture: its tilted cone leaves out the deep voxels of the
    # steep tilts (with the receive-axis span alone the cone is empty)
    G0 = dict(G0, apertureX=0.004, apertureY=0.004, apertureZ=0.0)
    Op = dict(Op, interpolation="lanczos4", flagMaskTransmit=True)
    kw = dict(n_sample=TILE.n_sample, device=dev)
    plan = RcaPlan.from_reference(H.flat_geometry(G0), T, P, Op, depth_run=TILE.nz, **kw)
    assert plan.depth_run == TILE.nz
    iq = H.iq_frame(TILE, 1, n_frame=5)
    x = iq.to(dev)
    out, mask = plan.forward(x, compound=False, nosum=True)
    torch.cuda.synchronize()
    ref, rmask = O.kernel_forward(iq.numpy().view(np.float32), H.grid_f32(G0), O.prepare(G0, T),
                                  H.oracle_params(TILE, TILE.halves[1]), O.INTERP_LANCZOS, 4, True,
                                  nosum=True)
    # the oracle leaves outputs of uninsonified (voxel, tilt) untouched (zeros here)
    r = H.rel_l2(out.cpu().numpy(), ref)
    say(f"NoSum 5 frames + mask: rel-L2 {r:.3e}")
    assert r <= 1e-5
    assert H.rel_l2(mask.cpu().numpy(), rmask) <= 1e-5
    m = mask.cpu().numpy()
    assert (m == 0).any() and (m > 0).any()          # the cone excludes some (voxel, tilt)
    acc = out.clone()
    plan.forward(x, compound=False, nosum=True, out=acc, want_mask=False, accumulate=True)
    torch.cuda.synchronize()
    assert H.rel_l2(acc.cpu().numpy(), 2 * out.cpu().numpy()) <= 1e-6
    h, s = quantize_fp16(x)
    o16, _ = plan.forward(h, s, compound=False, nosum=True, want_mask=False)
    torch.cuda.synchronize()
    r16 = H.rel_l2(o16.cpu().numpy(), ref)
    say(f"NoSum FP16: rel-L2 {r16:.3e}")
    assert r16 <= 3e-3
