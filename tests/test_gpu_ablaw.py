"""GPU parity of the receive aberration laws against the FP64 oracle.

Laws: per line (Hyp1_save/bfVirtualSourceGeneric_RCA_ab_1.cu:306-311), per
(tilt, line) (Hyp2/bfVirtualSourceGeneric_RCA_ab_2.cu:306-311) and per
(line, voxel) with receive and transmit sides
(exemple_paul/bfVirtualSourceGenericTxRx.cu:268-275, 292-294).  Each runs on
the untiled and the depth-tiled launch (where the per-voxel law widens the
sample window a workgroup stages), per tilt and compounded, FP16-I/Q, NoSum,
several launch shapes, the adjoint, and the drop-in entry point.  Bar as for
the other forward tests: relative L2 <= 1e-5 and envelope <= 0.1 dB vs the
FP64 oracle (3e-3 for FP16-I/Q); tiled vs untiled to FP32 summation order.
"""
import numpy as np
import pytest
import torch

from oracle import rca_oracle as O
from simu_rca_ab_amd import _lib, configs as C
from simu_rca_ab_amd.beamformer import bfVirtualSourceGeneric_RCA_TxRx
from simu_rca_ab_amd.plan import RcaPlan, quantize_fp16
from tests import helpers as H
from tests.test_gpu_forward import OKIND, check, say

pytestmark = pytest.mark.gpu

TOL = 4e-6
CFG = C.scaled(C.get("cfg1"), nx=5, ny=3, nz=128, n_scat=300)   # 15 columns, 2 depth tiles
MODE = {"channel": 1, "tx_channel": 2, "pixel_rx": 3, "pixel_tx": 3, "pixel_both": 3}
MODE_C = {"channel": 1, "tx_channel": 2, "pixel_channel": 3}


def laws(cfg, kind, seed=5):
    """(plan kwargs, oracle kwargs): delays of a few samples, random per entry.
    Arrays are in MATLAB order ([nCh, nTx] / [nPix, nCh]); both sides flatten
    them column-major."""
    rng = np.random.default_rng(seed)
    n_ch, n_tx, n_pix = cfg.n_elem, len(cfg.halves[0].theta), cfg.nx * cfg.ny * cfg.nz
    span = 3.0 / cfg.fs
    if kind == "channel":
        rx, tx = rng.uniform(-span, span, n_ch), None
    elif kind == "tx_channel":
        rx, tx = rng.uniform(-span, span, (n_ch, n_tx)), None
    else:
        rx = rng.uniform(-span, span, (n_pix, n_ch)) if kind != "pixel_tx" else None
        tx = rng.uniform(-span, span, (n_pix, n_ch)) if kind != "pixel_rx" else None
    f32 = lambda a: None if a is None else np.asarray(a, np.float32)   # noqa: E731
    rx, tx = f32(rx), f32(tx)
    mode = {1: "channel", 2: "tx_channel", 3: "pixel_channel"}[MODE[kind]]
    return (dict(ab_law=rx, ab_law_tx=tx, ab_law_mode=mode),
            dict(ablaw=rx, ablaw_tx=tx, ablaw_mode=MODE[kind]))


def plans(cfg, half_index, dev, interp, pkw):
    G, T, P, Op = H.case(cfg, half_index)
    Op = dict(Op, interpolation=interp)
    P = dict(P, AbLawRx=pkw["ab_law"], AbLawTx=pkw["ab_law_tx"], AbLawMode=pkw["ab_law_mode"])
    kw = dict(n_sample=cfg.n_sample, device=dev)
    flat = RcaPlan.from_reference(H.flat_geometry(G), T, P, Op, depth_run=0, **kw)
    tiled = RcaPlan.from_reference(H.flat_geometry(G), T, P, Op, depth_run=cfg.nz, **kw)
    assert flat.depth_run == 0 and tiled.depth_run == cfg.nz
    assert flat.desc.ab_law_mode == MODE_C[pkw["ab_law_mode"]]
    return flat, tiled, G, T


def oracle(cfg, half_index, G, T, iq, interp, okw, nosum=False):
    kind, a = OKIND[interp]
    return O.kernel_forward(iq.numpy().view(np.float32), H.grid_f32(G), O.prepare(G, T),
                            H.oracle_params(cfg, cfg.halves[half_index]), kind, a, False,
                            nosum=nosum, **okw)


@pytest.mark.parametrize("kind", list(MODE))
@pytest.mark.parametrize("half_index", [0, 1])
def test_law_tiled_untiled_oracle(dev, kind, half_index):
    say(f"case law {kind} half={half_index}: 2 frames, untiled + tiled, per tilt + compound")
    pkw, okw = laws(CFG, kind, seed=5 + half_index)
    flat, tiled, G, T = plans(CFG, half_index, dev, "lanczos4", pkw)
    iq = H.iq_frame(CFG, half_index, n_frame=2)
    x = iq.to(dev)
    ref, rmask = oracle(CFG, half_index, G, T, iq, "lanczos4", okw)
    for compound in (False, True):
        o_f, m_f = flat.forward(x, compound=compound)
        o_t, m_t = tiled.forward(x, compound=compound)
        torch.cuda.synchronize()
        r = H.rel_l2(o_t.cpu().numpy(), o_f.cpu().numpy())
        say(f"tiled vs untiled compound={compound}: rel-L2 {r:.3e}")
        assert r <= (TOL if compound else 1e-6)
        want = ref.sum(axis=1, keepdims=True) if compound else ref
        check(o_f.cpu().numpy(), want, f"untiled {kind} compound={compound}")
        check(o_t.cpu().numpy(), want, f"tiled {kind} compound={compound}")
        wm = rmask.sum(axis=0, keepdims=True) if compound else rmask
        assert H.rel_l2(m_t.cpu().numpy(), wm) <= 1e-5


@pytest.mark.parametrize("kind", ["tx_channel", "pixel_both"])
def test_law_launch_shapes_fp16_nosum(dev, kind):
    say(f"case law {kind}: V x FB x LDS shapes (5 frames), FP16-I/Q, NoSum")
    pkw, okw = laws(CFG, kind, seed=9)
    flat, tiled, G, T = plans(CFG, 0, dev, "lanczos3", pkw)
    iq = H.iq_frame(CFG, 0, n_frame=5)
    x = iq.to(dev)
    base = flat.forward(x)[0].cpu().numpy()
    try:
        for v, fb, kib in ((2, 1, 0), (1, 4, 0), (1, 8, 0), (1, 1, 16), (1, 8, 64)):
            _lib.set_tuning(v, fb, 0, kib)
            r = H.rel_l2(tiled.forward(x)[0].cpu().numpy(), base)
            say(f"V={v} FB={fb} LDS={kib} KiB: rel-L2 vs untiled {r:.3e}")
            assert r <= TOL, (v, fb, kib)
    finally:
        _lib.set_tuning(0, 0, 0, 0)
    ref, _ = oracle(CFG, 0, G, T, iq, "lanczos3", okw)
    check(base, ref.sum(axis=1, keepdims=True), f"{kind} 5 frames vs oracle")
    h, s = quantize_fp16(x)
    o16 = tiled.forward(h, s)[0].cpu().numpy()
    check(o16, ref.sum(axis=1, keepdims=True), f"{kind} FP16-I/Q vs oracle", rel=3e-3, env=0.1)
    n_t = tiled.forward(x[:1], compound=False, nosum=True, want_mask=False)[0]
    n_f = flat.forward(x[:1], compound=False, nosum=True, want_mask=False)[0]
    torch.cuda.synchronize()
    assert H.rel_l2(n_t.cpu().numpy(), n_f.cpu().numpy()) <= 1e-6
    refn, _ = oracle(CFG, 0, G, T, iq[:1], "lanczos3", okw, nosum=True)
    check(n_t.cpu().numpy(), refn, f"{kind} NoSum vs oracle")


@pytest.mark.parametrize("kind", ["channel", "tx_channel", "pixel_both"])
def test_law_adjoint(dev, kind):
    say(f"case law {kind}: adjoint tiled + untiled vs FP64 oracle adjoint")
    pkw, okw = laws(CFG, kind, seed=13)
    flat, tiled, G, T = plans(CFG, 1, dev, "lanczos4", pkw)
    rng = np.random.default_rng(17)
    nf = 2
    g = torch.from_numpy((rng.standard_normal((nf, 1, flat.n_pix)) +
                          1j * rng.standard_normal((nf, 1, flat.n_pix))).astype(np.complex64)).to(dev)
    a_f = flat.adjoint(g).cpu().numpy()
    a_t = tiled.adjoint(g).cpu().numpy()
    kind_o, a = OKIND["lanczos4"]
    ref = O.kernel_adjoint(g.cpu().numpy(), H.grid_f32(G), O.prepare(G, T),
                           H.oracle_params(CFG, CFG.halves[1]), nf, kind_o, a, **okw)
    for name, o in (("untiled", a_f), ("tiled", a_t)):
        r = H.rel_l2(o, ref)
        say(f"{name} adjoint vs FP64 oracle: rel-L2 {r:.3e}")
        assert r <= 1e-5


def test_txrx_dropin(dev):
    cfg = C.scaled(CFG, nx=4, ny=3, nz=64)
    say("case drop-in bfVirtualSourceGeneric_RCA_TxRx (laws [s1,s2,s3,nCh]), numpy in/out")
    G, T, P, Op = H.case(cfg, 0)
    Op = dict(Op, interpolation="lanczos4", flagCompounding=False)
    pkw, okw = laws(cfg, "pixel_both", seed=21)
    shape4 = (cfg.nz, cfg.nx, cfg.ny, cfg.n_elem)
    Prm = dict(P, AbLawRx=pkw["ab_law"].reshape(shape4, order="F"),
               AbLawTx=pkw["ab_law_tx"].reshape(shape4, order="F"))
    iq = H.iq_frame(cfg, 0)
    from simu_rca_ab_amd import synth
    data = synth.to_matlab_layout(iq.numpy())
    dataBf, mask = bfVirtualSourceGeneric_RCA_TxRx(data, G, T, Prm, Op, device=dev)
    ref, _ = oracle(cfg, 0, G, T, iq, "lanczos4", okw)
    got = dataBf[..., 0].reshape(-1, dataBf.shape[3], order="F").T   # [nTx, nPix]
    check(got, ref[0], "TxRx drop-in vs oracle")
    # empty laws: plain beamforming
    plain, _ = bfVirtualSourceGeneric_RCA_TxRx(data, G, T, dict(P, AbLawRx=[], AbLawTx=[]), Op,
                                               device=dev)
    ref0, _ = oracle(cfg, 0, G, T, iq, "lanczos4", {})
    check(plain[..., 0].reshape(-1, plain.shape[3], order="F").T, ref0[0], "TxRx empty laws")
