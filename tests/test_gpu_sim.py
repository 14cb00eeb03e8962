"""GPU parity of the SIMUS-RCA simulator kernels against the FP64 oracle.

Bar: relative L2 <= 1e-6 (FP32 fields and sums; every phase formed in FP64
and range-reduced) for the transmit field, the receive field and the
time-domain output, on small seeded cases covering the flat and focused
elevation terms, the baffle, the lens, attenuation, steered sub-elements,
every sample mode, several frames and the receive-aberration variant.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import simus_oracle as S
from simu_rca_ab_amd import _lib
from simu_rca_ab_amd.simulator import RcaSimulator, SimUS_CUDARCA, SimUS_CUDARCA_ab
from tests import helpers as H
from tests.simus_cases import rca_case

pytestmark = pytest.mark.gpu

REL = 1e-6


def say(*a):
    print(*a, flush=True)


def _steer(T, rng):
    """Give the sub-elements small azimuth/elevation angles."""
    T = dict(T)
    e = np.array(T["ElementPos"], np.float32)
    e[:, 3, :] = rng.uniform(-0.2, 0.2, e[:, 3, :].shape)
    e[:, 4, :] = rng.uniform(-0.1, 0.1, e[:, 4, :].shape)
    T["ElementPos"] = e
    return T


VARIANTS = {
    "flat": dict(),
    "focused": dict(rfocus=0.02),
    "baffle_lens": dict(baffle=1, lens=3.1e-7),
    "no_attenuation": dict(attenuation=0.0),
    "steered": dict(rfocus=0.025),
}


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_fields_vs_oracle(dev, variant):
    say(f"case sim fields {variant}: txfield and rxfield (reference ABI) vs FP64 oracle")
    M, TT, TR, TX, RX = rca_case(n_ele=12, n_sub=3, n_tx=3, n_scat=7, **VARIANTS[variant])
    if variant == "steered":
        rng = np.random.default_rng(5)
        TT, TR = _steer(TT, rng), _steer(TR, rng)
    P = S.prepare(M, TT, TR, TX, RX)
    pos = M["scattererPos"][:, :, :3]
    c = float(np.float32(M["speedOfSound"]))
    ref = S.txfield(pos[:, :3, :], TT["ElementPos"], P["k"], P["pulse"], P["ir_tx"], P["delay"],
                    P["apod"], *S.transducer_scalars(TT), c)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)   # noqa: E731
    n_scat, nf, n_tx = pos.shape[0], P["nf"], 3
    n_ele, n_sub = TT["ElementPos"].shape[0], TT["ElementPos"].shape[2]
    out = torch.empty((n_tx, n_scat, nf), dtype=torch.complex64, device=dev)
    r_scat = d(pos[:, :3, :].transpose(2, 1, 0))
    e_tx = d(np.asarray(TT["ElementPos"], np.float32).transpose(2, 1, 0))
    k, pulse, ir = d(P["k"]), d(P["pulse"]), d(P["ir_tx"])
    delay, apod = d(P["delay"].T), d(P["apod"].T)
    vp = ctypes.c_void_p
    lib = _lib.load()
    _lib.check(lib.rca_sim_txfield(vp(out.data_ptr()), vp(r_scat.data_ptr()), vp(e_tx.data_ptr()),
                                   vp(k.data_ptr()), vp(pulse.data_ptr()), vp(ir.data_ptr()),
                                   vp(delay.data_ptr()), vp(apod.data_ptr()),
                                   *S.transducer_scalars(TT), c, n_scat, n_ele, nf, n_tx, n_sub, None,
                                   vp(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    got = out.cpu().numpy().transpose(2, 1, 0)
    r = H.rel_l2(got, ref)
    say(f"txfield {variant}: rel-L2 {r:.3e}")
    assert r <= REL
    # receive field of tilt 1's scatterers
    ref_rx = S.rxfield(pos[:, :3, 1], TR["ElementPos"], P["k"], P["ir_rx"], *S.transducer_scalars(TR), c)
    rx = torch.empty((n_scat, TR["ElementPos"].shape[0], nf), dtype=torch.complex64, device=dev)
    rs1 = d(pos[:, :3, 1].T)
    e_rx = d(np.asarray(TR["ElementPos"], np.float32).transpose(2, 1, 0))
    ir_rx = d(P["ir_rx"])
    _lib.check(lib.rca_sim_rxfield(vp(rx.data_ptr()), vp(rs1.data_ptr()), vp(e_rx.data_ptr()),
                                   vp(k.data_ptr()), vp(ir_rx.data_ptr()), *S.transducer_scalars(TR), c,
                                   n_scat, TR["ElementPos"].shape[0], nf, n_sub,
                                   vp(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    got_rx = rx.cpu().numpy().transpose(2, 1, 0)
    r = H.rel_l2(got_rx, ref_rx)
    say(f"rxfield {variant}: rel-L2 {r:.3e}")
    assert r <= REL


@pytest.mark.parametrize("mode,variant", [("BS100BW", "flat"), ("NS200BW", "focused"),
                                          ("BS50BW", "baffle_lens")])
def test_simulate_vs_oracle(dev, mode, variant):
    say(f"case SimUS_CUDARCA {mode} {variant}: 2 frames x 3 tilts, drop-in vs FP64 oracle")
    M, TT, TR, TX, RX = rca_case(n_ele=10, n_sub=2, n_tx=3, n_scat=6, n_frame=2, n_sample=160,
                                 mode=mode, **VARIANTS[variant])
    raw = SimUS_CUDARCA(M, TT, TR, TX, RX, {"decreaseFactorGPU": 1}, device=dev)
    ref = S.simulate(M, TT, TR, TX, RX)
    assert raw.shape == ref.shape and raw.dtype == np.float32
    r = H.rel_l2(raw, ref)
    say(f"Raw {mode} {variant}: rel-L2 {r:.3e}  shape {raw.shape}")
    assert r <= REL
    assert np.abs(ref).max() > 0


def test_simulate_ab_and_nan_scrub(dev):
    say("case SimUS_CUDARCA_ab (per-element receive law) + a scatterer on a transmit sub-element")
    M, TT, TR, TX, RX = rca_case(n_ele=8, n_sub=2, n_tx=2, n_scat=5, n_sample=128)
    pos = M["scattererPos"].copy()
    pos[0, :3, 0] = TT["ElementPos"][3, :3, 1]      # r = 0: the transmit field is NaN -> 0
    M = dict(M, scattererPos=pos)
    rng = np.random.default_rng(8)
    law = {"Delay": rng.uniform(-60e-9, 60e-9, 8), "Apod": rng.uniform(0.5, 1.0, 8)}
    raw, raw_ab = SimUS_CUDARCA_ab(M, TT, TR, TX, RX, law, device=dev)
    ref, ref_ab = S.simulate(M, TT, TR, TX, RX, AbLaw=law)
    assert np.isfinite(raw).all() and np.isfinite(ref).all()
    for name, g, w in (("Raw", raw, ref), ("Raw_ab", raw_ab, ref_ab)):
        r = H.rel_l2(g, w)
        say(f"{name}: rel-L2 {r:.3e}")
        assert r <= REL
    assert H.rel_l2(raw_ab, raw) > 1e-2   # the law acts


def test_frame_api_matches_dropin(dev):
    say("case RcaSimulator.frame on device tensors == drop-in")
    M, TT, TR, TX, RX = rca_case(n_ele=6, n_sub=2, n_tx=2, n_scat=4, n_frame=2)
    sim = RcaSimulator(M, TT, TR, TX, RX, device=dev)
    raw, _ = sim.run(M["scattererPos"])
    torch.cuda.synchronize()
    want = SimUS_CUDARCA(M, TT, TR, TX, RX, device=dev)
    np.testing.assert_array_equal(raw.cpu().numpy().transpose(2, 1, 0), want)
