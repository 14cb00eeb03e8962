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


def test_simulated_point_focuses_through_beamformer(dev):
    """The simulator's Raw feeds the beamformer drop-in as real data
    (IQ = RF(2n) - j RF(2n+1)): a point scatterer simulated on the GPU and
    beamformed on the GPU focuses at its voxel, and the GPU chain equals the
    oracle chain (simulator oracle -> beamformer oracle)."""
    from oracle import rca_oracle as O
    from simu_rca_ab_amd.beamformer import bfVirtualSourceGeneric_RCA
    say("case simulator -> beamformer: one point scatterer, 32+32 elements, 7 tilts +-15 deg")
    n_ele, n_tx, fc, c = 32, 7, 7.5e6, 1540.0
    M, TT, TR, TX, RX = rca_case(n_ele=n_ele, n_sub=8, n_tx=n_tx, n_scat=1, n_sample=512, fc=fc,
                                 attenuation=0.0)
    lam = c / fc
    xe = (np.arange(n_ele) - (n_ele - 1) / 2) * lam / 2
    p0 = np.array([1.5 * lam / 2, -2.0 * lam / 2, 30 * lam])
    pos = np.zeros((1, 4, n_tx), np.float32)
    pos[0, :3, :] = p0[:, None]
    pos[0, 3, :] = 1.0
    M = dict(M, scattererPos=pos)
    th = np.deg2rad(np.linspace(-15, 15, n_tx))
    for r in RX:
        r["Apod"] = np.ones(n_ele)
    for t, a in zip(TX, th):
        t["Apod"] = np.ones(n_ele)
        t["Delay"] = xe * np.sin(a) / c - (xe * np.sin(a) / c).min()
    raw = SimUS_CUDARCA(M, TT, TR, TX, RX, device=dev)
    raw_ref = S.simulate(M, TT, TR, TX, RX)
    assert H.rel_l2(raw, raw_ref) <= REL
    # time origin of the simulated RF as the beamformer sees it: the sample
    # index of a delay tau is fs (s tau + tg) - the wrapper's frequency grid
    # spacing fs/(2(Nf-1)) stretches time by s = Nf/(Nf-1), and the spectra
    # add the group delays of the pulse and transmit response and subtract
    # those of the conjugated receive response and input filter
    P = S.prepare(M, TT, TR, TX, RX)
    s = P["nf"] / (P["nf"] - 1)
    tg = (np.size(TX[0]["pulseShape"]) - 1) / 2 / 250e6 + 16 / 250e6 - 16 / 250e6 - 2 / P["fs"]
    t0 = (s - 1) * 2 * p0[2] / c - s * P["t_start"] + tg
    gx = p0[0] + (np.arange(9) - 4) * lam / 2
    gy = p0[1] + (np.arange(9) - 4) * lam / 2
    gz = p0[2] + (np.arange(13) - 6) * lam / 4
    Z, X, Y = np.meshgrid(gz, gx, gy, indexing="ij")
    G = {"gridBfX": X, "gridBfY": Y, "gridBfZ": Z, "posEleX": np.zeros(n_ele), "posEleY": xe,
         "posEleZ": np.zeros(n_ele)}
    T = {"thetaX": th, "thetaY": np.zeros(n_tx), "flagReceiveX": False}
    Par = {"samplingFrequency": P["fs"] / P["sub"], "demodulationFrequency": fc, "speedOfSound": c,
           "startTime": t0, "fnumber": 1.0}
    Op = {"interpolation": "lanczos4", "flagCompounding": True, "flagMaskTransmit": False}
    vol, _ = bfVirtualSourceGeneric_RCA(raw[:, :, None, :], G, T, Par, Op, device=dev)
    ref, _ = O.bf(raw_ref[:, :, None, :], G, T, Par, Op)
    r = H.rel_l2(vol, ref)
    say(f"GPU chain vs oracle chain: rel-L2 {r:.3e}")
    assert r <= 1e-5
    v = np.abs(vol[..., 0, 0])
    peak = np.unravel_index(np.argmax(v), v.shape)
    say(f"peak voxel {peak}, scatterer voxel (6, 4, 4), centre/max {v[6, 4, 4] / v.max():.3f}")
    assert all(abs(a - b) <= 1 for a, b in zip(peak, (6, 4, 4)))
    assert v[6, 4, 4] >= 0.9 * v.max()


def test_superposition_and_scaling_full_probe(dev):
    """Size-independent properties on the full probe of the dataset generator
    (128 + 128 elements x 128 sub-elements, 21 tilts, 640 frequencies): the
    output is additive over scatterer sets, and doubling every amplitude
    doubles it (a power-of-two scale commutes with FP32 rounding, so the
    difference is expected to be exactly zero)."""
    import bench
    say("case full probe 128+128 x 128 sub-elements, 21 tilts: Raw(A u B) = Raw(A) + Raw(B); Raw(2a) = 2 Raw(a)")
    M, TT, TR, TX, RX = bench.sim_case(n_scat=48)
    pos = M["scattererPos"]
    raw = SimUS_CUDARCA(M, TT, TR, TX, RX, device=dev)
    ra = SimUS_CUDARCA(dict(M, scattererPos=np.ascontiguousarray(pos[:20])), TT, TR, TX, RX, device=dev)
    rb = SimUS_CUDARCA(dict(M, scattererPos=np.ascontiguousarray(pos[20:])), TT, TR, TX, RX, device=dev)
    r = H.rel_l2(ra + rb, raw)
    say(f"superposition: rel-L2 {r:.3e}  shape {raw.shape}")
    assert np.isfinite(raw).all() and np.abs(raw).max() > 0
    assert r <= 1e-5
    p2 = pos.copy()
    p2[:, 3, :] *= 2
    raw2 = SimUS_CUDARCA(dict(M, scattererPos=p2), TT, TR, TX, RX, device=dev)
    d = np.abs(raw2 - 2 * raw).max()
    say(f"scaling: max |Raw(2a) - 2 Raw(a)| = {d:.3e} (bitwise equal: {d == 0})")
    assert d <= 1e-7 * np.abs(raw).max()
