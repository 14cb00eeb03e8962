"""GPU parity of the forward HIP kernels against the FP64 oracle.

Bar (FP32 I/Q, FP32 accumulation): relative L2 <= 1e-5 and envelope error
<= 0.1 dB (voxels within -40 dB of the maximum) against the FP64 oracle on the
same seeded inputs; masks to 1e-5 relative.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import rca_oracle as O
from simu_rca_ab_amd import _lib, configs as C, synth
from simu_rca_ab_amd.beamformer import bfVirtualSourceGeneric_RCA
from simu_rca_ab_amd.plan import RcaPlan
from tests import helpers as H

pytestmark = pytest.mark.gpu

REL = 1e-5
ENV_DB = 0.1
OKIND = {"quadratic": (O.INTERP_QUADRATIC, 3), "lanczos": (O.INTERP_LANCZOS_REF, 3),
         "lanczos3": (O.INTERP_LANCZOS, 3), "lanczos4": (O.INTERP_LANCZOS, 4)}


def say(*a):
    print(*a, flush=True)


def plan_for(cfg, half_index, dev, interp=None, mask_tx=False, G=None, depth_run=0, **pkw):
    """Plan of one half on a flattened grid: untiled (depth_run 0) unless a
    depth hint is given (the drop-in and the bench pass Nz)."""
    G0, T, P, Op = H.case(cfg, half_index)
    G = G or G0
    Op = dict(Op, flagMaskTransmit=mask_tx)
    if interp:
        Op["interpolation"] = interp
    P = dict(P, **pkw)
    return RcaPlan.from_reference(H.flat_geometry(G), T, P, Op, n_sample=cfg.n_sample, device=dev,
                                  depth_run=depth_run), G, T


def oracle_fwd(cfg, half_index, iq, interp, mask_tx=False, G=None, ablaw=None, ablaw_mode=0):
    G0, T, P, Op = H.case(cfg, half_index)
    G = G or G0
    prep = O.prepare(G, T)
    kind, a = OKIND[interp]
    return O.kernel_forward(iq.numpy().view(np.float32), H.grid_f32(G), prep,
                            H.oracle_params(cfg, cfg.halves[half_index]), kind, a, mask_tx,
                            ablaw=ablaw, ablaw_mode=ablaw_mode)


def check(got, want, what, rel=REL, env=ENV_DB):
    r = H.rel_l2(got, want)
    e = H.envelope_db_err(got, want)
    say(f"{what}: rel-L2 {r:.3e}  envelope {e:.4f} dB")
    assert r <= rel, what
    assert e <= env, what


def test_cfg0_full_size_vs_oracle(dev):
    cfg = C.get("cfg0")
    say("case cfg0 full 64x64x128, 11 tilts, 64 ch, 1024 samples, lanczos3")
    iq = H.iq_frame(cfg)
    plan, G, T = plan_for(cfg, 0, dev)
    out, mask = plan.forward(iq.to(dev))
    torch.cuda.synchronize()
    ref, rmask = oracle_fwd(cfg, 0, iq, "lanczos3")
    check(out[0, 0].cpu().numpy(), ref[0].sum(axis=0), "cfg0 full")
    assert H.rel_l2(mask[0].cpu().numpy(), rmask.sum(axis=0)) <= REL


SMALL = C.scaled(C.get("cfg1"), nx=6, ny=5, nz=40, n_scat=300)


@pytest.mark.parametrize("interp", ["quadratic", "lanczos", "lanczos3", "lanczos4"])
@pytest.mark.parametrize("half_index", [0, 1])
@pytest.mark.parametrize("compound", [True, False])
def test_matrix_small(dev, interp, half_index, compound):
    say(f"case small cfg1-geometry interp={interp} half={half_index} compound={compound}")
    iq = H.iq_frame(SMALL, half_index, n_frame=2)
    plan, G, T = plan_for(SMALL, half_index, dev, interp)
    out, mask = plan.forward(iq.to(dev), compound=compound)
    torch.cuda.synchronize()
    ref, rmask = oracle_fwd(SMALL, half_index, iq, interp)
    if compound:
        ref, rmask = ref.sum(axis=1, keepdims=True), rmask.sum(axis=0, keepdims=True)
    check(out.cpu().numpy(), ref, f"{interp}/{half_index}/{compound}")
    assert H.rel_l2(mask.cpu().numpy(), rmask) <= REL


def test_launch_shapes_agree(dev):
    cfg = C.scaled(SMALL, nx=7, ny=3, nz=37)      # nPix = 777, not a multiple of 256
    say("case launch shapes V x FB, 5 frames, nPix=777")
    iq = H.iq_frame(cfg, 0, n_frame=5).to(dev)
    plan, G, T = plan_for(cfg, 0, dev, "lanczos4")
    outs = {}
    try:
        for v in (1, 2, 4):
            for fb in (1, 2, 4):
                _lib.set_tuning(v, fb, 0)
                outs[(v, fb)] = plan.forward(iq)[0].cpu().numpy()
    finally:
        _lib.set_tuning(0, 0, 0)
    base = outs[(1, 1)]
    for k, o in outs.items():
        assert H.rel_l2(o, base) <= 1e-6, k
    ref, _ = oracle_fwd(cfg, 0, iq.cpu(), "lanczos4")
    check(base, ref.sum(axis=1, keepdims=True), "shapes")


def test_mask_transmit(dev):
    cfg = C.scaled(SMALL, nx=12, ny=10, nz=24)
    G, T, P, Op = H.case(cfg, 0)
    G = dict(G, apertureX=0.0254, apertureY=0.0254, apertureZ=0.0)
    say("case flagMaskTransmit=True with a square aperture")
    iq = H.iq_frame(cfg, 0)
    plan, _, _ = plan_for(cfg, 0, dev, "lanczos3", mask_tx=True, G=G)
    out, mask = plan.forward(iq.to(dev), compound=False)
    torch.cuda.synchronize()
    ref, rmask = oracle_fwd(cfg, 0, iq, "lanczos3", mask_tx=True, G=G)
    masked = (rmask == 0).mean()
    say(f"fraction of (voxel, tilt) outside the cone: {masked:.3f}")
    check(out.cpu().numpy(), ref, "mask_tx")
    np.testing.assert_array_equal(mask.cpu().numpy() == 0, rmask == 0)


def test_accumulate_two_halves(dev):
    cfg = C.scaled(SMALL, nx=8, ny=8, nz=16)
    say("case RCA frame = half X + half Y accumulated in place")
    iq0, iq1 = H.iq_frame(cfg, 0).to(dev), H.iq_frame(cfg, 1).to(dev)
    p0, _, _ = plan_for(cfg, 0, dev)
    p1, _, _ = plan_for(cfg, 1, dev)
    a, ma = p0.forward(iq0)
    b, mb = p1.forward(iq1)
    s = a.clone()
    ms = ma.clone()
    p1.forward(iq1, out=s, mask=ms, accumulate=True)
    torch.cuda.synchronize()
    np.testing.assert_allclose(s.cpu().numpy(), (a + b).cpu().numpy(), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(ms.cpu().numpy(), (ma + mb).cpu().numpy(), rtol=1e-6)


def test_cfg1_full_size_sampled_voxels(dev):
    cfg = C.get("cfg1")
    say("case cfg1 full 128x128x256, 2 x 16 tilts, 128 ch, 2048 samples, lanczos4, noise")
    iqs = synth.frame(cfg)
    out = None
    for hi in range(2):
        plan, G, T = plan_for(cfg, hi, dev)
        if out is None:
            out, mask = plan.forward(iqs[hi].to(dev))
        else:
            plan.forward(iqs[hi].to(dev), out=out, mask=mask, accumulate=True)
        del plan
    torch.cuda.synchronize()
    got_all = out[0, 0].cpu().numpy()
    assert np.isfinite(got_all).all()
    rng = np.random.default_rng(123)
    sel = np.sort(rng.choice(cfg.n_vox, 3000, replace=False))
    ref = 0
    for hi in range(2):
        G, T, P, Op = H.case(cfg, hi)
        Gs = dict(G)
        for k in ("gridBfX", "gridBfY", "gridBfZ"):
            Gs[k] = np.asarray(G[k]).ravel(order="F")[sel]
        prep = O.prepare(G, T)
        r, _ = O.kernel_forward(iqs[hi].numpy().view(np.float32), H.grid_f32(Gs), prep,
                                H.oracle_params(cfg, cfg.halves[hi]), O.INTERP_LANCZOS, 4)
        ref = ref + r[0].sum(axis=0)
    check(got_all[sel], ref, "cfg1 full (3000 sampled voxels)")


def test_cfg1_full_size_linearity_and_tilt_sum(dev):
    """Size-independent properties on the full config-1 half (8.6e9 updates
    per launch): the output is linear in the I/Q, and the compounded volume
    and mask equal the sums of the per-tilt ones."""
    cfg = C.get("cfg1")
    say("case cfg1 full half 0, depth-tiled launch: linearity A(2x - 0.5j y) and compound == "
        "sum of per-tilt outputs")
    x = synth.frame(cfg)[0].to(dev)
    g = torch.Generator(device="cpu").manual_seed(7)
    y = torch.complex(torch.randn(x.shape, generator=g), torch.randn(x.shape, generator=g)).to(dev)
    plan, _, _ = plan_for(cfg, 0, dev, depth_run=cfg.nz)
    ax, mask = plan.forward(x)
    ay, _ = plan.forward(y)
    az, _ = plan.forward((2.0 * x - 0.5j * y).contiguous())
    r = H.rel_l2(az.cpu().numpy(), (2.0 * ax - 0.5j * ay).cpu().numpy())
    say(f"linearity: rel-L2 {r:.3e}")
    assert r <= 1e-5
    per, pmask = plan.forward(x, compound=False)
    assert per.shape[1] == len(cfg.halves[0].theta)
    r = H.rel_l2(per.sum(dim=1).cpu().numpy(), ax[:, 0].cpu().numpy())
    rm = H.rel_l2(pmask.sum(dim=0).cpu().numpy(), mask[0].cpu().numpy())
    say(f"tilt sum: rel-L2 {r:.3e}, mask {rm:.3e}")
    assert r <= REL and rm <= REL


def test_golden_fixture(dev):
    from tests.golden import make_golden as MG
    say("case golden fixture rca_cfg0_tiny")
    ref = np.load(MG.PATH, allow_pickle=False)
    iq, G, T, P, Op = MG.inputs()
    plan = RcaPlan.from_reference(H.flat_geometry(G), T, P, Op, n_sample=MG.CFG.n_sample, device=dev)
    out, mask = plan.forward(torch.from_numpy(iq).to(dev))
    torch.cuda.synchronize()
    check(out[0, 0].cpu().numpy(), ref["dataBf_re"] + 1j * ref["dataBf_im"], "golden")
    assert H.rel_l2(mask[0].cpu().numpy(), ref["mask"]) <= REL


@pytest.mark.parametrize("mode", ["channel", "tx_channel"])
def test_aberration_law(dev, mode):
    cfg = C.scaled(SMALL, nx=6, ny=6, nz=24)
    say(f"case receive aberration law per {mode}")
    n = cfg.n_elem if mode == "channel" else 16 * cfg.n_elem
    law = (np.random.default_rng(3).standard_normal(n) * 20e-9).astype(np.float32)
    iq = H.iq_frame(cfg, 0)
    plan, G, T = plan_for(cfg, 0, dev, AbLawRx=law, AbLawMode=mode)
    out, _ = plan.forward(iq.to(dev), compound=False)
    torch.cuda.synchronize()
    ref, _ = oracle_fwd(cfg, 0, iq, "lanczos4", ablaw=law, ablaw_mode=1 if mode == "channel" else 2)
    check(out.cpu().numpy(), ref, f"ablaw {mode}")


@pytest.mark.parametrize("lanczos", [False, True])
def test_reference_kernel_abi(dev, lanczos):
    """bfVirtualSourceGeneric_RCA C entry point with the reference argument list."""
    cfg = C.scaled(C.get("cfg0"), nx=10, ny=6, nz=20)
    say(f"case reference ABI flagInterpolationLanczos={lanczos}")
    iq = H.iq_frame(cfg, 0, n_frame=2)
    G, T, P, Op = H.case(cfg, 0)
    prep = O.prepare(G, T)
    grid = H.grid_f32(G)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)
    dg = [t(g) for g in grid]
    de = [t(prep[k]) for k in ("ex", "ey", "ez")]
    dv = [t(prep[k]) for k in ("vx", "vy", "vz", "d0")]
    npix, ntx, nch, nf = grid[0].size, 11, 64, 2
    dataBf = torch.zeros(2 * npix * ntx * nf, dtype=torch.float32, device=dev)
    mask = torch.zeros(npix * ntx, dtype=torch.float32, device=dev)
    data = iq.to(dev).contiguous()
    lib = _lib.load()
    ap = [float(a) for a in prep["ap"]]
    vp = lambda x: ctypes.c_void_p(x.data_ptr())
    rc = lib.bfVirtualSourceGeneric_RCA(
        vp(dataBf), vp(mask), vp(data), *map(vp, dg), *map(vp, de), *map(vp, dv), 1, *ap,
        float(cfg.fs), float(cfg.fd), float(cfg.c), float(cfg.t0), float(cfg.fnumber),
        npix, nch, ntx, nf, cfg.n_sample, int(lanczos), 0,
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    _lib.check(rc)
    torch.cuda.synchronize()
    kind = O.INTERP_LANCZOS_REF if lanczos else O.INTERP_QUADRATIC
    ref, rmask = O.kernel_forward(iq.numpy().view(np.float32), grid, prep,
                                  H.oracle_params(cfg, cfg.halves[0]), kind, 3)
    got = dataBf.cpu().numpy().view(np.complex64).reshape(nf, ntx, npix)
    check(got, ref, "reference ABI")
    assert H.rel_l2(mask.cpu().numpy().reshape(ntx, npix), rmask) <= REL


def test_dropin_numpy_torch_real(dev):
    cfg = C.scaled(C.get("cfg0"), nx=9, ny=7, nz=20)
    say("case drop-in bfVirtualSourceGeneric_RCA: numpy complex / numpy real / torch GPU")
    iq = H.iq_frame(cfg, 0, n_frame=2)
    G, T, P, Op = H.case(cfg, 0)
    data = synth.to_matlab_layout(iq.numpy())
    ref, rmask = O.bf(data, G, T, P, Op)
    got, gmask = bfVirtualSourceGeneric_RCA(data, G, T, P, Op)
    assert got.shape == ref.shape == (20, 9, 7, 1, 2) and gmask.shape == rmask.shape
    check(got, ref, "drop-in numpy complex")
    assert H.rel_l2(gmask, rmask) <= REL
    rf = np.empty((2 * cfg.n_sample,) + data.shape[1:], np.float32)
    rf[0::2], rf[1::2] = data.real, -data.imag
    got2, _ = bfVirtualSourceGeneric_RCA(rf, G, T, P, Op)
    np.testing.assert_allclose(got2, got, rtol=1e-6, atol=1e-6)
    Gt = {k: (torch.from_numpy(np.asarray(v)).to(dev) if k.startswith("grid") else v)
          for k, v in G.items()}
    got3, gm3 = bfVirtualSourceGeneric_RCA(synth.to_matlab_layout(iq.to(dev)), Gt, T, P, Op)
    assert got3.is_cuda and tuple(got3.shape) == got.shape
    np.testing.assert_allclose(got3.cpu().numpy(), got, rtol=1e-6, atol=1e-6)
    Opn = dict(Op, flagCompounding=False)
    got4, gm4 = bfVirtualSourceGeneric_RCA(data, G, T, P, Opn)
    ref4, rm4 = O.bf(data, G, T, P, Opn)
    assert got4.shape == ref4.shape == (20, 9, 7, 11, 2)
    check(got4, ref4, "drop-in no compounding")


@pytest.mark.parametrize("interp", ["quadratic", "lanczos4"])
def test_nosum_per_line_outputs(dev, interp):
    cfg = C.scaled(SMALL, nx=5, ny=4, nz=16)
    say(f"case NoSum per-line outputs interp={interp}")
    iq = H.iq_frame(cfg, 0, n_frame=2)
    plan, G, T = plan_for(cfg, 0, dev, interp)
    out, mask = plan.forward(iq.to(dev), compound=False, nosum=True)
    torch.cuda.synchronize()
    G0, T0, P0, Op0 = H.case(cfg, 0)
    kind, a = OKIND[interp]
    ref, rmask = O.kernel_forward(iq.numpy().view(np.float32), H.grid_f32(G0), O.prepare(G0, T0),
                                  H.oracle_params(cfg, cfg.halves[0]), kind, a, False, nosum=True)
    assert tuple(out.shape) == ref.shape                     # [nCh, F, T, P]
    r = H.rel_l2(out.cpu().numpy(), ref)
    say(f"nosum rel-L2 {r:.3e}")
    assert r <= REL
    assert H.rel_l2(mask.cpu().numpy(), rmask) <= REL
    # the per-line outputs sum to the per-tilt output
    full, _ = plan.forward(iq.to(dev), compound=False)
    assert H.rel_l2(out.sum(dim=0).cpu().numpy(), full.cpu().numpy()) <= 1e-5


def test_dropin_variants(dev):
    from simu_rca_ab_amd.beamformer import (bfVirtualSourceGeneric_RCA_Nosum,
                                            bfVirtualSourceGeneric_RCA_Nosum_ab,
                                            bfVirtualSourceGeneric_RCA_ab_1,
                                            bfVirtualSourceGeneric_RCA_ab_2)
    cfg = C.scaled(C.get("cfg0"), nx=4, ny=3, nz=10)
    say("case drop-in variants NoSum / ab_1 / ab_2 / Nosum_ab")
    iq = H.iq_frame(cfg, 0, n_frame=2)
    G, T, P, Op = H.case(cfg, 0)
    data = synth.to_matlab_layout(iq.numpy())
    grid = H.grid_f32(G)
    prep = O.prepare(G, T)
    prm = H.oracle_params(cfg, cfg.halves[0])
    d, m = bfVirtualSourceGeneric_RCA_Nosum(data, G, T, P, Op)
    assert d.shape == (10, 4, 3, 11, 2, 64) and m.shape == (10, 4, 3, 11)
    ref, _ = O.kernel_forward(iq.numpy().view(np.float32), grid, prep, prm, O.INTERP_LANCZOS, 3,
                              nosum=True)
    # [P, T, F, C] (column-major voxel flatten) -> [C, F, T, P]
    check(np.moveaxis(d.reshape(-1, 11, 2, 64, order="F"), [3, 2, 1, 0], [0, 1, 2, 3]), ref,
          "Nosum drop-in")
    law1 = (np.random.default_rng(0).standard_normal(64) * 15e-9)
    law2 = (np.random.default_rng(1).standard_normal((64, 11)) * 15e-9)   # AbLawRx(:)[tx*64+ch]
    for fn, law, mode, nos in ((bfVirtualSourceGeneric_RCA_ab_1, law1, 1, False),
                               (bfVirtualSourceGeneric_RCA_ab_2, law2, 2, False),
                               (bfVirtualSourceGeneric_RCA_Nosum_ab, law1, 1, True)):
        d, _ = fn(data, G, T, dict(P, AbLawRx=law), Op)
        ref, _ = O.kernel_forward(iq.numpy().view(np.float32), grid, prep, prm, O.INTERP_LANCZOS, 3,
                                  nosum=nos, ablaw=np.asarray(law, np.float32).ravel(order="F"),
                                  ablaw_mode=mode)
        if nos:
            got = np.moveaxis(d.reshape(-1, 11, 2, 64, order="F"), [3, 2, 1, 0], [0, 1, 2, 3])
        else:
            got = d[..., 0, :].reshape(-1, 2, order="F").T
            ref = ref.sum(axis=1)
        check(got, ref, f"{fn.__name__}")


def test_plan_from_host_arrays(dev):
    """rca_plan_create: host FP64 arrays in, the wrapper's prep in C, one upload;
    rca_forward on its descriptor equals the Python plan and the oracle."""
    from tests.test_capi import _plan_params
    cfg = C.scaled(SMALL, nx=6, ny=6, nz=64)        # Nz >= 64: the depth hint applies
    say("case rca_plan_create / rca_forward(rca_plan_desc) / rca_plan_destroy")
    G, T, P, Op = H.case(cfg, 1)
    Gf = H.flat_geometry(G)
    keep = []
    p = _plan_params(Gf, T, cfg, keep)
    p.interp = _lib.INTERP["lanczos4"]
    p.depth_run = cfg.nz
    lib = _lib.load()
    handle = ctypes.c_void_p()
    _lib.check(lib.rca_plan_create(ctypes.byref(p), ctypes.byref(handle)))
    try:
        desc = lib.rca_plan_desc(handle)
        assert desc.contents.depth_run == cfg.nz and desc.contents.n_tx == 16
        iq = H.iq_frame(cfg, 1, n_frame=2).to(dev)
        out = torch.empty((2, 1, desc.contents.n_pix), dtype=torch.complex64, device=dev)
        _lib.check(lib.rca_forward(desc, ctypes.c_void_p(iq.data_ptr()), _lib.IQ_C64, None, 2,
                                   ctypes.c_void_p(out.data_ptr()), None, _lib.FLAG_COMPOUND,
                                   ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        torch.cuda.synchronize()
        plan, _, _ = plan_for(cfg, 1, dev, "lanczos4", depth_run=cfg.nz)
        assert plan.depth_run == cfg.nz
        want, _ = plan.forward(iq)
        torch.cuda.synchronize()
        r = H.rel_l2(out.cpu().numpy(), want.cpu().numpy())
        say(f"C plan vs Python plan (both depth-tiled): rel-L2 {r:.3e}")
        assert r <= 1e-6
        ref, _ = oracle_fwd(cfg, 1, iq.cpu(), "lanczos4")
        check(out.cpu().numpy(), ref.sum(axis=1, keepdims=True), "C plan vs oracle")
    finally:
        _lib.check(lib.rca_plan_destroy(handle))


def test_offsets_beyond_int32(dev):
    """260 frames of one config-1 half: 1.09e9 complex samples, so the
    reference's int32 float index 2 (c Ns + t Ns nCh + f Ns nCh nTx)
    (bfVirtualSourceGeneric_RCA.cu:315-317) would overflow in the last frames.
    Frames 0 and 259 carry the same synthetic I/Q, the others zeros: both
    volumes equal the single-frame result and an empty frame stays zero."""
    cfg = C.get("cfg1")
    nf = 260
    say(f"case cfg1 half 0 x {nf} frames (8.7 GB of I/Q, offsets past 2^31 floats)")
    one = synth.frame(cfg)[0].to(dev)                     # [1, T, C, Ns]
    nt, nc, ns = one.shape[1:]
    assert 2 * nf * nt * nc * ns > 2**31
    big = torch.zeros((nf, nt, nc, ns), dtype=torch.complex64, device=dev)
    big[0] = one[0]
    big[nf - 1] = one[0]
    plan, _, _ = plan_for(cfg, 0, dev)
    ref, _ = plan.forward(one, want_mask=False)
    out, _ = plan.forward(big, want_mask=False)
    torch.cuda.synchronize()
    for f in (0, nf - 1):
        r = H.rel_l2(out[f].cpu().numpy(), ref[0].cpu().numpy())
        say(f"frame {f}: rel-L2 vs the single-frame launch {r:.3e}")
        assert r <= 1e-6
    assert torch.count_nonzero(out[nf // 2]).item() == 0
    del big, out
