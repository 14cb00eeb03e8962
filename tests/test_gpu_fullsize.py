"""Full-size GPU parity of the launches the drop-in, the bench and serving run.

Grids given as [Nz, Nx, Ny] volumes (or a depth hint of Nz) make every
launch depth-tiled: a forward workgroup owns 16 depths x 16 neighbouring
columns and stages per (tilt chunk, line) only its sample window; the adjoint
owns 64 depths x 4V columns (V voxels per lane, 4 by default).  These tests run that launch at the workloads'
own sizes:

* config 0 (64+64 elements, 11 tilts, 64x64x128, a=3): the whole volume
  against the FP64 oracle;
* config 1 (128+128, 2 x 16 tilts, 128x128x256, a=4): whole workgroup tiles at
  the lateral corners, edges and centre, first / middle / last depth tile,
  against the oracle; the whole volume against the untiled launch, per tilt
  and compounded, both halves; the adjoint of both halves (dot-product
  identity, tiled vs untiled, gradient traces against the oracle adjoint);
* config 1 with each aberration law (per line, per (tilt, line), per (line,
  voxel) with 2 x 2.1 GB of delays): whole tiles against the oracle; NoSum
  on a 64-deep 128 x 32 slab (4.3 GB of per-line outputs): whole tiles
  against the oracle, the line sum against the per-tilt output;
* config 2 (config-1 geometry, 200-frame Doppler ensemble, 40 dB tissue
  slab): FP16-I/Q against the FP32 kernel on all 200 frames, both against the
  oracle on whole tiles of frames 0, 99 and 199;
* config 3 (serving, 128-element 2048-sample probe, a=4): requests inside its
  ranges through the real GPU backend, one of 160^2 x 320 voxels and 64
  tilts: bitwise alone vs in a mix, equal to the plain path, FP32 against the
  oracle on whole tiles, FP16 against FP32;
* config 4 (96+96, 16 tilts, 96x96x192, 8 frames, a=3): the forward against
  the oracle on whole tiles, the adjoint dot-product identity at full size,
  and gradient traces against the oracle adjoint restricted to their tilt and
  line.

Bars: FP32 rel-L2 <= 1e-5 and envelope error <= 0.1 dB (voxels within -40 dB
of the maximum); FP16-I/Q rel-L2 <= 3e-3 and envelope <= 0.1 dB against
FP32 on config 2; tiled vs untiled 4e-6 compounded (the tilt-chunk order of
the sum differs), 1e-6 per tilt; adjoint 1e-5.
"""
import numpy as np
import pytest
import torch

from oracle import rca_oracle as O
from simu_rca_ab_amd import configs as C, synth
from simu_rca_ab_amd.plan import RcaPlan, quantize_fp16
from tests import helpers as H
from tests.test_gpu_forward import check, say

pytestmark = pytest.mark.gpu

REL, ENV = 1e-5, 0.1


def tiled_plan(cfg, hi, dev, depth_run=None, interp=None):
    """The bench / drop-in plan: flattened grid with the depth hint Nz."""
    G, T, P, Op = H.case(cfg, hi)
    if interp:
        Op = dict(Op, interpolation=interp)
    plan = RcaPlan.from_reference(H.flat_geometry(G), T, P, Op, n_sample=cfg.n_sample, device=dev,
                                  depth_run=cfg.nz if depth_run is None else depth_run)
    return plan, G, T


def oracle_on(cfg, hi, iq, sel, interp=None, frames=None):
    """FP64 oracle, compounded over the half's tilts, on voxels ``sel``;
    iq: complex64 [F, T, C, Ns] (CPU)."""
    G, T, P, Op = H.case(cfg, hi)
    Gs = dict(G)
    for k in ("gridBfX", "gridBfY", "gridBfZ"):
        Gs[k] = np.asarray(G[k]).ravel(order="F")[sel]
    kind, a = {"lanczos3": (O.INTERP_LANCZOS, 3), "lanczos4": (O.INTERP_LANCZOS, 4)}[interp or cfg.interpolation]
    x = iq if frames is None else iq[frames]
    r, m = O.kernel_forward(np.ascontiguousarray(x.numpy()).view(np.float32), H.grid_f32(Gs),
                            O.prepare(G, T), H.oracle_params(cfg, cfg.halves[hi]), kind, a)
    return r.sum(axis=1), m.sum(axis=0)


def test_cfg0_full_volume_tiled(dev):
    cfg = C.get("cfg0")
    say("case cfg0 full 64x64x128 through the depth-tiled launch vs FP64 oracle (whole volume)")
    iq = H.iq_frame(cfg)
    plan, G, T = tiled_plan(cfg, 0, dev)
    assert plan.depth_run == cfg.nz
    out, mask = plan.forward(iq.to(dev))
    torch.cuda.synchronize()
    ref, rmask = oracle_on(cfg, 0, iq, np.arange(cfg.n_vox))
    check(out[0, 0].cpu().numpy(), ref[0], "cfg0 tiled full volume")
    assert H.rel_l2(mask[0].cpu().numpy(), rmask) <= REL


@pytest.fixture(scope="module")
def cfg1_iq():
    return synth.frame(C.get("cfg1"))


def test_cfg1_tiled_whole_tiles_vs_oracle(dev, cfg1_iq):
    cfg = C.get("cfg1")
    sel, blocks = H.edge_tiles(cfg.nx, cfg.ny, cfg.nz, n_random=12, seed=5)
    say(f"case cfg1 full, depth-tiled launch, both halves: {len(blocks)} whole workgroup tiles "
        f"({sel.size} voxels: corners, edges, centre x first/middle/last depth tile) vs FP64 oracle")
    out = mask = None
    ref = rmask = 0
    for hi in range(2):
        plan, _, _ = tiled_plan(cfg, hi, dev)
        if out is None:
            out, mask = plan.forward(cfg1_iq[hi].to(dev))
        else:
            plan.forward(cfg1_iq[hi].to(dev), out=out, mask=mask, accumulate=True)
        del plan
        r, m = oracle_on(cfg, hi, cfg1_iq[hi], sel)
        ref, rmask = ref + r[0], rmask + m
    torch.cuda.synchronize()
    got = out[0, 0].cpu().numpy()
    assert np.isfinite(got).all()
    check(got[sel], ref, "cfg1 tiled, whole tiles")
    assert H.rel_l2(mask[0].cpu().numpy()[sel], rmask) <= REL
    for b in blocks[:6]:   # per tile, the corner / edge tiles one by one
        v = np.searchsorted(sel, H.tile_voxels(cfg.n_vox, cfg.nz, b))
        r = H.rel_l2(got[sel][v], ref[v])
        say(f"  tile {b}: rel-L2 {r:.3e}")
        assert r <= REL


@pytest.mark.parametrize("hi", [0, 1])
def test_cfg1_tiled_vs_untiled_full_volume(dev, cfg1_iq, hi):
    cfg = C.get("cfg1")
    say(f"case cfg1 half {hi}: depth-tiled vs untiled launch on the whole 128x128x256 volume, "
        f"per tilt and compounded")
    x = cfg1_iq[hi].to(dev)
    tiled, _, _ = tiled_plan(cfg, hi, dev)
    flat, _, _ = tiled_plan(cfg, hi, dev, depth_run=0)
    assert tiled.depth_run == cfg.nz and flat.depth_run == 0
    for compound in (True, False):
        o_t, m_t = tiled.forward(x, compound=compound)
        o_f, m_f = flat.forward(x, compound=compound)
        torch.cuda.synchronize()
        r = H.rel_l2_t(o_t, o_f)
        rm = float(torch.linalg.vector_norm((m_t - m_f).double()) / torch.linalg.vector_norm(m_f.double()))
        e = H.envelope_db_err_t(o_t, o_f)
        say(f"compound={compound}: rel-L2 {r:.3e}, mask {rm:.3e}, envelope {e:.2e} dB")
        assert r <= (4e-6 if compound else 1e-6) and rm <= 4e-6 and e <= 1e-3
        del o_t, o_f, m_t, m_f
    torch.cuda.empty_cache()


@pytest.mark.parametrize("hi", [0, 1])
def test_cfg1_adjoint_full(dev, cfg1_iq, hi):
    """The adjoint at config 1's full size (the receive-X half runs the
    line-outer kernel, the receive-Y half the general one): the dot-product
    identity over the whole volume, the tiled launch against the untiled one,
    and gradient traces against the oracle adjoint restricted to their tilt
    and line."""
    cfg = C.get("cfg1")
    say(f"case cfg1 half {hi}: full-size adjoint (128x128x256, 16 tilts, 128 lines)")
    x = cfg1_iq[hi].to(dev)
    plan, G, T = tiled_plan(cfg, hi, dev)
    y_fwd, _ = plan.forward(x, want_mask=False)
    g = torch.Generator(device=dev).manual_seed(1100 + hi)
    yv = torch.randn((1, 1, plan.n_pix), dtype=torch.complex64, device=dev, generator=g)
    AHy = plan.adjoint(yv)
    torch.cuda.synchronize()
    lhs = torch.vdot(yv.flatten().to(torch.complex128), y_fwd.flatten().to(torch.complex128))
    rhs = torch.vdot(AHy.flatten().to(torch.complex128), x.flatten().to(torch.complex128))
    r = abs(complex(lhs - rhs)) / abs(complex(lhs))
    say(f"dot-product identity at full size: mismatch {r:.3e}")
    assert r <= REL
    flat, _, _ = tiled_plan(cfg, hi, dev, depth_run=0)
    r = H.rel_l2_t(AHy, flat.adjoint(yv))
    torch.cuda.synchronize()
    say(f"tiled vs untiled adjoint: rel-L2 {r:.3e}")
    assert r <= REL
    prep = O.prepare(G, T)
    grid = H.grid_f32(G)
    prm = H.oracle_params(cfg, cfg.halves[hi])
    yc = yv.cpu().numpy()
    for t, c in ((0, 0), (8, 64), (15, 127)):
        p1 = dict(prep)
        for k in ("ex", "ey", "ez"):
            p1[k] = np.ascontiguousarray(prep[k][c:c + 1])
        for k in ("vx", "vy", "vz", "d0"):
            p1[k] = np.ascontiguousarray(prep[k][t:t + 1])
        want = O.kernel_adjoint(yc, grid, p1, prm, 1, O.INTERP_LANCZOS, 4)[:, 0, 0]
        have = AHy[:, t, c].cpu().numpy()
        r = H.rel_l2(have, want)
        say(f"gradient trace (tilt {t}, line {c}): rel-L2 vs oracle adjoint {r:.3e}")
        assert r <= REL
    del plan, flat, AHy, y_fwd
    torch.cuda.empty_cache()


@pytest.mark.parametrize("kind,hi", [("channel", 0), ("tx_channel", 1), ("pixel_channel", 0)])
def test_cfg1_laws_whole_tiles(dev, cfg1_iq, kind, hi):
    """The aberration-law kernels at config 1's full size through the
    depth-tiled launch: per line (Hyp1), per (tilt, line) (Hyp2) and per
    (line, voxel) with receive and transmit arrays (TxRx, 2 x 2.1 GB), delays
    of up to 3 samples, random per entry; whole workgroup tiles (corners,
    edges, centre x first/middle/last depth tile) vs the FP64 oracle."""
    cfg = C.get("cfg1")
    n_ch, n_tx, n_pix = cfg.n_elem, len(cfg.halves[hi].theta), cfg.n_vox
    say(f"case cfg1 half {hi}: {kind} law, full size, depth-tiled, whole tiles vs oracle")
    rng = np.random.default_rng(310 + hi)
    span = 3.0 / cfg.fs

    def draw(shape):
        a = rng.random(shape, dtype=np.float32)
        a *= 2 * span
        a -= span
        return a

    sel, _ = H.edge_tiles(cfg.nx, cfg.ny, cfg.nz, n_random=4, seed=31)
    if kind == "channel":
        rx, tx = draw(n_ch), None
        okw = dict(ablaw=rx, ablaw_mode=1)
    elif kind == "tx_channel":
        rx, tx = draw((n_ch, n_tx)), None
        okw = dict(ablaw=rx, ablaw_mode=2)
    else:
        rx, tx = draw((n_pix, n_ch)), draw((n_pix, n_ch))   # MATLAB [nPix, nChannel]
        okw = dict(ablaw=np.ascontiguousarray(rx[sel]), ablaw_tx=np.ascontiguousarray(tx[sel]),
                   ablaw_mode=3)
    G, T, P, Op = H.case(cfg, hi)
    P = dict(P, AbLawRx=rx, AbLawTx=tx, AbLawMode=kind)
    plan = RcaPlan.from_reference(H.flat_geometry(G), T, P, Op, n_sample=cfg.n_sample, device=dev,
                                  depth_run=cfg.nz)
    del rx, tx
    assert plan.depth_run == cfg.nz
    out, mask = plan.forward(cfg1_iq[hi].to(dev))
    torch.cuda.synchronize()
    got = out[0, 0].cpu().numpy()[sel]
    gm = mask[0].cpu().numpy()[sel]
    del plan, out, mask
    torch.cuda.empty_cache()
    Gs = dict(G)
    for k in ("gridBfX", "gridBfY", "gridBfZ"):
        Gs[k] = np.asarray(G[k]).ravel(order="F")[sel]
    kind_o, a = O.INTERP_LANCZOS, 4
    r, m = O.kernel_forward(np.ascontiguousarray(cfg1_iq[hi].numpy()).view(np.float32), H.grid_f32(Gs),
                            O.prepare(G, T), H.oracle_params(cfg, cfg.halves[hi]), kind_o, a, **okw)
    check(got, r.sum(axis=1)[0], f"cfg1 half {hi} {kind} law, {sel.size} voxels in whole tiles")
    assert H.rel_l2(gm, m.sum(axis=0)) <= REL


@pytest.mark.parametrize("hi", [0, 1])
def test_cfg1_nosum_slab(dev, cfg1_iq, hi):
    """NoSum (one output per receive line, NoSum.cu:376-380) at the bench's
    variant shape: a 64-deep 128 x 32 slab of the config-1 grid, 16 tilts x
    128 lines (4.3 GB of per-line output), depth-tiled: whole tiles vs the
    oracle, and the per-line outputs summed over the lines equal to the
    per-tilt output over the whole slab."""
    cfg = C.get("cfg1")
    h = cfg.halves[hi]
    nx, ny, nz = cfg.nx, 32, 64
    say(f"case cfg1 half {hi}: NoSum on a {nz}-deep {nx}x{ny} slab, 16 tilts x 128 lines")
    x, y, z = C.grid_axes(cfg)
    G = C.geometry(cfg, h, C.meshgrid_zxy(x, y[:ny], z[:nz]))
    T = C.transmit(h)
    plan = RcaPlan.from_reference(H.flat_geometry(G), T, C.param(cfg), C.options(cfg),
                                  n_sample=cfg.n_sample, device=dev, depth_run=nz)
    assert plan.depth_run == nz
    xi = cfg1_iq[hi].to(dev)
    ns, _ = plan.forward(xi, compound=False, nosum=True, want_mask=False)   # [nCh, F, T, P]
    pt, _ = plan.forward(xi, compound=False, want_mask=False)               # [F, T, P]
    torch.cuda.synchronize()
    r = H.rel_l2_t(ns.sum(dim=0), pt)
    say(f"sum over lines of the NoSum outputs vs the per-tilt output: rel-L2 {r:.3e}")
    assert r <= 1e-6
    sel, _ = H.edge_tiles(nx, ny, nz, n_random=4, seed=17)
    got = ns[:, 0][..., torch.from_numpy(sel).to(dev)].cpu().numpy()        # [nCh, T, n_sel]
    del ns, pt, plan
    torch.cuda.empty_cache()
    Gs = dict(G)
    for k in ("gridBfX", "gridBfY", "gridBfZ"):
        Gs[k] = np.asarray(G[k]).ravel(order="F")[sel]
    ref, _ = O.kernel_forward(np.ascontiguousarray(cfg1_iq[hi].numpy()).view(np.float32),
                              H.grid_f32(Gs), O.prepare(G, T), H.oracle_params(cfg, h),
                              O.INTERP_LANCZOS, 4, nosum=True)
    check(got, ref[:, 0], f"cfg1 half {hi} NoSum slab, {sel.size} voxels x 16 tilts x 128 lines")


def test_cfg2_ensemble_200_frames(dev):
    cfg = C.get("cfg2")
    nf = cfg.n_frame
    say(f"case cfg2 Doppler ensemble, all {nf} frames (FP32 13.4 GB, FP16 6.7 GB of I/Q), "
        "depth-tiled launch: FP16 vs FP32 kernel on every frame, both vs the oracle on whole tiles")
    halves = synth.ensemble(cfg, n_frame=nf, device=dev)
    say("  ensemble synthesised")
    out32 = out16 = None
    for hi in range(2):
        plan, _, _ = tiled_plan(cfg, hi, dev)
        h, s = quantize_fp16(halves[hi])
        if out32 is None:
            out32, _ = plan.forward(halves[hi], want_mask=False)
            out16, _ = plan.forward(h, s, want_mask=False)
        else:
            plan.forward(halves[hi], out=out32, want_mask=False, accumulate=True)
            plan.forward(h, s, out=out16, want_mask=False, accumulate=True)
        del plan, h, s
        torch.cuda.synchronize()
        say(f"  half {hi} beamformed")
    assert torch.isfinite(out32).all() and torch.isfinite(out16).all()
    r = H.rel_l2_t(out16, out32)
    e = H.envelope_db_err_t(out16, out32)
    say(f"fp16 vs FP32 kernel, {nf} frames: rel-L2 {r:.3e}, envelope {e:.4f} dB")
    assert r <= 3e-3 and e <= ENV
    per = [H.rel_l2_t(out16[f], out32[f]) for f in range(nf)]
    say(f"  per frame: max rel-L2 {max(per):.3e} (frame {int(np.argmax(per))})")
    assert max(per) <= 3e-3
    # moving scatterers: the frames differ
    assert H.rel_l2_t(out32[nf - 1], out32[0]) > 1e-3
    sel, _ = H.edge_tiles(cfg.nx, cfg.ny, cfg.nz, n_random=4, seed=22)
    frames = [0, nf // 2 - 1, nf - 1]
    ref = 0
    for hi in range(2):
        ref = ref + oracle_on(cfg, hi, halves[hi][frames].cpu(), sel)[0]
    a32 = out32[frames, 0][:, torch.from_numpy(sel).to(dev)].cpu().numpy()
    a16 = out16[frames, 0][:, torch.from_numpy(sel).to(dev)].cpu().numpy()
    check(a32, ref, f"FP32 vs FP64 oracle, frames {frames} ({sel.size} voxels in whole tiles)")
    check(a16, ref, f"FP16 vs FP64 oracle, frames {frames}", rel=3e-3, env=ENV)
    del halves, out32, out16
    torch.cuda.empty_cache()


def _cfg3_backend(dev):
    from simu_rca_ab_amd.serving import GpuBackend
    return GpuBackend(dev)          # 128 elements, 2048 samples, Lanczos a=4


def test_cfg3_serving_shapes(dev):
    from simu_rca_ab_amd.serving import RcaServer, Request, RequestSpec, synth_request
    be = _cfg3_backend(dev)
    specs = [RequestSpec(160, 160, 320, 64, 2, "fp32"), RequestSpec(64, 64, 128, 8, 5, "fp16"),
             RequestSpec(100, 100, 200, 16, 3, "fp16"), RequestSpec(117, 117, 234, 32, 1, "fp32")]
    say("case cfg3 serving on the 128-element 2048-sample probe, Lanczos a=4: "
        + ", ".join(f"{s.nx}^2x{s.nz} {s.n_tilt} tilts {s.n_frame} fr {s.dtype}" for s in specs))
    reqs = [synth_request(i, s, be, dev, n_scat=100) for i, s in enumerate(specs)]
    fresh = lambda: [Request(r.rid, r.spec, r.halves) for r in reqs]
    alone = {}
    srv = RcaServer(be, budget_bytes=64 << 30).start()
    for r in fresh():
        srv.submit(r)
        alone[r.rid] = r.wait(300).numpy().copy()
    srv.close(60)
    say("  alone done")
    srv = RcaServer(be, budget_bytes=64 << 30, quantum=int(4e10)).start()
    mix = fresh()
    for r in mix:
        srv.submit(r)
    for r in mix:
        assert np.array_equal(r.wait(300).numpy(), alone[r.rid]), f"request {r.rid} differs alone vs mix"
    srv.close(60)
    say(f"  mix bitwise equal: {srv.stats.summary()['batches']} batches")
    for r in reqs:
        sp = r.spec
        got = alone[r.rid]
        out = None
        for h in range(2):
            tab = be._tilt_table(sp.n_tilt)[h]
            plan = RcaPlan(grid=be._grid(sp.nx, sp.ny, sp.nz), ele=tab["ele"], vs=tab["vs"],
                           min_dist=tab["d0"], receive_x=h == 0, aperture=tab["aperture"],
                           fs=be.fs, fd=be.fd, c=be.c, t0=0.0, fnumber=1.0, n_sample=be.n_sample,
                           interp="lanczos4", device=dev, depth_run=sp.nz)
            x = r.halves[h]
            args = (x[0].to(dev), x[1].to(dev)) if isinstance(x, tuple) else (x.to(dev),)
            o, _ = plan.forward(*args, out=out, accumulate=out is not None, want_mask=False)
            out = o
        rp = H.rel_l2(got, out[:, 0].cpu().numpy())
        say(f"request {r.rid} ({sp.dtype}): vs plain tiled path {rp:.2e}")
        # serving runs frames in blocks whose LDS layout sets the tilts per
        # chunk, so its tilt-sum order can differ from the plain launch's
        assert rp <= 4e-6
        if sp.dtype == "fp32":
            sel, _ = H.edge_tiles(sp.nx, sp.ny, sp.nz, n_random=3, seed=r.rid)
            ref = 0
            for h in range(2):
                tab = be._tilt_table(sp.n_tilt)[h]
                prep = {"ex": tab["ele"][0].cpu().numpy(), "ey": tab["ele"][1].cpu().numpy(),
                        "ez": tab["ele"][2].cpu().numpy(), "vx": tab["vs"][0].cpu().numpy(),
                        "vy": tab["vs"][1].cpu().numpy(), "vz": tab["vs"][2].cpu().numpy(),
                        "d0": tab["d0"].cpu().numpy(), "ap": np.float32(tab["aperture"])}
                grid = tuple(np.ascontiguousarray(g.cpu().numpy()[sel]) for g in be._grid(sp.nx, sp.ny, sp.nz))
                prm = dict(receive_x=h == 0, fs=np.float32(be.fs), fd=np.float32(be.fd),
                           c=np.float32(be.c), t0=np.float32(0), fnumber=np.float32(1), n_sample=be.n_sample)
                o, _ = O.kernel_forward(r.halves[h].numpy().view(np.float32), grid, prep, prm,
                                        O.INTERP_LANCZOS, 4)
                ref = ref + o.sum(axis=1)
            check(got[:, sel], ref, f"request {r.rid} FP32 vs FP64 oracle ({sel.size} voxels, whole tiles)")
    # FP16 vs FP32 on the same request: serve request 2's complex64 twin
    s2 = specs[2]
    twin = synth_request(2, RequestSpec(s2.nx, s2.ny, s2.nz, s2.n_tilt, s2.n_frame, "fp32"), be, dev,
                         n_scat=100)
    srv = RcaServer(be, budget_bytes=64 << 30).start()
    srv.submit(twin)
    b32 = twin.wait(300).numpy()
    srv.close(60)
    check(alone[2], b32, "request 2: FP16-I/Q vs FP32", rel=3e-3, env=ENV)


def test_cfg4_full(dev):
    cfg = C.get("cfg4")
    nf = cfg.n_frame
    say("case cfg4 full 96x96x192, 16 tilts, 96 lines, 8 frames, a=3, depth-tiled forward and adjoint")
    plan, G, T = tiled_plan(cfg, 0, dev)
    iq = H.iq_frame(cfg, 0, n_frame=nf)           # 8 seeded frames (different scatterers)
    x = iq.to(dev)
    y_fwd, _ = plan.forward(x, want_mask=False)
    torch.cuda.synchronize()
    sel, _ = H.edge_tiles(cfg.nx, cfg.ny, cfg.nz, n_random=6, seed=44)
    ref, _ = oracle_on(cfg, 0, iq, sel, frames=[0, nf - 1])
    got = y_fwd[[0, nf - 1], 0][:, torch.from_numpy(sel).to(dev)].cpu().numpy()
    check(got, ref, f"cfg4 forward, frames 0 and {nf - 1} ({sel.size} voxels in whole tiles)")
    # dot-product identity <A x, y> = <x, A^H y> at full size
    g = torch.Generator(device=dev).manual_seed(4044)
    yv = torch.randn((nf, 1, plan.n_pix), dtype=torch.complex64, device=dev, generator=g)
    AHy = plan.adjoint(yv)
    torch.cuda.synchronize()
    lhs = torch.vdot(yv.flatten().to(torch.complex128), y_fwd.flatten().to(torch.complex128))
    rhs = torch.vdot(AHy.flatten().to(torch.complex128), x.flatten().to(torch.complex128))
    r = abs(complex(lhs - rhs)) / abs(complex(lhs))
    say(f"dot-product identity at full size: mismatch {r:.3e}")
    assert r <= 1e-5
    # gradient traces vs the oracle adjoint restricted to their tilt and line
    prep = O.prepare(G, T)
    grid = H.grid_f32(G)
    prm = H.oracle_params(cfg, cfg.halves[0])
    yc = yv.cpu().numpy()
    for t, c in ((0, 0), (7, 47), (15, 95), (11, 20)):
        p1 = dict(prep)
        for k in ("ex", "ey", "ez"):
            p1[k] = np.ascontiguousarray(prep[k][c:c + 1])
        for k in ("vx", "vy", "vz", "d0"):
            p1[k] = np.ascontiguousarray(prep[k][t:t + 1])
        want = O.kernel_adjoint(yc, grid, p1, prm, nf, O.INTERP_LANCZOS, 3)[:, 0, 0]
        have = AHy[:, t, c].cpu().numpy()
        r = H.rel_l2(have, want)
        say(f"gradient trace (tilt {t}, line {c}), 8 frames: rel-L2 vs oracle adjoint {r:.3e}")
        assert r <= REL
