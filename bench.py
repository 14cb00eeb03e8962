"""Benchmark of the RCA delay-and-sum beamformer on one MI355X.

Workload (one "step"): one full RCA frame of the 128+128-element probe —
2 x 16 tilts (-12..12 deg) of 128-line traces with 2048 complex64 samples,
beamformed on 128x128x256 voxels with Lanczos a=4, the two halves summed in
place (two launches of the forward kernel).  Inputs are seeded synthetic I/Q
already resident in HBM when the timed region starts.

Metric: voxel*tilt*element updates per second, counted densely
(nVox * nTilt * nElem per frame).  Prints one JSON line.

Extra legs:
  * roofline of the dominant kernel (the forward DAS launch), which is bound
    by vector-ALU issue: VALU lane-instructions per launch (hardware counter
    profile of this kernel and workload) over the launch's mean duration from
    HIP events on its stream, against the issue ceiling measured by the FMA
    probe in the same run; HBM traffic, its fraction of 8 TB/s and the
    compulsory bytes beside it;
  * cpu_baseline: the FP32 CPU restatement of the reference kernel (OpenMP on
    every host core) on a bounded voxel sample of the same workload;
  * fp16: the FP16-I/Q mode on a 16-frame ensemble of the same geometry.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
# MIOpen: heuristic solver choice instead of an exhaustive search on first use
os.environ.setdefault("MIOPEN_FIND_MODE", "FAST")

HBM_PEAK_GBS = 8000.0
# MI355X L2: 8 XCDs x 16 channels x 128 B per clock at 2.4 GHz (derived from
# the L2 organisation; no published figure)
L2_PEAK_GBS = 8 * 16 * 128 * 2.4
# MI355X FP32 vector peak 157.3 TFLOP/s = 78.6 T lane-FMA/s (2 FLOP per FMA)
VALU_PEAK_NOMINAL_T = 78.6
TRAFFIC_PROFILE = ROOT / "profiles" / "das_fwd_cfg1_counters.md"


def _profile_value(key):
    """A ``key = value`` figure of the committed das_fwd counter summary
    (rocprofv3 --pmc cannot run inside the timed process), or None."""
    import re
    try:
        m = re.search(key + r"\s*=\s*([0-9.eE+-]+)", TRAFFIC_PROFILE.read_text())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def measured_traffic():
    """HBM bytes per das_fwd launch at config 1 (FETCH_SIZE/WRITE_SIZE)."""
    return _profile_value("traffic_bytes_per_launch")


def valu_probe(dev, n_blocks=2048, iters=4096, reps=5, mode=0):
    """Measured VALU issue ceiling in T lane-ops/s: the library's probe of
    independent instruction chains of one class (mode 0: v_fma_f32), 8 waves
    per SIMD on 256 CUs, HIP events on its stream."""
    import ctypes
    import torch
    from simu_rca_ab_amd import _lib
    lib = _lib.load()
    out = torch.empty(n_blocks * 256, dtype=torch.float32, device=dev)
    st = torch.cuda.current_stream(dev)
    lane = ctypes.c_longlong(0)
    run = lambda: _lib.check(lib.rca_probe_valu(mode, ctypes.c_void_p(out.data_ptr()), n_blocks, iters,
                                                 ctypes.c_void_p(st.cuda_stream), ctypes.byref(lane)))
    run()
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        run()
        b.record(st)
        torch.cuda.synchronize()
        ms.append(a.elapsed_time(b))
    return lane.value / (statistics.median(ms) * 1e-3) / 1e12, statistics.median(ms)


def roofline(cfg, mean_launch_ms, upd_launch, probe_t):
    """Roofline of the dominant kernel (das_fwd, one config-1 half per launch).
    The kernel is bound by VALU issue: achieved = VALU lane-instructions per
    launch (SQ_INSTS_VALU x 64 from the committed counter profile of this
    kernel on this workload) / the launch's mean duration measured here,
    against the FMA-chain probe's ceiling measured in the same process.  HBM
    is reported beside it from the measured traffic (FETCH_SIZE/WRITE_SIZE,
    copy-calibrated), with the compulsory bytes; the interpolation footprint
    (taps x 8 B per update) is not HBM traffic, since the windows are staged
    in LDS: its ratio to the measured traffic is the LDS reuse factor."""
    t = mean_launch_ms * 1e-3
    vinst = _profile_value("valu_insts_per_launch")
    traffic = measured_traffic()
    achieved = vinst * 64 / t / 1e12 if vinst else None
    algo = upd_launch * cfg.taps() * 8
    n_pix = cfg.n_vox
    iq = len(cfg.halves[0].theta) * cfg.n_elem * cfg.n_sample * 8
    # per launch, mean of the two halves: I/Q and grid read, output and mask
    # written (the second half also reads them to accumulate)
    compulsory = iq + 12 * n_pix + 1.5 * (8 * n_pix + 4 * n_pix)
    # L2 (TCC) traffic: read requests x 128 B + write requests x 64 B, the
    # request sizes a 512 MiB copy calibrates
    l2_rd, l2_wr = _profile_value("l2_read_bytes_per_launch"), _profile_value("l2_write_bytes_per_launch")
    l2 = None
    if l2_rd and l2_wr:
        l2 = {"achieved_gbs": (l2_rd + l2_wr) / t / 1e9, "peak_gbs": L2_PEAK_GBS,
              "frac": (l2_rd + l2_wr) / t / 1e9 / L2_PEAK_GBS, "read_bytes": l2_rd, "write_bytes": l2_wr,
              "hit_rate": _profile_value("l2_hit_rate"),
              "peak_source": "8 XCDs x 16 channels x 128 B/clock x 2.4 GHz (derived)"}
    return {"bound": "valu", "achieved": achieved, "peak": probe_t, "unit": "T VALU lane-instr/s",
            "frac": achieved / probe_t if achieved else None, "traffic": traffic,
            "peak_source": "measured in this run: rca_probe_valu mode 1 (8 independent v_fma_f32 "
                           "chains per lane with inline-constant operands, 8 waves/SIMD): one VALU "
                           "wave-instruction per ~2.2 cycles per SIMD; nominal 78.6 T lane-ops/s = "
                           "157.3 TFLOP/s FP32 / 2 (2 cycles per wave64 instruction at 2.4 GHz). "
                           "Instructions with two or three VGPR sources (the tap FMAs) issue at "
                           "about half that rate (ceiling_two_vgpr_sources, probe mode 0)",
            "peak_nominal": VALU_PEAK_NOMINAL_T, "frac_of_nominal": achieved / VALU_PEAK_NOMINAL_T if achieved else None,
            "valu_insts_per_launch": vinst, "valu_lane_instr_per_update": vinst * 64 / upd_launch if vinst else None,
            "source": "profiles/das_fwd_cfg1_counters.md (SQ_INSTS_VALU, FETCH_SIZE, WRITE_SIZE, TCC_READ, TCC_WRITE, TCC_HIT, TCC_MISS)",
            "kernel": "das_fwd<lanczos4, FP32, V=1, FB=1>", "launch_ms": mean_launch_ms,
            "updates_per_launch": upd_launch,
            "hbm": {"achieved_gbs": traffic / t / 1e9 if traffic else None, "peak_gbs": HBM_PEAK_GBS,
                    "frac": traffic / t / 1e9 / HBM_PEAK_GBS if traffic else None,
                    "compulsory_bytes": compulsory,
                    "traffic_over_compulsory": traffic / compulsory if traffic else None},
            "l2": l2,
            "lds_reuse_factor": algo / traffic if traffic else None,
            "interp_footprint_bytes_per_update": cfg.taps() * 8}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_plans(cfg, dev):
    from simu_rca_ab_amd import configs as C
    from simu_rca_ab_amd.plan import RcaPlan
    grid = C.meshgrid_zxy(*C.grid_axes(cfg))
    plans = []
    for h in cfg.halves:
        G = C.geometry(cfg, h, grid)
        for k in ("gridBfX", "gridBfY", "gridBfZ"):
            G[k] = G[k].ravel(order="F")
        plans.append(RcaPlan.from_reference(G, C.transmit(h), C.param(cfg), C.options(cfg),
                                            n_sample=cfg.n_sample, device=dev, depth_run=cfg.nz))
    return plans


def run_gpu(cfg, steps, warmup, dev):
    import torch
    from simu_rca_ab_amd import synth
    log(f"[bench] synthesising {cfg.name} I/Q on {dev}")
    iqs = [x.to(dev).contiguous() for x in synth.frame(cfg, device=dev)]
    plans = build_plans(cfg, dev)
    out = torch.empty((1, 1, plans[0].n_pix), dtype=torch.complex64, device=dev)
    mask = torch.empty((1, plans[0].n_pix), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    ev = []

    def step(record):
        for i, (p, iq) in enumerate(zip(plans, iqs)):
            if record:
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
            p.forward(iq, out=out, mask=mask, accumulate=i > 0)
            if record:
                b.record(stream)
                ev.append((a, b))

    log(f"[bench] warmup {warmup}")
    for _ in range(warmup):
        step(False)
    torch.cuda.synchronize()
    log(f"[bench] timing {steps} steps")
    t0 = time.perf_counter()
    for _ in range(steps):
        step(True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    launch_ms = [a.elapsed_time(b) for a, b in ev]
    return dt, launch_ms, plans, iqs, out


def run_fp16(cfg, dev, n_frame=16, reps=3):
    """FP16-I/Q ensemble throughput on one half-pair of the same geometry."""
    import torch
    from simu_rca_ab_amd import synth
    from simu_rca_ab_amd.plan import quantize_fp16
    plans = build_plans(cfg, dev)
    iqs = [x.to(dev) for x in synth.frame(cfg, device=dev)]
    # an ensemble of n_frame frames (frame content varies by a per-frame phase)
    res = {}
    for name, make in (("fp16", True), ("fp32", False)):
        data = []
        for iq in iqs:
            ph = torch.exp(1j * torch.linspace(0, 3, n_frame, device=dev)).to(torch.complex64)
            ens = (iq * ph[:, None, None, None]).contiguous()
            data.append(quantize_fp16(ens) if make else (ens, None))
        out = torch.empty((n_frame, 1, plans[0].n_pix), dtype=torch.complex64, device=dev)
        for _ in range(1):
            for i, (p, (x, s)) in enumerate(zip(plans, data)):
                p.forward(x, s, out=out, want_mask=False, accumulate=i > 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            for i, (p, (x, s)) in enumerate(zip(plans, data)):
                p.forward(x, s, out=out, want_mask=False, accumulate=i > 0)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        res[name] = cfg.updates_per_frame() * n_frame / dt
        del data, out
    return res


def run_cfg2(dev):
    """Config 2: the 200-frame Doppler ensemble on the config-1 geometry
    (scatterers moving up to 5 mm/s, 1 ms apart, a static tissue slab 40 dB
    brighter), both halves compounded, from FP16-I/Q (half2 + per-trace FP32
    scale, 6.7 GB) and from complex64 I/Q (13.4 GB); updates/s of each pass
    and the FP16 result's error against the FP32 one over all frames."""
    import torch
    from simu_rca_ab_amd import configs as C, synth
    from simu_rca_ab_amd.plan import quantize_fp16
    cfg = C.get("cfg2")
    nf = cfg.n_frame
    t0 = time.perf_counter()
    halves = synth.ensemble(cfg, device=dev)
    torch.cuda.synchronize()
    log(f"[bench] cfg2: {nf}-frame ensemble synthesised in {time.perf_counter() - t0:.1f} s")
    plans = build_plans(cfg, dev)
    q = [quantize_fp16(h) for h in halves]
    out32 = torch.empty((nf, 1, plans[0].n_pix), dtype=torch.complex64, device=dev)
    out16 = torch.empty_like(out32)

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t

    def run32():
        for i, (p, x) in enumerate(zip(plans, halves)):
            p.forward(x, out=out32, want_mask=False, accumulate=i > 0)

    def run16():
        for i, (p, (h, sc)) in enumerate(zip(plans, q)):
            p.forward(h, sc, out=out16, want_mask=False, accumulate=i > 0)

    t32 = timed(run32)
    t16 = timed(run16)
    upd = cfg.updates_per_frame() * nf
    d = (out16 - out32).to(torch.complex128)
    rel = float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(out32.to(torch.complex128)))
    per = torch.linalg.vector_norm(d[:, 0], dim=1) / torch.linalg.vector_norm(
        out32[:, 0].to(torch.complex128), dim=1)
    g, w = out16.abs().reshape(-1).double(), out32.abs().reshape(-1).double()
    keep = w >= w.max() * 1e-2
    env = float((20 * torch.log10(g[keep].clamp_min(1e-30)) - 20 * torch.log10(w[keep])).abs().max())
    log(f"[bench] cfg2: FP32 {upd / t32 / 1e9:.1f}, FP16 {upd / t16 / 1e9:.1f} G updates/s, "
        f"FP16 vs FP32 rel-L2 {rel:.2e}, envelope {env:.3f} dB")
    res = {"fp32_updates_per_s": upd / t32, "fp16_updates_per_s": upd / t16,
           "fp16_over_fp32": t32 / t16, "fp32_s": t32, "fp16_s": t16, "frames": nf,
           "updates": upd, "fp16_vs_fp32_rel_l2": rel, "fp16_vs_fp32_rel_l2_worst_frame": float(per.max()),
           "fp16_vs_fp32_envelope_db_40db": env,
           "workload": "cfg2: 128+128 elements, 2 x 16 tilts, 2048 samples, 128x128x256 voxels, "
                       "Lanczos a=4, 200 frames (1000 moving scatterers + 2000 static tissue "
                       "scatterers 40 dB brighter), I/Q resident in HBM; the FP64-oracle error on "
                       "whole tiles of frames 0, 99, 199 is in tests/test_gpu_fullsize.py"}
    del halves, q, out32, out16, d
    return res


def run_adjoint(dev, reps=3):
    """DAS forward and adjoint (the backward of the beamformer) on config 4:
    96+96 elements, 16 tilts, 96x96x192 voxels, 8 frames, Lanczos a=3; the
    adjoint with float atomics and the deterministic (fixed-point) one."""
    import torch
    from simu_rca_ab_amd import configs as C, synth
    cfg = C.get("cfg4")
    plan = build_plans(cfg, dev)[0]
    one = synth.frame(cfg, device=dev)[0].to(dev)
    ph = torch.exp(1j * torch.linspace(0, 2, cfg.n_frame, device=dev)).to(torch.complex64)
    iq = (one * ph[:, None, None, None]).contiguous()   # [8, nTx, nCh, Ns]
    g = torch.Generator(device=dev).manual_seed(4004)
    gout = torch.randn((cfg.n_frame, 1, plan.n_pix), dtype=torch.complex64, device=dev, generator=g)
    res = {}
    for name, fn in (("forward", lambda: plan.forward(iq, want_mask=False)),
                     ("adjoint", lambda: plan.adjoint(gout)),
                     ("adjoint_deterministic", lambda: plan.adjoint(gout, deterministic=True))):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        res[name + "_ms"] = dt * 1e3
        res[name + "_updates_per_s"] = plan.updates(cfg.n_frame) / dt
    res["workload"] = ("cfg4: 96+96 elements, 16 tilts -10..10 deg, 1024 samples, 96x96x192 voxels, "
                       "8 frames, Lanczos a=3")
    return res


def run_svd(dev, n_frame=200, reps=3):
    """SVD clutter filter of a config-2-shaped Doppler ensemble: 128x128x256
    voxels x 200 frames (synthetic: a rank-3 tissue part 40 dB above random
    blood), high-pass keeping eigenvectors 3..199."""
    import torch
    from simu_rca_ab_amd import clutterfilter as CF
    npix = 128 * 128 * 256
    g = torch.Generator(device=dev).manual_seed(2002)
    t = torch.arange(n_frame, device=dev, dtype=torch.float32)
    iq = torch.randn((n_frame, npix), dtype=torch.complex64, device=dev, generator=g)
    for r in range(3):   # tissue: slowly varying, 40 dB above
        sp = torch.randn(npix, dtype=torch.complex64, device=dev, generator=g) * 100.0
        iq += torch.exp(2j * torch.pi * 0.01 * (r + 1) * t).to(torch.complex64)[:, None] * sp
    keep = [False] * 3 + [True] * (n_frame - 3)
    out, ev = CF.svd_filter(iq, keep)   # warm-up: rocBLAS / rocSOLVER set-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out, ev = CF.svd_filter(iq, keep)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    ev = ev.cpu().numpy()
    gram_flop = 8.0 * npix * n_frame * n_frame / 2   # Gram, one triangle
    # rank-3 projection Y - (Y U) U^H: two complex MACs per (pixel, frame, vector)
    proj_flop = 2 * 8.0 * npix * n_frame * 3
    return {"ms": dt * 1e3, "pixels": npix, "frames": n_frame,
            "pixel_frames_per_s": npix * n_frame / dt,
            "fp64_tflops": (gram_flop + proj_flop) / dt / 1e12,
            "tissue_to_blood_eigengap": float(ev[2] / ev[3]),
            "workload": "cfg2 shape: 128x128x256 voxels x 200 frames, complex64 in/out, FP64 "
                        "Gram (hand-written MFMA) + eigh (rocSOLVER ZHEEVD) + rank-3 "
                        "projection (hand-written, one pass over the complex64 data)"}


def run_variants(cfg, dev, reps=3):
    """The beamformer variants on one config-1 half (16 tilts x 128 lines,
    FP32, one frame, Lanczos a=4): compounded output with each receive
    aberration law (per line, per (tilt, line), per (line, voxel) with both
    arrays), and the NoSum per-line output (the per-line output of the full
    grid would be 69 GB: a 64-deep 128x32 slab instead).  Dense updates/s; the
    plain kernel on the same half for comparison."""
    import numpy as np
    import torch
    from simu_rca_ab_amd import configs as C, synth
    from simu_rca_ab_amd.plan import RcaPlan
    h = cfg.halves[0]
    iq = synth.frame(cfg, device=dev)[0].to(dev).contiguous()

    def plan_for(nz, ny=None, **kw):
        x, y, z = C.grid_axes(cfg)
        grid = C.meshgrid_zxy(x, y[:ny], z[:nz])
        G = C.geometry(cfg, h, grid)
        for k in ("gridBfX", "gridBfY", "gridBfZ"):
            G[k] = G[k].ravel(order="F")
        P = dict(C.param(cfg), **kw)
        return RcaPlan.from_reference(G, C.transmit(h), P, C.options(cfg), n_sample=cfg.n_sample,
                                      device=dev, depth_run=nz)

    def timeit(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    rng = np.random.default_rng(77)
    n_ch, n_tx, nz = cfg.n_elem, len(h.theta), cfg.nz
    res = {}
    base = plan_for(nz)
    upd = base.n_pix * n_tx * n_ch
    out = torch.empty((1, 1, base.n_pix), dtype=torch.complex64, device=dev)
    t = timeit(lambda: base.forward(iq, out=out, want_mask=False))
    res["plain"] = {"ms": t * 1e3, "updates_per_s": upd / t}
    laws = {"law_per_line": dict(AbLawRx=(rng.standard_normal(n_ch) * 20e-9).astype(np.float32),
                                 AbLawMode="channel"),
            "law_per_tilt_line": dict(AbLawRx=(rng.standard_normal(n_tx * n_ch) * 20e-9).astype(np.float32),
                                      AbLawMode="tx_channel"),
            # [nChannel][nPix] device tensors (2.1 GB each)
            "law_per_line_voxel": dict(
                AbLawRx=torch.randn((n_ch, base.n_pix), device=dev) * 20e-9,
                AbLawTx=torch.randn((n_ch, base.n_pix), device=dev) * 20e-9,
                AbLawMode="pixel_channel")}
    for name, kw in laws.items():
        pl = plan_for(nz, **kw)
        t = timeit(lambda: pl.forward(iq, out=out, want_mask=False))
        res[name] = {"ms": t * 1e3, "updates_per_s": upd / t, "vs_plain": (upd / t) / res["plain"]["updates_per_s"]}
        del pl
        torch.cuda.empty_cache()
    del laws
    slab = plan_for(64, ny=32)
    n_out = torch.empty((n_ch, 1, n_tx, slab.n_pix), dtype=torch.complex64, device=dev)
    t = timeit(lambda: slab.forward(iq, compound=False, nosum=True, out=n_out, want_mask=False))
    u = slab.n_pix * n_tx * n_ch
    res["nosum_slab"] = {"ms": t * 1e3, "updates_per_s": u / t, "voxels": slab.n_pix,
                         "output_bytes": int(n_out.numel() * 8)}
    res["workload"] = ("cfg1 half 0 (128 lines, 16 tilts, 2048 samples, Lanczos a=4, FP32, 1 frame); "
                       "laws on 128x128x256 voxels, NoSum on a 64-deep 128x32 slab (262144 voxels)")
    return res


def sim_case(n_elem=128, n_sub=128, n_tilt=21, n_scat=500, n_sample=1024, fc=7.5e6, seed=4242):
    """SIMUS-RCA workload of one half of the dataset generator
    (save_simus_sh.m:44-110, 193-225): 128 row elements transmit and 128
    column elements receive, each element a line of 128 sub-elements across
    the aperture, 21 plane-wave tilts -10..10 deg, 350-500 scatterers per
    event (here 500), 1024 samples in BS100BW mode (fs = 4 fc), flat
    elevation, 0.5 dB/cm/MHz."""
    import numpy as np
    c = 1540.0
    pitch = 0.2e-3
    xe = (np.arange(n_elem) - (n_elem - 1) / 2) * pitch
    ys = (np.arange(n_sub) - (n_sub - 1) / 2) * pitch * n_elem / n_sub
    t_ir = np.arange(33) / 250e6
    ir = np.sin(2 * np.pi * fc * t_ir) * np.hanning(t_ir.size)

    def trans(row):
        pos = np.zeros((n_elem, 5, n_sub), np.float32)
        for u in range(n_sub):
            pos[:, 0 if row else 1, u] = xe
            pos[:, 1 if row else 0, u] = ys[u]
        return {"frequency": fc, "ElementPos": pos, "ElementWidth": 0.18e-3,
                "ElementHeight": n_elem * pitch, "ElementElevationFocus": 10000.0, "baffle": 0,
                "lensCorrection": 0.0, "impulseResponse": ir}

    tp = np.arange(int(2 / fc * 250e6)) / 250e6
    pulse = np.sin(2 * np.pi * fc * tp) * np.hanning(tp.size)
    th = np.deg2rad(np.linspace(-10, 10, n_tilt))
    TXSI = [{"pulseShape": pulse, "Delay": xe * np.sin(t) / c - (xe * np.sin(t) / c).min(),
             "Apod": np.ones(n_elem)} for t in th]
    lam = c / fc
    RXSI = [{"samplingFrequency": 4 * fc, "demodFrequency": fc, "sampleMode": "BS100BW",
             "startTime": 2 * 10 * lam / c, "numelSample": n_sample, "Apod": np.ones(n_elem),
             "InputFilter": np.array([0.1, 0.25, 0.3, 0.25, 0.1])} for _ in th]
    rng = np.random.default_rng(seed)
    span = n_elem * pitch
    pos = np.stack([rng.uniform(-span / 3, span / 3, n_scat), rng.uniform(-span / 3, span / 3, n_scat),
                    rng.uniform(25 * lam, 85 * lam, n_scat), rng.standard_normal(n_scat)], axis=1)
    pos = np.repeat(pos[:, :, None], n_tilt, axis=2).astype(np.float32)   # one frame
    Media = {"speedOfSound": c, "attenuation": 0.5, "scattererPos": pos}
    return Media, trans(True), trans(False), TXSI, RXSI


def run_sim(dev, reps=2):
    """One frame of the SIMUS-RCA generator (sim_case): transmit field, fused
    receive + sum over scatterers, inverse FFT and decimation.  Throughput in
    field terms (frequency x scatterer x element x sub-element evaluations,
    transmit plus receive) per second."""
    import torch
    from simu_rca_ab_amd.simulator import RcaSimulator
    Media, TT, TR, TXSI, RXSI = sim_case()
    sim = RcaSimulator(Media, TT, TR, TXSI, RXSI, device=dev)
    pos = torch.from_numpy(Media["scattererPos"].transpose(2, 1, 0).copy()).to(dev)
    raw = torch.empty((sim.n_tx, sim.n_ele_rx, 2 * sim.n_out), dtype=torch.float32, device=dev)
    sim.frame(pos, raw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.frame(pos, raw)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    nf, n_scat = sim.P["nf"], pos.shape[2]
    terms = sim.n_tx * nf * n_scat * sim.n_sub * (sim.n_ele_tx + sim.n_ele_rx)
    return {"ms_per_frame": dt * 1e3, "field_terms_per_s": terms / dt, "terms_per_frame": terms,
            "frequencies": nf, "nfft": sim.P["nfft"], "scatterers": int(n_scat),
            "workload": "SIMUS RCA half-frame: 128 row (tx) + 128 column (rx) elements x 128 "
                        "sub-elements, 21 tilts, 500 scatterers, 1024 samples BS100BW at 7.5 MHz, "
                        "flat elevation"}


def run_serving(dev, capacity, n_req=16, load=0.7, seed=3000):
    """Serving stream as config 3 states it: Poisson arrivals of requests on
    the resident 128+128-element probe (2048 samples, Lanczos a=4), each with
    grid N_xy ~ U{64..160}, N_z = 2 N_xy, tilts from {8, 16, 32, 64}, frames
    ~ U{1..100} and FP32 or FP16 I/Q (even odds), its own seeded scatterers;
    offered at ``load`` times ``capacity`` (updates/s).  The requests' I/Q is
    synthesised in HBM before the stream starts (as the on-GPU simulator
    would deliver it); latency is arrival -> beamformed volume in host
    memory."""
    import numpy as np
    import torch
    from simu_rca_ab_amd.serving import GpuBackend, RcaServer, RequestSpec, synth_request
    rng = np.random.Generator(np.random.PCG64(seed))
    be = GpuBackend(dev)
    specs = []
    for _ in range(n_req):
        nxy = int(rng.integers(64, 161))
        specs.append(RequestSpec(nxy, nxy, 2 * nxy, int(rng.choice([8, 16, 32, 64])),
                                 int(rng.integers(1, 101)), "fp16" if rng.random() < 0.5 else "fp32"))
    log(f"[bench] serving: synthesising {n_req} requests in HBM")
    reqs = [synth_request(i, sp, be, dev, n_scat=100, host=False) for i, sp in enumerate(specs)]
    torch.cuda.synchronize()
    upd = [sp.updates(be.n_elem) for sp in specs]
    mean_upd = float(np.mean(upd))
    rate = load * capacity / mean_upd
    gaps = rng.exponential(1.0 / rate, n_req)
    srv = RcaServer(be, budget_bytes=192 << 30).start()
    log(f"[bench] serving: {n_req} requests at {rate:.3f} req/s (offered load {load:.0%} of "
        f"{capacity / 1e9:.0f} G updates/s)")
    t = time.perf_counter()
    for i, (r, g) in enumerate(zip(reqs, gaps)):
        t += g
        while time.perf_counter() < t:
            time.sleep(min(1e-3, max(0.0, t - time.perf_counter())))
        srv.submit(r)
        if i % 4 == 3:
            log(f"[bench] serving: {i + 1} submitted")
    for r in reqs:
        r.wait(900)
    srv.close(60)
    st = srv.stats.summary()
    st.update({"arrival_rate_per_s": rate, "offered_load": load, "capacity_basis": capacity,
               "n_requests": n_req, "frames_per_request": "1-100 (uniform)",
               "grid": "N_xy ~ U{64..160}, N_z = 2 N_xy", "tilts": "{8, 16, 32, 64}",
               "mean_updates_per_request": mean_upd, "max_updates_per_request": float(max(upd)),
               "fp16_fraction": sum(sp.dtype == "fp16" for sp in specs) / n_req,
               "iq": "resident in HBM at arrival",
               "latency": "arrival -> beamformed volume in host memory"})
    del reqs
    return st


def run_serving_saturated(dev, capacity, n_req=12, seed=3100, frames=(4, 16), host=True,
                          compare=True):
    """FP16-I/Q serving at saturation: requests drawn like the serving
    workload (grid N_xy ~ U{64..160}, N_z = 2 N_xy, tilts in {8,16,32,64}),
    ``frames`` (inclusive range) frames each, all FP16, all submitted at once
    so the queue never runs dry; the sustained rate is compared with the
    plain FP32 kernel's.  I/Q in pinned host memory (uploaded by the server)
    or, with host=False, resident in HBM at arrival.  With ``compare`` one
    request is then served again with its complex64 I/Q, and its FP16 result
    is compared with that FP32 result."""
    import numpy as np
    import torch
    from simu_rca_ab_amd.serving import GpuBackend, RcaServer, RequestSpec, synth_request
    rng = np.random.Generator(np.random.PCG64(seed))
    be = GpuBackend(dev)
    specs = []
    for _ in range(n_req):
        nxy = int(rng.integers(64, 161))
        specs.append(RequestSpec(nxy, nxy, 2 * nxy, int(rng.choice([8, 16, 32, 64])),
                                 int(rng.integers(frames[0], frames[1] + 1)), "fp16"))
    log(f"[bench] saturated FP16 serving ({frames[0]}-{frames[1]} frames): synthesising {n_req} requests")
    reqs = [synth_request(i, sp, be, dev, n_scat=100, host=host) for i, sp in enumerate(specs)]
    torch.cuda.synchronize()
    srv = RcaServer(be, budget_bytes=(48 if host else 160) << 30).start()
    log(f"[bench] saturated FP16 serving: {n_req} requests submitted at once")
    for r in reqs:
        srv.submit(r)
    for r in reqs:
        r.wait(600)
    srv.close(60)
    st = srv.stats.summary()
    st.update({"plain_fp32_updates_per_s": capacity, "vs_plain_fp32": st["updates_per_s"] / capacity,
               "n_requests": n_req, "frames_per_request": f"{frames[0]}-{frames[1]}", "dtype": "fp16",
               "arrival": "all at t=0 (saturation)",
               "iq": "pinned host memory, uploaded by the server" if host else "resident in HBM at arrival",
               "latency": "arrival -> beamformed volume in host memory"})
    if not compare:
        del reqs
        return st
    s0 = specs[0]
    twin = synth_request(0, RequestSpec(s0.nx, s0.ny, s0.nz, s0.n_tilt, s0.n_frame, "fp32"), be, dev,
                         n_scat=100)
    srv2 = RcaServer(be, budget_bytes=48 << 30).start()
    srv2.submit(twin)
    twin.wait(600)
    srv2.close(60)
    a, b = reqs[0].result, twin.result
    rel = float((a - b).abs().norm() / b.abs().norm())
    st.update({"fp16_vs_fp32_rel_l2": rel, "rel_l2_request": f"request 0 ({s0.nx}x{s0.ny}x{s0.nz}, "
                                                          f"{s0.n_tilt} tilts, {s0.n_frame} frames)"})
    return st


def run_dbf(dev, steps=3):
    """Differentiable beamforming (96+96 elements, 16 tilts, 96x96x192 voxels,
    8 frames, Lanczos a=3) feeding a 4-layer complex 3-D CNN (32 channels)."""
    import torch
    from simu_rca_ab_amd import configs as C, dbf, synth
    cfg = C.get("cfg4")
    plan = build_plans(cfg, dev)[0]
    one = synth.frame(cfg, device=dev)[0]
    ph = torch.exp(1j * torch.linspace(0, 2, cfg.n_frame, device=dev)).to(torch.complex64)
    iq = (one * ph[:, None, None, None]).contiguous()
    r = dbf.time_cfg4(plan, iq, (cfg.nx, cfg.ny, cfg.nz), steps=steps, log=log)
    upd = r["updates_per_pass"]
    r["das_forward_updates_per_s"] = upd / (r["das_forward_ms"] * 1e-3)
    r["das_adjoint_updates_per_s"] = upd / (r["das_adjoint_ms"] * 1e-3)
    r["workload"] = ("cfg4: 96+96 elements, 16 tilts -10..10 deg, 1024 samples, 96x96x192 voxels, "
                     "8 frames, Lanczos a=3 -> 4 complex Conv3d (32 ch, 3^3, CReLU) -> MSE")
    return r


def cpu_baseline(cfg, target_s=6.0):
    """FP32 CPU restatement of the reference kernel on a voxel sample of the
    workload, OpenMP over every core this process may use."""
    import numpy as np
    from oracle import rca_oracle as O
    from simu_rca_ab_amd import configs as C, synth
    cores = len(os.sched_getaffinity(0))
    iqs = synth.frame(cfg)
    grid = C.meshgrid_zxy(*C.grid_axes(cfg))
    flat = [g.ravel(order="F").astype(np.float32) for g in grid]

    def run(nvox):
        t = 0.0
        sel = np.linspace(0, cfg.n_vox - 1, nvox).astype(np.int64)
        gs = tuple(np.ascontiguousarray(f[sel]) for f in flat)
        for hi, h in enumerate(cfg.halves):
            G, T = C.geometry(cfg, h, grid), C.transmit(h)
            prep = O.prepare(G, T)
            params = dict(receive_x=h.receive_x, fs=cfg.fs, fd=cfg.fd, c=cfg.c, t0=cfg.t0,
                          fnumber=cfg.fnumber, n_sample=cfg.n_sample)
            d = iqs[hi].numpy().view(np.float32)
            t1 = time.perf_counter()
            O.kernel_forward(d, gs, prep, params, O.INTERP_LANCZOS, 4, precision=32,
                             nthreads=cores)
            t += time.perf_counter() - t1
        return t

    probe = 2048
    run(probe)                                   # warm-up (page-in, thread pool)
    tp = run(probe)
    nvox = int(min(cfg.n_vox, max(probe, probe * target_s / max(tp, 1e-3) / 5)))
    times = [run(nvox) for _ in range(5)]
    t = statistics.median(times)
    upd = nvox * cfg.n_tilt * cfg.n_elem
    # one-core figure on a smaller sample (about 1 s per run)
    cores_all = cores
    cores = 1
    t1p = run(256)
    nv1 = int(max(256, 256 * 1.0 / max(t1p, 1e-3)))
    t1 = statistics.median([run(nv1) for _ in range(3)])
    cores = cores_all
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": upd / t, "unit": "voxel*tilt*element updates/s", "cores": cores,
            "kind": "port", "cpu_model": model,
            "one_core_value": nv1 * cfg.n_tilt * cfg.n_elem / t1,
            "sample": f"{nvox} of {cfg.n_vox} voxels (evenly spaced), all 32 tilts x 128 elements, "
                      f"1 frame; FP32 CPU restatement of bfVirtualSourceGeneric_RCA.cu, OpenMP "
                      f"{cores} threads, median of 5 runs after a warm-up ({t:.2f} s); one-core "
                      f"figure on {nv1} voxels, median of 3"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp16", action="store_true", help="skip the FP16-I/Q leg")
    ap.add_argument("--no-cfg2", action="store_true", help="skip the 200-frame config-2 leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-stream leg")
    ap.add_argument("--no-svd", action="store_true", help="skip the SVD clutter-filter leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint (cfg4) leg")
    ap.add_argument("--no-sim", action="store_true", help="skip the SIMUS-RCA simulator leg")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the NoSum / aberration-law variants leg")
    ap.add_argument("--no-dbf", action="store_true",
                    help="skip the differentiable-beamforming leg (config 4 DAS + complex CNN "
                         "train steps; its first run compiles MIOpen Conv3d kernels)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 512 MiB device copy (HBM-counter calibration)")
    ap.add_argument("--valu-probe", action="store_true",
                    help="only run the VALU FMA-chain probe (VALU-counter calibration)")
    args = ap.parse_args()
    if args.valu_probe:
        import torch
        t, ms = valu_probe(torch.device("cuda", 0), reps=2)
        print(json.dumps({"valu_probe_t_lane_fma_per_s": t, "ms": ms}))
        return
    if args.copy_probe:
        import torch
        n = 512 << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n, "copy_probe_bytes_written": n}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build runs on a single MI355X; "
              "multi-GPU sharding (frames/voxels across GPUs) is designed, not built.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from simu_rca_ab_amd import configs as C
    cfg = C.get("cfg1")
    dev = torch.device("cuda", 0)
    dt, launch_ms, plans, iqs, out = run_gpu(cfg, args.steps, args.warmup, dev)
    ms_step = dt * 1e3 / args.steps
    upd_step = cfg.updates_per_frame()
    value = upd_step / (dt / args.steps)
    mean_launch = statistics.mean(launch_ms)
    upd_launch = upd_step / len(cfg.halves)
    log(f"[bench] {value / 1e9:.2f} G updates/s, {ms_step:.2f} ms/step, "
        f"launch mean {mean_launch:.2f} ms")
    del plans, iqs, out
    torch.cuda.empty_cache()
    probe_t, probe_ms = valu_probe(dev, mode=1)
    probe_t0, _ = valu_probe(dev, mode=0)
    log(f"[bench] VALU probe: {probe_t:.1f} T lane-ops/s (one VGPR source), {probe_t0:.1f} "
        f"(two VGPR sources)")
    roof = roofline(cfg, mean_launch, upd_launch, probe_t)
    roof["ceiling_two_vgpr_sources"] = probe_t0
    extra = {}
    ens_rate = None
    if not args.no_fp16:
        log("[bench] FP16-I/Q ensemble leg (16 frames)")
        r = run_fp16(cfg, dev)
        extra["fp16_iq"] = {"value": r["fp16"], "fp32_same_ensemble": r["fp32"],
                            "unit": "voxel*tilt*element updates/s", "frames": 16,
                            "vs_plain_fp32_single_frame": r["fp16"] / value,
                            "fp16_over_fp32": r["fp16"] / r["fp32"]}
        ens_rate = 0.5 * (r["fp16"] + r["fp32"])
        torch.cuda.empty_cache()
    if not args.no_cfg2:
        log("[bench] config-2 leg (200-frame ensemble, FP16 and FP32)")
        extra["cfg2_ensemble"] = run_cfg2(dev)
        torch.cuda.empty_cache()
    if not args.no_adjoint:
        log("[bench] adjoint leg (config 4)")
        extra["adjoint"] = run_adjoint(dev)
        torch.cuda.empty_cache()
    if not args.no_svd:
        log("[bench] SVD clutter-filter leg (128x128x256 x 200 frames)")
        extra["svd_filter"] = run_svd(dev)
        torch.cuda.empty_cache()
    if not args.no_variants:
        log("[bench] NoSum / aberration-law variants leg")
        extra["variants"] = run_variants(cfg, dev)
        torch.cuda.empty_cache()
    if not args.no_sim:
        log("[bench] SIMUS-RCA simulator leg (one half-frame)")
        extra["simulator"] = run_sim(dev)
        torch.cuda.empty_cache()
    if not args.no_dbf:
        log("[bench] differentiable beamforming leg")
        extra["dbf"] = run_dbf(dev)
        torch.cuda.empty_cache()
    if not args.no_serving:
        log("[bench] serving leg (config 3)")
        extra["serving"] = run_serving(dev, ens_rate or value)
        torch.cuda.empty_cache()
        # the FP16 serving mode at saturation on the serving workload's own
        # frame counts (1-100 per request), I/Q resident at arrival
        log("[bench] saturated FP16 serving leg (frames 1-100)")
        extra["serving_fp16_saturated"] = run_serving_saturated(
            dev, value, n_req=6, seed=3200, frames=(1, 100), host=False, compare=False)
        torch.cuda.empty_cache()
        # small requests (4-16 frames) uploaded from pinned host memory, and
        # FP16 vs FP32 on one of them
        log("[bench] saturated FP16 serving leg (small requests, host I/Q)")
        extra["serving_fp16_saturated_small_requests"] = run_serving_saturated(dev, value)
        torch.cuda.empty_cache()
    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline leg")
        cpu = cpu_baseline(cfg)
        cpu["gpu_speedup"] = value / cpu["value"]
    line = {
        "metric": "voxel*tilt*element updates/s (FP32 drop-in, 1 frame)",
        "value": value, "unit": "voxel*tilt*element updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded analytic point scatterers + 20 dB noise)",
        "config": {"workload": "cfg1: RCA 128 row + 128 column elements, 2 x 16 tilts -12..12 deg, "
                               "2048-sample complex64 I/Q, 128x128x256 voxels, Lanczos a=4, 1 frame",
                   "voxels": cfg.n_vox, "tilts": cfg.n_tilt, "elements": cfg.n_elem,
                   "samples": cfg.n_sample, "interp": cfg.interpolation, "frames": 1},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
