"""Window cover of the depth-tiled kernels, checked on the CPU.

The forward and adjoint kernels stage per (tilt chunk, receive line) only the
sample window their workgroup's voxels can reach, and read each update's taps
from it; a tap outside the staged rows would be clamped to the row edge (LDS
index kept in bounds), a wrong value with no error.  ``window_check`` walks
every workgroup, tilt chunk, line and pass of a launch with the kernels' own
tile, bound, window and pass code (``rca_tile.h``) and counts the in-range
updates whose taps are not inside the rows of the one pass that takes them,
and the skipped lines that a voxel of the tile would have used.  Both must be
zero, for every LDS layout the launches use (FP32 single-frame and
frame-block, FP16 frame-block, the serving batch kernels, the adjoint), at the
full config-0/1/4 sizes, at serving shapes, with the three aberration-law
layouts, with windows longer than the rows (several passes), and without the
depth hint (whole-trace rows).  A window narrowed by one sample must fail.
"""
import numpy as np
import pytest

from simu_rca_ab_amd import configs as C
from tests import window_check as W

KEYS = ("blocks", "lines", "lines_skipped", "updates_in_range", "violations", "skipped_in_aperture",
        "multi_pass_lines", "max_passes", "tc", "wcap")


def _ok(r, what, passes=None):
    print(what, {k: r[k] for k in KEYS}, flush=True)
    assert r["updates_in_range"] > 0, what
    assert r["violations"] == 0, (what, r)
    assert r["skipped_in_aperture"] == 0, (what, r)
    if passes:
        assert r["max_passes"] >= passes, (what, r)


# the launch layouts: forward FP32 one frame (the drop-in), forward FP32 and
# FP16 frame blocks of 8 (ensembles), the serving batch kernels (FP32 one
# frame at 32 KiB, FP16 blocks of 4 at 48 KiB), the adjoint (4 frames x 4
# voxels per lane, the default; 2 x 2 for fewer frames; one voxel per lane;
# the line-outer kernel's windows: one tilt per window, 192-sample rows, the
# layout the 16 KiB budget gives)
LAYOUTS = {"fwd_f32": dict(), "fwd_f32_fb8": dict(fb=8), "fwd_f16_fb8": dict(sbytes=4, fb=8),
           "batch_f16_fb4": dict(sbytes=4, fb=4, lds_kib=48),
           "adjoint": dict(kernel="adjoint", fb=4, v=4), "adjoint_fb2_v2": dict(kernel="adjoint", fb=2, v=2),
           "adjoint_v1": dict(kernel="adjoint"),
           "adjoint_line_outer": dict(kernel="adjoint", fb=4, v=4, lds_kib=16)}


@pytest.mark.parametrize("interp", ["quadratic", "lanczos", "lanczos3", "lanczos4"])
@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_cfg0_full(interp, layout):
    _ok(W.check(C.get("cfg0"), 0, interp=interp, **LAYOUTS[layout]), f"cfg0 {interp} {layout}")


@pytest.mark.parametrize("half", [0, 1])
def test_cfg1_full_drop_in_layout(half):
    """128x128x256 voxels, 16 tilts x 128 lines, 2048 samples: every workgroup
    of the drop-in / bench launch (7.6e9 in-range updates per half)."""
    _ok(W.check(C.get("cfg1"), half), f"cfg1 half {half} drop-in")


def _sample_blocks(nblk, ndc, n_rand, seed):
    """Workgroups 0 .. ndc-1 and the last ndc (first / last column groups,
    every depth tile), plus a seeded random subset."""
    rng = np.random.default_rng(seed)
    b = np.concatenate([np.arange(ndc), np.arange(nblk - ndc, nblk),
                        rng.choice(nblk, size=min(n_rand, nblk), replace=False)])
    return np.unique(b)


@pytest.mark.parametrize("layout", ["fwd_f32_fb8", "fwd_f16_fb8", "batch_f16_fb4", "adjoint",
                                    "adjoint_line_outer"])
def test_cfg1_other_layouts(layout):
    cfg = C.get("cfg1")
    kw = LAYOUTS[layout]
    if layout.startswith("adjoint"):      # 64 depths x 16 columns
        nblk, ndc = (128 * 128 // 16) * 4, 4
    else:                        # 16 depths x 16 columns
        nblk, ndc = (128 * 128 // 16) * 16, 16
    for half in (0, 1):
        _ok(W.check(cfg, half, blocks=_sample_blocks(nblk, ndc, 3000, 5 + half), **kw),
            f"cfg1 half {half} {layout} (sampled workgroups)")


@pytest.mark.parametrize("layout", ["fwd_f32", "fwd_f32_fb8", "adjoint", "adjoint_fb2_v2", "adjoint_line_outer"])
def test_cfg4_full(layout):
    """96x96x192 voxels, 16 tilts x 96 lines, 1024 samples, Lanczos a=3."""
    _ok(W.check(C.get("cfg4"), 0, **LAYOUTS[layout]), f"cfg4 {layout}")


def _serving_cfg(nxy, nz, n_tilt):
    """A serving request's half on the 128-element, 2048-sample probe
    (tilts -12..12 deg, Lanczos a=4)."""
    th = np.deg2rad(np.linspace(-12, 12, n_tilt // 2))
    return C.RcaConfig(f"req{nxy}", 128, 2048, nxy, nxy, nz, [C.Half(True, th), C.Half(False, th)],
                       "lanczos4")


@pytest.mark.parametrize("nxy,nz,n_tilt", [(64, 128, 64), (100, 200, 8), (117, 234, 16),
                                           (160, 320, 64)])
@pytest.mark.parametrize("layout", ["fwd_f32", "batch_f16_fb4"])
def test_serving_shapes(nxy, nz, n_tilt, layout):
    """Serving grids 64^2x128 .. 160^2x320: column counts that are not a
    multiple of the tile width put columns of two grid rows in one tile, and
    depths that are not a multiple of 16 leave partial depth tiles."""
    cfg = _serving_cfg(nxy, nz, n_tilt)
    nblk = -(-nxy * nxy // 16) * -(-nz // 16)
    for half in (0, 1):
        _ok(W.check(cfg, half, blocks=_sample_blocks(nblk, -(-nz // 16), 2500, nxy + half),
                    **LAYOUTS[layout]), f"serving {nxy}^2x{nz} {n_tilt} tilts half {half} {layout}")


@pytest.mark.parametrize("layout", ["fwd_f32", "fwd_f16_fb8", "adjoint"])
def test_laws_widen_windows(layout):
    """Aberration laws of +-3 samples (more than any real law): per line, per
    (tilt, line) and per (line, voxel) with both sides."""
    cfg = C.get("cfg0")
    n_ch, n_tx, n_pix = cfg.n_elem, len(cfg.halves[0].theta), cfg.n_vox
    rng = np.random.default_rng(9)
    s = 3.0 / cfg.fs
    cases = {"per line": dict(ab=rng.uniform(-s, s, n_ch), ab_mode=1),
             "per (tilt, line)": dict(ab=rng.uniform(-s, s, n_tx * n_ch), ab_mode=2),
             "per (line, voxel)": dict(ab=rng.uniform(-s, s, n_ch * n_pix),
                                       abt=rng.uniform(-s, s, n_ch * n_pix), ab_mode=3)}
    for name, kw in cases.items():
        _ok(W.check(cfg, 0, **kw, **LAYOUTS[layout]), f"cfg0 law {name} {layout}")


@pytest.mark.parametrize("layout", ["fwd_f32", "fwd_f32_fb8", "fwd_f16_fb8", "adjoint"])
def test_windows_longer_than_rows(layout):
    """Depths 0.6 mm apart (a tile spans 9.6 mm forward, 38 mm in the adjoint):
    windows of several rows, staged in passes."""
    cfg = C.scaled(C.get("cfg1"), nx=20, ny=12, nz=128)
    x, y, _ = C.grid_axes(cfg)
    z = 5e-3 + np.arange(cfg.nz) * 0.6e-3
    X, Y, Z = C.meshgrid_zxy(x, y, z)
    grid = tuple(a.ravel(order="F") for a in (X, Y, Z))
    for half in (0, 1):
        _ok(W.check(cfg, half, grid=grid, depth_run=cfg.nz, **LAYOUTS[layout]),
            f"depth-coarse half {half} {layout}", passes=2)


@pytest.mark.parametrize("kib", [16, 64])
def test_lds_budgets(kib):
    """Tilts per chunk set by the LDS budget (rca_set_tuning)."""
    cfg = C.scaled(C.get("cfg1"), nx=40, ny=40)
    _ok(W.check(cfg, 1, lds_kib=kib), f"cfg1 40x40x256 forward LDS {kib} KiB")
    _ok(W.check(cfg, 1, lds_kib=kib, **LAYOUTS["adjoint"]), f"cfg1 40x40x256 adjoint LDS {kib} KiB")
    _ok(W.check(cfg, 1, lds_kib=kib, **LAYOUTS["adjoint_v1"]), f"cfg1 40x40x256 adjoint V=1 LDS {kib} KiB")


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_short_traces(layout):
    """Traces of as many samples as taps hold no in-range update; 24-sample
    traces whose start time is moved to the shallowest depth hold a few, all
    inside the rows (the shapes of the GPU short-trace tests)."""
    for interp, nt in (("quadratic", 3), ("lanczos3", 6), ("lanczos4", 8)):
        for half in (0, 1):
            cfg = C.scaled(C.get("cfg1"), nx=5, ny=4, nz=64, n_sample=nt)
            r = W.check(cfg, half, interp=interp, **LAYOUTS[layout])
            print(f"{interp} half {half} {nt} samples", {k: r[k] for k in KEYS}, flush=True)
            assert r["updates_in_range"] == 0 and r["violations"] == 0
            cfg = C.scaled(C.get("cfg1"), nx=5, ny=4, nz=64, n_sample=24, t0=-6.4e-6)
            _ok(W.check(cfg, half, interp=interp, **LAYOUTS[layout]),
                f"{interp} half {half} 24 samples from 6.4 us {layout}")


def test_no_depth_hint_whole_trace_rows():
    """Without the hint a workgroup owns 256 consecutive voxels and rows hold
    whole traces from the window start."""
    for layout in ("fwd_f32", "adjoint", "adjoint_v1"):
        _ok(W.check(C.get("cfg0"), 0, depth_run=0, **LAYOUTS[layout]), f"cfg0 no hint {layout}")


def test_narrowed_window_is_caught():
    """The check fails when the windows lose one sample of margin per side."""
    r = W.check(C.get("cfg0"), 0, shrink=1)
    assert r["violations"] > 0
    r = W.check(C.get("cfg0"), 0, shrink=2, **LAYOUTS["adjoint"])
    assert r["violations"] > 0
