"""Serving loop on the GPU: every request's volume is bitwise identical alone
and in a concurrent mix, equals the plain forward path (both halves,
compounded, the same depth tiling), and FP32 requests match the FP64 oracle
to 1e-5.  Requests with Nz >= 64 run the depth-tiled batch kernels (the
serving shapes all do); shallower ones the untiled path."""
import time

import numpy as np
import pytest

from oracle import rca_oracle as O
from simu_rca_ab_amd import configs as C
from simu_rca_ab_amd.plan import RcaPlan
from simu_rca_ab_amd.serving import GpuBackend, RcaServer, RequestSpec, synth_request
from tests import helpers as H
from tests.test_gpu_forward import say

pytestmark = pytest.mark.gpu

SPECS = [RequestSpec(16, 16, 32, 8, 3, "fp32"), RequestSpec(24, 20, 40, 16, 2, "fp16"),
         RequestSpec(12, 12, 24, 8, 1, "fp32"), RequestSpec(20, 16, 36, 16, 4, "fp16"),
         RequestSpec(16, 24, 30, 8, 2, "fp32"),
         # depth-tiled: 16x16x128 and 24x16x160 (partial column groups), both dtypes
         RequestSpec(16, 16, 128, 8, 3, "fp32"), RequestSpec(24, 16, 160, 16, 5, "fp16"),
         RequestSpec(24, 16, 160, 16, 2, "fp32"), RequestSpec(16, 16, 128, 8, 4, "fp16")]


@pytest.fixture(scope="module")
def backend(dev):
    return GpuBackend(dev, n_elem=64, n_sample=1024, interp="lanczos4", max_halves=32,
                      max_items=1 << 14)


@pytest.fixture(scope="module")
def requests(backend, dev):
    return [synth_request(i, s, backend, dev, n_scat=50) for i, s in enumerate(SPECS)]


def _fresh(reqs):
    from simu_rca_ab_amd.serving import Request
    return [Request(r.rid, r.spec, r.halves) for r in reqs]


def test_alone_vs_mix_bitwise(backend, requests):
    say("case serving: each request alone, then all five in a concurrent mix")
    alone = {}
    srv = RcaServer(backend, budget_bytes=8 << 30, quantum=int(2e8)).start()
    for r in _fresh(requests):
        srv.submit(r)
        alone[r.rid] = r.wait(120).numpy().copy()
    srv.close(60)
    srv = RcaServer(backend, budget_bytes=8 << 30, quantum=int(2e8)).start()
    mix = _fresh(requests)
    for r in mix:
        srv.submit(r)
        time.sleep(0.002)
    for r in mix:
        got = r.wait(120).numpy()
        assert np.array_equal(got, alone[r.rid]), f"request {r.rid} differs alone vs mix"
    srv.close(60)
    s = srv.stats.summary()
    say(f"mix: {s}")
    assert s["requests"] == len(mix) and s["batches"] >= 2


def test_serving_matches_plain_path_and_oracle(backend, requests, dev):
    say("case serving result vs RcaPlan forward and FP64 oracle")
    srv = RcaServer(backend, budget_bytes=8 << 30, quantum=int(2e8)).start()
    reqs = _fresh(requests)
    for r in reqs:
        srv.submit(r)
    for r in reqs:
        got = r.wait(120)
        spec = r.spec
        out = None
        for h in range(2):
            d = backend.desc(spec, h)
            tab = backend._tilt_table(spec.n_tilt)[h]
            plan = RcaPlan(grid=backend._grid(spec.nx, spec.ny, spec.nz), ele=tab["ele"], vs=tab["vs"],
                           min_dist=tab["d0"], receive_x=h == 0, aperture=tab["aperture"],
                           fs=backend.fs, fd=backend.fd, c=backend.c, t0=0.0, fnumber=1.0,
                           n_sample=backend.n_sample, interp="lanczos4", device=dev,
                           depth_run=spec.nz)
            assert plan.depth_run == (spec.nz if spec.nz >= 64 else plan.depth_run)
            x = r.halves[h]
            if isinstance(x, tuple):
                o, _ = plan.forward(x[0].to(dev), x[1].to(dev), out=out, accumulate=out is not None,
                                    want_mask=False)
            else:
                o, _ = plan.forward(x.to(dev), out=out, accumulate=out is not None, want_mask=False)
            out = o
        ref = out[:, 0].cpu().numpy()
        r_rel = H.rel_l2(got.numpy(), ref)
        say(f"request {r.rid} ({spec.dtype}): rel vs plain path {r_rel:.2e}")
        # frame blocks set the tilts per chunk: the tilt-sum order may differ
        assert r_rel <= 4e-6
        if spec.dtype == "fp32":
            # every voxel of the small grids; whole workgroup tiles of the tiled ones
            sel = (np.arange(spec.n_vox) if spec.nz < 64 else
                   H.edge_tiles(spec.nx, spec.ny, spec.nz, n_random=4, seed=r.rid)[0])
            tot = 0
            for h in range(2):
                tab = backend._tilt_table(spec.n_tilt)[h]
                prep = {"ex": tab["ele"][0].cpu().numpy(), "ey": tab["ele"][1].cpu().numpy(),
                        "ez": tab["ele"][2].cpu().numpy(), "vx": tab["vs"][0].cpu().numpy(),
                        "vy": tab["vs"][1].cpu().numpy(), "vz": tab["vs"][2].cpu().numpy(),
                        "d0": tab["d0"].cpu().numpy(), "ap": np.float32(tab["aperture"])}
                grid = tuple(np.ascontiguousarray(g.cpu().numpy()[sel])
                             for g in backend._grid(spec.nx, spec.ny, spec.nz))
                prm = dict(receive_x=h == 0, fs=np.float32(backend.fs), fd=np.float32(backend.fd),
                           c=np.float32(backend.c), t0=np.float32(0), fnumber=np.float32(1),
                           n_sample=backend.n_sample)
                o, _ = O.kernel_forward(r.halves[h].numpy().view(np.float32), grid, prep, prm,
                                        O.INTERP_LANCZOS, 4)
                tot = tot + o.sum(axis=1)
            rel = H.rel_l2(got.numpy()[:, sel], tot)
            say(f"request {r.rid}: rel vs FP64 oracle {rel:.2e} ({sel.size} voxels)")
            assert rel <= 1e-5
    srv.close(60)


def test_budget_rejection_on_gpu(backend, requests):
    say("case serving: request above the HBM budget is rejected, the rest run")
    reqs = _fresh(requests[:2])
    need = backend.request_bytes(reqs[1].spec)
    srv = RcaServer(backend, budget_bytes=need - 1, quantum=int(2e8)).start()
    for r in reqs:
        srv.submit(r)
    ok = [r for r in reqs if backend.request_bytes(r.spec) <= need - 1]
    for r in ok:
        assert r.wait(120) is not None
    with pytest.raises(RuntimeError, match="budget"):
        reqs[1].wait(120)
    srv.close(60)
