"""Scheduler / admission logic of the serving loop on the CPU, through a fake
backend with the GPU backend's interface (no GPU work)."""
import time

import numpy as np
import pytest

from simu_rca_ab_amd.serving import Request, RequestSpec, RcaServer


class _Ev:
    def __init__(self, t_ready):
        self.t = t_ready

    def query(self):
        return time.perf_counter() >= self.t

    def synchronize(self):
        while not self.query():
            time.sleep(1e-4)


class FakeBackend:
    """Executes batches 'instantly' after a simulated delay proportional to work;
    the result of frame f of request r is r*1000 + f in every voxel."""
    n_elem = 4

    def __init__(self, upd_per_s=1e9, fail_on_batch=None, fail_upload_of=None):
        self.rate = upd_per_s
        self.busy_until = 0.0
        self.batches = []
        self.fail_on_batch = fail_on_batch
        self.fail_upload_of = fail_upload_of
        self.live = {}

    def request_bytes(self, spec):
        return spec.n_vox * spec.n_frame * 8

    def upload(self, req):
        if req.rid == self.fail_upload_of:
            raise RuntimeError("injected out of memory")
        req.dev = {"out": np.zeros((req.spec.n_frame, req.spec.n_vox), np.complex64)}
        self.live[req.rid] = True
        return _Ev(time.perf_counter() + 1e-3)

    def run_batch(self, k, work):
        if self.fail_on_batch is not None and len(self.batches) == self.fail_on_batch:
            raise RuntimeError("injected launch failure")
        self.batches.append([(r.rid, f0, f1) for r, f0, f1 in work])
        upd = 0
        for r, f0, f1 in work:
            assert self.live.get(r.rid), "work issued for a released request"
            for f in range(f0, f1):
                r.dev["out"][f] = r.rid * 1000 + f
            upd += r.spec.updates(self.n_elem) // r.spec.n_frame * (f1 - f0)
        start = max(self.busy_until, time.perf_counter())
        self.busy_until = start + upd / self.rate
        return _Ev(self.busy_until), 0

    def download(self, req, after):
        return req.dev["out"].copy(), _Ev(after.t + 1e-4)

    def release(self, req):
        self.live[req.rid] = False
        req.dev = {}

    def max_frames_per_chunk(self, spec, quantum):
        per = spec.updates(self.n_elem) // spec.n_frame
        return max(1, quantum // per)


def _req(rid, nvox=1000, frames=4, tilts=8):
    return Request(rid, RequestSpec(nvox, 1, 1, tilts, frames), halves=[None, None])


def test_all_requests_complete_with_their_own_results():
    be = FakeBackend()
    srv = RcaServer(be, budget_bytes=10**9, quantum=10**6).start()
    reqs = [srv.submit(_req(i, nvox=500 + 100 * i, frames=1 + i % 5)) for i in range(12)]
    for r in reqs:
        out = r.wait(10)
        want = r.rid * 1000 + np.arange(r.spec.n_frame)[:, None]
        assert np.array_equal(out.real, np.broadcast_to(want, out.shape))
    srv.drain(5)
    srv.close(5)
    s = srv.stats.summary()
    assert s["requests"] == 12 and s["rejected"] == 0 and s["p99_ms"] >= s["p50_ms"] > 0


def test_admission_respects_budget_and_order():
    be = FakeBackend(upd_per_s=2e7)
    spec_bytes = 1000 * 4 * 8
    srv = RcaServer(be, budget_bytes=2 * spec_bytes, quantum=10**9).start()
    reqs = [srv.submit(_req(i)) for i in range(5)]
    for r in reqs:
        r.wait(20)
    srv.close(5)
    assert srv.stats.max_bytes <= 2 * spec_bytes
    assert srv.stats.max_in_flight <= 2
    adm = [r.admitted for r in reqs]
    assert adm == sorted(adm)                       # first come, first served
    assert reqs[2].admitted >= min(reqs[0].done, reqs[1].done) - 1e-3


def test_oversized_request_is_rejected():
    be = FakeBackend()
    srv = RcaServer(be, budget_bytes=1000, quantum=10**6).start()
    big = srv.submit(_req(0, nvox=10**5))
    small = srv.submit(_req(1, nvox=10, frames=1))
    with pytest.raises(RuntimeError, match="budget"):
        big.wait(5)
    assert small.wait(5) is not None
    srv.close(5)
    assert srv.stats.rejected == 1


def test_quantum_splits_large_requests_and_interleaves():
    be = FakeBackend(upd_per_s=1e6)          # 32 ms of simulated GPU time per frame
    per_frame = 1000 * 8 * 4
    srv = RcaServer(be, budget_bytes=10**9, quantum=2 * per_frame).start()
    a = srv.submit(_req(0, frames=8))
    time.sleep(0.02)
    b = srv.submit(_req(1, frames=2))
    a.wait(10), b.wait(10)
    srv.close(5)
    for batch in be.batches:
        frames = sum(f1 - f0 for _, f0, f1 in batch)
        assert frames <= 2
    # the late request finished before the long one: it joined a batch mid-way
    assert b.done < a.done


def test_backend_failure_reaches_waiters():
    """A failing launch fails every pending request, stops the scheduler
    without an unhandled thread exception, and later submits are refused."""
    be = FakeBackend(fail_on_batch=0)
    srv = RcaServer(be, budget_bytes=10**9, quantum=10**6).start()
    r = srv.submit(_req(0))
    with pytest.raises(RuntimeError, match="server failed"):
        r.wait(5)
    srv._thread.join(5)
    assert not srv._thread.is_alive()
    assert "injected launch failure" in srv.error
    with pytest.raises(RuntimeError, match="server failed"):
        srv.submit(_req(1))
    srv.close(5)


def test_upload_failure_fails_that_request_only():
    """An upload that raises (out of device memory) fails its own request and
    returns its budget; the server goes on with the others."""
    be = FakeBackend(fail_upload_of=1)
    srv = RcaServer(be, budget_bytes=10**9, quantum=10**6).start()
    reqs = [srv.submit(_req(i, frames=2)) for i in range(3)]
    with pytest.raises(RuntimeError, match="injected out of memory"):
        reqs[1].wait(5)
    for i in (0, 2):
        out = reqs[i].wait(5)
        assert np.all(out[1] == i * 1000 + 1)
    srv.drain(5)
    assert srv.in_use == 0 and srv.error is None
    srv.close(5)


def test_wait_timeout_raises():
    be = FakeBackend(upd_per_s=1e4)           # 3.2 s of simulated GPU time per frame
    srv = RcaServer(be, budget_bytes=10**9, quantum=10**9).start()
    r = srv.submit(_req(0, frames=1))
    with pytest.raises(TimeoutError):
        r.wait(0.05)
    srv.close(10)


def test_submit_after_close_raises():
    srv = RcaServer(FakeBackend(), budget_bytes=10**9, quantum=10**6).start()
    assert srv.submit(_req(0, frames=1)).wait(5) is not None
    srv.close(5)
    with pytest.raises(RuntimeError, match="closed"):
        srv.submit(_req(1))


def test_close_fails_queued_and_unfinished_requests():
    """close() lets in-flight batches finish and fails the rest: every waiter
    returns (a result or RuntimeError), none blocks."""
    be = FakeBackend(upd_per_s=2e6)           # 16 ms per frame
    per_frame = 1000 * 8 * 4
    budget = 1000 * 16 * 8                    # one 16-frame request at a time
    srv = RcaServer(be, budget_bytes=budget, quantum=2 * per_frame).start()
    reqs = [srv.submit(_req(i, frames=16)) for i in range(3)]
    time.sleep(0.05)
    srv.close(10)
    assert not srv._thread.is_alive()
    failed = 0
    for r in reqs:
        assert r.event.is_set()
        try:
            r.wait(0)
        except RuntimeError as e:
            assert "closed" in str(e)
            failed += 1
    assert failed >= 2                        # the queued ones at least
    assert not srv.queue and not srv.active


def test_gpu_backend_chunks_start_on_frame_blocks():
    """GpuBackend.max_frames_per_chunk (host logic only): chunks are a
    multiple of the batched kernels' frame block of the request (by dtype and
    depth tiling), at least one block; within the quantum when a block fits
    it, and within the work-list capacity (work items of 16x16-voxel tiles
    with depth tiles, 256 voxels without)."""
    import math
    from types import SimpleNamespace
    from simu_rca_ab_amd.serving import GpuBackend
    for fbt in (4, 8):
        fb = {("fp32", False): 1, ("fp16", False): 4, ("fp32", True): fbt, ("fp16", True): fbt}
        be = SimpleNamespace(n_elem=128, max_items=1 << 17, fb=fb)
        for nxy, nz, nt, nf in ((64, 128, 8, 100), (100, 200, 16, 40), (160, 320, 64, 7),
                                (128, 256, 32, 8), (24, 40, 8, 9), (117, 234, 32, 100)):
            for dt in ("fp32", "fp16"):
                sp = RequestSpec(nxy, nxy, nz, nt, nf, dt)
                q = int(1.5e10)
                f = GpuBackend.max_frames_per_chunk(be, sp, q)
                per = sp.updates(be.n_elem) // sp.n_frame
                b = fb[(dt, nz >= 64)]
                assert f >= b and f % b == 0
                if per * b <= q:
                    assert f * per <= q
                items = (math.ceil(nxy * nxy / 16) * math.ceil(nz / 16) if nz >= 64
                         else math.ceil(sp.n_vox / 256))
                assert (f // b) * items <= be.max_items or f == b
