"""Continuous-batching beamforming server on one GPU.

Requests are RCA frames of uneven shape: voxel grid (Nx, Ny, Nz), tilt count,
frame count and I/Q precision (FP32 complex64, or FP16 half2 with a per-trace
FP32 scale).  The server

* keeps the probe (element positions), the per-tilt-count virtual-source
  tables and the voxel grids resident on the GPU (grids cached by shape);
* admits queued requests first-come first-served whenever their device bytes
  (I/Q, output, mask) fit in the HBM budget; a request larger than the whole
  budget is rejected;
* uploads admitted requests on a copy stream while earlier batches run;
* cuts the runnable work into batches of at most ``quantum`` voxel-tilt-
  element updates, taking frame chunks from every runnable request, least
  attained service first, so requests that arrive while others run join the
  next batch ahead of long-running ones;
* launches every batch as a replay of one of two captured HIP graphs of the
  batched kernels (``rca_batch_launch``: one workgroup per work item, FP32
  and FP16 items in kernels of their own, an item's second RCA half in a
  launch after its first), whose work list is copied
  into a fixed device buffer right before the replay — two work-list slots
  keep one batch queued while the previous one runs;
* hands each finished request to a download thread, which copies its volume
  to pinned host memory and records its latency (arrival -> result in host
  memory); pinning a multi-GB result buffer takes long enough that doing it
  on the scheduler thread would leave the GPU without queued batches.

Per-voxel arithmetic does not depend on what else is in a batch, so a
request's result is bitwise identical alone and in any mix.

The scheduling logic talks to a backend object; :class:`GpuBackend` is the
real one, tests use a CPU fake with the same interface.
"""
from __future__ import annotations

import ctypes
import math
import queue
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, configs as C
from .plan import prepare_geometry


@dataclass
class RequestSpec:
    """Shape of one request: an RCA frame set with n_tilt tilts split over the
    two halves (row-transmit / column-transmit)."""
    nx: int
    ny: int
    nz: int
    n_tilt: int
    n_frame: int
    dtype: str = "fp32"        # 'fp32' | 'fp16'

    @property
    def n_vox(self):
        return self.nx * self.ny * self.nz

    def updates(self, n_elem: int) -> int:
        return self.n_vox * self.n_tilt * n_elem * self.n_frame


@dataclass
class Request:
    rid: int
    spec: RequestSpec
    halves: list                 # per half: complex64 [F, T/2, C, Ns] or (fp16 [...,2], scale)
    arrival: float = 0.0
    admitted: float = 0.0
    started: float = 0.0
    done: float = 0.0
    result: object = None        # host complex64 [F, nVox] (column-major voxel order)
    error: str | None = None
    # scheduler state
    dev: dict = field(default_factory=dict)
    frames_issued: int = 0
    frames_done: int = 0
    event: threading.Event = field(default_factory=threading.Event)

    def latency(self):
        return self.done - self.arrival

    def wait(self, timeout=None):
        """The beamformed volume; raises TimeoutError if it is not ready within
        ``timeout`` seconds and RuntimeError if the request failed (rejected,
        server closed or failed)."""
        if not self.event.wait(timeout):
            raise TimeoutError(f"request {self.rid} not finished after {timeout} s")
        if self.error:
            raise RuntimeError(self.error)
        return self.result


class ServerStats:
    def __init__(self):
        self.latencies = []
        self.updates = 0
        self.batches = 0
        self.rejected = 0
        self.t_first = None
        self.t_last = None
        self.max_in_flight = 0
        self.max_bytes = 0

    def summary(self):
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        span = (self.t_last - self.t_first) if self.t_first is not None and self.t_last else 0.0
        return {"requests": len(self.latencies), "rejected": self.rejected, "batches": self.batches,
                "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                "mean_ms": float(lat.mean() * 1e3), "updates": self.updates,
                "updates_per_s": self.updates / span if span > 0 else 0.0, "span_s": span,
                "max_requests_in_flight": self.max_in_flight, "max_device_bytes": self.max_bytes}


# ----------------------------------------------------------------------------
# GPU backend
# ----------------------------------------------------------------------------
class _RcaBatchEntry(ctypes.Structure):
    _fields_ = [("half", ctypes.POINTER(_lib.RcaDasDesc) * 2), ("n_half", ctypes.c_int),
                ("iq", ctypes.c_void_p * 2), ("trace_scale", ctypes.c_void_p * 2),
                ("iq_dtype", ctypes.c_int), ("n_frame", ctypes.c_int),
                ("frame_begin", ctypes.c_int), ("frame_end", ctypes.c_int),
                ("out", ctypes.c_void_p), ("mask", ctypes.c_void_p)]


class GpuBackend:
    """Resident geometry, HBM allocations, streams and the captured batch graphs."""

    def __init__(self, device, n_elem=128, n_sample=2048, interp="lanczos4",
                 max_halves=256, max_items=1 << 17, fs=C.FS, fd=C.F0, c=C.C_SOUND, t0=0.0,
                 fnumber=1.0, dxyz=C.LAMBDA / 2, z0=5e-3, max_tilt_deg=12.0):
        self.lib = _lib.load()
        lib = self.lib
        lib.rca_batch_workspace_bytes.restype = ctypes.c_size_t
        lib.rca_batch_workspace_bytes.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.rca_batch_encode.argtypes = [ctypes.POINTER(_RcaBatchEntry), ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                         ctypes.POINTER(ctypes.c_int)]
        lib.rca_batch_encode.restype = ctypes.c_int
        lib.rca_batch_launch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_void_p]
        lib.rca_batch_launch.restype = ctypes.c_int
        lib.rca_batch_frame_block.argtypes = [ctypes.c_int, ctypes.c_int]
        # frames run in blocks per workgroup (by dtype and depth tiling):
        # chunks start on a block
        self.fb = {(dt, tiled): int(lib.rca_batch_frame_block(code, int(tiled)))
                   for dt, code in (("fp32", _lib.IQ_C64), ("fp16", _lib.IQ_F16X2))
                   for tiled in (False, True)}
        self.device = torch.device(device)
        self.n_elem, self.n_sample, self.interp = n_elem, n_sample, interp
        self.max_halves, self.max_items = max_halves, max_items
        self.fs, self.fd, self.c, self.t0, self.fnumber = fs, fd, c, t0, fnumber
        self.dxyz, self.z0, self.max_tilt_deg = dxyz, z0, max_tilt_deg
        self.compute = torch.cuda.Stream(self.device)
        self.copy = torch.cuda.Stream(self.device)       # uploads
        self.d2h = torch.cuda.Stream(self.device)        # downloads (never delay an upload)
        self._grids = {}
        self._tilts = {}
        self._descs = {}
        self.resident_bytes = 0
        e = C.element_positions(n_elem)
        z = np.zeros_like(e)
        self.half_geo = [{"posEleX": e, "posEleY": z, "posEleZ": z},   # receive on X (columns)
                         {"posEleX": z, "posEleY": e, "posEleZ": z}]   # receive on Y (rows)
        self.ws_bytes = int(lib.rca_batch_workspace_bytes(max_halves, max_items))
        self.host_ws = [torch.empty(self.ws_bytes, dtype=torch.uint8).pin_memory() for _ in range(2)]
        self.dev_ws = [torch.zeros(self.ws_bytes, dtype=torch.uint8, device=self.device)
                       for _ in range(2)]
        self.graphs = [self._capture(k) for k in range(2)]

    # -- graphs ---------------------------------------------------------
    def _launch(self, k):
        rc = self.lib.rca_batch_launch(ctypes.c_void_p(self.dev_ws[k].data_ptr()), self.max_halves,
                                       self.max_items, _lib.INTERP[self.interp], self.n_sample,
                                       ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream))
        _lib.check(rc)

    def _capture(self, k):
        with torch.cuda.stream(self.compute):
            self._launch(k)                 # warm-up: empty work list (header = 0)
        self.compute.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=self.compute):
            self._launch(k)
        return g

    # -- resident geometry ----------------------------------------------
    def _grid(self, nx, ny, nz):
        key = (nx, ny, nz)
        if key not in self._grids:
            x = (np.arange(nx) - (nx - 1) / 2.0) * self.dxyz
            y = (np.arange(ny) - (ny - 1) / 2.0) * self.dxyz
            zz = self.z0 + np.arange(nz) * self.dxyz
            X, Y, Z = C.meshgrid_zxy(x, y, zz)
            g = [torch.from_numpy(a.ravel(order="F").astype(np.float32)).to(self.device)
                 for a in (X, Y, Z)]
            self._grids[key] = g
            self.resident_bytes += 12 * nx * ny * nz
        return self._grids[key]

    def _tilt_table(self, n_tilt):
        if n_tilt not in self._tilts:
            per = n_tilt // 2
            th = np.deg2rad(np.linspace(-self.max_tilt_deg, self.max_tilt_deg, per))
            zero = np.zeros_like(th)
            tabs = []
            for h, T in enumerate(({"thetaX": zero, "thetaY": th}, {"thetaX": th, "thetaY": zero})):
                prep = prepare_geometry(self.half_geo[h], T)
                tabs.append({"ele": [torch.from_numpy(a).to(self.device) for a in prep["ele"]],
                             "vs": [torch.from_numpy(a).to(self.device) for a in prep["vs"]],
                             "d0": torch.from_numpy(prep["d0"]).to(self.device),
                             "aperture": prep["aperture"]})
            self._tilts[n_tilt] = tabs
        return self._tilts[n_tilt]

    def desc(self, spec: RequestSpec, h: int):
        key = (spec.nx, spec.ny, spec.nz, spec.n_tilt, h)
        if key not in self._descs:
            g = self._grid(spec.nx, spec.ny, spec.nz)
            tab = self._tilt_table(spec.n_tilt)[h]
            d = _lib.RcaDasDesc()
            d.grid_x, d.grid_y, d.grid_z = (t.data_ptr() for t in g)
            d.ele_x, d.ele_y, d.ele_z = (t.data_ptr() for t in tab["ele"])
            d.vs_x, d.vs_y, d.vs_z = (t.data_ptr() for t in tab["vs"])
            d.min_dist = tab["d0"].data_ptr()
            d.ab_law_rx, d.ab_law_mode = None, 0
            d.n_pix, d.n_channel, d.n_tx, d.n_sample = spec.n_vox, self.n_elem, spec.n_tilt // 2, self.n_sample
            d.receive_x = int(h == 0)
            d.mask_transmit = 0
            d.interp = _lib.INTERP[self.interp]
            d.aperture_x, d.aperture_y, d.aperture_z = tab["aperture"]
            d.sampling_frequency, d.demodulation_frequency = self.fs, self.fd
            d.speed_of_sound, d.start_time, d.fnumber = self.c, self.t0, self.fnumber
            d.depth_run = spec.nz   # voxels are stored depth-fastest
            self._descs[key] = d
        return self._descs[key]

    # -- request memory ---------------------------------------------------
    def request_bytes(self, spec: RequestSpec) -> int:
        per_half = (spec.n_tilt // 2) * self.n_elem * self.n_sample * spec.n_frame
        iq = 2 * per_half * (8 if spec.dtype == "fp32" else 4)
        scale = 0 if spec.dtype == "fp32" else 2 * (spec.n_tilt // 2) * self.n_elem * spec.n_frame * 4
        return iq + scale + spec.n_vox * spec.n_frame * 8 + spec.n_vox * 4

    def upload(self, req: Request):
        """Allocate the request's device buffers and start the H2D copies
        (I/Q already on this device, e.g. written there by the simulator, is
        used in place)."""
        dev = {}
        # device-resident I/Q may still be in production on the default stream
        self.copy.wait_stream(torch.cuda.default_stream(self.device))
        with torch.cuda.stream(self.copy):
            dev["iq"], dev["scale"] = [], []
            for h in req.halves:
                if isinstance(h, tuple):
                    dev["iq"].append(h[0].to(self.device, non_blocking=True).contiguous())
                    dev["scale"].append(h[1].to(self.device, non_blocking=True).contiguous())
                else:
                    dev["iq"].append(h.to(self.device, non_blocking=True).contiguous())
                    dev["scale"].append(None)
            dev["out"] = torch.empty((req.spec.n_frame, 1, req.spec.n_vox), dtype=torch.complex64,
                                     device=self.device)
            dev["mask"] = torch.empty((1, req.spec.n_vox), dtype=torch.float32, device=self.device)
            ev = torch.cuda.Event()
            ev.record(self.copy)
        req.dev = dev
        return ev

    def run_batch(self, k, work):
        """work: list of (req, frame_begin, frame_end).  Encodes the work list
        into slot k, copies it to the device and replays graph k."""
        entries = (_RcaBatchEntry * len(work))()
        for i, (req, f0, f1) in enumerate(work):
            e = entries[i]
            for h in range(2):
                e.half[h] = ctypes.pointer(self.desc(req.spec, h))
                e.iq[h] = req.dev["iq"][h].data_ptr()
                e.trace_scale[h] = req.dev["scale"][h].data_ptr() if req.dev["scale"][h] is not None else None
            e.n_half = 2
            e.iq_dtype = _lib.IQ_C64 if req.spec.dtype == "fp32" else _lib.IQ_F16X2
            e.n_frame, e.frame_begin, e.frame_end = req.spec.n_frame, f0, f1
            e.out = req.dev["out"].data_ptr()
            e.mask = req.dev["mask"].data_ptr()
        n_items = ctypes.c_int(0)
        rc = self.lib.rca_batch_encode(entries, len(work), self.max_halves, self.max_items,
                                       self.n_sample, ctypes.c_void_p(self.host_ws[k].data_ptr()),
                                       self.ws_bytes, ctypes.byref(n_items))
        _lib.check(rc)
        with torch.cuda.stream(self.compute):
            for req in {id(r): r for r, _, _ in work}.values():
                self.compute.wait_event(req.dev["uploaded"])
            self.dev_ws[k].copy_(self.host_ws[k], non_blocking=True)
            self.graphs[k].replay()
            ev = torch.cuda.Event()
            ev.record(self.compute)
        return ev, n_items.value

    def download(self, req: Request, after):
        host = torch.empty((req.spec.n_frame, req.spec.n_vox), dtype=torch.complex64, pin_memory=True)
        with torch.cuda.stream(self.d2h):
            self.d2h.wait_event(after)
            host.copy_(req.dev["out"][:, 0], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.d2h)
        return host, ev

    def release(self, req: Request):
        req.dev = {}

    def max_frames_per_chunk(self, spec: RequestSpec, quantum: int) -> int:
        """Frames per work chunk: about ``quantum`` updates, within the work-list
        capacity, a multiple of the request's frame block (the batched kernels
        run frames in blocks - 8 with depth tiles - so chunks start on a
        block; the last chunk ends at the last frame)."""
        per_frame = spec.updates(self.n_elem) // spec.n_frame
        tiled = spec.nz >= 64                   # the descriptor's depth hint is Nz
        b = self.fb[(spec.dtype, tiled)]
        # work items per frame block: 16-depth x 16-column tiles, or 256 voxels
        items = (math.ceil(spec.nx * spec.ny / 16) * math.ceil(spec.nz / 16) if tiled
                 else math.ceil(spec.n_vox / 256))
        f = min(quantum // max(per_frame, 1), (self.max_items // max(items, 1)) * b)
        return max(b, f - f % b)


class RcaServer:
    """Scheduler loop over a backend (see module docstring)."""

    def __init__(self, backend, budget_bytes: int, quantum: int = int(1.5e10), idle_sleep=2e-4):
        self.be = backend
        self.budget = int(budget_bytes)
        self.quantum = int(quantum)
        self.idle_sleep = idle_sleep
        self.queue = deque()
        self.lock = threading.Lock()
        self.active = []                 # admitted, not finished
        self.in_use = 0
        self.stats = ServerStats()
        self.slots = [None, None]        # (event, work) per work-list slot
        self.next_slot = 0
        self._dl = queue.Queue()         # (req, batch event) to download, None ends
        self._dl_thread = None
        self._stop = False
        self._closed = False
        self._thread = None
        self.error = None                # repr of the exception that stopped the scheduler

    # -- client side ------------------------------------------------------
    def submit(self, req: Request, arrival: float | None = None):
        """Queue a request.  Raises RuntimeError once the server is closed or
        its scheduler has failed: such a request would never run."""
        req.arrival = time.perf_counter() if arrival is None else arrival
        with self.lock:
            if self.error is not None:
                raise RuntimeError(f"server failed: {self.error}")
            if self._closed:
                raise RuntimeError("server is closed")
            self.queue.append(req)
        return req

    def start(self):
        self._dl_thread = threading.Thread(target=self._download_loop, daemon=True)
        self._dl_thread.start()
        self._thread = threading.Thread(target=self._loop, daemon=True)
        self._thread.start()
        return self

    def close(self, timeout=None):
        """Stop accepting requests, let the batches on the GPU finish and
        deliver what completes, then fail every request still queued or
        unfinished (their waiters get RuntimeError instead of blocking)."""
        with self.lock:
            self._closed = True
        self._stop = True
        if self._thread is not None:
            self._thread.join(timeout)
            if self._thread.is_alive():
                return          # still retiring in-flight batches; it fails the rest on exit
        self._fail_pending("server closed before the request finished")

    def _fail_pending(self, msg):
        with self.lock:
            pending = list(self.queue) + list(self.active)
            self.queue.clear()
            self.active = []
            self.in_use = 0
        for r in pending:
            if r.event.is_set():
                continue
            r.error = msg
            r.done = time.perf_counter()
            try:
                self.be.release(r)
            except Exception:   # a failed backend may not release; the request is failed anyway
                pass
            r.event.set()

    def drain(self, timeout=None):
        """Wait until every submitted request has finished."""
        t0 = time.perf_counter()
        while True:
            with self.lock:
                busy = bool(self.queue) or bool(self.active)
            if not busy:
                return
            if timeout is not None and time.perf_counter() - t0 > timeout:
                raise TimeoutError("server did not drain")
            time.sleep(1e-3)

    # -- scheduler --------------------------------------------------------
    def _admit(self):
        with self.lock:
            while self.queue:
                req = self.queue[0]
                need = self.be.request_bytes(req.spec)
                if need > self.budget:
                    self.queue.popleft()
                    req.error = f"request {req.rid} needs {need} bytes, budget is {self.budget}"
                    req.done = time.perf_counter()
                    self.stats.rejected += 1
                    req.event.set()
                    continue
                if self.in_use + need > self.budget:
                    return
                self.queue.popleft()
                req.admitted = time.perf_counter()
                try:
                    ev = self.be.upload(req)
                except Exception as exc:   # e.g. out of device memory: this request fails
                    req.error = f"upload of request {req.rid} failed: {exc!r}"
                    req.done = time.perf_counter()
                    try:
                        self.be.release(req)
                    except Exception:
                        pass
                    req.event.set()
                    continue
                self.in_use += need
                req.dev_bytes = need
                req.dev["uploaded"] = ev
                self.active.append(req)
                self.stats.max_in_flight = max(self.stats.max_in_flight, len(self.active))
                self.stats.max_bytes = max(self.stats.max_bytes, self.in_use)

    def _build_batch(self):
        work, budget = [], self.quantum
        with self.lock:
            active = list(self.active)
        runnable = [r for r in active if r.frames_issued < r.spec.n_frame
                    and r.dev.get("uploaded") is not None and r.dev["uploaded"].query()]
        # least attained service first: a request that has had less GPU work so
        # far (e.g. one that just arrived) is served before long-running ones
        runnable.sort(key=lambda r: (r.frames_issued * (r.spec.updates(self.be.n_elem) // r.spec.n_frame),
                                     r.arrival))
        while runnable and budget > 0:
            nxt = []
            for r in runnable:
                per_frame = r.spec.updates(self.be.n_elem) // r.spec.n_frame
                fmax = self.be.max_frames_per_chunk(r.spec, self.quantum)
                f0 = r.frames_issued
                f1 = min(r.spec.n_frame, f0 + fmax)
                if work and per_frame * (f1 - f0) > budget:
                    continue
                work.append((r, f0, f1))
                r.frames_issued = f1
                if r.started == 0.0:
                    r.started = time.perf_counter()
                budget -= per_frame * (f1 - f0)
                if r.frames_issued < r.spec.n_frame:
                    nxt.append(r)
                if budget <= 0:
                    break
            runnable = nxt if budget > 0 else []
        return work

    def _retire(self, k):
        ev, work = self.slots[k]
        ev.synchronize()
        self.slots[k] = None
        now = time.perf_counter()
        with self.lock:
            for r, f0, f1 in work:
                r.frames_done += f1 - f0
                self.stats.updates += r.spec.updates(self.be.n_elem) // r.spec.n_frame * (f1 - f0)
            self.stats.t_last = max(self.stats.t_last or now, now)
        for r, f0, f1 in work:
            if r.frames_done == r.spec.n_frame:
                self._dl.put((r, ev))

    def _download_loop(self):
        """Download thread: volume to host memory, latency, release."""
        while True:
            item = self._dl.get()
            if item is None:
                return
            r, ev = item
            try:
                host, dev_ev = self.be.download(r, ev)
                dev_ev.synchronize()
                r.result = host
                r.done = time.perf_counter()
            except Exception as exc:   # the request fails; the scheduler goes on
                r.error = f"download failed: {exc!r}"
                r.done = time.perf_counter()
            try:
                self.be.release(r)
            except Exception:
                pass
            with self.lock:
                if r.error is None:
                    self.stats.latencies.append(r.latency())
                    self.stats.t_last = max(self.stats.t_last or r.done, r.done)
                self.in_use -= r.dev_bytes
                if r in self.active:
                    self.active.remove(r)
            r.event.set()

    def _stop_downloads(self):
        if self._dl_thread is not None:
            self._dl.put(None)
            self._dl_thread.join()

    def _loop(self):
        try:
            while not self._stop:
                self._admit()
                k = self.next_slot
                if self.slots[k] is not None:
                    if self.slots[k][0].query() or self.slots[1 - k] is not None:
                        self._retire(k)
                launched = False
                if self.slots[k] is None:
                    work = self._build_batch()
                    if work:
                        if self.stats.t_first is None:
                            self.stats.t_first = time.perf_counter()
                        ev, _ = self.be.run_batch(k, work)
                        self.slots[k] = (ev, work)
                        self.stats.batches += 1
                        self.next_slot = 1 - k
                        launched = True
                for j in (0, 1):
                    if self.slots[j] is not None and self.slots[j][0].query():
                        self._retire(j)
                if not launched:
                    time.sleep(self.idle_sleep)
            for j in (0, 1):
                if self.slots[j] is not None:
                    self._retire(j)
            self._stop_downloads()
            self._fail_pending("server closed before the request finished")
        except Exception as exc:  # surface scheduler failures to every waiter
            with self.lock:
                self.error = repr(exc)
                self._closed = True
            # batches still on the GPU use the requests' device buffers: let
            # them end before the buffers are released
            for j in (0, 1):
                if self.slots[j] is not None:
                    try:
                        self.slots[j][0].synchronize()
                    except Exception:
                        pass
            self._stop_downloads()
            self._fail_pending(f"server failed: {exc!r}")


# ----------------------------------------------------------------------------
# Synthetic request data
# ----------------------------------------------------------------------------
def synth_request(rid: int, spec: RequestSpec, backend, device, n_scat=200, seed=0, host=True):
    """Seeded analytic I/Q for a request (scatterers seeded by request id;
    frames differ by a seeded per-frame phase), in the layout the backend
    uploads: in pinned host memory, or left on ``device`` with host=False."""
    from . import synth
    from .plan import quantize_fp16
    cfg = C.RcaConfig(f"req{rid}", backend.n_elem, backend.n_sample, spec.nx, spec.ny, spec.nz,
                      [C.Half(True, np.deg2rad(np.linspace(-12, 12, spec.n_tilt // 2))),
                       C.Half(False, np.deg2rad(np.linspace(-12, 12, spec.n_tilt // 2)))],
                      backend.interp, n_frame=spec.n_frame, n_scat=n_scat,
                      scat_box=(spec.nx * backend.dxyz / 2, spec.ny * backend.dxyz / 2, backend.z0,
                                backend.z0 + spec.nz * backend.dxyz),
                      seeds=(seed + 7919 * rid, seed + 7919 * rid + 1, seed + 7919 * rid + 2))
    pos, amp = synth.scatterers(cfg)
    rng = np.random.Generator(np.random.PCG64(seed + 104729 * rid + 3))
    ph = torch.from_numpy(np.exp(1j * rng.uniform(0, 2 * np.pi, spec.n_frame)).astype(np.complex64)).to(device)
    halves = []
    for h in cfg.halves:
        one = synth.echoes(cfg, h, pos, amp, device)
        ens = (one[None] * ph[:, None, None, None]).contiguous()
        if spec.dtype == "fp16":
            hq, sc = quantize_fp16(ens)
            halves.append((hq.cpu().pin_memory(), sc.cpu().pin_memory()) if host else (hq, sc))
        else:
            halves.append(ens.cpu().pin_memory() if host else ens)
        del ens
    return Request(rid, spec, halves)
