"""SIMUS-style RCA simulator on the GPU (``include/rca_sim.h``).

Drop-in for ``MATLAB/lib/simulator/SimUS_CUDARCA.m`` and
``SimUS_CUDARCA_ab.m`` with the MATLAB structs as dicts (``TXSI`` / ``RXSI``
as lists of dicts, one per tilt):

* ``SimUS_CUDARCA(Media, TransTX, TransRX, TXSI, RXSI[, Options])`` -> Raw
* ``SimUS_CUDARCA_ab(Media, TransTX, TransRX, TXSI, RXSI, AbLaw[, Options])``
  -> (Raw, Raw_ab)

``Raw`` is float32 [2*numelSample/sub, nEleRX, nEvent] (F-ordered numpy, the
reference's layout): per sample mode, samples 1 and 2 of every ``sub`` of the
simulated RF (SimUS_CUDARCA.m:386-390).  ``Media.scattererPos`` is
[nScat, 4, nEvent] (x, y, z, backscatter; one set per transmit event, events
grouped by frame).  The host computes the time/frequency grid and the short
spectra of the pulse, impulse responses and input filter (numpy FFTs of a
few hundred samples, .m:136-238); the fields, the sum over scatterers, the
inverse transforms and the decimation run in the HIP library.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib

SIM_FS = 250e6   # sampling of the pulse and impulse responses (.m:146)
SUB = {"NS200BW": 2, "BS100BW": 4, "BS50BW": 8}   # .m:158-173
_ID = "SimUS_CUDARCA"


class SimInputError(ValueError):
    def __init__(self, tag: str, msg: str):
        super().__init__(msg)
        self.identifier = f"{_ID}:{tag}"


def _scalar32(x) -> float:
    return float(np.float32(np.ravel(np.asarray(x))[0]))


def prepare(Media, TransTX, TransRX, TXSI, RXSI):
    """Time/frequency grid, wavenumbers and spectra (SimUS_CUDARCA.m:136-243),
    in the wrapper's precision: single values where it makes them single."""
    fs = float(RXSI[0]["samplingFrequency"])
    mode = RXSI[0]["sampleMode"]
    if mode not in SUB:
        raise SimInputError("sampleModeInvalid", f"Sampling Mode Not implemented: {mode!r}")
    sub = SUB[mode]
    t_start = float(RXSI[0]["startTime"]) - 1.0 / float(TransTX["frequency"])
    pulse_dur = np.size(TXSI[0]["pulseShape"]) / SIM_FS
    max_delay = max(float(np.max(np.asarray(t["Delay"], np.float64))) for t in TXSI)
    tmax = float(RXSI[0]["numelSample"]) / fs + pulse_dur + max_delay + t_start
    nfft = int(128 * math.ceil(tmax * fs / 128))
    nf = nfft // 2
    f = np.linspace(0, fs / 2, nf).astype(np.float32)
    c = float(Media["speedOfSound"])
    agl = (2 * math.pi * f).astype(np.float32)              # single arithmetic, as in .m:184-190
    kr = (agl / c).astype(np.float32)
    ki = (float(Media["attenuation"]) / 8.7 * kr * c / (2 * math.pi * 1e4)).astype(np.float32)
    k = (kr + 1j * ki).astype(np.complex64)
    m = int(np.ceil(np.float32(nfft) * np.float32(SIM_FS) / np.float32(fs)))

    def spec(x, n):
        return np.fft.fft(np.asarray(x, np.float64).ravel(), n)[:nf].astype(np.complex64)

    rx = TransRX if TransRX else TransTX
    delay = np.stack([np.asarray(t["Delay"], np.float32).ravel() for t in TXSI])   # [nTx][nEle]
    apod = np.stack([np.asarray(t["Apod"], np.float32).ravel() for t in TXSI])
    return {"fs": fs, "sub": sub, "t_start": t_start, "nfft": nfft, "nf": nf, "agl": agl, "k": k,
            "pulse": spec(TXSI[0]["pulseShape"], m), "ir_tx": spec(TransTX["impulseResponse"], m),
            "ir_rx": spec(rx["impulseResponse"], m), "ipfilt": spec(RXSI[0]["InputFilter"], nfft),
            "delay": (delay - np.float32(t_start)).astype(np.float32), "apod": apod}


def _element_array(T, n_sub=None):
    """ElementPos [nEle, 5, nSub] (MATLAB) -> float32 [nSub][5][nEle]."""
    e = np.asarray(T["ElementPos"], np.float32)
    if e.ndim == 2:
        e = e[:, :, None]
    if n_sub is not None:
        if e.shape[2] < n_sub:
            raise SimInputError("subElementMismatch",
                                f"the receive array has {e.shape[2]} sub-elements, the transmit "
                                f"array {n_sub}; the wrapper loops over the transmit count")
        e = e[:, :, :n_sub]
    return np.ascontiguousarray(e.transpose(2, 1, 0))


def _transducer(T):
    return [_scalar32(T[k]) for k in ("ElementWidth", "ElementHeight", "ElementElevationFocus",
                                      "baffle", "lensCorrection")]


class RcaSimulator:
    """One (TransTX, TransRX, TXSI, RXSI) set resident on a GPU; ``frame``
    simulates the nTx events of one frame."""

    def __init__(self, Media, TransTX, TransRX, TXSI, RXSI, AbLaw=None, device=None):
        self.lib = _lib.load()
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        dev = self.device
        P = prepare(Media, TransTX, TransRX, TXSI, RXSI)
        self.P = P
        rx = TransRX if TransRX else TransTX
        e_tx = _element_array(TransTX)
        self.n_sub = e_tx.shape[0]
        e_rx = _element_array(rx, self.n_sub)      # .m:142-143: the transmit count
        self.n_ele_tx, self.n_ele_rx = e_tx.shape[2], e_rx.shape[2]
        self.n_tx = len(TXSI)
        if P["delay"].shape != (self.n_tx, self.n_ele_tx):
            raise SimInputError("DelayInvalid", "TXSI(t).Delay must have one value per transmit element")
        to = lambda a, dt=None: torch.from_numpy(np.ascontiguousarray(a)).to(dev)   # noqa: E731
        self.e_tx, self.e_rx = to(e_tx), to(e_rx)
        self.k = to(P["k"])
        self.pulse, self.ir_tx, self.ir_rx, self.ipfilt = (to(P[n]) for n in ("pulse", "ir_tx", "ir_rx", "ipfilt"))
        self.agl = to(P["agl"])
        self.delay, self.apod_tx = to(P["delay"]), to(P["apod"])
        self.apod_rx = to(np.stack([np.asarray(r["Apod"], np.float32).ravel() for r in RXSI[:self.n_tx]]))
        if self.apod_rx.shape[1] != self.n_ele_rx:
            raise SimInputError("ApodInvalid", "RXSI(t).Apod must have one value per receive element")
        self.c = float(np.float32(Media["speedOfSound"]))
        self.tr_tx, self.tr_rx = _transducer(TransTX), _transducer(rx)
        self.ab = None
        if AbLaw is not None:
            d = np.asarray(AbLaw["Delay"], np.float32).ravel()
            a = np.asarray(AbLaw.get("Apod", np.ones_like(d)), np.float32).ravel()
            if d.size != self.n_ele_rx or a.size != self.n_ele_rx:
                raise SimInputError("AbLawInvalid", "AbLaw.Delay/Apod need one value per receive element")
            self.ab = (to(d), to(a))
        self.n_sample = int(math.ceil(float(RXSI[0]["numelSample"])))
        self.sub = P["sub"]
        self.n_out = self.n_sample // self.sub

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def frame(self, pos, raw, raw_ab=None):
        """Simulate one frame.  pos: float32 [nTx][4][nScat] tensor on the
        device (x, y, z, bsc of each tilt's scatterers).  Writes raw (and
        raw_ab): float32 [nTx][nEleRX][2*n_out]."""
        lib, P = self.lib, self.P
        n_tx, n_scat = pos.shape[0], pos.shape[2]
        nf, nfft = P["nf"], P["nfft"]
        if n_tx != self.n_tx:
            raise ValueError(f"a frame has {self.n_tx} events")
        r_scat = pos[:, :3, :].contiguous()
        bsc = pos[:, 3, :].contiguous()
        tx = torch.empty((n_tx, n_scat, nf), dtype=torch.complex64, device=self.device)
        st = self._stream()
        vp = ctypes.c_void_p
        w, h, rf, bf, ln = self.tr_tx
        with torch.cuda.device(self.device):
            _lib.check(lib.rca_sim_txfield(
                vp(tx.data_ptr()), vp(r_scat.data_ptr()), vp(self.e_tx.data_ptr()), vp(self.k.data_ptr()),
                vp(self.pulse.data_ptr()), vp(self.ir_tx.data_ptr()), vp(self.delay.data_ptr()),
                vp(self.apod_tx.data_ptr()), w, h, rf, bf, ln, self.c, n_scat, self.n_ele_tx, nf, n_tx,
                self.n_sub, vp(bsc.data_ptr()), st))
            spec = torch.empty((n_tx, self.n_ele_rx, nfft), dtype=torch.complex64, device=self.device)
            spec_ab = torch.empty_like(spec) if (self.ab is not None and raw_ab is not None) else None
            a = _lib.RcaSimReceiveArgs()
            a.tx, a.r_scat, a.r_ele = tx.data_ptr(), r_scat.data_ptr(), self.e_rx.data_ptr()
            a.k, a.ir, a.ipfilt = self.k.data_ptr(), self.ir_rx.data_ptr(), self.ipfilt.data_ptr()
            a.agl, a.apod = self.agl.data_ptr(), self.apod_rx.data_ptr()
            if spec_ab is not None:
                a.ab_delay, a.ab_apod = self.ab[0].data_ptr(), self.ab[1].data_ptr()
                a.spec_ab = spec_ab.data_ptr()
            a.width, a.height, a.rfocus, a.baffle, a.lens = self.tr_rx
            a.speed_of_sound = self.c
            a.n_scat, a.n_ele, a.n_f, a.n_sub, a.n_tx, a.nfft = (n_scat, self.n_ele_rx, nf, self.n_sub,
                                                                 n_tx, nfft)
            a.spec = spec.data_ptr()
            _lib.check(lib.rca_sim_receive(ctypes.byref(a), st))
            for s, r in ((spec, raw), (spec_ab, raw_ab)):
                if s is None:
                    continue
                _lib.check(lib.rca_sim_spectra_to_raw(vp(s.data_ptr()), n_tx * self.n_ele_rx, nfft,
                                                      self.n_sample, self.sub, vp(r.data_ptr()), st))
        return raw, raw_ab

    def run(self, scattererPos):
        """All frames of Media.scattererPos [nScat, 4, nEvent]; returns device
        tensors [nEvent][nEleRX][2*n_out] (and the aberrated ones)."""
        pos = np.asarray(scattererPos, np.float32)
        if pos.ndim == 2:
            pos = pos[:, :, None]
        n_event = pos.shape[2]
        if n_event % self.n_tx:
            raise SimInputError("eventsInvalid", "the number of events must be a multiple of numel(TXSI)")
        dev = self.device
        raw = torch.zeros((n_event, self.n_ele_rx, 2 * self.n_out), dtype=torch.float32, device=dev)
        raw_ab = torch.zeros_like(raw) if self.ab is not None else None
        allpos = torch.from_numpy(np.ascontiguousarray(pos.transpose(2, 1, 0))).to(dev)   # [ev][4][S]
        for fr in range(n_event // self.n_tx):
            sl = slice(fr * self.n_tx, (fr + 1) * self.n_tx)
            self.frame(allpos[sl], raw[sl], raw_ab[sl] if raw_ab is not None else None)
        return raw, raw_ab


def _validate_options(Options):
    Options = dict(Options or {})
    if "decreaseFactorGPU" not in Options:
        Options["decreaseFactorGPU"] = 1                              # .m:122-123
    else:
        v = float(np.ravel(Options["decreaseFactorGPU"])[0])
        if math.floor(v) != v or v < 0:                               # .m:124-130
            raise SimInputError("decreaseFactorGPUInvalid",
                                "GPU decrease factor ('Options.decreaseFactorGPU') must be a positive integer!")
    return Options


def _to_matlab(raw):
    return raw.cpu().numpy().transpose(2, 1, 0)     # [2*n_out, nEle, nEvent], F-ordered


def SimUS_CUDARCA(*args, device=None):
    """``Raw = SimUS_CUDARCA(Media, TransTX, TransRX, TXSI, RXSI[, Options])``
    (SimUS_CUDARCA.m:1-390).  With no argument it builds the HIP library
    (the reference prints its PTX build line, .m:110-119)."""
    if not args:
        from . import build
        build.build_library()
        return None
    Media, TransTX, TransRX, TXSI, RXSI = args[:5]
    _validate_options(args[5] if len(args) > 5 else None)
    sim = RcaSimulator(Media, TransTX, TransRX, list(TXSI), list(RXSI), device=device)
    raw, _ = sim.run(Media["scattererPos"])
    return _to_matlab(raw)


def SimUS_CUDARCA_ab(*args, device=None):
    """``[Raw, Raw_ab] = SimUS_CUDARCA_ab(Media, TransTX, TransRX, TXSI, RXSI,
    AbLaw[, Options])``: Raw_ab has the receive field multiplied by
    AbLaw.Apod exp(j 2 pi f AbLaw.Delay) per receive element
    (SimUS_CUDARCA_ab.m:242-243, 365-382, 398-410)."""
    Media, TransTX, TransRX, TXSI, RXSI, AbLaw = args[:6]
    _validate_options(args[6] if len(args) > 6 else None)
    sim = RcaSimulator(Media, TransTX, TransRX, list(TXSI), list(RXSI), AbLaw=AbLaw, device=device)
    raw, raw_ab = sim.run(Media["scattererPos"])
    return _to_matlab(raw), _to_matlab(raw_ab)


__all__ = ["SimUS_CUDARCA", "SimUS_CUDARCA_ab", "RcaSimulator", "prepare", "SimInputError"]
