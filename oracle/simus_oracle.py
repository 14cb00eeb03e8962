"""ORACLE — test infrastructure only.

Only tests may import this module, and only as the checker.  The product
(``simu_rca_ab_amd.simulator`` over ``include/rca_sim.h``) never routes
through it.

FP64 numpy restatement of the SIMUS-style RCA simulator of
``MATLAB/lib/simulator/SimUS_CUDARCA.cu`` / ``.m`` (and the ``_ab`` variant):

* ``csinc``            SimUS_CUDARCA.cu:43-66   complex sinc, 0 for |x| >= pi
* ``dir_flat``         .cu:116-130              sinc(sin(theta)/2 * width * k) (x cos(theta) with baffle 1)
* ``dir_curv``         .cu:166-203              four-Gaussian elevation-focus term
* ``element_terms``    .cu:237-294 / 340-373    Green exp(jkr)/r x azimuth x elevation x lens
* ``txfield``          .cu:209-298              transmit field per (frequency, scatterer, tilt)
* ``rxfield``          .cu:301-379              receive field per (frequency, element, scatterer)
* ``prepare``          SimUS_CUDARCA.m:136-259  time/frequency grid, wavenumbers, spectra
* ``simulate``         .m:261-390 (+ _ab.m:242-243, 365-382, 398-410)

The kernels' inputs are the FP32 values the reference passes (``prepare``
rounds them the way the wrapper does); every function here promotes them to
FP64 and computes in FP64.
"""
from __future__ import annotations

import math

import numpy as np

# four-Gaussian expansion coefficients (SimUS_CUDARCA.cu:10-13), FP32 constants
# in the kernels
_f = lambda v: float(np.float32(v))   # noqa: E731
GA = np.array([complex(_f(a), _f(b)) for a, b in
               ((0.187, 0.275), (0.288, -1.954), (0.187, -0.275), (0.288, 1.954))])
GB = np.array([complex(_f(a), _f(b)) for a, b in
               ((4.558, -25.59), (8.598, -7.924), (4.558, 25.59), (8.598, 7.924))])
PI_REF = _f(3.14159265359)   # the kernels' float PI constant
SIM_FS = 250e6                              # pulse / impulse-response sampling (.m:146)


def csinc(x):
    """sin(x)/x of complex x, 1 at 0 and 0 where |x| >= pi (.cu:43-66)."""
    x = np.asarray(x, dtype=np.complex128)
    ax = np.abs(x)
    safe = np.where(ax > 0, x, 1.0)
    val = np.sin(safe) / safe
    return np.where(ax > 0, np.where(ax < PI_REF, val, 0.0), 1.0)


def dir_flat(theta, k, width, baffle):
    """Flat-element directivity (.cu:116-130)."""
    d = csinc(np.sin(theta) / 2 * width * k)
    if baffle == 1:
        d = d * np.cos(theta)
    return d


def dir_curv(r, y, k, height, rfocus):
    """Elevation-focused directivity: sum over the four Gaussians of
    A_g sqrt(pi/alpha_g) exp(beta^2/(4 alpha_g) + gamma) with
    alpha_g = B_g/h^2 + j k (r - rf)/(2 r rf), beta = -j k y/r,
    gamma = j k y^2/(2r) (.cu:166-203)."""
    r = np.asarray(r, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    k = np.asarray(k, dtype=np.complex128)
    alpha = 1j * k * ((r - rfocus) / (r * rfocus) * 0.5)
    beta = -1j * k * (y / r)
    gama = 1j * k * (y * y / (2 * r))
    out = 0
    for a, b in zip(GA, GB):
        ag = b / (height * height) + alpha
        out = out + a * (np.sqrt(PI_REF / ag) * np.exp(beta * beta / (4 * ag) + gama))
    return out


def element_terms(scat, ele, k, c, width, height, rfocus, baffle, lens):
    """Per-(frequency, scatterer, element) field of one transducer, summed over
    its sub-elements: Green x azimuth directivity x elevation directivity x
    lens phase (.cu:253-284, 340-367).

    scat [S, 3]; ele [E, 5, U] (x, y, z, azimuth, elevation per sub-element);
    k [F] complex.  Returns complex [F, S, E]."""
    scat = np.asarray(scat, np.float64)
    ele = np.asarray(ele, np.float64)
    k = np.asarray(k, np.complex128)
    w = k.real * float(c)
    out = np.zeros((k.size, scat.shape[0], ele.shape[0]), np.complex128)
    kk = k[:, None, None]
    for u in range(ele.shape[2]):
        dx = scat[:, None, 0] - ele[None, :, 0, u]
        dy = scat[:, None, 1] - ele[None, :, 1, u]
        dz = scat[:, None, 2] - ele[None, :, 2, u]
        r = np.sqrt(dx * dx + dy * dy + dz * dz)                        # .cu:262
        az = np.arcsin(dx / r) - ele[None, :, 3, u]                      # .cu:263
        el = np.arcsin(dy / r) - ele[None, :, 4, u]                      # .cu:264
        term = np.exp(1j * kk * r) / r                                   # .cu:267 (Green)
        term = term * dir_flat(az, kk, width, baffle)                    # .cu:268
        if rfocus > 10.0:
            term = term * dir_flat(el, kk, width, baffle)                # .cu:271 (width, as in the reference)
        else:
            term = term * dir_curv(r, r * np.sin(el), kk, height, rfocus)   # .cu:273
        term = term * np.exp(1j * w * lens)[:, None, None]              # .cu:275
        out += term
    return out


def txfield(r_scat, r_ele, k, pulse, ir, delay, apod, width, height, rfocus, baffle, lens, c):
    """Transmit field TX [F, S, T] (.cu:209-298).

    r_scat [S, 3, T] (each tilt's scatterers), r_ele [E, 5, U], k/pulse/ir
    [F], delay/apod [E, T] (delay already minus the receive start time)."""
    k = np.asarray(k, np.complex128)
    w = k.real * float(c)
    spec = np.conj(np.asarray(pulse, np.complex128)) * np.conj(np.asarray(ir, np.complex128))
    T = r_scat.shape[2]
    out = np.zeros((k.size, r_scat.shape[0], T), np.complex128)
    for t in range(T):
        g = element_terms(r_scat[:, :, t], r_ele, k, c, width, height, rfocus, baffle, lens)
        s = apod[None, :, t] * np.exp(1j * w[:, None] * np.asarray(delay, np.float64)[None, :, t])
        out[:, :, t] = (g * s[:, None, :]).sum(axis=2) * spec[:, None]   # .cu:278-288
    return out


def rxfield(r_scat, r_ele, k, ir, width, height, rfocus, baffle, lens, c):
    """Receive field RX [F, E, S] (.cu:301-379); r_scat [S, 3]."""
    g = element_terms(r_scat, r_ele, k, c, width, height, rfocus, baffle, lens)
    return np.transpose(g, (0, 2, 1)) * np.asarray(ir, np.complex128)[:, None, None]


def _f32(x):
    return np.asarray(x, dtype=np.float32)


def prepare(Media, TransTX, TransRX, TXSI, RXSI):
    """Time/frequency grid and the FP32 arrays the kernels receive
    (SimUS_CUDARCA.m:136-259).  FP32 operations are done in FP32, as the
    wrapper's single arithmetic does."""
    fs = float(RXSI[0]["samplingFrequency"])
    mode = RXSI[0]["sampleMode"]
    sub = {"NS200BW": 2, "BS100BW": 4, "BS50BW": 8}[mode]                # .m:158-173
    pulse_dur = np.size(TXSI[0]["pulseShape"]) / SIM_FS                   # .m:150
    t_start = RXSI[0]["startTime"] - 1.0 / TransTX["frequency"]           # .m:154
    max_delay = max(np.max(np.asarray(t["Delay"])) for t in TXSI)        # .m:156-159
    tmax = RXSI[0]["numelSample"] / fs + pulse_dur + max_delay + t_start  # .m:175
    nfft = int(128 * math.ceil(tmax * fs / 128))                          # .m:177
    nf = nfft // 2
    f = _f32(np.linspace(0, fs / 2, nf))                                  # .m:179
    agl = _f32(2 * np.pi * f)
    kr = _f32(agl / float(Media["speedOfSound"]))
    ki = _f32(float(Media["attenuation"]) / 8.7 * kr * float(Media["speedOfSound"]) / (2 * np.pi * 1e4))
    k = (kr.astype(np.complex64) + 1j * ki.astype(np.complex64)).astype(np.complex64)   # .m:186-190
    m = int(np.ceil(np.float32(nfft) * np.float32(SIM_FS) / np.float32(fs)))            # .m:205-207
    pulse = np.fft.fft(np.asarray(TXSI[0]["pulseShape"], np.float64).ravel(), m)[:nf].astype(np.complex64)
    ir_tx = np.fft.fft(np.asarray(TransTX["impulseResponse"], np.float64).ravel(), m)[:nf].astype(np.complex64)
    rx = TransRX if TransRX is not None else TransTX
    ir_rx = np.fft.fft(np.asarray(rx["impulseResponse"], np.float64).ravel(), m)[:nf].astype(np.complex64)
    ipf = np.fft.fft(np.asarray(RXSI[0]["InputFilter"], np.float64).ravel(), nfft)[:nf].astype(np.complex64)
    n_tx = len(TXSI)
    delay = np.stack([_f32(np.asarray(t["Delay"]).ravel()) for t in TXSI], axis=1)     # .m:240-243
    apod = np.stack([_f32(np.asarray(t["Apod"]).ravel()) for t in TXSI], axis=1)
    delay = _f32(delay - np.float32(t_start))                                           # .m:318
    return dict(fs=fs, sub=sub, t_start=t_start, nfft=nfft, nf=nf, f=f, agl=agl, k=k, pulse=pulse,
                ir_tx=ir_tx, ir_rx=ir_rx, ipfilt=ipf, delay=delay, apod=apod, n_tx=n_tx)


def transducer_scalars(T):
    """single(X(1)) of ElementWidth, ElementHeight, ElementElevationFocus,
    baffle, lensCorrection (.m:321-327, 347-353)."""
    return [float(np.float32(np.ravel(T[k_])[0])) for k_ in
            ("ElementWidth", "ElementHeight", "ElementElevationFocus", "baffle", "lensCorrection")]


def _ele(T):
    e = np.asarray(T["ElementPos"], np.float32)
    return e if e.ndim == 3 else e[:, :, None]


def simulate(Media, TransTX, TransRX, TXSI, RXSI, AbLaw=None):
    """SimUS_CUDARCA(Media, TransTX, TransRX, TXSI, RXSI) in FP64 (.m:136-390);
    with AbLaw also returns the receive-aberrated data (_ab.m:242-243,
    365-382).  Media.scattererPos [nScat, 4, nEvent]; returns Raw
    [2*numelSample/sub, nEleRX, nEvent] (and Raw_ab)."""
    P = prepare(Media, TransTX, TransRX, TXSI, RXSI)
    rx = TransRX if TransRX is not None else TransTX
    e_tx, e_rx = _ele(TransTX), _ele(rx)
    n_sub_rx = e_tx.shape[2]               # .m:142-143: the transmit array's count
    e_rx = e_rx[:, :, :n_sub_rx]
    pos = np.asarray(Media["scattererPos"], np.float32)
    if pos.ndim == 2:
        pos = pos[:, :, None]
    n_scat, _, n_event = pos.shape
    T = P["n_tx"]
    nfft, nf = P["nfft"], P["nf"]
    n_ele_rx = e_rx.shape[0]
    RF = np.zeros((nfft, n_ele_rx, n_event))
    RF_ab = np.zeros_like(RF)
    c = float(np.float32(Media["speedOfSound"]))
    prm_tx, prm_rx = transducer_scalars(TransTX), transducer_scalars(rx)
    if AbLaw is not None:
        law = (_f32(np.ravel(AbLaw["Apod"]))[None, :].astype(np.float64) *
               np.exp(1j * P["agl"].astype(np.float64)[:, None] *
                      _f32(np.ravel(AbLaw["Delay"])).astype(np.float64)[None, :]))   # _ab.m:369
    n_frame = -(-n_event // T)
    for fr in range(n_frame):
        ev = fr * T + np.arange(T)
        r_scat = pos[:, :3, ev]
        bsc = pos[:, 3, ev].astype(np.float64)                             # [S, T]
        TX = txfield(r_scat, e_tx, P["k"], P["pulse"], P["ir_tx"], P["delay"], P["apod"],
                     *prm_tx, c)
        TX = TX * bsc[None, :, :]                                          # .m:332
        TX[np.isnan(TX)] = 0                                               # .m:335
        for t in range(T):
            RX = rxfield(r_scat[:, :, t], e_rx, P["k"], P["ir_rx"], *prm_rx, c)
            RX = RX * np.asarray(RXSI[t]["Apod"], np.float64).ravel()[None, :, None]   # .m:358
            txrx = (TX[:, None, :, t] * RX).sum(axis=2) * P["ipfilt"].astype(np.complex128)[:, None]
            RF[:, :, ev[t]] += np.real(np.fft.ifft(np.conj(txrx), nfft, axis=0))         # .m:361-366
            if AbLaw is not None:
                RF_ab[:, :, ev[t]] += np.real(np.fft.ifft(np.conj(txrx * law), nfft, axis=0))
    out = [to_raw(RF, RXSI[0]["numelSample"], P["sub"])]
    if AbLaw is not None:
        out.append(to_raw(RF_ab, RXSI[0]["numelSample"], P["sub"]))
    return tuple(out) if len(out) > 1 else out[0]


def to_raw(RF, n_sample, sub):
    """Trim or zero-pad to numelSample rows, then keep samples 1 and 2 of every
    ``sub`` (.m:378-390): Raw(1:2:end) = RF(1:sub:end), Raw(2:2:end) = RF(2:sub:end)."""
    n_sample = int(math.ceil(n_sample))
    if RF.shape[0] > n_sample:
        RF = RF[:n_sample]
    else:
        RF = np.concatenate([RF, np.zeros((n_sample - RF.shape[0],) + RF.shape[1:])], axis=0)
    n_out = RF.shape[0] // sub
    raw = np.zeros((2 * n_out,) + RF.shape[1:])
    raw[0::2] = RF[0:n_out * sub:sub]
    raw[1::2] = RF[1:n_out * sub:sub]
    return raw
