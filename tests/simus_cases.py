"""Small SIMUS-RCA cases shared by the simulator tests: an RCA probe (row
elements transmit, column elements receive, each element a line of
sub-elements), a Gaussian-windowed pulse, Verasonics-style sequence dicts."""
from __future__ import annotations

import numpy as np


def rca_case(n_ele=8, n_sub=3, n_tx=2, n_scat=5, n_frame=1, n_sample=128, mode="BS100BW",
             fc=7.5e6, rfocus=10000.0, baffle=0, lens=0.0, attenuation=0.5, seed=3):
    """(Media, TransTX, TransRX, TXSI, RXSI) with FP32-representable values."""
    rng = np.random.default_rng(seed)
    c = 1540.0
    lam = c / fc
    pitch = lam / 2
    xe = (np.arange(n_ele) - (n_ele - 1) / 2) * pitch
    span = n_ele * pitch
    ys = (np.arange(n_sub) - (n_sub - 1) / 2) * span / max(n_sub, 1)

    def trans(row):
        pos = np.zeros((n_ele, 5, n_sub))
        for u in range(n_sub):
            pos[:, 0 if row else 1, u] = xe
            pos[:, 1 if row else 0, u] = ys[u]
        t = np.arange(33) / 250e6
        ir = np.sin(2 * np.pi * fc * t) * np.hanning(t.size)
        return {"frequency": fc, "ElementPos": pos.astype(np.float32),
                "ElementWidth": np.float32(0.9 * pitch), "ElementHeight": np.float32(span),
                "ElementElevationFocus": np.float32(rfocus), "baffle": baffle,
                "lensCorrection": np.float32(lens), "impulseResponse": ir}

    tp = np.arange(int(2 / fc * 250e6)) / 250e6
    pulse = np.sin(2 * np.pi * fc * tp) * np.hanning(tp.size)
    fs = 4 * fc if mode != "NS200BW" else 4 * fc
    theta = np.linspace(-5, 5, n_tx) * np.pi / 180 if n_tx > 1 else np.zeros(1)
    TXSI = [{"pulseShape": pulse, "Delay": (xe * np.sin(th) / c - (xe * np.sin(th) / c).min()),
             "Apod": np.hanning(n_ele + 2)[1:-1]} for th in theta]
    filt = np.array([0.1, 0.25, 0.3, 0.25, 0.1])
    RXSI = [{"samplingFrequency": fs, "demodFrequency": fc, "sampleMode": mode,
             "startTime": 10 * lam / c, "numelSample": n_sample,
             "Apod": np.linspace(0.8, 1.0, n_ele), "InputFilter": filt} for _ in range(n_tx)]
    depth = (12 + rng.uniform(0, 10, (n_scat, 1, n_tx * n_frame))) * lam
    lat = rng.uniform(-span / 3, span / 3, (n_scat, 2, n_tx * n_frame))
    bsc = rng.standard_normal((n_scat, 1, n_tx * n_frame))
    pos = np.concatenate([lat, depth, bsc], axis=1).astype(np.float32)
    Media = {"speedOfSound": c, "attenuation": attenuation, "scattererPos": pos}
    return Media, trans(True), trans(False), TXSI, RXSI
