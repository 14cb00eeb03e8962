"""ORACLE — test infrastructure only.

Only tests may import this module, and only as the checker.  The product
(``simu_rca_ab_amd.clutterfilter`` over ``rca_svd_filter``) never routes
through it.

FP64 numpy restatement of the SVD clutter filter:

* ``Python/clutterfilter/SVDClutterFilter.py:10-23`` (``SVDClutterFilter``):
  Casorati matrix of the F-ordered pixels (``af.moddims``, :15), Gram
  ``conj(IQ.T) @ IQ`` (:16), SVD of the Gram (:16), kept singular vectors
  ``u[:, int(ncut[0]):int(ncut[1])]`` with ``ncut = wcut * Nt`` (:12, :17),
  ``(IQ @ V) @ conj(V.T)`` (:18), reshape back (:19); the mask argument is not
  used; returns the filtered data and the singular values.
* ``MATLAB/lib/clutterFilter/SVDClutterFilter.m:1-45``: ``w`` must lie in
  [0, 1] (:6-8), ``Ncut = round(w * Nt)`` (:13), data divided by
  ``sqrt(Mask)`` (:21, default ones :15-17), Gram and SVD in double (:23),
  bands 'low' ``1:Ncut(1)``, 'high' ``Ncut(1):end``, 'stop'
  ``[1:Ncut(1), Ncut(2):end]``, default 'band' ``Ncut(1):Ncut(2)`` (1-based,
  :26-44), output not re-weighted (:44-45).

The Gram matrix is Hermitian positive semi-definite, so its SVD is its
eigen-decomposition: singular values = eigenvalues in descending order.
"""
from __future__ import annotations

import math

import numpy as np


def casorati(iq: np.ndarray) -> np.ndarray:
    """[..., Nt] -> [Npix, Nt] with pixels in F order (SVDClutterFilter.py:15,
    SVDClutterFilter.m:22)."""
    iq = np.asarray(iq)
    return iq.reshape(-1, iq.shape[-1], order="F")


def gram_eig(Y: np.ndarray):
    """Eigenvalues (descending) and eigenvectors (columns, same order) of
    Y^H Y in FP64 (SVDClutterFilter.py:16, .m:23)."""
    Y = np.asarray(Y, dtype=np.complex128)
    G = Y.conj().T @ Y
    w, V = np.linalg.eigh(G)
    return w[::-1].copy(), V[:, ::-1].copy()


def project(Y: np.ndarray, keep) -> tuple:
    """Y V_k V_k^H for the kept eigenvector indices (descending order).
    Returns (filtered, eigenvalues)."""
    Y = np.asarray(Y, dtype=np.complex128)
    w, V = gram_eig(Y)
    keep = np.asarray(keep, dtype=bool)
    Vk = V[:, keep]
    return (Y @ Vk) @ Vk.conj().T, w


def keep_python(wcut, nt: int) -> np.ndarray:
    """``u[:, int(ncut[0]):int(ncut[1])]`` with ``ncut = wcut * Nt``
    (SVDClutterFilter.py:12, :17)."""
    ncut = np.asarray(wcut, dtype=np.float64) * nt
    keep = np.zeros(nt, bool)
    keep[int(ncut[0]):int(ncut[1])] = True
    return keep


def matlab_round(x):
    """MATLAB round: halves away from zero."""
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def keep_matlab(w, nt: int, kind: str = "band") -> np.ndarray:
    """Columns of ``Eig_vect`` that SVDClutterFilter.m:26-44 keeps, as a
    0-based mask (MATLAB ranges are 1-based and inclusive)."""
    w = np.atleast_1d(np.asarray(w, dtype=np.float64))
    if w.min() < 0 or w.max() > 1:
        raise ValueError("w must be is the normalized cutoff frequency [0 1]")
    n = matlab_round(w * nt).astype(int)

    def start(i):   # a range that starts at Ncut(i): MATLAB indices are >= 1
        if n[i] < 1:
            raise IndexError("Array indices must be positive integers")
        return n[i] - 1

    keep = np.zeros(nt, bool)
    if kind == "low":
        keep[:n[0]] = True                       # 1:Ncut(1)
    elif kind == "high":
        keep[start(0):] = True                   # Ncut(1):end
    elif kind == "stop":
        keep[:n[0]] = True                       # [1:Ncut(1), Ncut(2):end]
        keep[start(1):] = True
    else:
        keep[start(0):n[1]] = True               # Ncut(1):Ncut(2)
    return keep


def svd_filter_python(IQ, wcut):
    """SVDClutterFilter.py:10-23 in FP64: (IQf with IQ's shape, singular values)."""
    IQ = np.asarray(IQ)
    nt = IQ.shape[-1]
    out, w = project(casorati(IQ), keep_python(wcut, nt))
    return out.reshape(IQ.shape, order="F"), w


def svd_filter_matlab(IQ, Mask, w, kind="band"):
    """SVDClutterFilter.m:1-45 in FP64: (IQf with IQ's shape, eigenvalues)."""
    IQ = np.asarray(IQ)
    nt = IQ.shape[-1]
    keep = keep_matlab(w, nt, kind)
    Y = casorati(IQ).astype(np.complex128)
    if Mask is not None and np.size(Mask):
        Y = Y / np.sqrt(np.asarray(Mask, dtype=np.float64).reshape(-1, order="F"))[:, None]
    out, ev = project(Y, keep)
    return out.reshape(IQ.shape, order="F"), ev


def tissue_blood_ensemble(npix: int, nt: int, rank: int = 3, tissue_db: float = 40.0,
                          seed: int = 7):
    """Synthetic Doppler ensemble: a rank-``rank`` slowly varying tissue part
    ``tissue_db`` above unit-power random blood.  Returns (Y, tissue, blood)."""
    rng = np.random.default_rng(seed)
    t = np.arange(nt)
    spatial = (rng.standard_normal((npix, rank)) + 1j * rng.standard_normal((npix, rank)))
    temporal = np.stack([np.exp(2j * math.pi * (0.03 * (r + 1)) * t) for r in range(rank)], 1)
    tissue = spatial @ temporal.T * 10 ** (tissue_db / 20) / math.sqrt(2 * rank)
    blood = (rng.standard_normal((npix, nt)) + 1j * rng.standard_normal((npix, nt))) / math.sqrt(2)
    return tissue + blood, tissue, blood
