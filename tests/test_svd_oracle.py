"""Pins the FP64 SVD clutter-filter oracle (oracle/svd_oracle.py) on what the
mathematics and the reference's index conventions fix (CPU only)."""
import numpy as np
import pytest

from oracle import svd_oracle as S


def rand(shape, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(shape) + 1j * r.standard_normal(shape)


def test_casorati_is_f_order():
    IQ = rand((5, 4, 6), 1)
    Y = S.casorati(IQ)
    for t in range(6):
        np.testing.assert_array_equal(Y[:, t], IQ[:, :, t].ravel(order="F"))
    assert Y[5 * 1 + 2, 3] == IQ[2, 1, 3]   # p = iz + Nz*ix


def test_eigenvalues_are_squared_singular_values():
    Y = rand((300, 20), 2)
    w, V = S.gram_eig(Y)
    sv = np.linalg.svd(Y, compute_uv=False)
    np.testing.assert_allclose(w, sv ** 2, rtol=1e-10)
    assert np.all(np.diff(w) <= 0)
    np.testing.assert_allclose(w.sum(), np.vdot(Y, Y).real, rtol=1e-12)
    np.testing.assert_allclose(V.conj().T @ V, np.eye(20), atol=1e-12)


def test_projector_properties_and_complement():
    Y = rand((200, 16), 3)
    keep = np.zeros(16, bool)
    keep[[1, 2, 5, 9, 15]] = True
    a, _ = S.project(Y, keep)
    b, _ = S.project(Y, ~keep)
    np.testing.assert_allclose(a + b, Y, atol=1e-11)
    _, V = S.gram_eig(Y)
    P = V[:, keep] @ V[:, keep].conj().T
    np.testing.assert_allclose(P @ P, P, atol=1e-12)
    np.testing.assert_allclose(P, P.conj().T, atol=1e-12)
    np.testing.assert_allclose(np.trace(P).real, keep.sum(), atol=1e-12)
    all_, _ = S.project(Y, np.ones(16, bool))
    np.testing.assert_allclose(all_, Y, atol=1e-11)
    none, _ = S.project(Y, np.zeros(16, bool))
    assert np.abs(none).max() == 0


def test_projection_equals_right_singular_subspace():
    # an independent route: Y = U S W^H; keeping the first k columns of W
    Y = rand((150, 12), 4)
    _, s, Wh = np.linalg.svd(Y, full_matrices=False)
    k = 5
    ref = Y @ Wh[:k].conj().T @ Wh[:k]
    keep = np.zeros(12, bool)
    keep[:k] = True
    got, _ = S.project(Y, keep)
    np.testing.assert_allclose(got, ref, atol=1e-10)


def test_high_pass_removes_a_low_rank_tissue_signal():
    Y, tissue, blood = S.tissue_blood_ensemble(2000, 64, rank=3, tissue_db=40.0)
    keep = np.ones(64, bool)
    keep[:3] = False
    got, w = S.project(tissue, keep)            # tissue alone: rank 3, fully removed
    assert np.linalg.norm(got) <= 1e-9 * np.linalg.norm(tissue)
    got, w = S.project(Y, keep)                 # with blood: tissue energy gone
    assert w[2] > 100 * w[3]                    # 40 dB gap in the spectrum
    # the tissue's temporal subspace is suppressed (> 40 dB) and the output
    # keeps the blood energy outside that subspace
    t = np.arange(64)
    Q, _ = np.linalg.qr(np.stack([np.exp(2j * np.pi * 0.03 * (r + 1) * t) for r in range(3)], 1))
    assert np.linalg.norm(got @ Q) ** 2 <= 1e-4 * np.linalg.norm(tissue) ** 2
    outside = np.linalg.norm(blood - (blood @ Q) @ Q.conj().T) ** 2
    assert abs(np.linalg.norm(got) ** 2 / outside - 1) <= 1e-2


def test_python_and_matlab_band_conventions():
    nt = 200
    k = S.keep_python([0.1, 1.0], nt)          # u[:, 20:200]
    assert k.sum() == 180 and k[20] and not k[19]
    k = S.keep_matlab([0.1, 1.0], nt)          # Eig_vect(:, 20:200), 1-based
    assert k.sum() == 181 and k[19] and not k[18]
    assert S.keep_matlab([0.1], nt, "low").sum() == 20
    k = S.keep_matlab([0.1], nt, "high")
    assert k.sum() == 181 and k[19]
    k = S.keep_matlab([0.1, 0.9], nt, "stop")
    assert k[:20].all() and not k[20:179].any() and k[179:].all()
    assert S.keep_matlab([0.0125], 200, "low").sum() == 3   # round(2.5) = 3, halves away from 0
    with pytest.raises(ValueError):
        S.keep_matlab([-0.1, 0.5], nt)
    with pytest.raises(IndexError):
        S.keep_matlab([0.0, 0.5], nt)          # Eig_vect(:, 0:100) is an error in MATLAB


def test_matlab_mask_weighting_is_linear():
    IQ = rand((10, 8, 16), 5)
    a, ea = S.svd_filter_matlab(IQ, None, [0.2, 0.8])
    b, eb = S.svd_filter_matlab(IQ, np.full((10, 8), 4.0), [0.2, 0.8])
    np.testing.assert_allclose(b, a / 2, atol=1e-12)
    np.testing.assert_allclose(eb, ea / 4, rtol=1e-12)
    c, _ = S.svd_filter_python(IQ, [0.2, 0.8])   # python: u[:, 3:12]
    d, _ = S.project(S.casorati(IQ), S.keep_python([0.2, 0.8], 16))
    np.testing.assert_allclose(S.casorati(c), d, atol=1e-12)
