"""Pins the FP64 SIMUS-RCA simulator oracle (CPU, no GPU needed).

The reference ships no tests or fixtures for the simulator; the oracle is
pinned by what the mathematics fixes:
  * the complex sinc equals numpy's sinc on the real axis and sin(x)/x off it,
    with the reference's truncation at |x| >= pi;
  * the four-Gaussian elevation term equals the Gaussian-beam integral it
    closes, by dense numerical quadrature (sign of every exponent, the
    sqrt(pi/alpha) factor);
  * on axis (zero azimuth and elevation) the field of a point element is
    Green's function exp(jkr)/r with attenuation exp(-Im(k) r);
  * off axis the azimuth directivity of a flat element is sinc(k w sin/2);
    in elevation the flat branch uses the width (as the reference does) and
    the height only matters to the focused branch;
  * a transmit delay d multiplies the transmit field by exp(j w d);
  * the output is linear in the backscatter coefficients and sums over
    scatterers; a receive aberration law of zero delay and apodisation a
    scales the output by a;
  * real(ifft) of the half spectrum equals an explicit DFT sum, and the
    sample-mode decimation keeps samples 1 and 2 of every ``sub``;
  * the time/frequency grid follows the wrapper's closed forms.
"""
import math

import numpy as np
import pytest

from oracle import simus_oracle as S
from tests.simus_cases import rca_case


def test_csinc_real_axis_and_truncation():
    x = np.linspace(-5, 5, 2001)
    want = np.where(np.abs(x) < S.PI_REF, np.sinc(x / np.pi), 0.0)
    np.testing.assert_allclose(S.csinc(x).real, want, atol=1e-15)
    z = np.array([0.3 + 0.2j, -1.1 + 0.05j, 2.0 - 0.4j])
    np.testing.assert_allclose(S.csinc(z), (np.exp(1j * z) - np.exp(-1j * z)) / (2j * z), rtol=1e-13)
    assert S.csinc(np.array([0.0 + 0j]))[0] == 1
    assert S.csinc(np.array([3.2 + 0.0j]))[0] == 0


@pytest.mark.parametrize("r,y,rf", [(0.03, 0.002, 0.02), (0.015, -0.004, 0.025), (0.04, 0.0, 0.04)])
def test_dir_curv_is_the_gaussian_beam_integral(r, y, rf):
    k = 2 * np.pi * 7.5e6 / 1540 + 3.0j
    h = 3e-3
    u = np.linspace(-6 * h, 6 * h, 400001)
    w = sum(a * np.exp(-b * (u / h) ** 2) for a, b in zip(S.GA, S.GB))
    phase = np.exp(-1j * k * (r - rf) / (2 * r * rf) * u ** 2 - 1j * k * y * u / r)
    integral = np.trapezoid(w * phase, u) * np.exp(1j * k * y * y / (2 * r))
    # sqrt(PI_REF/alpha) uses the kernels' FP32 pi: a relative offset of 5e-9
    np.testing.assert_allclose(S.dir_curv(r, y, k, h, rf), integral, rtol=1e-6)


def _one_element(width=1e-4, az=0.0, el=0.0):
    return np.array([[[0.0], [0.0], [0.0], [az], [el]]])


@pytest.mark.parametrize("rfocus", [10000.0, 0.02])
def test_on_axis_point_element_is_greens_function(rfocus):
    k = np.array([2 * np.pi * 5e6 / 1540 + 0.7j, 2 * np.pi * 8e6 / 1540 + 1.1j])
    z = 0.023
    g = S.element_terms(np.array([[0.0, 0.0, z]]), _one_element(), k, 1540.0, 2e-4, 2e-3, rfocus,
                        0, 0.0)[:, 0, 0]
    if rfocus > 10:
        np.testing.assert_allclose(np.abs(g) * z, np.exp(-k.imag * z), rtol=1e-13)
        np.testing.assert_allclose(np.angle(g), np.angle(np.exp(1j * k.real * z)), atol=1e-9)
    else:   # times the focused elevation term of y = 0
        np.testing.assert_allclose(g, np.exp(1j * k * z) / z * S.dir_curv(z, 0.0, k, 2e-3, rfocus),
                                   rtol=1e-12)


def test_flat_directivity_off_axis_and_width_in_elevation():
    k = np.array([2 * np.pi * 6e6 / 1540 + 0.0j])
    w = 1.2e-4
    r = 0.02
    for th in np.deg2rad([3.0, 11.0, 25.0]):
        pos = np.array([[r * np.sin(th), 0.0, r * np.cos(th)]])
        g = S.element_terms(pos, _one_element(), k, 1540.0, w, 5e-3, 10000.0, 0, 0.0)[0, 0, 0]
        want = np.sinc(np.sin(th) * w * k.real[0] / 2 / np.pi) / r
        np.testing.assert_allclose(abs(g), abs(want), rtol=1e-12)
        # same angle in elevation: the flat branch uses the width, not the height
        pos_e = np.array([[0.0, r * np.sin(th), r * np.cos(th)]])
        ge = S.element_terms(pos_e, _one_element(), k, 1540.0, w, 5e-3, 10000.0, 0, 0.0)[0, 0, 0]
        np.testing.assert_allclose(ge, g, rtol=1e-12)
        ge2 = S.element_terms(pos_e, _one_element(), k, 1540.0, w, 9e-3, 10000.0, 0, 0.0)[0, 0, 0]
        assert ge2 == ge
        gf = S.element_terms(pos_e, _one_element(), k, 1540.0, w, 9e-3, 0.02, 0, 0.0)[0, 0, 0]
        gf2 = S.element_terms(pos_e, _one_element(), k, 1540.0, w, 5e-3, 0.02, 0, 0.0)[0, 0, 0]
        assert abs(gf - gf2) > 1e-3 * abs(gf)
    # baffle 1 multiplies by cos(theta) per axis
    th = np.deg2rad(17.0)
    pos = np.array([[r * np.sin(th), 0.0, r * np.cos(th)]])
    g0 = S.element_terms(pos, _one_element(), k, 1540.0, w, 5e-3, 10000.0, 0, 0.0)[0, 0, 0]
    g1 = S.element_terms(pos, _one_element(), k, 1540.0, w, 5e-3, 10000.0, 1, 0.0)[0, 0, 0]
    np.testing.assert_allclose(g1, g0 * np.cos(th), rtol=1e-12)


def test_transmit_delay_and_lens_are_phases():
    M, TT, TR, TX, RX = rca_case(n_ele=4, n_sub=2, n_tx=1, n_scat=3)
    P = S.prepare(M, TT, TR, TX, RX)
    pos = M["scattererPos"][:, :3, :1]
    prm = S.transducer_scalars(TT)
    c = float(np.float32(M["speedOfSound"]))
    a = S.txfield(pos, TT["ElementPos"], P["k"], P["pulse"], P["ir_tx"], P["delay"], P["apod"], *prm, c)
    d = 3.7e-7
    b = S.txfield(pos, TT["ElementPos"], P["k"], P["pulse"], P["ir_tx"], P["delay"].astype(np.float64) + d, P["apod"], *prm, c)
    w = P["k"].real.astype(np.float64) * c
    np.testing.assert_allclose(b, a * np.exp(1j * w * d)[:, None, None], rtol=1e-11, atol=1e-9 * np.abs(a).max())
    prm_l = list(prm)
    prm_l[4] = 2.5e-7
    e = S.txfield(pos, TT["ElementPos"], P["k"], P["pulse"], P["ir_tx"], P["delay"], P["apod"], *prm_l, c)
    np.testing.assert_allclose(e, a * np.exp(1j * w * 2.5e-7)[:, None, None], rtol=1e-11,
                               atol=1e-9 * np.abs(a).max())


def test_linearity_superposition_and_ablaw():
    M, TT, TR, TX, RX = rca_case(n_ele=6, n_sub=2, n_tx=2, n_scat=4)
    pos = M["scattererPos"].copy()
    r_all = S.simulate(M, TT, TR, TX, RX)
    parts = 0
    for s in range(pos.shape[0]):
        one = pos.copy()
        one[np.arange(pos.shape[0]) != s, 3, :] = 0
        parts = parts + S.simulate(dict(M, scattererPos=one), TT, TR, TX, RX)
    np.testing.assert_allclose(parts, r_all, rtol=1e-10, atol=1e-10 * np.abs(r_all).max())
    twice = pos.copy()
    twice[:, 3, :] *= 2
    np.testing.assert_allclose(S.simulate(dict(M, scattererPos=twice), TT, TR, TX, RX), 2 * r_all,
                               rtol=1e-12, atol=1e-12 * np.abs(r_all).max())
    raw, rab = S.simulate(M, TT, TR, TX, RX, AbLaw={"Delay": np.zeros(6), "Apod": np.full(6, 0.5)})
    np.testing.assert_allclose(rab, 0.5 * raw, rtol=1e-12, atol=1e-12 * np.abs(raw).max())
    assert np.abs(raw).max() > 0


def test_inverse_transform_and_decimation():
    rng = np.random.default_rng(4)
    n, nf = 64, 32
    X = rng.standard_normal(nf) + 1j * rng.standard_normal(nf)
    t = np.arange(n)
    direct = np.array([sum((X[i] * np.exp(2j * np.pi * i * m / n)).real for i in range(nf)) / n
                       for m in t])
    np.testing.assert_allclose(np.real(np.fft.ifft(X, n)), direct, atol=1e-13)
    RF = np.arange(40, dtype=float).reshape(20, 2)
    raw = S.to_raw(RF, 16, 4)            # trimmed to 16 rows, 4 kept in pairs
    np.testing.assert_array_equal(raw[:, 0], [0, 2, 8, 10, 16, 18, 24, 26])
    raw = S.to_raw(RF, 24, 8)            # zero-padded to 24 rows
    np.testing.assert_array_equal(raw[:, 1], [1, 3, 17, 19, 33, 35])


def test_prepare_closed_forms():
    M, TT, TR, TX, RX = rca_case(n_sample=200)
    P = S.prepare(M, TT, TR, TX, RX)
    fs = RX[0]["samplingFrequency"]
    t_start = RX[0]["startTime"] - 1 / TT["frequency"]
    tmax = 200 / fs + np.size(TX[0]["pulseShape"]) / 250e6 + max(np.max(t["Delay"]) for t in TX) + t_start
    assert P["nfft"] == 128 * math.ceil(tmax * fs / 128) and P["nf"] == P["nfft"] // 2
    np.testing.assert_allclose(P["f"], np.linspace(0, fs / 2, P["nf"]), rtol=1e-7)
    np.testing.assert_allclose(P["k"].real, 2 * np.pi * P["f"].astype(float) / 1540, rtol=1e-6)
    np.testing.assert_allclose(P["k"].imag, 0.5 / 8.7 * P["f"].astype(float) / 1e4, rtol=1e-6)
    assert P["sub"] == 4
    np.testing.assert_allclose(P["delay"][:, 1], np.float32(TX[1]["Delay"] - t_start), rtol=1e-6)
