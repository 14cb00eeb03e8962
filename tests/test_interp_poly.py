"""The polynomial interpolation weights the HIP kernels evaluate
(simu_rca_ab_amd/csrc/rca_interp_poly.h) against the closed form
sinc(x) * sinc(x / W) of each interpolation kind, with the kernels' own tap
assignment and evaluation order: for d = u - floor(u), e = d - 1/2,
u' = e^2 and p = 1 .. HI, with the even and odd parts stored in monic form
lead * M(u'), M(u') = u'^D + a_{D-1} u'^{D-1} + ... + a_0,

    q[p + LO]     = fma(+e * leadO_p, M_O(u'), leadE_p * M_E(u'))   (tap k = p)
    q[1 - p + LO] = fma(-e * leadO_p, M_O(u'), leadE_p * M_E(u'))   (tap k = 1 - p)

(rca_das.hip weight_pair, horner_monic).  The header states the bound per
kind: 2e-7 in FP32 for the 6-tap kinds (degree 4), 5e-7 for Lanczos a=4
(degree 3, chosen for speed: a weight error of 4.3e-7 against the FP32
kernel's 1e-5 relative-L2 bar).  The FP64 evaluation shows the fit itself,
the FP32 evaluation (one rounding per step, as the kernels' instructions)
the whole error.
"""
import re
from pathlib import Path

import numpy as np
import pytest

HDR = Path(__file__).resolve().parent.parent / "simu_rca_ab_amd" / "csrc" / "rca_interp_poly.h"
KINDS = {"RCA_INTERP_LANCZOS_REF": (2, 3, 2, 4, 2e-7), "RCA_INTERP_LANCZOS3": (2, 3, 3, 4, 2e-7),
         "RCA_INTERP_LANCZOS4": (3, 4, 4, 3, 5e-7)}   # LO, HI, W, degree, FP32 bound


def _tables():
    txt = HDR.read_text()
    out = {}
    for kind in KINDS:
        blk = txt[txt.index(f"struct InterpPoly<{kind}>"):]
        blk = blk[:blk.index("\n};\n") + 4]
        arr = {"deg": int(re.search(r"kDeg = (\d+);", blk).group(1))}
        num = lambda t: [float(v.strip().rstrip("f")) for v in t.split(",")]
        for name in ("kEven", "kOdd"):
            arr[name + "Lead"] = np.array(num(re.search(name + r"Lead\[\d+\] = \{(.*?)\};", blk).group(1)))
            body = re.search(name + r"\[\d+\]\[\d+\] = \{(.*?)\};", blk, re.S).group(1)
            arr[name] = np.array([num(r) for r in re.findall(r"\{([^{}]*)\}", body)])
        out[kind] = arr
    return out


def _monic(a, u, dtype):
    r = (u + dtype(a[-1])).astype(dtype)
    for c in a[-2::-1]:
        r = (r.astype(np.float64) * u + np.float64(c)).astype(dtype)
    return r


def _weights(tab, lo, hi, d, dtype):
    e = (d - 0.5).astype(dtype)
    u = (e * e).astype(dtype)
    q = np.zeros((lo + hi + 1, d.size), dtype=dtype)
    for p in range(1, hi + 1):
        pe = (_monic(tab["kEven"][p - 1], u, dtype) * dtype(tab["kEvenLead"][p - 1])).astype(dtype)
        eo = (e * dtype(tab["kOddLead"][p - 1])).astype(np.float64)
        mo = _monic(tab["kOdd"][p - 1], u, dtype).astype(np.float64)
        q[p + lo] = (eo * mo + pe).astype(dtype)
        q[1 - p + lo] = (-eo * mo + pe).astype(dtype)
    return q


@pytest.mark.parametrize("kind", list(KINDS))
def test_polynomial_weights_match_closed_form(kind):
    lo, hi, w, deg, bound = KINDS[kind]
    tab = _tables()[kind]
    assert tab["deg"] == deg
    assert tab["kEven"].shape == (hi, deg) and tab["kOdd"].shape == (hi, deg)
    assert tab["kEvenLead"].shape == (hi,) and tab["kOddLead"].shape == (hi,)
    d = np.linspace(0.0, 1.0, 20001, endpoint=False)
    k = np.arange(-lo, hi + 1)[:, None]
    x = d[None, :] - k
    want = np.sinc(x) * np.sinc(x / w)
    err64 = np.abs(_weights(tab, lo, hi, d, np.float64) - want).max()
    err32 = np.abs(_weights(tab, lo, hi, d.astype(np.float32), np.float32).astype(np.float64) - want).max()
    assert err64 <= 0.9 * bound, err64
    assert err32 <= bound, err32


def test_a_swapped_tap_assignment_fails():
    """Mirroring the pairs (q[p+LO] <-> q[1-p+LO]) must break the match."""
    lo, hi, w = KINDS["RCA_INTERP_LANCZOS4"][:3]
    tab = _tables()["RCA_INTERP_LANCZOS4"]
    d = np.linspace(0.05, 0.45, 101)
    q = _weights(tab, lo, hi, d, np.float64)[::-1]
    x = d[None, :] - np.arange(-lo, hi + 1)[:, None]
    assert np.abs(q - np.sinc(x) * np.sinc(x / w)).max() > 1e-2
