"""The polynomial interpolation weights the HIP kernels evaluate
(simu_rca_ab_amd/csrc/rca_interp_poly.h) against the closed form
sinc(x) * sinc(x / W) of each interpolation kind, with the kernels' own tap
assignment: for d = u - floor(u), e = d - 1/2, u' = e^2 and p = 1 .. HI,

    q[p + LO]     = E_p(u') + e * O_p(u')   (tap k = p,     offset d - p)
    q[1 - p + LO] = E_p(u') - e * O_p(u')   (tap k = 1 - p, offset d - 1 + p)

(rca_das.hip weight_pair).  The header states a bound of 2e-7 in FP32; the
FP64 evaluation shows the fit itself, the FP32 evaluation (Horner steps
rounded to FP32, as the kernels' FMAs are) the whole error.
"""
import re
from pathlib import Path

import numpy as np
import pytest

HDR = Path(__file__).resolve().parent.parent / "simu_rca_ab_amd" / "csrc" / "rca_interp_poly.h"
KINDS = {"RCA_INTERP_LANCZOS_REF": (2, 3, 2), "RCA_INTERP_LANCZOS3": (2, 3, 3),
         "RCA_INTERP_LANCZOS4": (3, 4, 4)}   # LO, HI, W


def _tables():
    txt = HDR.read_text()
    out = {}
    for kind in KINDS:
        blk = txt[txt.index(f"struct InterpPoly<{kind}>"):]
        blk = blk[:blk.index("\n};\n") + 4]
        arr = {}
        for name in ("kEven", "kOdd"):
            body = re.search(name + r"\[\d+\]\[\d+\] = \{(.*?)\};", blk, re.S).group(1)
            rows = re.findall(r"\{([^{}]*)\}", body)
            arr[name] = np.array([[float(v.strip().rstrip("f")) for v in r.split(",")] for r in rows])
        out[kind] = arr
    return out


def _weights(tab, lo, hi, d, dtype):
    e = (d - 0.5).astype(dtype)
    u = (e * e).astype(dtype)
    q = np.zeros((lo + hi + 1, d.size), dtype=dtype)
    for p in range(1, hi + 1):
        ev = np.full_like(u, dtype(tab["kEven"][p - 1][-1]))
        od = np.full_like(u, dtype(tab["kOdd"][p - 1][-1]))
        for c in tab["kEven"][p - 1][-2::-1]:
            ev = (u * ev + dtype(c)).astype(dtype)
        for c in tab["kOdd"][p - 1][-2::-1]:
            od = (u * od + dtype(c)).astype(dtype)
        po = (od * e).astype(dtype)
        q[p + lo] = ev + po
        q[1 - p + lo] = ev - po
    return q


@pytest.mark.parametrize("kind", list(KINDS))
def test_polynomial_weights_match_closed_form(kind):
    lo, hi, w = KINDS[kind]
    tab = _tables()[kind]
    assert tab["kEven"].shape == (hi, 5) and tab["kOdd"].shape == (hi, 5)
    d = np.linspace(0.0, 1.0, 20001, endpoint=False)
    k = np.arange(-lo, hi + 1)[:, None]
    x = d[None, :] - k
    want = np.sinc(x) * np.sinc(x / w)
    err64 = np.abs(_weights(tab, lo, hi, d, np.float64) - want).max()
    err32 = np.abs(_weights(tab, lo, hi, d.astype(np.float32), np.float32).astype(np.float64) - want).max()
    assert err64 <= 1.5e-7, err64
    assert err32 <= 3e-7, err32


def test_a_swapped_tap_assignment_fails():
    """Mirroring the pairs (q[p+LO] <-> q[1-p+LO]) must break the match."""
    lo, hi, w = KINDS["RCA_INTERP_LANCZOS4"]
    tab = _tables()["RCA_INTERP_LANCZOS4"]
    d = np.linspace(0.05, 0.45, 101)
    q = _weights(tab, lo, hi, d, np.float64)[::-1]
    x = d[None, :] - np.arange(-lo, hi + 1)[:, None]
    assert np.abs(q - np.sinc(x) * np.sinc(x / w)).max() > 1e-2
