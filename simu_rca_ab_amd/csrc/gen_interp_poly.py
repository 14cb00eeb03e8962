"""Generate rca_interp_poly.h: polynomial forms of the Lanczos-type
interpolation weights used by the DAS kernels.

The weight of tap k at fractional sample position d in [0, 1) is
q_k(d) = L(d - k), L(x) = sinc(x) * sinc(x / W) (sinc(x) = sin(pi x)/(pi x);
bfVirtualSourceGeneric_RCA.cu:79-82, 331-336 with W = 2, taps k = -2..3;
textbook Lanczos a = 3, 4 with W = a, taps k = 1-a..a).  With e = d - 1/2 the
taps pair up: q_{1-k}(e) = q_k(-e), so for p = 1..HI

    q_p     = E_p(e^2) + e * O_p(e^2)
    q_{1-p} = E_p(e^2) - e * O_p(e^2)

with E_p and O_p polynomials of degree DEG in u = e^2 (u in [0, 1/4]).  They
are least-squares fits at Chebyshev nodes of e in [0, 1/2], close to the
minimax fits; the coefficients are rounded to FP32 and the fit is checked in
FP32 Horner arithmetic against the FP64 closed form.  Run:

    python simu_rca_ab_amd/csrc/gen_interp_poly.py > simu_rca_ab_amd/csrc/rca_interp_poly.h
"""
from __future__ import annotations

import numpy as np

DEG = 4                       # u^0 .. u^4 for both parts (e^8, e^9)
KINDS = {1: ("RCA_INTERP_LANCZOS_REF", 2, 3), 2: ("RCA_INTERP_LANCZOS3", 3, 3),
         3: ("RCA_INTERP_LANCZOS4", 4, 4)}   # kind id: (name, W, HI)
NODES = 400
BOUND = 2.0e-7                # max |error| allowed in FP32 evaluation


def lanczos(x, w):
    return np.sinc(x) * np.sinc(x / w)


def fit(w, p):
    t = np.cos(np.pi * (np.arange(NODES) + 0.5) / NODES)
    e = 0.25 * (t + 1.0)
    u = e * e
    f = lambda s: lanczos(s + 0.5 - p, w)
    even = (f(e) + f(-e)) / 2
    odd = (f(e) - f(-e)) / 2
    a = np.linalg.lstsq(np.vander(u, DEG + 1, increasing=True), even, rcond=None)[0]
    b = np.linalg.lstsq(e[:, None] * np.vander(u, DEG + 1, increasing=True), odd, rcond=None)[0]
    return a.astype(np.float32), b.astype(np.float32)


def horner32(c, u):
    r = np.full_like(u, c[-1])
    for x in c[-2::-1]:
        r = (r.astype(np.float64) * u + np.float64(x)).astype(np.float32)   # one rounding per FMA
    return r


def check(w, hi, coef):
    d = np.linspace(0.0, 1.0, 40001, endpoint=False).astype(np.float32)
    e = (d - np.float32(0.5)).astype(np.float32)
    u = (e * e).astype(np.float32)
    worst = 0.0
    for p in range(1, hi + 1):
        a, b = coef[p - 1]
        pe = horner32(a, u)
        po = (horner32(b, u) * e).astype(np.float32)
        for k, q in ((p, pe + po), (1 - p, pe - po)):
            worst = max(worst, float(np.abs(q.astype(np.float64) - lanczos(d.astype(np.float64) - k, w)).max()))
    return worst


def main():
    out = ["// Generated by gen_interp_poly.py: polynomial forms of the Lanczos-type",
           "// interpolation weights (see that file).  kEven[p-1][j], kOdd[p-1][j]: the",
           "// coefficient of u^j of E_p and O_p; taps p = 1..HI and 1-p.",
           "#ifndef RCA_INTERP_POLY_H", "#define RCA_INTERP_POLY_H", "",
           f"constexpr int kPolyDeg = {DEG};", "",
           "template <int KIND> struct InterpPoly;", ""]
    for kid, (name, w, hi) in KINDS.items():
        coef = [fit(w, p) for p in range(1, hi + 1)]
        err = check(w, hi, coef)
        assert err <= BOUND, (name, err)
        out.append(f"// {name}: sinc(x) sinc(x/{w}), taps {1 - hi}..{hi}; max |error| {err:.2e} (FP32)")
        out.append(f"template <> struct InterpPoly<{name}> {{")
        for tag, idx in (("kEven", 0), ("kOdd", 1)):
            rows = ", ".join("{" + ", ".join(f"{float(c):.9e}f" for c in coef[p][idx]) + "}"
                             for p in range(hi))
            out.append(f"  static constexpr float {tag}[{hi}][{DEG + 1}] = {{{rows}}};")
        out.append("};")
        out.append("")
    out.append("#endif  // RCA_INTERP_POLY_H")
    print("\n".join(out))


if __name__ == "__main__":
    main()
