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
minimax fits.  Each is stored in monic form, E_p(u) = lead * M(u) with
M(u) = u^DEG + a_{DEG-1} u^{DEG-1} + ... + a_0, because the kernels evaluate M
by Horner steps that start with an add of a literal (one register operand)
instead of a move and a two-register FMA, and fold the lead into one multiply
(of the even part) and the e factor (of the odd part):

    q_p = fma(+e * leadO, M_O(u), leadE * M_E(u)),  q_{1-p} likewise with -e.

The coefficients are rounded to FP32 and the fit is checked in that FP32
arithmetic (one rounding per FMA) against the FP64 closed form.  Lanczos a=4
uses degree 3 (error 4.3e-7, a fifth of the FP32 update cost of degree 4 in
the kernels' weights); the 6-tap kinds keep degree 4 (degree 3 would give
1e-6 and 3e-6).  Run:

    python simu_rca_ab_amd/csrc/gen_interp_poly.py > simu_rca_ab_amd/csrc/rca_interp_poly.h
"""
from __future__ import annotations

import numpy as np

# kind id: (name, W, HI, degree in u, max |error| allowed in FP32 evaluation)
KINDS = {1: ("RCA_INTERP_LANCZOS_REF", 2, 3, 4, 2.0e-7), 2: ("RCA_INTERP_LANCZOS3", 3, 3, 4, 2.0e-7),
         3: ("RCA_INTERP_LANCZOS4", 4, 4, 3, 5.0e-7)}
NODES = 400


def lanczos(x, w):
    return np.sinc(x) * np.sinc(x / w)


def fit(w, p, deg):
    t = np.cos(np.pi * (np.arange(NODES) + 0.5) / NODES)
    e = 0.25 * (t + 1.0)
    u = e * e
    f = lambda s: lanczos(s + 0.5 - p, w)
    even = (f(e) + f(-e)) / 2
    odd = (f(e) - f(-e)) / 2
    a = np.linalg.lstsq(np.vander(u, deg + 1, increasing=True), even, rcond=None)[0]
    b = np.linalg.lstsq(e[:, None] * np.vander(u, deg + 1, increasing=True), odd, rcond=None)[0]
    return monic(a), monic(b)


def monic(c):
    """(lead, a_0 .. a_{deg-1}) in FP32 with c(u) = lead * (u^deg + sum a_j u^j)."""
    return np.float32(c[-1]), (c[:-1] / c[-1]).astype(np.float32)


def monic32(a, u):
    """M(u) in the kernels' FP32 steps: an add, then FMAs (one rounding each)."""
    r = (u.astype(np.float64) + np.float64(a[-1])).astype(np.float32)
    for x in a[-2::-1]:
        r = (r.astype(np.float64) * u + np.float64(x)).astype(np.float32)
    return r


def check(w, hi, coef):
    d = np.linspace(0.0, 1.0, 40001, endpoint=False).astype(np.float32)
    e = (d - np.float32(0.5)).astype(np.float32)
    u = (e * e).astype(np.float32)
    worst = 0.0
    for p in range(1, hi + 1):
        (le, a), (lo, b) = coef[p - 1]
        pe = (monic32(a, u) * le).astype(np.float32)
        eo = (e * lo).astype(np.float32).astype(np.float64)
        mo = monic32(b, u).astype(np.float64)
        for k, q in ((p, (eo * mo + pe).astype(np.float32)), (1 - p, (-eo * mo + pe).astype(np.float32))):
            worst = max(worst, float(np.abs(q.astype(np.float64) - lanczos(d.astype(np.float64) - k, w)).max()))
    return worst


def main():
    out = ["// Generated by gen_interp_poly.py: polynomial forms of the Lanczos-type",
           "// interpolation weights (see that file).  Pair p = 1..HI (taps p and 1-p):",
           "// E_p(u) = kEvenLead[p-1] * (u^kDeg + sum_j kEven[p-1][j] u^j), O_p likewise.",
           "#ifndef RCA_INTERP_POLY_H", "#define RCA_INTERP_POLY_H", "",
           "template <int KIND> struct InterpPoly;", ""]
    f = lambda c: f"{float(c):.9e}f"
    for kid, (name, w, hi, deg, bound) in KINDS.items():
        coef = [fit(w, p, deg) for p in range(1, hi + 1)]
        err = check(w, hi, coef)
        assert err <= bound, (name, err)
        out.append(f"// {name}: sinc(x) sinc(x/{w}), taps {1 - hi}..{hi}, degree {deg} in u;"
                   f" max |error| {err:.2e} (FP32)")
        out.append(f"template <> struct InterpPoly<{name}> {{")
        out.append(f"  static constexpr int kDeg = {deg};")
        for tag, idx in (("kEven", 0), ("kOdd", 1)):
            out.append(f"  static constexpr float {tag}Lead[{hi}] = {{"
                       + ", ".join(f(coef[p][idx][0]) for p in range(hi)) + "};")
            rows = ", ".join("{" + ", ".join(f(c) for c in coef[p][idx][1]) + "}" for p in range(hi))
            out.append(f"  static constexpr float {tag}[{hi}][{deg}] = {{{rows}}};")
        out.append("};")
        out.append("")
    out.append("#endif  // RCA_INTERP_POLY_H")
    print("\n".join(out))


if __name__ == "__main__":
    main()
