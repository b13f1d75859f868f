"""Writes j1_modphase.hpp: the FP32 coefficients of J1(x) for x >= 4 in
modulus-phase form, used by the disk branch of the fast form-factor kernel.

With H1(x) = J1(x) + i Y1(x) = M(x) e^{i theta(x)} and y = 1/x:

    M(x) = m(y) sqrt(y),   theta(x) = x - 3 pi/4 + 2 pi p(y),

m and p smooth on y in [0, 1/4] (m(0) = sqrt(2/pi), p(0) = 0).  Both are
fitted as degree-7 polynomials in y by least squares on 200,000 points
(relative error for m, absolute for p), against scipy.special.j1 / y1 in
FP64.  The residuals are printed: m relative 2.3e-9, p 1e-10 revolutions.

Run from the repository root:  python pointint_amd/csrc/gen_j1_modphase.py
"""
import os

import numpy as np
from scipy.special import j1, y1

DEG = 7
X0 = 4.0


def fit():
    ymax = 1.0 / X0
    y = np.linspace(0.0, ymax, 200001)[1:]
    x = 1.0 / y
    z = (j1(x) + 1j * y1(x)) * np.exp(-1j * (x - 0.75 * np.pi))
    m = np.abs(z) * np.sqrt(x)
    p = np.angle(z) / (2.0 * np.pi)
    y = np.concatenate([[0.0], y])
    m = np.concatenate([[np.sqrt(2.0 / np.pi)], m])
    p = np.concatenate([[0.0], p])
    s = y / ymax  # scaled variable for conditioning
    V = np.vander(s, DEG + 1, increasing=True)
    cm = np.linalg.lstsq(V / m[:, None], np.ones_like(m), rcond=None)[0]
    # p has no constant term: fit p(y) / y's polynomial through the columns s^1..s^DEG
    cp = np.linalg.lstsq(V[:, 1:], p, rcond=None)[0]
    em = np.max(np.abs(V @ cm - m) / m)
    ep = np.max(np.abs(V[:, 1:] @ cp - p))
    scale = ymax ** np.arange(DEG + 1)
    return cm / scale, np.concatenate([[0.0], cp / scale[1:]]), em, ep


def main():
    cm, cp, em, ep = fit()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "j1_modphase.hpp")
    fmt = lambda c: ", ".join("%.9ef" % v for v in c)
    with open(out, "w") as f:
        f.write("// Generated by gen_j1_modphase.py: J1(x) = m(y) sqrt(y) cos(x - 3 pi/4 + 2 pi p(y)),\n")
        f.write("// y = 1/x, x >= 4; coefficients of y^0..y^%d (least squares against\n" % DEG)
        f.write("// scipy.special.j1 / y1: m relative error %.1e, p %.1e revolutions).\n" % (em, ep))
        f.write("#pragma once\n\nnamespace pii {\n\n")
        f.write("constexpr int kJ1ModPhaseDeg = %d;\n" % DEG)
        f.write("constexpr float kJ1Mod[%d] = {%s};\n" % (DEG + 1, fmt(cm)))
        f.write("constexpr float kJ1Phase[%d] = {%s};\n" % (DEG + 1, fmt(cp)))
        f.write("\n}  // namespace pii\n")
    print(f"wrote {out}: m rel {em:.2e}, p abs {ep:.2e} rev")


if __name__ == "__main__":
    main()
