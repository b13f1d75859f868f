"""The FP32 Bessel factor J1(x)/x for x >= 4 of the disk kernel: the
modulus-phase coefficients in pointint_amd/csrc/j1_modphase.hpp, evaluated in
an FP32 emulation of the kernel's operation order, against
scipy.special.j1 (FP64).  Pins the constants and the phase reduction."""
import os
import re

import numpy as np
import pytest

HDR = os.path.join(os.path.dirname(os.path.abspath(__file__)), os.pardir, "pointint_amd", "csrc",
                   "j1_modphase.hpp")
f32 = np.float32


def _coef(name):
    txt = open(HDR).read()
    m = re.search(name + r"\[\d+\] = \{([^}]*)\}", txt)
    return np.array([float(v.strip().rstrip("f")) for v in m.group(1).split(",")], dtype=np.float32)


def j1ox_large_f32(x):
    """Vectorised FP32 emulation (fma = exact product-sum, then one rounding)."""
    x = np.asarray(x, dtype=np.float32)
    cm, cp = _coef("kJ1Mod"), _coef("kJ1Phase")
    hi = f32(1.0 / (2.0 * np.pi))
    lo = f32(1.0 / (2.0 * np.pi) - float(hi))
    y = (f32(1.0) / x).astype(np.float32)
    sy = (1.0 / np.sqrt(x.astype(np.float64))).astype(np.float32)
    fma = lambda a, b, c: (a.astype(np.float64) * b + c).astype(np.float32)
    m = np.full_like(x, cm[-1])
    p = np.full_like(x, cp[-1])
    for i in range(len(cm) - 2, -1, -1):
        m = fma(m, y, cm[i])
        p = fma(p, y, cp[i])
    r1 = (x * hi).astype(np.float32)
    e1 = (x.astype(np.float64) * float(hi) - r1).astype(np.float32)
    f = ((r1 - np.rint(r1)).astype(np.float32) +
         (fma(x, lo, e1) + (p - f32(0.375))).astype(np.float32)).astype(np.float32)
    f = (f - np.rint(f)).astype(np.float32)
    c = np.cos(2.0 * np.pi * f.astype(np.float64)).astype(np.float32)
    return (m * (y * sy) * c).astype(np.float32)


def test_j1_modphase_against_scipy():
    sp = pytest.importorskip("scipy.special")
    x = np.concatenate([np.linspace(4.0, 12.0, 20001), np.geomspace(12.0, 1e4, 20001)]).astype(np.float32)
    got = j1ox_large_f32(x).astype(np.float64)
    xd = x.astype(np.float64)
    ref = sp.j1(xd) / xd
    err = np.abs(got - ref)
    env = np.sqrt(2.0 / np.pi) * xd ** -1.5
    assert err.max() < 1e-7, err.max()
    assert (err / env).max() < 1e-6, (err / env).max()
    # continuity with the small-argument side: J1(4)/4
    assert abs(j1ox_large_f32(np.array([4.0]))[0] - sp.j1(4.0) / 4.0) < 5e-8


def test_header_matches_generator():
    """The header holds what gen_j1_modphase.py computes (to FP32 rounding)."""
    pytest.importorskip("scipy.special")
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "gen", os.path.join(os.path.dirname(HDR), "gen_j1_modphase.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    cm, cp, em, ep = gen.fit()
    assert np.allclose(_coef("kJ1Mod"), cm.astype(np.float32), rtol=1e-6, atol=1e-12)
    assert np.allclose(_coef("kJ1Phase"), cp.astype(np.float32), rtol=1e-6, atol=1e-12)
    assert em < 1e-8 and ep < 1e-9
