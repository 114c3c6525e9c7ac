"""Pin the oracle's SPH kernels (src/core/kernel.cpp:112-234) with the
reference's own known-answer tests and closed forms.

* unit integral over the support by Monte Carlo, |I - 1| < 0.05
  (src/tests/test_kernel.cpp:184-205; 65,536 samples, h = 1);
* central finite differences of W match dW and d2W within 1e-3 at h = 10,
  e = 0.01 over 1,024 random radii (src/tests/test_kernel.cpp:207-266);
* closed forms at r = 0 and the cutoff rules.
A wrong coefficient, power or sign in W, dW or d2W fails one of these.
"""
import math

import numpy as np
import pytest


def _mc_integral(oracle, name, samples=65536, seed=7):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1.0, 1.0, size=(samples, 3))
    d = (u * u).sum(axis=1)
    inside = (d - 1.0 < 0) | (np.abs(d - 1.0) < 1e-8)
    acc = sum(oracle.kernel(name, 1.0, math.sqrt(x)) for x in d[inside])
    return acc * 8.0 / samples


@pytest.mark.parametrize("name", ["poly6_W", "spiky_W"])
def test_unit_integral_monte_carlo(oracle, name):
    assert abs(_mc_integral(oracle, name) - 1.0) < 0.05


def test_unit_integral_quadrature(oracle):
    # 4 pi int_0^h W(r) r^2 dr = 1 to quadrature accuracy (tighter than the MC KAT)
    for name in ("poly6_W", "spiky_W"):
        for h in (0.02, 1.0, 3.0):
            r = np.linspace(0.0, h, 20001)
            w = np.array([oracle.kernel(name, h, x) for x in r])
            val = 4.0 * math.pi * np.trapezoid(w * r * r, r)
            assert abs(val - 1.0) < 2e-6, (name, h, val)


def test_derivatives_finite_difference(oracle):
    rng = np.random.default_rng(3)
    e, rad = 0.01, 10.0
    n = 0
    while n < 1024:
        p = (2.0 * rng.uniform(size=3) - 1.0) * rad
        d = float(np.linalg.norm(p))
        if d > rad - e:
            continue
        n += 1
        for w, dw, d2w in (("poly6_W", "poly6_dW", "poly6_d2W"), ("spiky_W", "spiky_dW", "spiky_d2W")):
            Wp, W0, Wm = (oracle.kernel(w, rad, d + e), oracle.kernel(w, rad, d),
                          oracle.kernel(w, rad, d - e))
            fd1 = (Wp - Wm) / (2.0 * e)
            fd2 = (Wp - 2.0 * W0 + Wm) / (e * e)
            if w == "spiky_W" and d < e:
                continue  # spiky is not smooth at r = 0
            assert abs(fd1 - oracle.kernel(dw, rad, d)) < 1e-3
            assert abs(fd2 - oracle.kernel(d2w, rad, d)) < 1e-3


def test_closed_forms_at_origin(oracle):
    for h in (0.02, 0.5, 1.0):
        assert oracle.kernel("poly6_W", h, 0.0) == pytest.approx(315.0 / (64.0 * math.pi * h ** 3), rel=1e-15)
        assert oracle.kernel("spiky_W", h, 0.0) == pytest.approx(15.0 / (math.pi * h ** 3), rel=1e-15)
        assert oracle.kernel("spiky_dW", h, 0.0) == pytest.approx(-45.0 / (math.pi * h ** 4), rel=1e-15)
        assert oracle.kernel("spiky_d2W", h, 0.0) == pytest.approx(90.0 / (math.pi * h ** 5), rel=1e-15)


def test_cutoff_rules(oracle):
    h = 0.02
    # IsZero(of) || of > 0 -> 0, with of = d^2 - h^2 (poly6) or d - h (spiky)
    assert oracle.kernel("poly6_W", h, h) == 0.0
    assert oracle.kernel("spiky_W", h, h) == 0.0
    assert oracle.kernel("poly6_W", h, 1.1 * h) == 0.0
    assert oracle.kernel("spiky_dW", h, 1.1 * h) == 0.0
    # poly6 cutoff margin is 1e-8 in d^2: just inside by 2e-8 in d^2 is non-zero
    r_in = math.sqrt(h * h - 2e-8)
    assert oracle.kernel("poly6_W", h, r_in) > 0.0
    r_edge = math.sqrt(h * h - 5e-9)
    assert oracle.kernel("poly6_W", h, r_edge) == 0.0
    assert oracle.lib().bbo_within_std(r_in, h) == 1
    assert oracle.lib().bbo_within_std(r_edge, h) == 0
    assert oracle.lib().bbo_within_spiky(h - 2e-8, h) == 1
    assert oracle.lib().bbo_within_spiky(h - 5e-9, h) == 0


def test_gradient_sign_points_inward(oracle):
    # spiky dW < 0 inside the support, so gradW(r, (xj-xi)/r) = -dW dir points to j
    for r in (0.001, 0.01, 0.019):
        assert oracle.kernel("spiky_dW", 0.02, r) < 0
