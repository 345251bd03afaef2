"""Pins the oracle's SPH kernels (rx_kernel.h) with closed forms, independent of the formulas' retyping."""
import math

import numpy as np
import pytest


def _ball_integral(f, h, n=4000):
    # int_0^h f(r) 4 pi r^2 dr by Gauss-Legendre
    x, w = np.polynomial.legendre.leggauss(n)
    r = 0.5 * h * (x + 1.0)
    return float(np.sum(w * 0.5 * h * np.array([f(ri) for ri in r]) * 4.0 * math.pi * r * r))


@pytest.mark.parametrize("h", [0.02, 0.1, 1.0])
def test_poly6_normalised(orc, h):
    a = orc.orc_kernel_coef(0, h, 1)
    assert _ball_integral(lambda r: orc.orc_kernel_poly6(r, h, a), h) == pytest.approx(1.0, rel=1e-9)
    assert orc.orc_kernel_poly6(h, h, a) == 0.0
    assert orc.orc_kernel_poly6(1.0001 * h, h, a) == 0.0


@pytest.mark.parametrize("h", [0.02, 0.1])
def test_spiky_gradient_is_derivative_of_normalised_spiky(orc, h):
    # W_spiky = 15/(pi h^6)(h-r)^3 integrates to 1; its radial derivative is -45/(pi h^6)(h-r)^2, so the
    # oracle's gradient factor a (h-r)^2 / r times r must equal dW/dr.
    aw = 15.0 / (math.pi * h ** 6)
    w = lambda r: aw * (h - r) ** 3
    assert _ball_integral(w, h) == pytest.approx(1.0, rel=1e-9)
    ag = orc.orc_kernel_coef(1, h, 2)
    for r in np.linspace(0.05 * h, 0.95 * h, 7):
        e = 1e-7 * h
        dwdr = (w(r + e) - w(r - e)) / (2 * e)
        assert orc.orc_kernel_spiky_grad(r, h, ag) * r == pytest.approx(dwdr, rel=1e-6)
    assert orc.orc_kernel_spiky_grad(0.0, h, ag) == 0.0
    assert orc.orc_kernel_spiky_grad(h, h, ag) == 0.0


@pytest.mark.parametrize("h", [0.02, 0.1])
def test_visc_laplacian_closed_form(orc, h):
    # Mueller 2003 viscosity kernel: laplacian = 45/(pi h^6) (h - r); also the Laplacian of
    # W_visc = 15/(2 pi h^3)(-q^3/2 + q^2 + 1/(2q) - 1) computed numerically.
    al = orc.orc_kernel_coef(2, h, 3)
    a1 = 15.0 / (2 * math.pi * h ** 3)
    W = lambda r: a1 * (-(r / h) ** 3 / 2 + (r / h) ** 2 + h / (2 * r) - 1)
    for r in np.linspace(0.2 * h, 0.9 * h, 5):
        e = 1e-5 * h
        lap = (W(r + e) - 2 * W(r) + W(r - e)) / e ** 2 + 2 / r * (W(r + e) - W(r - e)) / (2 * e)
        assert orc.orc_kernel_visc_lap(r, h, al) == pytest.approx(lap, rel=1e-4)
        assert orc.orc_kernel_visc_lap(r, h, al) == pytest.approx(45 / (math.pi * h ** 6) * (h - r), rel=1e-12)


def test_rest_density_lattice_count(orc):
    # spacing h/2: the lattice points within r <= h are 33 (0, 6 at s, 12 at s sqrt2, 8 at s sqrt3, 6 at 2s)
    h = 0.02
    s = 0.5 * h
    a = orc.orc_kernel_coef(0, h, 1)
    pts = [(0, 1), (s, 6), (s * math.sqrt(2), 12), (s * math.sqrt(3), 8), (2 * s, 6)]
    expect = sum(c * orc.orc_kernel_poly6(r, h, a) for r, c in pts)
    assert orc.orc_rest_density(1.0, h, s) == pytest.approx(expect, rel=1e-12)
