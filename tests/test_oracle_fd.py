"""Pins the oracle's 6th-order FD operators (oracle/oracle.h, restating
src/utils/modelsolver.c:101-327) against what the mathematics fixes:

* closed forms of the discrete operators on Fourier modes sin(k x), sin(a x + b y);
* exactness on polynomials (first derivative up to degree 6, second and cross up to degree 7).

A wrong weight, a dropped i = 3 term, a swapped diagonal pencil or a mixed-up axis spacing fails
one of these.
"""
import numpy as np
import pytest

import pyoracle

H = np.array([0.11, 0.13, 0.09], dtype=np.longdouble)  # distinct spacings catch axis mix-ups


def grid_coords(n, h):
    m = n + 6
    idx = np.arange(m) - 3
    z, y, x = np.meshgrid(idx * h[2], idx * h[1], idx * h[0], indexing="ij")
    return x, y, z


def d1_factor(k, h):
    k, h = np.longdouble(k), np.longdouble(h)
    return (1.5 * np.sin(k * h) - 0.3 * np.sin(2 * k * h) + np.sin(3 * k * h) / 30.0) / h


def d2_factor(k, h):
    k, h = np.longdouble(k), np.longdouble(h)
    return -(49.0 / 18.0 - 3 * np.cos(k * h) + 0.3 * np.cos(2 * k * h) - np.cos(3 * k * h) / 45.0) / h**2


def cross_factor(a, b, ha, hb):
    a, b, ha, hb = (np.longdouble(v) for v in (a, b, ha, hb))
    c = np.array([0.0, 270.0, -27.0, 2.0], dtype=np.longdouble) / 720
    return sum(-4 * c[i] * np.sin(i * a * ha) * np.sin(i * b * hb) for i in (1, 2, 3)) / (ha * hb)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 2e-12), (np.longdouble, 1e-15)])
def test_fourier_closed_forms(dtype, tol):
    n = 10
    x, y, z = (np.asarray(v, dtype=dtype) for v in grid_coords(n, H))
    kx, ky, kz = 1.7, 2.3, 0.9
    cases = [
        # field, derivative index, expected factor times the analytic companion
        (np.sin(kx * x), 1, d1_factor(kx, H[0]) * np.cos(kx * x)),
        (np.sin(kx * x), 4, d2_factor(kx, H[0]) * np.sin(kx * x)),
        (np.sin(ky * y), 2, d1_factor(ky, H[1]) * np.cos(ky * y)),
        (np.sin(ky * y), 5, d2_factor(ky, H[1]) * np.sin(ky * y)),
        (np.sin(kz * z), 3, d1_factor(kz, H[2]) * np.cos(kz * z)),
        (np.sin(kz * z), 6, d2_factor(kz, H[2]) * np.sin(kz * z)),
        (np.sin(kx * x + ky * y), 7, cross_factor(kx, ky, H[0], H[1]) * np.sin(kx * x + ky * y)),
        (np.sin(kx * x + kz * z), 8, cross_factor(kx, kz, H[0], H[2]) * np.sin(kx * x + kz * z)),
        (np.sin(ky * y + kz * z), 9, cross_factor(ky, kz, H[1], H[2]) * np.sin(ky * y + kz * z)),
    ]
    inv = (1.0 / H).astype(np.float64)
    rng = np.random.default_rng(7)
    for field, which, expected in cases:
        f = np.ascontiguousarray(field, dtype=dtype)
        scale = np.max(np.abs(expected)) + 1e-300
        for _ in range(6):
            i, j, k = rng.integers(3, n + 3, size=3)
            got = pyoracle.derivs(f, inv, i, j, k)[which]
            assert abs(float(got) - float(expected[k, j, i])) <= tol * scale, (which, i, j, k)
    # The cross derivative of a field that depends on one axis only vanishes.
    f = np.ascontiguousarray(np.sin(kx * x), dtype=dtype)
    out = pyoracle.derivs(f, inv, 5, 6, 7)
    assert abs(float(out[7])) < 1e-9 and abs(float(out[8])) < 1e-9 and abs(float(out[9])) < 1e-9


def _poly(rng, deg, x, y, z):
    """Random polynomial of total degree <= deg and its exact derivatives."""
    terms = [(a, b, c) for a in range(deg + 1) for b in range(deg + 1) for c in range(deg + 1) if a + b + c <= deg]
    coef = rng.uniform(-1, 1, size=len(terms))
    f = np.zeros_like(x)
    d = {k: np.zeros_like(x) for k in ("x", "y", "z", "xx", "yy", "zz", "xy", "xz", "yz")}

    def mono(a, b, c):
        return (x**a if a >= 0 else 0) * (y**b if b >= 0 else 0) * (z**c if c >= 0 else 0)

    for (a, b, c), w in zip(terms, coef):
        f += w * mono(a, b, c)
        if a >= 1:
            d["x"] += w * a * mono(a - 1, b, c)
        if b >= 1:
            d["y"] += w * b * mono(a, b - 1, c)
        if c >= 1:
            d["z"] += w * c * mono(a, b, c - 1)
        if a >= 2:
            d["xx"] += w * a * (a - 1) * mono(a - 2, b, c)
        if b >= 2:
            d["yy"] += w * b * (b - 1) * mono(a, b - 2, c)
        if c >= 2:
            d["zz"] += w * c * (c - 1) * mono(a, b, c - 2)
        if a >= 1 and b >= 1:
            d["xy"] += w * a * b * mono(a - 1, b - 1, c)
        if a >= 1 and c >= 1:
            d["xz"] += w * a * c * mono(a - 1, b, c - 1)
        if b >= 1 and c >= 1:
            d["yz"] += w * b * c * mono(a, b - 1, c - 1)
    return f, d


def test_polynomial_exactness():
    n = 6
    x, y, z = grid_coords(n, H)
    x, y, z = (np.asarray(v, dtype=np.longdouble) + np.longdouble(0.25) for v in (x, y, z))
    rng = np.random.default_rng(11)
    names = ["x", "y", "z", "xx", "yy", "zz", "xy", "xz", "yz"]
    # degree 6: every operator is exact
    f, d = _poly(rng, 6, x, y, z)
    f = np.ascontiguousarray(f, dtype=np.longdouble)
    inv = (1.0 / H).astype(np.float64)
    for (i, j, k) in [(3, 3, 3), (5, 4, 8), (8, 8, 8), (4, 7, 6)]:
        out = pyoracle.derivs(f, inv, i, j, k)
        assert float(out[0]) == pytest.approx(float(f[k, j, i]), rel=0, abs=0)
        for q, name in enumerate(names):
            exp = float(d[name][k, j, i])
            assert abs(float(out[q + 1]) - exp) <= 1e-11 * (1 + abs(exp)), (name, i, j, k)
