"""Pins the oracle's MHD right-hand side (oracle/oracle.h, restating
src/utils/modelsolver.c:537-777) against an independent formulation of the same PDEs.

On polynomial fields of total degree <= 3 every 6th-order FD operator is exact, so the discrete
RHS at each interior cell must equal the continuous RHS evaluated analytically. The continuous
RHS here is derived symbolically (sympy) from physically written forms that differ from the
oracle's algebra:
  momentum    -(u.grad)u - grad(p)/rho + (J x B)/rho + div(2 rho nu S)/rho + zeta grad(div u),
              with p = rho cs^2 / gamma and J = curl(curl A) / mu0
  continuity  -div(rho u) / rho
  induction   u x curl(A) + eta lap(A)
  entropy     -u.grad(s) + (eta mu0 J^2 + 2 rho nu S:S + zeta rho (div u)^2) / (rho T)
              + div(K grad T) / (rho T),  T = exp(lnT)
A dropped term, a wrong sign or coefficient, or a transposed tensor index in the oracle changes
the result far beyond the tolerance.
"""
import numpy as np
import pytest

import pyoracle

sympy = pytest.importorskip("sympy")

H = (0.11, 0.13, 0.09)
PAR = dict(nu=0.013, eta=0.007, cs=0.9, gamma=5.0 / 3.0, zeta=0.02, mu0=1.3, cp=1.1, lnT0=1.2,
           lnrho0=1.3)
K_THERMAL = 0.001


def real_params():
    from astaroth_amd.abi import RP, NUM_REAL_PARAMS

    rp = np.zeros(NUM_REAL_PARAMS)
    rp[RP["AC_dsx"]], rp[RP["AC_dsy"]], rp[RP["AC_dsz"]] = H
    rp[RP["AC_nu_visc"]] = PAR["nu"]
    rp[RP["AC_eta"]] = PAR["eta"]
    rp[RP["AC_cs_sound"]] = PAR["cs"]
    rp[RP["AC_gamma"]] = PAR["gamma"]
    rp[RP["AC_zeta"]] = PAR["zeta"]
    rp[RP["AC_mu0"]] = PAR["mu0"]
    rp[RP["AC_cp_sound"]] = PAR["cp"]
    rp[RP["AC_lnT0"]] = PAR["lnT0"]
    rp[RP["AC_lnrho0"]] = PAR["lnrho0"]
    return rp


def symbolic_rhs(seed=3):
    x, y, z = sympy.symbols("x y z", real=True)
    X = (x, y, z)
    rng = np.random.default_rng(seed)

    def poly(scale):
        terms = [(a, b, c) for a in range(4) for b in range(4) for c in range(4) if a + b + c <= 3]
        return sum(sympy.Rational(int(rng.integers(-99, 100)), 100) * scale * x**a * y**b * z**c
                   for (a, b, c) in terms)

    lnrho = poly(sympy.Rational(1, 10))
    u = [poly(sympy.Rational(1, 20)) for _ in range(3)]
    A = [poly(sympy.Rational(1, 10)) for _ in range(3)]
    s = poly(sympy.Rational(1, 20))

    nu, eta, cs, gamma, zeta, mu0, cp, lnT0, lnrho0 = (sympy.nsimplify(PAR[k]) for k in
                                                        ("nu", "eta", "cs", "gamma", "zeta", "mu0", "cp", "lnT0", "lnrho0"))
    d = sympy.diff
    rho = sympy.exp(lnrho)

    def grad(f):
        return [d(f, v) for v in X]

    def div(v):
        return sum(d(v[i], X[i]) for i in range(3))

    def curl(v):
        return [d(v[2], y) - d(v[1], z), d(v[0], z) - d(v[2], x), d(v[1], x) - d(v[0], y)]

    def cross(a, b):
        return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]

    def lap(f):
        return sum(d(f, v, 2) for v in X)

    divu = div(u)
    S = [[sympy.Rational(1, 2) * (d(u[i], X[j]) + d(u[j], X[i])) - (sympy.Rational(1, 3) * divu if i == j else 0)
          for j in range(3)] for i in range(3)]
    cs2 = cs**2 * sympy.exp(gamma * s / cp + (gamma - 1) * (lnrho - lnrho0))
    p = rho * cs2 / gamma
    B = curl(A)
    J = [c / mu0 for c in curl(B)]
    JxB = cross(J, B)
    visc = [sum(d(2 * rho * nu * S[i][j], X[j]) for j in range(3)) / rho for i in range(3)]
    mom = [-sum(u[j] * d(u[i], X[j]) for j in range(3)) - d(p, X[i]) / rho + JxB[i] / rho + visc[i]
           + zeta * d(divu, X[i]) for i in range(3)]
    cont = -div([rho * ui for ui in u]) / rho
    ind = [cross(u, B)[i] + eta * lap(A[i]) for i in range(3)]
    T = sympy.exp(lnT0 + gamma * s / cp + (gamma - 1) * (lnrho - lnrho0))
    SS = sum(S[i][j] ** 2 for i in range(3) for j in range(3))
    heating = eta * mu0 * sum(Ji**2 for Ji in J) + 2 * rho * nu * SS + zeta * rho * divu**2
    Kt = sympy.Rational(1, 1000)
    heat = div([Kt * g for g in grad(T)]) / (rho * T)
    ent = -sum(u[i] * d(s, X[i]) for i in range(3)) + heating / (rho * T) + heat

    fields = [lnrho, u[0], u[1], u[2], A[0], A[1], A[2], s]
    rates = [cont, mom[0], mom[1], mom[2], ind[0], ind[1], ind[2], ent]
    f_fields = [sympy.lambdify(X, e, "numpy") for e in fields]
    f_rates = [sympy.lambdify(X, e, "numpy") for e in rates]
    return f_fields, f_rates


def test_rhs_matches_independent_formulation():
    n = 5
    m = n + 6
    idx = np.arange(m) - 3
    zc, yc, xc = np.meshgrid(idx * H[2] + 0.3, idx * H[1] - 0.2, idx * H[0] + 0.1, indexing="ij")
    f_fields, f_rates = symbolic_rhs()
    fields = []
    for ff in f_fields:
        v = np.broadcast_to(np.asarray(ff(xc, yc, zc), dtype=np.float64), xc.shape)
        fields.append(np.ascontiguousarray(v, dtype=np.longdouble))
    rates = pyoracle.rhs(fields, real_params())
    names = ["lnrho", "ux", "uy", "uz", "ax", "ay", "az", "entropy"]
    sl = (slice(3, -3),) * 3
    for q in range(8):
        exp = np.broadcast_to(np.asarray(f_rates[q](xc, yc, zc), dtype=np.float64), xc.shape)[sl]
        got = rates[q][sl].astype(np.float64)
        scale = np.max(np.abs(exp))
        assert scale > 0
        err = np.max(np.abs(got - exp)) / scale
        assert err < 1e-12, (names[q], err)


def test_rhs_detects_planted_mistake():
    """The comparison is sharp: dropping the Lorentz force (a plausible mistake) is detected."""
    n = 5
    m = n + 6
    idx = np.arange(m) - 3
    zc, yc, xc = np.meshgrid(idx * H[2] + 0.3, idx * H[1] - 0.2, idx * H[0] + 0.1, indexing="ij")
    f_fields, f_rates = symbolic_rhs()
    fields = [np.ascontiguousarray(np.broadcast_to(np.asarray(ff(xc, yc, zc), dtype=np.float64), xc.shape),
                                   dtype=np.longdouble) for ff in f_fields]
    rp = real_params()
    # mu0 -> infinity removes J entirely from the oracle (Lorentz force and Ohmic heating)
    from astaroth_amd.abi import RP

    rp[RP["AC_mu0"]] = 1e300
    rates = pyoracle.rhs(fields, rp)
    sl = (slice(3, -3),) * 3
    exp = np.asarray(f_rates[1](xc, yc, zc), dtype=np.float64)[sl]
    got = rates[1][sl].astype(np.float64)
    assert np.max(np.abs(got - exp)) / np.max(np.abs(exp)) > 1e-6
