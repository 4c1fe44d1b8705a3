"""Pins of the oracle's stencil and split-flux primitives
(src/globals.h:69-95, src/cuda_derivs.h:30-312) against closed forms and
independent derivations."""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O

# Known-answer vectors: the coefficient tables as written in src/globals.h:69-95
REF_TABLES = {
    1: ([-1.0 / 2.0], [1.0, -2.0]),
    2: ([1.0 / 12.0, -2.0 / 3.0], [-1.0 / 12.0, 4.0 / 3.0, -5.0 / 2.0]),
    3: ([-1.0 / 60.0, 3.0 / 20.0, -3.0 / 4.0], [1.0 / 90.0, -3.0 / 20.0, 3.0 / 2.0, -49.0 / 18.0]),
    4: ([1.0 / 280.0, -4.0 / 105.0, 1.0 / 5.0, -4.0 / 5.0],
        [-1.0 / 560.0, 8.0 / 315.0, -1.0 / 5.0, 8.0 / 5.0, -205.0 / 72.0]),
}


@pytest.mark.parametrize("s", [1, 2, 3, 4])
def test_coefficients_equal_reference_tables_bitwise(s):
    cF, cS = O.coeffs(s)
    rF, rS = REF_TABLES[s]
    assert cF.tolist() == rF
    assert cS.tolist() == rS


def _taylor_weights(s, deriv):
    """Central weights w_k (k=-s..s) exact for polynomials of degree 2s,
    by solving the Taylor (Vandermonde) system in exact rationals."""
    n = 2 * s + 1
    ks = list(range(-s, s + 1))
    A = [[Fraction(k) ** p for k in ks] for p in range(n)]
    b = [Fraction(0)] * n
    b[deriv] = Fraction(1 if deriv < 2 else 2)  # p! for p = deriv (1! or 2!)
    # Gaussian elimination
    M = [row[:] + [b[i]] for i, row in enumerate(A)]
    for c in range(n):
        piv = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[piv] = M[piv], M[c]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    return {k: M[i][n] / M[i][i] for i, k in enumerate(ks)}


@pytest.mark.parametrize("s", [1, 2, 3, 4])
def test_coefficients_match_taylor_system(s):
    cF, cS = O.coeffs(s)
    w1 = _taylor_weights(s, 1)
    w2 = _taylor_weights(s, 2)
    for it in range(s):
        # cF[it] multiplies f[i+it-s] - f[i+s-it]
        assert cF[it] == float(w1[it - s])
        assert float(w1[s - it]) == -cF[it]
        assert cS[it] == float(w2[it - s]) == float(w2[s - it])
    assert cS[s] == float(w2[0])
    # rows of the second-derivative operator sum to zero (constants map to 0)
    assert abs(2 * sum(cS[:s]) + cS[s]) < 4e-15
    assert sum(w2.values()) == 0


@pytest.mark.parametrize("s", [1, 2, 3, 4])
def test_modified_wavenumbers(s):
    """D1 sin(kx) = (2/h) sum_l a_l sin(l k h) cos(kx); D2 cos(kx) = ... exactly."""
    n = 32
    L = 2 * np.pi
    h = L / n
    x = (np.arange(n) + 0.5) * h
    cF, cS = O.coeffs(s)
    a = {s - it: -cF[it] for it in range(s)}  # weight on f[i+l]
    b = {s - it: cS[it] for it in range(s)}
    for k in (1, 3, 7):
        d1 = O.pencil_d1(np.sin(k * x), s, 1.0 / h)
        lam1 = sum(2 * a[l] * np.sin(l * k * h) for l in a) / h
        np.testing.assert_allclose(d1, lam1 * np.cos(k * x), rtol=0, atol=1e-12 * k)
        d2 = O.pencil_d2(np.cos(k * x), s, 1.0 / h ** 2)
        lam2 = (cS[s] + sum(2 * b[l] * np.cos(l * k * h) for l in b)) / h ** 2
        np.testing.assert_allclose(d2, lam2 * np.cos(k * x), rtol=0, atol=1e-10 * k * k)


@pytest.mark.parametrize("s", [1, 2, 3, 4])
def test_convergence_order(s):
    def err(n):
        L = 2 * np.pi
        h = L / n
        x = (np.arange(n) + 0.5) * h
        f = np.exp(np.sin(x))
        d1 = O.pencil_d1(f, s, 1.0 / h)
        d2 = O.pencil_d2(f, s, 1.0 / h ** 2)
        e1 = np.max(np.abs(d1 - np.cos(x) * f))
        e2 = np.max(np.abs(d2 - (np.cos(x) ** 2 - np.sin(x)) * f))
        return e1, e2
    n0 = {1: 32, 2: 32, 3: 48, 4: 48}[s]
    a, b = err(n0), err(2 * n0)
    p1, p2 = np.log2(a[0] / b[0]), np.log2(a[1] / b[1])
    assert p1 > 2 * s - 0.5 and p2 > 2 * s - 0.5, (p1, p2)


def _D(f, s, d):
    return O.pencil_d1(f, s, d)


@pytest.mark.parametrize("s", [1, 2, 3, 4])
def test_split_flux_equals_kennedy_gruber_expansion(s):
    """fluxQuad = -1/2[D(fg) + f Dg + g Df];
    fluxCube = -1/4[D(fgh) + f D(gh) + g D(fh) + h D(fg) + gh Df + fh Dg + fg Dh]
    with D the central first-derivative operator of the same order."""
    rng = np.random.default_rng(s)
    n = 40
    d = 1.7
    f, g, h = (rng.uniform(0.5, 2.0, n) for _ in range(3))
    quad = O.pencil_flux(f, g, None, s, d)
    kq = -0.5 * (_D(f * g, s, d) + f * _D(g, s, d) + g * _D(f, s, d))
    np.testing.assert_allclose(quad, kq, rtol=0, atol=1e-13 * np.max(np.abs(kq)))
    cube = O.pencil_flux(f, g, h, s, d)
    kc = -0.25 * (_D(f * g * h, s, d) + f * _D(g * h, s, d) + g * _D(f * h, s, d) + h * _D(f * g, s, d)
                  + g * h * _D(f, s, d) + f * h * _D(g, s, d) + f * g * _D(h, s, d))
    np.testing.assert_allclose(cube, kc, rtol=0, atol=1e-13 * np.max(np.abs(kc)))


@pytest.mark.parametrize("s", [1, 2, 3, 4])
def test_split_flux_special_cases(s):
    rng = np.random.default_rng(10 + s)
    n = 36
    d = 0.9
    f = rng.uniform(-1, 1, n)
    one = np.ones(n)
    # g = h = 1 reduces the cubic flux to -D f; quadratic with g = 1 too
    np.testing.assert_allclose(O.pencil_flux(f, one, one, s, d), -_D(f, s, d), atol=1e-13)
    np.testing.assert_allclose(O.pencil_flux(f, one, None, s, d), -_D(f, s, d), atol=1e-13)
    # constant fields: exactly zero; telescoping: sum over the period is zero
    c = 1.3 * one
    assert np.max(np.abs(O.pencil_flux(c, c, c, s, d))) == 0.0
    g, h = rng.uniform(0.5, 1.5, n), rng.uniform(0.5, 1.5, n)
    assert abs(np.sum(O.pencil_flux(f, g, h, s, d))) < 1e-13 * n
    assert abs(np.sum(O.pencil_flux(f, g, None, s, d))) < 1e-13 * n


def test_splitmix64_known_answers():
    from tests.cases import splitmix64_uniform
    # published SplitMix64 outputs: seed 0 -> e220a8397b1dcdaf, 6e789e6aa1b965f4
    u = splitmix64_uniform(0, 2)
    assert u[0] == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53
    assert u[1] == (0x6E789E6AA1B965F4 >> 11) * 2.0 ** -53
    # seed 1234567 -> 6457827717110365317, 3203168211198807973
    v = splitmix64_uniform(1234567, 2)
    assert v[0] == (6457827717110365317 >> 11) * 2.0 ** -53
    assert v[1] == (3203168211198807973 >> 11) * 2.0 ** -53
