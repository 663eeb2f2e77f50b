"""Pins of the oracle's reconstruction (reconstruction.c restated in
oracle/orc_recon.c).

Expected values are derived independently of the code under test:
  * the basis is checked for zero mean and orthogonality on the unit cell;
  * candidate polynomials are obtained by solving the cell-average equations
    of each stencil (numpy linear algebra);
  * smoothness indicators are the Jiang-Shu integrals of squared derivatives,
    evaluated by Gauss quadrature of the fitted polynomials;
  * weights follow the WENO definition w_k ~ gamma_k / (beta_k + eps)^4 with
    the reference's gamma = (1, 100, 1) / (1,1,1,1), eps = 1e-12.
Exactness for quadratics (weno3) and cubics in the linear-weight regime
(weno4) pins the stencils whatever the weights.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_py as orc

GX, GW = np.polynomial.legendre.leggauss(6)
GX, GW = GX / 2.0, GW / 2.0  # on [-1/2, 1/2]
EPS = 1e-12


def basis(x, y, k):
    return orc.lib().orc_basis(x, y, k)


def test_basis_zero_mean_and_orthogonal():
    n = 10
    G = np.zeros((n, n))
    for k in range(n):
        for l in range(n):
            G[k, l] = sum(wx * wy * basis(x, y, k) * basis(x, y, l) for x, wx in zip(GX, GW) for y, wy in zip(GX, GW))
    for k in range(1, n):
        mean = sum(wx * wy * basis(x, y, k) for x, wx in zip(GX, GW) for y, wy in zip(GX, GW))
        assert abs(mean) < 1e-15
    off = G - np.diag(np.diag(G))
    assert np.max(np.abs(off)) < 1e-15
    assert np.all(np.diag(G) > 0)


def recon(N, ux, uy, uxy):
    co = np.zeros(10)
    orc.lib().orc_reconstruct(orc.dptr(np.array(ux, float)), orc.dptr(np.array(uy, float)),
                              orc.dptr(np.array(uxy, float)), N, orc.dptr(co))
    return co


def stencil_from_field(f):
    """f(i, j) -> cell average; returns the 5+5+5 stencil arrays at (0,0)."""
    ux = [f(i, 0) for i in (-2, -1, 0, 1, 2)]
    uy = [f(0, j) for j in (-2, -1, 0, 1, 2)]
    uxy = [f(0, 0), f(1, 1), f(1, -1), f(-1, 1), f(-1, -1)]
    return ux, uy, uxy


def quad_avg(i, j, a):
    """Cell average over [i-1/2,i+1/2]x[j-1/2,j+1/2] of sum a_k basis_k."""
    return sum(wx * wy * sum(a[k] * basis(x + i, y + j, k) for k in range(len(a)))
               for x, wx in zip(GX, GW) for y, wy in zip(GX, GW))


@pytest.mark.parametrize("seed", range(5))
def test_weno3_exact_for_quadratics(seed):
    a = np.random.default_rng(seed).normal(size=6) * np.array([1, 3, -2, 5, 4, 7])
    ux, uy, uxy = stencil_from_field(lambda i, j: quad_avg(i, j, a))
    co = recon(2, ux, uy, uxy)
    assert np.allclose(co[:6], a, rtol=0, atol=1e-11 * np.max(np.abs(a)))


def fit_quadratic(cells, vals):
    """a + b k + c k^2 = cell average of a + b x + c (x^2 - 1/12) over cell k."""
    A = np.array([[1.0, k, k * k] for k in cells])
    return np.linalg.solve(A, np.array(vals, float))


def beta_1d(b, c):
    """Jiang-Shu indicator: integral over the cell of p'^2 + p''^2 for
    p = b x + c (x^2 - 1/12), by Gauss quadrature."""
    return sum(w * ((b + 2 * c * x) ** 2 + (2 * c) ** 2) for x, w in zip(GX, GW))


def weno_dir(u):
    """Expected (slope, curvature) of the directional WENO from u(-2..2)."""
    stencils = [(-2, -1, 0), (-1, 0, 1), (0, 1, 2)]
    gam = np.array([1.0, 100.0, 1.0])
    bc = [fit_quadratic(s, [u[k + 2] for k in s])[1:] for s in stencils]
    beta = np.array([beta_1d(b, c) for b, c in bc])
    w = gam / (beta + EPS) ** 4
    w /= w.sum()
    return sum(w[k] * bc[k][0] for k in range(3)), sum(w[k] * bc[k][1] for k in range(3))


def weno_cross(u0, c1, c2, c3, c4, corners):
    """Expected xy coefficient: corner candidates from the cell average of
    the corner cell, indicators = integral of |Hessian|^2."""
    fs, betas = [], []
    for (sx, sy), uc in corners.items():
        f = (uc - u0 - c1 * sx - c2 * sy - c3 - c4) / (sx * sy)
        fs.append(f)
        betas.append((2 * c3) ** 2 + (2 * c4) ** 2 + f ** 2)
    w = 1.0 / (np.array(betas) + EPS) ** 4
    w /= w.sum()
    return float(np.dot(w, fs))


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("scale", [1.0, 1e-3, 1e3])
def test_weno3_nonlinear_weights(seed, scale):
    rng = np.random.default_rng(100 + seed)
    vals = {}

    def f(i, j):
        if (i, j) not in vals:
            vals[(i, j)] = scale * (rng.normal() + (3.0 if i + 0.3 * j > 0.5 else 0.0))
        return vals[(i, j)]

    ux, uy, uxy = stencil_from_field(f)
    co = recon(2, ux, uy, uxy)
    c1, c3 = weno_dir(ux)
    c2, c4 = weno_dir(uy)
    corners = {(1, 1): uxy[1], (1, -1): uxy[2], (-1, 1): uxy[3], (-1, -1): uxy[4]}
    c5 = weno_cross(uxy[0], co[1], co[2], co[3], co[4], corners)
    expect = np.array([uxy[0], c1, c2, c3, c4, c5])
    tol = 1e-11 * max(np.max(np.abs(expect)), scale)
    assert np.max(np.abs(co[:6] - expect)) <= tol


def test_weno3_linear_weights_regime():
    """Amplitudes far below sqrt(eps): weights reduce to (1,100,1)/102."""
    u = 1e-12 * np.array([(-2.0) ** 3, -1.0, 0.0, 1.0, 8.0])  # cubic cell values
    co = recon(2, u, np.zeros(5), [0, 0, 0, 0, 0])
    bc = [fit_quadratic(s, [u[k + 2] for k in s])[1:] for s in [(-2, -1, 0), (-1, 0, 1), (0, 1, 2)]]
    lin = (bc[0][0] + 100 * bc[1][0] + bc[2][0]) / 102.0
    assert abs(co[1] - lin) <= 1e-6 * abs(lin)
    assert co[1] != bc[1][0]  # the one-sided candidates contribute


def test_weno3_eno_property_at_a_step():
    """Left stencil smooth, others cross a jump: the result is the smooth
    candidate (weights ~ (beta+eps)^-4)."""
    u = [1.0, 1.0, 1.0, 2.0, 2.0]
    co = recon(2, u, [1.0] * 5, [1.0, 2.0, 2.0, 1.0, 1.0])
    assert abs(co[1]) < 1e-40 and abs(co[3]) < 1e-40


def test_minmod_and_weno2():
    L = orc.lib()
    assert L.orc_minmod(1.0, -2.0) == 0.0
    assert L.orc_minmod(-3.0, -2.0) == -2.0
    assert L.orc_minmod(0.5, 2.0) == 0.5
    assert L.orc_minmod(0.0, 2.0) == 0.0
    co = recon(1, [9, 1.0, 2.0, 4.0, 9], [9, 3.0, 2.0, 1.5, 9], [2.0, 0, 0, 0, 0])
    assert list(co[:3]) == [2.0, 1.0, -0.5]
    co = recon(0, [0] * 5, [0] * 5, [7.5, 0, 0, 0, 0])
    assert co[0] == 7.5


def cubic_avg(i, j, a):
    return quad_avg(i, j, a)


@pytest.mark.parametrize("seed", range(4))
def test_weno4_exact_for_quadratics(seed):
    a = np.zeros(10)
    a[:6] = np.random.default_rng(seed).normal(size=6)
    ux, uy, uxy = stencil_from_field(lambda i, j: cubic_avg(i, j, a))
    co = recon(3, ux, uy, uxy)
    assert np.allclose(co, a, rtol=0, atol=1e-10 * np.max(np.abs(a)))


@pytest.mark.parametrize("seed", range(4))
def test_weno4_fourth_order_polynomial_in_linear_regime(seed):
    """Tiny cubic data: CWENO weights are the linear ones, so the optimal
    fourth-order polynomial is returned exactly (10 coefficients)."""
    a = 1e-11 * np.random.default_rng(50 + seed).normal(size=10)
    ux, uy, uxy = stencil_from_field(lambda i, j: cubic_avg(i, j, a))
    co = recon(3, ux, uy, uxy)
    assert np.allclose(co, a, rtol=0, atol=1e-6 * np.max(np.abs(a)))


# ---- CWENO4 (reconstruction.c:223-357) in the nonlinear-weight regime -------
#
# Every candidate polynomial is obtained here from the cell averages it must
# reproduce (a linear solve on quadrature cell averages of the basis), and
# every smoothness indicator is the Jiang-Shu sum over all derivatives
# D^a, 1 <= |a| <= degree, of the cell integral of (D^a p)^2, evaluated by
# Gauss quadrature of the fitted polynomial.  No constant of the closed-form
# indicators (13/3, 7/6, 0.1, 39.05, 4.7) or of the candidate formulas
# (41/60, 11/120, ...) enters the expected values.

CELLS = [(0, 0), (-2, 0), (-1, 0), (1, 0), (2, 0), (0, -2), (0, -1), (0, 1), (0, 2),
         (1, 1), (1, -1), (-1, 1), (-1, -1)]


def cell_avg_matrix(nb):
    """A[c, k]: average of basis k over cell CELLS[c] (exact: degree <= 3)."""
    return np.array([[quad_avg(i, j, np.eye(nb)[k]) for k in range(nb)] for (i, j) in CELLS])


def functional(terms):
    """Linear functional on the vector of the 13 stencil averages."""
    v = np.zeros(len(CELLS))
    for cell, wgt in terms:
        v[CELLS.index(cell)] += wgt
    return v


def avg(cell):
    return functional([(cell, 1.0)])


# The conditions that define each candidate: cell averages it reproduces and,
# for the corner information, the combinations of corner averages it matches.
R4_CONDITIONS = [
    avg((0, 0)), avg((1, 0)), avg((-1, 0)), avg((0, 1)), avg((0, -1)),
    functional([((2, 0), 1.0), ((-2, 0), -1.0)]),
    functional([((0, 2), 1.0), ((0, -2), -1.0)]),
    functional([((1, 1), 1.0), ((1, -1), -1.0), ((-1, 1), -1.0), ((-1, -1), 1.0)]),
    functional([((1, 1), 1.0), ((-1, 1), 1.0), ((1, -1), -1.0), ((-1, -1), -1.0)]),
    functional([((1, 1), 1.0), ((1, -1), 1.0), ((-1, 1), -1.0), ((-1, -1), -1.0)]),
]
R3_CONDITIONS = [
    # central
    [avg((0, 0)), avg((1, 0)), avg((-1, 0)), avg((0, 1)), avg((0, -1)),
     functional([((1, 1), 1.0), ((1, -1), -1.0), ((-1, 1), -1.0), ((-1, -1), 1.0)])],
    # left-biased in x
    [avg((0, 0)), avg((-1, 0)), avg((-2, 0)), avg((0, 1)), avg((0, -1)),
     functional([((-1, 1), 1.0), ((-1, -1), -1.0)])],
    # right-biased in x
    [avg((0, 0)), avg((1, 0)), avg((2, 0)), avg((0, 1)), avg((0, -1)),
     functional([((1, 1), 1.0), ((1, -1), -1.0)])],
    # biased downwards in y
    [avg((0, 0)), avg((0, -1)), avg((0, -2)), avg((1, 0)), avg((-1, 0)),
     functional([((1, -1), 1.0), ((-1, -1), -1.0)])],
    # biased upwards in y
    [avg((0, 0)), avg((0, 1)), avg((0, 2)), avg((1, 0)), avg((-1, 0)),
     functional([((1, 1), 1.0), ((-1, 1), -1.0)])],
]


def fit(conditions, nb, data):
    A = cell_avg_matrix(nb)
    M = np.array([c @ A for c in conditions])
    return np.linalg.solve(M, np.array([c @ data for c in conditions]))


_MONO = None


def monomial_coeffs(a):
    """Coefficients c[m, n] of x^m y^n (m + n <= 3) of sum_k a_k basis_k,
    by exact interpolation of the basis at a 4 x 4 point set."""
    global _MONO
    if _MONO is None:
        pts = [(x, y) for x in (-0.4, -0.1, 0.2, 0.45) for y in (-0.35, -0.05, 0.15, 0.4)]
        mons = [(m, n) for m in range(4) for n in range(4) if m + n <= 3]
        V = np.array([[x ** m * y ** n for (m, n) in mons] for (x, y) in pts])
        B = np.array([[basis(x, y, k) for k in range(10)] for (x, y) in pts])
        sol = np.linalg.lstsq(V, B, rcond=None)[0]  # (10 monomials) x (10 basis)
        _MONO = (mons, sol)
    mons, sol = _MONO
    c = np.zeros((4, 4))
    for r, (m, n) in enumerate(mons):
        c[m, n] = sol[r, : len(a)] @ a
    return c


def jiang_shu(a, degree):
    """sum over 1 <= |alpha| <= degree of the cell integral of (D^alpha p)^2."""
    P = np.polynomial.polynomial
    c = monomial_coeffs(a)
    total = 0.0
    for ax in range(degree + 1):
        for ay in range(degree + 1 - ax):
            if ax + ay == 0:
                continue
            d = P.polyder(P.polyder(c, ax, axis=0), ay, axis=1) if ay else P.polyder(c, ax, axis=0)
            total += sum(wx * wy * P.polyval2d(x, y, d) ** 2 for x, wx in zip(GX, GW) for y, wy in zip(GX, GW))
    return total


def cweno4_expected(ux, uy, uxy):
    data = np.zeros(len(CELLS))
    vals = {(i, 0): ux[i + 2] for i in range(-2, 3)}
    vals.update({(0, j): uy[j + 2] for j in range(-2, 3)})
    vals.update({(1, 1): uxy[1], (1, -1): uxy[2], (-1, 1): uxy[3], (-1, -1): uxy[4]})
    for c, cell in enumerate(CELLS):
        data[c] = vals[cell]
    a4 = fit(R4_CONDITIONS, 10, data)
    a3 = [np.concatenate([fit(cond, 6, data), np.zeros(4)]) for cond in R3_CONDITIONS]
    g_hi = g_lo = 0.85
    gam4 = g_hi
    gam3 = np.array([(1 - g_hi) * g_lo] + [0.25 * (1 - g_hi) * (1 - g_lo)] * 4)
    w4 = gam4 / (jiang_shu(a4, 3) + EPS) ** 4
    w3 = np.array([gam3[k] / (jiang_shu(a3[k], 2) + EPS) ** 4 for k in range(5)])
    s = w4 + w3.sum()
    w4, w3 = w4 / s, w3 / s
    opt = a4 - sum(gam3[k] * a3[k] for k in range(5))
    return (w4 / gam4) * opt + sum(w3[k] * a3[k] for k in range(5)), a4, a3


def test_cweno4_candidates_reproduce_cubics():
    """Sanity of the fitting: on cubic cell averages the fourth-order
    candidate is the cubic, each third-order one its quadratic part's fit."""
    a = np.random.default_rng(7).normal(size=10)
    ux, uy, uxy = stencil_from_field(lambda i, j: cubic_avg(i, j, a))
    _, a4, _ = cweno4_expected(ux, uy, uxy)
    assert np.allclose(a4, a, rtol=0, atol=1e-12 * np.max(np.abs(a)))
    q = np.zeros(10)
    q[:6] = a[:6]
    ux, uy, uxy = stencil_from_field(lambda i, j: cubic_avg(i, j, q))
    _, _, a3 = cweno4_expected(ux, uy, uxy)
    for cand in a3:
        assert np.allclose(cand, q, rtol=0, atol=1e-12 * np.max(np.abs(q)))


def test_jiang_shu_indicator_closed_forms():
    """The quadrature indicator against hand integrals on single modes:
    x -> 1, x^2 - 1/12 -> 13/3, xy -> 7/6, x(x^2 - 3/20) -> 39.06 (incl. its
    cross term with x), y(x^2 - 1/12) -> 4.7."""
    for k, want in [(1, 1.0), (3, 13.0 / 3.0), (5, 7.0 / 6.0), (6, 39.06), (8, 4.7), (9, 4.7), (7, 39.06)]:
        a = np.zeros(10)
        a[k] = 1.0
        assert abs(jiang_shu(a, 3) - want) < 1e-13 * want


@pytest.mark.parametrize("seed", range(10))
@pytest.mark.parametrize("scale", [1.0, 1e-2, 1e2])
def test_cweno4_nonlinear_weights(seed, scale):
    """Data with a jump through the stencil (and smooth data of moderate
    amplitude for odd seeds): weights far from linear, so every term of every
    indicator moves the result."""
    rng = np.random.default_rng(300 + seed)
    vals = {}

    def f(i, j):
        if (i, j) not in vals:
            if seed % 2:
                vals[(i, j)] = scale * (np.sin(0.9 * i + 0.4 * j + seed) + 0.3 * np.cos(1.3 * j - 0.2 * i))
            else:
                vals[(i, j)] = scale * (rng.normal() + (2.5 if 0.7 * i - j > 0.4 else 0.0))
        return vals[(i, j)]

    ux, uy, uxy = stencil_from_field(f)
    co = recon(3, ux, uy, uxy)
    expect, _, _ = cweno4_expected(ux, uy, uxy)
    tol = 1e-11 * max(np.max(np.abs(expect)), scale)
    assert np.max(np.abs(co - expect)) <= tol, (co, expect)
