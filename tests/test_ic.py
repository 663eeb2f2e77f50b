"""Initial-state layer (hype_initial_state, host C; initialize_solution.c:21-81)
and the Rankine-Hugoniot post-shock state of the shock-bubble configuration."""
import numpy as np
import pytest

import hype2d as H


def test_gauss_average_of_uniform_state_is_exact():
    p = H.make_problem(H.EULER, 8, 6, domain=(0, 1, 0, 0.75), gamma1=1.4)
    U = H.initial_state(p, H.IC_UNIFORM, [1.3, 0.2, -0.1, 2.0])
    E = 2.0 / 0.4 + 0.5 * 1.3 * (0.04 + 0.01)
    assert np.allclose(U, [1.3, 1.3 * 0.2, -1.3 * 0.1, E], rtol=1e-14, atol=0)  # 25 weights sum to 1 +- ulps


def test_quadrants_on_aligned_grid():
    ne, nw, sw, se = [1.5, 0, 0, 1.5], [0.5323, 1.206, 0, 0.3], [0.138, 1.206, 1.206, 0.029], [0.5323, 0, 1.206, 0.3]
    p = H.make_problem(H.EULER, 16, 16, domain=(0, 1, 0, 1), gamma1=1.4)
    U = H.initial_state(p, H.IC_QUADRANTS, [0.5, 0.5] + ne + nw + sw + se)

    def cons(V):
        r, u, v, pr = V
        return np.array([r, r * u, r * v, pr / 0.4 + 0.5 * r * (u * u + v * v)])

    assert np.allclose(U[12, 12], cons(ne), rtol=1e-15)
    assert np.allclose(U[12, 3], cons(nw), rtol=1e-15)
    assert np.allclose(U[3, 3], cons(sw), rtol=1e-15)
    assert np.allclose(U[3, 12], cons(se), rtol=1e-15)


def test_dmda_coordinates_stretch_nonperiodic_axes():
    """DMDASetUniformCoordinates(x_min+h/2, x_max+h/2) spaces M centres with
    M-1 gaps on a non-periodic axis (hype.c:132-135)."""
    ne = nw = sw = se = [1.0, 0, 0, 1.0]
    right = [2.0, 0, 0, 1.0]
    args = [0.95, -1.0] + right + nw + sw + right  # x >= 0.95 -> density 2
    p = H.make_problem(H.EULER, 10, 10, domain=(0, 1, 0, 1), gamma1=1.4)
    exact = H.initial_state(p, H.IC_QUADRANTS, args, coord=H.COORD_EXACT)
    dmda = H.initial_state(p, H.IC_QUADRANTS, args, coord=H.COORD_DMDA)
    # last cell straddles x = 0.95 at its centre: the 5-point Gauss weights at
    # non-negative offsets (0.28444 + 0.23931 + 0.11846) see density 2
    assert abs(exact[5, 9, 0] - (1.0 + 0.28444444444444444 + 0.23931433524968326 + 0.11846344252809456)) < 1e-12
    assert abs(dmda[5, 9, 0] - 2.0) < 1e-14  # stretched centre 1.05: cell wholly right of it
    p.bc[0] = p.bc[1] = H.BC_PERIODIC
    per = H.initial_state(p, H.IC_QUADRANTS, args, coord=H.COORD_DMDA)
    assert np.array_equal(per, exact)


@pytest.mark.parametrize("mach", [1.1, 1.22, 2.0])
def test_rankine_hugoniot_jump_conditions(mach):
    g, r0, p0 = 1.4, 1.4, 1.0e5
    r1, u1, p1 = H.rankine_hugoniot(g, r0, p0, mach)
    W = mach * np.sqrt(g * p0 / r0)
    # shock frame: w0 = W, w1 = W - u1
    w0, w1 = W, W - u1
    assert abs(r0 * w0 - r1 * w1) < 1e-12 * r0 * w0
    assert abs((r0 * w0 ** 2 + p0) - (r1 * w1 ** 2 + p1)) < 1e-12 * p1
    h0 = g / (g - 1) * p0 / r0 + 0.5 * w0 ** 2
    h1 = g / (g - 1) * p1 / r1 + 0.5 * w1 ** 2
    assert abs(h0 - h1) < 1e-12 * h0
    assert r1 > r0 and p1 > p0 and u1 > 0


def test_shock_bubble_state():
    nx, ny = 128, 64
    h = 2.0 / nx
    p = H.make_problem(H.KAPILA, nx, ny, domain=(0, 2, 0, 1), gamma1=1.4, gamma2=1.648,
                       bc=(H.BC_TRANSMISSIVE, H.BC_TRANSMISSIVE, H.BC_REFLECTIVE, H.BC_REFLECTIVE))
    U = H.initial_state(p, H.IC_SHOCK_BUBBLE, [1.22, 0.5, 1.0, 0.5, 0.25, 1e-6, 2 * h, 1.4, 0.25463, 1e5])
    a1 = U[..., 5]
    assert a1.min() >= 1e-6 - 1e-18 and a1.max() <= 1 - 1e-6 + 1e-15
    assert a1[32, 64] < 1e-5  # bubble centre is helium
    r1, u1, p1 = H.rankine_hugoniot(1.4, 1.4, 1e5, 1.22)
    rho = U[32, 5, 0] + U[32, 5, 1]
    assert abs(U[32, 5, 2] / rho - u1) < 1e-9 * u1  # post-shock gas moves at u1


def test_water_air_is_seeded_and_bounded():
    nx, ny = 64, 32
    h = 2.0 / nx
    p = H.make_problem(H.BAER_NUNZIATO, nx, ny, prim=False, domain=(0, 2, 0, 1), gamma1=4.4, gamma2=1.4, pi1=6e8)
    args = [0.5, 1e-6, 2 * h, 1.0, 0.3, 0.1, 1e9, 1e5, 1000.0, 1.0, 8, 0.005, 2.0]
    a = H.initial_state(p, H.IC_WATER_AIR, args, seed=7)
    b = H.initial_state(p, H.IC_WATER_AIR, args, seed=7)
    c = H.initial_state(p, H.IC_WATER_AIR, args, seed=8)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    phi = a[..., 8]
    assert phi.min() >= 1e-6 - 1e-18 and phi.max() <= 1 - 1e-6 + 1e-15
    assert phi[2, 10] > 0.99 and phi[-2, 10] < 0.01  # water below, air above


def test_smooth_vortex_profile():
    n = 64
    p = H.make_problem(H.KAPILA, n, n, domain=(-3, 3, -3, 3), gamma1=4.0, gamma2=1.4, pi1=20.0,
                       bc=(H.BC_PERIODIC,) * 4)
    U = H.initial_state(p, H.IC_SMOOTH_VORTEX)
    from math import erfc
    x = -3 + (np.arange(n) + 0.5) * 6 / n
    X, Y = np.meshgrid(x, x)
    r = np.hypot(X, Y)
    a = 0.25 + 0.25 * np.vectorize(erfc)((r - 1.0) / 0.3)
    assert np.max(np.abs(U[..., 5] - a)) < 0.01  # cell average vs centre value, O(h^2)


def test_problem_options_validated():
    p = H.make_problem(H.EULER, 8, 6, domain=(0, 1, 0, 0.75), gamma1=1.4, fma=True)
    assert p.options == H.OPT_FMA
    H.initial_state(p, H.IC_UNIFORM, [1.0, 0.0, 0.0, 1.0])  # validated, accepted
    for ok in (H.OPT_ONE_WAVE_TILES, H.OPT_SPLIT_TILES | H.OPT_FMA):
        p.options = ok
        H.initial_state(p, H.IC_UNIFORM, [1.0, 0.0, 0.0, 1.0])
    for bad in (0x8, H.OPT_ONE_WAVE_TILES | H.OPT_SPLIT_TILES):  # unknown bit; two layouts at once
        p.options = bad
        with pytest.raises(H.HypeError):
            H.initial_state(p, H.IC_UNIFORM, [1.0, 0.0, 0.0, 1.0])
