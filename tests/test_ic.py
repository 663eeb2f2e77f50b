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


# ---- the reference's Kapila test cases (kapila_5eqn.c:565-996) -------------
#
# Pins: a cell wholly inside a region averages to that region's conserved
# state, computed here from the stiffened-gas mixture energy
# rho e = sum_k alpha_k (p + gamma_k pi_k) / (gamma_k - 1); the areas of the
# circular regions and the position of the Richtmyer-Meshkov interface from
# the volume fraction field; the smooth bubble-collapse profile against its
# centre values; the case settings against the reference's header comments.

EPS5 = 1.0e-5


def kapila_cons(p, V):
    r1, r2, u, v, pr, a1 = V
    a2 = 1.0 - a1
    rho = a1 * r1 + a2 * r2
    rhoe = a1 * (pr + p.gamma1 * p.pi1) / (p.gamma1 - 1.0) + a2 * (pr + p.gamma2 * p.pi2) / (p.gamma2 - 1.0)
    return np.array([a1 * r1, a2 * r2, rho * u, rho * v, rhoe + 0.5 * rho * (u * u + v * v), a1])


def centres(p):
    h = (p.x_max - p.x_min) / p.nx
    x = p.x_min + (np.arange(p.nx) + 0.5) * h
    y = p.y_min + (np.arange(p.ny) + 0.5) * h
    return x, y, h


def case(kind, nx):
    p, t_end = H.case_problem(kind, nx, 3)
    ny = int(round(nx * (p.y_max - p.y_min) / (p.x_max - p.x_min)))
    p, t_end = H.case_problem(kind, nx, ny)
    return p, t_end, H.initial_state(p, kind)


def inner(mask, h, x, y, inside):
    """cells whose whole square satisfies `inside` (tested at its corners and centre)"""
    out = np.ones(mask.shape, bool)
    for dx in (-0.5, 0.0, 0.5):
        for dy in (-0.5, 0.0, 0.5):
            X, Y = np.meshgrid(x + dx * h, y + dy * h)
            out &= inside(X, Y)
    return out


def fraction_area(U, h, a_in, a_out):
    return float(np.sum((U[..., 5] - a_out) / (a_in - a_out))) * h * h


T_, R_, P_, I_ = H.BC_TRANSMISSIVE, H.BC_REFLECTIVE, H.BC_PERIODIC, H.BC_INLET


@pytest.mark.parametrize("kind,dom,bc,eos,t_end", [
    (H.IC_SMOOTH_VORTEX, (-3, 3, -3, 3), (P_, P_, P_, P_), (4.0, 1.4, 20.0, 0.0), 6.0),
    (H.IC_INTERFACE_ADVECTION, (0, 1, -0.05, 0.05), (T_, T_, T_, T_), (4.4, 1.4, 6.0e8, 0.0), 0.26 / 150.0),
    (H.IC_AIR_HELIUM, (0, 0.356, 0, 0.089), (T_, T_, R_, R_), (1.4, 1.648, 0.0, 0.0), 674.0e-6),
    (H.IC_AIR_R22, (0, 0.356, 0, 0.089), (T_, T_, R_, R_), (1.4, 1.249, 0.0, 0.0), 417.0e-6),
    (H.IC_RICHTMYER_MESHKOV, (0, 4, -0.5, 0.5), (T_, T_, T_, T_), (1.4, 1.093, 0.0, 0.0), 7.82),
    (H.IC_UNDERWATER_EXPLOSION, (-2, 2, -1.5, 1.5), (T_, T_, R_, T_), (4.4, 1.4, 6000.0, 0.0), 0.05),
    (H.IC_WATER_CYLINDER, (0, 0.12, -0.04, 0.04), (T_, T_, T_, T_), (6.12, 1.4, 3.43e8, 0.0), 70.0e-6),
    (H.IC_AIR_CAVITY, (0, 12, 0, 12), (T_, T_, R_, R_), (4.4, 1.4, 6000.0, 0.0), 0.045),
    (H.IC_BUBBLE_COLLAPSE, (0, 10, -2.5, 2.5), (R_, T_, T_, T_), (4.4, 1.4, 6000.0, 0.0), 0.02),
    (H.IC_AIR_JET, (0, 30, -15, 15), (I_, T_, R_, R_), (4.4, 1.4, 6000.0, 0.0), 1000.0),
    (H.IC_WATER_JET, (0, 1, -0.25, 0.25), (I_, T_, T_, T_), (6.12, 1.4, 3.43e8, 0.0), 6.0e-6),
])
def test_case_settings(kind, dom, bc, eos, t_end):
    p, te = H.case_problem(kind, 64, 64)
    assert p.system == H.KAPILA
    assert (p.x_min, p.x_max, p.y_min, p.y_max) == dom
    assert tuple(p.bc) == bc
    assert (p.gamma1, p.gamma2, p.pi1, p.pi2) == eos
    assert te == t_end
    assert p.nx == 64 and p.order == 2 and p.recons_prim == 1


def test_case_setup_rejects_other_kinds():
    p = H.make_problem(H.KAPILA, 8, 8)
    for bad in (H.IC_UNIFORM, H.IC_QUADRANTS, H.IC_SHOCK_BUBBLE, H.IC_WATER_AIR, 15, -1):
        with pytest.raises(H.HypeError):
            H.case_problem(bad, 8, 8)
    with pytest.raises(H.HypeError):  # the test cases are Kapila states
        H.initial_state(H.make_problem(H.EULER, 8, 8), H.IC_AIR_CAVITY)


def test_interface_advection():
    p, _, U = case(H.IC_INTERFACE_ADVECTION, 40)  # cell faces at x = 0.5
    left = kapila_cons(p, (1000.0, 1.2, 150.0, 0.0, 101325.0, 0.01))
    right = kapila_cons(p, (1000.0, 1.2, 150.0, 0.0, 101325.0, 0.99))
    assert np.allclose(U[:, :20], left, rtol=1e-14, atol=0)
    assert np.allclose(U[:, 20:], right, rtol=1e-14, atol=0)


@pytest.mark.parametrize("kind,r2", [(H.IC_AIR_HELIUM, 0.25463), (H.IC_AIR_R22, 4.41540)])
def test_shock_bubble_cases(kind, r2):
    p, _, U = case(kind, 712)  # h = 5e-4
    x, y, h = centres(p)
    post = kapila_cons(p, (1.92691, r2, -114.42, 0.0, 1.56980e5, 1.0 - EPS5))
    pre = kapila_cons(p, (1.4, r2, 0.0, 0.0, 1.0e5, 1.0 - EPS5))
    bub = kapila_cons(p, (1.4, r2, 0.0, 0.0, 1.0e5, EPS5))
    in_b = lambda X, Y: (X - 0.245) ** 2 + (Y - 0.0455) ** 2 <= 0.025 ** 2
    out_b = lambda X, Y: ~in_b(X, Y)
    m_post = inner(U[..., 0], h, x, y, lambda X, Y: (X >= 0.275) & out_b(X, Y))
    m_pre = inner(U[..., 0], h, x, y, lambda X, Y: (X < 0.275) & out_b(X, Y))
    m_bub = inner(U[..., 0], h, x, y, in_b)
    assert m_post.sum() > 1000 and m_pre.sum() > 1000 and m_bub.sum() > 1000
    assert np.allclose(U[m_post], post, rtol=1e-14, atol=0)
    assert np.allclose(U[m_pre], pre, rtol=1e-14, atol=0)
    assert np.allclose(U[m_bub], bub, rtol=1e-14, atol=0)
    area = fraction_area(U, h, EPS5, 1.0 - EPS5)
    assert abs(area - np.pi * 0.025 ** 2) < 0.1 * 2 * np.pi * 0.025 * h


def test_richtmyer_meshkov_interface():
    p, _, U = case(H.IC_RICHTMYER_MESHKOV, 320)  # h = 1/80
    x, y, h = centres(p)
    # light gas (alpha1 = eps) for x <= xs(y); count per row
    light = (U[..., 5] - (1.0 - EPS5)) / (EPS5 - (1.0 - EPS5))
    xs = 2.9 + 0.1 * np.sin(2.0 * np.pi * (y + 0.25))
    assert np.max(np.abs(light.sum(axis=1) * h - xs)) < 0.5 * h
    beyond = kapila_cons(p, (1.4112, 5.0, -0.3613, 0.0, 1.6272 / 1.4, 1.0 - EPS5))
    m = x - 0.5 * h > 3.2
    assert np.allclose(U[:, m], beyond, rtol=1e-14, atol=0)


def test_underwater_explosion():
    p, _, U = case(H.IC_UNDERWATER_EXPLOSION, 400)  # h = 0.01
    x, y, h = centres(p)
    air = kapila_cons(p, (1.0, 1.225e-03, 0.0, 0.0, 1.01325, EPS5))
    water = kapila_cons(p, (1.0, 1.225e-03, 0.0, 0.0, 1.01325, 1.0 - EPS5))
    gas = kapila_cons(p, (1.0, 1.25, 0.0, 0.0, 1.0e4, EPS5))
    in_b = lambda X, Y: X ** 2 + (Y + 0.3) ** 2 <= 0.12 ** 2
    assert np.allclose(U[inner(U[..., 0], h, x, y, lambda X, Y: Y >= 0.0)], air, rtol=1e-14, atol=0)
    assert np.allclose(U[inner(U[..., 0], h, x, y, lambda X, Y: (Y < 0.0) & ~in_b(X, Y))], water, rtol=1e-14,
                       atol=0)
    assert np.allclose(U[inner(U[..., 0], h, x, y, in_b)], gas, rtol=1e-14, atol=0)
    below = y < -0.05
    area = fraction_area(U[below], h, EPS5, 1.0 - EPS5)
    assert abs(area - np.pi * 0.12 ** 2) < 0.1 * 2 * np.pi * 0.12 * h


def test_water_cylinder():
    p, _, U = case(H.IC_WATER_CYLINDER, 600)  # h = 2e-4
    x, y, h = centres(p)
    cyl = lambda X, Y: (X - 0.04) ** 2 + Y ** 2 < 1.1e-2 ** 2
    shocked = kapila_cons(p, (1.0, 3.7579, 574.57, 0.0, 6.6189e5, EPS5))
    still = kapila_cons(p, (1.0, 1.17, 0.0, 0.0, 1.01e5, EPS5))
    water = kapila_cons(p, (1000.0, 1.0, 0.0, 0.0, 1.01e5, 1.0 - EPS5))
    assert np.allclose(U[inner(U[..., 0], h, x, y, lambda X, Y: X < 0.028)], shocked, rtol=1e-14, atol=0)
    assert np.allclose(U[inner(U[..., 0], h, x, y, lambda X, Y: (X >= 0.028) & ~cyl(X, Y))], still, rtol=1e-14,
                       atol=0)
    assert np.allclose(U[inner(U[..., 0], h, x, y, cyl)], water, rtol=1e-14, atol=0)
    area = fraction_area(U, h, 1.0 - EPS5, EPS5)
    assert abs(area - np.pi * 1.1e-2 ** 2) < 0.1 * 2 * np.pi * 1.1e-2 * h


def test_air_cavity():
    p, _, U = case(H.IC_AIR_CAVITY, 240)  # h = 0.05
    x, y, h = centres(p)
    cav = lambda X, Y: (X - 6.0) ** 2 + (Y - 6.0) ** 2 < 9.0
    shock = kapila_cons(p, (1.325, 0.0012, -68.525, 0.0, 19153.0, 1.0 - EPS5))
    water = kapila_cons(p, (1.0, 0.0012, 0.0, 0.0, 1.0, 1.0 - EPS5))
    air = kapila_cons(p, (1.0, 0.0012, 0.0, 0.0, 1.0, EPS5))
    assert np.allclose(U[inner(U[..., 0], h, x, y, lambda X, Y: X > 11.4)], shock, rtol=1e-14, atol=0)
    assert np.allclose(U[inner(U[..., 0], h, x, y, lambda X, Y: (X <= 11.4) & ~cav(X, Y))], water, rtol=1e-14,
                       atol=0)
    assert np.allclose(U[inner(U[..., 0], h, x, y, cav)], air, rtol=1e-14, atol=0)
    area = fraction_area(U, h, EPS5, 1.0 - EPS5)
    assert abs(area - 9.0 * np.pi) < 0.1 * 2 * np.pi * 3.0 * h


def test_bubble_collapse_profile():
    p, _, U = case(H.IC_BUBBLE_COLLAPSE, 1000)  # h = 0.01, the tanh width is 0.06
    x, y, h = centres(p)
    X, Y = np.meshgrid(x, y)
    R = np.hypot(X, Y)
    th = np.tanh((R - 1.0) / (3.0 * 10.0 / 500.0))
    a1 = 0.5 * ((1e-5 + (1 - 1e-5)) + ((1 - 1e-5) - 1e-5) * th)
    assert np.max(np.abs(U[..., 5] - a1)) < 5e-3  # cell average vs centre: O(h^2 a'')
    far = R > 2.0
    pr = 0.5 * (101.0 + 99.0 * th[far])
    Vfar = np.stack([np.ones_like(pr), np.full_like(pr, 1e-2), 0 * pr, 0 * pr, pr, a1[far]], axis=-1)
    want = np.array([kapila_cons(p, V) for V in Vfar[:: 97]])
    assert np.allclose(U[far][:: 97], want, rtol=1e-9, atol=0)


@pytest.mark.parametrize("kind,V", [
    (H.IC_AIR_JET, (997.0, 1.225, 0.0, 0.0, 1.013e5, 1.0 - EPS5)),
    (H.IC_WATER_JET, (1.0, 1.0, 0.0, 0.0, 1.0e5, EPS5)),
])
def test_jet_cases_uniform(kind, V):
    p, _, U = case(kind, 60)
    assert p.bc[0] == H.BC_INLET
    assert np.allclose(U, kapila_cons(p, V), rtol=1e-14, atol=0)
