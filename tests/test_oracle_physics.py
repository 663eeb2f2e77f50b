"""Pins of the oracle's point functions (euler.c, kapila_5eqn.c, baer_nunziato.c,
riemann.c restated in oracle/orc_physics.c) against independent textbook forms:
stiffened-gas energies, physical fluxes, Wood's mixture sound speed, the Kapila
K-term identity K = a1 (rho c^2 / (rho1 c1^2) - 1), Rusanov consistency and the
pairwise cancellation of the Baer-Nunziato interface terms."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as orc

EOS = {
    orc.EULER: dict(g1=1.4, g2=1.4, p1=0.0, p2=0.0),
    orc.KAPILA: dict(g1=4.4, g2=1.4, p1=6.0e3, p2=0.0),
    orc.BN: dict(g1=4.4, g2=1.4, p1=6.0e8, p2=0.0),
}


def prob(system, prim=True, quirks=0, **kw):
    e = dict(EOS[system])
    e.update(kw)
    return orc.problem(system, 8, 8, prim=prim, quirks=quirks, **e)


def vec(n):
    return np.zeros(n, dtype=np.float64)


def call_vec(fn, P, x, nout):
    out = vec(nout)
    fn(C.byref(P), orc.dptr(np.ascontiguousarray(x, dtype=np.float64)), orc.dptr(out))
    return out


def prim_sample(system, rng):
    if system == orc.EULER:
        return np.array([rng.uniform(0.1, 2), rng.normal(), rng.normal(), rng.uniform(0.1, 3)])
    if system == orc.KAPILA:
        return np.array([rng.uniform(500, 1500), rng.uniform(0.5, 2), rng.normal(scale=50), rng.normal(scale=50),
                         rng.uniform(1e4, 1e6), rng.uniform(0.01, 0.99)])
    return np.array([rng.uniform(900, 1100), rng.normal(), rng.normal(), rng.uniform(1e5, 1e9),
                     rng.uniform(0.5, 2), rng.normal(), rng.normal(), rng.uniform(1e4, 1e6), rng.uniform(0.01, 0.99)])


def cons_textbook(system, V, e):
    """Conserved variables from primitives with the stiffened-gas law written
    out independently: rho e_k = (p + g_k pi_k)/(g_k - 1)."""
    if system == orc.EULER:
        r, u, v, p = V
        return np.array([r, r * u, r * v, p / (e["g1"] - 1) + 0.5 * r * (u * u + v * v)])
    if system == orc.KAPILA:
        r1, r2, u, v, p, a1 = V
        a2 = 1 - a1
        rho = a1 * r1 + a2 * r2
        rhoe = a1 * (p + e["g1"] * e["p1"]) / (e["g1"] - 1) + a2 * (p + e["g2"] * e["p2"]) / (e["g2"] - 1)
        return np.array([a1 * r1, a2 * r2, rho * u, rho * v, rhoe + 0.5 * rho * (u * u + v * v), a1])
    r1, u1, v1, p1, r2, u2, v2, p2, f = V
    E1 = (p1 + e["g1"] * e["p1"]) / (e["g1"] - 1) + 0.5 * r1 * (u1 * u1 + v1 * v1)
    E2 = (p2 + e["g2"] * e["p2"]) / (e["g2"] - 1) + 0.5 * r2 * (u2 * u2 + v2 * v2)
    return np.array([f * r1, f * r1 * u1, f * r1 * v1, f * E1, (1 - f) * r2, (1 - f) * r2 * u2, (1 - f) * r2 * v2,
                     (1 - f) * E2, f])


def flux_textbook(system, V, e, nx, ny):
    """Physical flux and |u_n| + c (Wood's mixture speed for Kapila)."""
    Q = cons_textbook(system, V, e)
    if system == orc.EULER:
        r, u, v, p = V
        un = nx * u + ny * v
        F = np.array([r * un, r * u * un + nx * p, r * v * un + ny * p, un * (Q[3] + p)])
        return F, abs(un) + np.sqrt(e["g1"] * p / r)
    if system == orc.KAPILA:
        r1, r2, u, v, p, a1 = V
        a2 = 1 - a1
        rho = a1 * r1 + a2 * r2
        un = nx * u + ny * v
        c1s = e["g1"] * (p + e["p1"]) / r1
        c2s = e["g2"] * (p + e["p2"]) / r2
        rc2 = 1.0 / (a1 / (r1 * c1s) + a2 / (r2 * c2s))  # Wood: 1/(rho c^2) = sum a_k/(rho_k c_k^2)
        F = np.array([a1 * r1 * un, a2 * r2 * un, rho * u * un + nx * p, rho * v * un + ny * p, (Q[4] + p) * un, 0])
        return F, abs(un) + np.sqrt(rc2 / rho)
    r1, u1, v1, p1, r2, u2, v2, p2, f = V
    un1, un2 = nx * u1 + ny * v1, nx * u2 + ny * v2
    F = np.array([Q[0] * un1, f * (r1 * u1 * un1 + nx * p1), f * (r1 * v1 * un1 + ny * p1), un1 * (Q[3] + f * p1),
                  Q[4] * un2, (1 - f) * (r2 * u2 * un2 + nx * p2), (1 - f) * (r2 * v2 * un2 + ny * p2),
                  un2 * (Q[7] + (1 - f) * p2), 0])
    s1 = abs(un1) + np.sqrt(e["g1"] * (p1 + e["p1"]) / r1)
    s2 = abs(un2) + np.sqrt(e["g2"] * (p2 + e["p2"]) / r2)
    return F, max(s1, s2)


def rel(a, b):
    return np.max(np.abs(a - b) / (np.abs(b) + 1e-300 + np.max(np.abs(b)) * 1e-14))


@pytest.mark.parametrize("system", [0, 1, 2])
def test_prim2cons_textbook_and_round_trip(system):
    L = orc.lib()
    rng = np.random.default_rng(system)
    P = prob(system)
    nv = orc.NVAR[system]
    for _ in range(50):
        V = prim_sample(system, rng)
        Q = call_vec(L.orc_prim2cons, P, V, nv)
        assert rel(Q, cons_textbook(system, V, EOS[system])) < 1e-13
        V2 = call_vec(L.orc_cons2prim, P, Q, nv)
        assert rel(V2, V) < 1e-9 if system == orc.BN else rel(V2, V) < 1e-12


def test_kapila_equal_phases_reduce_to_euler():
    L = orc.lib()
    P = prob(orc.KAPILA, g1=1.4, g2=1.4, p1=0.0, p2=0.0)
    rng = np.random.default_rng(7)
    for _ in range(20):
        Q = np.array([rng.uniform(0.1, 1), rng.uniform(0.1, 1), rng.normal(), rng.normal(), 0.0, rng.uniform(0.05, 0.95)])
        rho = Q[0] + Q[1]
        Q[4] = 5.0 + 0.5 * (Q[2] ** 2 + Q[3] ** 2) / rho
        V = call_vec(L.orc_cons2prim, P, Q, 6)
        p_euler = 0.4 * (Q[4] - 0.5 * (Q[2] ** 2 + Q[3] ** 2) / rho)
        assert abs(V[4] - p_euler) <= 1e-13 * p_euler


@pytest.mark.parametrize("system", [0, 1, 2])
@pytest.mark.parametrize("prim", [True, False])
@pytest.mark.parametrize("nxny", [(1.0, 0.0), (0.0, 1.0)])
def test_flux_and_wave_speed_textbook(system, prim, nxny):
    L = orc.lib()
    rng = np.random.default_rng(10 + system)
    P = prob(system, prim=prim)
    nv = orc.NVAR[system]
    bad = C.c_int()
    for _ in range(30):
        V = prim_sample(system, rng)
        F = vec(nv)
        if prim:
            s = L.orc_flux_prim(C.byref(P), orc.dptr(V), nxny[0], nxny[1], orc.dptr(F), C.byref(bad))
        else:
            Q = cons_textbook(system, V, EOS[system])
            s = L.orc_flux(C.byref(P), orc.dptr(Q), nxny[0], nxny[1], orc.dptr(F), C.byref(bad))
        Ft, st = flux_textbook(system, V, EOS[system], *nxny)
        scale = np.max(np.abs(Ft))
        assert np.max(np.abs(F - Ft)) <= 1e-9 * scale
        assert abs(s - st) <= 1e-12 * st
        assert bad.value == 0


@pytest.mark.parametrize("system", [0, 1, 2])
@pytest.mark.parametrize("prim", [True, False])
def test_rusanov_consistency(system, prim):
    """Equal states: numerical flux = physical flux, jump term = 0."""
    L = orc.lib()
    rng = np.random.default_rng(20 + system)
    P = prob(system, prim=prim)
    nv = orc.NVAR[system]
    bad = C.c_int()
    for nx_, ny_ in ((1.0, 0.0), (0.0, 1.0)):
        V = prim_sample(system, rng)
        X = V if prim else cons_textbook(system, V, EOS[system])
        F, D, Fp = vec(nv), vec(nv), vec(nv)
        s = L.orc_riemann(C.byref(P), orc.dptr(X), orc.dptr(X), nx_, ny_, orc.dptr(F), orc.dptr(D), C.byref(bad))
        Ft, st = flux_textbook(system, V, EOS[system], nx_, ny_)
        assert np.max(np.abs(F - Ft)) <= 1e-9 * np.max(np.abs(Ft))
        assert np.all(D == 0.0)
        assert abs(s - st) <= 1e-12 * st


def test_rusanov_dissipation_uses_conserved_jump():
    """F = (F_L + F_R)/2 - s/2 (Q_R - Q_L) with s = max(s_L, s_R), also on the
    primitive path (riemann.c:123-134)."""
    L = orc.lib()
    rng = np.random.default_rng(3)
    for prim in (True, False):
        P = prob(orc.EULER, prim=prim)
        VL, VR = prim_sample(0, rng), prim_sample(0, rng)
        QL, QR = cons_textbook(0, VL, EOS[0]), cons_textbook(0, VR, EOS[0])
        XL, XR = (VL, VR) if prim else (QL, QR)
        F, D = vec(4), vec(4)
        bad = C.c_int()
        s = L.orc_riemann(C.byref(P), orc.dptr(XL), orc.dptr(XR), 1.0, 0.0, orc.dptr(F), orc.dptr(D), C.byref(bad))
        FL, sL = flux_textbook(0, VL, EOS[0], 1, 0)
        FR, sR = flux_textbook(0, VR, EOS[0], 1, 0)
        assert s == max(sL, sR) or abs(s - max(sL, sR)) < 1e-14 * s
        expect = 0.5 * (FL + FR) - 0.5 * s * (QR - QL)
        assert np.max(np.abs(F - expect)) <= 1e-12 * np.max(np.abs(expect))


def test_kapila_ncp_k_identity():
    """B_5 = u a_x + v a_y - K (u_x + v_y) with K = a1 (rho c^2/(rho1 c1^2) - 1),
    c the Wood speed: an identity independent of the reference's fraction."""
    L = orc.lib()
    P = prob(orc.KAPILA)
    e = EOS[orc.KAPILA]
    rng = np.random.default_rng(5)
    for _ in range(30):
        V = prim_sample(orc.KAPILA, rng)
        gx, gy = rng.normal(size=6), rng.normal(size=6)
        B = vec(6)
        L.orc_ncp_prim(C.byref(P), orc.dptr(V), orc.dptr(gx), orc.dptr(gy), orc.dptr(B))
        r1, r2, u, v, p, a1 = V
        a2 = 1 - a1
        rho = a1 * r1 + a2 * r2
        c1s, c2s = e["g1"] * (p + e["p1"]) / r1, e["g2"] * (p + e["p2"]) / r2
        rc2 = 1.0 / (a1 / (r1 * c1s) + a2 / (r2 * c2s))
        K = a1 * (rc2 / (r1 * c1s) - 1.0)
        expect = u * gx[5] + v * gy[5] - K * (gx[2] + gy[3])
        assert np.all(B[:5] == 0.0)
        assert abs(B[5] - expect) <= 1e-10 * (abs(u * gx[5]) + abs(v * gy[5]) + abs(K * (gx[2] + gy[3])))


def test_kapila_ncp_conservative_quotient_rule():
    """Conservative path: u_x from momentum and density gradients by the
    quotient rule gives the same B_5 as the primitive form."""
    L = orc.lib()
    P, Pp = prob(orc.KAPILA, prim=False), prob(orc.KAPILA)
    e = EOS[orc.KAPILA]
    rng = np.random.default_rng(6)
    for _ in range(20):
        V = prim_sample(orc.KAPILA, rng)
        Q = cons_textbook(orc.KAPILA, V, e)
        dV_x, dV_y = rng.normal(size=6) * 1e-3, rng.normal(size=6) * 1e-3
        # consistent conserved gradients: exact directional derivatives of
        # prim2cons by the complex step
        hs = 1e-30
        gQx = cons_textbook(1, V + 1j * hs * dV_x, e).imag / hs
        gQy = cons_textbook(1, V + 1j * hs * dV_y, e).imag / hs
        B, Bp = vec(6), vec(6)
        L.orc_ncp(C.byref(P), orc.dptr(Q), orc.dptr(gQx), orc.dptr(gQy), orc.dptr(B))
        L.orc_ncp_prim(C.byref(Pp), orc.dptr(V), orc.dptr(dV_x), orc.dptr(dV_y), orc.dptr(Bp))
        assert abs(B[5] - Bp[5]) <= 1e-9 * (abs(Bp[5]) + 1e-12)


@pytest.mark.parametrize("prim", [True, False])
def test_bn_interface_terms_cancel(prim):
    """u_I = u_1, p_I = p_2; momentum and energy exchange cancel between phases."""
    L = orc.lib()
    P = prob(orc.BN, prim=prim)
    e = EOS[orc.BN]
    rng = np.random.default_rng(8)
    for _ in range(20):
        V = prim_sample(orc.BN, rng)
        X = V if prim else cons_textbook(orc.BN, V, e)
        gx, gy = rng.normal(size=9), rng.normal(size=9)
        B = vec(9)
        (L.orc_ncp_prim if prim else L.orc_ncp)(C.byref(P), orc.dptr(X), orc.dptr(gx), orc.dptr(gy), orc.dptr(B))
        pI, uI, vI = V[7], V[1], V[2]
        assert B[1] == -B[5] and B[2] == -B[6] and B[3] == -B[7]
        assert B[0] == 0.0 and B[4] == 0.0
        assert abs(B[5] - pI * gx[8]) <= 1e-9 * abs(pI * gx[8]) + 1e-300
        assert abs(B[8] - (uI * gx[8] + vI * gy[8])) <= 1e-9 * (abs(uI * gx[8]) + abs(vI * gy[8]))


def test_bn_primitive_flux_p2_quirk():
    """baer_nunziato.c:271 reads p_2 = V[8]; the quirk switch reproduces it."""
    L = orc.lib()
    rng = np.random.default_rng(9)
    V = prim_sample(orc.BN, rng)
    bad = C.c_int()
    F1, F0 = vec(9), vec(9)
    L.orc_flux_prim(C.byref(prob(orc.BN, quirks=orc.QUIRK_BN_P2_IS_PHI)), orc.dptr(V), 1.0, 0.0, orc.dptr(F1),
                    C.byref(bad))
    L.orc_flux_prim(C.byref(prob(orc.BN)), orc.dptr(V), 1.0, 0.0, orc.dptr(F0), C.byref(bad))
    f, r2, u2 = V[8], V[4], V[5]
    assert abs(F1[5] - (1 - f) * (r2 * u2 * u2 + V[8])) <= 1e-12 * abs(F1[5])
    assert abs(F0[5] - (1 - f) * (r2 * u2 * u2 + V[7])) <= 1e-12 * abs(F0[5])


def test_admissibility_thresholds():
    L = orc.lib()
    rng = np.random.default_rng(11)
    # primitive checks (euler.c:292-310, kapila_5eqn.c:450-479, baer_nunziato.c:350-385)
    for system in (0, 1, 2):
        P = prob(system)
        V = prim_sample(system, rng)
        assert L.orc_pad_prim(C.byref(P), orc.dptr(V)) == 1
        W = V.copy()
        W[0] = 1e-15  # Kapila tests the mixture density: make both phases vanish
        if system == orc.KAPILA:
            W[1] = 1e-15
        assert L.orc_pad_prim(C.byref(P), orc.dptr(W)) == 0
    P = prob(orc.KAPILA)
    V = prim_sample(orc.KAPILA, rng)
    for a in (-1e-9, 1.0 + 1e-9):
        W = V.copy()
        W[5] = a
        assert L.orc_pad_prim(C.byref(P), orc.dptr(W)) == 0
    for a in (0.0, 1.0):
        W = V.copy()
        W[5] = a
        assert L.orc_pad_prim(C.byref(P), orc.dptr(W)) == 1
    W = V.copy()
    W[4] = -P.p1 - 1.0  # negative mixture sound speed squared
    assert L.orc_pad_prim(C.byref(P), orc.dptr(W)) == 0
    # conservative Kapila check tests alpha_1 < 1 || alpha_1 > 1 (kapila_5eqn.c:278)
    Pc = prob(orc.KAPILA, prim=False)
    Q = cons_textbook(orc.KAPILA, prim_sample(orc.KAPILA, rng), EOS[orc.KAPILA])
    assert L.orc_pad(C.byref(Pc), orc.dptr(Q)) == 0
    # Euler: p < 1e-12 fails
    P = prob(orc.EULER)
    V = np.array([1.0, 0.0, 0.0, 1e-13])
    assert L.orc_pad_prim(C.byref(P), orc.dptr(V)) == 0
    Q = cons_textbook(0, np.array([1.0, 0.3, 0.0, 1.0]), EOS[0])
    assert L.orc_pad(C.byref(P), orc.dptr(Q)) == 1


def test_flux_reports_first_failing_check():
    L = orc.lib()
    bad = C.c_int()
    F = vec(9)
    # BN: negative gas density is reported before negative gas pressure
    V = prim_sample(orc.BN, np.random.default_rng(12))
    V[4] = -1.0
    V[7] = -1e9
    L.orc_flux_prim(C.byref(prob(orc.BN)), orc.dptr(V), 1.0, 0.0, orc.dptr(F), C.byref(bad))
    assert bad.value == 5
    V = np.array([1.0, 0.0, 0.0, -1.0])
    L.orc_flux_prim(C.byref(prob(orc.EULER)), orc.dptr(V), 1.0, 0.0, orc.dptr(F), C.byref(bad))
    assert bad.value == 2
