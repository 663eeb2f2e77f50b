"""Whole-RHS pins of the oracle (rhs_function_prim.c / rhs_function.c restated
in oracle/orc_rhs.c):

  * free-stream preservation: a uniform state gives RHS = 0 to round-off
    (weno3's one-sided slopes, e.g. -1.5u + 2u - 0.5u, cancel only to an ulp);
  * discrete conservation with periodic boundaries;
  * boundary index maps: a reflective (periodic, transmissive) boundary equals
    the interior of a larger transmissive domain built by mirroring (tiling,
    replicating) the data -- bitwise, as the stencil sees identical values;
  * mirror and transpose symmetry, 1D reduction;
  * the conservative Kapila path always falls back to first order
    (kapila_5eqn.c:278): its RHS equals a first-order Rusanov assembly;
  * dt = CFL h / (2 (|u| + c)) for a uniform state.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_py as orc
import states

T, R, P = orc.BC_TRANSMISSIVE, orc.BC_REFLECTIVE, orc.BC_PERIODIC
EOS = {0: (1.4, 1.4, 0.0, 0.0), 1: (1.4, 1.648, 0.0, 0.0), 2: (4.4, 1.4, 6e8, 0.0)}
VX = {0: (1,), 1: (2,), 2: (1, 5)}
VY = {0: (2,), 1: (3,), 2: (2, 6)}


def mk(system, nx, ny, prim=True, N=2, bc=(T, T, T, T), h=None, quirks=0):
    g1, g2, p1, p2 = EOS[system]
    return orc.problem(system, nx, ny, N=N, prim=prim, h=h or 1.0 / 64, cfl=0.5, bc=bc, g1=g1, g2=g2, p1=p1, p2=p2,
                       quirks=quirks)


def rhs(Pr, U):
    r = orc.rhs(Pr, U, want_fallback=True)
    assert r["rc"] == 0, r["err"].kind
    return r


def uniform(system, nx, ny):
    V = {0: [1.3, 0.4, -0.2, 2.0], 1: [1.2, 0.3, 15.0, -7.0, 1.1e5, 0.4],
         2: [1000.0, 1.0, -0.5, 2e5, 1.2, 0.5, 0.2, 1e5, 0.6]}[system]
    L = orc.lib()
    Q = np.zeros(orc.NVAR[system])
    L.orc_prim2cons(C.byref(mk(system, 4, 4)), orc.dptr(np.array(V, float)), orc.dptr(Q))
    return np.broadcast_to(Q, (ny, nx, len(Q))).copy(), np.array(V)


@pytest.mark.parametrize("system", [0, 1, 2])
@pytest.mark.parametrize("prim", [True, False])
@pytest.mark.parametrize("N", [0, 1, 2, 3])
@pytest.mark.parametrize("bc", [(T, T, T, T), (R, R, R, R), (P, P, P, P), (P, T, R, P)])
def test_free_stream_preservation(system, prim, N, bc):
    U, V = uniform(system, 9, 7)
    if R in bc:  # a reflective wall needs zero velocity for a uniform state to be steady
        for c in VX[system] + VY[system]:
            U[..., c] = 0.0
    Pr = mk(system, 9, 7, prim, N, bc)
    r = rhs(Pr, U)
    # scale of each component's flux divergence: ((|u| + c) |Q_c| + p) / h
    s_max = 0.5 * Pr.h / (2 * r["dt"])
    p_max = max(V[i] for i in {0: (3,), 1: (4,), 2: (3, 7)}[system])
    scale = (s_max * np.max(np.abs(U), axis=(0, 1)) + p_max) / Pr.h
    assert np.all(np.abs(r["rhs"]) <= 1e-14 * scale)


@pytest.mark.parametrize("system", [0, 1, 2])
@pytest.mark.parametrize("prim", [True, False])
def test_uniform_dt_closed_form(system, prim):
    U, V = uniform(system, 6, 5)
    h = 1.0 / 64
    r = rhs(mk(system, 6, 5, prim, 2, h=h), U)
    g1, g2, p1, p2 = EOS[system]
    if system == 0:
        s = max(abs(V[1]), abs(V[2])) + np.sqrt(g1 * V[3] / V[0])
    elif system == 1:
        a1, a2 = V[5], 1 - V[5]
        rho = a1 * V[0] + a2 * V[1]
        rc2 = 1 / (a1 / (g1 * (V[4] + p1)) + a2 / (g2 * (V[4] + p2)))
        s = max(abs(V[2]), abs(V[3])) + np.sqrt(rc2 / rho)
    else:
        s = max(max(abs(V[1]), abs(V[2])) + np.sqrt(g1 * (V[3] + p1) / V[0]),
                max(abs(V[5]), abs(V[6])) + np.sqrt(g2 * (V[7] + p2) / V[4]))
    assert abs(r["dt"] - 0.5 * h / (2 * s)) <= 1e-13 * r["dt"]


@pytest.mark.parametrize("system", [0, 1, 2])
@pytest.mark.parametrize("prim", [True, False])
def test_periodic_conservation(system, prim):
    U = states.state(system, 24, 20, seed=3, sharp=False)
    r = rhs(mk(system, 24, 20, prim, 2, (P, P, P, P), quirks=0), U)
    R_ = r["rhs"]
    cons = {0: [(0,), (1,), (2,), (3,)], 1: [(0,), (1,), (2,), (3,), (4,)],
            2: [(0,), (4,), (1, 5), (2, 6), (3, 7)]}[system]
    for group in cons:
        tot = sum(R_[..., c].sum() for c in group)
        scale = sum(np.abs(R_[..., c]).sum() for c in group) + 1e-300
        assert abs(tot) <= 1e-12 * scale, (group, tot, scale)


def mirror_x(U, system):
    M = U[:, ::-1, :].copy()
    for c in VX[system]:
        M[..., c] = -M[..., c]
    return M


def mirror_y(U, system):
    M = U[::-1, :, :].copy()
    for c in VY[system]:
        M[..., c] = -M[..., c]
    return M


@pytest.mark.parametrize("system", [0, 1, 2])
@pytest.mark.parametrize("prim", [True, False])
def test_reflective_equals_mirrored_domain(system, prim):
    """Reflective left and bottom walls == interior of a domain extended by
    the mirror images (corner block mirrored twice)."""
    nx, ny = 14, 12
    U = states.state(system, nx, ny, seed=4, sharp=False)
    bc = (R, T, R, T)
    r = rhs(mk(system, nx, ny, prim, 2, bc), U)
    top = np.concatenate([mirror_x(U, system), U], axis=1)
    big = np.concatenate([mirror_y(top, system), top], axis=0)
    rb = rhs(mk(system, 2 * nx, 2 * ny, prim, 2, (T, T, T, T)), big)
    assert np.array_equal(rb["rhs"][ny:, nx:], r["rhs"])
    assert np.array_equal(rb["fallback"][ny:, nx:], r["fallback"])


@pytest.mark.parametrize("system", [0, 1, 2])
def test_periodic_equals_tiled_domain(system):
    nx, ny = 10, 9
    U = states.state(system, nx, ny, seed=5, sharp=False)
    r = rhs(mk(system, nx, ny, True, 2, (P, P, P, P)), U)
    big = np.tile(U, (3, 3, 1))
    rb = rhs(mk(system, 3 * nx, 3 * ny, True, 2, (T, T, T, T)), big)
    assert np.array_equal(rb["rhs"][ny:2 * ny, nx:2 * nx], r["rhs"])


@pytest.mark.parametrize("system", [0, 1, 2])
def test_transmissive_equals_replicated_edges(system):
    nx, ny = 11, 8
    U = states.state(system, nx, ny, seed=6, sharp=False)
    r = rhs(mk(system, nx, ny, True, 2, (T, T, T, T)), U)
    big = np.pad(U, ((3, 3), (3, 3), (0, 0)), mode="edge")
    rb = rhs(mk(system, nx + 6, ny + 6, True, 2, (T, T, T, T)), big)
    assert np.array_equal(rb["rhs"][3:-3, 3:-3], r["rhs"])


def rel(a, b):
    return np.max(np.abs(a - b)) / (np.max(np.abs(b)) + 1e-300)


@pytest.mark.parametrize("system", [0, 1, 2])
def test_mirror_symmetry(system):
    U = states.state(system, 16, 12, seed=7, sharp=True)
    bc = (R, R, T, T)
    r = rhs(mk(system, 16, 12, True, 2, bc), U)
    rm = rhs(mk(system, 16, 12, True, 2, bc), mirror_x(U, system))
    assert rel(rm["rhs"], mirror_x(r["rhs"], system)) < 1e-12


@pytest.mark.parametrize("system", [0, 1, 2])
def test_transpose_symmetry(system):
    n = 14
    U = states.state(system, n, n, seed=8, sharp=True)
    r = rhs(mk(system, n, n, True, 2, (T, R, P, P)), U)
    perm = list(range(orc.NVAR[system]))
    for a, b in zip(VX[system], VY[system]):
        perm[a], perm[b] = b, a
    Ut = U.transpose(1, 0, 2)[..., perm].copy()
    rt = rhs(mk(system, n, n, True, 2, (P, P, T, R)), Ut)
    assert rel(rt["rhs"], r["rhs"].transpose(1, 0, 2)[..., perm]) < 1e-12


def test_one_dimensional_reduction():
    """y-uniform data without y-velocity: RHS independent of j, zero y-momentum."""
    U = states.state(0, 20, 1, seed=9, sharp=True)
    U[..., 2] = 0.0
    U2 = np.repeat(U, 7, axis=0)
    r = rhs(mk(0, 20, 7, True, 2, (T, T, P, P)), U2)["rhs"]
    assert np.all(r[..., 2] == 0.0)
    assert np.all(r == r[:1])


def test_conservative_kapila_is_first_order_rusanov():
    """kapila_5eqn.c:278 rejects every node with alpha_1 != 1, so the
    conservative path zeroes all slopes (rhs_function.c:416): face states are
    cell averages and the volume term vanishes.  Compare with a first-order
    Rusanov + jump-term assembly built from the pinned point Riemann solver."""
    L = orc.lib()
    nx, ny = 9, 7
    h = 1.0 / 64
    U = states.state(1, nx, ny, seed=10, sharp=True)
    Pr = mk(1, nx, ny, False, 2, (T, T, T, T), h=h)
    r = rhs(Pr, U)
    assert np.all(r["fallback"] == 1)
    G = np.pad(U, ((1, 1), (1, 1), (0, 0)), mode="edge")
    F = np.zeros((ny, nx + 1, 6))
    D = np.zeros_like(F)
    Gy = np.zeros((ny + 1, nx, 6))
    Ey = np.zeros_like(Gy)
    bad = C.c_int()
    for j in range(ny):
        for i in range(nx + 1):
            f, d = np.zeros(6), np.zeros(6)
            L.orc_riemann(C.byref(Pr), orc.dptr(G[j + 1, i].copy()), orc.dptr(G[j + 1, i + 1].copy()), 1.0, 0.0,
                          orc.dptr(f), orc.dptr(d), C.byref(bad))
            F[j, i], D[j, i] = f, d  # two Gauss points with weight 1/2 on a constant state
    for j in range(ny + 1):
        for i in range(nx):
            f, d = np.zeros(6), np.zeros(6)
            L.orc_riemann(C.byref(Pr), orc.dptr(G[j, i + 1].copy()), orc.dptr(G[j + 1, i + 1].copy()), 0.0, 1.0,
                          orc.dptr(f), orc.dptr(d), C.byref(bad))
            Gy[j, i], Ey[j, i] = f, d
    expect = (-(F[:, 1:] - F[:, :-1] + 0.5 * (D[:, 1:] + D[:, :-1])) / h
              - (Gy[1:] - Gy[:-1] + 0.5 * (Ey[1:] + Ey[:-1])) / h)
    assert rel(r["rhs"], expect) < 1e-13


def test_inadmissible_state_reports_lowest_face():
    nx, ny = 12, 10
    U = states.state(0, nx, ny, seed=11, sharp=False)
    U[4, 5, 3] = -abs(U[4, 5, 3])
    U[8, 2, 3] = -abs(U[8, 2, 3])
    res = orc.rhs(mk(0, nx, ny, True, 2, (T, T, T, T), h=0.1), U)
    assert res["rc"] == orc.lib().orc_nvar(0) + 1  # ORC_E_INADMISSIBLE == 5
    e = res["err"]
    # lowest x-face touching the first bad cell (row 4): face i=5 (left face
    # of cell 5), x = x_min + 5 h, y = y_min + 4 h
    assert e.kind == 2 and abs(e.x - 0.5) < 1e-12 and abs(e.y - 0.4) < 1e-12
