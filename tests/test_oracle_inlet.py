"""The Kapila left inlet in the oracle (rhs_function_prim.c:135-144,
rhs_function.c:135-144; the profile of kapila_5eqn.c:289-320, 485-513).

Pins that do not restate the formula:
  * the profile: a Mach-2 air jet, M * sqrt(gamma_2 p / rho_air) with p = 1.013
    and rho_air = 1, peaks at twice that speed on the axis and vanishes at
    |y| = 1 and beyond, symmetric in y; the volume fraction is exactly the mean
    of its limits at |y| = 1 and tends to 1e-6 inside the jet, 1 - 1e-6 outside;
  * the ghost fill: on the conservative path the inlet writes PDEPrim2Cons of
    that state into the ghost cells, so a grid whose first column already holds
    it must give bitwise the RHS of a transmissive left boundary (which copies
    the first column), with transmissive or reflective bottom and top;
  * the primitive path: the same to rounding (its first column goes through
    cons2prim);
  * other boundaries and systems have no inlet in the reference.
"""
import math

import numpy as np
import pytest

import oracle_py as orc
import states

T, R, P, I = orc.BC_TRANSMISSIVE, orc.BC_REFLECTIVE, orc.BC_PERIODIC, orc.BC_INLET
EOS = dict(g1=4.4, g2=1.4, p1=6000.0, p2=0.0)


def prob(prim, bc, nx=24, ny=20, N=2, system=orc.KAPILA):
    return orc.problem(system, nx, ny, N=N, prim=prim, x_min=0.0, y_min=-1.5, h=3.0 / ny, bc=bc, **EOS)


def test_inlet_profile_properties():
    P = prob(True, (I, T, T, T))
    c_air = math.sqrt(1.4 * 1.013 / 1.0)
    V0 = orc.inlet_prim(P, 0.0)
    assert V0[0] == 1000.0 and V0[1] == 1.0 and V0[3] == 0.0 and V0[4] == 1.013
    assert V0[2] == pytest.approx(2.0 * 2.0 * c_air, rel=1e-15)  # twice the mean speed M c on the axis
    for y in (0.3, 0.77, 0.999):
        a, b = orc.inlet_prim(P, y), orc.inlet_prim(P, -y)
        assert np.array_equal(a, b)  # even in y
        assert a[2] == pytest.approx(V0[2] * (1.0 - y * y), rel=1e-14)  # parabolic
    for y in (1.0, -1.0, 1.2, -3.0):
        assert orc.inlet_prim(P, y)[2] == 0.0
    assert orc.inlet_prim(P, 1.0)[5] == 0.5  # tanh(0): the mean of 1e-6 and 1 - 1e-6
    assert orc.inlet_prim(P, 0.0)[5] == pytest.approx(1e-6, abs=2e-5)  # air inside the jet
    assert orc.inlet_prim(P, 3.0)[5] == pytest.approx(1.0 - 1e-6, abs=1e-9)  # water outside
    # the profile is monotone across the interface
    ys = np.linspace(0.0, 2.0, 41)
    a = np.array([orc.inlet_prim(P, y)[5] for y in ys])
    assert np.all(np.diff(a) > 0)


def column_from_inlet(P, U):
    """U with its first column set to the inlet state of each row."""
    U = U.copy()
    for j in range(P.ny):
        U[j, 0] = orc.prim2cons(P, orc.inlet_prim(P, P.y_min + (j + 0.5) * P.h))
    return U


@pytest.mark.parametrize("ybc", [(T, T), (R, R), (R, T)])
@pytest.mark.parametrize("N", [1, 2, 3])
def test_inlet_equals_transmissive_when_the_first_column_is_the_inlet(ybc, N):
    Pi = prob(False, (I, T) + ybc, N=N)
    Pt = prob(False, (T, T) + ybc, N=N)
    U = column_from_inlet(Pi, states.inlet_state(orc, Pi, seed=3))
    a = orc.rhs(Pi, U, want_fallback=True)
    b = orc.rhs(Pt, U, want_fallback=True)
    assert a["rc"] == b["rc"] == 0
    assert np.array_equal(a["rhs"], b["rhs"]) and a["dt"] == b["dt"]
    assert np.array_equal(a["fallback"], b["fallback"])


@pytest.mark.parametrize("ybc", [(T, T), (R, R)])
def test_inlet_prim_path_equals_transmissive_to_rounding(ybc):
    Pi = prob(True, (I, T) + ybc)
    Pt = prob(True, (T, T) + ybc)
    U = column_from_inlet(Pi, states.inlet_state(orc, Pi, seed=4))
    a, b = orc.rhs(Pi, U), orc.rhs(Pt, U)
    assert a["rc"] == b["rc"] == 0
    for c in range(6):
        s = max(np.max(np.abs(b["rhs"][..., c])), 1e-300)
        assert np.max(np.abs(a["rhs"][..., c] - b["rhs"][..., c])) <= 1e-10 * s


def test_inlet_differs_from_transmissive_on_a_general_state():
    # guards against an inlet that silently falls back to a copy
    Pi, Pt = prob(False, (I, T, T, T)), prob(False, (T, T, T, T))
    U = states.inlet_state(orc, Pi, seed=5)
    a, b = orc.rhs(Pi, U), orc.rhs(Pt, U)
    assert not np.array_equal(a["rhs"][:, :3], b["rhs"][:, :3])
    assert np.array_equal(a["rhs"][:, 6:], b["rhs"][:, 6:])  # beyond the stencil reach of the left ghosts


@pytest.mark.parametrize("system,bc", [(orc.KAPILA, (T, I, T, T)), (orc.KAPILA, (T, T, I, T)),
                                       (orc.EULER, (I, T, T, T)), (orc.BN, (I, T, T, T))])
def test_inlet_elsewhere_is_unsupported(system, bc):
    g = dict(EOS)
    P = orc.problem(system, 8, 8, N=2, prim=True, bc=bc, **g)
    U = states.state(system, 8, 8, seed=1)
    assert orc.rhs(P, U)["rc"] == orc.E_UNSUPPORTED
