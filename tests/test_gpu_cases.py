"""The reference's Kapila test cases (kapila_5eqn.c:565-996) in the settings
their header comments give (hype_case_setup), a few SSP-RK steps on the
device against the oracle's solve loop (hype.c:337-384) on the same initial
state.  AirJet and WaterJet drive the left inlet (rhs_function_prim.c:135-144);
AirCavity is the case the reference's InitialCondition selects
(initialize_solution.c:14).  Bar: bitwise."""
import numpy as np
import pytest

import oracle_py as orc

pytestmark = pytest.mark.gpu


def run_case(hip, kind, nx, ny, steps, prim=True, order=2):
    p, t_end = hip.case_problem(kind, nx, ny, order=order, prim=prim)
    U0 = hip.initial_state(p, kind)
    op = orc.problem(orc.KAPILA, nx, ny, N=order, prim=prim, domain=(p.x_min, p.x_max, p.y_min, p.y_max),
                     cfl=p.cfl, bc=tuple(p.bc), g1=p.gamma1, g2=p.gamma2, p1=p.pi1, p2=p.pi2)
    ref = orc.solve(op, U0, steps, t_end=t_end)
    with hip.Context(p) as ctx:
        ctx.set_state(U0)
        ctx.begin(0.0, t_end, steps, hip.TS_CLIP)
        ctx.advance(steps + 1)
        st = ctx.sync(allow_inadmissible=True)
        U = ctx.get_state()
    print(f"[case] kind={kind} {nx}x{ny} prim={prim} N={order} steps={st.steps}/{ref['steps']} "
          f"rc={st.err}/{ref['rc']} t={st.t:.6e}", flush=True)
    assert st.err == ref["rc"]
    assert st.steps == ref["steps"] and st.t == ref["t"]
    assert np.array_equal(U, ref["U"])
    return st


@pytest.mark.parametrize("kind,nx,ny", [
    (13, 120, 120),   # AirJet: [0,30] x [-15,15], inlet left
    (14, 160, 80),    # WaterJet: [0,1] x [-0.25,0.25], inlet left
    (11, 96, 96),     # AirCavity: [0,12]^2
])
@pytest.mark.parametrize("prim", [True, False])
def test_case_steps_match_oracle(hip, kind, nx, ny, prim):
    run_case(hip, kind, nx, ny, 6, prim=prim)


@pytest.mark.parametrize("kind,nx,ny", [
    (5, 200, 20), (6, 160, 40), (7, 160, 40), (8, 160, 40), (9, 120, 90), (10, 120, 80), (12, 160, 80),
])
def test_other_cases_steps_match_oracle(hip, kind, nx, ny):
    run_case(hip, kind, nx, ny, 4)
