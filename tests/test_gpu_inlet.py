"""The Kapila left inlet on the device against the oracle: the inlet ghost
columns (rhs_function_prim.c:135-144, rhs_function.c:135-144) with every
bottom/top boundary kind, both paths, orders 1-3, several tilings (split tiles,
persistent queue), N steps, and inside a batch and a server.  Bar: bitwise,
as everywhere else on this path."""
import numpy as np
import pytest

import oracle_py as orc
import states

pytestmark = pytest.mark.gpu

T, R, P, I = orc.BC_TRANSMISSIVE, orc.BC_REFLECTIVE, orc.BC_PERIODIC, orc.BC_INLET
EOS = states.INLET_EOS


def pair(hip, nx, ny, bc, prim, order):
    h = 3.0 / ny
    p = hip.make_problem(hip.KAPILA, nx, ny, order=order, prim=prim, domain=(0.0, nx * h, -1.5, 1.5), cfl=0.5,
                         bc=bc, **EOS)
    op = orc.problem(orc.KAPILA, nx, ny, N=order, prim=prim, x_min=0.0, y_min=-1.5, h=h, bc=bc, g1=EOS["gamma1"],
                     g2=EOS["gamma2"], p1=EOS["pi1"], p2=EOS["pi2"])
    return p, op


@pytest.mark.parametrize("ybc", [(T, T), (R, R), (P, P), (R, T)])
@pytest.mark.parametrize("prim", [True, False])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_inlet_rhs_matches_oracle(hip, ybc, prim, order):
    p, op = pair(hip, 70, 36, (I, T) + ybc, prim, order)
    U = states.inlet_state(orc, op, seed=7)
    ref = orc.rhs(op, U, want_fallback=True)
    with hip.Context(p) as ctx:
        Rg, dt, mask, rc = ctx.rhs(U, fallback=True, allow_inadmissible=True)
    print(f"[inlet rhs] ybc={ybc} prim={prim} N={order} rc={rc}/{ref['rc']} fallback={int(mask.sum())}", flush=True)
    assert rc == ref["rc"]
    assert np.array_equal(Rg, ref["rhs"])
    assert dt == ref["dt"]
    assert np.array_equal(mask, ref["fallback"])


@pytest.mark.parametrize("hrows,grid,layout", [(1, 0, 0), (2, 0, 0), (5, 0, 0), (3, 9, 0), (64, 0, 0)])
def test_inlet_rhs_tilings(hip, hrows, grid, layout):
    p, op = pair(hip, 130, 40, (I, R, P, P), True, 2)
    U = states.inlet_state(orc, op, seed=8)
    ref = orc.rhs(op, U)
    with hip.Context(p) as ctx:
        ctx.set_tiling(hrows, grid)
        Rg, dt, rc = ctx.rhs(U, allow_inadmissible=True)
    assert rc == ref["rc"]
    assert np.array_equal(Rg, ref["rhs"]) and dt == ref["dt"]


@pytest.mark.parametrize("prim,order,ybc", [(True, 2, (R, R)), (False, 2, (T, T)), (True, 3, (P, P)),
                                            (True, 1, (R, T)), (False, 2, (R, R))])
def test_inlet_steps_match_oracle(hip, prim, order, ybc):
    """Ten steps; where the oracle meets an inadmissible state (the inlet's
    jump against the interior can drive the conservative path there), the
    device stops after the same step and reports the same face."""
    p, op = pair(hip, 64, 30, (I, T) + ybc, prim, order)
    U0 = states.inlet_state(orc, op, seed=9)
    ref = orc.solve(op, U0, 10)
    with hip.Context(p) as ctx:
        ctx.set_state(U0)
        ctx.begin(0.0, float("inf"), 10, hip.TS_CLIP)
        ctx.advance(10)
        st = ctx.sync(allow_inadmissible=True)
        U = ctx.get_state()
        dts = ctx.dt_history(st.steps)
        msg, x, y = ctx.last_error() if st.err else ("", 0.0, 0.0)
    print(f"[inlet steps] prim={prim} N={order} ybc={ybc} steps={st.steps}/{ref['steps']} rc={st.err}/{ref['rc']} "
          f"{msg!r}", flush=True)
    assert st.err == ref["rc"]
    assert st.steps == ref["steps"]
    assert st.t == ref["t"]
    assert np.array_equal(dts, ref["dts"])
    assert np.array_equal(U, ref["U"], equal_nan=True)
    if ref["rc"]:
        assert (x, y) == (ref["err"].x, ref["err"].y)


def test_inlet_in_batch_and_server(hip):
    """An inlet instance among others keeps its table through the batch slab
    and the server's arena; each result equals its solo run bitwise."""
    p1, op1 = pair(hip, 64, 30, (I, T, R, R), True, 2)
    p2, op2 = pair(hip, 48, 40, (I, R, P, P), False, 2)
    eos = states.EOS[1]
    p3 = hip.make_problem(hip.KAPILA, 40, 30, order=2, prim=True, domain=(0.0, 1.0, 0.0, 0.75), cfl=0.5,
                          bc=(T, T, R, R), **eos)
    U0s = [states.inlet_state(orc, op1, seed=10), states.inlet_state(orc, op2, seed=11), states.state(1, 40, 30, seed=12)]
    probs, steps = [p1, p2, p3], [7, 5, 6]
    solo = []
    for p, U0, n in zip(probs, U0s, steps):
        with hip.Context(p) as ctx:
            solo.append(ctx.solve(U0, n))
    with hip.Batch(probs) as b:
        for i, U in enumerate(U0s):
            b.set_state(i, U)
        b.begin(max_steps=np.array(steps, dtype=np.int32))
        b.advance(max(steps) + 1)
        st = b.sync()
        for i in range(3):
            assert st[i].steps == solo[i][1].steps
            assert np.array_equal(b.get_state(i), solo[i][0]), f"batch instance {i}"
    with hip.Server(hbm_budget_bytes=1 << 26) as srv:
        ids = [srv.submit(p, U0, max_steps=n) for p, U0, n in zip(probs, U0s, steps)]
        srv.drain()
        assert srv.bytes_in_use() == 0
        for i, rid in enumerate(ids):
            U, info = srv.result(rid)
            assert info.status.steps == steps[i]
            assert np.array_equal(U, solo[i][0]), f"served request {i}"


@pytest.mark.parametrize("system,bc", [(0, (I, T, T, T)), (2, (I, T, T, T)), (1, (T, I, T, T)),
                                       (1, (T, T, T, I))])
def test_inlet_elsewhere_is_unsupported(hip, system, bc):
    eos = states.EOS[system]
    p = hip.make_problem(system, 20, 20, order=2, prim=True, bc=bc, **eos)
    with pytest.raises(hip.HypeError) as e:
        hip.Context(p)
    assert e.value.code == hip.E_UNSUPPORTED
