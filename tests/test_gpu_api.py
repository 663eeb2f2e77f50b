"""Entry points of the C ABI not covered elsewhere: the primitive pass
(hype_cons2prim, the device ComputePrimitiveVariables of post_process.c:165-200)
against the oracle's restatement, timed advances, and a batch whose tile
height is set by the caller."""
import numpy as np
import pytest

import oracle_py as orc
import states

pytestmark = pytest.mark.gpu

T, R = orc.BC_TRANSMISSIVE, orc.BC_REFLECTIVE


def pair(hip, system, nx, ny, prim=True):
    eos = states.EOS[system]
    dom = (0.0, 1.0, 0.0, ny / nx)
    p = hip.make_problem(system, nx, ny, order=2, prim=prim, domain=dom, cfl=0.5, bc=(T, T, R, R), **eos)
    op = orc.problem(system, nx, ny, N=2, prim=True, domain=dom, cfl=0.5, bc=(T, T, R, R), g1=eos["gamma1"],
                     g2=eos["gamma2"], p1=eos["pi1"], p2=eos["pi2"])
    return p, op


@pytest.mark.parametrize("system", [0, 1, 2])
def test_cons2prim_matches_oracle(hip, system):
    p, op = pair(hip, system, 97, 41)
    U = states.state(system, 97, 41, seed=21)
    ref = orc.rhs(op, U, want_prim=True)["W"]  # the primitive pass of RHSFunctionPrim (Ctx->W)
    with hip.Context(p) as ctx:
        W = ctx.cons2prim(U)
    assert np.array_equal(W, ref, equal_nan=True)


def test_advance_timed_matches_untimed(hip):
    p, _ = pair(hip, 1, 80, 44)
    U0 = states.state(1, 80, 44, seed=22)
    with hip.Context(p) as ctx:
        U1, s1 = ctx.solve(U0, 8)
        ctx.set_state(U0)
        ctx.begin(0.0, float("inf"), 8, hip.TS_CLIP)
        ms = ctx.advance_timed(8)
        s2 = ctx.sync()
        U2 = ctx.get_state()
    assert ms > 0.0
    assert s1.steps == s2.steps == 8 and s1.t == s2.t
    assert np.array_equal(U1, U2)


@pytest.mark.parametrize("rows", [1, 3, 11])
def test_batch_rows_set_by_caller(hip, rows):
    probs, U0s = [], []
    for k, (s, nx, ny) in enumerate([(1, 70, 40), (0, 45, 33), (2, 40, 24)]):
        p, _ = pair(hip, s, nx, ny, prim=(s != 2))
        probs.append(p)
        U0s.append(states.state(s, nx, ny, seed=30 + k))
    with hip.Batch(probs) as b:
        b.set_rows(rows)
        assert b.tiling()[0] == rows
        for i, U in enumerate(U0s):
            b.set_state(i, U)
        b.begin(max_steps=np.array([4, 5, 6], dtype=np.int32))
        ms = b.advance_timed(7)
        st = b.sync()
        got = [b.get_state(i) for i in range(3)]
    assert ms > 0.0
    for i, (p, U0, n) in enumerate(zip(probs, U0s, (4, 5, 6))):
        with hip.Context(p) as ctx:
            U, s1 = ctx.solve(U0, n)
        assert st[i].steps == s1.steps == n
        assert np.array_equal(got[i], U)
