"""End-to-end known answer: convergence of the Kapila SmoothVortex
(kapila_5eqn.c:527-562) with WENO3 + SSPRK(4,3) on the device.

The reference's only published numbers are the L2/Linf errors of alpha_1
after one period (Results/5-Eqn-Convergence.ods, sheet "N=2"; ErrorNorms,
post_process.c:74-160, compares component 5 only, L2 norm divided by N_x).  The run configuration
behind them (CFL, EOS constants) was not recorded, so absolute errors are
not comparable; the spreadsheet's orders rise from 1.5 towards 3 (2.92 at
256->512, 2.99 at 512->1024) and that trend is what is checked here.  The
coarsest run is also compared bit for bit with the CPU oracle."""
import numpy as np
import pytest

import oracle_py as orc

pytestmark = pytest.mark.gpu

P = orc.BC_PERIODIC
EOS = dict(gamma1=4.0, gamma2=1.4, pi1=20.0, pi2=0.0)


def vortex(hip, n):
    p = hip.make_problem(hip.KAPILA, n, n, order=2, prim=True, domain=(-3.0, 3.0, -3.0, 3.0), cfl=0.9,
                         bc=(P,) * 4, **EOS)
    return p, hip.initial_state(p, hip.IC_SMOOTH_VORTEX)


def run_period(hip, p, U0, t_end=6.0):
    with hip.Context(p) as ctx:
        ctx.set_state(U0)
        ctx.begin(0.0, t_end, 1 << 30, hip.TS_CLIP | hip.TS_GRAPH)
        ctx.advance(int(2.5 * p.nx) + 8)
        st = ctx.sync()
        while not st.done:
            ctx.advance(64)
            st = ctx.sync()
        return ctx.get_state(), st


def test_smooth_vortex_coarse_run_matches_oracle(hip):
    p, U0 = vortex(hip, 48)
    U, st = run_period(hip, p, U0)
    op = orc.problem(1, 48, 48, N=2, prim=True, domain=(-3, 3, -3, 3), cfl=0.9, bc=(P,) * 4,
                     g1=4.0, g2=1.4, p1=20.0, p2=0.0)
    ref = orc.solve(op, U0, 100000, t_end=6.0, clip=True)
    assert st.t == 6.0 and st.steps == ref["steps"]
    assert np.array_equal(U, ref["U"])


def test_smooth_vortex_converges_at_third_order(hip):
    errs = []
    for n in (64, 128, 256, 512, 1024):
        p, U0 = vortex(hip, n)
        U, st = run_period(hip, p, U0)
        assert st.t == 6.0 and st.err == hip.OK
        e = U[..., 5] - U0[..., 5]
        errs.append((np.sqrt(np.sum(e * e)) / n, np.max(np.abs(e))))
        print(f"[vortex] N={n} steps={st.steps} L2={errs[-1][0]:.7e} Linf={errs[-1][1]:.4e}", flush=True)
    l2 = np.array([e[0] for e in errs])
    orders = np.log2(l2[:-1] / l2[1:])
    print(f"[vortex] L2 orders {np.round(orders, 3)}", flush=True)
    assert np.all(l2[1:] < l2[:-1])
    assert orders[-1] > 2.7 and orders[-2] > 2.4  # spreadsheet: 2.92, 2.99 at the two finest pairs
