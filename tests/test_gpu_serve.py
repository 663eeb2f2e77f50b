"""Served mode (hype_server_*): requests admitted under an HBM budget and
advanced with continuous batching must finish bitwise equal to the same
solve run alone (hype_ctx; per instance the loop of hype.c:337-384),
whatever else was resident, and the budget must hold at every round."""
import numpy as np
import pytest

import oracle_py as orc
import states

pytestmark = pytest.mark.gpu

T, R, P = orc.BC_TRANSMISSIVE, orc.BC_REFLECTIVE, orc.BC_PERIODIC

# (system, prim, order, nx, ny, bc, seed, steps, t_end)
REQS = [
    (1, True, 2, 96, 40, (T, T, R, R), 1, 14, np.inf),
    (0, True, 2, 64, 64, (T, T, T, T), 2, 1000, 2.0e-2),
    (2, False, 2, 50, 30, (T, T, R, R), 3, 9, np.inf),
    (1, False, 1, 40, 24, (R, R, P, P), 4, 21, np.inf),
    (0, False, 1, 130, 20, (P, P, P, P), 5, 17, np.inf),
    (1, True, 2, 33, 70, (P, P, R, R), 6, 12, np.inf),
    (0, True, 2, 48, 48, (R, R, R, R), 7, 25, np.inf),
    (1, True, 2, 120, 36, (T, T, R, R), 8, 6, np.inf),
    (2, False, 2, 64, 32, (P, P, T, T), 9, 11, np.inf),
]


def problem(hip, system, prim, order, nx, ny, bc, quirks=0):
    eos = states.EOS[system]
    return hip.make_problem(system, nx, ny, order=order, prim=prim, domain=(0.0, 1.0, 0.0, ny / nx), cfl=0.5,
                            bc=bc, quirks=quirks, **eos)


def solo(hip, p, U0, steps, t_end=float("inf")):
    with hip.Context(p) as ctx:
        ctx.set_state(U0)
        ctx.begin(0.0, t_end, steps, hip.TS_CLIP)
        ctx.advance(steps + 1)
        st = ctx.sync(allow_inadmissible=True)
        return ctx.get_state(), st


def state_bytes(p):
    return p.nx * p.ny * {0: 4, 1: 6, 2: 9}[p.system] * 8


@pytest.mark.parametrize("steps_per_round,policy,cap", [
    (0, 0, -1), (3, 0, 0), (3, 0, 6000), (0, 1, 6000), (2, 1, -1),
])
def test_served_results_equal_solo(hip, steps_per_round, policy, cap):
    """cap: cells advanced per round (instances beyond it wait resident);
    policy 0 = arrival order, 1 = least remaining work first."""
    probs = [problem(hip, *r[:6]) for r in REQS]
    U0s = [states.state(r[0], r[3], r[4], seed=r[6]) for r in REQS]
    biggest = max(3 * state_bytes(p) for p in probs)
    budget = 2 * biggest + 4096  # a few instances at a time: queueing and slot reuse
    ids = []
    with hip.Server(hbm_budget_bytes=budget, max_active=4, steps_per_round=steps_per_round,
                    max_active_cells=cap, policy=policy) as srv:
        peak_bytes, peak_active = 0, 0
        for k, (p, U0, r) in enumerate(zip(probs, U0s, REQS)):
            ids.append(srv.submit(p, U0, t_end=r[8], max_steps=r[7]))
            if k % 3 == 2:  # arrivals interleaved with rounds
                a, q = srv.poll()
                peak_active = max(peak_active, a)
                peak_bytes = max(peak_bytes, srv.bytes_in_use())
        while True:
            a, q = srv.poll()
            peak_active = max(peak_active, a)
            peak_bytes = max(peak_bytes, srv.bytes_in_use())
            if a == 0 and q == 0:
                break
        assert srv.bytes_in_use() == 0
        results = [srv.result(i) for i in ids]
    print(f"[serve] budget={budget} peak_bytes={peak_bytes} peak_active={peak_active}", flush=True)
    assert 0 < peak_bytes <= budget and 1 <= peak_active <= 4
    for k, (p, U0, r) in enumerate(zip(probs, U0s, REQS)):
        U, info = results[k]
        ref, rst = solo(hip, p, U0, r[7], t_end=float(r[8]))
        assert info.state == hip.REQ_DONE and info.status.err == hip.OK
        assert info.status.steps == rst.steps and info.status.t == rst.t, f"request {k}"
        assert info.t_submit <= info.t_start <= info.t_done
        assert np.array_equal(U, ref), f"request {k} differs from its solo run"


def test_served_matches_oracle(hip):
    r = REQS[0]
    p = problem(hip, *r[:6])
    U0 = states.state(r[0], r[3], r[4], seed=r[6])
    eos = states.EOS[r[0]]
    op = orc.problem(r[0], r[3], r[4], N=r[2], prim=r[1], domain=(0.0, 1.0, 0.0, r[4] / r[3]), cfl=0.5, bc=r[5],
                     g1=eos["gamma1"], g2=eos["gamma2"], p1=eos["pi1"], p2=eos["pi2"])
    ref = orc.solve(op, U0, r[7])
    with hip.Server(hbm_budget_bytes=1 << 24) as srv:
        i = srv.submit(p, U0, max_steps=r[7])
        j = srv.submit(problem(hip, *REQS[2][:6]), states.state(2, 50, 30, seed=3), max_steps=5)
        srv.drain()
        U, info = srv.result(i)
        assert srv.info(j).state == hip.REQ_DONE
    assert info.status.steps == ref["steps"] and info.status.t == ref["t"]
    assert np.array_equal(U, ref["U"])


def test_request_larger_than_budget_is_refused(hip):
    p = problem(hip, 1, True, 2, 96, 40, (T, T, R, R))
    with hip.Server(hbm_budget_bytes=3 * state_bytes(p) - 1024) as srv:
        with pytest.raises(hip.HypeError) as e:
            srv.submit(p, states.state(1, 96, 40, seed=1), max_steps=3)
        assert e.value.code == hip.E_NOMEM


def test_failing_request_is_reported_and_isolated(hip):
    bad = problem(hip, 2, True, 2, 80, 44, (T, T, R, R), quirks=hip.QUIRK_BN_P2_IS_PHI)
    good = problem(hip, 0, True, 2, 64, 64, (T, T, T, T))
    Ub, Ug = states.state(2, 80, 44, seed=11), states.state(0, 64, 64, seed=2)
    with hip.Server(hbm_budget_bytes=1 << 24) as srv:
        ib = srv.submit(bad, Ub, max_steps=8)
        ig = srv.submit(good, Ug, max_steps=8)
        srv.drain()
        (Ubo, infb), (Ugo, infg) = srv.result(ib), srv.result(ig)
    _, rst = solo(hip, bad, Ub, 8)
    assert infb.status.err == hip.E_INADMISSIBLE and rst.err == hip.E_INADMISSIBLE
    assert infb.status.steps == rst.steps
    ref, _ = solo(hip, good, Ug, 8)
    assert infg.status.err == hip.OK and np.array_equal(Ugo, ref)


def test_burst_of_fresh_tiles_beyond_running_table(hip):
    """One round admits more fresh tiles (priming table) than the running
    subset holds and than the initial table capacity (1024 tiles): the
    stepping table must stay intact while the priming table grows.  Five
    Euler 256^2 requests with 2-row tiles are 640 tiles each; the cell cap
    lets one instance run at a time."""
    p = problem(hip, 0, True, 2, 256, 256, (T, T, T, T))
    U0s = [states.state(0, 256, 256, seed=40 + k) for k in range(5)]
    steps = [5, 7, 4, 6, 3]
    with hip.Server(hbm_budget_bytes=1 << 30, max_active=8, steps_per_round=2, tile_rows=2,
                    max_active_cells=70000, policy=0) as srv:
        ids = [srv.submit(p, U, max_steps=n) for U, n in zip(U0s, steps)]
        rounds = 0
        while True:
            a, q = srv.poll()
            rounds += 1
            if a == 0 and q == 0:
                break
        results = [srv.result(i) for i in ids]
    print(f"[serve-burst] rounds={rounds}", flush=True)
    for k, (U0, n) in enumerate(zip(U0s, steps)):
        U, info = results[k]
        ref, rst = solo(hip, p, U0, n)
        assert info.status.err == hip.OK and info.status.steps == rst.steps == n
        assert np.array_equal(U, ref), f"request {k} differs from its solo run"
