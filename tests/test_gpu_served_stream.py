"""The served configuration's request stream, replayed through hype_server:
the first requests of the seeded Poisson stream the benchmark uses (20
requests/s, mix 50 % Euler / 35 % Kapila / 15 % Baer-Nunziato, sides
log-uniform on [64, 2048], 100-2000 requested steps turned into an end time),
submitted at their arrival times while earlier ones run.  Bar: every result
is bitwise the same request solved alone (hype_ctx), with the same step
count and status, and the HBM budget (512 MiB, less than the prefix needs at
once) is never exceeded."""
import time

import numpy as np
import pytest

import oracle_py as orc
import workloads

pytestmark = pytest.mark.gpu

SEED, RATE, PREFIX = 2024, 20.0, 16


def solo(hip, p, U0, t_end, cap):
    with hip.Context(p) as ctx:
        ctx.set_state(U0)
        ctx.begin(0.0, t_end, cap, hip.TS_CLIP | hip.TS_GRAPH)
        ctx.advance(cap // 2)
        st = ctx.sync(allow_inadmissible=True)
        while not st.done:
            ctx.advance(32)
            st = ctx.sync(allow_inadmissible=True)
        return ctx.get_state(), st


@pytest.mark.parametrize("policy", ["shortest", "fifo"])
def test_served_stream_prefix_equals_solo(hip, policy):
    reqs = workloads.served_trace(np.random.default_rng(SEED), PREFIX, RATE)
    pol = hip.SERVE_SHORTEST if policy == "shortest" else hip.SERVE_FIFO
    budget = 512 << 20  # below the prefix's total: admission waits for HBM
    print(f"[served-stream] {policy}: {PREFIX} requests, "
          f"{sum(r[1].nx * r[1].ny * r[4] // 2 for r in reqs) / 1e9:.2f} G requested cell-steps", flush=True)
    with hip.Server(hbm_budget_bytes=budget, policy=pol) as srv:
        base = srv.now()
        ids, k, peak = [], 0, 0
        while k < len(reqs) or srv.poll() != (0, 0):
            now = srv.now() - base
            while k < len(reqs) and reqs[k][0] <= now:
                _, p, U0, t_end, cap = reqs[k]
                ids.append(srv.submit(p, U0, t_end=t_end, max_steps=cap))
                k += 1
            peak = max(peak, srv.bytes_in_use())
            if k < len(reqs) and srv.poll() == (0, 0):
                time.sleep(0.005)
        res = [srv.result(i) for i in ids]
    assert peak <= budget
    for j, ((U, inf), (_, p, U0, t_end, cap)) in enumerate(zip(res, reqs)):
        Us, st = solo(hip, p, U0, t_end, cap)
        print(f"[served-stream] request {j}: system {p.system} side {p.nx} steps {st.steps} err {st.err}",
              flush=True)
        assert inf.status.steps == st.steps and inf.status.err == st.err, f"request {j}: status differs"
        assert inf.status.t == st.t, f"request {j}: final time differs"
        assert np.array_equal(U, Us, equal_nan=True), f"request {j}: state differs from its solo run"


# requests of the stream that end on an inadmissible state; 253 is the first,
# a 65 x 65 Baer-Nunziato water-air interface whose gas pressure turns
# negative at step 904 (8 of the first 1024 requests do, all Baer-Nunziato)
INADMISSIBLE = 253
NEIGHBOURS = (249, 250, 252, 254, 257)  # small requests that arrive around it


def test_served_stream_inadmissible_request(hip):
    """The stream's first inadmissible request, served among its small
    neighbours, stops at the same step with the same state as its solo run;
    the solo run stops where the oracle's solve loop does (step, offending
    face and state), and the neighbours finish unaffected."""
    draws = workloads.served_draws(np.random.default_rng(SEED), INADMISSIBLE + 5, RATE)
    ks = sorted(NEIGHBOURS + (INADMISSIBLE,))
    reqs = {k: workloads.served_request(*draws[k][1:]) for k in ks}
    with hip.Server(policy=hip.SERVE_SHORTEST) as srv:
        ids = {k: srv.submit(p, U0, t_end=t_end, max_steps=cap) for k, (p, U0, t_end, cap) in reqs.items()}
        srv.drain()
        res = {k: srv.result(i) for k, i in ids.items()}
    for k in ks:
        p, U0, t_end, cap = reqs[k]
        Us, st = solo(hip, p, U0, t_end, cap)
        U, inf = res[k]
        print(f"[served-stream] request {k}: system {p.system} side {p.nx} steps {st.steps} err {st.err}",
              flush=True)
        assert inf.status.steps == st.steps and inf.status.err == st.err, f"request {k}: status differs"
        assert np.array_equal(U, Us, equal_nan=True), f"request {k}: state differs from its solo run"
        assert (st.err != 0) == (k == INADMISSIBLE), f"request {k}: unexpected status {st.err}"
    p, U0, t_end, cap = reqs[INADMISSIBLE]
    op = orc.problem(p.system, p.nx, p.ny, N=p.order, prim=bool(p.recons_prim),
                     domain=(p.x_min, p.x_max, p.y_min, p.y_max), cfl=p.cfl, bc=tuple(p.bc), g1=p.gamma1,
                     g2=p.gamma2, p1=p.pi1, p2=p.pi2, quirks=p.quirks)
    ref = orc.solve(op, U0, cap, t_end=t_end, clip=True, threaded=True)
    with hip.Context(p) as ctx:
        ctx.set_state(U0)
        ctx.begin(0.0, t_end, cap, hip.TS_CLIP | hip.TS_GRAPH)
        ctx.advance(cap)
        st = ctx.sync(allow_inadmissible=True)
        msg, x, y = ctx.last_error()
        U = ctx.get_state()
    print(f"[served-stream] request {INADMISSIBLE}: {msg!r}; oracle steps {ref['steps']} kind {ref['err'].kind} "
          f"at ({ref['err'].x}, {ref['err'].y})", flush=True)
    assert ref["rc"] == hip.E_INADMISSIBLE and st.err == hip.E_INADMISSIBLE
    assert st.steps == ref["steps"] == 904
    assert (x, y) == (ref["err"].x, ref["err"].y)
    assert np.array_equal(U, ref["U"], equal_nan=True)
    assert np.array_equal(res[INADMISSIBLE][0], U, equal_nan=True)
