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
