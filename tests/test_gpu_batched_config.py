"""The batched configuration exactly as the benchmark times it: 64 Kapila
shock-bubbles (seed 2024, sides log-uniform on [128, 1024], Mach on
[1.1, 2.0], radius on [0.1, 0.3], centre jittered by 0.1), 400 SSPRK(4,3)
steps each, advanced together by hype_batch with the tile height of the
tile-shape model.  Bar: every instance is bitwise the same instance solved
alone (hype_ctx), and the smallest and largest instances are bitwise the
oracle's solve loop (tests/golden/batched_sha256.json, written by
tests/golden/make_batched_hashes.py)."""
import hashlib
import json
import os

import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "batched_sha256.json")) as f:
    GOLDEN = json.load(f)


def sha(U):
    return hashlib.sha256(np.ascontiguousarray(U, dtype=np.float64).tobytes()).hexdigest()


def test_batched_config_equals_solo_and_oracle(hip):
    seed, count, steps = GOLDEN["seed"], GOLDEN["count"], GOLDEN["steps"]
    assert steps == workloads.BATCHED_STEPS
    insts = workloads.batched(seed, count)
    print(f"[batched-config] {count} instances, {sum(p.nx * p.ny for p, _ in insts)} cells, {steps} steps",
          flush=True)
    with hip.Batch([p for p, _ in insts]) as b:
        for i, (_, U0) in enumerate(insts):
            b.set_state(i, U0)
        b.begin(max_steps=steps)
        b.advance(steps + 2)  # steps past the budget are no-ops
        st = b.sync()
        out = [b.get_state(i) for i in range(count)]
        rows, tiles, groups = b.tiling()
    print(f"[batched-config] tile rows {rows}, {tiles} tiles, {groups} variant(s)", flush=True)
    assert all(s.steps == steps and s.err == 0 for s in st)
    for key, rec in GOLDEN["instances"].items():
        i = int(key)
        assert insts[i][0].nx == rec["side"]
        assert st[i].steps == rec["steps"] and st[i].t == rec["t"]
        assert sha(out[i]) == rec["sha256"], f"instance {i} differs from the oracle"
    for i, (p, U0) in enumerate(insts):
        with hip.Context(p) as ctx:
            U, s = ctx.solve(U0, steps, flags=hip.TS_CLIP | hip.TS_GRAPH)
        assert s.steps == steps and s.t == st[i].t, f"instance {i}: time differs from its solo run"
        assert np.array_equal(U, out[i]), f"instance {i} differs from its solo run"
