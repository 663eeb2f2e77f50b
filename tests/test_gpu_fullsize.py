"""The benchmark workloads at full size, solved on the device, against the
oracle's final states (tests/golden/fullsize_sha256.json, written by
tests/golden/make_fullsize_hashes.py): Kapila shock-bubble 1024x512 over 500
SSPRK(4,3) steps, Euler quadrants 256^2 over 200, Baer-Nunziato water-air
2048x1024 over 300.  Bar: the same final time, step count and state bytes."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import workloads  # noqa: E402

pytestmark = pytest.mark.gpu

with open(os.path.join(ROOT, "tests", "golden", "fullsize_sha256.json")) as f:
    GOLDEN = json.load(f)


@pytest.mark.parametrize("wname", sorted(GOLDEN))
def test_full_size_run_is_bitwise_the_oracle(hip, wname):
    rec = GOLDEN[wname]
    p, U0, steps, name = workloads.single(rec["config"])
    assert name == wname and steps == rec["steps"]
    with hip.Context(p) as ctx:
        U, st = ctx.solve(U0, steps, flags=hip.TS_CLIP | hip.TS_GRAPH)
    h = hashlib.sha256(np.ascontiguousarray(U, dtype=np.float64).tobytes()).hexdigest()
    print(f"[fullsize] {wname} steps={st.steps} t={st.t!r} sha256={h}", flush=True)
    assert st.steps == rec["steps"]
    assert st.t == rec["t"]
    assert h == rec["sha256"]
