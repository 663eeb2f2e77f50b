"""Writes batched_sha256.json: the oracle's final states of the smallest and
the largest instance of the batched set (64 Kapila shock-bubbles, seed 2024,
sides log-uniform on [128, 1024]; SSPRK(4,3) with the reference's dt lag,
400 steps each), as sha256 of the FP64 state bytes with the final time and
step count.  tests/test_gpu_batched_config.py holds the batched run to them
bit for bit.  The oracle needs about 3 minutes on 8 cores.

    python tests/golden/make_batched_hashes.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "hype-parallel_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import oracle_py as orc  # noqa: E402
import workloads  # noqa: E402

OUT = os.path.join(HERE, "batched_sha256.json")
SEED, COUNT = 2024, 64


def main():
    draws = workloads.batched_draws(SEED, COUNT)
    order = sorted(range(COUNT), key=lambda i: draws[i][0])
    data = {"seed": SEED, "count": COUNT, "steps": workloads.BATCHED_STEPS, "instances": {}}
    for i in (order[0], order[-1]):
        p, U0 = workloads.kapila_bubble(*draws[i])
        op = orc.problem(p.system, p.nx, p.ny, N=p.order, prim=bool(p.recons_prim),
                         domain=(p.x_min, p.x_max, p.y_min, p.y_max), cfl=p.cfl, bc=tuple(p.bc),
                         g1=p.gamma1, g2=p.gamma2, p1=p.pi1, p2=p.pi2, quirks=p.quirks)
        r = orc.solve(op, U0, workloads.BATCHED_STEPS, clip=True, threaded=True)
        assert r["rc"] == 0
        h = hashlib.sha256(np.ascontiguousarray(r["U"], dtype=np.float64).tobytes()).hexdigest()
        data["instances"][str(i)] = {"side": p.nx, "steps": r["steps"], "t": r["t"], "sha256": h}
        print(i, data["instances"][str(i)], flush=True)
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
