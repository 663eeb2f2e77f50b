"""Writes fullsize_sha256.json: the oracle's final states of the three
benchmark workloads (Kapila shock-bubble 1024x512 for 500 steps, Euler
quadrants 256^2 for 200 steps, Baer-Nunziato water-air 2048x1024 for 300
steps, SSPRK(4,3) with the reference's dt lag), as sha256 of the FP64 state
bytes with the final time and step count.  tests/test_gpu_fullsize.py holds
the device to them bit for bit.  The oracle needs about 7 minutes (Kapila)
and 26 minutes (Baer-Nunziato) on 8 cores.

    python tests/golden/make_fullsize_hashes.py [workload ...]
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "hype-parallel_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import workloads  # noqa: E402
import oracle_py as orc  # noqa: E402

OUT = os.path.join(HERE, "fullsize_sha256.json")


def oracle_final(name):
    p, U0, steps, wname = workloads.single(name)
    op = orc.problem(p.system, p.nx, p.ny, N=p.order, prim=bool(p.recons_prim),
                     domain=(p.x_min, p.x_max, p.y_min, p.y_max), cfl=p.cfl, bc=tuple(p.bc),
                     g1=p.gamma1, g2=p.gamma2, p1=p.pi1, p2=p.pi2, quirks=p.quirks)
    r = orc.solve(op, U0, steps, clip=True, threaded=True)
    assert r["rc"] == 0
    h = hashlib.sha256(np.ascontiguousarray(r["U"], dtype=np.float64).tobytes()).hexdigest()
    return wname, {"config": name, "steps": r["steps"], "t": r["t"], "sha256": h}


def main():
    names = sys.argv[1:] or ["euler", "kapila", "bn"]
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    for name in names:
        wname, rec = oracle_final(name)
        data[wname] = rec
        print(wname, rec, flush=True)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
