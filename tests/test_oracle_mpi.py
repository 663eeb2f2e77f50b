"""The benchmark's MPI-rank CPU baseline (oracle/orc_mpi_baseline.c) against
the one-process oracle (orc_solve): the same priming call and SSP-RK steps on
a grid split in row blocks over 1, 3 and 4 ranks end in the same state, bit
for bit.  The fastest wave sits inside the first rank's rows, so the faces of
the halo rows' outer edges, which also enter each rank's CFL maximum, cannot
set dt.  This checks the halo exchange and the block bookkeeping of the
program whose timing the benchmark reports."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_py as orc
import states

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "lib", "orc_mpi_baseline")


def mpiexec():
    for cand in (shutil.which("mpiexec"), "/opt/conda/bin/mpiexec"):
        if cand and os.path.exists(cand):
            return cand
    return None


def air_state(nx, ny, seed):
    """Air only (alpha_1 = 1 - 1e-6), smooth random perturbations, and a fast
    stream (u + 600 m/s) in rows 3-5, inside the first block of any split
    into at most 4 ranks."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    one = np.ones_like(X)
    r1 = 1.4 + 0.05 * np.sin(2 * np.pi * Y) + 0.005 * rng.standard_normal((ny, nx))
    p = 1.0e5 * (1 + 0.02 * np.sin(2 * np.pi * (X - Y)))
    u = 20.0 * np.sin(2 * np.pi * Y) + 2.0 * rng.standard_normal((ny, nx))
    u[3:6, 8:14] += 600.0
    v = 10.0 * np.cos(2 * np.pi * X)
    return states.kapila_cons(r1, 0.25463 * one, u, v, p, (1 - 1e-6) * one, 1.4, 1.648, 0.0, 0.0)


@pytest.mark.skipif(not os.path.exists(EXE) or mpiexec() is None, reason="no MPI installation here")
@pytest.mark.parametrize("ranks", [1, 3, 4])
def test_mpi_baseline_matches_one_process(ranks, tmp_path):
    nx, ny, steps = 40, 36, 3
    bc = (orc.BC_TRANSMISSIVE, orc.BC_TRANSMISSIVE, orc.BC_REFLECTIVE, orc.BC_REFLECTIVE)
    U0 = air_state(nx, ny, seed=11)
    op = orc.problem(orc.KAPILA, nx, ny, N=2, prim=True, bc=bc, g1=1.4, g2=1.648)
    state = tmp_path / "u0.bin"
    out = tmp_path / "u1.bin"
    np.ascontiguousarray(U0, dtype=np.float64).tofile(state)
    args = [mpiexec(), "-n", str(ranks), EXE, str(state), str(orc.KAPILA), "2", "1", str(nx), str(ny),
            "0.0", "0.0", repr(op.h), "0.5", *[str(b) for b in bc], "1.4", "1.648", "0.0", "0.0", "0",
            "0", str(steps), str(out)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split()[:4] == ["ranks", str(ranks), "steps", str(steps)]
    U = np.fromfile(out, dtype=np.float64).reshape(ny, nx, 6)
    ref = orc.solve(op, U0, steps + 1, clip=False)  # the untimed first step, then the timed ones
    assert ref["rc"] == 0 and ref["steps"] == steps + 1
    assert np.array_equal(U, ref["U"])
