"""The device's SSP-RK stage table (hyped_stage_table, the StageDesc the
stage kernel follows) replayed on the CPU: the buffer algebra of every
stage, with a test RHS in place of the kernel's L(U), must give the oracle's
TSSSP step (orc_ssp_step, PETSc's RKS2 / RKS3 / RK104 as in hype.c:362-382)
bit for bit.  Needs the library (host function), not a GPU."""
import ctypes as C

import numpy as np
import pytest

import hype2d as H
from test_oracle_time import oracle_step


def stage_table(order):
    L = H.lib()
    L.hyped_stage_table.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    rows, s = [], 0
    while True:
        ib = (C.c_int * 5)()
        cf = (C.c_double * 3)()
        n = L.hyped_stage_table(order, s, ib, cf)
        if n < 0:
            break
        rows.append((list(ib), list(cf)))
        s += 1
        if s == n:
            break
    return rows


def replay(order, f, u0, dt):
    """Buffers 0 = U, 1 = A, 2 = B, updated exactly as the stage kernel does."""
    buf = [np.array(u0, dtype=np.float64), np.zeros(len(u0)), np.zeros(len(u0))]
    for (i_in, i_out, xb, zb, op), (a, g, bden) in stage_table(order):
        uin = buf[i_in].copy()
        rhs = f(uin)
        cdt = dt / bden
        if op == 0:
            buf[i_out] = uin + cdt * rhs
        elif op == 2:
            w0 = uin + cdt * rhs
            w1 = a * buf[i_out] + g * w0
            buf[i_out] = w1
            buf[1] = 15.0 * w1 + -5.0 * w0
        else:
            buf[i_out] = a * buf[xb] + cdt * rhs + g * buf[zb]
    return buf[0]


@pytest.mark.parametrize("order,nstages", [(1, 2), (2, 4), (3, 10)])
def test_stage_table_replays_the_oracle_step(order, nstages):
    assert len(stage_table(order)) == nstages
    rng = np.random.default_rng(order)
    u0 = rng.uniform(-1.0, 1.0, 7)
    f = lambda u: -u * u + 0.1 * u[::-1] + 0.3
    dt = 0.037
    got = replay(order, f, u0, dt)
    exp, _, _ = oracle_step(order, lambda t, u: f(u), 0.0, u0, dt)
    assert np.array_equal(got, exp)
