"""Batched mode (hype_batch_*): every instance of a mixed batch must end
bitwise equal to the same instance solved alone (hype_ctx), and to the CPU
oracle's solve loop (hype.c:337-384 per instance), whatever the mix, tile
height or the failure of another instance."""
import numpy as np
import pytest

import oracle_py as orc
import states

pytestmark = pytest.mark.gpu

T, R, P = orc.BC_TRANSMISSIVE, orc.BC_REFLECTIVE, orc.BC_PERIODIC

# (system, prim, order, nx, ny, bc, seed)
MIX = [
    (1, True, 2, 96, 40, (T, T, R, R), 1),
    (1, True, 2, 33, 70, (P, P, R, R), 2),
    (0, True, 2, 64, 64, (T, T, T, T), 3),
    (2, False, 2, 50, 30, (T, T, R, R), 4),
    (1, False, 1, 40, 24, (R, R, P, P), 5),
    (0, False, 1, 130, 20, (P, P, P, P), 6),
    (1, True, 2, 200, 12, (T, R, T, R), 7),
]


def problem(hip, system, prim, order, nx, ny, bc, quirks=0, cfl=0.5):
    eos = states.EOS[system]
    return hip.make_problem(system, nx, ny, order=order, prim=prim, domain=(0.0, 1.0, 0.0, ny / nx), cfl=cfl,
                            bc=bc, quirks=quirks, **eos)


def solo(hip, p, U0, steps, t_end=float("inf")):
    with hip.Context(p) as ctx:
        ctx.set_state(U0)
        ctx.begin(0.0, t_end, steps, hip.TS_CLIP)
        ctx.advance(steps + 1)
        st = ctx.sync(allow_inadmissible=True)
        return ctx.get_state(), st


def run_batch(hip, probs, U0s, steps, t_end=None, tile_rows=0, extra=3):
    with hip.Batch(probs, tile_rows=tile_rows) as b:
        for i, U in enumerate(U0s):
            b.set_state(i, U)
        b.begin(t_end=t_end, max_steps=steps)
        b.advance(max(np.max(steps), 1) + extra)
        st = b.sync(allow_inadmissible=True)
        out = [b.get_state(i) for i in range(len(probs))]
        tiling = b.tiling()
    return out, st, tiling


@pytest.mark.parametrize("tile_rows", [0, 5])
def test_mixed_batch_equals_solo(hip, tile_rows):
    probs, U0s, steps = [], [], []
    for k, (s, pr, o, nx, ny, bc, seed) in enumerate(MIX):
        probs.append(problem(hip, s, pr, o, nx, ny, bc))
        U0s.append(states.state(s, nx, ny, seed=seed))
        steps.append(6 + 2 * k)  # every instance its own step budget
    steps = np.array(steps, dtype=np.int32)
    out, st, tiling = run_batch(hip, probs, U0s, steps, tile_rows=tile_rows)
    print(f"[batch] rows/tiles/groups={tiling}", flush=True)
    assert tiling[2] == 5  # five kernel variants in the mix
    for i, p in enumerate(probs):
        ref, rst = solo(hip, p, U0s[i], int(steps[i]))
        assert st[i].steps == rst.steps == steps[i]
        assert st[i].t == rst.t and st[i].dt == rst.dt and st[i].done
        assert np.array_equal(out[i], ref), f"instance {i} differs from its solo run"


def test_batch_matches_oracle(hip):
    cases = [(1, True, 2, 48, 30, (T, T, R, R), 21), (2, False, 2, 40, 24, (P, P, R, R), 22),
             (0, True, 1, 36, 36, (R, R, R, R), 23), (1, True, 3, 44, 28, (T, T, R, R), 24)]
    probs, U0s = [], []
    for s, pr, o, nx, ny, bc, seed in cases:
        probs.append(problem(hip, s, pr, o, nx, ny, bc))
        U0s.append(states.state(s, nx, ny, seed=seed))
    steps = np.array([7, 9, 11, 5], dtype=np.int32)
    out, st, _ = run_batch(hip, probs, U0s, steps)
    for i, (s, pr, o, nx, ny, bc, seed) in enumerate(cases):
        eos = states.EOS[s]
        op = orc.problem(s, nx, ny, N=o, prim=pr, domain=(0.0, 1.0, 0.0, ny / nx), cfl=0.5, bc=bc,
                         g1=eos["gamma1"], g2=eos["gamma2"], p1=eos["pi1"], p2=eos["pi2"])
        ref = orc.solve(op, U0s[i], int(steps[i]), clip=True)
        assert st[i].steps == ref["steps"] and st[i].t == ref["t"]
        assert np.array_equal(out[i], ref["U"])


def test_batch_per_instance_end_time(hip):
    """Each instance stops at its own t_end (MATCHSTEP clip) while the others run on."""
    probs = [problem(hip, 0, True, 2, 64, 64, (T,) * 4), problem(hip, 0, True, 2, 64, 64, (T,) * 4),
             problem(hip, 1, True, 2, 80, 40, (T, T, R, R))]
    U0s = [states.state(0, 64, 64, seed=31), states.state(0, 64, 64, seed=32), states.state(1, 80, 40, seed=33)]
    t_end = np.array([1.0e-2, 2.5e-2, np.inf])
    steps = np.array([1000, 1000, 9], dtype=np.int32)
    with hip.Batch(probs) as b:
        for i, U in enumerate(U0s):
            b.set_state(i, U)
        b.begin(t_end=t_end, max_steps=steps)
        for _ in range(200):
            b.advance(4)
            st = b.sync()
            if all(s.done for s in st):
                break
        out = [b.get_state(i) for i in range(3)]
    for i in range(3):
        ref, rst = solo(hip, probs[i], U0s[i], int(steps[i]), t_end=float(t_end[i]))
        assert st[i].steps == rst.steps and st[i].t == rst.t
        if np.isfinite(t_end[i]):
            assert st[i].t == t_end[i]
        assert np.array_equal(out[i], ref)


def test_failing_instance_does_not_disturb_others(hip):
    """The primitive BN flux quirk (baer_nunziato.c:271) drives this state
    inadmissible in the first step; the instance stops with the reference's
    message, the rest of the batch completes unchanged."""
    bad = problem(hip, 2, True, 2, 80, 44, (T, T, R, R), quirks=hip.QUIRK_BN_P2_IS_PHI)
    good = problem(hip, 1, True, 2, 56, 32, (T, T, R, R))
    Ub, Ug = states.state(2, 80, 44, seed=11), states.state(1, 56, 32, seed=42)
    out, st, _ = run_batch(hip, [good, bad], [Ug, Ub], np.array([8, 8], dtype=np.int32))
    assert st[1].err == hip.E_INADMISSIBLE and st[1].done and st[1].steps < 8
    assert st[0].err == hip.OK and st[0].steps == 8
    ref, _ = solo(hip, good, Ug, 8)
    assert np.array_equal(out[0], ref)
    _, rst = solo(hip, bad, Ub, 8)
    assert rst.err == hip.E_INADMISSIBLE and rst.steps == st[1].steps


def test_autotune_keeps_results(hip):
    probs = [problem(hip, 1, True, 2, 72, 44, (T, T, R, R)), problem(hip, 0, True, 2, 50, 50, (P,) * 4)]
    U0s = [states.state(1, 72, 44, seed=51), states.state(0, 50, 50, seed=52)]
    with hip.Batch(probs) as b:
        for i, U in enumerate(U0s):  # the tuner times steps of the current states
            b.set_state(i, U)
        rows, ms = b.autotune([3, 8, 16], steps=2)  # the Euler instance: both layouts tried
        assert rows in (3, 8, 16) and ms > 0
        assert b.tiling()[0] == rows and b.layout() in (-1, 1)
        for i, U in enumerate(U0s):
            b.set_state(i, U)
        b.begin(max_steps=np.array([5, 5], dtype=np.int32))
        b.advance(6)
        b.sync()
        out = [b.get_state(i) for i in range(2)]
    for i in range(2):
        ref, _ = solo(hip, probs[i], U0s[i], 5)
        assert np.array_equal(out[i], ref)


@pytest.mark.parametrize("tile_rows", [0, 4, 7])
def test_batch_with_forced_tile_layouts(hip, tile_rows):
    """Euler instances forced onto one-wave tiles and onto split tiles run as
    two variants of one batch (two graph branches, persistent queues of one-
    and two-wave workgroups); each ends bitwise equal to its solo run and to
    the oracle."""
    cases = [(0, True, 2, 70, 44, (T, T, R, R), 31, 0x2), (0, True, 2, 70, 44, (T, T, R, R), 31, 0x4),
             (0, False, 3, 45, 30, (P, P, R, R), 32, 0x4), (0, True, 1, 90, 17, (R, T, P, P), 33, 0x2)]
    probs, U0s, steps = [], [], []
    for k, (s, pr, o, nx, ny, bc, seed, opt) in enumerate(cases):
        p = problem(hip, s, pr, o, nx, ny, bc)
        p.options = opt
        probs.append(p)
        U0s.append(states.state(s, nx, ny, seed=seed))
        steps.append(6 + max(k - 1, 0))  # the first two: the same problem, the same budget
    out, st, tiling = run_batch(hip, probs, U0s, np.array(steps, dtype=np.int32), tile_rows=tile_rows)
    print(f"[batch layouts] tile_rows={tile_rows} tiling={tiling}", flush=True)
    assert tiling[2] == 4  # (prim, order, layout) differ: four variants
    for i, p in enumerate(probs):
        U, s1 = solo(hip, p, U0s[i], steps[i])
        assert st[i].steps == s1.steps == steps[i]
        assert st[i].t == s1.t
        assert np.array_equal(out[i], U), f"instance {i}"
    # the two layouts of the same problem agree bitwise too
    assert np.array_equal(out[0], out[1])
    s, pr, o, nx, ny, bc, seed, _ = cases[0]
    eos = states.EOS[s]
    op = orc.problem(s, nx, ny, N=o, prim=pr, domain=(0.0, 1.0, 0.0, ny / nx), cfl=0.5, bc=bc, g1=eos["gamma1"],
                     g2=eos["gamma2"], p1=eos["pi1"], p2=eos["pi2"])
    ref = orc.solve(op, U0s[0], steps[0])
    assert np.array_equal(out[0], ref["U"])


def test_advance_before_begin_is_refused(hip):
    """The time states are written by begin; stepping before it is an error
    (and the zeroed states make any stage launch a no-op)."""
    p = problem(hip, 1, True, 2, 48, 30, (T, T, R, R))
    with hip.Batch([p, p]) as b:
        with pytest.raises(hip.HypeError) as e:
            b.advance(2)
        assert e.value.code == hip.E_INVALID
        U0 = states.state(1, 48, 30, seed=3)
        b.set_state(0, U0)
        b.set_state(1, U0)
        b.begin(max_steps=np.array([3, 3], dtype=np.int32))
        b.advance(4)
        st = b.sync()
        assert st[0].steps == st[1].steps == 3
        assert np.array_equal(b.get_state(0), b.get_state(1))
