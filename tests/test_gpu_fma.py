"""HYPE_OPT_FMA: the stage kernels built with FMA contraction.  Not bitwise
equal to the uncontracted reference arithmetic, but held to the FP64
tolerances of the reference path: relative 1e-12 per RHS evaluation and
1e-10 after N steps (normwise per component, as in the parity tests),
against the CPU oracle and, at full benchmark size, against the exact
kernels (which are bitwise the oracle)."""
import numpy as np
import pytest

import oracle_py as orc
import states

pytestmark = pytest.mark.gpu

T, R, P = orc.BC_TRANSMISSIVE, orc.BC_REFLECTIVE, orc.BC_PERIODIC
RHS_TOL = 1e-12
STEP_TOL = 1e-10


def rel_err(a, b):
    out = 0.0
    for c in range(a.shape[-1]):
        d = np.max(np.abs(a[..., c] - b[..., c]))
        s = max(np.max(np.abs(b[..., c])), 1e-300)
        out = max(out, d / s)
    return out


# momentum pairs per system: (rho u, rho v) and, for Baer-Nunziato, both phases
MOMENTA = {0: [(1, 2)], 1: [(2, 3)], 2: [(1, 2), (5, 6)]}


def rel_err_vec(a, b, system):
    """rel_err with each momentum pair normalised by the maximum of its
    vector magnitude: a transverse momentum that is still zero physically
    (rounding noise only) is then compared on the flow's momentum scale."""
    scale = [max(np.max(np.abs(b[..., c])), 1e-300) for c in range(b.shape[-1])]
    for (cu, cv) in MOMENTA[system]:
        m = max(np.max(np.hypot(b[..., cu], b[..., cv])), 1e-300)
        scale[cu] = scale[cv] = m
    return max(np.max(np.abs(a[..., c] - b[..., c])) / scale[c] for c in range(b.shape[-1]))


def pair(hip, system, prim, order, nx, ny, bc, quirks=0):
    eos = states.EOS[system]
    dom = (0.0, 1.0, 0.0, ny / nx)
    p = hip.make_problem(system, nx, ny, order=order, prim=prim, domain=dom, cfl=0.5, bc=bc, quirks=quirks,
                         fma=True, **eos)
    op = orc.problem(system, nx, ny, N=order, prim=prim, domain=dom, cfl=0.5, bc=bc, g1=eos["gamma1"],
                     g2=eos["gamma2"], p1=eos["pi1"], p2=eos["pi2"], quirks=quirks)
    return p, op


@pytest.mark.parametrize("system", [0, 1, 2])
@pytest.mark.parametrize("prim", [True, False])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_fma_rhs_within_tolerance(hip, system, prim, order):
    nx, ny = 75, 40
    quirks = orc.QUIRK_BN_P2_IS_PHI if system == 2 else 0
    p, op = pair(hip, system, prim, order, nx, ny, (T, T, R, R), quirks)
    U = states.state(system, nx, ny, seed=1)
    ref = orc.rhs(op, U, want_fallback=True)
    with hip.Context(p) as ctx:
        Rg, dt, mask, rc = ctx.rhs(U, fallback=True, allow_inadmissible=True)
    err = rel_err(Rg, ref["rhs"])
    print(f"[fma rhs] sys={system} prim={prim} N={order} rel={err:.2e}", flush=True)
    assert rc == ref["rc"]
    assert err <= RHS_TOL
    assert abs(dt - ref["dt"]) <= RHS_TOL * ref["dt"]
    assert np.array_equal(mask, ref["fallback"])


@pytest.mark.parametrize("system,prim,order", [(0, True, 2), (1, True, 2), (2, False, 2), (1, True, 3)])
def test_fma_steps_within_tolerance(hip, system, prim, order):
    nx, ny, steps = 80, 44, 12
    p, op = pair(hip, system, prim, order, nx, ny, (T, T, R, R))
    U0 = states.state(system, nx, ny, seed=11)
    ref = orc.solve(op, U0, steps)
    with hip.Context(p) as ctx:
        U, st = ctx.solve(U0, steps)
    err = rel_err(U, ref["U"])
    print(f"[fma steps] sys={system} N={order} rel={err:.2e} t={st.t!r}/{ref['t']!r}", flush=True)
    assert st.steps == ref["steps"] == steps
    # t accumulates dt from the wave-speed maximum: the stiffened-gas pressure
    # (gamma - 1)(E - k) - gamma pi cancels about four digits at pi = 6e8
    assert abs(st.t - ref["t"]) <= STEP_TOL * ref["t"]
    assert err <= STEP_TOL


def test_fma_benchmark_workload_matches_exact_kernels(hip):
    """The Kapila shock-bubble of the benchmark, 1024 x 512, 500 steps: the
    contracted kernels against the exact ones (bitwise the oracle).  The
    shock does not reach the bubble in 500 steps, so rho v is rounding noise
    of the pressure equilibrium across the curved interface; it is compared
    on the momentum scale |(rho u, rho v)|."""
    nx, ny = 1024, 512
    h = 2.0 / nx
    out = {}
    for fma in (False, True):
        p = hip.make_problem(hip.KAPILA, nx, ny, order=2, prim=True, domain=(0.0, 2.0, 0.0, 1.0), cfl=0.5,
                             bc=(hip.BC_TRANSMISSIVE, hip.BC_TRANSMISSIVE, hip.BC_REFLECTIVE, hip.BC_REFLECTIVE),
                             gamma1=1.4, gamma2=1.648, fma=fma)
        U0 = hip.initial_state(p, hip.IC_SHOCK_BUBBLE, [1.22, 0.5, 1.0, 0.5, 0.25, 1e-6, 2.0 * h, 1.4, 0.25463, 1e5])
        with hip.Context(p) as ctx:
            out[fma] = ctx.solve(U0, 500, flags=hip.TS_CLIP | hip.TS_GRAPH)
    (Ue, ste), (Uf, stf) = out[False], out[True]
    err = rel_err_vec(Uf, Ue, 1)
    print(f"[fma bench] rel={err:.2e} per-component={rel_err(Uf, Ue):.2e} max|rho v|={np.max(np.abs(Ue[..., 3])):.2e} "
          f"t={stf.t!r}/{ste.t!r}", flush=True)
    assert stf.steps == ste.steps == 500
    assert abs(stf.t - ste.t) <= 1e-10 * ste.t
    assert err <= STEP_TOL


def test_fma_and_exact_instances_share_a_batch(hip):
    """Both kernel builds in one batch (two graph branches); each instance
    equals its solo run in its own mode bitwise."""
    probs, U0s = [], []
    for fma in (False, True):
        p, _ = pair(hip, 1, True, 2, 64, 40, (T, T, R, R))
        p.options = hip.OPT_FMA if fma else 0
        probs.append(p)
        U0s.append(states.state(1, 64, 40, seed=21))
    with hip.Batch(probs) as b:
        for i, U in enumerate(U0s):
            b.set_state(i, U)
        b.begin(max_steps=np.array([6, 6], dtype=np.int32))
        b.advance(7)
        st = b.sync()
        assert b.tiling()[2] == 2
        got = [b.get_state(i) for i in range(2)]
    for i, p in enumerate(probs):
        with hip.Context(p) as ctx:
            U, s = ctx.solve(U0s[i], 6)
        assert st[i].steps == s.steps == 6
        assert np.array_equal(got[i], U)
    assert not np.array_equal(got[0], got[1])  # the two builds really differ
