"""Device RHS against the CPU oracle on identical states (one evaluation).

Bar: relative error <= 1e-12 per evaluation (normwise per component).  The
device path keeps the reference's operation order without FMA contraction,
so the tests also report (and require) bitwise equality.
"""
import numpy as np
import pytest

import oracle_py as orc
import states

pytestmark = pytest.mark.gpu

T, R, P = orc.BC_TRANSMISSIVE, orc.BC_REFLECTIVE, orc.BC_PERIODIC
BCS = {
    "trans": (T, T, T, T),
    "refl_tb": (T, T, R, R),
    "refl_all": (R, R, R, R),
    "periodic": (P, P, P, P),
    "per_x_refl_y": (P, P, R, R),
    "mixed": (R, T, P, P),
}
REL_TOL = 1e-12


def rel_err(a, b):
    """max over components of ||a_c - b_c||_inf / max(||b_c||_inf, tiny)."""
    nv = a.shape[-1]
    out = 0.0
    for c in range(nv):
        d = np.max(np.abs(a[..., c] - b[..., c]))
        s = max(np.max(np.abs(b[..., c])), 1e-300)
        out = max(out, d / s)
    return out


def check_rhs(hip, system, prim, order, bc, nx, ny, seed, hrows=None, grid=0, quirks=0, sharp=True,
              rough=False, options=0):
    eos = states.EOS[system]
    U = states.rough_state(system, nx, ny, seed) if rough else states.state(system, nx, ny, seed=seed, sharp=sharp)
    dom = (0.0, 1.0, 0.0, ny / nx)
    p = hip.make_problem(system, nx, ny, order=order, prim=prim, domain=dom, cfl=0.5, bc=bc,
                         quirks=quirks, **eos)
    p.options = options
    op = orc.problem(system, nx, ny, N=order, prim=prim, domain=dom, cfl=0.5, bc=bc,
                     g1=eos["gamma1"], g2=eos["gamma2"], p1=eos["pi1"], p2=eos["pi2"], quirks=quirks)
    ref = orc.rhs(op, U, want_fallback=True)
    with hip.Context(p) as ctx:
        if hrows is not None:
            ctx.set_tiling(hrows, grid)
        R, dt, mask, rc = ctx.rhs(U, fallback=True, allow_inadmissible=True)
    print(f"[rhs] sys={system} prim={prim} N={order} bc={bc} {nx}x{ny} seed={seed} hrows={hrows} "
          f"rc={rc}/{ref['rc']} fallback_cells={int(mask.sum())}", flush=True)
    assert rc == ref["rc"]
    err = rel_err(R, ref["rhs"])
    assert err <= REL_TOL, f"relative error {err:.3e}"
    assert np.array_equal(R, ref["rhs"]), "not bitwise equal"
    assert dt == ref["dt"]
    assert np.array_equal(mask, ref["fallback"])
    return mask


@pytest.mark.parametrize("system", [0, 1, 2])
@pytest.mark.parametrize("prim", [True, False])
@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_rhs_systems_paths_orders(hip, system, prim, order):
    quirks = orc.QUIRK_BN_P2_IS_PHI if system == 2 else 0
    check_rhs(hip, system, prim, order, BCS["refl_tb"], 75, 40, seed=1, quirks=quirks)


@pytest.mark.parametrize("bcname", list(BCS))
@pytest.mark.parametrize("system", [0, 1, 2])
def test_rhs_boundaries(hip, bcname, system):
    check_rhs(hip, system, True, 2, BCS[bcname], 66, 35, seed=2)


@pytest.mark.parametrize("hrows,grid", [(1, 0), (3, 0), (7, 0), (64, 0), (5, 16), (2, 3)])
def test_rhs_tilings(hip, hrows, grid):
    check_rhs(hip, 1, True, 2, BCS["refl_tb"], 130, 29, seed=3, hrows=hrows, grid=grid)


# tile layouts: one wave per tile, or split at the middle row over two waves
# (tile heights 1 to 5 cover the unsplit one-row tile, even and odd splits;
# grid > 0 runs the persistent queue with two-wave workgroups)
LAYOUTS = {"one": 0x2, "split": 0x4}


@pytest.mark.parametrize("layout", ["one", "split"])
@pytest.mark.parametrize("hrows,grid", [(1, 0), (2, 0), (3, 0), (4, 0), (5, 0), (64, 0), (3, 7)])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_rhs_euler_tile_layouts(hip, layout, hrows, grid, order):
    check_rhs(hip, 0, True, order, BCS["mixed"], 130, 29, seed=5, hrows=hrows, grid=grid,
              options=LAYOUTS[layout])


@pytest.mark.parametrize("system,layout", [(1, "one"), (2, "split")])
def test_rhs_layout_not_built_is_unsupported(hip, system, layout):
    eos = states.EOS[system]
    p = hip.make_problem(system, 40, 20, order=2, prim=True, bc=BCS["trans"], **eos)
    p.options = LAYOUTS[layout]
    with pytest.raises(hip.HypeError) as e:
        hip.Context(p)
    assert e.value.code == hip.E_UNSUPPORTED


@pytest.mark.parametrize("nx,ny", [(3, 3), (4, 7), (61, 5), (62, 3), (63, 4), (124, 9), (200, 3)])
def test_rhs_ragged_sizes(hip, nx, ny):
    check_rhs(hip, 1, True, 2, BCS["mixed"], nx, ny, seed=4)


@pytest.mark.parametrize("system", [0, 1, 2])
@pytest.mark.parametrize("prim", [True, False])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_rhs_fallback_exercised(hip, system, prim, order):
    quirks = orc.QUIRK_BN_P2_IS_PHI if system == 2 else 0
    mask = check_rhs(hip, system, prim, order, BCS["refl_tb"], 96, 80, seed=5, quirks=quirks, rough=True)
    if order >= 2:  # minmod slopes (N=1) never leave the admissible set
        assert mask.sum() > 0, "state should trigger the admissibility fallback somewhere"


def test_rhs_bn_quirk_off(hip):
    check_rhs(hip, 2, True, 2, BCS["trans"], 70, 33, seed=6, quirks=0)


def test_rhs_large(hip):
    check_rhs(hip, 1, True, 2, BCS["refl_tb"], 512, 256, seed=7)


KIND_TEXT = {1: "Negative density", 2: "Negative pressure", 3: "Negative square speed of sound",
             4: "Negative density (solid)", 5: "Negative density (gas)", 6: "Negative pressure (solid)",
             7: "Negative pressure (gas)"}


@pytest.mark.parametrize("system,prim", [(0, True), (0, False), (1, True), (2, True), (2, False)])
def test_rhs_inadmissible_reports_first_face(hip, system, prim):
    """A cell with negative pressure: MPI_Abort in the reference becomes
    HYPE_E_INADMISSIBLE with the reference's message and the lowest face."""
    nx, ny = 40, 30
    eos = states.EOS[system]
    U = states.state(system, nx, ny, seed=8, sharp=False).copy()
    e_idx = {0: 3, 1: 4, 2: 7}[system]
    U[17, 23, e_idx] *= -0.5 if system != 2 else -1.0  # energy -> negative pressure
    bc = BCS["trans"]
    dom = (0.0, 1.0, 0.0, ny / nx)
    quirks = 0
    p = hip.make_problem(system, nx, ny, order=2, prim=prim, domain=dom, bc=bc, quirks=quirks, **eos)
    op = orc.problem(system, nx, ny, N=2, prim=prim, domain=dom, bc=bc, g1=eos["gamma1"], g2=eos["gamma2"],
                     p1=eos["pi1"], p2=eos["pi2"], quirks=quirks)
    ref = orc.rhs(op, U)
    assert ref["rc"] == hip.E_INADMISSIBLE
    with hip.Context(p) as ctx:
        R, dt, rc = ctx.rhs(U, allow_inadmissible=True)
        msg, x, y = ctx.last_error()
    print(f"[inadmissible] {msg!r} oracle kind={ref['err'].kind} x={ref['err'].x} y={ref['err'].y}", flush=True)
    assert rc == hip.E_INADMISSIBLE
    assert msg.startswith(KIND_TEXT[ref["err"].kind])
    assert (x, y) == (ref["err"].x, ref["err"].y)
    assert np.array_equal(R, ref["rhs"], equal_nan=True)


def _rhs_pair(hip, system, prim, U, bc=(T, T, T, T)):
    ny, nx = U.shape[:2]
    eos = states.EOS[system]
    dom = (0.0, 1.0, 0.0, ny / nx)
    p = hip.make_problem(system, nx, ny, order=2, prim=prim, domain=dom, cfl=0.5, bc=bc, **eos)
    op = orc.problem(system, nx, ny, N=2, prim=prim, domain=dom, cfl=0.5, bc=bc,
                     g1=eos["gamma1"], g2=eos["gamma2"], p1=eos["pi1"], p2=eos["pi2"])
    ref = orc.rhs(op, U)
    with hip.Context(p) as ctx:
        R, dt, rc = ctx.rhs(U, allow_inadmissible=True)
    assert rc == ref["rc"]
    assert np.array_equal(R, ref["rhs"]) and dt == ref["dt"]


@pytest.mark.parametrize("prim", [True, False])
def test_weno_division_range_fallback(hip, prim):
    """The WENO weights use a shared-reciprocal division that is exact only
    for divisors up to 2^380; larger smoothness indicators (pressure jumps of
    1e20 here, IS ~ 1e40) take the plain-division path lane by lane.  Both
    paths, mixed within one wave, must match the oracle bit for bit."""
    nx, ny = 70, 24
    g = states.EOS[0]["gamma1"]
    X = (np.arange(nx) + 0.5) / nx
    rho = np.ones((ny, nx)) + 0.1 * np.sin(7 * X)[None, :]
    u = np.zeros((ny, nx))
    v = 0.01 * np.cos(5 * X)[None, :] * np.ones((ny, 1))
    pr = np.where(X[None, :] < 0.5, 1.0e20, 1.0) * (1.0 + 0.1 * np.cos(3 * X)[None, :]) * np.ones((ny, 1))
    U = states.euler_cons(rho, u, v, pr, g)
    _rhs_pair(hip, 0, prim, U)


def test_small_magnitudes_keep_fast_path_exact(hip):
    """State scaled by 1e-6 (still admissible): every smoothness indicator is
    far below the 1e-12 floor, so the weights sit near their linear values."""
    U = states.state(0, 64, 32, seed=5)
    U = U * 1.0e-6
    _rhs_pair(hip, 0, True, U)


def test_large_grid_reports_first_face(hip):
    """A grid past 2^28 face points (8200 x 8200) switches the error word to
    36-bit face keys: the first failing face is still found and located.
    Uniform Euler state with one negative-pressure cell at (i0, j0): the
    minmod slopes around it vanish, so the first inadmissible face in the
    reference's loop order is that cell's left x-face."""
    n, i0, j0 = 8200, 8001, 8123
    g = states.EOS[0]["gamma1"]
    U = np.empty((n, n, 4))
    U[..., 0], U[..., 1], U[..., 2], U[..., 3] = 1.0, 0.0, 0.0, 1.0 / (g - 1.0)
    U[j0, i0, 3] = -1.0 / (g - 1.0)  # p = -1
    p = hip.make_problem(hip.EULER, n, n, order=1, prim=True, domain=(0.0, 1.0, 0.0, 1.0), gamma1=g)
    with hip.Context(p) as ctx:
        R, dt, rc = ctx.rhs(U, allow_inadmissible=True)
        msg, x, y = ctx.last_error()
    h = 1.0 / n
    print(f"[large] {msg!r} x={x} y={y}", flush=True)
    assert rc == hip.E_INADMISSIBLE
    assert msg.startswith(KIND_TEXT[2]) and "-1.0" in msg
    assert (x, y) == (0.0 + i0 * h, 0.0 + j0 * h)
    R[j0 - 1:j0 + 2, i0 - 1:i0 + 2] = 0.0
    assert not R.any()  # uniform elsewhere: zero RHS
