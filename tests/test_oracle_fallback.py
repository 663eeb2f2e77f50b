"""Pin of the admissibility fallback of the oracle's RHS
(rhs_function_prim.c:382-421 on the primitive path, rhs_function.c:382-421 on
the conservative path), restated in oracle/orc_rhs.c.

On a grid whose state does not vary in y, the Euler RHS reduces to a 1D
finite-volume update that is rebuilt here independently of the oracle's RHS
loop:
  * directional WENO3 slopes and curvatures from fitted candidate quadratics
    and Jiang-Shu indicators (test_oracle_recon.weno_dir); the y slopes and
    the cross term vanish for y-uniform data;
  * node values at the 8 face Gauss nodes, the admissibility test of the
    path (density and pressure floors, euler.c:292-310 / :244-262);
  * where a node fails, the replacement polynomial: cell mean plus the two
    minmod slopes with every higher coefficient zeroed on the primitive path
    (rhs_function_prim.c:416), cell mean alone on the conservative path
    (rhs_function.c:416);
  * Rusanov fluxes at the x-faces from orc_riemann (pinned separately), and
    the update -(F_{i+1/2} - F_{i-1/2}) / h; the y-fluxes of identical rows
    cancel exactly.
The fallback mask and the whole RHS must match the oracle.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_py as orc
from test_oracle_recon import weno_dir

GAMMA = 1.4
RHO_FLOOR, PRS_FLOOR = 1e-14, 1e-12  # hype.h:50-51
G = 0.28867513459481287             # face Gauss point (hype.c:213-230)
NODES = [(0.5, -G), (0.5, G), (-0.5, -G), (-0.5, G), (-G, 0.5), (G, 0.5), (-G, -0.5), (G, -0.5)]


def basis(x, y, k):
    return orc.lib().orc_basis(x, y, k)


def state(seed, nx, ny):
    rng = np.random.default_rng(seed)
    rho = 10 ** rng.uniform(-2, 1, nx)
    u = rng.normal(0.0, 1.0, nx)
    p = 10 ** rng.uniform(-5, 1, nx)
    U = np.zeros((ny, nx, 4))
    U[:, :, 0] = rho
    U[:, :, 1] = rho * u
    U[:, :, 3] = p / (GAMMA - 1.0) + 0.5 * rho * u * u
    return U


def admissible(prim, V):
    if prim:
        rho, p = V[0], V[3]
    else:
        rho = V[0]
        p = (GAMMA - 1.0) * (V[3] - 0.5 * (V[1] * V[1] + V[2] * V[2]) / V[0])
    return rho >= RHO_FLOOR and p >= PRS_FLOOR


def minmod(a, b):
    if a * b < 0.0:
        return 0.0
    return a if abs(a) < abs(b) else b


def expected_rhs(P, cells):
    """cells: (nx, 4) values of one row in the path's variables.  Returns the
    expected RHS row (conserved, nx x 4) and the fallback mask."""
    prim = bool(P.prim)
    nx = cells.shape[0]
    ext = np.concatenate([np.repeat(cells[:1], 3, 0), cells, np.repeat(cells[-1:], 3, 0)])  # transmissive
    coef, mask = {}, {}
    for i in range(-1, nx + 1):
        a = np.zeros((4, 6))
        for c in range(4):
            u = ext[i + 1: i + 6, c]  # cells i-2 .. i+2
            a[c, 0] = u[2]
            a[c, 1], a[c, 3] = weno_dir(u)
        nodes = [np.array([sum(a[c, k] * basis(x, y, k) for k in range(6)) for c in range(4)]) for x, y in NODES]
        ok = all(admissible(prim, v) for v in nodes)
        if not ok:
            fb = np.zeros((4, 6))
            for c in range(4):
                fb[c, 0] = ext[i + 3, c]
                if prim:  # minmod slopes kept, k >= 3 zeroed
                    fb[c, 1] = minmod(ext[i + 4, c] - ext[i + 3, c], ext[i + 3, c] - ext[i + 2, c])
                    fb[c, 2] = 0.0  # minmod of the (zero) y differences
            a = fb
        coef[i], mask[i] = a, not ok

    def face(i, x):
        return np.array([sum(coef[i][c, k] * basis(x, G, k) for k in range(6)) for c in range(4)])

    L = orc.lib()
    F = np.zeros((nx + 1, 4))
    D = np.zeros((nx + 1, 4))
    for f in range(nx + 1):  # face between cells f-1 and f
        VL, VR = face(f - 1, 0.5), face(f, -0.5)
        Ff, Df, bad = np.zeros(orc.ORC_MAXVAR_PY), np.zeros(orc.ORC_MAXVAR_PY), C.c_int(0)
        L.orc_riemann(C.byref(P), orc.dptr(np.ascontiguousarray(VL)), orc.dptr(np.ascontiguousarray(VR)),
                      1.0, 0.0, orc.dptr(Ff), orc.dptr(Df), C.byref(bad))
        F[f], D[f] = Ff[:4], Df[:4]
    R = -(F[1:] - F[:-1] + 0.5 * (D[1:] + D[:-1])) / P.h
    return R, np.array([mask[i] for i in range(nx)]), np.max(np.abs(F))


@pytest.mark.parametrize("prim", [True, False])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_fallback_cells_and_rhs(prim, seed):
    nx, ny = 40, 3
    U = state(seed, nx, ny)
    P = orc.problem(orc.EULER, nx, ny, N=2, prim=prim, h=1.0 / nx,
                    bc=(orc.BC_TRANSMISSIVE, orc.BC_TRANSMISSIVE, orc.BC_PERIODIC, orc.BC_PERIODIC), g1=GAMMA)
    ref = orc.rhs(P, U, want_prim=True, want_fallback=True)
    assert ref["rc"] == orc.OK
    row = ref["W"][1] if prim else U[1]
    R, mask, fscale = expected_rhs(P, row)
    # both kinds of cell occur, and some fallback cell has a non-zero minmod
    # slope, so the k >= 1 / k >= 3 distinction is visible
    assert 0 < mask.sum() < nx
    assert np.array_equal(ref["fallback"][1].astype(bool), mask)
    for j in range(ny):
        assert np.array_equal(ref["rhs"][j], ref["rhs"][1])
    tol = 1e-12 * fscale / P.h
    assert np.max(np.abs(ref["rhs"][1] - R)) <= tol


def test_prim_fallback_keeps_nonzero_slopes():
    """The state of the test above has fallback cells whose minmod slopes are
    non-zero: the two rules give different face values there."""
    nx = 40
    U = state(0, nx, 3)
    P = orc.problem(orc.EULER, nx, 3, N=2, prim=True, h=1.0 / nx,
                    bc=(orc.BC_TRANSMISSIVE, orc.BC_TRANSMISSIVE, orc.BC_PERIODIC, orc.BC_PERIODIC), g1=GAMMA)
    ref = orc.rhs(P, U, want_prim=True, want_fallback=True)
    W = ref["W"][1]
    slopes = [max(abs(minmod(W[min(i + 1, nx - 1), c] - W[i, c], W[i, c] - W[max(i - 1, 0), c])) for c in range(4))
              for i in range(nx) if ref["fallback"][1][i]]
    assert sum(s > 1e-3 for s in slopes) >= 3
