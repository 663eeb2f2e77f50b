"""Pins of the oracle's time integration (PETSc TSSSP as driven by
hype.c:337-384 and post_process.c:19, restated in oracle/orc_time.c).

  * Each scheme equals its Butcher tableau (derived here from the low-storage
    Shu-Osher forms of Ketcheson 2008) on a non-autonomous nonlinear ODE,
    which also pins the stage times;
  * the linear amplification factors R(z) are the known polynomials;
  * the solve loop uses the dt of the last RHS call of the previous step
    (priming call for step 1) and lands on t_end when clipping.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_py as orc
import states

CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double),
                 C.POINTER(C.c_double))


def tableau(N):
    if N == 0:
        return np.array([[0.0]]), np.array([1.0]), np.array([0.0])
    if N == 1:
        return np.array([[0, 0], [1.0, 0]]), np.array([0.5, 0.5]), np.array([0.0, 1.0])
    if N == 2:
        A = np.zeros((4, 4))
        A[1, 0] = 0.5
        A[2, :2] = 0.5
        A[3, :3] = 1.0 / 6
        return A, np.array([1 / 6, 1 / 6, 1 / 6, 0.5]), np.array([0, 0.5, 1.0, 0.5])
    A = np.zeros((10, 10))
    for i in range(1, 5):
        A[i, :i] = 1.0 / 6
    for i in range(5, 10):
        A[i, :5] = 1.0 / 15
        A[i, 5:i] = 1.0 / 6
    c = A.sum(axis=1)
    return A, np.full(10, 0.1), c


def butcher_step(N, f, t0, u, dt):
    A, b, c = tableau(N)
    k = []
    for i in range(len(b)):
        ui = u + dt * sum(A[i, j] * k[j] for j in range(i)) if i else u.copy()
        k.append(f(t0 + c[i] * dt, ui))
    return u + dt * sum(b[i] * k[i] for i in range(len(b)))


def oracle_step(N, f, t0, u, dt):
    n = len(u)
    U = np.array(u, dtype=np.float64)
    work = np.zeros(3 * n)
    last = C.c_double()
    times = []

    def cb(ctx, t, Up, Fp, dtp):
        x = np.ctypeslib.as_array(Up, shape=(n,)).copy()
        times.append(t)
        np.ctypeslib.as_array(Fp, shape=(n,))[:] = f(t, x)
        dtp[0] = 0.25 + len(times)  # a marker: the step reports the last call's value
        return 0

    cfun = CB(cb)
    rc = orc.lib().orc_ssp_step(N, n, orc.dptr(U), t0, dt, C.cast(cfun, C.c_void_p), None, orc.dptr(work),
                                C.byref(last))
    assert rc == 0
    return U, times, last.value


@pytest.mark.parametrize("N", [0, 1, 2, 3])
def test_schemes_match_butcher_tableaux(N):
    f = lambda t, u: np.cos(3.0 * t) - u * u + 0.1 * u[::-1]
    u0 = np.array([0.3, -0.7, 1.1])
    t0, dt = 0.4, 0.05
    got, times, last = oracle_step(N, f, t0, u0, dt)
    exp = butcher_step(N, f, t0, u0, dt)
    assert np.max(np.abs(got - exp)) <= 1e-15 * 10
    A, b, c = tableau(N)
    assert np.allclose(times, t0 + c * dt, rtol=0, atol=1e-15)
    assert last == 0.25 + len(b)


@pytest.mark.parametrize("N,poly", [
    (0, [1, 1]),
    (1, [1, 1, 1 / 2]),
    (2, [1, 1, 1 / 2, 1 / 6, 1 / 48]),
])
def test_linear_amplification(N, poly):
    lam, dt = -0.8, 0.7
    got, _, _ = oracle_step(N, lambda t, u: lam * u, 0.0, np.array([1.0]), dt)
    z = lam * dt
    assert abs(got[0] - sum(a * z ** k for k, a in enumerate(poly))) < 1e-15


def test_rk104_fourth_order():
    lam = -1.0
    errs = []
    for dt in (0.2, 0.1):
        got, _, _ = oracle_step(3, lambda t, u: lam * u, 0.0, np.array([1.0]), dt)
        errs.append(abs(got[0] - np.exp(lam * dt)))
    assert 4.5 < np.log2(errs[0] / errs[1]) < 5.5  # local error O(dt^5)


def test_solve_uses_lagged_dt_and_clips_final_step():
    """dt of step k+1 = dt written by the last RHS call of step k (stage-4
    input u_3), the first from the priming call (hype.c:343,356)."""
    system, nx, ny = 1, 24, 16
    eos = states.EOS[system]
    U0 = states.state(system, nx, ny, seed=1)
    P = orc.problem(system, nx, ny, N=2, domain=(0, 1, 0, ny / nx), cfl=0.5, bc=(4, 4, 3, 3),
                    g1=eos["gamma1"], g2=eos["gamma2"], p1=eos["pi1"], p2=eos["pi2"])
    L = orc.lib()
    ctx = L.orc_ctx_create(C.byref(P))
    n = nx * ny * 6
    calls = []

    def cb(_c, t, Up, Fp, dtp):
        err = orc.OrcError()
        rc = L.orc_rhs_ctx(ctx, t, Up, Fp, dtp, None, None, C.byref(err))
        calls.append(dtp[0])
        return rc

    cfun = CB(cb)
    U = U0.copy().reshape(-1)
    R = np.zeros(n)
    dt0 = C.c_double()
    L.orc_rhs_ctx(ctx, 0.0, orc.dptr(U), orc.dptr(R), C.byref(dt0), None, None, C.byref(orc.OrcError()))
    dts, dt, t = [], dt0.value, 0.0
    work = np.zeros(3 * n)
    for _ in range(5):
        last = C.c_double()
        L.orc_ssp_step(2, n, orc.dptr(U), t, dt, C.cast(cfun, C.c_void_p), None, orc.dptr(work), C.byref(last))
        dts.append(dt)
        t += dt
        assert last.value == calls[-1]
        dt = last.value
    L.orc_ctx_destroy(ctx)
    ref = orc.solve(P, U0, 5)
    assert np.array_equal(ref["dts"], np.array(dts))
    assert np.array_equal(ref["U"].reshape(-1), U)
    # clipping: stop exactly at t_end, the last step shortened
    t_end = sum(dts[:3]) + 0.25 * dts[3]
    clipped = orc.solve(P, U0, 100, t_end=t_end, clip=True)
    assert clipped["steps"] == 4 and clipped["t"] == t_end
    assert clipped["dts"][3] == t_end - sum(dts[:3])
