"""hype_rhs_noalloc, the RHS without a context or any allocation, through
the C ABI on device buffers, against hype_rhs of a context on the same
device buffers: the same RHS bits, dt and status, on two calls in a row
with the second on a smaller grid (the scratch keeps the first call's
contents), for every system and both paths, with the left inlet (its table
in the scratch), and on an inadmissible state.  Grids beyond the scratch
are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

import states

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost
T, R = 4, 3      # HYPE_BC_TRANSMISSIVE, HYPE_BC_REFLECTIVE


@pytest.fixture(scope="module")
def hiprt(hip):
    rt = C.CDLL("libamdhip64.so")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


class DevBuf:
    def __init__(self, rt, nbytes):
        self.rt, self.n, self.p = rt, nbytes, C.c_void_p()
        assert rt.hipMalloc(C.byref(self.p), nbytes) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.nbytes == self.n
        assert self.rt.hipMemcpy(self.p, a.ctypes.data, self.n, H2D) == 0

    def get(self, shape):
        out = np.empty(shape, dtype=np.float64)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data, self.p, self.n, D2H) == 0
        return out

    def free(self):
        self.rt.hipFree(self.p)


def problem(hip, system, prim, nx, ny, bc=(T, T, R, R)):
    eos = states.EOS[system]
    return hip.make_problem(system, nx, ny, order=2, prim=prim, domain=(0.0, 1.0, 0.0, ny / nx), cfl=0.5,
                            bc=bc, **eos)


def both(hip, rt, p, U):
    """(RHS, dt, rc, err x/y) of hype_rhs_noalloc and of a context's hype_rhs,
    both on device buffers holding U."""
    shape = U.shape
    dU, dA, dB = DevBuf(rt, U.nbytes), DevBuf(rt, U.nbytes), DevBuf(rt, U.nbytes)
    try:
        dU.put(U)
        dA.put(np.full(shape, np.nan))
        dB.put(np.full(shape, np.nan))
        dt_n, rc_n, xy_n = hip.rhs_noalloc(p, dU.p.value, dA.p.value, allow_inadmissible=True)
        RA = dA.get(shape)
        with hip.Context(p) as ctx:
            dt = C.c_double()
            rc = ctx.L.hype_rhs(ctx.h, 0.0, C.cast(dU.p, C.POINTER(C.c_double)),
                                C.cast(dB.p, C.POINTER(C.c_double)), C.byref(dt), hip.MEM_DEVICE)
            xy = ctx.last_error()[1:] if rc == hip.E_INADMISSIBLE else None
        RB = dB.get(shape)
    finally:
        dU.free()
        dA.free()
        dB.free()
    return (RA, dt_n, rc_n, xy_n), (RB, dt.value, rc, xy)


CASES = [  # (system, prim, first grid, second (smaller) grid)
    (1, True, (300, 170), (61, 23)),
    (1, False, (150, 96), (40, 40)),
    (0, True, (256, 256), (33, 65)),
    (0, False, (120, 80), (64, 9)),
    (2, True, (200, 100), (50, 30)),
    (2, False, (190, 130), (70, 11)),
]


@pytest.mark.parametrize("system,prim,g1,g2", CASES)
def test_noalloc_equals_context_rhs_twice(hip, hiprt, system, prim, g1, g2):
    for k, (nx, ny) in enumerate((g1, g2)):
        p = problem(hip, system, prim, nx, ny)
        U = states.state(system, nx, ny, seed=11 + k)
        print(f"[noalloc] system={system} prim={prim} call {k + 1}: {nx}x{ny}", flush=True)
        (RA, dtA, rcA, _), (RB, dtB, rcB, _) = both(hip, hiprt, p, U)
        assert rcA == rcB == hip.OK
        assert dtA == dtB
        assert np.array_equal(RA, RB), f"call {k + 1}: RHS differs from hype_rhs"


def test_noalloc_left_inlet_twice(hip, hiprt):
    import oracle_py as orc

    for k, (nx, ny) in enumerate(((160, 160), (48, 48))):
        p, _ = hip.case_problem(hip.IC_AIR_JET, nx, ny)
        assert p.bc[0] == hip.BC_INLET
        op = orc.problem(p.system, nx, ny, N=2, prim=True, domain=(p.x_min, p.x_max, p.y_min, p.y_max),
                         cfl=p.cfl, bc=tuple(p.bc), g1=p.gamma1, g2=p.gamma2, p1=p.pi1, p2=p.pi2)
        U = states.inlet_state(orc, op, seed=3 + k)
        print(f"[noalloc] inlet call {k + 1}: {nx}x{ny}", flush=True)
        (RA, dtA, rcA, _), (RB, dtB, rcB, _) = both(hip, hiprt, p, U)
        assert rcA == rcB == hip.OK
        assert dtA == dtB
        assert np.array_equal(RA, RB)


def test_noalloc_inadmissible_then_smaller(hip, hiprt):
    nx, ny = 40, 30
    U = states.state(0, nx, ny, seed=8, sharp=False).copy()
    U[17, 23, 3] *= -0.5  # negative pressure
    p = problem(hip, 0, True, nx, ny, bc=(T, T, T, T))
    print("[noalloc] inadmissible 40x30", flush=True)
    (RA, dtA, rcA, xyA), (RB, dtB, rcB, xyB) = both(hip, hiprt, p, U)
    assert rcA == rcB == hip.E_INADMISSIBLE
    assert xyA == xyB
    assert np.array_equal(RA, RB, equal_nan=True)
    # the next, smaller call starts from a clean error slot
    p2 = problem(hip, 0, True, 24, 20, bc=(T, T, T, T))
    U2 = states.state(0, 24, 20, seed=9)
    print("[noalloc] admissible 24x20 after it", flush=True)
    (RA, dtA, rcA, _), (RB, dtB, rcB, _) = both(hip, hiprt, p2, U2)
    assert rcA == rcB == hip.OK and dtA == dtB
    assert np.array_equal(RA, RB)


def test_noalloc_refuses_grids_beyond_the_scratch(hip, hiprt):
    p, _ = hip.case_problem(hip.IC_AIR_JET, 64, 64)
    p.ny = 13637  # (ny + 6) * 6 doubles of inlet table exceed the scratch
    d = DevBuf(hiprt, 64)
    try:
        dt = C.c_double()
        rc = hip.lib().hype_rhs_noalloc(C.byref(p), 0.0, C.cast(d.p, C.POINTER(C.c_double)),
                                        C.cast(d.p, C.POINTER(C.c_double)), C.byref(dt), None, None)
    finally:
        d.free()
    assert rc == hip.E_TOO_LARGE
