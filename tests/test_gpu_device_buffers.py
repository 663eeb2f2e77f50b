"""The C ABI on device buffers (HYPE_MEM_DEVICE): the RHS callback, the
state transfers of a context and of a batch, against the host-buffer path,
bit for bit.  Device memory comes straight from the HIP runtime."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as orc
import states

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


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


def problem(hip, system, prim, nx=96, ny=48):
    eos = states.EOS[system]
    quirks = orc.QUIRK_BN_P2_IS_PHI if system == 2 and prim else 0
    return hip.make_problem(system, nx, ny, order=2, prim=prim, domain=(0.0, 2.0, 0.0, 1.0), cfl=0.5,
                            bc=(hip.BC_TRANSMISSIVE, hip.BC_TRANSMISSIVE, hip.BC_REFLECTIVE, hip.BC_REFLECTIVE),
                            quirks=quirks, **eos)


@pytest.mark.parametrize("system,prim", [(1, True), (2, False), (0, True)])
def test_rhs_on_device_buffers_equals_host_path(hip, hiprt, system, prim):
    p = problem(hip, system, prim)
    U = states.state(system, p.nx, p.ny, seed=5)
    with hip.Context(p) as ctx:
        R_h, dt_h, _ = ctx.rhs(U)
        Ud, Rd = DevBuf(hiprt, U.nbytes), DevBuf(hiprt, U.nbytes)
        try:
            Ud.put(U)
            dt_d = ctx.rhs_device(Ud.p.value, Rd.p.value)
            R_d = Rd.get(U.shape)
        finally:
            Ud.free()
            Rd.free()
    assert dt_d == dt_h
    assert np.array_equal(R_d, R_h)


def test_context_state_through_device_buffers(hip, hiprt):
    p = problem(hip, 1, True)
    U0 = states.state(1, p.nx, p.ny, seed=6)
    with hip.Context(p) as ctx:
        U_h, st_h = ctx.solve(U0, 6)
    d_in, d_out = DevBuf(hiprt, U0.nbytes), DevBuf(hiprt, U0.nbytes)
    try:
        d_in.put(U0)
        with hip.Context(p) as ctx:
            assert ctx.L.hype_set_state(ctx.h, C.cast(d_in.p, C.POINTER(C.c_double)), hip.MEM_DEVICE) == hip.OK
            ctx.begin(0.0, float("inf"), 6, hip.TS_CLIP)
            ctx.advance(6)
            st = ctx.sync()
            assert ctx.L.hype_get_state(ctx.h, C.cast(d_out.p, C.POINTER(C.c_double)), hip.MEM_DEVICE) == hip.OK
            U_d = d_out.get(U0.shape)
    finally:
        d_in.free()
        d_out.free()
    assert st.steps == st_h.steps == 6 and st.t == st_h.t
    assert np.array_equal(U_d, U_h)


def test_batch_state_through_device_buffers(hip, hiprt):
    probs = [problem(hip, 1, True, 64, 40), problem(hip, 0, True, 50, 30)]
    U0s = [states.state(1, 64, 40, seed=7), states.state(0, 50, 30, seed=8)]
    bufs = [DevBuf(hiprt, U.nbytes) for U in U0s]
    outs = [DevBuf(hiprt, U.nbytes) for U in U0s]
    try:
        for b, U in zip(bufs, U0s):
            b.put(U)
        with hip.Batch(probs) as bt:
            for i, b in enumerate(bufs):
                assert bt.L.hype_batch_set_state(bt.h, i, C.cast(b.p, C.POINTER(C.c_double)), hip.MEM_DEVICE) == hip.OK
            bt.begin(max_steps=np.array([5, 7], dtype=np.int32))
            bt.advance(8)
            st = bt.sync()
            for i, o in enumerate(outs):
                assert bt.L.hype_batch_get_state(bt.h, i, C.cast(o.p, C.POINTER(C.c_double)), hip.MEM_DEVICE) == hip.OK
            got = [o.get(U.shape) for o, U in zip(outs, U0s)]
    finally:
        for b in bufs + outs:
            b.free()
    for i, (p, U0, n) in enumerate(zip(probs, U0s, (5, 7))):
        with hip.Context(p) as ctx:
            U, s1 = ctx.solve(U0, n)
        assert st[i].steps == s1.steps == n
        assert np.array_equal(got[i], U)
