"""ctypes binding of libhype2d.so (the C ABI declared in include/hype2d.h).

Mirrors the reference's callback interface: ``Context.rhs(U, t)`` is the
RHSFunctionPrim / RHSFunction call (HyPE-2D/rhs_function_prim.c:12,
rhs_function.c:12) on the DMDA AoS layout, returning the RHS and the dt side
effect; ``Context.solve`` drives the device-resident SSP-RK loop of
hype.c:337-384.  The binding refuses to run without the compiled library:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libhype2d.so")

OK, E_INVALID, E_UNSUPPORTED, E_DEVICE, E_NOMEM, E_INADMISSIBLE, E_BUSY, E_NOTFOUND, E_TOO_LARGE = range(9)
EULER, KAPILA, BAER_NUNZIATO = 0, 1, 2
BC_WALL, BC_INLET, BC_PERIODIC, BC_REFLECTIVE, BC_TRANSMISSIVE = range(5)
QUIRK_BN_P2_IS_PHI = 0x1
OPT_FMA = 0x1
OPT_ONE_WAVE_TILES = 0x2  # force one wave per tile
OPT_SPLIT_TILES = 0x4  # force tiles split over a two-wave workgroup
MEM_HOST, MEM_DEVICE = 0, 1
TS_CLIP, TS_GRAPH = 0x1, 0x2
NVAR = {EULER: 4, KAPILA: 6, BAER_NUNZIATO: 9}


class HypeError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"hype2d error {code}: {msg}")
        self.code = code
        self.msg = msg


class Problem(C.Structure):
    _fields_ = [
        ("system", C.c_int),
        ("order", C.c_int),
        ("recons_prim", C.c_int),
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("x_min", C.c_double),
        ("x_max", C.c_double),
        ("y_min", C.c_double),
        ("y_max", C.c_double),
        ("cfl", C.c_double),
        ("bc", C.c_int * 4),
        ("gamma1", C.c_double),
        ("gamma2", C.c_double),
        ("pi1", C.c_double),
        ("pi2", C.c_double),
        ("quirks", C.c_uint),
        ("options", C.c_uint),
    ]


class Status(C.Structure):
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("steps", C.c_int), ("done", C.c_int), ("err", C.c_int)]


class BatchOpts(C.Structure):
    _fields_ = [("tile_rows", C.c_int)]


class ServerOpts(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_size_t),
        ("max_active", C.c_int),
        ("steps_per_round", C.c_int),
        ("tile_rows", C.c_int),
        ("round_ms", C.c_double),
        ("max_active_cells", C.c_longlong),
        ("policy", C.c_int),
    ]


SERVE_FIFO, SERVE_SHORTEST = 0, 1


class Request(C.Structure):
    _fields_ = [
        ("problem", Problem),
        ("U0", C.POINTER(C.c_double)),
        ("U_out", C.POINTER(C.c_double)),
        ("t0", C.c_double),
        ("t_end", C.c_double),
        ("max_steps", C.c_int),
        ("flags", C.c_uint),
    ]


REQ_QUEUED, REQ_RUNNING, REQ_DONE = 0, 1, 2


class RequestInfo(C.Structure):
    _fields_ = [
        ("status", Status),
        ("state", C.c_int),
        ("t_submit", C.c_double),
        ("t_start", C.c_double),
        ("t_done", C.c_double),
        ("bytes", C.c_size_t),
    ]


def make_problem(system, nx, ny, *, order=2, prim=True, domain=(0.0, 1.0, 0.0, 1.0), cfl=0.5,
                 bc=(BC_TRANSMISSIVE,) * 4, gamma1=1.4, gamma2=1.4, pi1=0.0, pi2=0.0, quirks=0, fma=False):
    p = Problem()
    p.system, p.order, p.recons_prim = system, order, 1 if prim else 0
    p.nx, p.ny = nx, ny
    p.x_min, p.x_max, p.y_min, p.y_max = domain
    p.cfl = cfl
    for k in range(4):
        p.bc[k] = bc[k]
    p.gamma1, p.gamma2, p.pi1, p.pi2 = gamma1, gamma2, pi1, pi2
    p.quirks = quirks
    p.options = OPT_FMA if fma else 0
    return p


_lib = None


def lib():
    """Load libhype2d.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HypeError(E_DEVICE, f"{LIB_PATH} is missing; build it with `make -C hype-parallel_amd`")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    dp = P(C.c_double)
    vp = C.c_void_p
    L.hype_nvar.argtypes = [C.c_int]
    L.hype_ctx_create.argtypes = [P(Problem), P(vp)]
    L.hype_ctx_destroy.argtypes = [vp]
    L.hype_ctx_destroy.restype = None
    L.hype_rhs.argtypes = [vp, C.c_double, dp, dp, dp, C.c_int]
    L.hype_rhs_diag.argtypes = [vp, C.c_double, dp, dp, dp, P(C.c_ubyte), C.c_int]
    L.hype_rhs_noalloc.argtypes = [P(Problem), C.c_double, dp, dp, dp, dp, vp]
    L.hype_cons2prim.argtypes = [vp, dp, dp, C.c_int]
    L.hype_set_state.argtypes = [vp, dp, C.c_int]
    L.hype_get_state.argtypes = [vp, dp, C.c_int]
    L.hype_begin.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.c_uint]
    L.hype_advance.argtypes = [vp, C.c_int]
    L.hype_sync.argtypes = [vp, P(Status)]
    L.hype_dt_history.argtypes = [vp, dp, C.c_int]
    L.hype_last_error.argtypes = [vp, dp, dp]
    L.hype_last_error.restype = C.c_char_p
    L.hype_stream.argtypes = [vp]
    L.hype_stream.restype = vp
    L.hype_advance_timed.argtypes = [vp, C.c_int, dp]
    L.hype_initial_state.argtypes = [P(Problem), C.c_int, dp, C.c_int, C.c_ulonglong, C.c_int, dp]
    L.hype_rankine_hugoniot.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, dp, dp, dp]
    L.hype_rankine_hugoniot.restype = None
    L.hype_case_setup.argtypes = [C.c_int, P(Problem), dp]
    L.hype_ctx_set_tiling.argtypes = [vp, C.c_int, C.c_int]
    L.hype_ctx_tiling.argtypes = [vp, P(C.c_int), P(C.c_int), P(C.c_int)]
    ip = P(C.c_int)
    L.hype_batch_create.argtypes = [P(Problem), C.c_int, P(BatchOpts), P(vp)]
    L.hype_batch_destroy.argtypes = [vp]
    L.hype_batch_destroy.restype = None
    L.hype_batch_set_state.argtypes = [vp, C.c_int, dp, C.c_int]
    L.hype_batch_get_state.argtypes = [vp, C.c_int, dp, C.c_int]
    L.hype_batch_begin.argtypes = [vp, dp, dp, ip, C.c_uint]
    L.hype_batch_advance.argtypes = [vp, C.c_int]
    L.hype_batch_advance_timed.argtypes = [vp, C.c_int, dp]
    L.hype_batch_sync.argtypes = [vp, P(Status)]
    L.hype_batch_last_error.argtypes = [vp, dp, dp]
    L.hype_batch_last_error.restype = C.c_char_p
    L.hype_batch_tiling.argtypes = [vp, ip, ip, ip]
    L.hype_batch_set_rows.argtypes = [vp, C.c_int]
    L.hype_batch_autotune.argtypes = [vp, ip, C.c_int, C.c_int, ip, dp]
    L.hype_batch_layout.argtypes = [vp, ip]
    L.hype_server_create.argtypes = [P(ServerOpts), P(vp)]
    L.hype_server_destroy.argtypes = [vp]
    L.hype_server_destroy.restype = None
    L.hype_server_submit.argtypes = [vp, P(Request), ip]
    L.hype_server_poll.argtypes = [vp, ip, ip]
    L.hype_server_drain.argtypes = [vp]
    L.hype_server_info.argtypes = [vp, C.c_int, P(RequestInfo)]
    L.hype_server_now.argtypes = [vp]
    L.hype_server_now.restype = C.c_double
    L.hype_server_bytes_in_use.argtypes = [vp]
    L.hype_server_bytes_in_use.restype = C.c_size_t
    L.hype_write_vec_binary.argtypes = [C.c_char_p, dp, C.c_long, C.c_int]
    L.hype_read_vec_binary.argtypes = [C.c_char_p, dp, C.c_long]
    L.hype_write_vts.argtypes = [C.c_char_p, P(Problem), dp, C.c_int]
    _lib = L
    return L


def exported_symbols():
    """Names declared in include/hype2d.h (checked by the CPU tests)."""
    return [
        "hype_nvar", "hype_ctx_create", "hype_ctx_destroy", "hype_rhs", "hype_rhs_diag", "hype_rhs_noalloc",
        "hype_cons2prim", "hype_set_state", "hype_get_state", "hype_begin", "hype_advance",
        "hype_sync", "hype_dt_history", "hype_last_error", "hype_stream",
        "hype_ctx_set_tiling", "hype_ctx_tiling", "hype_advance_timed", "hype_initial_state",
        "hype_rankine_hugoniot", "hype_case_setup", "hype_batch_create", "hype_batch_destroy", "hype_batch_set_state",
        "hype_batch_get_state", "hype_batch_begin", "hype_batch_advance", "hype_batch_advance_timed",
        "hype_batch_sync", "hype_batch_last_error", "hype_batch_tiling", "hype_batch_set_rows",
        "hype_batch_autotune", "hype_batch_layout", "hype_server_create", "hype_server_destroy", "hype_server_submit",
        "hype_server_poll", "hype_server_drain", "hype_server_info", "hype_server_now",
        "hype_server_bytes_in_use", "hype_write_vec_binary", "hype_read_vec_binary", "hype_write_vts",
    ]


IC_UNIFORM, IC_QUADRANTS, IC_SHOCK_BUBBLE, IC_WATER_AIR, IC_SMOOTH_VORTEX = range(5)
# the reference's Kapila test cases (kapila_5eqn.c:565-996), no parameters
(IC_INTERFACE_ADVECTION, IC_AIR_HELIUM, IC_AIR_R22, IC_RICHTMYER_MESHKOV, IC_UNDERWATER_EXPLOSION,
 IC_WATER_CYLINDER, IC_AIR_CAVITY, IC_BUBBLE_COLLAPSE, IC_AIR_JET, IC_WATER_JET) = range(5, 15)
COORD_EXACT, COORD_DMDA = 0, 1


def initial_state(problem, kind, params=(), seed=0, coord=COORD_EXACT):
    """Cell-averaged initial state (hype_initial_state), shape (ny, nx, nvar)."""
    L = lib()
    nv = NVAR[problem.system]
    U = np.empty((problem.ny, problem.nx, nv), dtype=np.float64)
    a = np.ascontiguousarray(params, dtype=np.float64)
    if a.size == 0:
        a = np.zeros(1)
    rc = L.hype_initial_state(C.byref(problem), kind, _dptr(a), int(len(params)), seed, coord, _dptr(U))
    if rc != OK:
        raise HypeError(rc, "hype_initial_state")
    return U


def case_problem(kind, nx, ny, *, order=2, prim=True, cfl=0.5):
    """Problem of a Kapila test case as its header states it
    (hype_case_setup): domain, boundaries and EOS; returns (problem, t_end)."""
    p = make_problem(KAPILA, nx, ny, order=order, prim=prim, cfl=cfl)
    t_end = C.c_double()
    rc = lib().hype_case_setup(kind, C.byref(p), C.byref(t_end))
    if rc != OK:
        raise HypeError(rc, "hype_case_setup")
    return p, t_end.value


def write_restart(path, U, block_size=None):
    """State U in PETSc's binary Vec format (restart1.bin of the reference)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    bs = U.shape[-1] if block_size is None else block_size
    rc = lib().hype_write_vec_binary(os.fsencode(path), _dptr(U), U.size, bs)
    if rc != OK:
        raise HypeError(rc, f"hype_write_vec_binary({path})")


def read_restart(path, shape):
    U = np.empty(shape, dtype=np.float64)
    rc = lib().hype_read_vec_binary(os.fsencode(path), _dptr(U), U.size)
    if rc != OK:
        raise HypeError(rc, f"hype_read_vec_binary({path})")
    return U


def write_vts(path, problem, W, coord=COORD_DMDA):
    """Primitive state W as a VTK XML structured grid (sol-*.vts of the reference)."""
    W = np.ascontiguousarray(W, dtype=np.float64)
    rc = lib().hype_write_vts(os.fsencode(path), C.byref(problem), _dptr(W), coord)
    if rc != OK:
        raise HypeError(rc, f"hype_write_vts({path})")


def rankine_hugoniot(gamma, rho0, p0, mach):
    L = lib()
    r, u, p = C.c_double(), C.c_double(), C.c_double()
    L.hype_rankine_hugoniot(gamma, rho0, p0, mach, C.byref(r), C.byref(u), C.byref(p))
    return r.value, u.value, p.value


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def rhs_noalloc(problem, U_ptr, R_ptr, t=0.0, stream=None, allow_inadmissible=False):
    """hype_rhs_noalloc on device buffers (addresses U_ptr, R_ptr): the RHS
    without a context or any allocation.  Returns (dt, rc, (x, y) of the
    first inadmissible face or None)."""
    dt = C.c_double()
    xy = (C.c_double * 2)()
    rc = lib().hype_rhs_noalloc(C.byref(problem), t, C.cast(U_ptr, C.POINTER(C.c_double)),
                                C.cast(R_ptr, C.POINTER(C.c_double)), C.byref(dt),
                                C.cast(xy, C.POINTER(C.c_double)), stream)
    if rc == OK:
        return dt.value, rc, None
    if rc == E_INADMISSIBLE and allow_inadmissible:
        return dt.value, rc, (xy[0], xy[1])
    raise HypeError(rc, "hype_rhs_noalloc")


class Context:
    """One grid on the device (hype_ctx)."""

    def __init__(self, problem: Problem):
        self.L = lib()
        self.problem = problem
        self.nv = NVAR[problem.system]
        self.shape = (problem.ny, problem.nx, self.nv)
        h = C.c_void_p()
        rc = self.L.hype_ctx_create(C.byref(problem), C.byref(h))
        if rc != OK:
            raise HypeError(rc, "hype_ctx_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.hype_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what, allow_inadmissible=False):
        if rc == OK or (allow_inadmissible and rc == E_INADMISSIBLE):
            return rc
        msg = what
        if rc == E_INADMISSIBLE:
            msg += ": " + self.last_error()[0]
        raise HypeError(rc, msg)

    def last_error(self):
        x, y = C.c_double(), C.c_double()
        s = self.L.hype_last_error(self.h, C.byref(x), C.byref(y))
        return (s.decode() if s else ""), x.value, y.value

    def rhs(self, U, t=0.0, fallback=False, allow_inadmissible=False):
        """RHS of the host state U (shape ny x nx x nvar). Returns (RHS, dt[, mask], rc)."""
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.shape)
        R = np.empty_like(U)
        dt = C.c_double()
        if fallback:
            m = np.zeros(self.shape[:2], dtype=np.uint8)
            rc = self.L.hype_rhs_diag(self.h, t, _dptr(U), _dptr(R), C.byref(dt),
                                      m.ctypes.data_as(C.POINTER(C.c_ubyte)), MEM_HOST)
            self._check(rc, "hype_rhs_diag", allow_inadmissible)
            return R, dt.value, m, rc
        rc = self.L.hype_rhs(self.h, t, _dptr(U), _dptr(R), C.byref(dt), MEM_HOST)
        self._check(rc, "hype_rhs", allow_inadmissible)
        return R, dt.value, rc

    def rhs_device(self, U_ptr, R_ptr, t=0.0):
        dt = C.c_double()
        rc = self.L.hype_rhs(self.h, t, C.cast(U_ptr, C.POINTER(C.c_double)),
                             C.cast(R_ptr, C.POINTER(C.c_double)), C.byref(dt), MEM_DEVICE)
        self._check(rc, "hype_rhs(device)")
        return dt.value

    def cons2prim(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.shape)
        W = np.empty_like(U)
        self._check(self.L.hype_cons2prim(self.h, _dptr(U), _dptr(W), MEM_HOST), "hype_cons2prim")
        return W

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.shape)
        self._check(self.L.hype_set_state(self.h, _dptr(U), MEM_HOST), "hype_set_state")

    def get_state(self):
        U = np.empty(self.shape, dtype=np.float64)
        self._check(self.L.hype_get_state(self.h, _dptr(U), MEM_HOST), "hype_get_state")
        return U

    def begin(self, t0=0.0, t_end=float("inf"), max_steps=1 << 30, flags=TS_CLIP):
        self._check(self.L.hype_begin(self.h, t0, t_end, max_steps, flags), "hype_begin")

    def advance(self, nsteps):
        self._check(self.L.hype_advance(self.h, nsteps), "hype_advance")

    def sync(self, allow_inadmissible=False):
        st = Status()
        rc = self.L.hype_sync(self.h, C.byref(st))
        self._check(rc, "hype_sync", allow_inadmissible)
        return st

    def dt_history(self, n):
        d = np.zeros(n, dtype=np.float64)
        self._check(self.L.hype_dt_history(self.h, _dptr(d), n), "hype_dt_history")
        return d

    def advance_timed(self, nsteps):
        ms = C.c_double()
        self._check(self.L.hype_advance_timed(self.h, nsteps, C.byref(ms)), "hype_advance_timed")
        return ms.value

    def set_tiling(self, hrows, grid=0):
        self._check(self.L.hype_ctx_set_tiling(self.h, hrows, grid), "hype_ctx_set_tiling")

    def tiling(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._check(self.L.hype_ctx_tiling(self.h, C.byref(a), C.byref(b), C.byref(c)), "hype_ctx_tiling")
        return a.value, b.value, c.value

    def solve(self, U0, steps, t0=0.0, t_end=float("inf"), flags=TS_CLIP, chunk=None):
        """Advance U0 by `steps` steps (or to t_end); returns (U, status)."""
        self.set_state(U0)
        self.begin(t0, t_end, steps, flags)
        self.advance(steps)
        st = self.sync()
        return self.get_state(), st


class Batch:
    """Many independent instances advanced together (hype_batch)."""

    def __init__(self, problems, tile_rows=0):
        self.L = lib()
        self.problems = list(problems)
        n = len(self.problems)
        arr = (Problem * n)(*self.problems)
        opts = BatchOpts(tile_rows)
        h = C.c_void_p()
        rc = self.L.hype_batch_create(arr, n, C.byref(opts), C.byref(h))
        if rc != OK:
            raise HypeError(rc, "hype_batch_create")
        self.h = h
        self.shapes = [(p.ny, p.nx, NVAR[p.system]) for p in self.problems]

    def __len__(self):
        return len(self.problems)

    def close(self):
        if getattr(self, "h", None):
            self.L.hype_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what, allow_inadmissible=False):
        if rc == OK or (allow_inadmissible and rc == E_INADMISSIBLE):
            return rc
        msg = what
        if rc == E_INADMISSIBLE:
            msg += ": " + self.last_error()[0]
        raise HypeError(rc, msg)

    def last_error(self):
        x, y = C.c_double(), C.c_double()
        s = self.L.hype_batch_last_error(self.h, C.byref(x), C.byref(y))
        return (s.decode() if s else ""), x.value, y.value

    def set_state(self, i, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.shapes[i])
        self._check(self.L.hype_batch_set_state(self.h, i, _dptr(U), MEM_HOST), "hype_batch_set_state")

    def get_state(self, i):
        U = np.empty(self.shapes[i], dtype=np.float64)
        self._check(self.L.hype_batch_get_state(self.h, i, _dptr(U), MEM_HOST), "hype_batch_get_state")
        return U

    def begin(self, t_end=None, max_steps=None, t0=None, flags=TS_CLIP):
        n = len(self)
        te = np.full(n, np.inf) if t_end is None else np.ascontiguousarray(np.broadcast_to(t_end, (n,)), dtype=np.float64)
        ms = np.full(n, 1 << 30, dtype=np.int32) if max_steps is None else \
            np.ascontiguousarray(np.broadcast_to(max_steps, (n,)), dtype=np.int32)
        t0p = None
        if t0 is not None:
            t0a = np.ascontiguousarray(np.broadcast_to(t0, (n,)), dtype=np.float64)
            t0p = _dptr(t0a)
        self._check(self.L.hype_batch_begin(self.h, t0p, _dptr(te), ms.ctypes.data_as(C.POINTER(C.c_int)), flags),
                    "hype_batch_begin")

    def advance(self, nsteps):
        self._check(self.L.hype_batch_advance(self.h, nsteps), "hype_batch_advance")

    def advance_timed(self, nsteps):
        ms = C.c_double()
        self._check(self.L.hype_batch_advance_timed(self.h, nsteps, C.byref(ms)), "hype_batch_advance_timed")
        return ms.value

    def sync(self, allow_inadmissible=False):
        st = (Status * len(self))()
        rc = self.L.hype_batch_sync(self.h, st)
        self._check(rc, "hype_batch_sync", allow_inadmissible)
        return list(st)

    def tiling(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._check(self.L.hype_batch_tiling(self.h, C.byref(a), C.byref(b), C.byref(c)), "hype_batch_tiling")
        return a.value, b.value, c.value

    def set_rows(self, rows):
        self._check(self.L.hype_batch_set_rows(self.h, rows), "hype_batch_set_rows")

    def layout(self):
        """Tile layout of the instances built for both (-1 default one wave
        per tile, 1 split tiles)."""
        v = C.c_int()
        self._check(self.L.hype_batch_layout(self.h, C.byref(v)), "hype_batch_layout")
        return v.value

    def autotune(self, candidates, steps=4):
        c = np.ascontiguousarray(candidates, dtype=np.int32)
        r, ms = C.c_int(), C.c_double()
        self._check(self.L.hype_batch_autotune(self.h, c.ctypes.data_as(C.POINTER(C.c_int)), len(c), steps,
                                               C.byref(r), C.byref(ms)), "hype_batch_autotune")
        return r.value, ms.value


class Server:
    """Request loop with continuous batching (hype_server)."""

    def __init__(self, hbm_budget_bytes=0, max_active=0, steps_per_round=0, tile_rows=0, round_ms=0.0,
                 max_active_cells=0, policy=SERVE_FIFO):
        self.L = lib()
        o = ServerOpts(hbm_budget_bytes, max_active, steps_per_round, tile_rows, round_ms, max_active_cells, policy)
        h = C.c_void_p()
        rc = self.L.hype_server_create(C.byref(o), C.byref(h))
        if rc != OK:
            raise HypeError(rc, "hype_server_create")
        self.h = h
        self._keep = {}  # request id -> (U0, U_out) kept alive until done

    def close(self):
        if getattr(self, "h", None):
            self.L.hype_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, problem, U0, t_end=float("inf"), max_steps=1 << 30, t0=0.0, flags=TS_CLIP):
        shape = (problem.ny, problem.nx, NVAR[problem.system])
        U0 = np.ascontiguousarray(U0, dtype=np.float64).reshape(shape)
        out = np.empty(shape, dtype=np.float64)
        r = Request(problem, _dptr(U0), _dptr(out), t0, t_end, max_steps, flags)
        rid = C.c_int()
        rc = self.L.hype_server_submit(self.h, C.byref(r), C.byref(rid))
        if rc != OK:
            raise HypeError(rc, "hype_server_submit")
        self._keep[rid.value] = (U0, out)
        return rid.value

    def poll(self):
        a, q = C.c_int(), C.c_int()
        rc = self.L.hype_server_poll(self.h, C.byref(a), C.byref(q))
        if rc != OK:
            raise HypeError(rc, "hype_server_poll")
        return a.value, q.value

    def drain(self):
        rc = self.L.hype_server_drain(self.h)
        if rc != OK:
            raise HypeError(rc, "hype_server_drain")

    def info(self, rid):
        inf = RequestInfo()
        rc = self.L.hype_server_info(self.h, rid, C.byref(inf))
        if rc != OK:
            raise HypeError(rc, "hype_server_info")
        return inf

    def result(self, rid):
        """(U, info) of a finished request; the result buffer is released."""
        inf = self.info(rid)
        if inf.state != REQ_DONE:
            raise HypeError(E_BUSY, f"request {rid} not finished")
        U0, out = self._keep.pop(rid)
        return out, inf

    def now(self):
        return self.L.hype_server_now(self.h)

    def bytes_in_use(self):
        return self.L.hype_server_bytes_in_use(self.h)
