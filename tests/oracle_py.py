"""ctypes binding of the CPU oracle (oracle/lib/liboracle*.so).

Test infrastructure: used only by tests, the smoke check and the benchmark's
CPU-baseline leg, as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

EULER, KAPILA, BN = 0, 1, 2
BC_WALL, BC_INLET, BC_PERIODIC, BC_REFLECTIVE, BC_TRANSMISSIVE = range(5)
OK, E_INVALID, E_UNSUPPORTED, E_INADMISSIBLE = 0, 1, 2, 5  # oracle.h status codes
ORC_MAXVAR_PY = 16
NVAR = {EULER: 4, KAPILA: 6, BN: 9}
QUIRK_BN_P2_IS_PHI = 0x1


class OrcProblem(C.Structure):
    _fields_ = [
        ("system", C.c_int),
        ("N", C.c_int),
        ("prim", C.c_int),
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("x_min", C.c_double),
        ("y_min", C.c_double),
        ("h", C.c_double),
        ("cfl", C.c_double),
        ("bc", C.c_int * 4),
        ("g1", C.c_double),
        ("g2", C.c_double),
        ("p1", C.c_double),
        ("p2", C.c_double),
        ("quirks", C.c_uint),
    ]


class OrcError(C.Structure):
    _fields_ = [
        ("code", C.c_int),
        ("kind", C.c_int),
        ("x", C.c_double),
        ("y", C.c_double),
        ("value", C.c_double),
        ("index", C.c_long),
    ]


_libs = {}


def lib(threaded=False):
    name = "liboracle_omp.so" if threaded else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(ORACLE_DIR, "lib", name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle`")
    L = C.CDLL(path)
    P = C.POINTER
    dp = P(C.c_double)
    L.orc_rhs.argtypes = [P(OrcProblem), C.c_double, dp, dp, dp, dp, P(C.c_ubyte), P(OrcError)]
    L.orc_solve.argtypes = [P(OrcProblem), dp, C.c_double, C.c_double, C.c_int, C.c_int,
                            P(C.c_int), dp, dp, P(OrcError)]
    L.orc_ctx_create.argtypes = [P(OrcProblem)]
    L.orc_ctx_create.restype = C.c_void_p
    L.orc_ctx_destroy.argtypes = [C.c_void_p]
    L.orc_ctx_destroy.restype = None
    L.orc_rhs_ctx.argtypes = [C.c_void_p, C.c_double, dp, dp, dp, dp, P(C.c_ubyte), P(OrcError)]
    L.orc_cons2prim.argtypes = [P(OrcProblem), dp, dp]
    L.orc_cons2prim.restype = None
    L.orc_prim2cons.argtypes = [P(OrcProblem), dp, dp]
    L.orc_prim2cons.restype = None
    L.orc_inlet_prim.argtypes = [P(OrcProblem), C.c_double, dp]
    L.orc_inlet_prim.restype = None
    L.orc_flux.argtypes = [P(OrcProblem), dp, C.c_double, C.c_double, dp, P(C.c_int)]
    L.orc_flux.restype = C.c_double
    L.orc_flux_prim.argtypes = [P(OrcProblem), dp, C.c_double, C.c_double, dp, P(C.c_int)]
    L.orc_flux_prim.restype = C.c_double
    L.orc_ncp.argtypes = [P(OrcProblem), dp, dp, dp, dp]
    L.orc_ncp.restype = None
    L.orc_ncp_prim.argtypes = [P(OrcProblem), dp, dp, dp, dp]
    L.orc_ncp_prim.restype = None
    L.orc_pad.argtypes = [P(OrcProblem), dp]
    L.orc_pad_prim.argtypes = [P(OrcProblem), dp]
    L.orc_riemann.argtypes = [P(OrcProblem), dp, dp, C.c_double, C.c_double, dp, dp, P(C.c_int)]
    L.orc_riemann.restype = C.c_double
    L.orc_reconstruct.argtypes = [dp, dp, dp, C.c_int, dp]
    L.orc_reconstruct.restype = None
    L.orc_basis.argtypes = [C.c_double, C.c_double, C.c_int]
    L.orc_basis.restype = C.c_double
    L.orc_minmod.argtypes = [C.c_double, C.c_double]
    L.orc_minmod.restype = C.c_double
    L.orc_ssp_step.argtypes = [C.c_int, C.c_long, dp, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                               dp, dp]
    _libs[name] = L
    return L


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def inlet_prim(P, y):
    """Primitive state of the Kapila left inlet at height y."""
    V = np.zeros(ORC_MAXVAR_PY, dtype=np.float64)
    lib().orc_inlet_prim(C.byref(P), float(y), dptr(V))
    return V[:NVAR[P.system]].copy()


def prim2cons(P, V):
    """PDEPrim2Cons of one primitive state."""
    nv = NVAR[P.system]
    V = np.ascontiguousarray(V, dtype=np.float64)
    Q = np.zeros(nv, dtype=np.float64)
    lib().orc_prim2cons(C.byref(P), dptr(V), dptr(Q))
    return Q


def problem(system, nx, ny, *, N=2, prim=True, x_min=0.0, y_min=0.0, h=None, cfl=0.5,
            bc=(BC_TRANSMISSIVE,) * 4, g1=1.4, g2=1.4, p1=0.0, p2=0.0, quirks=0, domain=None):
    p = OrcProblem()
    p.system, p.N, p.prim, p.nx, p.ny = system, N, 1 if prim else 0, nx, ny
    if domain is not None:
        x_min, x_max, y_min, _ = domain
        h = (x_max - x_min) / nx
    p.x_min, p.y_min = x_min, y_min
    p.h = h if h is not None else 1.0 / nx
    p.cfl = cfl
    for k in range(4):
        p.bc[k] = bc[k]
    p.g1, p.g2, p.p1, p.p2, p.quirks = g1, g2, p1, p2, quirks
    return p


def rhs(P, U, t=0.0, threaded=False, want_prim=False, want_fallback=False):
    """Returns dict(rhs, dt, rc, err, W?, fallback?)."""
    L = lib(threaded)
    nv = NVAR[P.system]
    U = np.ascontiguousarray(U, dtype=np.float64).reshape(P.ny, P.nx, nv)
    R = np.zeros_like(U)
    dt = C.c_double()
    err = OrcError()
    W = np.zeros_like(U) if want_prim else None
    fb = np.zeros((P.ny, P.nx), dtype=np.uint8) if want_fallback else None
    rc = L.orc_rhs(C.byref(P), t, dptr(U), dptr(R), C.byref(dt),
                   dptr(W) if W is not None else None,
                   fb.ctypes.data_as(C.POINTER(C.c_ubyte)) if fb is not None else None, C.byref(err))
    return dict(rhs=R, dt=dt.value, rc=rc, err=err, W=W, fallback=fb)


def solve(P, U0, steps, t0=0.0, t_end=float("inf"), clip=True, threaded=False):
    L = lib(threaded)
    nv = NVAR[P.system]
    U = np.array(U0, dtype=np.float64, copy=True).reshape(P.ny, P.nx, nv)
    nsteps = C.c_int()
    tout = C.c_double()
    dts = np.zeros(max(steps, 1), dtype=np.float64)
    err = OrcError()
    rc = L.orc_solve(C.byref(P), dptr(U), t0, t_end, steps, 1 if clip else 0, C.byref(nsteps),
                     C.byref(tout), dptr(dts), C.byref(err))
    return dict(U=U, rc=rc, steps=nsteps.value, t=tout.value, dts=dts[: nsteps.value], err=err)


class RhsCtx:
    """Reusable oracle scratch (for timing)."""

    def __init__(self, P, threaded=True):
        self.L = lib(threaded)
        self.P = P
        self.h = self.L.orc_ctx_create(C.byref(P))
        if not self.h:
            raise RuntimeError("orc_ctx_create failed")

    def __call__(self, U, R):
        dt = C.c_double()
        err = OrcError()
        rc = self.L.orc_rhs_ctx(self.h, 0.0, dptr(U), dptr(R), C.byref(dt), None, None, C.byref(err))
        return rc, dt.value

    def close(self):
        if self.h:
            self.L.orc_ctx_destroy(self.h)
            self.h = None
