"""Seeded test states (conserved variables, AoS [j][i][c]) for the parity tests.

Each generator mixes smooth fields with sharp interfaces and shocks so the
nonlinear WENO weights, the minmod fallback and the reflective/periodic
boundaries are all exercised.
"""
from __future__ import annotations

import numpy as np

EULER, KAPILA, BN = 0, 1, 2


def _grid(nx, ny):
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    return np.meshgrid(x, y)  # arrays of shape (ny, nx)


def euler_cons(rho, u, v, p, g):
    E = p / (g - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E], axis=-1)


def kapila_cons(r1, r2, u, v, p, a1, g1, g2, p1, p2):
    a2 = 1.0 - a1
    rho = a1 * r1 + a2 * r2
    e = a1 * (p + g1 * p1) / (g1 - 1.0) + a2 * (p + g2 * p2) / (g2 - 1.0)
    E = e + 0.5 * rho * (u * u + v * v)
    return np.stack([a1 * r1, a2 * r2, rho * u, rho * v, E, a1], axis=-1)


def bn_cons(r1, u1, v1, pp1, r2, u2, v2, pp2, phi, g1, g2, p1, p2):
    q0 = phi * r1
    q4 = (1.0 - phi) * r2
    E1 = 0.5 * (u1 * u1 + v1 * v1) + (pp1 + g1 * p1) / (r1 * (g1 - 1.0))
    E2 = 0.5 * (u2 * u2 + v2 * v2) + (pp2 + g2 * p2) / (r2 * (g2 - 1.0))
    return np.stack([q0, q0 * u1, q0 * v1, q0 * E1, q4, q4 * u2, q4 * v2, q4 * E2, phi], axis=-1)


def euler_state(nx, ny, seed=0, g=1.4, sharp=True):
    rng = np.random.default_rng(seed)
    X, Y = _grid(nx, ny)
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * (X + 2 * Y)) + 0.05 * rng.standard_normal((ny, nx))
    u = 0.3 * np.cos(2 * np.pi * Y) + 0.05 * rng.standard_normal((ny, nx))
    v = -0.2 * np.sin(2 * np.pi * X)
    p = 1.0 + 0.1 * np.cos(2 * np.pi * X * Y)
    if sharp:  # quadrant-style jumps
        q = (X > 0.5) & (Y > 0.4)
        rho = np.where(q, 0.138, rho)
        p = np.where(q, 0.029, p)
        u = np.where(X < 0.3, u + 1.206, u)
    return euler_cons(rho, u, v, np.abs(p) + 0.01, g)


def kapila_state(nx, ny, seed=0, g1=1.4, g2=1.648, p1=0.0, p2=0.0, eps=1e-6, sharp=True):
    rng = np.random.default_rng(seed)
    X, Y = _grid(nx, ny)
    r = np.hypot(X - 0.55, Y - 0.5)
    width = 0.02 if sharp else 0.15
    a1 = eps + (1 - 2 * eps) * 0.5 * (1 + np.tanh((r - 0.25) / width))
    if sharp:
        a1 = np.where(r < 0.12, eps, a1)  # a sharp core: triggers the fallback
    r1 = 1.4 + 0.1 * np.sin(2 * np.pi * Y) + 0.01 * rng.standard_normal((ny, nx))
    r2 = 0.25463 * (1 + 0.05 * np.cos(2 * np.pi * X))
    p = 1.0e5 * (1 + 0.02 * np.sin(2 * np.pi * (X - Y)))
    u = 30.0 * np.sin(2 * np.pi * Y) + 5.0 * rng.standard_normal((ny, nx))
    v = 20.0 * np.cos(2 * np.pi * X)
    if sharp:  # a Mach-1.22 shock on the left
        s = X < 0.2
        r1 = np.where(s, 1.92691, r1)
        p = np.where(s, 1.5698e5, p)
        u = np.where(s, 114.42, u)
    return kapila_cons(r1, r2, u, v, p, a1, g1, g2, p1, p2)


def bn_state(nx, ny, seed=0, g1=4.4, g2=1.4, p1=6e8, p2=0.0, eps=1e-6, sharp=True):
    rng = np.random.default_rng(seed)
    X, Y = _grid(nx, ny)
    iface = 0.5 + 0.03 * np.sin(2 * np.pi * 3 * X + rng.uniform(0, 2 * np.pi))
    width = 0.01 if sharp else 0.1
    phi = eps + (1 - 2 * eps) * 0.5 * (1 - np.tanh((Y - iface) / width))  # water below
    r1 = 1000.0 * (1 + 0.01 * np.sin(2 * np.pi * X))
    r2 = 1.0 + 0.1 * np.cos(2 * np.pi * Y)
    pp = 1.0e5 * (1 + 0.1 * np.sin(2 * np.pi * (X + Y)))
    if sharp:
        pp = np.where(np.hypot(X - 0.5, Y - 0.3) < 0.1, 1.0e9, pp)
    u1 = 2.0 * np.sin(2 * np.pi * Y) + 0.1 * rng.standard_normal((ny, nx))
    v1 = 1.0 * np.cos(2 * np.pi * X)
    u2 = u1 + 0.5 * rng.standard_normal((ny, nx))
    v2 = v1 - 0.3
    return bn_cons(r1, u1, v1, pp, r2, u2, v2, pp * 1.01, phi, g1, g2, p1, p2)


def state(system, nx, ny, seed=0, sharp=True):
    if system == EULER:
        return euler_state(nx, ny, seed, sharp=sharp)
    if system == KAPILA:
        return kapila_state(nx, ny, seed, sharp=sharp)
    return bn_state(nx, ny, seed, sharp=sharp)


# EOS constants used with each generator
EOS = {
    EULER: dict(gamma1=1.4, gamma2=1.4, pi1=0.0, pi2=0.0),
    KAPILA: dict(gamma1=1.4, gamma2=1.648, pi1=0.0, pi2=0.0),
    BN: dict(gamma1=4.4, gamma2=1.4, pi1=6e8, pi2=0.0),
}


def rough_state(system, nx, ny, seed=0):
    """Salt-and-pepper jumps (near-vacuum density, or volume fractions jumping
    between 1e-6 and 1-1e-6) inside a disc: the reconstruction leaves the
    admissible set at some nodes, so the minmod fallback fires."""
    rng = np.random.default_rng(seed + 100)
    X, Y = _grid(nx, ny)
    reg = np.hypot(X - 0.5, Y - 0.5) < 0.3
    salt = reg & (rng.random((ny, nx)) < 0.3)
    one = np.ones_like(X)
    if system == EULER:
        rho = np.where(salt, 1e-3, 1.0)
        p = np.where(reg, 1e-3, 1.0)
        return euler_cons(rho, 0.5 * one, 0.1 * one, p, 1.4)
    if system == KAPILA:
        a1 = np.where(salt, 1e-6, 1 - 1e-6)
        return kapila_cons(1.4 * one, 0.25463 * one, 10.0 * one, -3.0 * one, 1e5 * one, a1, 1.4, 1.648, 0.0, 0.0)
    phi = np.where(salt, 1e-6, 1 - 1e-6)
    return bn_cons(1000.0 * one, 0.2 * one, 0.0 * one, 1e5 * one, one, 0.1 * one, 0.3 * one, 1e5 * one, phi,
                   4.4, 1.4, 6e8, 0.0)


# Kapila jet setting of the reference's left inlet: water (gamma 4.4, pi 6000)
# and air (gamma 1.4), the EOS constants of hype.h:42-45
INLET_EOS = dict(gamma1=4.4, gamma2=1.4, pi1=6000.0, pi2=0.0)


def inlet_state(orc, P, seed=0, amp=0.05):
    """Kapila state near the left-inlet profile: every column carries the
    inlet's primitive state at the row's height, smoothly perturbed (density,
    velocity and pressure within a few per cent), converted with the oracle's
    PDEPrim2Cons.  P is an oracle problem (x_min, y_min, h, nx, ny)."""
    rng = np.random.default_rng(seed)
    nx, ny = P.nx, P.ny
    x = P.x_min + (np.arange(nx) + 0.5) * P.h
    y = P.y_min + (np.arange(ny) + 0.5) * P.h
    k1, k2 = rng.uniform(0.5, 2.0, 2)
    ph = rng.uniform(0, 2 * np.pi, 3)
    U = np.empty((ny, nx, 6))
    for j in range(ny):
        V0 = orc.inlet_prim(P, y[j])
        for i in range(nx):
            s = np.sin(k1 * x[i] + ph[0]) * np.cos(k2 * y[j] + ph[1])
            V = V0.copy()
            V[0] *= 1.0 + amp * s
            V[1] *= 1.0 + amp * np.cos(k1 * x[i] - ph[2])
            V[2] += amp * 3.0 * s
            V[3] += amp * 2.0 * np.sin(k2 * y[j] + ph[2])
            V[4] *= 1.0 + amp * s
            U[j, i] = orc.prim2cons(P, V)
    return U
