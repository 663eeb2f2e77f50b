"""The benchmark workloads: problem descriptions and initial states.

Shared by the benchmark and the tests so that a test of a configuration runs
exactly the instances the benchmark times.  All states are synthetic (cell
averages of closed-form initial conditions, hype_initial_state); the random
draws come from numpy's PCG64 with a stated seed.

* ``single(name)``: the three fixed grids,
  - "kapila": Kapila air-helium shock-bubble, 1024x512 on [0,2]x[0,1], 500 steps;
  - "euler":  Euler four-quadrant Riemann problem, 256x256 on [0,1]^2, 200 steps;
  - "bn":     Baer-Nunziato water-air interface, 2048x1024 on [0,2]x[0,1], 300 steps.
* ``batched(seed, count)``: 64 Kapila shock-bubbles, sides log-uniform on
  [128, 1024], Mach uniform on [1.1, 2.0], bubble radius uniform on
  [0.1, 0.3], centre jittered by +-0.1; 400 steps each.
* ``served_trace(rng, count, rate)``: a Poisson request stream, mix 50 %
  Euler / 35 % Kapila / 15 % Baer-Nunziato, sides log-uniform on [64, 2048],
  100-2000 steps turned into an end time.
"""
from __future__ import annotations

import numpy as np

import hype2d as H

BATCHED_STEPS = 400
QUADRANTS = [0.5, 0.5, 1.5, 0.0, 0.0, 1.5, 0.5323, 1.206, 0.0, 0.3, 0.138, 1.206, 1.206, 0.029,
             0.5323, 0.0, 1.206, 0.3]


def single(name):
    """(problem, initial state, steps, workload name) of a fixed grid."""
    if name == "kapila":
        nx, ny = 1024, 512
        h = 2.0 / nx
        p = H.make_problem(H.KAPILA, nx, ny, order=2, prim=True, domain=(0.0, 2.0, 0.0, 1.0), cfl=0.5,
                           bc=(H.BC_TRANSMISSIVE, H.BC_TRANSMISSIVE, H.BC_REFLECTIVE, H.BC_REFLECTIVE),
                           gamma1=1.4, gamma2=1.648, pi1=0.0, pi2=0.0)
        params = [1.22, 0.5, 1.0, 0.5, 0.25, 1e-6, 2.0 * h, 1.4, 0.25463, 1.0e5]
        U0 = H.initial_state(p, H.IC_SHOCK_BUBBLE, params)
        return p, U0, 500, "kapila_shock_bubble_1024x512"
    if name == "euler":
        nx = ny = 256
        p = H.make_problem(H.EULER, nx, ny, order=2, prim=True, domain=(0.0, 1.0, 0.0, 1.0), cfl=0.5,
                           bc=(H.BC_TRANSMISSIVE,) * 4, gamma1=1.4)
        U0 = H.initial_state(p, H.IC_QUADRANTS, QUADRANTS)
        return p, U0, 200, "euler_quadrants_256x256"
    if name == "bn":
        nx, ny = 2048, 1024
        h = 2.0 / nx
        p = H.make_problem(H.BAER_NUNZIATO, nx, ny, order=2, prim=False, domain=(0.0, 2.0, 0.0, 1.0),
                           cfl=0.5, bc=(H.BC_TRANSMISSIVE,) * 4, gamma1=4.4, gamma2=1.4, pi1=6.0e8, pi2=0.0)
        params = [0.5, 1e-6, 2.0 * h, 1.0, 0.3, 0.1, 1.0e9, 1.0e5, 1000.0, 1.0, 8, 0.005, 2.0]
        U0 = H.initial_state(p, H.IC_WATER_AIR, params, seed=2024)
        return p, U0, 300, "baer_nunziato_water_air_2048x1024"
    raise ValueError(f"unknown workload {name}")


def kapila_bubble(n, mach, R, cx, cy):
    """Square n x n Kapila shock-bubble on [0,1]^2 (air-helium, reflective
    top and bottom), shock at x = 0.05 moving +x."""
    h = 1.0 / n
    p = H.make_problem(H.KAPILA, n, n, order=2, prim=True, domain=(0.0, 1.0, 0.0, 1.0), cfl=0.5,
                       bc=(H.BC_TRANSMISSIVE, H.BC_TRANSMISSIVE, H.BC_REFLECTIVE, H.BC_REFLECTIVE),
                       gamma1=1.4, gamma2=1.648)
    U0 = H.initial_state(p, H.IC_SHOCK_BUBBLE, [mach, 0.05, cx, cy, R, 1e-6, 2.0 * h, 1.4, 0.25463, 1.0e5])
    return p, U0


def batched_draws(seed=2024, count=64):
    """(side, Mach, radius, cx, cy) of every instance of the batched set."""
    rng = np.random.default_rng(seed)
    sides = np.exp(rng.uniform(np.log(128), np.log(1024), count)).astype(int)
    return [(int(n), rng.uniform(1.1, 2.0), rng.uniform(0.1, 0.3), 0.5 + rng.uniform(-0.1, 0.1),
             0.5 + rng.uniform(-0.1, 0.1)) for n in sides]


def batched(seed=2024, count=64):
    """The batched set: a list of (problem, initial state)."""
    return [kapila_bubble(*d) for d in batched_draws(seed, count)]


def batched_name(count=64):
    return f"batched_kapila_{count}x_loguniform_128_1024"


def dt_estimate(p, U):
    """Initial-dt estimate CFL*h/(2*max(|u|+c)) from cell averages (frozen
    sound speeds); used only to turn a requested step count into an end time."""
    g1, g2, p1, p2 = p.gamma1, p.gamma2, p.pi1, p.pi2
    if p.system == H.EULER:
        rho = U[..., 0]
        u, v = U[..., 1] / rho, U[..., 2] / rho
        pr = (g1 - 1.0) * (U[..., 3] - 0.5 * rho * (u * u + v * v))
        c = np.sqrt(np.maximum(g1 * pr / rho, 0.0))
        s = np.maximum(np.abs(u), np.abs(v)) + c
    elif p.system == H.KAPILA:
        rho = U[..., 0] + U[..., 1]
        u, v, a1 = U[..., 2] / rho, U[..., 3] / rho, U[..., 5]
        a2 = 1.0 - a1
        e = U[..., 4] - 0.5 * rho * (u * u + v * v)
        pr = (e - a1 * g1 * p1 / (g1 - 1.0) - a2 * g2 * p2 / (g2 - 1.0)) / (a1 / (g1 - 1.0) + a2 / (g2 - 1.0))
        c2 = (U[..., 0] * g1 * (pr + p1) / np.maximum(U[..., 0] / a1, 1e-300)
              + U[..., 1] * g2 * (pr + p2) / np.maximum(U[..., 1] / np.maximum(a2, 1e-300), 1e-300)) / rho
        s = np.maximum(np.abs(u), np.abs(v)) + np.sqrt(np.maximum(c2, 0.0))
    else:
        phi = U[..., 8]
        s = 0.0
        for q, a, g, pi in ((U[..., 0:4], phi, g1, p1), (U[..., 4:8], 1.0 - phi, g2, p2)):
            r = q[..., 0] / a
            u, v = q[..., 1] / q[..., 0], q[..., 2] / q[..., 0]
            pr = (q[..., 3] / q[..., 0] - 0.5 * (u * u + v * v)) * r * (g - 1.0) - g * pi
            c = np.sqrt(np.maximum(g * (pr + pi) / r, 0.0))
            s = np.maximum(s, np.maximum(np.abs(u), np.abs(v)) + c)
    h = (p.x_max - p.x_min) / p.nx
    return p.cfl * h / (2.0 * float(np.max(s)))


def served_draws(rng, count, rate):
    """Arrival time, system, side, steps and BN seed of every request of the
    stream, without building the states.  Each request's draws follow the
    previous request's, so a shorter stream is a prefix of a longer one."""
    t = 0.0
    out = []
    for k in range(count):
        t += rng.exponential(1.0 / rate)
        n = int(np.exp(rng.uniform(np.log(64), np.log(2048))))
        steps = int(rng.integers(100, 2001))
        u = rng.uniform()
        system = H.EULER if u < 0.5 else (H.KAPILA if u < 0.85 else H.BAER_NUNZIATO)
        seed = int(rng.integers(1 << 62)) if system == H.BAER_NUNZIATO else 0
        out.append((float(t), system, n, steps, seed))
    return out


def served_request(system, n, steps, seed):
    """(problem, initial state, end time, step cap) of one request: the end
    time is the requested step count times the initial-dt estimate; the cap
    is twice the requested count."""
    h = 1.0 / n
    if system == H.EULER:
        p = H.make_problem(H.EULER, n, n, order=2, prim=True, domain=(0.0, 1.0, 0.0, 1.0), cfl=0.5,
                           bc=(H.BC_TRANSMISSIVE,) * 4, gamma1=1.4)
        U0 = H.initial_state(p, H.IC_QUADRANTS, QUADRANTS)
    elif system == H.KAPILA:
        p, U0 = kapila_bubble(n, 1.22, 0.25, 0.5, 0.5)
    else:
        p = H.make_problem(H.BAER_NUNZIATO, n, n, order=2, prim=False, domain=(0.0, 1.0, 0.0, 1.0),
                           cfl=0.5, bc=(H.BC_TRANSMISSIVE,) * 4, gamma1=4.4, gamma2=1.4, pi1=6.0e8)
        U0 = H.initial_state(p, H.IC_WATER_AIR, [0.5, 1e-6, 2.0 * h, 0.5, 0.3, 0.1, 1.0e9, 1.0e5, 1000.0,
                                                 1.0, 8, 0.005, 1.0], seed=seed)
    t_end = steps * dt_estimate(p, U0)
    return p, U0, t_end, 2 * steps


def served_trace(rng, count, rate):
    """The request stream: a list of (arrival, problem, U0, t_end, step cap)."""
    return [(arr,) + served_request(system, n, steps, seed)
            for arr, system, n, steps, seed in served_draws(rng, count, rate)]


def served_name(rate, count):
    return f"served_poisson_{rate:g}rps_{count}req_mix_E50_K35_B15_side64_2048"
