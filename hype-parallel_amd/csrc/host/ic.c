/*
 * ic.c -- initial states for the benchmark configurations and the
 * reference's smooth-vortex test, cell-averaged like InitializeSolution
 * (initialize_solution.c:21-81): 5x5 Gauss-Legendre points per cell, the
 * pointwise state converted to conserved variables (PDEPrim2Cons) before
 * averaging.
 *
 * Cell centres are x_min + (i+1/2) h.  HYPE_COORD_DMDA reproduces the
 * reference's DMDASetUniformCoordinates(x_min+h/2, x_max+h/2) call
 * (hype.c:132-135): for a non-periodic axis DMDA spaces M points over that
 * interval with M-1 gaps, so the sampling centres are stretched by M/(M-1).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "common.h"

/* ---- conserved variables from primitives (same formulas as PDEPrim2Cons) */
static void prim2cons(const hype_problem *p, const double *V, double *Q)
{
    if (p->system == HYPE_EULER) { /* euler.c:29-38 */
        double e = (V[3]) / (p->gamma1 - 1.0);
        double k = 0.5 * V[0] * (V[1] * V[1] + V[2] * V[2]);
        Q[0] = V[0];
        Q[1] = V[0] * V[1];
        Q[2] = V[0] * V[2];
        Q[3] = k + e;
    } else if (p->system == HYPE_KAPILA) { /* kapila_5eqn.c:41-61 */
        double a1 = V[5], a2 = 1.0 - a1;
        double rho = a1 * V[0] + a2 * V[1];
        double G1m1 = p->gamma1 - 1.0, G2m1 = p->gamma2 - 1.0;
        double e = (V[4] * (a1 * G2m1 + a2 * G1m1) + a1 * p->gamma1 * p->pi1 * G2m1 +
                    a2 * p->gamma2 * p->pi2 * G1m1) /
                   (G1m1 * G2m1);
        Q[0] = a1 * V[0];
        Q[1] = a2 * V[1];
        Q[2] = rho * V[2];
        Q[3] = rho * V[3];
        Q[4] = e + 0.5 * rho * (V[2] * V[2] + V[3] * V[3]);
        Q[5] = a1;
    } else { /* baer_nunziato.c:47-60 */
        Q[0] = V[8] * V[0];
        Q[1] = Q[0] * V[1];
        Q[2] = Q[0] * V[2];
        Q[3] = Q[0] * (0.5 * (V[1] * V[1] + V[2] * V[2]) +
                       (V[3] + p->gamma1 * p->pi1) / (V[0] * (p->gamma1 - 1.0)));
        Q[4] = (1.0 - V[8]) * V[4];
        Q[5] = Q[4] * V[5];
        Q[6] = Q[4] * V[6];
        Q[7] = Q[4] * (0.5 * (V[5] * V[5] + V[6] * V[6]) +
                       (V[7] + p->gamma2 * p->pi2) / (V[4] * (p->gamma2 - 1.0)));
        Q[8] = V[8];
    }
}

/* ---- left inlet of the Kapila system ------------------------------------
 * InletBCPrim (kapila_5eqn.c:485-513); InletBC (:289-320) is PDEPrim2Cons of
 * the same state.  The profile depends on the row's height only, so the host
 * tabulates it once per ghosted row with the C library's sqrt and tanh, the
 * functions the reference calls. */
static void inlet_prim(const hype_problem *p, double y, double *V)
{
    const double M = 2.0;
    const double rho_air = 1.0;
    const double rho_water = 1000.0;
    const double Pr = 1.013;
    const double U_avg = M * sqrt(p->gamma2 * Pr / rho_air);
    const double phi_R = 1.0 - 1.0e-6;
    const double phi_L = 1.0e-6;
    const double PHI = 0.5 * (phi_R + phi_L);
    const double dif_phi = (phi_L - phi_R) / 2.0;
    const double hh = 30.0 / 500.0;
    const double epsilon = 3.0 * hh;
    const double r_0 = 1.0;
    const double r = fabs(y);
    V[0] = rho_water;
    V[1] = rho_air;
    V[2] = (r < 1.0) ? 2.0 * U_avg * (1.0 - y * y) : 0.0;
    V[3] = 0.0;
    V[4] = Pr;
    V[5] = PHI + dif_phi * tanh(-(r - r_0) / epsilon);
}

size_t hype_inlet_doubles(const hype_problem *p)
{
    if (p->bc[0] != HYPE_BC_INLET) return 0;
    return (size_t)(p->ny + 6) * (size_t)hype_nvar(p->system);
}

void hype_inlet_table(const hype_problem *p, double *out)
{
    const double h = (p->x_max - p->x_min) / (double)(p->nx); /* hype.c:57 */
    const int nv = hype_nvar(p->system);
    int j;
    for (j = -3; j < p->ny + 3; ++j) {
        /* cell-centre height of the (ghost) row, rhs_function_prim.c:138 */
        const double y_loc = p->y_min + ((double)(j) + 0.5) * h;
        double V[9];
        double *d = out + (size_t)(j + 3) * nv;
        int c;
        inlet_prim(p, y_loc, V);
        if (p->recons_prim)
            for (c = 0; c < nv; ++c) d[c] = V[c];
        else
            prim2cons(p, V, d);
    }
}

/* ---- splitmix64 / xoshiro256** (seeded parameters) -------------------- */
static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double hype_rng_uniform(uint64_t *state)
{
    return (double)(splitmix64(state) >> 11) * (1.0 / 9007199254740992.0);
}

/* ---- pointwise initial conditions (primitive variables) ---------------- */
typedef struct {
    const hype_problem *p;
    int kind;
    const double *a; /* parameters */
    double modes[3][8]; /* interface modes of the water-air case */
} ic_ctx;

/* Mach-M shock moving +x into gas (rho0, p0, at rest): Rankine-Hugoniot
 * post-shock density, velocity and pressure */
void hype_rankine_hugoniot(double gamma, double rho0, double p0, double M, double *rho1,
                           double *u1, double *p1)
{
    double c0 = sqrt(gamma * p0 / rho0);
    double W = M * c0;
    *p1 = p0 * (1.0 + 2.0 * gamma / (gamma + 1.0) * (M * M - 1.0));
    *rho1 = rho0 * (gamma + 1.0) * M * M / ((gamma - 1.0) * M * M + 2.0);
    *u1 = W * (1.0 - rho0 / *rho1);
}

static void point_state(const ic_ctx *C, double x, double y, double *V)
{
    const hype_problem *p = C->p;
    const double *a = C->a;
    switch (C->kind) {
    case HYPE_IC_UNIFORM: {
        memcpy(V, a, (size_t)hype_nvar(p->system) * sizeof(double));
        break;
    }
    case HYPE_IC_QUADRANTS: { /* a: xs, ys, NE[4], NW[4], SW[4], SE[4] (rho,u,v,p) */
        int q = (x >= a[0]) ? ((y >= a[1]) ? 0 : 3) : ((y >= a[1]) ? 1 : 2);
        memcpy(V, a + 2 + 4 * q, 4 * sizeof(double));
        break;
    }
    case HYPE_IC_SHOCK_BUBBLE: {
        /* a: mach, x_shock, cx, cy, R, eps, width, rho_air, rho_he, p0.
         * Phase 1 is air (gamma1), phase 2 helium (gamma2). */
        double rho_a = a[7], rho_h = a[8], p0 = a[9];
        double r = sqrt((x - a[2]) * (x - a[2]) + (y - a[3]) * (y - a[3]));
        double alpha = a[5] + (1.0 - 2.0 * a[5]) * 0.5 * (1.0 + tanh((r - a[4]) / a[6]));
        V[0] = rho_a;
        V[1] = rho_h;
        V[2] = 0.0;
        V[3] = 0.0;
        V[4] = p0;
        V[5] = alpha;
        if (x < a[1]) {
            double r1, u1, pp;
            hype_rankine_hugoniot(p->gamma1, rho_a, p0, a[0], &r1, &u1, &pp);
            V[0] = r1;
            V[2] = u1;
            V[4] = pp;
        }
        break;
    }
    case HYPE_IC_WATER_AIR: {
        /* a: y_iface, eps, width, cx, cy, r_burst, p_burst, p0, rho_w, rho_a,
         * nmodes, amplitude, domain length.  Phase 1 water (below the
         * interface), phase 2 air. */
        double eta = a[0];
        int k, nm = (int)a[10];
        for (k = 0; k < nm; ++k)
            eta += a[11] * sin(2.0 * M_PI * (k + 1) * x / a[12] + C->modes[0][k]);
        double phi = a[1] + (1.0 - 2.0 * a[1]) * 0.5 * (1.0 - tanh((y - eta) / a[2]));
        double r = sqrt((x - a[3]) * (x - a[3]) + (y - a[4]) * (y - a[4]));
        double pr = (r < a[5]) ? a[6] : a[7];
        V[0] = a[8];
        V[1] = 0.0;
        V[2] = 0.0;
        V[3] = pr;
        V[4] = a[9];
        V[5] = 0.0;
        V[6] = 0.0;
        V[7] = pr;
        V[8] = phi;
        break;
    }
    case HYPE_IC_SMOOTH_VORTEX: { /* kapila_5eqn.c:527-562 */
        const double alpha_min = 0.25, alpha_max = 0.75, omega = M_PI / 3.0;
        const double p_atm = 1.0, k_eps = 0.3, R = 1.0, delta = 0.1;
        const double rho1 = 1000.0, rho2 = 1.0, u0 = 1.0, v0 = 1.0;
        double r = sqrt(x * x + y * y);
        V[0] = rho1 + delta * rho1 * sin(omega * (2.0 * x + y)) * cos(omega * (x - 2.0 * y));
        V[1] = rho2 + delta * rho2 * sin(omega * (x - 2.0 * y)) * cos(omega * (2.0 * x + y));
        V[2] = u0;
        V[3] = v0;
        V[4] = p_atm;
        V[5] = alpha_min + 0.5 * (alpha_max - alpha_min) * erfc((r / R - 1.0) / k_eps);
        break;
    }
    default: memset(V, 0, HYPE_MAXVAR * sizeof(double));
    }
}

int hype_initial_state(const hype_problem *p, int kind, const double *params, int nparams,
                       unsigned long long seed, int coord_mode, double *U)
{
    /* initialize_solution.c:41-43 */
    static const double x_gp5[] = {0.0000000000000000, -0.2692346550528416, 0.2692346550528416,
                                   -0.4530899229693320, 0.4530899229693320};
    static const double w_gp5[] = {0.28444444444444444, 0.23931433524968326, 0.23931433524968326,
                                   0.11846344252809456, 0.11846344252809456};
    static const int need[] = {0, 18, 10, 13, 0};
    ic_ctx C;
    int nv = 0, i, j, l, m, c;
    double h;
    HYPE_TRY(hype_check_problem(p, &nv));
    if (!U || kind < 0 || kind > HYPE_IC_SMOOTH_VORTEX) return HYPE_E_INVALID;
    if (kind == HYPE_IC_UNIFORM && nparams < nv) return HYPE_E_INVALID;
    if (kind != HYPE_IC_UNIFORM && nparams < need[kind]) return HYPE_E_INVALID;
    if ((kind == HYPE_IC_QUADRANTS && p->system != HYPE_EULER) ||
        ((kind == HYPE_IC_SHOCK_BUBBLE || kind == HYPE_IC_SMOOTH_VORTEX) && p->system != HYPE_KAPILA) ||
        (kind == HYPE_IC_WATER_AIR && p->system != HYPE_BAER_NUNZIATO))
        return HYPE_E_INVALID;
    memset(&C, 0, sizeof(C));
    C.p = p;
    C.kind = kind;
    C.a = params;
    if (kind == HYPE_IC_WATER_AIR) {
        uint64_t s = seed;
        int k;
        if ((int)params[10] > 8 || (int)params[10] < 0) return HYPE_E_INVALID;
        for (k = 0; k < 8; ++k) C.modes[0][k] = 2.0 * M_PI * hype_rng_uniform(&s);
    }
    h = (p->x_max - p->x_min) / (double)p->nx; /* hype.c:57 */
    for (j = 0; j < p->ny; ++j)
        for (i = 0; i < p->nx; ++i) {
            double integral[HYPE_MAXVAR], V[HYPE_MAXVAR], Q[HYPE_MAXVAR];
            double xc = hype_cell_coord(p, 0, i, coord_mode);
            double yc = hype_cell_coord(p, 1, j, coord_mode);
            for (c = 0; c < nv; ++c) integral[c] = 0.0;
            for (l = 0; l < 5; ++l)
                for (m = 0; m < 5; ++m) {
                    point_state(&C, xc + h * x_gp5[l], yc + h * x_gp5[m], V);
                    prim2cons(p, V, Q);
                    for (c = 0; c < nv; ++c) integral[c] += w_gp5[l] * w_gp5[m] * Q[c];
                }
            for (c = 0; c < nv; ++c) U[((size_t)j * p->nx + i) * nv + c] = integral[c];
        }
    return HYPE_OK;
}
