/*
 * orc_physics.c -- point functions of the three PDE systems (oracle; test
 * infrastructure only).  Each function restates the reference function named
 * in its comment, keeping the reference's expression trees so that the device
 * path can be compared bit for bit.  The EOS constants are runtime values
 * (P->g1, g2, p1, p2) where the reference has GAMMA_1/2, PI_1/2 (hype.h:42-45).
 *
 * Inadmissible-state checks that call MPI_Abort in the reference set *bad to
 * a kind code instead (first failing check in the reference's order):
 *   1 negative density           euler.c:60-64 / 212-216
 *   2 negative pressure          euler.c:66-70 / 218-222
 *   3 negative square sound speed kapila_5eqn.c:95-99 / 353-357
 *   4 negative density (solid)   baer_nunziato.c:91-95 / 281-285
 *   5 negative density (gas)     baer_nunziato.c:97-101 / 287-291
 *   6 negative pressure (solid)  baer_nunziato.c:103-107 / 293-297
 *   7 negative pressure (gas)    baer_nunziato.c:109-113 / 299-303
 */
#include <math.h>
#include "oracle.h"
#include "orc_internal.h"

int orc_nvar(int system)
{
    switch (system) {
    case ORC_EULER: return 4;
    case ORC_KAPILA: return 6;
    case ORC_BN: return 9;
    default: return 0;
    }
}

static double pmax(double a, double b) { return (a < b) ? b : a; } /* PetscMax */

static void setbad(int *bad, double *val, int kind, double v)
{
    if (bad && *bad == 0) {
        *bad = kind;
        if (val) *val = v;
    }
}

/* ---------------------------------------------------------------- Euler -- */

/* euler.c:14-23 */
static void euler_cons2prim(const orc_problem *P, const double *Q, double *V)
{
    double irho = 1.0 / Q[0];
    V[0] = Q[0];
    V[1] = irho * Q[1];
    V[2] = irho * Q[2];
    V[3] = (P->g1 - 1.0) * (Q[3] - 0.5 * irho * (Q[1] * Q[1] + Q[2] * Q[2]));
}

/* euler.c:29-38 */
static void euler_prim2cons(const orc_problem *P, const double *V, double *Q)
{
    double e = (V[3]) / (P->g1 - 1.0);
    double k = 0.5 * V[0] * (V[1] * V[1] + V[2] * V[2]);
    Q[0] = V[0];
    Q[1] = V[0] * V[1];
    Q[2] = V[0] * V[2];
    Q[3] = k + e;
}

/* euler.c:44-84 */
static double euler_flux(const orc_problem *P, const double *Q, double nx,
                         double ny, double *F, int *bad, double *bval)
{
    double rho = Q[0];
    double irho = 1.0 / rho;
    double u = irho * Q[1];
    double v = irho * Q[2];
    double un = nx * u + ny * v;
    double rhoun = rho * un;
    double E = Q[3];
    double p = (P->g1 - 1.0) * (E - 0.5 * rho * (u * u + v * v));
    if (rho < ORC_RHO_FLOOR) setbad(bad, bval, 1, rho);
    if (p < ORC_PRS_FLOOR) setbad(bad, bval, 2, p);
    F[0] = rhoun;
    F[1] = rhoun * u + nx * p;
    F[2] = rhoun * v + ny * p;
    F[3] = un * (E + p);
    return fabs(un) + sqrt(P->g1 * p * irho);
}

/* euler.c:197-236 */
static double euler_flux_prim(const orc_problem *P, const double *V, double nx,
                              double ny, double *F, int *bad, double *bval)
{
    double rho = V[0];
    double u = V[1];
    double v = V[2];
    double un = nx * u + ny * v;
    double rhoun = rho * un;
    double p = V[3];
    double E = p / (P->g1 - 1.0) + 0.5 * rho * (u * u + v * v);
    if (rho < ORC_RHO_FLOOR) setbad(bad, bval, 1, rho);
    if (p < ORC_PRS_FLOOR) setbad(bad, bval, 2, p);
    F[0] = rhoun;
    F[1] = rhoun * u + nx * p;
    F[2] = rhoun * v + ny * p;
    F[3] = un * (E + p);
    return fabs(un) + sqrt(P->g1 * p / rho);
}

/* euler.c:169-191 */
static int euler_pad(const orc_problem *P, const double *Q)
{
    int PAD = 1;
    double rho = Q[0];
    double irho = 1.0 / rho;
    double u = irho * Q[1];
    double v = irho * Q[2];
    double E = Q[3];
    double p = (P->g1 - 1.0) * (E - 0.5 * rho * (u * u + v * v));
    if (rho < ORC_RHO_FLOOR) PAD = 0;
    if (p < ORC_PRS_FLOOR) PAD = 0;
    return PAD;
}

/* euler.c:292-310 */
static int euler_pad_prim(const double *V)
{
    int PAD = 1;
    if (V[0] < ORC_RHO_FLOOR) PAD = 0;
    if (V[3] < ORC_PRS_FLOOR) PAD = 0;
    return PAD;
}

/* --------------------------------------------------------------- Kapila -- */

/* kapila_5eqn.c:14-35 */
static void kapila_cons2prim(const orc_problem *P, const double *Q, double *V)
{
    double alpha_1 = Q[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho = Q[0] + Q[1];
    double u = Q[2] / rho;
    double v = Q[3] / rho;
    double E = Q[4];
    double e = E - 0.5 * rho * (u * u + v * v);
    double G1m1 = P->g1 - 1.0;
    double G2m1 = P->g2 - 1.0;
    double p = (e * G1m1 * G2m1 - alpha_1 * P->g1 * P->p1 * G2m1 -
                alpha_2 * P->g2 * P->p2 * G1m1) /
               (alpha_1 * G2m1 + alpha_2 * G1m1);
    V[0] = Q[0] / alpha_1;
    V[1] = Q[1] / alpha_2;
    V[2] = u;
    V[3] = v;
    V[4] = p;
    V[5] = alpha_1;
}

/* kapila_5eqn.c:41-61 */
static void kapila_prim2cons(const orc_problem *P, const double *V, double *Q)
{
    double alpha_1 = V[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho = alpha_1 * V[0] + alpha_2 * V[1];
    double u = V[2];
    double v = V[3];
    double p = V[4];
    double G1m1 = P->g1 - 1.0;
    double G2m1 = P->g2 - 1.0;
    double e = (p * (alpha_1 * G2m1 + alpha_2 * G1m1) +
                alpha_1 * P->g1 * P->p1 * G2m1 +
                alpha_2 * P->g2 * P->p2 * G1m1) /
               (G1m1 * G2m1);
    double E = e + 0.5 * rho * (u * u + v * v);
    Q[0] = alpha_1 * V[0];
    Q[1] = alpha_2 * V[1];
    Q[2] = rho * V[2];
    Q[3] = rho * V[3];
    Q[4] = E;
    Q[5] = alpha_1;
}

/* kapila_5eqn.c:67-113 */
static double kapila_flux(const orc_problem *P, const double *Q, double nx,
                          double ny, double *F, int *bad, double *bval)
{
    double alpha_1 = Q[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho_1 = Q[0] / alpha_1;
    double rho_2 = Q[1] / alpha_2;
    double rho = Q[0] + Q[1];
    double u = Q[2] / rho;
    double v = Q[3] / rho;
    double E = Q[4];
    double e = E - 0.5 * rho * (u * u + v * v);
    double G1m1 = P->g1 - 1.0;
    double G2m1 = P->g2 - 1.0;
    double p = (e * G1m1 * G2m1 - alpha_1 * P->g1 * P->p1 * G2m1 -
                alpha_2 * P->g2 * P->p2 * G1m1) /
               (alpha_1 * G2m1 + alpha_2 * G1m1);
    double un = u * nx + v * ny;
    double a1_sq = P->g1 * (p + P->p1) / rho_1;
    double a2_sq = P->g2 * (p + P->p2) / rho_2;
    double a_sq = (rho_1 * a1_sq * rho_2 * a2_sq) /
                  (rho * (alpha_2 * rho_1 * a1_sq + alpha_1 * rho_2 * a2_sq));
    if (a_sq < ORC_PRS_FLOOR) setbad(bad, bval, 3, a_sq);
    F[0] = alpha_1 * rho_1 * un;
    F[1] = alpha_2 * rho_2 * un;
    F[2] = rho * u * un + p * nx;
    F[3] = rho * v * un + p * ny;
    F[4] = (E + p) * un;
    F[5] = 0.0;
    return fabs(un) + sqrt(a_sq);
}

/* kapila_5eqn.c:164-198 */
static void kapila_ncp(const orc_problem *P, const double *Q, const double *gx,
                       const double *gy, double *B)
{
    double alpha_1 = Q[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho_1 = Q[0] / alpha_1;
    double rho_2 = Q[1] / alpha_2;
    double rho = Q[0] + Q[1];
    double irho = 1.0 / rho;
    double u = Q[2] / rho;
    double v = Q[3] / rho;
    double E = Q[4];
    double e = E - 0.5 * rho * (u * u + v * v);
    double G1m1 = P->g1 - 1.0;
    double G2m1 = P->g2 - 1.0;
    double p = (e * G1m1 * G2m1 - alpha_1 * P->g1 * P->p1 * G2m1 -
                alpha_2 * P->g2 * P->p2 * G1m1) /
               (alpha_1 * G2m1 + alpha_2 * G1m1);
    double a1_sq = P->g1 * (p + P->p1) / rho_1;
    double a2_sq = P->g2 * (p + P->p2) / rho_2;
    double K = alpha_1 * alpha_2 * (rho_2 * a2_sq - rho_1 * a1_sq) /
               (alpha_1 * rho_2 * a2_sq + alpha_2 * rho_1 * a1_sq);
    double alpha1_x = gx[5];
    double alpha1_y = gy[5];
    double u_x = irho * irho * (gx[2] * rho - Q[2] * (gx[0] + gx[1]));
    double v_y = irho * irho * (gy[3] * rho - Q[3] * (gy[0] + gy[1]));
    B[0] = 0.0;
    B[1] = 0.0;
    B[2] = 0.0;
    B[3] = 0.0;
    B[4] = 0.0;
    B[5] = u * alpha1_x + v * alpha1_y - K * (u_x + v_y);
}

/* kapila_5eqn.c:247-283 (note the alpha_1 test at :278, quirk kept) */
static int kapila_pad(const orc_problem *P, const double *Q)
{
    double alpha_1 = Q[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho_1 = Q[0] / alpha_1;
    double rho_2 = Q[1] / alpha_2;
    double rho = Q[0] + Q[1];
    double u = Q[2] / rho;
    double v = Q[3] / rho;
    double E = Q[4];
    double e = E - 0.5 * rho * (u * u + v * v);
    double G1m1 = P->g1 - 1.0;
    double G2m1 = P->g2 - 1.0;
    double p = (e * G1m1 * G2m1 - alpha_1 * P->g1 * P->p1 * G2m1 -
                alpha_2 * P->g2 * P->p2 * G1m1) /
               (alpha_1 * G2m1 + alpha_2 * G1m1);
    double a1_sq = P->g1 * (p + P->p1) / rho_1;
    double a2_sq = P->g2 * (p + P->p2) / rho_2;
    double a_sq = (rho_1 * a1_sq * rho_2 * a2_sq) /
                  (rho * (alpha_2 * rho_1 * a1_sq + alpha_1 * rho_2 * a2_sq));
    if (rho < ORC_RHO_FLOOR) return 0;
    if (a_sq < ORC_PRS_FLOOR) return 0;
    if (alpha_1 < 1.0 || alpha_1 > 1.0) return 0;
    return 1;
}

/* kapila_5eqn.c:326-371 */
static double kapila_flux_prim(const orc_problem *P, const double *V,
                               double nx, double ny, double *F, int *bad,
                               double *bval)
{
    double alpha_1 = V[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho_1 = V[0];
    double rho_2 = V[1];
    double rho = alpha_1 * rho_1 + alpha_2 * rho_2;
    double u = V[2];
    double v = V[3];
    double p = V[4];
    double G1m1 = P->g1 - 1.0;
    double G2m1 = P->g2 - 1.0;
    double e = (p * (alpha_1 * G2m1 + alpha_2 * G1m1) +
                alpha_1 * P->g1 * P->p1 * G2m1 +
                alpha_2 * P->g2 * P->p2 * G1m1) /
               (G1m1 * G2m1);
    double E = e + 0.5 * rho * (u * u + v * v);
    double un = u * nx + v * ny;
    double a1_sq = P->g1 * (p + P->p1) / rho_1;
    double a2_sq = P->g2 * (p + P->p2) / rho_2;
    double a_sq = (rho_1 * a1_sq * rho_2 * a2_sq) /
                  (rho * (alpha_2 * rho_1 * a1_sq + alpha_1 * rho_2 * a2_sq));
    if (a_sq < ORC_PRS_FLOOR) setbad(bad, bval, 3, a_sq);
    F[0] = alpha_1 * rho_1 * un;
    F[1] = alpha_2 * rho_2 * un;
    F[2] = rho * u * un + p * nx;
    F[3] = rho * v * un + p * ny;
    F[4] = (E + p) * un;
    F[5] = 0.0;
    return fabs(un) + sqrt(a_sq);
}

/* kapila_5eqn.c:417-444 */
static void kapila_ncp_prim(const orc_problem *P, const double *V,
                            const double *gx, const double *gy, double *B)
{
    double alpha_1 = V[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho_1 = V[0];
    double rho_2 = V[1];
    double u = V[2];
    double v = V[3];
    double p = V[4];
    double phi_x = gx[5];
    double phi_y = gy[5];
    double u_x = gx[2];
    double v_y = gy[3];
    double a_1_sq = P->g1 * (p + P->p1) / rho_1;
    double a_2_sq = P->g2 * (p + P->p2) / rho_2;
    double K = (alpha_1 * alpha_2 * (rho_2 * a_2_sq - rho_1 * a_1_sq)) /
               (alpha_1 * rho_2 * a_2_sq + alpha_2 * rho_1 * a_1_sq);
    B[0] = 0.0;
    B[1] = 0.0;
    B[2] = 0.0;
    B[3] = 0.0;
    B[4] = 0.0;
    B[5] = u * phi_x + v * phi_y - K * (u_x + v_y);
}

/* kapila_5eqn.c:450-479 */
static int kapila_pad_prim(const orc_problem *P, const double *V)
{
    double alpha_1 = V[5];
    double alpha_2 = 1.0 - alpha_1;
    double rho_1 = V[0];
    double rho_2 = V[1];
    double rho = alpha_1 * rho_1 + alpha_2 * rho_2;
    double p = V[4];
    double a1_sq = P->g1 * (p + P->p1) / rho_1;
    double a2_sq = P->g2 * (p + P->p2) / rho_2;
    double a_sq = (rho_1 * a1_sq * rho_2 * a2_sq) /
                  (rho * (alpha_2 * rho_1 * a1_sq + alpha_1 * rho_2 * a2_sq));
    if (rho < ORC_RHO_FLOOR) return 0;
    if (a_sq < ORC_PRS_FLOOR) return 0;
    if (alpha_1 < 0.0 || alpha_1 > 1.0) return 0;
    return 1;
}

/* ------------------------------------------------------- Baer-Nunziato -- */

/* baer_nunziato.c:14-41 */
static void bn_cons2prim(const orc_problem *P, const double *Q, double *V)
{
    double phi_1 = Q[8];
    double phi_2 = 1.0 - Q[8];
    double rho_1 = Q[0] / phi_1;
    double rho_2 = Q[4] / phi_2;
    double u_1 = Q[1] / Q[0];
    double v_1 = Q[2] / Q[0];
    double u_2 = Q[5] / Q[4];
    double v_2 = Q[6] / Q[4];
    double E_1 = Q[3] / phi_1;
    double E_2 = Q[7] / phi_2;
    double p_1 = (P->g1 - 1.0) * (E_1 - 0.5 * rho_1 * (u_1 * u_1 + v_1 * v_1)) -
                 P->g1 * P->p1;
    double p_2 = (P->g2 - 1.0) * (E_2 - 0.5 * rho_2 * (u_2 * u_2 + v_2 * v_2)) -
                 P->g2 * P->p2;
    V[0] = rho_1;
    V[1] = u_1;
    V[2] = v_1;
    V[3] = p_1;
    V[4] = rho_2;
    V[5] = u_2;
    V[6] = v_2;
    V[7] = p_2;
    V[8] = phi_1;
}

/* baer_nunziato.c:47-60 */
static void bn_prim2cons(const orc_problem *P, const double *V, double *Q)
{
    Q[0] = V[8] * V[0];
    Q[1] = Q[0] * V[1];
    Q[2] = Q[0] * V[2];
    Q[3] = Q[0] * (0.5 * (V[1] * V[1] + V[2] * V[2]) +
                   (V[3] + P->g1 * P->p1) / (V[0] * (P->g1 - 1.0)));
    Q[4] = (1.0 - V[8]) * V[4];
    Q[5] = Q[4] * V[5];
    Q[6] = Q[4] * V[6];
    Q[7] = Q[4] * (0.5 * (V[5] * V[5] + V[6] * V[6]) +
                   (V[7] + P->g2 * P->p2) / (V[4] * (P->g2 - 1.0)));
    Q[8] = V[8];
}

/* baer_nunziato.c:66-131 */
static double bn_flux(const orc_problem *P, const double *Q, double nx,
                      double ny, double *F, int *bad, double *bval)
{
    double phi_1 = Q[8];
    double phi_2 = 1.0 - Q[8];
    double rho_1 = Q[0] / phi_1;
    double rho_2 = Q[4] / phi_2;
    double u_1 = Q[1] / Q[0];
    double v_1 = Q[2] / Q[0];
    double u_2 = Q[5] / Q[4];
    double v_2 = Q[6] / Q[4];
    double E_1 = Q[3] / phi_1;
    double E_2 = Q[7] / phi_2;
    double p_1 = (P->g1 - 1.0) * (E_1 - 0.5 * rho_1 * (u_1 * u_1 + v_1 * v_1)) -
                 P->g1 * P->p1;
    double p_2 = (P->g2 - 1.0) * (E_2 - 0.5 * rho_2 * (u_2 * u_2 + v_2 * v_2)) -
                 P->g2 * P->p2;
    double un_1 = nx * u_1 + ny * v_1;
    double un_2 = nx * u_2 + ny * v_2;
    if (rho_1 < ORC_RHO_FLOOR) setbad(bad, bval, 4, rho_1);
    if (rho_2 < ORC_RHO_FLOOR) setbad(bad, bval, 5, rho_2);
    if ((p_1 + P->p1) < ORC_PRS_FLOOR) setbad(bad, bval, 6, p_1 + P->p1);
    if ((p_2 + P->p2) < ORC_PRS_FLOOR) setbad(bad, bval, 7, p_2 + P->p2);
    F[0] = phi_1 * rho_1 * un_1;
    F[1] = phi_1 * (rho_1 * u_1 * un_1 + nx * p_1);
    F[2] = phi_1 * (rho_1 * v_1 * un_1 + ny * p_1);
    F[3] = phi_1 * un_1 * (E_1 + p_1);
    F[4] = phi_2 * rho_2 * un_2;
    F[5] = phi_2 * (rho_2 * u_2 * un_2 + nx * p_2);
    F[6] = phi_2 * (rho_2 * v_2 * un_2 + ny * p_2);
    F[7] = phi_2 * un_2 * (E_2 + p_2);
    F[8] = 0.0;
    double s_1 = fabs(un_1) + sqrt(P->g1 * (p_1 + P->p1) / rho_1);
    double s_2 = fabs(un_2) + sqrt(P->g2 * (p_2 + P->p2) / rho_2);
    return pmax(s_1, s_2);
}

/* baer_nunziato.c:137-165 (u_I = u_1, p_I = p_2) */
static void bn_ncp(const orc_problem *P, const double *Q, const double *gx,
                   const double *gy, double *B)
{
    double phi_2 = 1.0 - Q[8];
    double rho_2 = Q[4] / phi_2;
    double u_1 = Q[1] / Q[0];
    double v_1 = Q[2] / Q[0];
    double u_2 = Q[5] / Q[4];
    double v_2 = Q[6] / Q[4];
    double E_2 = Q[7] / phi_2;
    double p_I = (P->g2 - 1.0) * (E_2 - 0.5 * rho_2 * (u_2 * u_2 + v_2 * v_2)) -
                 P->g2 * P->p2;
    double phi_x = gx[8];
    double phi_y = gy[8];
    B[0] = 0.0;
    B[1] = -p_I * phi_x;
    B[2] = -p_I * phi_y;
    B[3] = -p_I * (u_1 * phi_x + v_1 * phi_y);
    B[4] = 0.0;
    B[5] = p_I * phi_x;
    B[6] = p_I * phi_y;
    B[7] = p_I * (u_1 * phi_x + v_1 * phi_y);
    B[8] = u_1 * phi_x + v_1 * phi_y;
}

/* baer_nunziato.c:207-250 */
static int bn_pad(const orc_problem *P, const double *Q)
{
    int PAD = 1;
    double phi_1 = Q[8];
    double phi_2 = 1.0 - Q[8];
    double rho_1 = Q[0] / phi_1;
    double rho_2 = Q[4] / phi_2;
    double u_1 = Q[1] / Q[0];
    double v_1 = Q[2] / Q[0];
    double u_2 = Q[5] / Q[4];
    double v_2 = Q[6] / Q[4];
    double E_1 = Q[3] / phi_1;
    double E_2 = Q[7] / phi_2;
    double p_1 = (P->g1 - 1.0) * (E_1 - 0.5 * rho_1 * (u_1 * u_1 + v_1 * v_1)) -
                 P->g1 * P->p1;
    double p_2 = (P->g2 - 1.0) * (E_2 - 0.5 * rho_2 * (u_2 * u_2 + v_2 * v_2)) -
                 P->g2 * P->p2;
    if (rho_1 < ORC_RHO_FLOOR) PAD = 0;
    if (rho_2 < ORC_RHO_FLOOR) PAD = 0;
    if ((p_1 + P->p1) < ORC_PRS_FLOOR) PAD = 0;
    if ((p_2 + P->p2) < ORC_PRS_FLOOR) PAD = 0;
    if (phi_1 < 0.0 || phi_1 > 1.0) PAD = 0;
    return PAD;
}

/* baer_nunziato.c:256-321 (p_2 = V[8] at :271 unless the quirk is off) */
static double bn_flux_prim(const orc_problem *P, const double *V, double nx,
                           double ny, double *F, int *bad, double *bval)
{
    double phi_1 = V[8];
    double phi_2 = 1.0 - V[8];
    double rho_1 = V[0];
    double rho_2 = V[4];
    double u_1 = V[1];
    double v_1 = V[2];
    double u_2 = V[5];
    double v_2 = V[6];
    double p_1 = V[3];
    double p_2 = (P->quirks & ORC_QUIRK_BN_P2_IS_PHI) ? V[8] : V[7];
    double E_1 = (p_1 + P->g1 * P->p1) / (P->g1 - 1.0) +
                 0.5 * rho_1 * (u_1 * u_1 + v_1 * v_1);
    double E_2 = (p_2 + P->g2 * P->p2) / (P->g2 - 1.0) +
                 0.5 * rho_2 * (u_2 * u_2 + v_2 * v_2);
    double un_1 = nx * u_1 + ny * v_1;
    double un_2 = nx * u_2 + ny * v_2;
    if (rho_1 < ORC_RHO_FLOOR) setbad(bad, bval, 4, rho_1);
    if (rho_2 < ORC_RHO_FLOOR) setbad(bad, bval, 5, rho_2);
    if ((p_1 + P->p1) < ORC_PRS_FLOOR) setbad(bad, bval, 6, p_1 + P->p1);
    if ((p_2 + P->p2) < ORC_PRS_FLOOR) setbad(bad, bval, 7, p_2 + P->p2);
    F[0] = phi_1 * rho_1 * un_1;
    F[1] = phi_1 * (rho_1 * u_1 * un_1 + nx * p_1);
    F[2] = phi_1 * (rho_1 * v_1 * un_1 + ny * p_1);
    F[3] = phi_1 * un_1 * (E_1 + p_1);
    F[4] = phi_2 * rho_2 * un_2;
    F[5] = phi_2 * (rho_2 * u_2 * un_2 + nx * p_2);
    F[6] = phi_2 * (rho_2 * v_2 * un_2 + ny * p_2);
    F[7] = phi_2 * un_2 * (E_2 + p_2);
    F[8] = 0.0;
    double s_1 = fabs(un_1) + sqrt(P->g1 * (p_1 + P->p1) / rho_1);
    double s_2 = fabs(un_2) + sqrt(P->g2 * (p_2 + P->p2) / rho_2);
    return pmax(s_1, s_2);
}

/* baer_nunziato.c:327-344 */
static void bn_ncp_prim(const double *V, const double *gx, const double *gy,
                        double *B)
{
    double u_I = V[1];
    double v_I = V[2];
    double p_I = V[7];
    double phi_x = gx[8];
    double phi_y = gy[8];
    B[0] = 0.0;
    B[1] = -p_I * phi_x;
    B[2] = -p_I * phi_y;
    B[3] = -p_I * (u_I * phi_x + v_I * phi_y);
    B[4] = 0.0;
    B[5] = p_I * phi_x;
    B[6] = p_I * phi_y;
    B[7] = p_I * (u_I * phi_x + v_I * phi_y);
    B[8] = u_I * phi_x + v_I * phi_y;
}

/* baer_nunziato.c:350-385 */
static int bn_pad_prim(const orc_problem *P, const double *V)
{
    int PAD = 1;
    double phi_1 = V[8];
    double rho_1 = V[0];
    double rho_2 = V[4];
    double p_1 = V[3];
    double p_2 = V[7];
    if (rho_1 < ORC_RHO_FLOOR) PAD = 0;
    if (rho_2 < ORC_RHO_FLOOR) PAD = 0;
    if ((p_1 + P->p1) < ORC_PRS_FLOOR) PAD = 0;
    if ((p_2 + P->p2) < ORC_PRS_FLOOR) PAD = 0;
    if (phi_1 < 0.0 || phi_1 > 1.0) PAD = 0;
    return PAD;
}

/* ------------------------------------------------------- Kapila inlet -- */

/* Left inlet of the Kapila system (kapila_5eqn.c:485-513; InletBC at :289-320
 * is PDEPrim2Cons of the same state): a Mach-2 water jet of parabolic
 * velocity profile in |y| < 1, air pressure 1.013, and a tanh volume-fraction
 * profile of width 3*30/500 around |y| = 1.  Independent of x and t. */
void orc_inlet_prim(const orc_problem *P, double y, double *V)
{
    const double M = 2.0;
    const double rho_air = 1.0;
    const double rho_water = 1000.0;
    const double Pr = 1.013;
    const double U_avg = M * sqrt(P->g2 * Pr / rho_air);
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

/* ------------------------------------------------------------ dispatch -- */

void orc_cons2prim(const orc_problem *P, const double *Q, double *V)
{
    switch (P->system) {
    case ORC_EULER: euler_cons2prim(P, Q, V); break;
    case ORC_KAPILA: kapila_cons2prim(P, Q, V); break;
    default: bn_cons2prim(P, Q, V); break;
    }
}

void orc_prim2cons(const orc_problem *P, const double *V, double *Q)
{
    switch (P->system) {
    case ORC_EULER: euler_prim2cons(P, V, Q); break;
    case ORC_KAPILA: kapila_prim2cons(P, V, Q); break;
    default: bn_prim2cons(P, V, Q); break;
    }
}

double orc_flux_x(const orc_problem *P, const double *Q, double nx, double ny,
                  double *F, int *bad, double *bval)
{
    switch (P->system) {
    case ORC_EULER: return euler_flux(P, Q, nx, ny, F, bad, bval);
    case ORC_KAPILA: return kapila_flux(P, Q, nx, ny, F, bad, bval);
    default: return bn_flux(P, Q, nx, ny, F, bad, bval);
    }
}

double orc_flux_prim_x(const orc_problem *P, const double *V, double nx,
                       double ny, double *F, int *bad, double *bval)
{
    switch (P->system) {
    case ORC_EULER: return euler_flux_prim(P, V, nx, ny, F, bad, bval);
    case ORC_KAPILA: return kapila_flux_prim(P, V, nx, ny, F, bad, bval);
    default: return bn_flux_prim(P, V, nx, ny, F, bad, bval);
    }
}

double orc_flux(const orc_problem *P, const double *Q, double nx, double ny,
                double *F, int *bad)
{
    if (bad) *bad = 0;
    return orc_flux_x(P, Q, nx, ny, F, bad, 0);
}

double orc_flux_prim(const orc_problem *P, const double *V, double nx,
                     double ny, double *F, int *bad)
{
    if (bad) *bad = 0;
    return orc_flux_prim_x(P, V, nx, ny, F, bad, 0);
}

/* euler.c:142-148 (NCP of the Euler system is identically zero) */
void orc_ncp(const orc_problem *P, const double *Q, const double *gx,
             const double *gy, double *B)
{
    switch (P->system) {
    case ORC_EULER: B[0] = B[1] = B[2] = B[3] = 0.0; break;
    case ORC_KAPILA: kapila_ncp(P, Q, gx, gy, B); break;
    default: bn_ncp(P, Q, gx, gy, B); break;
    }
}

/* euler.c:278-286 */
void orc_ncp_prim(const orc_problem *P, const double *V, const double *gx,
                  const double *gy, double *B)
{
    switch (P->system) {
    case ORC_EULER: B[0] = B[1] = B[2] = B[3] = 0.0; break;
    case ORC_KAPILA: kapila_ncp_prim(P, V, gx, gy, B); break;
    default: bn_ncp_prim(V, gx, gy, B); break;
    }
}

int orc_pad(const orc_problem *P, const double *Q)
{
    switch (P->system) {
    case ORC_EULER: return euler_pad(P, Q);
    case ORC_KAPILA: return kapila_pad(P, Q);
    default: return bn_pad(P, Q);
    }
}

int orc_pad_prim(const orc_problem *P, const double *V)
{
    switch (P->system) {
    case ORC_EULER: return euler_pad_prim(V);
    case ORC_KAPILA: return kapila_pad_prim(P, V);
    default: return bn_pad_prim(P, V);
    }
}

/* riemann.c:52-105 (conservative) and riemann.c:112-140 (primitive):
 * Rusanov flux + path-conservative jump D = B(mean)*(R-L). */
double orc_riemann_x(const orc_problem *P, const double *L, const double *R,
                     double nx, double ny, double *F, double *D, int *bad,
                     double *bval)
{
    double FL[ORC_MAXVAR], FR[ORC_MAXVAR], gx[ORC_MAXVAR], gy[ORC_MAXVAR];
    double av[ORC_MAXVAR];
    int nv = orc_nvar(P->system), c;
    if (P->prim) {
        double QL[ORC_MAXVAR], QR[ORC_MAXVAR];
        orc_prim2cons(P, L, QL);
        orc_prim2cons(P, R, QR);
        double sl = orc_flux_prim_x(P, L, nx, ny, FL, bad, bval);
        double sr = orc_flux_prim_x(P, R, nx, ny, FR, bad, bval);
        double s = pmax(sl, sr);
        for (c = 0; c < nv; ++c) {
            av[c] = 0.5 * (R[c] + L[c]);
            gx[c] = nx * (R[c] - L[c]);
            gy[c] = ny * (R[c] - L[c]);
            F[c] = 0.5 * (FR[c] + FL[c] - s * (QR[c] - QL[c]));
        }
        orc_ncp_prim(P, av, gx, gy, D);
        return s;
    } else {
        double sl = orc_flux_x(P, L, nx, ny, FL, bad, bval);
        double sr = orc_flux_x(P, R, nx, ny, FR, bad, bval);
        double s = pmax(sl, sr);
        for (c = 0; c < nv; ++c) {
            av[c] = 0.5 * (R[c] + L[c]);
            gx[c] = nx * (R[c] - L[c]);
            gy[c] = ny * (R[c] - L[c]);
            F[c] = 0.5 * (FR[c] + FL[c] - s * (R[c] - L[c]));
        }
        orc_ncp(P, av, gx, gy, D);
        return s;
    }
}

double orc_riemann(const orc_problem *P, const double *L, const double *R,
                   double nx, double ny, double *F, double *D, int *bad)
{
    if (bad) *bad = 0;
    return orc_riemann_x(P, L, R, nx, ny, F, D, bad, 0);
}
