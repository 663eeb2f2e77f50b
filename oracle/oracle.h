/*
 * oracle.h -- FP64 CPU oracle for the HyPE-2D finite-volume right-hand side.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * RHS callbacks (HyPE-2D/rhs_function_prim.c, HyPE-2D/rhs_function.c), their
 * point functions (euler.c, kapila_5eqn.c, baer_nunziato.c, reconstruction.c,
 * riemann.c) and of the PETSc TSSSP stage arithmetic the reference drives them
 * with (hype.c:337-384).  It keeps the reference's pass structure (primitive
 * pass, ghosted local array with boundary fill, reconstruction into face
 * scratch, x-face and y-face Riemann passes, assembly, CFL) and its expression
 * order, so that results can be compared bit for bit with the device path.
 *
 * Only tests, the smoke check and the benchmark's CPU-baseline leg may call
 * it, and only as the checker.  Nothing in the product links or loads it.
 *
 * Pinning: see the tests under tests/test_oracle_*.py.  Each function below is
 * checked against closed forms, invariants or special cases that fix it
 * (polynomial exactness, free-stream preservation, discrete conservation,
 * mirror/transpose symmetry, round trips, linear-stability polynomials).
 * Parity unpinned: the PETSc arithmetic order of VecAXPY/VecAXPBYPCZ inside
 * TSSSP (PETSc is not vendored and no reference test pins it) -- restated
 * from PETSc's published TSSSP algorithms (Ketcheson 2008, pseudocode 2/3).
 */
#ifndef HYPE_ORACLE_H
#define HYPE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* PDE systems (compile-time #define in hype.h:33-35; runtime here) */
enum { ORC_EULER = 0, ORC_KAPILA = 1, ORC_BN = 2 };

/* boundary kinds, same order as enum bndry_type (hype.h:76) */
enum { ORC_BC_WALL = 0, ORC_BC_INLET = 1, ORC_BC_PERIODIC = 2,
       ORC_BC_REFLECTIVE = 3, ORC_BC_TRANSMISSIVE = 4 };

/* error codes (shared numbering with the product C ABI) */
enum { ORC_OK = 0, ORC_E_INVALID = 1, ORC_E_UNSUPPORTED = 2,
       ORC_E_INADMISSIBLE = 5 };

/* quirk bits */
#define ORC_QUIRK_BN_P2_IS_PHI 0x1u /* baer_nunziato.c:271 p_2 = V[8] */

typedef struct {
    int system;            /* ORC_EULER / ORC_KAPILA / ORC_BN */
    int N;                 /* degree of approximation 0..3 (AppCtx.N) */
    int prim;              /* 1: RHSFunctionPrim, 0: RHSFunction */
    int nx, ny;            /* AppCtx.N_x, N_y */
    double x_min, y_min;   /* lower-left corner of the domain */
    double h;              /* AppCtx.h = (x_max-x_min)/N_x (hype.c:57) */
    double cfl;            /* AppCtx.CFL */
    int bc[4];             /* left, right, bottom, top */
    double g1, g2, p1, p2; /* GAMMA_1, GAMMA_2, PI_1, PI_2 (hype.h:42-45) */
    unsigned quirks;
} orc_problem;

typedef struct {
    int code;              /* ORC_OK or ORC_E_INADMISSIBLE */
    int kind;              /* which check fired (see orc_physics.c) */
    double x, y;           /* face coordinates, as printed by the reference */
    double value;          /* the offending value, as printed */
    long index;            /* smallest face key (deterministic choice) */
} orc_error;

int orc_nvar(int system);

/* One RHS evaluation: U (nx*ny*nvar, AoS [j][i][c]) -> RHS (same layout),
 * *dt_out = CFL*h/(2*s_max).  W (optional) receives the primitive pass
 * (Ctx->W side effect of RHSFunctionPrim).  fallback (optional, nx*ny bytes)
 * receives 1 where the admissibility check replaced the reconstruction of an
 * owned cell.  Returns ORC_OK, ORC_E_INADMISSIBLE (err filled; the RHS is
 * still computed, where the reference would have called MPI_Abort), or
 * ORC_E_UNSUPPORTED / ORC_E_INVALID. */
int orc_rhs(const orc_problem *P, double t, const double *U, double *RHS,
            double *dt_out, double *W, unsigned char *fallback,
            orc_error *err);

/* generic RHS callback for the time stepper */
typedef int (*orc_rhs_fn)(void *ctx, double t, const double *U, double *F,
                          double *dt_out);

/* One time step of the scheme PETSc uses for order N (hype.c:362-382):
 * N=0 TSEULER, N=1 TSSSPRKS2 (2 stages), N=2 TSSSPRKS3 (4 stages),
 * N=3 TSSSPRK104.  U is advanced in place; work holds 3*n doubles;
 * *dt_last gets the dt written by the last RHS call of the step. */
int orc_ssp_step(int N, long n, double *U, double t0, double dt,
                 orc_rhs_fn f, void *ctx, double *work, double *dt_last);

/* Whole solve: priming RHS call (hype.c:343), then steps with the dt lag of
 * MonitorFunction (post_process.c:19).  Stops after max_steps steps or when
 * t >= t_end.  clip=1 shortens a step that would pass t_end.  dts (optional,
 * max_steps entries) receives the dt of every step taken. */
int orc_solve(const orc_problem *P, double *U, double t0, double t_end,
              int max_steps, int clip, int *steps_done, double *t_out,
              double *dts, orc_error *err);

/* Point functions exported for the pinning tests. */
void orc_cons2prim(const orc_problem *P, const double *Q, double *V);
void orc_prim2cons(const orc_problem *P, const double *V, double *Q);
/* primitive state of the Kapila left inlet at height y (kapila_5eqn.c:485-513) */
void orc_inlet_prim(const orc_problem *P, double y, double *V);
double orc_flux(const orc_problem *P, const double *Q, double nx, double ny,
                double *F, int *bad);
double orc_flux_prim(const orc_problem *P, const double *V, double nx,
                     double ny, double *F, int *bad);
void orc_ncp(const orc_problem *P, const double *Q, const double *gx,
             const double *gy, double *B);
void orc_ncp_prim(const orc_problem *P, const double *V, const double *gx,
                  const double *gy, double *B);
int orc_pad(const orc_problem *P, const double *Q);
int orc_pad_prim(const orc_problem *P, const double *V);
double orc_riemann(const orc_problem *P, const double *L, const double *R,
                   double nx, double ny, double *F, double *D, int *bad);
void orc_reconstruct(const double *ux, const double *uy, const double *uxy,
                     int N, double *coeffs);
double orc_basis(double x, double y, int k);
double orc_minmod(double a, double b);

#ifdef __cplusplus
}
#endif
#endif
