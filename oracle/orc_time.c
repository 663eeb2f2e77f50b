/*
 * orc_time.c -- SSP Runge-Kutta steps and the solve loop (oracle; test
 * infrastructure only).
 *
 * The reference delegates time integration to PETSc TSSSP (hype.c:362-382):
 *   N=0 TSEULER, N=1 TSSSPRKS2 with 2 stages, N=2 TSSSPRKS3 with 4 stages,
 *   N=3 TSSSPRK104.
 * PETSc is not vendored; these are restated from PETSc's published TSSSP
 * algorithms (Ketcheson 2008, pseudocode 2 for RKS2/RKS3 and pseudocode 3 for
 * RK104), with PETSc's vector kernels evaluated in their source order:
 *   VecAXPY(y,a,x):            y = y + a*x
 *   VecAYPX(y,b,x):            y = x + b*y
 *   VecAXPBY(y,a,b,x):         y = a*x + b*y
 *   VecAXPBYPCZ(z,a,b,c,x,y):  z = a*x + b*y + c*z   (c*z dropped when c==0)
 * The dt of each step is the value the last RHS call of the previous step
 * wrote (MonitorFunction, post_process.c:19); the first step uses the
 * priming call of hype.c:343.
 */
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include "orc_internal.h"

void *orc_ctx_create(const orc_problem *P);
void orc_ctx_destroy(void *C);
int orc_rhs_ctx(void *vc, double t, const double *U, double *RHS,
                double *dt_out, double *W, unsigned char *fallback,
                orc_error *err);

static void axpy(long n, double *y, double a, const double *x)
{
    long i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) y[i] = y[i] + a * x[i];
}

static void aypx(long n, double *y, double b, const double *x)
{
    long i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) y[i] = x[i] + b * y[i];
}

static void axpby(long n, double *y, double a, double b, const double *x)
{
    long i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) y[i] = a * x[i] + b * y[i];
}

static void axpbypcz(long n, double *z, double a, double b, double c,
                     const double *x, const double *y)
{
    long i;
    if (c == 0.0) {
#pragma omp parallel for schedule(static)
        for (i = 0; i < n; ++i) z[i] = a * x[i] + b * y[i];
    } else {
#pragma omp parallel for schedule(static)
        for (i = 0; i < n; ++i) z[i] = a * x[i] + b * y[i] + c * z[i];
    }
}

int orc_ssp_step(int N, long n, double *U, double t0, double dt,
                 orc_rhs_fn f, void *ctx, double *work, double *dt_last)
{
    double *w0 = work, *w1 = work + n, *F = work + 2 * n;
    int rc = ORC_OK, r1;
    size_t bytes = (size_t)n * sizeof(double);
#define CALL(tt, X)                                   \
    do {                                              \
        r1 = f(ctx, (tt), (X), F, dt_last);           \
        if (r1 != ORC_OK && rc == ORC_OK) rc = r1;    \
        if (r1 != ORC_OK && r1 != ORC_E_INADMISSIBLE) \
            return r1;                                \
    } while (0)

    if (N == 0) {
        /* TSEULER: update = F(U); update = U + dt*update; U = update */
        CALL(t0, U);
        aypx(n, F, dt, U);
        memcpy(U, F, bytes);
    } else if (N == 1) {
        /* TSSSPRKS2, s = 2 */
        const int s = 2;
        int i;
        memcpy(w0, U, bytes);
        for (i = 0; i < s - 1; i++) {
            double stage_time = t0 + dt * (i / (s - 1.));
            CALL(stage_time, w0);
            axpy(n, w0, dt / (s - 1.), F);
        }
        CALL(t0 + dt, w0);
        axpbypcz(n, U, (s - 1.) / s, dt / s, 1. / s, w0, F);
    } else if (N == 2) {
        /* TSSSPRKS3, s = 4: n = 2, r = 2 */
        const int s = 4, nn = 2, r = s - nn;
        int i = 0;
        memcpy(w0, U, bytes);
        for (i = 0; i < (nn - 1) * (nn - 2) / 2; i++) {
            double c = (i < nn * (nn + 1) / 2) ? 1. * i / (s - nn) : (1. * i - nn) / (s - nn);
            CALL(t0 + c * dt, w0);
            axpy(n, w0, dt / r, F);
        }
        memcpy(w1, w0, bytes);
        for (; i < nn * (nn + 1) / 2 - 1; i++) {
            double c = (i < nn * (nn + 1) / 2) ? 1. * i / (s - nn) : (1. * i - nn) / (s - nn);
            CALL(t0 + c * dt, w0);
            axpy(n, w0, dt / r, F);
        }
        {
            double c = (i < nn * (nn + 1) / 2) ? 1. * i / (s - nn) : (1. * i - nn) / (s - nn);
            CALL(t0 + c * dt, w0);
            axpbypcz(n, w0, 1. * nn / (2 * nn - 1), (nn - 1.) * dt / (r * (2 * nn - 1)),
                     (nn - 1.) / (2 * nn - 1), w1, F);
            i++;
        }
        for (; i < s; i++) {
            double c = (i < nn * (nn + 1) / 2) ? 1. * i / (s - nn) : (1. * i - nn) / (s - nn);
            CALL(t0 + c * dt, w0);
            axpy(n, w0, dt / r, F);
        }
        memcpy(U, w0, bytes);
    } else if (N == 3) {
        /* TSSSPRK104 */
        const double c[10] = {0, 1. / 6, 2. / 6, 3. / 6, 4. / 6, 2. / 6, 3. / 6, 4. / 6, 5. / 6, 1};
        int i;
        memcpy(w0, U, bytes);
        for (i = 0; i < 5; i++) {
            CALL(t0 + c[i] * dt, w0);
            axpy(n, w0, dt / 6, F);
        }
        axpbypcz(n, w1, 1. / 25, 9. / 25, 0, U, w0);
        axpby(n, w0, 15, -5, w1);
        for (; i < 9; i++) {
            CALL(t0 + c[i] * dt, w0);
            axpy(n, w0, dt / 6, F);
        }
        CALL(t0 + dt, w0);
        axpbypcz(n, w1, 3. / 5, dt / 10, 1, w0, F);
        memcpy(U, w1, bytes);
    } else {
        return ORC_E_INVALID;
    }
#undef CALL
    return rc;
}

typedef struct {
    void *C;
    orc_error *err;
} solve_ctx;

static int solve_rhs(void *vctx, double t, const double *U, double *F,
                     double *dt_out)
{
    solve_ctx *S = (solve_ctx *)vctx;
    orc_error e;
    int rc = orc_rhs_ctx(S->C, t, U, F, dt_out, NULL, NULL, &e);
    if (rc == ORC_E_INADMISSIBLE && S->err && S->err->code == ORC_OK) *S->err = e;
    return rc;
}

int orc_solve(const orc_problem *P, double *U, double t0, double t_end,
              int max_steps, int clip, int *steps_done, double *t_out,
              double *dts, orc_error *err)
{
    long n = (long)P->nx * P->ny * orc_nvar(P->system);
    void *C = orc_ctx_create(P);
    double *work, *F, dt = 0.0, t = t0;
    int steps = 0, rc;
    solve_ctx S;
    if (!C) return ORC_E_INVALID;
    work = (double *)malloc((size_t)3 * n * sizeof(double));
    F = (double *)malloc((size_t)n * sizeof(double));
    if (!work || !F) { free(work); free(F); orc_ctx_destroy(C); return ORC_E_INVALID; }
    if (err) memset(err, 0, sizeof(*err));
    S.C = C;
    S.err = err;
    /* priming call (hype.c:343): only dt is used */
    rc = solve_rhs(&S, t0, U, F, &dt);
    while (rc == ORC_OK && steps < max_steps && t < t_end) {
        double dt_step = dt, dt_next = dt;
        if (clip && t + dt_step > t_end) dt_step = t_end - t;
        rc = orc_ssp_step(P->N, n, U, t, dt_step, solve_rhs, &S, work, &dt_next);
        if (dts) dts[steps] = dt_step;
        t = t + dt_step;
        dt = dt_next;
        steps++;
    }
    if (steps_done) *steps_done = steps;
    if (t_out) *t_out = t;
    free(work);
    free(F);
    orc_ctx_destroy(C);
    return rc;
}
