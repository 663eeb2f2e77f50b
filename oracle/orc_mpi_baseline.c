/*
 * orc_mpi_baseline.c -- the oracle's RHS and TSSSP step run with MPI ranks,
 * the way the reference runs its PETSc path: a DMDA decomposition with
 * width-3 halos (hype.c:108-120), one halo exchange per RHS call
 * (DMGlobalToLocal, rhs_function_prim.c:44-45) and MPI_Allreduce(MIN) of dt
 * (rhs_function_prim.c:617-623).
 *
 * TEST INFRASTRUCTURE ONLY: the benchmark's CPU-baseline leg starts it under
 * mpiexec to time the oracle on the host cores.  Nothing in the product links
 * or runs it.
 *
 * Rows are split in blocks over the ranks (1D decomposition in y).  Each RHS
 * call swaps 3 rows with each neighbour and evaluates orc_rhs on the block
 * plus its halo rows; the stencil of an owned row reaches 3 rows, so the
 * owned rows see only real data.  The sub-grid's outer rows at a rank
 * boundary get transmissive ghosts, whose faces also enter that rank's CFL
 * maximum, so dt can differ from the one-rank run.  Global y boundaries keep
 * their kinds; periodic ones are not supported.
 *
 * usage: orc_mpi_baseline STATE system N prim nx ny x_min y_min h cfl
 *                         bc0 bc1 bc2 bc3 g1 g2 p1 p2 quirks budget_s
 *                         [steps [OUT]]
 *   STATE: the initial state, nx*ny*nvar doubles, AoS [j][i][c]
 *   steps: time exactly this many steps after the first (default: as many
 *          as fit budget_s, 1 to 60, from the timed first step)
 *   OUT:   rank 0 writes the final state there (for the decomposition test)
 * Rank 0 prints "ranks R steps K seconds S".
 */
#include <math.h>
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define HALO 3

typedef struct {
    orc_problem sub; /* the rank's block plus its halo rows */
    long row;        /* doubles per grid row */
    int own;         /* owned rows */
    int lo, hi;      /* halo rows below and above (0 at a global boundary) */
    int rank;
    double *ext, *rext;
} block;

static int rhs_cb(void *vc, double t, const double *U, double *F, double *dt_out) {
    block *b = (block *)vc;
    const long row = b->row;
    MPI_Request rq[4];
    int nr = 0;
    memcpy(b->ext + (long)b->lo * row, U, sizeof(double) * row * b->own);
    if (b->lo) { /* neighbour below: its top rows come in, my bottom rows go out */
        MPI_Irecv(b->ext, (int)(HALO * row), MPI_DOUBLE, b->rank - 1, 1, MPI_COMM_WORLD, &rq[nr++]);
        MPI_Isend((void *)U, (int)(HALO * row), MPI_DOUBLE, b->rank - 1, 0, MPI_COMM_WORLD, &rq[nr++]);
    }
    if (b->hi) { /* neighbour above */
        MPI_Irecv(b->ext + (long)(b->lo + b->own) * row, (int)(HALO * row), MPI_DOUBLE, b->rank + 1, 0,
                  MPI_COMM_WORLD, &rq[nr++]);
        MPI_Isend((void *)(U + (long)(b->own - HALO) * row), (int)(HALO * row), MPI_DOUBLE, b->rank + 1, 1,
                  MPI_COMM_WORLD, &rq[nr++]);
    }
    MPI_Status stat[4];
    MPI_Waitall(nr, rq, stat);
    double dt = INFINITY;
    orc_error err;
    orc_rhs(&b->sub, t, b->ext, b->rext, &dt, NULL, NULL, &err); /* timing only: status ignored */
    memcpy(F, b->rext + (long)b->lo * row, sizeof(double) * row * b->own);
    MPI_Allreduce(&dt, dt_out, 1, MPI_DOUBLE, MPI_MIN, MPI_COMM_WORLD);
    return 0;
}

int main(int argc, char **argv) {
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    if (argc < 21) {
        if (rank == 0) fprintf(stderr, "usage: see orc_mpi_baseline.c\n");
        MPI_Finalize();
        return 2;
    }
    orc_problem P;
    memset(&P, 0, sizeof(P));
    P.system = atoi(argv[2]);
    P.N = atoi(argv[3]);
    P.prim = atoi(argv[4]);
    P.nx = atoi(argv[5]);
    P.ny = atoi(argv[6]);
    P.x_min = atof(argv[7]);
    P.y_min = atof(argv[8]);
    P.h = atof(argv[9]);
    P.cfl = atof(argv[10]);
    for (int k = 0; k < 4; ++k) P.bc[k] = atoi(argv[11 + k]);
    P.g1 = atof(argv[15]);
    P.g2 = atof(argv[16]);
    P.p1 = atof(argv[17]);
    P.p2 = atof(argv[18]);
    P.quirks = (unsigned)strtoul(argv[19], NULL, 0);
    const double budget = atof(argv[20]);
    const int fixed_steps = argc > 21 ? atoi(argv[21]) : 0;
    const int nv = orc_nvar(P.system);
    const int base = P.ny / size, rem = P.ny % size;
    if (nv <= 0 || base < HALO || P.bc[2] == ORC_BC_PERIODIC || P.bc[3] == ORC_BC_PERIODIC) {
        if (rank == 0) fprintf(stderr, "unsupported decomposition\n");
        MPI_Finalize();
        return 2;
    }
    block b;
    b.rank = rank;
    b.row = (long)P.nx * nv;
    b.own = base + (rank < rem ? 1 : 0);
    const int j0 = rank * base + (rank < rem ? rank : rem);
    b.lo = rank > 0 ? HALO : 0;
    b.hi = rank < size - 1 ? HALO : 0;
    b.sub = P;
    b.sub.ny = b.own + b.lo + b.hi;
    b.sub.y_min = P.y_min + (double)(j0 - b.lo) * P.h;
    if (b.lo) b.sub.bc[2] = ORC_BC_TRANSMISSIVE;
    if (b.hi) b.sub.bc[3] = ORC_BC_TRANSMISSIVE;
    const long n = b.row * b.own;
    double *U = malloc(sizeof(double) * n);
    double *F = malloc(sizeof(double) * n);
    double *work = malloc(sizeof(double) * 3 * n);
    b.ext = malloc(sizeof(double) * b.row * b.sub.ny);
    b.rext = malloc(sizeof(double) * b.row * b.sub.ny);
    FILE *fp = fopen(argv[1], "rb");
    int ok = U && F && work && b.ext && b.rext && fp && fseek(fp, (long)j0 * b.row * (long)sizeof(double), SEEK_SET) == 0 &&
             fread(U, sizeof(double), (size_t)n, fp) == (size_t)n;
    if (fp) fclose(fp);
    int all_ok = 0;
    MPI_Allreduce(&ok, &all_ok, 1, MPI_INT, MPI_MIN, MPI_COMM_WORLD);
    if (!all_ok) {
        if (rank == 0) fprintf(stderr, "cannot read the state\n");
        MPI_Finalize();
        return 1;
    }
    /* priming call (hype.c:343), then steps with the lagged dt */
    double dt, dt_last, t = 0.0;
    rhs_cb(&b, t, U, F, &dt);
    MPI_Barrier(MPI_COMM_WORLD);
    double t0 = MPI_Wtime();
    orc_ssp_step(P.N, n, U, t, dt, rhs_cb, &b, work, &dt_last);
    t += dt;
    dt = dt_last;
    MPI_Barrier(MPI_COMM_WORLD);
    const double t1 = MPI_Wtime() - t0;
    int k = fixed_steps;
    if (k <= 0) {
        k = (int)lround(budget / (t1 > 1e-6 ? t1 : 1e-6));
        k = k < 1 ? 1 : (k > 60 ? 60 : k);
    }
    MPI_Bcast(&k, 1, MPI_INT, 0, MPI_COMM_WORLD);
    MPI_Barrier(MPI_COMM_WORLD);
    t0 = MPI_Wtime();
    for (int s = 0; s < k; ++s) {
        orc_ssp_step(P.N, n, U, t, dt, rhs_cb, &b, work, &dt_last);
        t += dt;
        dt = dt_last;
    }
    MPI_Barrier(MPI_COMM_WORLD);
    const double tk = MPI_Wtime() - t0;
    if (rank == 0) printf("ranks %d steps %d seconds %.6f\n", size, k, tk);
    if (argc > 22) { /* gather the final state on rank 0 */
        int *cnt = malloc(sizeof(int) * size), *off = malloc(sizeof(int) * size);
        double *all = rank == 0 ? malloc(sizeof(double) * b.row * P.ny) : NULL;
        for (int r = 0; r < size; ++r) {
            cnt[r] = (int)(b.row * (base + (r < rem ? 1 : 0)));
            off[r] = (int)(b.row * (r * base + (r < rem ? r : rem)));
        }
        MPI_Gatherv(U, (int)n, MPI_DOUBLE, all, cnt, off, MPI_DOUBLE, 0, MPI_COMM_WORLD);
        if (rank == 0) {
            FILE *fo = fopen(argv[22], "wb");
            if (fo) {
                fwrite(all, sizeof(double), (size_t)(b.row * P.ny), fo);
                fclose(fo);
            }
            free(all);
        }
        free(cnt);
        free(off);
    }
    free(U);
    free(F);
    free(work);
    free(b.ext);
    free(b.rext);
    MPI_Finalize();
    return 0;
}
