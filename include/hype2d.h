/*
 * hype2d.h -- C ABI of the MI355X-native HyPE-2D finite-volume right-hand side.
 *
 * The device RHS is a drop-in for the reference's PETSc TSRHSFunction callbacks
 *   PetscErrorCode RHSFunctionPrim(TS, PetscReal t, Vec U, Vec RHS, void *ctx)
 *   PetscErrorCode RHSFunction    (TS, PetscReal t, Vec U, Vec RHS, void *ctx)
 * (HyPE-2D/hype.h:305-306, rhs_function_prim.c:12, rhs_function.c:12) on the
 * same DMDA cell-centred layout: AoS records of nvar doubles, row-major
 * [j][i][c] over the owned nx*ny cells (hype.h:68-70, hype.c:108-120).  The
 * ghost frame of width 3 is synthesised on the device from the boundary kinds
 * (replacing DMGlobalToLocal and the boundary loops, rhs_function_prim.c:44-315).
 *
 * All arithmetic is IEEE FP64.  Plain pointers and sizes only; no HIP or torch
 * types appear in this header.  One context per host thread; device work of a
 * context runs on its own HIP stream.
 */
#ifndef HYPE2D_H
#define HYPE2D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define HYPE_OK 0
#define HYPE_E_INVALID 1      /* bad argument or problem description */
#define HYPE_E_UNSUPPORTED 2  /* combination the reference leaves undefined */
#define HYPE_E_DEVICE 3       /* HIP runtime failure */
#define HYPE_E_NOMEM 4        /* device or host allocation failed */
#define HYPE_E_INADMISSIBLE 5 /* replaces MPI_Abort on an inadmissible state
                                 (kapila_5eqn.c:95-99,353-357; euler.c:60-70,
                                 212-222; baer_nunziato.c:91-113,281-303) */
#define HYPE_E_BUSY 6         /* server: no slot or HBM for the request yet */
#define HYPE_E_NOTFOUND 7     /* server: unknown request id */
#define HYPE_E_TOO_LARGE 8    /* hype_rhs_noalloc: the grid's tables exceed the
                                 library's fixed device scratch */

/* ---- PDE systems (compile-time #define in hype.h:33-35, runtime here) -- */
#define HYPE_EULER 0          /* (rho, rho u, rho v, E), nvar 4 (euler.c) */
#define HYPE_KAPILA 1         /* (a1 r1, a2 r2, rho u, rho v, E, a1), nvar 6 */
#define HYPE_BAER_NUNZIATO 2  /* (p1 r1, p1 r1 u1, p1 r1 v1, p1 E1,
                                 p2 r2, ..., p1), nvar 9 */

/* ---- boundary kinds, same values as enum bndry_type (hype.h:76) ------- */
#define HYPE_BC_WALL 0         /* no implementation in the reference: rejected */
#define HYPE_BC_INLET 1        /* left boundary of the Kapila system only: the
                                  reference's fixed jet profile (InletBC /
                                  InletBCPrim, kapila_5eqn.c:289-320,485-513) */
#define HYPE_BC_PERIODIC 2
#define HYPE_BC_REFLECTIVE 3
#define HYPE_BC_TRANSMISSIVE 4

/* ---- quirk switches ---------------------------------------------------- */
/* baer_nunziato.c:271 reads p_2 = V[8] (the volume fraction) in the
 * primitive-path flux.  Set: reproduce it (reference parity).  Clear: p_2 =
 * V[7] (physical). */
#define HYPE_QUIRK_BN_P2_IS_PHI 0x1u

/* ---- memory location of caller buffers -------------------------------- */
#define HYPE_MEM_HOST 0
#define HYPE_MEM_DEVICE 1

/* ---- problem description: the fields of AppCtx the RHS reads
 *      (hype.h:203-252) ------------------------------------------------- */
typedef struct hype_problem {
    int system;            /* HYPE_EULER / HYPE_KAPILA / HYPE_BAER_NUNZIATO */
    int order;             /* AppCtx.N, 0..3 (reconstruction degree) */
    int recons_prim;       /* AppCtx.ReconsPrimitive: 1 RHSFunctionPrim, 0 RHSFunction */
    int nx, ny;            /* AppCtx.N_x, N_y (>= 3) */
    double x_min, x_max;   /* h = (x_max - x_min)/nx for both axes (hype.c:57) */
    double y_min, y_max;
    double cfl;            /* AppCtx.CFL */
    int bc[4];             /* left, right, bottom, top (HYPE_BC_*) */
    double gamma1, gamma2; /* GAMMA_1, GAMMA_2 (hype.h:42-43); Euler uses gamma1 */
    double pi1, pi2;       /* PI_1, PI_2 (hype.h:44-45) */
    unsigned quirks;       /* HYPE_QUIRK_* */
    unsigned options;      /* HYPE_OPT_* */
} hype_problem;

/* ---- numerics options --------------------------------------------------- */
/* Run the kernels built with FMA contraction: a*b+c fused wherever the
 * compiler can (about 10 % faster on MI355X).  Results stay within the
 * FP64 tolerances of the reference path (relative 1e-12 per RHS evaluation)
 * but are no longer bitwise equal to the uncontracted arithmetic; a
 * threshold decision on a value within rounding of it (admissibility, WENO
 * range guard) can then differ.  Default off: bitwise reference arithmetic. */
#define HYPE_OPT_FMA 0x1u

/* ---- tile layout ----------------------------------------------------------
 * The stage kernel sweeps tiles of up to 62 columns.  A tile runs either on
 * one wave, or split at its middle row over the two waves of a workgroup,
 * which saves one reconstructed ring row per wave.  By default
 * hype_ctx_create picks the faster layout for the grid (Kapila: split;
 * Baer-Nunziato: one wave; Euler: split for small grids, one wave when the
 * grid fills two waves per SIMD), and batches and servers use split tiles
 * for Kapila and one wave per tile otherwise.  These bits force a layout; one
 * the system's kernel is not built for is HYPE_E_UNSUPPORTED.  Results are
 * bitwise the same either way. */
#define HYPE_OPT_ONE_WAVE_TILES 0x2u
#define HYPE_OPT_SPLIT_TILES 0x4u

/* number of components of a system (the reference's nVar, hype.h:37) */
int hype_nvar(int system);

/* ======================================================================
 * Single-instance context
 * ====================================================================== */
typedef struct hype_ctx hype_ctx;

/* Allocates the device state for one grid (replaces DMDACreate2d + the
 * scratch of hype.c:108-128,235-251).  Returns HYPE_OK or an error code. */
int hype_ctx_create(const hype_problem *p, hype_ctx **out);
void hype_ctx_destroy(hype_ctx *c);

/* The RHS callback (rhs_function_prim.c:12-626 / rhs_function.c:12-625).
 * U and RHS: nx*ny*nvar doubles, AoS [j][i][c], both host or both device
 * (mem = HYPE_MEM_*).  *dt_out (may be NULL) receives CFL*h/(2*s_max), the
 * Ctx->dt side effect (rhs_function_prim.c:620-623).  t is accepted for
 * signature parity; the supported boundaries do not depend on it.
 * Returns HYPE_E_INADMISSIBLE where the reference would call MPI_Abort; the
 * RHS is still written and hype_last_error() describes the first offending
 * face (lowest face index). */
int hype_rhs(hype_ctx *c, double t, const double *U, double *RHS,
             double *dt_out, int mem);

/* Same, plus diagnostics: fallback (nx*ny bytes, host) receives 1 for every
 * owned cell whose reconstruction failed the admissibility check and fell
 * back to minmod slopes (rhs_function_prim.c:393-421). */
int hype_rhs_diag(hype_ctx *c, double t, const double *U, double *RHS,
                  double *dt_out, unsigned char *fallback, int mem);

/* hype_rhs on caller-owned device buffers, without allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It creates no context and calls no
 * device or host allocator; the instance descriptor, time state, tile table
 * and left-inlet table live in a fixed device scratch of 1 MiB per GPU that
 * the library holds until the process ends.  U and RHS: device pointers
 * (nx*ny*nvar doubles, AoS [j][i][c]); stream: a hipStream_t, or NULL for
 * the null stream.  The call waits for its stream and returns the same
 * RHS bits, *dt_out and status as hype_rhs on the same problem; on
 * HYPE_E_INADMISSIBLE, err_xy (may be NULL) receives the face coordinates
 * hype_last_error would give.  Limits: the tile table holds 16384 tiles
 * (a grid the tile-shape model would cut finer gets taller tiles, with the
 * same results; nx > 62 * 16384 = 1015808 is HYPE_E_TOO_LARGE), and the
 * left-inlet table (ny + 6) * nvar <= 81856 doubles (Kapila: ny <= 13636,
 * else HYPE_E_TOO_LARGE).  Every call shares the one scratch of its GPU and
 * leaves its contents behind; calls in one process are serialised by the
 * library, and two processes each hold their own scratch. */
int hype_rhs_noalloc(const hype_problem *p, double t, const double *U, double *RHS,
                     double *dt_out, double *err_xy, void *stream);

/* Primitive pass (ComputePrimitiveVariables, post_process.c:165-200). */
int hype_cons2prim(hype_ctx *c, const double *U, double *W, int mem);

/* ---- device-resident time integration -------------------------------- */
/* Stepping options */
#define HYPE_TS_CLIP 0x1u      /* shorten the step that would pass t_end
                                  (TS_EXACTFINALTIME_MATCHSTEP, hype.c:359) */
#define HYPE_TS_GRAPH 0x2u     /* replay captured hipGraphs per step */

typedef struct hype_status {
    double t;                  /* simulated time reached */
    double dt;                 /* dt of the last step taken */
    int steps;                 /* steps taken since hype_begin */
    int done;                  /* t >= t_end or steps >= max_steps */
    int err;                   /* HYPE_OK or HYPE_E_INADMISSIBLE */
} hype_status;

/* Copy a state in/out of the context (same layout as hype_rhs). */
int hype_set_state(hype_ctx *c, const double *U, int mem);
int hype_get_state(hype_ctx *c, double *U, int mem);

/* Start a solve on the current state: priming RHS call for dt (hype.c:343,
 * 356), time t0, stop at t_end or after max_steps steps.  The scheme is the
 * one the reference selects for order N (hype.c:362-382): N=1 SSPRK(2,2),
 * N=2 SSPRK(4,3); N=0 (TSEULER) is not stepped on the device and returns
 * HYPE_E_UNSUPPORTED (hype_rhs supports it).  dt of step k+1 comes from the last RHS
 * evaluation of step k (MonitorFunction, post_process.c:19), computed and
 * consumed on the device. */
int hype_begin(hype_ctx *c, double t0, double t_end, int max_steps,
               unsigned flags);

/* Enqueue nsteps steps (asynchronous; steps past t_end are no-ops). */
int hype_advance(hype_ctx *c, int nsteps);

/* hype_advance bracketed by HIP events on the context stream; *ms receives
 * the device time of the nsteps steps (waits for them to finish). */
int hype_advance_timed(hype_ctx *c, int nsteps, double *ms);

/* Wait for queued work and report progress. */
int hype_sync(hype_ctx *c, hype_status *st);

/* dt of every step taken so far (up to n entries). */
int hype_dt_history(hype_ctx *c, double *dts, int n);

/* Human-readable description of the last inadmissible state, with the face
 * coordinates the reference prints (x_min + i*h, y_min + j*h). */
const char *hype_last_error(hype_ctx *c, double *x, double *y);

/* Device stream of the context as an opaque handle (hipStream_t). */
void *hype_stream(hype_ctx *c);

/* ======================================================================
 * Initial states: cell averages by 5x5 Gauss-Legendre quadrature of the
 * pointwise state in conserved variables (InitializeSolution,
 * initialize_solution.c:21-81).
 * ====================================================================== */
#define HYPE_IC_UNIFORM 0       /* params: the nvar primitive values */
#define HYPE_IC_QUADRANTS 1     /* Euler: xs, ys, then (rho,u,v,p) of NE, NW, SW, SE */
#define HYPE_IC_SHOCK_BUBBLE 2  /* Kapila: mach, x_shock, cx, cy, R, eps, width,
                                   rho_air, rho_he, p0 (air = phase 1) */
#define HYPE_IC_WATER_AIR 3     /* Baer-Nunziato: y_iface, eps, width, cx, cy,
                                   r_burst, p_burst, p0, rho_water, rho_air,
                                   nmodes (<= 8), amplitude, length; mode phases
                                   drawn from seed (water = phase 1, below) */
#define HYPE_IC_SMOOTH_VORTEX 4 /* Kapila SmoothVortex (kapila_5eqn.c:527-562) */
/* the reference's other Kapila test cases (no parameters) */
#define HYPE_IC_INTERFACE_ADVECTION 5   /* kapila_5eqn.c:572 */
#define HYPE_IC_AIR_HELIUM 6            /* kapila_5eqn.c:606 */
#define HYPE_IC_AIR_R22 7               /* kapila_5eqn.c:656 */
#define HYPE_IC_RICHTMYER_MESHKOV 8     /* kapila_5eqn.c:705 */
#define HYPE_IC_UNDERWATER_EXPLOSION 9  /* kapila_5eqn.c:753 */
#define HYPE_IC_WATER_CYLINDER 10       /* kapila_5eqn.c:817 */
#define HYPE_IC_AIR_CAVITY 11           /* kapila_5eqn.c:867, the case
                                           InitialCondition selects
                                           (initialize_solution.c:14) */
#define HYPE_IC_BUBBLE_COLLAPSE 12      /* kapila_5eqn.c:918 */
#define HYPE_IC_AIR_JET 13              /* kapila_5eqn.c:957, left inlet */
#define HYPE_IC_WATER_JET 14            /* kapila_5eqn.c:982, left inlet */

#define HYPE_COORD_EXACT 0      /* cell centres x_min + (i+1/2) h */
#define HYPE_COORD_DMDA 1       /* the centres DMDASetUniformCoordinates gives the
                                   reference (hype.c:132-135) */

int hype_initial_state(const hype_problem *p, int kind, const double *params,
                       int nparams, unsigned long long seed, int coord_mode,
                       double *U);

/* The setting a Kapila test case's header comment states
 * (kapila_5eqn.c:519-990): sets system, domain, boundary kinds and EOS
 * constants of *p (grid, order, path, CFL and options are left to the
 * caller; the grid must give square cells) and, if t_end is not NULL, the
 * case's final time.  kind: HYPE_IC_SMOOTH_VORTEX .. HYPE_IC_WATER_JET. */
int hype_case_setup(int kind, hype_problem *p, double *t_end);

/* Post-shock state of a Mach-M shock moving +x into gas at rest
 * (Rankine-Hugoniot, ideal/stiffened gas with zero stiffness). */
void hype_rankine_hugoniot(double gamma, double rho0, double p0, double M,
                           double *rho1, double *u1, double *p1);

/* ---- on-disk formats -------------------------------------------------- */
/* Restart file in PETSc's binary Vec format (VecView/VecLoad on a binary
 * viewer, post_process.c:47-63, hype.c:316-330): big-endian int32 class id
 * 1211214, int32 length n, n big-endian doubles; block_size > 1 also writes
 * <path>.info with "-vecload_block_size <block_size>".  The state layout is
 * the DMDA natural ordering, i.e. [j][i][c] on one rank. */
int hype_write_vec_binary(const char *path, const double *U, long n, int block_size);
/* HYPE_E_NOTFOUND if the file is missing, HYPE_E_INVALID if its header does
 * not describe a Vec of length n */
int hype_read_vec_binary(const char *path, double *U, long n);

/* Primitive state W (as hype_cons2prim returns it) as a VTK XML structured
 * grid, one scalar array per component, points at the cell centres
 * (coord_mode HYPE_COORD_*): the sol-*.vts snapshots of MonitorFunction
 * (post_process.c:29-41) that plot.py reads. */
int hype_write_vts(const char *path, const hype_problem *p, const double *W, int coord_mode);

/* ---- tiling (used by the autotuner and the tests) ------------------- */
/* Rows per tile of the line-sweep kernel and persistent grid size (0 = one
 * workgroup per tile).  Results do not depend on either. */
int hype_ctx_set_tiling(hype_ctx *c, int hrows, int grid);
int hype_ctx_tiling(hype_ctx *c, int *hrows, int *ntiles, int *grid);

/* ======================================================================
 * Batched mode: many independent instances (any mix of systems, orders,
 * paths and grid sizes) advanced together.  The reference runs one grid
 * per program (hype.c:14-456); each instance here follows exactly the
 * single-instance solve of hype_begin/hype_advance (same dt sequence, same
 * bits), with its own CFL dt, end time and step budget.  One step of the
 * batch is one persistent launch per RK stage per kernel variant over a
 * flattened tile queue of all instances, replayed from a captured hipGraph.
 * ====================================================================== */
typedef struct hype_batch hype_batch;

typedef struct hype_batch_opts {
    int tile_rows;         /* rows per tile; 0 = picked by the tile-shape model */
} hype_batch_opts;

int hype_batch_create(const hype_problem *probs, int n, const hype_batch_opts *opts,
                      hype_batch **out);
void hype_batch_destroy(hype_batch *b);
int hype_batch_set_state(hype_batch *b, int i, const double *U, int mem);
int hype_batch_get_state(hype_batch *b, int i, double *U, int mem);
/* t0 may be NULL (all zero); flags: HYPE_TS_CLIP */
int hype_batch_begin(hype_batch *b, const double *t0, const double *t_end,
                     const int *max_steps, unsigned flags);
/* enqueue nsteps steps of every instance; finished instances idle */
int hype_batch_advance(hype_batch *b, int nsteps);
int hype_batch_advance_timed(hype_batch *b, int nsteps, double *ms);
/* status of every instance (st may be NULL); returns HYPE_E_INADMISSIBLE if
 * any instance stopped on an inadmissible state (first such instance is
 * described by hype_batch_last_error) */
int hype_batch_sync(hype_batch *b, hype_status *st);
const char *hype_batch_last_error(hype_batch *b, double *x, double *y);
int hype_batch_tiling(hype_batch *b, int *rows, int *tiles, int *groups);
int hype_batch_set_rows(hype_batch *b, int rows);
/* Measured tuning of the tile shape and LDS footprint: times `steps` steps
 * at every candidate tile height and, where an instance's system is built
 * for both tile layouts (one wave per tile with one LDS window, or each
 * tile split over a two-wave workgroup with two windows and exchange
 * areas: Euler), in both layouts, and keeps the fastest pair.  Advances the
 * current states: set them again afterwards. */
int hype_batch_autotune(hype_batch *b, const int *cand, int ncand, int steps,
                        int *best_rows, double *best_ms);
/* tile layout of the instances built for both: -1 the default (one wave
 * per tile), 1 split tiles (as hype_batch_autotune left it) */
int hype_batch_layout(hype_batch *b, int *layout);

/* ======================================================================
 * Served mode: a stream of independent solve requests with continuous
 * batching.  Requests are admitted while instance slots and the HBM budget
 * allow; each round retires finished instances, admits queued ones and
 * advances the resident instances that fit an active-cell cap (in arrival
 * order, or least remaining work first) by a few steps.  A request's
 * result is bitwise the result of the same solve run alone.  The reference
 * has no counterpart (one problem per program run, hype.c:14-456).
 * ====================================================================== */
typedef struct hype_server hype_server;

typedef struct hype_server_opts {
    size_t hbm_budget_bytes; /* device bytes for instance states (3 per
                                instance: U and two stage buffers);
                                0 = three quarters of the free memory */
    int max_active;          /* instance slots; 0 = 512 */
    int steps_per_round;     /* 0 = as many as fit round_ms */
    int tile_rows;           /* 0 = tile-shape model at every admission */
    double round_ms;         /* target GPU time of a round; 0 = 20 ms */
    long long max_active_cells; /* cells advanced per round: resident
                                   instances beyond it wait in HBM (one
                                   always runs); 0 = enough to fill the
                                   device, < 0 = no cap */
    int policy;              /* HYPE_SERVE_FIFO or HYPE_SERVE_SHORTEST */
} hype_server_opts;

/* Admission and, under the cell cap, the order instances are advanced in */
#define HYPE_SERVE_FIFO 0     /* arrival order */
#define HYPE_SERVE_SHORTEST 1 /* least remaining work (cells x nvar x steps) first */

typedef struct hype_request {
    hype_problem problem;    /* order 1 or 2 */
    const double *U0;        /* host initial state, read at admission */
    double *U_out;           /* host result buffer (may be NULL), written at
                                completion; both must stay valid until then */
    double t0, t_end;
    int max_steps;
    unsigned flags;          /* HYPE_TS_CLIP */
} hype_request;

#define HYPE_REQ_QUEUED 0
#define HYPE_REQ_RUNNING 1
#define HYPE_REQ_DONE 2

typedef struct hype_request_info {
    hype_status status;      /* final status once done */
    int state;               /* HYPE_REQ_* */
    double t_submit, t_start, t_done; /* server clock (hype_server_now), s */
    size_t bytes;            /* device bytes held while running */
} hype_request_info;

int hype_server_create(const hype_server_opts *opts, hype_server **out);
void hype_server_destroy(hype_server *s);
/* queue a request (HYPE_E_NOMEM if it can never fit the budget) */
int hype_server_submit(hype_server *s, const hype_request *r, int *id);
/* one round: retire, admit, enqueue steps; returns the resident and queued
 * counts.  Two rounds stay in flight: a poll waits only for the round before
 * the last one, whose time states it acts on, while the device runs the
 * last one. */
int hype_server_poll(hype_server *s, int *active, int *queued);
/* poll until every submitted request is done */
int hype_server_drain(hype_server *s);
int hype_server_info(hype_server *s, int id, hype_request_info *info);
double hype_server_now(hype_server *s);
size_t hype_server_bytes_in_use(hype_server *s);

#ifdef __cplusplus
}
#endif
#endif /* HYPE2D_H */
