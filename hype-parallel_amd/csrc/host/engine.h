/* engine.h -- multi-instance stepping engine shared by the batched and the
 * served drivers (host side, C).
 *
 * Instances live in slots of a fixed-capacity table on the device.  Every
 * kernel variant in use (system x path x order x layout) has a group: a
 * tile table on the device whose length is also read from device memory, a
 * persistent-queue counter and a branch stream.  One time step of all
 * instances is one persistent launch per RK stage per group, whose
 * workgroups pull tiles tallest first; the launches of all groups are
 * captured once into a hipGraph with one parallel branch per group, and
 * replayed.  Admitting or retiring instances rewrites the tables between
 * replays, never the graph.
 */
#ifndef HYPE_ENGINE_H
#define HYPE_ENGINE_H

#include "common.h"

#define HYPE_MAX_GROUPS 48  /* system x path x order 1..3 x contraction, plus inlet builds */

/* eng_init flags */
#define ENG_ASYNC 0x1u      /* table and time-state uploads do not wait for the
                               stream: the caller orders them (served rounds) */

typedef struct {
    hype_variant v;
    int nstages;
    hype_tile *tiles_d;     /* stepping table (capacity tile_cap) */
    int *ntiles_d;          /* its device-resident length */
    hype_tile *prime_d;     /* priming table of newly started instances */
    int *nprime_d;
    int tile_cap;           /* capacity of the stepping table */
    int prime_cap;          /* capacity of the priming table */
    hype_tile *stage_tiles[2];  /* pinned staging of the two tables, per */
    hype_tile *stage_prime[2];  /* upload parity (capacities as above) */
    unsigned *queue_d;      /* persistent-queue counters {next, finished} of
                               the stepping table, then of the priming table */
    int ntiles, nprime;     /* host mirrors */
    int grid;               /* persistent workgroups (all resident at once) */
    int bwaves;             /* waves per workgroup (2: tile halves on two waves) */
    void *stream;           /* graph branch */
    void *ev_done;
} eng_group;

typedef struct {
    int cap;                /* instance slots */
    hype_inst *inst_h;
    hype_inst *inst_d;
    hype_dstate *st_d;      /* one per slot */
    hype_problem *P;
    int *used, *group_of, *nsy, *fresh;
    int *run;               /* used slots advanced by eng_step (default all) */
    int target_rows;        /* common tile height of all instances */
    int layout;             /* tile layout where a system has both: -1 the
                               default (one wave per tile), 0 one wave, 1 split */
    eng_group g[HYPE_MAX_GROUPS];
    int ng;
    void *stream;           /* main stream */
    void *ev_fork;
    void *graph;            /* one captured step */
    int ncu;
    unsigned flags;         /* ENG_* */
    int par;                /* staging parity of the next uploads */
    hype_inst *inst_stage[2];   /* pinned staging: descriptors, fresh time states */
    hype_dstate *st_stage[2];
} hype_engine;

int eng_init(hype_engine *E, int cap, int target_rows, unsigned flags);
/* Uploads go through pinned staging buffers of two parities.  An ENG_ASYNC
 * engine's caller flips the parity (eng_flip) once per round and makes sure
 * the uploads of the round before the last have completed before it stages
 * new ones; other engines wait for every upload before returning. */
void eng_flip(hype_engine *E);
void eng_free(hype_engine *E);
/* place an instance in a slot (its U, A, B device buffers) */
int eng_attach(hype_engine *E, int slot, const hype_problem *p, double *U, double *A, double *B,
               const double *inlet);
int eng_detach(hype_engine *E, int slot);
/* tile layout of the instances whose system is built for both (-1 default,
 * 0 one wave per tile, 1 split tiles); moves them between groups, taking
 * effect at the next eng_commit */
int eng_set_layout(hype_engine *E, int layout);
/* include (on = 1) or leave out a used slot from the steps; takes effect at
 * the next eng_commit.  A left-out instance keeps its state and time. */
int eng_set_running(hype_engine *E, int slot, int on);
/* recompute the tile tables of every group from the running slots */
int eng_commit(hype_engine *E);
/* initialise the time state of the fresh slots and run their priming call
 * (clip: per slot, may be NULL; at most one call per staging parity) */
int eng_start(hype_engine *E, const int *slots, int n, const double *t0, const double *t_end,
              const int *max_steps, const int *clip, double **hist, const int *hist_len);
/* enqueue nsteps steps of every used slot (graph replays) */
int eng_step(hype_engine *E, int nsteps);
int eng_sync(hype_engine *E);
int eng_read_states(hype_engine *E, hype_dstate *out);
/* tile height the model picks for the used slots */
int eng_pick_rows(hype_engine *E);
int eng_set_rows(hype_engine *E, int rows);
int eng_total_tiles(hype_engine *E);

#endif
