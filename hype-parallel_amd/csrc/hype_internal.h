/*
 * hype_internal.h -- data shared by the C host layer and the HIP device layer.
 *
 * Plain C structs only.  The host layer (csrc/host) fills them; the device
 * layer (csrc/hip) launches kernels over them through the thin C ABI
 * declared at the bottom (hyped_ functions).
 */
#ifndef HYPE_INTERNAL_H
#define HYPE_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HYPE_MAXVAR 9

/* strip geometry of the line-sweep kernel: one wavefront per tile, lanes
 * cover the ring column on each side plus HYPE_TILE_W owned columns */
#define HYPE_WAVE 64
#define HYPE_TILE_W (HYPE_WAVE - 2)

/* device-resident time state of one solve */
typedef struct hype_dstate {
    double t;              /* time at the start of the current step */
    double dt;             /* dt of the current step (written by stage 0) */
    double t_end;
    double pad0;
    unsigned long long smax[2];   /* CFL max wave speed slots, double bits */
    unsigned long long err[2];    /* packed first inadmissible face of the
                                     step of that parity, ~0 = none; the
                                     priming call records into err[1] */
    unsigned long long smax_rhs;  /* slots of single RHS evaluations */
    unsigned long long err_rhs;
    int step;              /* steps completed */
    int cur;               /* index of the step in flight (written by stage 0) */
    int max_steps;
    int done;
    int clip;
    int hist_len;
    double *dt_hist;       /* optional: dt of every step */
} hype_dstate;

/* per-instance descriptor read by the kernels */
typedef struct hype_inst {
    double *buf[4];        /* U (solution), A, B, C (stage work vectors) */
    hype_dstate *st;       /* device time state */
    int nx, ny;
    int bc[4];
    int nsx, nsy;          /* strips x segments of the tiling */
    double h, r1_h, cfl;
    double x_min, y_min;
    double g1, g2, p1, p2;
    unsigned quirks;
    int keybits;           /* face-key bits of the packed error word (28 or 36) */
    const double *inlet;   /* left inlet: ghost values per row j = -3 .. ny+2
                              (nvar each, in the path's variables), or NULL */
} hype_inst;

/* tile: instance, first owned column and row, owned width (<= HYPE_TILE_W)
 * and height, flags (bit 0 = leader tile of its instance, which updates the
 * time state) */
typedef struct hype_tile {
    int inst, x0, y0, w, h, flags;
} hype_tile;

/* packed error word, most significant first: call sequence within the step
 * (4 bits: 0 priming or single RHS, 1.. stage; 10 stages at most), face key
 * (K bits), check kind (4 bits), the offending value as the top 56-K bits of
 * its float32 image.  K = 28 when the grid has fewer than 2^28 face points
 * (value mantissa 19 bits), 36 beyond.  The smallest word is the lowest face
 * of the earliest failing call, as the reference aborts there. */
#define HYPE_ERR_SEQ_SHIFT 60
#define HYPE_ERR_KEYBITS_SMALL 28
#define HYPE_ERR_KEYBITS_LARGE 36

/* kernel modes */
#define HYPE_MODE_STAGE 0
#define HYPE_MODE_RHS 1
#define HYPE_MODE_PRIME 2

/* kernel variant: system x path x order */
typedef struct hype_variant {
    int system, prim, order;
    int fma;               /* 1: the FMA-contracted kernel build (HYPE_OPT_FMA) */
    int pair;              /* 1: each tile split over a two-wave workgroup */
    int inlet;             /* 1: the build with left-inlet ghost columns */
} hype_variant;

typedef struct hype_launch {
    hype_variant v;
    const hype_inst *insts;    /* device array */
    const hype_tile *tiles;    /* device array */
    int ntiles;                /* tile count, unless ntiles_dev is set */
    const int *ntiles_dev;     /* device-resident tile count (tables that
                                  change between graph replays) */
    int mode;
    int stage;
    int grid;                  /* persistent blocks; 0 = one per tile */
    unsigned *queue;           /* device {next, finished} counters */
    const double *rhs_in;      /* RHS/PRIME mode: input state override */
    double *rhs_out;           /* RHS mode: output */
    unsigned char *fallback;   /* RHS mode: optional per-cell fallback mask */
} hype_launch;

/* device scratch of the allocation-free RHS (hype_rhs_noalloc, scratch.hip):
 * the instance descriptor at offset 0, the time state at STATE_OFF, the tile
 * table at TILES_OFF (at most MAX_TILES tiles), the left-inlet table at
 * INLET_OFF up to the end */
#define HYPE_SCRATCH_BYTES (1u << 20)
#define HYPE_SCRATCH_STATE_OFF 256
#define HYPE_SCRATCH_TILES_OFF 512
#define HYPE_SCRATCH_MAX_TILES 16384
#define HYPE_SCRATCH_INLET_OFF (HYPE_SCRATCH_TILES_OFF + HYPE_SCRATCH_MAX_TILES * sizeof(hype_tile))
#define HYPE_SCRATCH_INLET_DOUBLES ((HYPE_SCRATCH_BYTES - HYPE_SCRATCH_INLET_OFF) / sizeof(double))

/* number of stages of the scheme for order N */
int hyped_num_stages(int order);
/* stage s of the order-N scheme: ibuf = {in, out, xb, zb, op}, coef = {a, g,
 * bden}; returns the number of stages, or -1 */
int hyped_stage_table(int order, int s, int *ibuf, double *coef);

/* ---- thin device ABI (implemented in csrc/hip/device.hip) -------------- */
int hyped_init(void);
int hyped_malloc(void **p, size_t bytes);
int hyped_free(void *p);
/* pinned host memory, for copies that do not block the host */
int hyped_host_alloc(void **p, size_t bytes);
int hyped_host_free(void *p);
int hyped_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int hyped_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int hyped_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
int hyped_fill_u64(void *dst, unsigned long long value, size_t count, void *stream);
int hyped_stream_create(void **s);
int hyped_stream_destroy(void *s);
int hyped_stream_sync(void *s);
int hyped_launch(const hype_launch *L, void *stream);
int hyped_cons2prim(int system, const double *U, double *W, long ncells,
                    double g1, double g2, double p1, double p2, void *stream);
int hyped_copy(double *dst, const double *src, long n, void *stream);
int hyped_mem_info(size_t *free_bytes, size_t *total_bytes);
int hyped_num_cus(void);
/* resident workgroups per CU of the stage kernel of a variant */
int hyped_blocks_per_cu(const hype_variant *v);
/* waves per workgroup of that kernel: 2 with split tiles (the two halves of
 * each tile on their own waves, see the stage kernel), else 1 */
int hyped_block_waves(const hype_variant *v);
/* tile layouts a system's stage kernel is built for: bit 0 one wave per
 * tile, bit 1 split tiles */
int hyped_pair_modes(int system);
const char *hyped_last_error(void);
/* this device's address of the scratch array and its size (looked up on the
 * first call on each device, with that device current) */
int hyped_scratch(void **ptr, size_t *bytes);

/* HIP events on a stream */
int hyped_event_create(void **ev);
int hyped_event_destroy(void *ev);
int hyped_event_record(void *ev, void *stream);
int hyped_event_elapsed_ms(void *a, void *b, float *ms);
int hyped_event_sync(void *ev);
int hyped_event_create_sync(void **ev);   /* no timing: fork/join in graphs */
int hyped_stream_wait(void *stream, void *ev);

/* graph capture of a sequence of launches on one stream */
int hyped_graph_begin(void *stream);
int hyped_graph_end(void *stream, void **graph_exec);
int hyped_graph_launch(void *graph_exec, void *stream);
int hyped_graph_destroy(void *graph_exec);

#ifdef __cplusplus
}
#endif
#endif
