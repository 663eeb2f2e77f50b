/* common.h -- helpers shared by the C host files. */
#ifndef HYPE_HOST_COMMON_H
#define HYPE_HOST_COMMON_H

#include "../../../include/hype2d.h"
#include "../hype_internal.h"

#define HYPE_TRY(x)                  \
    do {                             \
        int rc_ = (x);               \
        if (rc_ != HYPE_OK) return rc_; \
    } while (0)

/* validate a problem description; fills nvar */
int hype_check_problem(const hype_problem *p, int *nvar);

/* strips of width <= HYPE_TILE_W covering nx columns */
int hype_strips(int nx);

/* number of row segments for a grid on a device with `simds` SIMDs and
 * `occ` resident waves of the kernel per SIMD; pair: each tile runs on two
 * waves, one per half (tile-shape model in common.c) */
int hype_pick_segments(int nx, int ny, int simds, int occ, int extra_waves, int pair);
/* left inlet (Kapila): doubles of the per-row table (0 without an inlet),
 * and the table itself, rows j = -3 .. ny+2 of nvar values each, primitive
 * or conserved as the problem's path stores ghost cells */
size_t hype_inlet_doubles(const hype_problem *p);
void hype_inlet_table(const hype_problem *p, double *out);
/* tile layouts a problem may use (bit 0: one wave per tile, bit 1: split
 * tiles), given those its system's kernel is built for */
int hype_layouts(const hype_problem *p, int built);
/* tile layout (*pair: 1 split tiles on two-wave workgroups, 0 one wave per
 * tile) and row segments of a single grid: the tile-shape model's cheapest
 * among the layouts the system's kernel is built for; *slots (may be NULL)
 * receives the resident waves of that kernel on the device */
int hype_pick_layout(const hype_problem *p, int *pair, int *nsy, int *slots);
/* row-iterations of one wave for a tile of h rows */
double hype_tile_rows(int h, int pair);
/* modelled time of a launch of `waves` waves of `per` row-iterations each */
double hype_tile_cost(long waves, double per, int simds, int occ);

/* build the tile list of one instance (nsx strips x nsy segments, near-equal
 * splits) into tiles (capacity cap); returns the number of tiles, the first
 * flagged as leader, or -1 */
int hype_make_tiles(int inst, int nx, int ny, int nsx, int nsy, hype_tile *tiles, int cap);

/* devices dispatch workgroups round-robin over this many XCDs (MI355X) */
#define HYPE_NXCD 8
/* reorder a row-major tile list so each XCD gets a contiguous band (in
 * place, or from src into dst) */
int hype_xcd_order(hype_tile *tiles, int n, int nxcd);
void hype_xcd_order_into(const hype_tile *src, hype_tile *dst, int n, int nxcd);

/* face-key bits of the packed error word for an nx x ny grid */
int hype_key_bits(int nx, int ny);

/* decode the packed error word into a message and face coordinates */
void hype_decode_error(unsigned long long word, int nx, int ny, int ngf,
                       double x_min, double y_min, double h, char *msg,
                       size_t cap, double *x, double *y);

/* x (axis 0) or y (axis 1) coordinate of cell i (HYPE_COORD_* mode) */
double hype_cell_coord(const hype_problem *p, int axis, int i, int coord_mode);

/* fill the device-independent fields of an instance descriptor */
void hype_fill_inst(hype_inst *I, const hype_problem *p, int nsy);

#endif
