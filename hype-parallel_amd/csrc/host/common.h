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

/* default rows per tile for a grid, given the device's CU count */
int hype_default_rows(int nx, int ny, int system, int ncu);

/* build the tile list of one instance into tiles (capacity cap); returns
 * the number of tiles (first one flagged as leader) */
int hype_make_tiles(int inst, int nx, int ny, int hrows, hype_tile *tiles, int cap);
int hype_count_tiles(int nx, int ny, int hrows);

/* decode the packed error word into a message and face coordinates */
void hype_decode_error(unsigned long long word, int nx, int ny, int ngf,
                       double x_min, double y_min, double h, char *msg,
                       size_t cap, double *x, double *y);

/* fill the device-independent fields of an instance descriptor */
void hype_fill_inst(hype_inst *I, const hype_problem *p, int hrows);

#endif
