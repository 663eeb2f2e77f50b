/* orc_internal.h -- shared constants of the oracle (test infrastructure). */
#ifndef HYPE_ORC_INTERNAL_H
#define HYPE_ORC_INTERNAL_H
#include "oracle.h"

#define ORC_MAXVAR 9
#define ORC_MAXDOF 10

/* hype.h:50-52 */
#define ORC_PRS_FLOOR 1.0e-12
#define ORC_RHO_FLOOR 1.0e-14
#define ORC_SMALL_NUM 1.0e-12

double orc_flux_x(const orc_problem *P, const double *Q, double nx, double ny,
                  double *F, int *bad, double *bval);
double orc_flux_prim_x(const orc_problem *P, const double *V, double nx,
                       double ny, double *F, int *bad, double *bval);
double orc_riemann_x(const orc_problem *P, const double *L, const double *R,
                     double nx, double ny, double *F, double *D, int *bad,
                     double *bval);
#endif
