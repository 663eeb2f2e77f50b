// sweep_fc_bn.hip -- the Baer-Nunziato stage kernels of the contracted build
// (sweep_fc.hip), compiled apart from the Euler and Kapila ones as in the
// exact build (sweep_bn.hip).
#define HYPE_KNS hype_fc
#define HYPE_CONTRACTED_BUILD 1
#define HYPE_SWEEP_BN 1
#include "sweep.hip"
