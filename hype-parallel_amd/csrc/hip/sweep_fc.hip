// sweep_fc.hip -- the stage kernel again, compiled with FMA contraction
// (-ffp-contract=fast) into its own namespace.  Selected per problem with
// HYPE_OPT_FMA: a*b+c is fused wherever the compiler can, which is faster and
// no less accurate, but no longer bitwise equal to the uncontracted
// reference arithmetic (see include/hype2d.h).  This unit holds Euler and
// Kapila; sweep_fc_bn.hip holds Baer-Nunziato.
#define HYPE_KNS hype_fc
#define HYPE_CONTRACTED_BUILD 1
#include "sweep.hip"
