// sweep_bn.hip -- the Baer-Nunziato stage kernels of the exact build, compiled
// apart from the Euler and Kapila ones so that each part gets its own
// scheduler options (hype-parallel_amd/Makefile).  hype::launch_stage and
// hype::blocks_per_cu in sweep.hip hand Baer-Nunziato launches to this unit.
#define HYPE_SWEEP_BN 1
#include "sweep.hip"
