// csd_bisect8.hip -- instantiations of the standalone bisection kernel for
// configurations of at most 8 dimensions.
#define CSD_BISECT_DEFINE_KERNEL
#include "csd_bisect.h"

void csd::launch_bisect_dp8(const BisectLaunch &L, int var, int split) {
  launch_bisect_dp<8>(L, var, split);
}
