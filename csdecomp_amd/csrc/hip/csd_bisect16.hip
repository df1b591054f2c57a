// csd_bisect16.hip -- instantiations of the standalone bisection kernel for
// configurations of at most 16 dimensions.
#define CSD_BISECT_DEFINE_KERNEL
#include "csd_bisect.h"

void csd::launch_bisect_dp16(const BisectLaunch &L, int var, int split) {
  launch_bisect_dp<16>(L, var, split);
}
