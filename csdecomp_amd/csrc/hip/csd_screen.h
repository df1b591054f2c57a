// csd_screen.h -- screened checking of fast-path plants (csd_screen.hip).
#pragma once
#include "csd_internal.h"

namespace csd {
// FK + screened pair tests for every configuration, then the full check of
// the undecided ones, `range` configurations per survivor workgroup
// (a multiple of 64, at most 16384).
void launch_screened(int var, const csd_plant *pl, const float *q, int64_t n, int64_t sn,
                     int64_t sj, int32_t range, uint8_t *out, hipStream_t st);
// screening applies to fast-path plants of 4 sphere slots per body
bool screen_supported(int var);
// Chooses the plant's screen from the pair-test hit bits of up to 4,096
// configurations sampled evenly from q; *undecided = the sample fraction the
// screen leaves undecided. Synchronizes `st`.
int build_screen(int var, csd_plant *pl, const float *q, int64_t n, int64_t sn, int64_t sj,
                 hipStream_t st, double *undecided);
}  // namespace csd
