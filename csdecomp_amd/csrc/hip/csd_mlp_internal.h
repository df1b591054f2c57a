// csd_mlp_internal.h -- surrogate inference on points generated on the
// device, used by the EI-ZO driver to rank particles.
#pragma once
#include "csd_internal.h"

struct csd_mlp;

namespace csd {

// Row (slot, i, l), slot < num_slots, i < cand, l < L: the point at fraction
// (l + 1) / (L + 1) from the projection of particle i of region
// active[slot] onto its segment towards the particle. Rows past a region's
// particle count repeat its last particle (their outputs are unused).
struct SegmentPoints {
  const float *part, *ls, *le;  // [R][U][dim], [R][dim], [R][dim]
  const int32_t *active, *u;    // [slots], [R]
  int U, dim, L, cand;
  // device-resident rounds: the number of live slots, read on the device
  // (rows of later slots are skipped); null: every slot is live
  const uint32_t *live_slots = nullptr;
};

// out [num_slots][cand][L]: the surrogate's output at each point
int mlp_predict_segment_points(const csd_mlp *m, const SegmentPoints &s, int num_slots,
                               float *out, hipStream_t st);

// out [num_blocks * 256]: the surrogate's output at the samples of an EI-ZO
// staging buffer, num_blocks segments of 256 samples as [dim][256]
// live_slots (device, may be null): only the first *live_slots slots of
// blocks_per_slot blocks each are predicted
int mlp_predict_staged(const csd_mlp *m, const float *stage, int num_blocks, int dim, float *out,
                       hipStream_t st, const uint32_t *live_slots = nullptr,
                       int blocks_per_slot = 1);

// EI-ZO workspaces that hold the surrogate register here (delta +1 on
// attach, -1 on detach); csd_mlp_destroy refuses while the count is positive,
// so a workspace never keeps a freed model.
void mlp_attach(const csd_mlp *m, int delta);

}  // namespace csd
