// Batched fused CTU step over a whole box list (one launch per stage).
#pragma once

#include <vector>

#include "mc_common.h"

namespace mc {

struct FusedLevel;

FusedLevel *fused_create(const std::vector<Box> &boxes, const int *n_cell, int bc, const double *dx);
void fused_destroy(FusedLevel *F);
// Estimate dt (min over cells, before the CFL factor) of the box-arena state S into *dt_dev.
int fused_est_dt(FusedLevel *F, const double *S_arena, double *dt_dev, hipStream_t s);
// One Castro::advance: S_new = clean(clean(S_old) + dt * hydro_source) on every box.
int fused_advance(FusedLevel *F, const double *S_old_arena, double *S_new_arena, double dt, hipStream_t s);
// Castro::post_timestep's clean_state on the box arena.
int fused_clean(FusedLevel *F, double *S_arena, hipStream_t s);

} // namespace mc
