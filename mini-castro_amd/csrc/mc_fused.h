// Batched fused CTU step: every box of a level, or every member of an
// ensemble of independent instances, in one launch per stage.
#pragma once

#include <vector>

#include "mc_common.h"

namespace mc {

struct FusedLevel;

struct FusedInstanceSpec {
    int n[3];
    double dx[3];
    int bc;  // 0 periodic, 1 outflow
};

// Stages of a step.  FST_TFLUX is FST_TRACE and FST_FLUX1 in one launch,
// FST_TRANS1C and FST_FINALC are FST_TRANS1 and FST_FINAL computed by
// cells; which of them a step runs is its kernel variant
// (fused_set_variant).
enum {
    FST_EST = 0, FST_PRIM, FST_FLAT, FST_TRACE, FST_FLUX1, FST_TRANS1, FST_FINAL, FST_UPDATE, FST_TFLUX, FST_TRANS1C,
    FST_FINALC, FST_N
};

// slab > 0: the face stages run in z-slabs of that many cell planes, so the
// face arrays hold one slab (and its ring) instead of whole instances; 0: whole
// instances.  fused_auto_slab picks 0 when whole-instance arrays fit in the
// free device memory with room to spare, else the deepest slab that fits.
FusedLevel *fused_create_set(const std::vector<FusedInstanceSpec> &spec, int slab = 0);
// single level: the box list only describes the state arena for import/export
FusedLevel *fused_create(const std::vector<Box> &boxes, const int *n_cell, int bc, const double *dx, int slab = 0);
size_t fused_bytes_for(const std::vector<FusedInstanceSpec> &spec, int slab);  // device bytes of an engine
int fused_auto_slab(const std::vector<FusedInstanceSpec> &spec);
int fused_slab(const FusedLevel *F);
void fused_destroy(FusedLevel *F);
int fused_import(FusedLevel *F, const double *arena, hipStream_t s);
int fused_export(FusedLevel *F, double *arena, hipStream_t s);
double *fused_state(FusedLevel *F, int m);  // instance m state (20 comps, cell layout)
long fused_ncell(FusedLevel *F, int m);
long fused_next(const FusedLevel *F, int m);  // component stride of instance m's extended arrays
int fused_ninst(FusedLevel *F);
int fused_estimate(FusedLevel *F, hipStream_t s);
int fused_steps(FusedLevel *F, int nsteps, double stop_time, hipStream_t s, bool use_graph);
int fused_steps_timed(FusedLevel *F, int nsteps, double stop_time, hipStream_t s, double *stage_ms, int *launches);
int fused_advance_dt(FusedLevel *F, double dt, hipStream_t s);
int fused_recompute_dt(FusedLevel *F, double dt, hipStream_t s);
int fused_set_clock(FusedLevel *F, int m, double dt, double time, int steps, hipStream_t s);
int fused_read_clock(FusedLevel *F, int m, double *dt, double *time, int *steps, hipStream_t s);
int fused_read_est(FusedLevel *F, int m, double *est, hipStream_t s);
// a brick of 64..256 threads (a multiple of 64); for FST_TFLUX,
// FST_TRANS1C and FST_FINALC the triple (W, L, X) with L, X >= 2, W*L of
// 64..256 threads and X dividing W*L
bool fused_tiling_ok(int stage, int bx, int by, int bz);
void fused_set_tiling(FusedLevel *F, int stage, int bx, int by, int bz);  // ignores an invalid tiling
// Brick order of a stage: 0 raster (x, then y, then z), or groups of zg
// brick planes in z (x, then z within the group, then y).
bool fused_brick_order_ok(int zg);
void fused_set_brick_order(FusedLevel *F, int stage, int zg);
int fused_brick_order(const FusedLevel *F, int stage);
void fused_get_tiling(FusedLevel *F, int stage, int *bxyz);
int fused_run_stage(FusedLevel *F, int stage, hipStream_t s);
int fused_autotune(FusedLevel *F, int reps, hipStream_t s, int *bricks, double *best_ms);
// Kernel variants of a step (bit flags; 0, the default, runs K_tflux,
// K_trans1c and K_final): FV_SPLIT_TRACE runs K_trace + K_flux1 instead of
// K_tflux, FV_FACE_TRANS1 K_trans1 instead of K_trans1c, FV_CELL_FINAL
// K_finalc instead of K_final.  Results are the same bit for bit.
enum { FV_SPLIT_TRACE = 1, FV_FACE_TRANS1 = 2, FV_CELL_FINAL = 4, FV_ALL = 7 };
void fused_set_variant(FusedLevel *F, int flags);
// true: the step ends after the advance's clean_state (no post_timestep clean,
// no estdt), for a coarse level whose post_timestep follows a reflux
void fused_set_defer_post(FusedLevel *F, bool defer);
int fused_variant(const FusedLevel *F);

// The arrays a step writes besides the state (primitives, flattening,
// traces, fluxes).  A caller may give the engine another set to write into,
// to keep a step's intermediates instead of recomputing them later.
struct FusedArrays {
    double *U1, *U1X, *Q, *QX, *FL, *DV;
    double *SM[3], *SP[3], *FD[3], *FT[6], *FN[3];
};
size_t fused_intermediate_bytes(const FusedLevel *F);
int fused_alloc_intermediates(const FusedLevel *F, FusedArrays &S);  // all null on failure
void fused_free_intermediates(FusedArrays &S);
FusedArrays fused_get_intermediates(const FusedLevel *F);
void fused_set_intermediates(FusedLevel *F, const FusedArrays &S);  // drops a captured graph

} // namespace mc
