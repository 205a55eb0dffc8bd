// Batched fused CTU path (placeholder until the fused kernels land).
#include "mc_fused.h"
#include "mc_internal.h"

namespace mc {

struct FusedLevel {
    int dummy;
};

FusedLevel *fused_create(const std::vector<Box> &, const int *, int, const double *)
{
    set_error("fused mode is not built yet");
    return nullptr;
}
void fused_destroy(FusedLevel *F) { delete F; }
int fused_est_dt(FusedLevel *, const double *, double *, hipStream_t) { return -1; }
int fused_advance(FusedLevel *, const double *, double *, double, hipStream_t) { return -1; }
int fused_clean(FusedLevel *, double *, hipStream_t) { return -1; }

} // namespace mc
