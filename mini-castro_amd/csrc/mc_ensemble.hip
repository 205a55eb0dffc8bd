// Ensembles of independent Sedov instances advanced together by the fused
// engine: one launch per stage covers every instance, each instance keeps its
// own time-step control (Castro::computeNewDt per instance).
#include <stdio.h>

#include <vector>

#include "../../include/mc_castro.h"
#include "mc_fused.h"
#include "mc_internal.h"

using namespace mc;

struct mc_ensemble {
    int ninst = 0;
    std::vector<FusedInstanceSpec> spec;
    std::vector<double> prob_lo, prob_hi;
    FusedLevel *F = nullptr;
    hipStream_t stream = nullptr;
};

extern "C" {

mc_ensemble_t *mc_ensemble_create(int ninst, const int *n_cell, const double *prob_lo, const double *prob_hi,
                                  const int *bc)
{
    return mc_ensemble_create_slabs(ninst, n_cell, prob_lo, prob_hi, bc, -1);
}

mc_ensemble_t *mc_ensemble_create_slabs(int ninst, const int *n_cell, const double *prob_lo, const double *prob_hi,
                                        const int *bc, int planes)
{
    if (planes < -1) {
        set_error("mc_ensemble_create_slabs: planes must be -1 (by memory), 0 (whole instances) or positive");
        return nullptr;
    }
    if (ninst <= 0) {
        set_error("mc_ensemble_create: no instances");
        return nullptr;
    }
    if (!eos_ready()) {
        set_error("mc_ensemble_create: EOS table not initialised");
        return nullptr;
    }
    mc_ensemble_t *E = new mc_ensemble_t();
    E->ninst = ninst;
    for (int m = 0; m < ninst; ++m) {
        FusedInstanceSpec s;
        for (int d = 0; d < 3; ++d) {
            s.n[d] = n_cell[3 * m + d];
            if (s.n[d] <= 0) {
                set_error("mc_ensemble_create: empty instance");
                delete E;
                return nullptr;
            }
            s.dx[d] = (prob_hi[3 * m + d] - prob_lo[3 * m + d]) / (double)s.n[d];
            E->prob_lo.push_back(prob_lo[3 * m + d]);
            E->prob_hi.push_back(prob_hi[3 * m + d]);
        }
        s.bc = bc[m];
        E->spec.push_back(s);
    }
    if (check_hip(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking), "hipStreamCreate") ||
        !(E->F = fused_create_set(E->spec, planes < 0 ? fused_auto_slab(E->spec) : planes))) {
        mc_ensemble_destroy(E);
        return nullptr;
    }
    return E;
}

void mc_ensemble_destroy(mc_ensemble_t *E)
{
    if (!E) return;
    if (E->stream) (void)hipStreamSynchronize(E->stream);
    if (E->F) fused_destroy(E->F);
    if (E->stream) (void)hipStreamDestroy(E->stream);
    delete E;
}

int mc_ensemble_ninst(const mc_ensemble_t *E) { return E->ninst; }
int mc_ensemble_slab(const mc_ensemble_t *E) { return fused_slab(E->F); }

long mc_fused_bytes(int ninst, const int *n_cell, int planes)
{
    std::vector<FusedInstanceSpec> spec(ninst > 0 ? ninst : 0);
    for (int m = 0; m < ninst; ++m)
        for (int d = 0; d < 3; ++d) {
            spec[m].n[d] = n_cell[3 * m + d];
            spec[m].dx[d] = 1.0;
        }
    return (long)fused_bytes_for(spec, planes);
}

double *mc_ensemble_state_ptr(mc_ensemble_t *E, int m)
{
    if (m < 0 || m >= E->ninst) return nullptr;
    return fused_state(E->F, m);
}

// Castro::initData per instance (Source/Castro.cpp:80-117) with the
// instance's own problem constants (Source/Prob.F90:11-15).
int mc_ensemble_init_data(mc_ensemble_t *E, const double *p_ambient, const double *dens_ambient,
                          const double *exp_energy, const double *r_init, int nsub)
{
    for (int m = 0; m < E->ninst; ++m)
        if (!cubic_cells(E->spec[m].dx)) {
            set_error("mc_ensemble_init_data: We don't support dx != dy != dz");
            return -1;
        }
    void *prev = mc_get_stream();
    mc_set_stream(E->stream);
    for (int m = 0; m < E->ninst; ++m) {
        const FusedInstanceSpec &s = E->spec[m];
        const int lo[3] = {0, 0, 0}, hi[3] = {s.n[0] - 1, s.n[1] - 1, s.n[2] - 1};
        const ProblemConsts pc = {p_ambient[m], dens_ambient[m], exp_energy[m], r_init[m], nsub};
        initdata_with(pc, lo, hi, fused_state(E->F, m), lo, hi, s.dx, &E->prob_lo[3 * m], &E->prob_hi[3 * m]);
        if (fused_set_clock(E->F, m, 0.0, 0.0, 0, E->stream)) break;
    }
    mc_set_stream(prev);
    if (check_hip(hipGetLastError(), "ensemble init") || fused_estimate(E->F, E->stream)) return -1;
    return check_hip(hipStreamSynchronize(E->stream), "ensemble init sync");
}

int mc_ensemble_set_state(mc_ensemble_t *E, int m, const double *host)
{
    if (m < 0 || m >= E->ninst) return -1;
    const long n = fused_ncell(E->F, m) * NVAR;
    if (check_hip(hipMemcpyAsync(fused_state(E->F, m), host, n * sizeof(double), hipMemcpyHostToDevice, E->stream),
                  "ensemble set_state"))
        return -1;
    if (fused_estimate(E->F, E->stream)) return -1;
    return check_hip(hipStreamSynchronize(E->stream), "ensemble set_state sync");
}

int mc_ensemble_get_state(mc_ensemble_t *E, int m, double *host)
{
    if (m < 0 || m >= E->ninst) return -1;
    const long n = fused_ncell(E->F, m) * NVAR;
    if (check_hip(hipMemcpyAsync(host, fused_state(E->F, m), n * sizeof(double), hipMemcpyDeviceToHost, E->stream),
                  "ensemble get_state"))
        return -1;
    return check_hip(hipStreamSynchronize(E->stream), "ensemble get_state sync");
}

int mc_ensemble_run(mc_ensemble_t *E, int nsteps, double stop_time, int use_graph)
{
    if (fused_steps(E->F, nsteps, stop_time, E->stream, use_graph != 0)) return -1;
    return check_hip(hipStreamSynchronize(E->stream), "ensemble run");
}

int mc_ensemble_autotune(mc_ensemble_t *E, int reps, int *bricks, double *best_ms)
{
    return fused_autotune(E->F, reps, E->stream, bricks, best_ms);
}

int mc_ensemble_set_stage_variant(mc_ensemble_t *E, int flags)
{
    if (flags & ~FV_ALL) {
        set_error("mc_ensemble_set_stage_variant: unknown flags");
        return -1;
    }
    fused_set_variant(E->F, flags);
    return 0;
}

int mc_ensemble_stage_variant(const mc_ensemble_t *E) { return fused_variant(E->F); }

int mc_ensemble_clock(mc_ensemble_t *E, int m, double *dt, double *time, int *steps)
{
    if (m < 0 || m >= E->ninst) return -1;
    return fused_read_clock(E->F, m, dt, time, steps, E->stream);
}

} // extern "C"
