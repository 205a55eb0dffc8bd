// Library-internal helpers shared by the translation units.
#pragma once

#include <math.h>

#include "mc_common.h"
#include "mc_eos.h"

namespace mc {

hipStream_t cur_stream();
const EosTable &eos_table();
bool eos_ready();
void set_error(const char *msg);
int check_hip(hipError_t e, const char *what);

// Castro::initData's guard (Source/Castro.cpp:95-100): only dx = dy = dz
inline bool cubic_cells(const double *dx)
{
    const double SMALL = 1.e-13;
    return !((fabs(dx[0] - dx[1]) > SMALL * dx[0]) || (fabs(dx[0] - dx[2]) > SMALL * dx[0]));
}

// Problem constants of initdata (Source/Prob.F90:11-15).  The drop-in
// initdata reads the process-wide set (mc_set_problem), as the reference
// reads its probdata module; the level and ensemble drivers pass their own,
// so drivers initialised from different host threads do not share them.
struct ProblemConsts {
    double p_ambient, dens_ambient, exp_energy, r_init;
    int nsub;
};
void initdata_with(const ProblemConsts &pc, const int *lo, const int *hi, double *state, const int *state_lo,
                   const int *state_hi, const double *dx, const double *problo, const double *probhi);

// Asynchronous estdt over box bx of a state FAB on sbox: min-reduces the
// CFL time step of every cell into *dt_dev (no synchronisation).
void launch_estdt(const Box &bx, const double *state, const Box &sbox, const double *dx, double *dt_dev,
                  hipStream_t s);

} // namespace mc
