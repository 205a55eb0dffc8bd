// Coupling of a coarse and a fine AMR level: the flux register that the
// reference keeps per fine level (AMReX FluxRegister, filled at
// Source/Castro_advance.cpp:82-88, applied by Castro::reflux,
// Source/Castro.cpp:451-486) and the average-down of Castro::avgDown
// (Source/Castro.cpp:488-505).
#pragma once

#include <vector>

#include "mc_common.h"

namespace mc {

// One face of a coarsened fine box: the coarse faces on its low or high
// boundary in direction dir, as a face box (the dir index is fixed).
struct RegFace {
    Box fb;          // face box, coarse face index space
    int fine_box;    // index into the register's fine box list
    int dir, side;   // side 0: low face, 1: high face
    long off;        // offset (doubles) of this face's NVAR components in the register buffer
    long n;          // face cells
};

} // namespace mc

struct mc_flux_register {
    int ratio[3];
    int crse_n[3];
    int periodic;
    std::vector<mc::Box> fine;    // fine boxes, fine index space
    std::vector<mc::RegFace> faces;  // (fine box, dir, side) in that order
    long total = 0;               // doubles in the register buffer
    double *reg = nullptr;        // device: every face's NVAR components, Fortran order per face
    unsigned char *keep = nullptr;  // device: 0 where a face cell is shared with another fine box
    // The coarse level writes the register on its stream, the fine level on
    // its own: every operation waits for the last one if that ran on another
    // stream (an event recorded after each operation).
    hipEvent_t last = nullptr;
    hipStream_t last_stream = nullptr;
    bool used = false;
};

namespace mc {

// the register operations on stream s (device pointers, caller's FABs)
int flux_register_crse_init(mc_flux_register *R, const double *flux, const Box &fbox, int dir, double mult,
                            hipStream_t s);
int flux_register_fine_add(mc_flux_register *R, int fine_box, const double *flux, const Box &fbox, int dir,
                           double mult, hipStream_t s);
// FineAdd of every fine box of the register from the fused engine of the
// fine level (its final fluxes of the step just enqueued): dt and the face
// areas of the three directions as store_flux takes them
struct FusedLevel;
int flux_register_fine_add_fused(mc_flux_register *R, const FusedLevel *F, double dt, const double *area,
                                 double mult, hipStream_t s);
// CrseInit of every register face from the fused engine of the coarse level
int flux_register_crse_init_fused(mc_flux_register *R, const FusedLevel *F, double dt, const double *area,
                                  double mult, hipStream_t s);
int flux_register_clear_internal(mc_flux_register *R, hipStream_t s);
int flux_register_fill(mc_flux_register *R, double v, hipStream_t s);
// vol null: every cell's volume is vconst
int flux_register_reflux(mc_flux_register *R, double *state, const Box &sbox, const Box &valid, const double *vol,
                         const Box &vbox, double scale, hipStream_t s, double vconst = 0.0);
int average_down(const double *fine, const Box &fbox, double *crse, const Box &cbox, const Box &region,
                 const int *ratio, int ncomp, hipStream_t s);

} // namespace mc
