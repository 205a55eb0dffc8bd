/*
 * mini-castro-amd: MI355X (gfx950) CTU hydro kernels behind mini-Castro's
 * C ABI.
 *
 * Part 1 declares, with the same names and argument lists, the extern "C"
 * entry points a mini-Castro host binds (Source/Castro_F.H:11-161).  A FAB
 * argument is the triplet (double* p, const int* p_lo, const int* p_hi) of
 * BL_FORT_FAB_ARG_3D: Fortran order, i fastest, component slowest
 * (Source/Castro.H:19).  Semantics are those of the reference's OpenACC /
 * OpenMP-offload builds (Source/Castro.H:23-24): the host calls each entry
 * point once per box with the whole box as (lo, hi) and DEVICE pointers; the
 * kernel is launched asynchronously on the stream given to mc_set_stream
 * (default: the null stream).  lo/hi and the FAB bounds are read during the
 * call and never retained.  Component maps are those of
 * Source/Castro_nd.F90:14-57 (1-based there, idir is 1-based here as well).
 *
 * Part 2 declares extensions that have no counterpart in the reference: a
 * stream setter, table loading from an explicit path or from memory, a
 * problem-constant setter, batch EOS evaluation, and the level driver that
 * advances a whole box list (the role of Castro::advance) either with the
 * reference's per-box kernel sequence or with the batched fused path.
 *
 * Errors: the reference routines return void and so do the Part 1 entry
 * points; a failure is printed and recorded for mc_last_error().  Part 2
 * calls return 0 on success and a negative code on failure.
 */
#ifndef MC_CASTRO_H
#define MC_CASTRO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ===================== Part 1: Source/Castro_F.H ===================== */

/* Source/Castro_F.H:11-13 (network.F90 has no state: no-ops) */
void network_init(void);
void network_finalize(void);

/* Source/Castro_F.H:15-17 -> Source/eos.F90:580-734, 1063-1101.
   eos_init reads ./helm_table.dat and aborts if it cannot (eos.F90:637-641). */
void eos_init(void);
void eos_finalize(void);

/* Source/Castro_F.H:19-23 -> Source/Castro_hydro_nd.F90:73-170 */
void ctoprim(const int *lo, const int *hi,
             const double *u, const int *u_lo, const int *u_hi,
             double *q, const int *q_lo, const int *q_hi,
             double *qaux, const int *qa_lo, const int *qa_hi);

/* Source/Castro_F.H:25-39 -> Source/Castro_nd.F90:69-297 (clean_state kernels) */
void enforce_minimum_density(const int *lo, const int *hi, double *state, const int *state_lo, const int *state_hi);
void normalize_species(const int *lo, const int *hi, double *state, const int *state_lo, const int *state_hi);
void reset_internal_e(const int *lo, const int *hi, double *state, const int *state_lo, const int *state_hi);
void compute_temp(const int *lo, const int *hi, double *state, const int *state_lo, const int *state_hi);

/* Source/Castro_F.H:41-45 -> Source/timestep.F90:12-72.  *dt = min(*dt, CFL
   time step of the box).  dt may be host, managed or device memory; the call
   returns after the stream has drained. */
void estdt(const int *lo, const int *hi,
           const double *state, const int *state_lo, const int *state_hi,
           const double *dx, double *dt);

/* Source/Castro_F.H:47-52 -> Source/Castro_hydro_nd.F90:301-340 */
void store_flux(const int *lo, const int *hi,
                double *flux_out, const int *flux_out_lo, const int *flux_out_hi,
                double *flux_in, const int *flux_in_lo, const int *flux_in_hi,
                double *area, const int *area_lo, const int *area_hi,
                const double dt);

/* Source/Castro_F.H:54-70 -> Source/Castro_hydro_nd.F90:344-441 */
void fill_hydro_source(const int *lo, const int *hi,
                       const double *state, const int *state_lo, const int *state_hi,
                       const double *q, const int *q_lo, const int *q_hi,
                       double *source, const int *source_lo, const int *source_hi,
                       const double *flux0, const int *flux0_lo, const int *flux0_hi,
                       const double *flux1, const int *flux1_lo, const int *flux1_hi,
                       const double *flux2, const int *flux2_lo, const int *flux2_hi,
                       const double *qe0, const int *qe0_lo, const int *qe0_hi,
                       const double *qe1, const int *qe1_lo, const int *qe1_hi,
                       const double *qe2, const int *qe2_lo, const int *qe2_hi,
                       const double *area0, const int *area0_lo, const int *area0_hi,
                       const double *area1, const int *area1_lo, const int *area1_hi,
                       const double *area2, const int *area2_lo, const int *area2_hi,
                       const double *volume, const int *volume_lo, const int *volume_hi,
                       const double *dx, const double dt);

/* Source/Castro_F.H:72-75 -> Source/Castro_hydro_nd.F90:248-297 */
void normalize_species_fluxes(const int *lo, const int *hi, double *flux, const int *flux_lo, const int *flux_hi);

/* Source/Castro_F.H:77-83 -> Source/Castro_hydro_nd.F90:174-245 */
void apply_av(const int *lo, const int *hi, const int idir, const double *dx,
              double *div, const int *div_lo, const int *div_hi,
              double *Sborder, const int *Sborder_lo, const int *Sborder_hi,
              double *flux, const int *flux_lo, const int *flux_hi);

/* Source/Castro_F.H:85-96 -> Source/trans.F90:14-243 */
void trans1(const int *lo, const int *hi, const int idir1, const int idir2,
            const double *q2m, const int *q2m_lo, const int *q2m_hi,
            double *q2mo, const int *q2mo_lo, const int *q2mo_hi,
            const double *q2p, const int *q2p_lo, const int *q2p_hi,
            double *q2po, const int *q2po_lo, const int *q2po_hi,
            const double *qaux, const int *qaux_lo, const int *qaux_hi,
            const double *f1, const int *f1_lo, const int *f1_hi,
            const double *q1, const int *q1_lo, const int *q1_hi,
            const double cdtdx);

/* Source/Castro_F.H:98-111 -> Source/trans.F90:250-505 */
void trans2(const int *lo, const int *hi, const int idir1, const int idir2, const int idir3,
            const double *qm, const int *qm_lo, const int *qm_hi,
            double *qmo, const int *qmo_lo, const int *qmo_hi,
            const double *qp, const int *qp_lo, const int *qp_hi,
            double *qpo, const int *qpo_lo, const int *qpo_hi,
            const double *qaux, const int *qaux_lo, const int *qaux_hi,
            const double *fyz, const int *fyz_lo, const int *fyz_hi,
            const double *fzy, const int *fzy_lo, const int *fzy_hi,
            const double *qy, const int *qy_lo, const int *qy_hi,
            const double *qz, const int *qz_lo, const int *qz_hi,
            const double hdtdx, const double cdtdy, const double cdtdz);

/* Source/Castro_F.H:113-121 -> Source/riemann.F90:10-331 */
void compute_flux(const int *lo, const int *hi,
                  const double *ql, const int *ql_lo, const int *ql_hi,
                  const double *qr, const int *qr_lo, const int *qr_hi,
                  double *flux, const int *flux_lo, const int *flux_hi,
                  double *q_int, const int *q_int_lo, const int *q_int_hi,
                  double *qe, const int *qe_lo, const int *qe_hi,
                  const double *qaux, const int *qaux_lo, const int *qaux_hi,
                  const int idir);

/* Source/Castro_F.H:123-127 -> Source/Castro_hydro_nd.F90:11-69 */
void divu(const int *lo, const int *hi,
          double *q, const int *q_lo, const int *q_hi,
          const double *dx,
          double *div, const int *div_lo, const int *div_hi);

/* Source/Castro_F.H:129-140 -> Source/ppm.F90:309-743 */
void trace_ppm(const int *lo, const int *hi, const int *vlo, const int *vhi, const int idir,
               const double *q, const int *q_lo, const int *q_hi,
               const double *qaux, const int *qaux_lo, const int *qaux_hi,
               double *qm, const int *qm_lo, const int *qm_hi,
               double *qp, const int *qp_lo, const int *qp_hi,
               const int *domlo, const int *domhi, const double *dx, const double dt);

/* Source/Castro_F.H:142-146 -> Source/Prob.F90:19-129; the problem
   constants (Prob.F90:11-15) come from mc_set_problem, defaulting to the
   reference's values. */
void initdata(const int *lo, const int *hi, double *state, const int *state_lo, const int *state_hi,
              const double *dx, const double *problo, const double *probhi);

/* Source/Castro_F.H:148-152 -> Source/Castro_nd.F90:301-345 */
void denerror(const int *lo, const int *hi,
              int8_t *tag, const int *tag_lo, const int *tag_hi,
              const double *den, const int *den_lo, const int *den_hi,
              const int8_t tagval, const int8_t clearval);

/* Source/Castro_F.H:154-159 -> Source/Castro_nd.F90:349-400.  blast_mass and
   blast_radius are host or managed memory and are accumulated into; the
   call synchronises. */
void calculate_blast_radius(const int *lo, const int *hi,
                            double *state, const int *state_lo, const int *state_hi,
                            const double *dx, const double *problo, const double *probhi,
                            double *blast_mass, double *blast_radius, const double max_density);

/* ===================== Part 2: extensions ===================== */

/* Allocation-free calculate_blast_radius.  No counterpart in the reference:
   it is for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc).
   Same arguments, arithmetic and results (bit for bit) as
   calculate_blast_radius above, but the block partial sums live in a
   device array of fixed size that the library holds on every GPU it runs on
   until the process ends, so the call allocates nothing.  A box may hold at
   most MC_NOALLOC_MAX_CELLS cells; a larger one is refused (-1, see
   mc_last_error) before anything is launched.  Every call on one GPU shares
   that GPU's array: calls that overlap on one GPU (several host threads or
   streams at once) race and must be serialised by the caller.  Returns 0 on
   success; synchronises the stream. */
#define MC_NOALLOC_MAX_CELLS 4194304
int mc_calculate_blast_radius_noalloc(const int *lo, const int *hi,
                                      double *state, const int *state_lo, const int *state_hi,
                                      const double *dx, const double *problo, const double *probhi,
                                      double *blast_mass, double *blast_radius, const double max_density);

/* Stream for subsequent Part 1 launches of the calling thread (hipStream_t). */
void mc_set_stream(void *stream);
void *mc_get_stream(void);
/* Text of the last error recorded on the calling thread. */
const char *mc_last_error(void);

/* Table loading: a helm_table.dat text file or its ".bin" cache; 21 host
   columns (541 x 201 each, i fastest, column order of Source/eos.F90:82-94);
   or the table the library generates (ideal electron gas, layout of
   Source/eos.F90:626-670). */
int mc_eos_init_from(const char *path);
int mc_eos_init_arrays(const double *tab21);
int mc_eos_init_generated(void);

/* Device batch EOS: mode 1 (rho,T), 2 (rho,e), 3 (rho,p); ev holds e or p.
   out (device, 10 doubles per point): T p e cs gam1 dpde dpdr_e dedT iat jat. */
int mc_eos_eval_batch(int n, int mode, const double *rho, const double *T, const double *ev,
                      const double *abar, const double *zbar, double *out);

/* EOS Newton solves (rho,e and rho,p inputs, every kernel of the library)
   that ended after 100 iterations without converging.  The reference passes
   them silently (Source/eos.F90:193,550) and so do the results here; the
   count is the diagnostic.  Synchronises the device; reset != 0 clears it. */
int mc_eos_newton_failures(unsigned long long *count, int reset);

/* Problem constants used by initdata (Source/Prob.F90:11-15): one
   process-wide set, like the reference's probdata module.  The level and
   ensemble init_data calls take their own constants and leave this set
   alone, so drivers initialised from different host threads do not race. */
void mc_set_problem(double p_ambient, double dens_ambient, double exp_energy, double r_init, int nsub);

/* ---- level driver: one AMR level, a list of boxes tiling the domain ----
   MC_MODE_PLAIN replays the reference's per-box kernel sequence through the
   Part 1 entry points; MC_MODE_FUSED advances all boxes with one launch per
   stage (bit-identical results).  In fused mode mc_level_advance includes
   post_timestep's clean_state and mc_level_post_timestep does nothing,
   except on the coarse side of a flux register (mc_level_set_flux_registers),
   whose post_timestep clean follows the reflux and average-down. */
typedef struct mc_level mc_level_t;
enum { MC_BC_PERIODIC = 0, MC_BC_OUTFLOW = 1 };
enum { MC_MODE_PLAIN = 0, MC_MODE_FUSED = 1 };

/* boxes: nbox x 6 ints (lo x,y,z, hi x,y,z) tiling [0, n_cell). */
mc_level_t *mc_level_create(const int *n_cell, const double *prob_lo, const double *prob_hi, int bc,
                            int nbox, const int *boxes);
void mc_level_destroy(mc_level_t *L);
int mc_level_set_mode(mc_level_t *L, int mode);
int mc_level_nbox(const mc_level_t *L);
/* S_new of box b (20 components, no ghosts, Fortran order) */
double *mc_level_state_ptr(mc_level_t *L, int b);
int mc_level_set_state(mc_level_t *L, int b, const double *host);
int mc_level_get_state(mc_level_t *L, int b, double *host);
/* Castro::initData with initdata on every box (Source/Castro.cpp:80-117) */
int mc_level_init_data(mc_level_t *L, double p_ambient, double dens_ambient, double exp_energy,
                       double r_init, int nsub);
double mc_level_est_time_step(mc_level_t *L);   /* Source/Castro.cpp:175-216 */
double mc_level_initial_dt(mc_level_t *L);      /* Source/Castro.cpp:162-173 */
int mc_level_advance(mc_level_t *L, double dt); /* Source/Castro_advance.cpp:11-100 */
int mc_level_post_timestep(mc_level_t *L);      /* Source/Castro.cpp:343-369 */
/* Amr::coarseTimeStep: computeInitialDt/computeNewDt, advance,
   post_timestep.  Returns the dt taken (negative on error). */
double mc_level_coarse_step(mc_level_t *L, double stop_time);
/* nsteps coarse steps; in fused mode dt stays on the device between steps and
   the step may be replayed from a captured hipGraph (use_graph != 0). */
int mc_level_run(mc_level_t *L, int nsteps, double stop_time, int use_graph);
/* Fused mode: nsteps eager steps with HIP events around every stage launch on
   the level's stream; stage_ms[0..11] / launches[0..11] receive per-stage
   totals (stages as in mc_level_set_tiling; index 11: the time-step control
   kernels).  A stage the step does not run reports 0 launches. */
int mc_level_run_timed(mc_level_t *L, int nsteps, double stop_time, double *stage_ms, int *launches);
int mc_level_steps(const mc_level_t *L);
/* Castro::post_timestep's blast-radius diagnostic (Source/Castro.cpp:371-424):
   max_density = S_new.max(Density); mass / radius sums of
   calculate_blast_radius over every box; the printed radius is
   (blast_radius / blast_mass) / 1e5 km. */
int mc_level_blast_radius(mc_level_t *L, double *blast_mass, double *blast_radius, double *max_density);
double mc_level_time(const mc_level_t *L);
double mc_level_dt(const mc_level_t *L);
int mc_level_sync(mc_level_t *L);
/* The fused engine keeps its face-stage arrays (traces, fluxes: about
   3.1 KB per cell) for the whole level, or for one z-slab of `planes` cell
   planes at a time, the face stages and the update then running slab after
   slab (about 0.35 KB per cell plus 3.1 KB per cell of one slab and its
   two ring planes).  Results are the same bit for bit.  planes = -1 (the
   default) keeps the whole level when that takes at most half the free
   device memory, else picks the deepest slab that does; 0 forces the whole
   level.  Changing it rebuilds the engine (state and clock kept).
   mc_level_fused_slab returns the planes in use (0: whole level). */
int mc_level_set_fused_slab(mc_level_t *L, int planes);
int mc_level_fused_slab(mc_level_t *L);
/* Fused-mode tiling: stage 0..10 = est, prim, flat, trace, flux1, trans1,
   final, update, tflux, trans1c, finalc; a brick of bx*by*bz threads
   (multiple of 64, <= 256).  tflux (trace and flux1 in one launch), trans1c
   and finalc (trans1 and final by cells), whose bricks overlap by one cell
   along their direction, take (W, L, X): bricks (W, L, 1) along y,
   (W, 1, L) along z, (X, W*L/X, 1) along x; L, X >= 2, W*L threads
   (multiple of 64, <= 256), X dividing W*L.  tflux stages its bricks'
   primitives in LDS with two cells either side: (L + 4) W <= 448 and
   (X + 4) W L / X <= 448.  trans1 stages its d-fluxes in LDS (the brick
   widened by one cell in two directions, 7 doubles a cell), at most 64 KB. */
int mc_level_set_tiling(mc_level_t *L, int stage, int bx, int by, int bz);
int mc_level_get_tiling(mc_level_t *L, int stage, int *bxyz);
/* Order in which a stage's bricks are launched: 0 raster (x fastest, then
   y, then z), or zgroup = 2..64 brick planes per group in z (x fastest,
   then z within the group, then y), which keeps a brick's z-neighbours
   close in time.  Never changes results.
   mc_level_brick_order returns it (-1 on error). */
int mc_level_set_brick_order(mc_level_t *L, int stage, int zgroup);
int mc_level_brick_order(mc_level_t *L, int stage);
/* Launch one fused stage on the current device state (for timing).  The
   prim stage replaces the state by its clean and the update stage advances
   it, both in place; the other stages leave it unchanged. */
int mc_level_run_stage(mc_level_t *L, int stage);
/* Fused mode: time every candidate brick of every stage of every kernel
   variant (reps launches each) on the data of one coarse step and keep the
   fastest, then the brick order (zgroup 0, 2, 4, 8) of each stage, then
   select the faster kernel of each alternative
   (the stage variant below; mc_level_stage_variant reports it); the state
   and the clock are restored.  bricks (11 x 3 ints) and best_ms (11) receive every stage's
   tiling and the tuned stages' time per launch (0 for est, which is not
   tuned); either may be NULL. */
int mc_level_autotune(mc_level_t *L, int reps, int *bricks, double *best_ms);
/* Fused mode: kernel variants of the step, bit flags (0, the default, runs
   the stages tflux, trans1c and final): MC_VARIANT_SPLIT_TRACE runs the
   traces and the first-stage fluxes as two launches (trace, flux1) instead
   of one (tflux); MC_VARIANT_FACE_TRANS1 the transverse-flux stage face by
   face (trans1) instead of cell by cell (trans1c); MC_VARIANT_CELL_FINAL
   the final-flux stage cell by cell (finalc) instead of face by face
   (final).  Results are the same bit for bit. */
#define MC_VARIANT_SPLIT_TRACE 1
#define MC_VARIANT_FACE_TRANS1 2
#define MC_VARIANT_CELL_FINAL 4
int mc_level_set_stage_variant(mc_level_t *L, int flags);
int mc_level_stage_variant(mc_level_t *L);  /* the flags in use, -1 on error */

/* ---- AMR coupling of a coarse and a fine level ----
   The flux register of a fine level (AMReX FluxRegister as the reference
   uses it: defined at Source/Castro.cpp:69, CrseInit / FineAdd at
   Source/Castro_advance.cpp:82-88, ClearInternalBorders + Reflux in
   Castro::reflux, Source/Castro.cpp:451-486) and the average-down of
   Castro::avgDown (Source/Castro.cpp:488-505, amrex::average_down).
   fine_boxes: nfine x 6 ints (lo, hi) in the fine index space, each
   coarsenable by ratio and inside the refined domain crse_n_cell * ratio.
   Fluxes are what store_flux writes (dt * F * face area), face-centred FABs
   in direction dir (0, 1, 2), NVAR components, Fortran order.  All pointers
   are device pointers; calls run on the stream of mc_set_stream. */
typedef struct mc_flux_register mc_flux_register_t;
mc_flux_register_t *mc_flux_register_create(int nfine, const int *fine_boxes, const int *ratio,
                                            const int *crse_n_cell, int bc);
void mc_flux_register_destroy(mc_flux_register_t *R);
/* doubles held: per fine box, direction and side (low, high) in that order,
   the coarse faces of the coarsened box on that side x NVAR, Fortran order */
long mc_flux_register_size(const mc_flux_register_t *R);
int mc_flux_register_get(mc_flux_register_t *R, double *host);
int mc_flux_register_set_val(mc_flux_register_t *R, double v);
/* register := mult * flux on every register face the coarse flux FAB holds */
int mc_flux_register_crse_init(mc_flux_register_t *R, const double *flux, const int *flux_lo, const int *flux_hi,
                               int dir, double mult);
/* register += mult * flux summed over the ratio^2 fine faces of each coarse
   face, on both dir-faces of fine box fine_box (its flux FAB must hold them) */
int mc_flux_register_fine_add(mc_flux_register_t *R, int fine_box, const double *flux, const int *flux_lo,
                              const int *flux_hi, int dir, double mult);
/* zero the faces where two fine boxes meet (periodic images included) */
int mc_flux_register_clear_internal_borders(mc_flux_register_t *R);
/* state(c) += -scale * reg / vol(c) for the coarse cell c just below a low
   face and += scale * reg / vol(c) just above a high face, for the cells c
   of box (lo, hi); faces in the order x, y, z low, then x, y, z high */
int mc_flux_register_reflux(mc_flux_register_t *R, double *state, const int *state_lo, const int *state_hi,
                            const int *lo, const int *hi, const double *vol, const int *vol_lo, const int *vol_hi,
                            double scale);
/* crse(c) = (sum of the ratio^3 fine cells of c) / ratio^3 for the coarse
   cells c of box (lo, hi), ncomp components */
int mc_average_down(const double *fine, const int *fine_lo, const int *fine_hi, double *crse, const int *crse_lo,
                    const int *crse_hi, const int *lo, const int *hi, const int *ratio, int ncomp);
/* crse_side is the register of the next finer level, which this level's
   steps initialise with their store_flux output (mult -1); fine_side is this
   level's own register, to which the steps of its boxes that are fine boxes
   of the register add their fluxes (mult +1).  NULL detaches.  A fused-mode
   level feeds the register from its final fluxes (whole-level face arrays,
   no z-slabs); as a coarse side its step stops after the advance's
   clean_state, and mc_level_post_timestep applies post_timestep's clean
   after the reflux and average-down, as Castro::post_timestep does.  A
   level with registers steps through mc_level_advance and
   mc_level_post_timestep (mc_level_run refuses). */
int mc_level_set_flux_registers(mc_level_t *L, mc_flux_register_t *crse_side, mc_flux_register_t *fine_side);
/* Castro::reflux into this coarse level's state, then the register is zeroed */
int mc_level_reflux(mc_level_t *L, mc_flux_register_t *R);
/* Castro::avgDown from the given boxes of the fine level (all if nfine < 0);
   either level may be in fused mode */
int mc_level_average_down(mc_level_t *fine, mc_level_t *crse, const int *ratio, int nfine, const int *fine_box_ids);

/* ---- batched advance over caller-owned FABs ----
   Castro::advance followed by post_timestep's clean_state
   (Source/Castro_advance.cpp:11-100, Source/Castro.cpp:343-369) for every box
   of a single level at once, replacing the per-box construct_hydro_source
   loop (Source/Castro_hydro.cpp:31-397) with one launch per stage over all
   boxes.  The caller keeps its MultiFabs: each work item names a valid box
   and the device FABs holding it, S_old (read) and S_new (written), each as
   (pointer, FAB lo, FAB hi) in Fortran order with NVAR components, i
   fastest; a FAB may be larger than its box (ghost cells are neither read
   nor written) and S_new may alias S_old.  The boxes must tile [0, n_cell)
   exactly once.  Ghost data come from the boxes themselves (periodic images
   or first-order extrapolation at the domain boundary), as FillPatch
   provides them (Source/Castro_advance.cpp:62).  S_new equals what the
   per-box path (and mc_level in either mode) computes, bit for bit.
   dt_est (may be NULL) receives Castro::estTimeStep of S_new, cfl * min
   dx/(c+|u|) (Source/Castro.cpp:175-216).  Runs on the stream given to
   mc_set_stream and returns after it has drained.  The engine built for the
   last geometry and box list is kept for the next call (calls from several
   threads are serialised); mc_advance_batch_release frees it. */
typedef struct {
    int lo[3], hi[3];                               /* valid box */
    const double *s_old; int old_lo[3], old_hi[3];  /* S_old FAB */
    double *s_new; int new_lo[3], new_hi[3];        /* S_new FAB */
} mc_batch_box_t;
int mc_advance_batch(const int *n_cell, const double *prob_lo, const double *prob_hi, int bc, int nbox,
                     const mc_batch_box_t *work, double dt, double *dt_est);
void mc_advance_batch_release(void);

/* ---- ensembles of independent single-box instances ----
   Every instance is its own domain [0, n_cell) with its own problem constants
   and time-step control; the fused engine advances all of them with one
   launch per stage.  Instance state: 20 components, Fortran order over the
   whole instance box (i fastest, component slowest).  The fused engine (this,
   MC_MODE_FUSED levels, mc_advance_batch, the adjoint) takes instances of
   fewer than 2^29 cells; with its face arrays covering whole instances (no
   z-slabs, and always for the adjoint) also fewer than 2^28 cells including
   one ghost layer, (nx+2)(ny+2)(nz+2) < 2^28.  Larger ones are refused with
   an error before any device memory is taken. */
typedef struct mc_ensemble mc_ensemble_t;
/* n_cell, prob_lo, prob_hi: 3 per instance; bc: one per instance */
mc_ensemble_t *mc_ensemble_create(int ninst, const int *n_cell, const double *prob_lo, const double *prob_hi,
                                  const int *bc);
/* the same with the z-slab planes of the engine (as mc_level_set_fused_slab;
   the default of mc_ensemble_create is -1) */
mc_ensemble_t *mc_ensemble_create_slabs(int ninst, const int *n_cell, const double *prob_lo, const double *prob_hi,
                                        const int *bc, int planes);
int mc_ensemble_slab(const mc_ensemble_t *E);
/* device bytes the fused engine holds for these instances (state, cell
   arrays, face arrays) with z-slabs of `planes` cell planes (0: whole
   instances); a host computation, no device needed */
long mc_fused_bytes(int ninst, const int *n_cell, int planes);
void mc_ensemble_destroy(mc_ensemble_t *E);
int mc_ensemble_ninst(const mc_ensemble_t *E);
double *mc_ensemble_state_ptr(mc_ensemble_t *E, int m);
/* problem constants: one value per instance */
int mc_ensemble_init_data(mc_ensemble_t *E, const double *p_ambient, const double *dens_ambient,
                          const double *exp_energy, const double *r_init, int nsub);
int mc_ensemble_set_state(mc_ensemble_t *E, int m, const double *host);
int mc_ensemble_get_state(mc_ensemble_t *E, int m, double *host);
/* nsteps coarse steps of every instance, each with its own dt */
int mc_ensemble_run(mc_ensemble_t *E, int nsteps, double stop_time, int use_graph);
int mc_ensemble_clock(mc_ensemble_t *E, int m, double *dt, double *time, int *steps);
/* as mc_level_autotune, over every instance */
int mc_ensemble_autotune(mc_ensemble_t *E, int reps, int *bricks, double *best_ms);
/* as mc_level_set_stage_variant / mc_level_stage_variant */
int mc_ensemble_set_stage_variant(mc_ensemble_t *E, int flags);
int mc_ensemble_stage_variant(const mc_ensemble_t *E);

/* ---- discrete adjoint of the CTU step (one single-box instance) ----
   Forward: U0 -> U_N over nsteps steps of the fused path with a frozen dt
   sequence, checkpointing the state at the start of every step.  Backward:
   given the adjoint of U_N, recompute each step from its checkpoint and
   apply the hand-written reverse of every stage, returning the adjoint of
   U0.  The differentiated components are rho, rho u, rho v, rho w, rho E
   and rho e; UTEMP (only the EOS Newton initial guess) and the species
   (whose influence through abar/zbar after normalize_species is of the
   order of its 1e-30 floor) are held constant.  All state arrays are
   device pointers, NVAR x cells in Fortran order; calls run on the stream
   given to mc_set_stream and return after it has drained. */
typedef struct mc_adjoint mc_adjoint_t;
mc_adjoint_t *mc_adjoint_create(const int *n_cell, const double *prob_lo, const double *prob_hi, int bc,
                                int max_steps);
void mc_adjoint_destroy(mc_adjoint_t *A);
/* compute_dt != 0: dts[0..nsteps-1] receive dt_0 = cfl * estdt(U0) and
   dt_{n+1} = min(cfl * estdt(U_{n+1}), 1.1 dt_n); otherwise dts are used. */
int mc_adjoint_forward(mc_adjoint_t *A, const double *U0, double *U_out, int nsteps, double *dts, int compute_dt);
/* Ub0 = (dU_N/dU0)^T Ub_final for the last forward pass */
int mc_adjoint_backward(mc_adjoint_t *A, const double *Ub_final, double *Ub0);
int mc_adjoint_steps(const mc_adjoint_t *A);
/* The reverse pass needs every step's intermediates (primitives, traces,
   fluxes).  By default, when max_steps sets of them fit in half the free
   device memory, the forward pass keeps them (one set per step) and the
   reverse pass reads them; otherwise each step is recomputed from its
   checkpoint.  keep = 0 selects recomputation, keep != 0 keeping (an error
   if the sets do not fit).  Either change drops the last forward pass. */
int mc_adjoint_keep_intermediates(mc_adjoint_t *A, int keep);
int mc_adjoint_keeps_intermediates(const mc_adjoint_t *A);

#ifdef __cplusplus
}
#endif
#endif /* MC_CASTRO_H */
