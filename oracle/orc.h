/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C restatement of mini-Castro's unsplit CTU hydro path, used as the
 * checker for the HIP product.  Only tests/, the smoke check and the
 * benchmark's CPU-baseline leg may load it; the product never links it.
 *
 * Every routine follows the reference Fortran/C++ routine named in its
 * comment (file:line relative to the reference tree) statement by statement,
 * in the same floating-point evaluation order, compiled with
 * -ffp-contract=off so that no multiply-add is fused.
 *
 * Conventions:
 *   - component indices are 0-based (URHO = 0 ... ) where the reference is
 *     1-based (Source/Castro_nd.F90:14-57);
 *   - spatial indices are the global AMReX cell indices, with every array
 *     passed as (ptr, lo[3], hi[3]) in Fortran order: i fastest, component
 *     slowest (Source/Castro.H:19).
 */
#ifndef ORC_H
#define ORC_H

/* Scalar type of the restated arithmetic: double; the tangent-linear build
   (compiled as C++ with ORC_TL) substitutes a forward-mode dual number, the
   op-count build (C++ with ORC_CNT) a double that counts its operations. */
#ifdef ORC_TL
#include "orc_dual.h"
typedef orc_dual real;
#elif defined(ORC_CNT)
#include "orc_count.h"
typedef orc_cnt real;
#else
typedef double real;
#endif

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_NSPEC = 13, ORC_NVAR = 20, ORC_QVAR = 21, ORC_NQAUX = 4, ORC_NGDNV = 6 };
enum { URHO = 0, UMX, UMY, UMZ, UEDEN, UEINT, UTEMP, UFS };
enum { QRHO = 0, QU, QV, QW, QGAME, QPRES, QREINT, QTEMP, QFS };
enum { QGAMC = 0, QC, QDPDR, QDPDE };
enum { GDRHO = 0, GDU, GDV, GDW, GDPRES, GDGAME };
enum { EOS_INPUT_RT = 1, EOS_INPUT_RE = 2, EOS_INPUT_RP = 3 };

/* eos_t, Source/eos.F90:18-58, plus the last table indices and the Newton
   iteration count (diagnostics the reference does not expose). */
typedef struct {
    real rho, T, p, e, h, s, abar, zbar;
    real dpdT, dpdr, dedT, dedr, dhdT, dhdr, dsdT, dsdr, dpde, dpdr_e;
    real cv, cp, xne, xnp, eta, pele, ppos, mu, mu_e, y_e, gam1, cs;
    real dpdA, dpdZ, dedA, dedZ;
    int iat, jat, niter, pad_;
} orc_eos_t;

/* EOS (Source/eos.F90) */
int orc_eos_init(const char *path);           /* text table, Source/eos.F90:580-734 */
void orc_eos_init_arrays(const double *tab);   /* 21 x 541 x 201, column order of eos.F90:82-94 */
void orc_eos_finalize(void);
void orc_eos(int input, orc_eos_t *state);     /* Source/eos.F90:141-576 */
void orc_eos_grid(double *t201, double *d541); /* the temperature / density grids */

/* Per-box kernels, same argument lists as Source/Castro_F.H */
void orc_ctoprim(const int *lo, const int *hi,
                 const real *u, const int *u_lo, const int *u_hi,
                 real *q, const int *q_lo, const int *q_hi,
                 real *qaux, const int *qa_lo, const int *qa_hi);
void orc_divu(const int *lo, const int *hi,
              const real *q, const int *q_lo, const int *q_hi,
              const real *dx, real *div, const int *d_lo, const int *d_hi);
void orc_uflatten(int i, int j, int k, const real *q, const int *q_lo, const int *q_hi,
                  real *flatn);
void orc_ppm_reconstruct(const real *s5, real flatn, real *sm, real *sp);
void orc_ppm_int_profile(real sm, real sp, real sc, real u, real c, real dtdx,
                         real *Ip3, real *Im3);
void orc_trace_ppm(const int *lo, const int *hi, const int *vlo, const int *vhi, int idir,
                   const real *q, const int *q_lo, const int *q_hi,
                   const real *qaux, const int *qa_lo, const int *qa_hi,
                   real *qm, const int *qm_lo, const int *qm_hi,
                   real *qp, const int *qp_lo, const int *qp_hi,
                   const int *domlo, const int *domhi, const real *dx, real dt);
void orc_compute_flux(const int *lo, const int *hi,
                      const real *ql, const int *ql_lo, const int *ql_hi,
                      const real *qr, const int *qr_lo, const int *qr_hi,
                      real *flx, const int *f_lo, const int *f_hi,
                      real *qint, const int *qi_lo, const int *qi_hi,
                      real *qgdnv, const int *qg_lo, const int *qg_hi,
                      const real *qaux, const int *qa_lo, const int *qa_hi, int idir);
void orc_trans1(const int *lo, const int *hi, int idir1, int idir2,
                const real *q2m, const int *q2m_lo, const int *q2m_hi,
                real *q2mo, const int *q2mo_lo, const int *q2mo_hi,
                const real *q2p, const int *q2p_lo, const int *q2p_hi,
                real *q2po, const int *q2po_lo, const int *q2po_hi,
                const real *qaux, const int *qa_lo, const int *qa_hi,
                const real *f1, const int *f1_lo, const int *f1_hi,
                const real *q1, const int *q1_lo, const int *q1_hi, real cdtdx);
void orc_trans2(const int *lo, const int *hi, int idir1, int idir2, int idir3,
                const real *qm, const int *qm_lo, const int *qm_hi,
                real *qmo, const int *qmo_lo, const int *qmo_hi,
                const real *qp, const int *qp_lo, const int *qp_hi,
                real *qpo, const int *qpo_lo, const int *qpo_hi,
                const real *qaux, const int *qa_lo, const int *qa_hi,
                const real *f2, const int *f2_lo, const int *f2_hi,
                const real *f3, const int *f3_lo, const int *f3_hi,
                const real *q2, const int *q2_lo, const int *q2_hi,
                const real *q3, const int *q3_lo, const int *q3_hi,
                real cdtdx1, real cdtdx2, real cdtdx3);
void orc_apply_av(const int *lo, const int *hi, int idir, const real *dx,
                  const real *div, const int *d_lo, const int *d_hi,
                  const real *u, const int *u_lo, const int *u_hi,
                  real *flux, const int *f_lo, const int *f_hi);
void orc_normalize_species_fluxes(const int *lo, const int *hi,
                                  real *flux, const int *f_lo, const int *f_hi);
void orc_store_flux(const int *lo, const int *hi,
                    real *fout, const int *fo_lo, const int *fo_hi,
                    const real *fin, const int *fi_lo, const int *fi_hi,
                    const real *area, const int *a_lo, const int *a_hi, real dt);
void orc_fill_hydro_source(const int *lo, const int *hi,
                           const real *u, const int *u_lo, const int *u_hi,
                           const real *q, const int *q_lo, const int *q_hi,
                           real *src, const int *s_lo, const int *s_hi,
                           const real *f1, const int *f1_lo, const int *f1_hi,
                           const real *f2, const int *f2_lo, const int *f2_hi,
                           const real *f3, const int *f3_lo, const int *f3_hi,
                           const real *qx, const int *qx_lo, const int *qx_hi,
                           const real *qy, const int *qy_lo, const int *qy_hi,
                           const real *qz, const int *qz_lo, const int *qz_hi,
                           const real *a1, const int *a1_lo, const int *a1_hi,
                           const real *a2, const int *a2_lo, const int *a2_hi,
                           const real *a3, const int *a3_lo, const int *a3_hi,
                           const real *vol, const int *v_lo, const int *v_hi,
                           const real *dx, real dt);
void orc_estdt(const int *lo, const int *hi, const real *u, const int *u_lo, const int *u_hi,
               const real *dx, real *dt);
void orc_enforce_minimum_density(const int *lo, const int *hi, real *u, const int *u_lo, const int *u_hi);
void orc_normalize_species(const int *lo, const int *hi, real *u, const int *u_lo, const int *u_hi);
void orc_reset_internal_e(const int *lo, const int *hi, real *u, const int *u_lo, const int *u_hi);
void orc_compute_temp(const int *lo, const int *hi, real *u, const int *u_lo, const int *u_hi);

/* Problem set-up and diagnostics (Source/Prob.F90, Source/Castro_nd.F90:301-400) */
typedef struct {
    real p_ambient, dens_ambient, exp_energy, r_init;
    int nsub, pad_;
} orc_prob_t;
void orc_initdata(const int *lo, const int *hi, real *u, const int *u_lo, const int *u_hi,
                  const real *dx, const real *problo, const real *probhi,
                  const orc_prob_t *prob);
void orc_calculate_blast_radius(const int *lo, const int *hi,
                                const real *u, const int *u_lo, const int *u_hi,
                                const real *dx, const real *problo, const real *probhi,
                                real *blast_mass, real *blast_radius, real max_density);
void orc_denerror(const int *lo, const int *hi, signed char *tag, const int *t_lo, const int *t_hi,
                  const real *den, const int *d_lo, const int *d_hi,
                  signed char set, signed char clear);

/* Level driver (Source/Castro_advance.cpp, Castro_hydro.cpp, Castro.cpp) */
typedef struct orc_level orc_level_t;
enum { ORC_BC_PERIODIC = 0, ORC_BC_OUTFLOW = 1 };
/* boxes: nbox x 6 ints (lo xyz, hi xyz) tiling [0, n_cell) */
orc_level_t *orc_level_create(const int *n_cell, const real *prob_lo, const real *prob_hi,
                              int bc, int nbox, const int *boxes, int nthreads);
void orc_level_destroy(orc_level_t *L);
int orc_level_nbox(const orc_level_t *L);
real *orc_level_state(orc_level_t *L, int ibox);    /* S_new of box ibox, 20 comps, no ghosts */
void orc_level_init_data(orc_level_t *L, const orc_prob_t *prob);
real orc_level_est_time_step(orc_level_t *L);        /* Castro.cpp:175-216 */
real orc_level_initial_dt(orc_level_t *L);           /* Castro.cpp:162-173 (init_shrink) */
void orc_level_advance(orc_level_t *L, real dt);     /* Castro_advance.cpp:11-100 */
void orc_level_post_timestep(orc_level_t *L);          /* Castro.cpp:343-369 */
void orc_level_clean_state_new(orc_level_t *L);        /* Castro.cpp:546-593 on S_new */
/* One coarse time step, Amr::coarseTimeStep semantics: returns dt taken. */
real orc_level_coarse_step(orc_level_t *L, real stop_time);
int orc_level_steps(const orc_level_t *L);
real orc_level_time(const orc_level_t *L);
real orc_level_dt(const orc_level_t *L);
/* hydro source of the last advance, for debugging (box ibox) */
real *orc_level_source(orc_level_t *L, int ibox);

/* AMR coupling of a coarse and a fine level (oracle/orc_amr.c) */
typedef struct orc_flux_register orc_flux_register_t;
orc_flux_register_t *orc_flux_register_create(int nfine, const int *fine_boxes, const int *ratio,
                                              const int *crse_n, int bc);
void orc_flux_register_destroy(orc_flux_register_t *R);
long orc_flux_register_size(const orc_flux_register_t *R);
real *orc_flux_register_data(orc_flux_register_t *R);
void orc_flux_register_set_val(orc_flux_register_t *R, real v);
void orc_flux_register_crse_init(orc_flux_register_t *R, const real *flux, const int *flux_lo,
                                 const int *flux_hi, int dir, real mult);
void orc_flux_register_fine_add(orc_flux_register_t *R, int fine_box, const real *flux, const int *flux_lo,
                                const int *flux_hi, int dir, real mult);
void orc_flux_register_clear_internal_borders(orc_flux_register_t *R);
void orc_flux_register_reflux(orc_flux_register_t *R, real *state, const int *state_lo, const int *state_hi,
                              const int *lo, const int *hi, const real *vol, const int *vol_lo,
                              const int *vol_hi, real scale);
void orc_average_down(const real *fine, const int *fine_lo, const int *fine_hi, real *crse, const int *crse_lo,
                      const int *crse_hi, const int *lo, const int *hi, const int *ratio, int ncomp);
int orc_flux_register_find(const orc_flux_register_t *R, const int *lo, const int *hi);
void orc_level_set_flux_registers(orc_level_t *L, orc_flux_register_t *crse_side, orc_flux_register_t *fine_side);
void orc_level_reflux(orc_level_t *L, orc_flux_register_t *R);
void orc_level_average_down(orc_level_t *fine, orc_level_t *crse, const int *ratio, int nfine, const int *ids);

#ifdef __cplusplus
}
#endif
#endif
