/*
 * ugks_oracle.h -- FP64 CPU restatement of the reference Hermite-UGKS time
 * step (2d/src and 1d/src Fortran sources).  TEST INFRASTRUCTURE ONLY: it is the
 * checker the parity tests, the smoke check and the CPU-baseline leg of the
 * benchmark compare the HIP library against.  The product never links,
 * loads or calls it.
 *
 * The restatement keeps the reference's data structures (per-cell arrays
 * with a ghost ring, per-face fluxes, serial update loop) and its
 * expression order where that is cheap, so it agrees with the compiled
 * reference to round-off.  It takes the same hugks_desc as the library.
 */
#ifndef UGKS_ORACLE_H
#define UGKS_ORACLE_H
#include "hugks.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle2d oracle2d;
typedef struct oracle1d oracle1d;

/* ---- 2D solver ---- */
oracle2d *oracle2d_create(const hugks_desc *d);
void oracle2d_destroy(oracle2d *o);
/* one reference step; returns 0 */
int oracle2d_step(oracle2d *o, hugks_step_info *info);
/* w [4*ncell], h,b [ncell*nv*nv], cell c = i + nx*j */
void oracle2d_get(const oracle2d *o, double *w, double *h, double *b);
void oracle2d_set(oracle2d *o, const double *w, const double *h, const double *b);
void oracle2d_time(const oracle2d *o, double *simtime, int32_t *iter, double *dt);
/* heat flux per cell [2*ncell] of get_heat_flux(h,b,u,v,prim(w)) */
void oracle2d_heat_flux(const oracle2d *o, double *q);

/* ---- 2D operators exposed for unit tests (all FP64, nn = nv*nv) ---- */
/* calc_flux_mtforcev (dir 1, vertical face) / calc_flux_mtforceh (dir 2).
 * cell arrays: h,b [nn], sh,sb [2*nn] (x slope then y slope), w[4],
 * len[2] = cell lengths (dx, dy).  Outputs flux[4], fh, fb [nn], already
 * multiplied by face_len as in the reference. */
void oracle2d_flux_interior(const hugks_desc *d, int dir, double dt,
                            double fn, double ft, double face_len,
                            const double *hL, const double *bL,
                            const double *shL, const double *sbL,
                            const double *wL, const double *lenL,
                            const double *hR, const double *bR,
                            const double *shR, const double *sbR,
                            const double *wR, const double *lenR,
                            double *flux, double *fh, double *fb);
/* symmetric_hermite_force_term_2d (variant 0), symmetric_b (1),
 * new_expansion_force_term (2) */
void oracle2d_hermite_force(const hugks_desc *d, int variant, int order,
                            const double *h, const double *b,
                            const double *prim, double gx, double gy,
                            double *force_h, double *force_b,
                            double *err_h, double *err_b);
/* boundary fluxes: kind HUGKS_BC_SPECULAR / HUGKS_BC_DIFFUSE, dir 1 (vertical
 * face, x-normal) or 2, rot +1 (min side) / -1 (max side) */
void oracle2d_flux_wall(const hugks_desc *d, int kind, int dir, int rot,
                        double tw, double dt, double face_len,
                        const double *h, const double *b,
                        const double *sh, const double *sb, double len,
                        double *flux, double *fh, double *fb);

/* ---- 1D solver ---- */
oracle1d *oracle1d_create(const hugks_desc *d);
void oracle1d_destroy(oracle1d *o);
int oracle1d_step(oracle1d *o, hugks_step_info *info);
void oracle1d_get(const oracle1d *o, double *w, double *h, double *b);
void oracle1d_set(oracle1d *o, const double *w, const double *h, const double *b);
void oracle1d_time(const oracle1d *o, double *simtime, int32_t *iter, double *dt);
void oracle1d_heat_flux(const oracle1d *o, double *q);
/* cal_fv: dvh, dvb [nv] for prim (rho, U, lambda) */
void oracle1d_hermite_force(const hugks_desc *d, int order, const double *h,
                            const double *b, const double *prim,
                            double *dvh, double *dvb);

#ifdef __cplusplus
}
#endif
#endif
