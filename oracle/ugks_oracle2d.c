/*
 * ugks_oracle2d.c -- FP64 CPU restatement of the reference 2D Hermite-UGKS
 * step.  TEST INFRASTRUCTURE ONLY (see ugks_oracle.h); never part of the
 * product.  Each routine cites the reference lines it follows.
 *
 * Generalisations of the reference's case switch (both reproduce the
 * reference exactly for its case 5 and case 6):
 *   - boundaries come from desc.bc[] instead of case_id: a periodic pair
 *     uses ghost copies and centred slopes everywhere (case 5 x), any wall
 *     uses one-sided slopes in its boundary cells (case 5 y, case 6);
 *   - accelerations come from desc.accel_* (the caller evaluates the
 *     reference's face_acceleration, 2d/src/flux.f90:12-26, and the cell
 *     acceleration, 2d/src/solver.f90:291-297).
 */
#include "ugks_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MNUM 6
#define MTUM 6

struct oracle2d {
    hugks_desc d;
    int nx, ny, nv, nn;
    double *uspace, *vspace, *weight;  /* [nn] each, 2d/src/velocity.f90:124-130 */
    /* cells with one ghost layer, (i,j) in [0,nx+1]x[0,ny+1],
     * 2d/src/initialization.f90:39 */
    double *w;              /* [4 * ncg] */
    double *h, *b;          /* [nn * ncg] */
    double *sh, *sb;        /* [2 * nn * ncg], dir-major within a cell */
    /* faces: vface(i,j) i in [1,nx+1], j in [1,ny]; hface(i,j) i in [1,nx],
     * j in [1,ny+1], 2d/src/initialization.f90:40-41 */
    double *vflux, *vfh, *vfb;
    double *hflux, *hfh, *hfb;
    double *accel_cell, *accel_v, *accel_h;
    double dt, simtime;
    int iter;
    double res[4];
    /* scratch */
    double *t0, *t1, *t2, *t3, *t4, *t5, *t6, *t7;
};

#define CI(o, i, j) ((size_t)(i) + (size_t)((o)->nx + 2) * (size_t)(j))
#define VI(o, i, j) ((size_t)((i) - 1) + (size_t)((o)->nx + 1) * (size_t)((j) - 1))
#define HI(o, i, j) ((size_t)((i) - 1) + (size_t)((o)->nx) * (size_t)((j) - 1))

/* ------------------------------------------------------------------ */
/* gas theory, 2d/src/gastheory2d.f90                                   */
/* ------------------------------------------------------------------ */

/* reduced_maxwell, 2d/src/gastheory2d.f90:6-13 */
static void reduced_maxwell(const hugks_desc *d, int nn, double *H, double *B,
                            const double *vn, const double *vt, const double *prim)
{
    for (int k = 0; k < nn; ++k) {
        double cu = vn[k] - prim[1], cv = vt[k] - prim[2];
        H[k] = prim[0] * (prim[3] / d->pi) * exp(-prim[3] * (cu * cu + cv * cv));
        B[k] = H[k] * d->innerk / (2.0 * prim[3]);
    }
}

/* shakhov_part, 2d/src/gastheory2d.f90:15-26 */
static void shakhov_part(const hugks_desc *d, int nn, const double *H, const double *B,
                         const double *vn, const double *vt, const double *qf,
                         const double *prim, double *Hp, double *Bp)
{
    double c = d->shakhov_coef * (1.0 - d->prandtl) * prim[3] * prim[3] / prim[0];
    for (int k = 0; k < nn; ++k) {
        double cu = vn[k] - prim[1], cv = vt[k] - prim[2];
        double cq = cu * qf[0] + cv * qf[1];
        double c2 = 2.0 * prim[3] * (cu * cu + cv * cv);
        Hp[k] = c * cq * (c2 + d->innerk - 5.0) * H[k];
        Bp[k] = c * cq * (c2 + d->innerk - 3.0) * B[k];
    }
}

/* get_conserved, 2d/src/gastheory2d.f90:28-36 */
static void get_conserved(const hugks_desc *d, const double *prim, double *w)
{
    w[0] = prim[0];
    w[1] = prim[0] * prim[1];
    w[2] = prim[0] * prim[2];
    w[3] = 0.5 * prim[0] / prim[3] / (d->gamma - 1.0) +
           0.5 * prim[0] * (prim[1] * prim[1] + prim[2] * prim[2]);
}

/* get_primary, 2d/src/gastheory2d.f90:38-46 */
static void get_primary(const hugks_desc *d, const double *w, double *prim)
{
    prim[0] = w[0];
    prim[1] = w[1] / w[0];
    prim[2] = w[2] / w[0];
    prim[3] = 0.5 * w[0] / (d->gamma - 1.0) /
              (w[3] - 0.5 * (w[1] * w[1] + w[2] * w[2]) / w[0]);
}

/* global_frame, 2d/src/gastheory2d.f90:48-57 */
static void global_frame(double *w, double cosa, double sina)
{
    double a = w[1], b = w[2];
    w[1] = a * cosa - b * sina;
    w[2] = a * sina + b * cosa;
}

/* local_frame, 2d/src/gastheory2d.f90:59-68 */
static void local_frame(const double *w, double cosa, double sina, double *out)
{
    out[0] = w[0];
    out[1] = w[1] * cosa + w[2] * sina;
    out[2] = w[2] * cosa - w[1] * sina;
    out[3] = w[3];
}

/* get_tau, 2d/src/gastheory2d.f90:87-97 */
static double get_tau(const hugks_desc *d, const double *prim)
{
    if (d->tau_model == HUGKS_TAU_MU_OVER_RHO) return d->muref / prim[0];
    return d->muref * 2.0 * pow(prim[3], 1.0 - d->omega) / prim[0];
}

/* get_heat_flux, 2d/src/gastheory2d.f90:121-129 */
static void get_heat_flux(int nn, const double *weight, const double *h, const double *b,
                          const double *vn, const double *vt, const double *prim, double *qf)
{
    double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
    for (int k = 0; k < nn; ++k) {
        double cu = vn[k] - prim[1], cv = vt[k] - prim[2];
        double c2 = cu * cu + cv * cv;
        s1 += weight[k] * cu * c2 * h[k];
        s2 += weight[k] * cu * b[k];
        s3 += weight[k] * cv * c2 * h[k];
        s4 += weight[k] * cv * b[k];
    }
    qf[0] = 0.5 * (s1 + s2);
    qf[1] = 0.5 * (s3 + s4);
}

/* ------------------------------------------------------------------ */
/* flux helpers, 2d/src/flux.f90:691-804                                */
/* ------------------------------------------------------------------ */

/* micro_slope, 2d/src/flux.f90:711-722 */
static void micro_slope(const hugks_desc *d, const double *prim, const double *sw, double *a)
{
    double K = d->innerk;
    a[3] = 4.0 * prim[3] * prim[3] / (K + 2.0) / prim[0] *
           (2.0 * sw[3] - 2.0 * prim[1] * sw[1] - 2.0 * prim[2] * sw[2] +
            sw[0] * (prim[1] * prim[1] + prim[2] * prim[2] - 0.5 * (K + 2.0) / prim[3]));
    a[2] = 2.0 * prim[3] / prim[0] * (sw[2] - prim[2] * sw[0]) - prim[2] * a[3];
    a[1] = 2.0 * prim[3] / prim[0] * (sw[1] - prim[1] * sw[0]) - prim[1] * a[3];
    a[0] = sw[0] / prim[0] - prim[1] * a[1] - prim[2] * a[2] -
           0.5 * (prim[1] * prim[1] + prim[2] * prim[2] + 0.5 * (K + 2.0) / prim[3]) * a[3];
}

/* calc_moment, 2d/src/flux.f90:730-762 */
static void calc_moment(const hugks_desc *d, const double *prim, double *Mu, double *Mv,
                        double *Mxi, double *MuL, double *MuR)
{
    double U = prim[1], V = prim[2], lam = prim[3], K = d->innerk;
    MuL[0] = 0.5 * erfc(-sqrt(lam) * U);
    MuL[1] = U * MuL[0] + 0.5 * exp(-lam * U * U) / sqrt(d->pi * lam);
    MuR[0] = 0.5 * erfc(sqrt(lam) * U);
    MuR[1] = U * MuR[0] - 0.5 * exp(-lam * U * U) / sqrt(d->pi * lam);
    for (int i = 2; i <= MNUM; ++i) {
        MuL[i] = U * MuL[i - 1] + 0.5 * (i - 1) * MuL[i - 2] / lam;
        MuR[i] = U * MuR[i - 1] + 0.5 * (i - 1) * MuR[i - 2] / lam;
    }
    for (int i = 0; i <= MNUM; ++i) Mu[i] = MuL[i] + MuR[i];
    Mv[0] = 1.0;
    Mv[1] = V;
    for (int i = 2; i <= MTUM; ++i) Mv[i] = V * Mv[i - 1] + 0.5 * (i - 1) * Mv[i - 2] / lam;
    Mxi[0] = 1.0;
    Mxi[1] = 0.5 * K / lam;
    Mxi[2] = (K * K + 2.0 * K) / (4.0 * lam * lam);
}

/* moment_uv, 2d/src/flux.f90:772-782 */
static void moment_uv(const double *Mu, const double *Mv, const double *Mxi,
                      int alpha, int beta, int theta, double *m)
{
    m[0] = Mu[alpha] * Mv[beta] * Mxi[theta / 2];
    m[1] = Mu[alpha + 1] * Mv[beta] * Mxi[theta / 2];
    m[2] = Mu[alpha] * Mv[beta + 1] * Mxi[theta / 2];
    m[3] = 0.5 * (Mu[alpha + 2] * Mv[beta] * Mxi[theta / 2] +
                  Mu[alpha] * Mv[beta + 2] * Mxi[theta / 2] +
                  Mu[alpha] * Mv[beta] * Mxi[(theta + 2) / 2]);
}

/* moment_au, 2d/src/flux.f90:792-804 */
static void moment_au(const double *a, const double *Mu, const double *Mv, const double *Mxi,
                      int alpha, int beta, double *m)
{
    double t1[4], t2[4], t3[4], t4[4], t5[4], t6[4];
    moment_uv(Mu, Mv, Mxi, alpha + 0, beta + 0, 0, t1);
    moment_uv(Mu, Mv, Mxi, alpha + 1, beta + 0, 0, t2);
    moment_uv(Mu, Mv, Mxi, alpha + 0, beta + 1, 0, t3);
    moment_uv(Mu, Mv, Mxi, alpha + 2, beta + 0, 0, t4);
    moment_uv(Mu, Mv, Mxi, alpha + 0, beta + 2, 0, t5);
    moment_uv(Mu, Mv, Mxi, alpha + 0, beta + 0, 2, t6);
    for (int c = 0; c < 4; ++c)
        m[c] = a[0] * t1[c] + a[1] * t2[c] + a[2] * t3[c] + 0.5 * a[3] * t4[c] +
               0.5 * a[3] * t5[c] + 0.5 * a[3] * t6[c];
}

/* add_force_compatibility, 2d/src/flux.f90:691-705 */
static void add_force_compatibility(int nn, const double *weight, double *sw, const double *prim,
                                    const double *vn, const double *vt, const double *H0,
                                    const double *B0, double fn, double ft)
{
    double s2 = 0, s3 = 0, s4 = 0;
    for (int k = 0; k < nn; ++k) {
        double g = fn * (vn[k] - prim[1]) + ft * (vt[k] - prim[2]);
        s2 += weight[k] * vn[k] * prim[3] * g * H0[k];
        s3 += weight[k] * vt[k] * prim[3] * g * H0[k];
        s4 += weight[k] * prim[3] * g * ((vn[k] * vn[k] + vt[k] * vt[k]) * H0[k] + B0[k]);
    }
    sw[1] = sw[1] + 2.0 * s2;
    sw[2] = sw[2] + 2.0 * s3;
    sw[3] = sw[3] + s4;
}

/* ------------------------------------------------------------------ */
/* interior face flux: calc_flux_mtforcev / calc_flux_mtforceh,         */
/* 2d/src/flux.f90:28-248 and 250-472 (h differs only in the sign of    */
/* the tangential slope, 307-311).                                      */
/* ------------------------------------------------------------------ */
static void flux_interior(const hugks_desc *d, int nn, const double *weight,
                          const double *uspace, const double *vspace, int dir, double dt,
                          double fn, double ft, double face_len,
                          const double *hL, const double *bL, const double *shL,
                          const double *sbL, const double *wL, const double *lenL,
                          const double *hR, const double *bR, const double *shR,
                          const double *sbR, const double *wR, const double *lenR,
                          double *flux, double *fh, double *fb, double *scratch)
{
    double cosa = (dir == 1) ? 1.0 : 0.0, sina = (dir == 1) ? 0.0 : 1.0;
    double *vn = scratch, *vt = scratch + nn, *h = scratch + 2 * nn, *b = scratch + 3 * nn;
    double *shn = scratch + 4 * nn, *sht = scratch + 5 * nn, *sbn = scratch + 6 * nn;
    double *sbt = scratch + 7 * nn, *H0 = scratch + 8 * nn, *B0 = scratch + 9 * nn;
    double *Hp = scratch + 10 * nn, *Bp = scratch + 11 * nn;
    int *delta = (int *)(scratch + 12 * nn);
    int dn = dir - 1, dt_ = 2 - dir; /* slope indices (DIRC, mod(DIRC,2)+1) - 1 */
    double tsign = (dir == 1) ? 1.0 : -1.0;

    /* flux.f90:68-74 */
    for (int k = 0; k < nn; ++k) {
        vn[k] = uspace[k] * cosa + vspace[k] * sina;
        vt[k] = vspace[k] * cosa - uspace[k] * sina;
        double vf = vn[k] - 0.5 * fn * dt;
        delta[k] = (int)((copysign(1.0, vf) + 1.0) / 2.0);
    }
    /* reconstruction, flux.f90:80-87 (307-311 for horizontal faces) */
    double hlL = 0.5 * lenL[dir - 1], hlR = 0.5 * lenR[dir - 1];
    for (int k = 0; k < nn; ++k) {
        double dl = (double)delta[k], dr = (double)(1 - delta[k]);
        h[k] = (hL[k] + hlL * shL[dn * nn + k]) * dl + (hR[k] - hlR * shR[dn * nn + k]) * dr;
        b[k] = (bL[k] + hlL * sbL[dn * nn + k]) * dl + (bR[k] - hlR * sbR[dn * nn + k]) * dr;
        shn[k] = shL[dn * nn + k] * dl + shR[dn * nn + k] * dr;
        sht[k] = tsign * shL[dt_ * nn + k] * dl + tsign * shR[dt_ * nn + k] * dr;
        sbn[k] = sbL[dn * nn + k] * dl + sbR[dn * nn + k] * dr;
        sbt[k] = tsign * sbL[dt_ * nn + k] * dl + tsign * sbR[dt_ * nn + k] * dr;
    }
    /* macroscopic variables at the interface, flux.f90:93-102 */
    double w[4] = {0, 0, 0, 0}, e1 = 0, e2 = 0;
    for (int k = 0; k < nn; ++k) {
        w[0] += weight[k] * h[k];
        w[1] += weight[k] * vn[k] * h[k];
        w[2] += weight[k] * vt[k] * h[k];
        e1 += weight[k] * (vn[k] * vn[k] + vt[k] * vt[k]) * h[k];
        e2 += weight[k] * b[k];
    }
    w[3] = 0.5 * (e1 + e2);
    double prim[4], qf[2];
    get_primary(d, w, prim);
    get_heat_flux(nn, weight, h, b, vn, vt, prim, qf);

    /* aL, aR, flux.f90:107-111 */
    double sw[4], loc[4], aL[4], aR[4], ab[4], aT[4];
    local_frame(wL, cosa, sina, loc);
    for (int c = 0; c < 4; ++c) sw[c] = (w[c] - loc[c]) / (0.5 * lenL[dir - 1]);
    micro_slope(d, prim, sw, aL);
    local_frame(wR, cosa, sina, loc);
    for (int c = 0; c < 4; ++c) sw[c] = (loc[c] - w[c]) / (0.5 * lenR[dir - 1]);
    micro_slope(d, prim, sw, aR);

    /* tangential slope b, flux.f90:116-120 */
    {
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0;
        for (int k = 0; k < nn; ++k) {
            s0 += weight[k] * sht[k];
            s1 += weight[k] * vn[k] * sht[k];
            s2 += weight[k] * vt[k] * sht[k];
            s3 += weight[k] * (vn[k] * vn[k] + vt[k] * vt[k]) * sht[k];
            s4 += weight[k] * sbt[k];
        }
        sw[0] = s0; sw[1] = s1; sw[2] = s2; sw[3] = 0.5 * (s3 + s4);
    }
    micro_slope(d, prim, sw, ab);

    /* time slope, flux.f90:126-135 */
    double Mu[MNUM + 1], MuL[MNUM + 1], MuR[MNUM + 1], Mv[MTUM + 1], Mxi[3];
    double MauL[4], MauR[4], Mbu[4], MauT[4], Mau0[4];
    calc_moment(d, prim, Mu, Mv, Mxi, MuL, MuR);
    moment_au(aL, MuL, Mv, Mxi, 1, 0, MauL);
    moment_au(aR, MuR, Mv, Mxi, 1, 0, MauR);
    moment_au(ab, Mu, Mv, Mxi, 0, 1, Mbu);
    for (int c = 0; c < 4; ++c) sw[c] = -prim[0] * (MauL[c] + MauR[c] + Mbu[c]);
    reduced_maxwell(d, nn, H0, B0, vn, vt, prim);
    add_force_compatibility(nn, weight, sw, prim, vn, vt, H0, B0, fn, ft);
    micro_slope(d, prim, sw, aT);

    /* collision time and time-integration coefficients, flux.f90:140-147 */
    double tau = get_tau(d, prim), Mt[7];
    double ex = exp(-dt / tau);
    Mt[4] = tau * (1.0 - ex);
    Mt[5] = -tau * dt * ex + tau * Mt[4];
    Mt[1] = dt - Mt[4];
    Mt[2] = -tau * Mt[1] + Mt[5];
    Mt[3] = dt * dt / 2.0 - tau * Mt[1];
    Mt[6] = tau * tau * tau - 0.5 * ex * tau * (dt * dt + 2.0 * dt * tau + 2.0 * tau * tau);

    /* conserved flux from g0, flux.f90:152-158 */
    moment_uv(Mu, Mv, Mxi, 1, 0, 0, Mau0);
    moment_au(aL, MuL, Mv, Mxi, 2, 0, MauL);
    moment_au(aR, MuR, Mv, Mxi, 2, 0, MauR);
    moment_au(ab, Mu, Mv, Mxi, 1, 1, Mbu);
    moment_au(aT, Mu, Mv, Mxi, 1, 0, MauT);
    for (int c = 0; c < 4; ++c)
        flux[c] = Mt[1] * prim[0] * Mau0[c] + Mt[2] * prim[0] * (MauL[c] + MauR[c] + Mbu[c]) +
                  Mt[3] * prim[0] * MauT[c];
    /* gravity terms related to g0, flux.f90:164-172 */
    moment_au(aL, MuL, Mv, Mxi, 1, 0, MauL);
    moment_au(aR, MuR, Mv, Mxi, 1, 0, MauR);
    for (int c = 0; c < 4; ++c)
        flux[c] = flux[c] - tau * Mt[2] * prim[0] * fn * (MauL[c] + MauR[c]) -
                  Mt[6] * prim[0] * fn * (MauL[c] + MauR[c]);
    moment_au(ab, Mu, Mv, Mxi, 0, 1, Mbu);
    for (int c = 0; c < 4; ++c)
        flux[c] = flux[c] - tau * Mt[2] * prim[0] * ft * Mbu[c] - Mt[6] * prim[0] * ft * Mbu[c];

    /* g+ and f0 parts, flux.f90:178-202 */
    reduced_maxwell(d, nn, H0, B0, vn, vt, prim);
    shakhov_part(d, nn, H0, B0, vn, vt, qf, prim, Hp, Bp);
    double S[40];
    memset(S, 0, sizeof(S));
    for (int k = 0; k < nn; ++k) {
        double wk = weight[k], un = vn[k], ut = vt[k], c2 = un * un + ut * ut;
        S[0] += wk * un * Hp[k];
        S[1] += wk * un * h[k];
        S[2] += wk * un * un * shn[k];
        S[3] += wk * ut * un * sht[k];
        S[4] += wk * un * un * Hp[k];
        S[5] += wk * un * un * h[k];
        S[6] += wk * un * un * un * shn[k];
        S[7] += wk * un * ut * un * sht[k];
        S[8] += wk * ut * un * Hp[k];
        S[9] += wk * ut * un * h[k];
        S[10] += wk * ut * un * un * shn[k];
        S[11] += wk * ut * ut * un * sht[k];
        S[12] += wk * un * c2 * Hp[k];
        S[13] += wk * un * Bp[k];
        S[14] += wk * un * c2 * h[k];
        S[15] += wk * un * b[k];
        S[16] += wk * un * un * c2 * shn[k];
        S[17] += wk * un * un * sbn[k];
        S[18] += wk * ut * un * c2 * sht[k];
        S[19] += wk * ut * un * sbt[k];
        S[20] += wk * un * shn[k];
        S[21] += wk * un * un * shn[k];
        S[22] += wk * ut * un * shn[k];
        S[23] += wk * c2 * un * shn[k];
        S[24] += wk * un * sbn[k];
        S[25] += wk * un * sht[k];
        S[26] += wk * un * un * sht[k];
        S[27] += wk * ut * un * sht[k];
        S[28] += wk * c2 * un * sht[k];
        S[29] += wk * un * sbt[k];
    }
    flux[0] = flux[0] + Mt[1] * S[0] + Mt[4] * S[1] - Mt[5] * S[2] - Mt[5] * S[3];
    flux[1] = flux[1] + Mt[1] * S[4] + Mt[4] * S[5] - Mt[5] * S[6] - Mt[5] * S[7];
    flux[2] = flux[2] + Mt[1] * S[8] + Mt[4] * S[9] - Mt[5] * S[10] - Mt[5] * S[11];
    flux[3] = flux[3] + Mt[1] * 0.5 * (S[12] + S[13]) + Mt[4] * 0.5 * (S[14] + S[15]) -
              Mt[5] * 0.5 * (S[16] + S[17]) - Mt[5] * 0.5 * (S[18] + S[19]);
    flux[0] = flux[0] + fn * Mt[6] * S[20];
    flux[1] = flux[1] + fn * Mt[6] * S[21];
    flux[2] = flux[2] + fn * Mt[6] * S[22];
    flux[3] = flux[3] + fn * Mt[6] * 0.5 * (S[23] + S[24]);
    flux[0] = flux[0] + ft * Mt[6] * S[25];
    flux[1] = flux[1] + ft * Mt[6] * S[26];
    flux[2] = flux[2] + ft * Mt[6] * S[27];
    flux[3] = flux[3] + ft * Mt[6] * 0.5 * (S[28] + S[29]);

    /* distribution flux, flux.f90:207-236 */
    for (int k = 0; k < nn; ++k) {
        double un = vn[k], ut = vt[k], c2 = un * un + ut * ut;
        double H = H0[k], Bm = B0[k];
        double dl = (double)delta[k], dr = (double)(1 - delta[k]);
        double gLh = aL[0] * H + aL[1] * un * H + aL[2] * ut * H + 0.5 * aL[3] * (c2 * H + Bm);
        double gRh = aR[0] * H + aR[1] * un * H + aR[2] * ut * H + 0.5 * aR[3] * (c2 * H + Bm);
        double gbh = ab[0] * H + ab[1] * un * H + ab[2] * ut * H + 0.5 * ab[3] * (c2 * H + Bm);
        double gTh = aT[0] * H + aT[1] * un * H + aT[2] * ut * H + 0.5 * aT[3] * (c2 * H + Bm);
        double gLb = aL[0] * Bm + aL[1] * un * Bm + aL[2] * ut * Bm + 0.5 * aL[3] * (c2 * Bm + Mxi[2] * H);
        double gRb = aR[0] * Bm + aR[1] * un * Bm + aR[2] * ut * Bm + 0.5 * aR[3] * (c2 * Bm + Mxi[2] * H);
        double gbb = ab[0] * Bm + ab[1] * un * Bm + ab[2] * ut * Bm + 0.5 * ab[3] * (c2 * Bm + Mxi[2] * H);
        double gTb = aT[0] * Bm + aT[1] * un * Bm + aT[2] * ut * Bm + 0.5 * aT[3] * (c2 * Bm + Mxi[2] * H);
        double xh = Mt[1] * un * (H + Hp[k]) + Mt[2] * un * un * gLh * dl + Mt[2] * un * un * gRh * dr +
                    Mt[2] * un * ut * gbh + Mt[3] * un * gTh + Mt[4] * un * h[k] -
                    Mt[5] * un * un * shn[k] - Mt[5] * un * ut * sht[k];
        double xb = Mt[1] * un * (Bm + Bp[k]) + Mt[2] * un * un * gLb * dl + Mt[2] * un * un * gRb * dr +
                    Mt[2] * un * ut * gbb + Mt[3] * un * gTb + Mt[4] * un * b[k] -
                    Mt[5] * un * un * sbn[k] - Mt[5] * un * ut * sbt[k];
        xh = xh - tau * Mt[2] * fn * un * gLh * dl - tau * Mt[2] * fn * un * gRh * dr + Mt[6] * fn * un * shn[k];
        xb = xb - tau * Mt[2] * fn * un * gLb * dl - tau * Mt[2] * fn * un * gRb * dr + Mt[6] * fn * un * sbn[k];
        xh = xh - tau * Mt[2] * ft * un * gbh + Mt[6] * ft * un * sht[k];
        xb = xb - tau * Mt[2] * ft * un * gbb + Mt[6] * ft * un * sbt[k];
        fh[k] = face_len * xh;
        fb[k] = face_len * xb;
    }
    /* final flux, flux.f90:242-246 */
    global_frame(flux, cosa, sina);
    for (int c = 0; c < 4; ++c) flux[c] = face_len * flux[c];
}

/* ------------------------------------------------------------------ */
/* wall fluxes, 2d/src/flux.f90:474-689                                 */
/* ------------------------------------------------------------------ */
static void flux_wall(const hugks_desc *d, int nn, int nv, const double *weight,
                      const double *uspace, const double *vspace, int kind, int dir,
                      int rot, double tw, double dt, double face_len, const double *hc,
                      const double *bc, const double *shc, const double *sbc, double len,
                      double *flux, double *fh, double *fb, double *scratch)
{
    double cosa = (dir == 1) ? 1.0 : 0.0, sina = (dir == 1) ? 0.0 : 1.0;
    double *vn = scratch, *vt = scratch + nn, *h = scratch + 2 * nn, *b = scratch + 3 * nn;
    double *H0 = scratch + 4 * nn, *B0 = scratch + 5 * nn;
    double *th = scratch + 6 * nn, *tb = scratch + 7 * nn;
    int *delta = (int *)(scratch + 8 * nn);
    int dn = dir - 1;
    for (int k = 0; k < nn; ++k) {
        vn[k] = uspace[k] * cosa + vspace[k] * sina;
        vt[k] = vspace[k] * cosa - uspace[k] * sina;
        delta[k] = (int)((copysign(1.0, vn[k]) * rot + 1.0) / 2.0);
        h[k] = hc[k] - rot * 0.5 * len * shc[dn * nn + k];
        b[k] = bc[k] - rot * 0.5 * len * sbc[dn * nn + k];
    }
    if (kind == HUGKS_BC_DIFFUSE) {
        /* calc_flux_boundary, flux.f90:504-526; bc = (0,0,0,1/Tw) */
        double bcw[4] = {0.0, 0.0, 0.0, 1.0 / tw}, prim[4];
        local_frame(bcw, cosa, sina, prim);
        double R1 = 0, R2 = 0;
        for (int k = 0; k < nn; ++k) {
            R1 += weight[k] * vn[k] * h[k] * (double)(1 - delta[k]);
            double cu = vn[k] - prim[1], cv = vt[k] - prim[2];
            R2 += weight[k] * vn[k] * exp(-prim[3] * (cu * cu + cv * cv)) * (double)delta[k];
        }
        R2 = (prim[3] / d->pi) * R2;
        prim[0] = -R1 / R2;
        reduced_maxwell(d, nn, H0, B0, vn, vt, prim);
        for (int k = 0; k < nn; ++k) {
            h[k] = H0[k] * delta[k] + h[k] * (1 - delta[k]);
            b[k] = B0[k] * delta[k] + b[k] * (1 - delta[k]);
        }
    } else {
        /* calc_flux_specular_vertical / _horizon, flux.f90:589-596 / 659-666 */
        for (int jv = 0; jv < nv; ++jv)
            for (int iu = 0; iu < nv; ++iu) {
                int src = (dir == 1) ? (nv - 1 - iu) + nv * jv : iu + nv * (nv - 1 - jv);
                th[iu + nv * jv] = h[src];
                tb[iu + nv * jv] = b[src];
            }
        for (int k = 0; k < nn; ++k) {
            h[k] = th[k] * delta[k] + h[k] * (1 - delta[k]);
            b[k] = tb[k] * delta[k] + b[k] * (1 - delta[k]);
        }
    }
    /* flux, flux.f90:531-547 */
    double s[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < nn; ++k) {
        s[0] += weight[k] * vn[k] * h[k];
        s[1] += weight[k] * vn[k] * vn[k] * h[k];
        s[2] += weight[k] * vn[k] * vt[k] * h[k];
        s[3] += weight[k] * vn[k] * ((vn[k] * vn[k] + vt[k] * vt[k]) * h[k] + b[k]);
        fh[k] = dt * face_len * (vn[k] * h[k]);
        fb[k] = dt * face_len * (vn[k] * b[k]);
    }
    flux[0] = s[0]; flux[1] = s[1]; flux[2] = s[2]; flux[3] = 0.5 * s[3];
    global_frame(flux, cosa, sina);
    for (int c = 0; c < 4; ++c) flux[c] = dt * face_len * flux[c];
}

/* ------------------------------------------------------------------ */
/* Hermite force operators                                              */
/* ------------------------------------------------------------------ */
static void hermite_basis(int nn, const double *x, int max_order, double *basis)
{
    /* build_probabilists_hermite, 2d/src/hermite_symmetric2d.f90:220-233;
     * basis[n*nn + k] */
    for (int k = 0; k < nn; ++k) basis[k] = 1.0;
    if (max_order == 0) return;
    for (int k = 0; k < nn; ++k) basis[nn + k] = x[k];
    for (int n = 2; n <= max_order; ++n)
        for (int k = 0; k < nn; ++k)
            basis[n * nn + k] = x[k] * basis[(n - 1) * nn + k] - (double)(n - 1) * basis[(n - 2) * nn + k];
}

static void force_symmetric(const hugks_desc *d, int nn, const double *weight,
                            const double *uspace, const double *vspace, int order, int variant_b,
                            const double *h, const double *b, const double *prim, double gx,
                            double gy, double *force_h, double *force_b, double *err_h,
                            double *err_b)
{
    /* symmetric_hermite_force_term_2d, 2d/src/hermite_symmetric2d.f90:23-104,
     * and the symmetric-b form 120-218 */
    double rt = 0.5 / prim[3]; /* Rc/prim(4), Rc = 0.5 (2d/src/data2d.f90:54) */
    double inv_sqrt_rt = 1.0 / sqrt(rt), jac = 1.0 / rt;
    double *ux = malloc(sizeof(double) * nn), *uy = malloc(sizeof(double) * nn);
    double *om = malloc(sizeof(double) * nn);
    double *hx = malloc(sizeof(double) * nn * (order + 2));
    double *hy = malloc(sizeof(double) * nn * (order + 2));
    double *rh = calloc(nn, sizeof(double)), *rb = calloc(nn, sizeof(double));
    double *ch = calloc((order + 2) * (order + 2), sizeof(double));
    double *cb = calloc((order + 2) * (order + 2), sizeof(double));
    double *fact = malloc(sizeof(double) * (order + 2));
    int ld = order + 2;
    for (int k = 0; k < nn; ++k) {
        ux[k] = (uspace[k] - prim[1]) / sqrt(rt);
        uy[k] = (vspace[k] - prim[2]) / sqrt(rt);
        om[k] = exp(-0.5 * (ux[k] * ux[k] + uy[k] * uy[k])) / (2.0 * d->pi);
    }
    hermite_basis(nn, ux, order + 1, hx);
    hermite_basis(nn, uy, order + 1, hy);
    fact[0] = 1.0;
    for (int n = 1; n <= order + 1; ++n) fact[n] = (double)n * fact[n - 1];
    for (int n = 0; n <= order; ++n)
        for (int r = 0; r <= n; ++r) {
            int s = n - r;
            double a = 0, c = 0;
            for (int k = 0; k < nn; ++k) {
                a += weight[k] * h[k] * hx[r * nn + k] * hy[s * nn + k];
                c += weight[k] * b[k] * hx[r * nn + k] * hy[s * nn + k];
            }
            ch[r * ld + s] = jac * a;
            cb[r * ld + s] = jac * c;
        }
    for (int k = 0; k < nn; ++k) { force_h[k] = 0.0; force_b[k] = 0.0; }
    for (int n = 0; n <= order; ++n)
        for (int r = 0; r <= n; ++r) {
            int s = n - r;
            double den = fact[r] * fact[s];
            for (int k = 0; k < nn; ++k) {
                double p = hx[r * nn + k] * hy[s * nn + k];
                rh[k] = rh[k] + om[k] * ch[r * ld + s] * p / den;
                rb[k] = rb[k] + om[k] * cb[r * ld + s] * p / den;
            }
        }
    if (!variant_b) {
        for (int n = 0; n <= order; ++n)
            for (int r = 0; r <= n; ++r) {
                int s = n - r;
                double den = fact[r] * fact[s];
                for (int k = 0; k < nn; ++k) {
                    double g = gx * hx[(r + 1) * nn + k] * hy[s * nn + k] +
                               gy * hx[r * nn + k] * hy[(s + 1) * nn + k];
                    force_h[k] = force_h[k] - inv_sqrt_rt * om[k] * ch[r * ld + s] / den * g;
                    force_b[k] = force_b[k] - inv_sqrt_rt * om[k] * cb[r * ld + s] / den * g;
                }
            }
    } else {
        for (int n = 0; n <= order; ++n)
            for (int r = 0; r <= n + 1; ++r) {
                int s = n + 1 - r;
                double bh = 0.0, bb = 0.0;
                if (r > 0) { bh += (double)r * gx * ch[(r - 1) * ld + s]; bb += (double)r * gx * cb[(r - 1) * ld + s]; }
                if (s > 0) { bh += (double)s * gy * ch[r * ld + s - 1]; bb += (double)s * gy * cb[r * ld + s - 1]; }
                bh = bh / (double)(n + 1);
                bb = bb / (double)(n + 1);
                double den = fact[r] * fact[s];
                for (int k = 0; k < nn; ++k) {
                    double p = hx[r * nn + k] * hy[s * nn + k];
                    force_h[k] = force_h[k] - inv_sqrt_rt * om[k] * (double)(n + 1) * bh * p / den;
                    force_b[k] = force_b[k] - inv_sqrt_rt * om[k] * (double)(n + 1) * bb * p / den;
                }
            }
    }
    double e1 = 0, e2 = 0, n1 = 0, n2 = 0;
    for (int k = 0; k < nn; ++k) {
        e1 += (h[k] - rh[k]) * (h[k] - rh[k]);
        e2 += (b[k] - rb[k]) * (b[k] - rb[k]);
        n1 += h[k] * h[k];
        n2 += b[k] * b[k];
    }
    if (err_h) *err_h = sqrt(e1) / fmax(sqrt(n1), 2.2250738585072014e-308);
    if (err_b) *err_b = sqrt(e2) / fmax(sqrt(n2), 2.2250738585072014e-308);
    free(ux); free(uy); free(om); free(hx); free(hy); free(rh); free(rb);
    free(ch); free(cb); free(fact);
}

static void force_legacy(const hugks_desc *d, int nn, const double *weight, const double *uspace,
                         const double *vspace, int order, const double *h, const double *b,
                         const double *prim, double gx, double gy, double *fh, double *fb)
{
    /* new_expansion_force_term, 2d/src/gravity.f90:189-273 */
    double theta = 1.0 / prim[3];
    double *vx = malloc(sizeof(double) * nn), *vy = malloc(sizeof(double) * nn);
    double *w = malloc(sizeof(double) * nn);
    double *hpx = malloc(sizeof(double) * nn * (order + 2));
    double *hpy = malloc(sizeof(double) * nn * (order + 2));
    double *frac = malloc(sizeof(double) * (order + 1));
    double *dn = malloc(sizeof(double) * (order + 1) * (order + 1));
    for (int k = 0; k < nn; ++k) {
        vx[k] = (uspace[k] - prim[1]) / sqrt(theta) * sqrt(2.0);
        vy[k] = (vspace[k] - prim[2]) / sqrt(theta) * sqrt(2.0);
        w[k] = 1.0 / 2.0 / d->pi * exp(-(vx[k] * vx[k] + vy[k] * vy[k]) / 2.0);
    }
    frac[0] = 1.0;
    for (int i = 1; i <= order; ++i) frac[i] = (double)i * frac[i - 1];
    for (int k = 0; k < nn; ++k) {
        hpx[k] = 1.0; hpy[k] = 1.0; hpx[nn + k] = vx[k]; hpy[nn + k] = vy[k];
    }
    if (order >= 1)
        for (int i = 2; i <= order + 1; ++i)
            for (int k = 0; k < nn; ++k) {
                hpx[i * nn + k] = vx[k] * hpx[(i - 1) * nn + k] - (i - 1.0) * hpx[(i - 2) * nn + k];
                hpy[i * nn + k] = vy[k] * hpy[(i - 1) * nn + k] - (i - 1.0) * hpy[(i - 2) * nn + k];
            }
    for (int pass = 0; pass < 2; ++pass) {
        const double *f = pass ? b : h;
        double *out = pass ? fb : fh;
        for (int i = 0; i <= order; ++i)
            for (int j = 0; j <= i; ++j) {
                double s = 0;
                for (int k = 0; k < nn; ++k) s += f[k] * hpx[j * nn + k] * hpy[(i - j) * nn + k] * weight[k];
                dn[i * (order + 1) + j] = s * 2.0 / theta;
            }
        for (int k = 0; k < nn; ++k) out[k] = 0.0;
        for (int i = 0; i <= order; ++i)
            for (int j = 0; j <= i; ++j)
                for (int k = 0; k < nn; ++k)
                    out[k] = out[k] - sqrt(2.0 / theta) * w[k] * 1.0 / frac[j] / frac[i - j] *
                                          (gx * hpx[(j + 1) * nn + k] * hpy[(i - j) * nn + k] +
                                           gy * hpx[j * nn + k] * hpy[(i - j + 1) * nn + k]) *
                                          dn[i * (order + 1) + j];
    }
    free(vx); free(vy); free(w); free(hpx); free(hpy); free(frac); free(dn);
}

/* ------------------------------------------------------------------ */
/* solver phases, 2d/src/solver.f90                                     */
/* ------------------------------------------------------------------ */
static double *CH(oracle2d *o, int i, int j) { return o->h + (size_t)o->nn * CI(o, i, j); }
static double *CB(oracle2d *o, int i, int j) { return o->b + (size_t)o->nn * CI(o, i, j); }
static double *CSH(oracle2d *o, int i, int j) { return o->sh + (size_t)2 * o->nn * CI(o, i, j); }
static double *CSB(oracle2d *o, int i, int j) { return o->sb + (size_t)2 * o->nn * CI(o, i, j); }
static double *CW(oracle2d *o, int i, int j) { return o->w + (size_t)4 * CI(o, i, j); }

/* timestep, 2d/src/solver.f90:13-61 */
static void timestep(oracle2d *o)
{
    const hugks_desc *d = &o->d;
    double temp = 0.0, area = d->dx * d->dy;
    for (int j = 1; j <= o->ny; ++j)
        for (int i = 1; i <= o->nx; ++i) {
            double prim[4];
            get_primary(d, CW(o, i, j), prim);
            double sos = sqrt(0.5 * d->gamma / prim[3]);
            double pu = fmax(d->umax_dt, fabs(prim[1])) + sos;
            double pv = fmax(d->vmax_dt, fabs(prim[2])) + sos;
            if (d->dt_rule == HUGKS_DT_2D_WALL_NORMAL) temp = fmax(temp, pv / d->dy);
            else temp = fmax(temp, (d->dy * pu + d->dx * pv) / area);
        }
    temp = fmax(temp, d->force_rate);
    double tmin = 1.0 / temp;
    o->dt = d->cfl * tmin / 1.1;
    if (d->clamp_dt) o->dt = fmin(o->dt, d->maxtime - o->simtime);
}

/* bc_periodic_2d, 2d/src/boundary.f90:16-53 */
static void periodic_ghosts(oracle2d *o, int dirc)
{
    size_t nn = o->nn;
    if (dirc == 1) {
        for (int j = 1; j <= o->ny; ++j) {
            int pairs[2][2] = {{0, o->nx}, {o->nx + 1, 1}};
            for (int p = 0; p < 2; ++p) {
                int g = pairs[p][0], s = pairs[p][1];
                memcpy(CW(o, g, j), CW(o, s, j), 4 * sizeof(double));
                memcpy(CH(o, g, j), CH(o, s, j), nn * sizeof(double));
                memcpy(CB(o, g, j), CB(o, s, j), nn * sizeof(double));
                memcpy(CSH(o, g, j), CSH(o, s, j), 2 * nn * sizeof(double));
                memcpy(CSB(o, g, j), CSB(o, s, j), 2 * nn * sizeof(double));
            }
        }
    } else {
        for (int i = 1; i <= o->nx; ++i) {
            int pairs[2][2] = {{0, o->ny}, {o->ny + 1, 1}};
            for (int p = 0; p < 2; ++p) {
                int g = pairs[p][0], s = pairs[p][1];
                memcpy(CW(o, i, g), CW(o, i, s), 4 * sizeof(double));
                memcpy(CH(o, i, g), CH(o, i, s), nn * sizeof(double));
                memcpy(CB(o, i, g), CB(o, i, s), nn * sizeof(double));
                memcpy(CSH(o, i, g), CSH(o, i, s), 2 * nn * sizeof(double));
                memcpy(CSB(o, i, g), CSB(o, i, s), 2 * nn * sizeof(double));
            }
        }
    }
}

/* van Leer slope of one array, 2d/src/solver.f90:146-148 */
static void vanleer(int nn, double smv, const double *L, const double *N, const double *R,
                    double dl, double dr, double *out)
{
    for (int k = 0; k < nn; ++k) {
        double sL = (N[k] - L[k]) / dl, sR = (R[k] - N[k]) / dr;
        out[k] = (copysign(1.0, sR) + copysign(1.0, sL)) * fabs(sR) * fabs(sL) /
                 (fabs(sR) + fabs(sL) + smv);
    }
}

/* interp_inner / interp_boundary, 2d/src/solver.f90:125-159; the conserved
 * slope sw is not read by any later phase and is not formed */
static void interp_cell(oracle2d *o, int iL, int jL, int iN, int jN, int iR, int jR, int dirc,
                        int inner)
{
    size_t nn = o->nn;
    double len = (dirc == 1) ? o->d.dx : o->d.dy;
    double *shN = CSH(o, iN, jN) + (dirc - 1) * nn, *sbN = CSB(o, iN, jN) + (dirc - 1) * nn;
    if (inner) {
        double dl = 0.5 * len + 0.5 * len;
        vanleer(o->nn, o->d.smv, CH(o, iL, jL), CH(o, iN, jN), CH(o, iR, jR), dl, dl, shN);
        vanleer(o->nn, o->d.smv, CB(o, iL, jL), CB(o, iN, jN), CB(o, iR, jR), dl, dl, sbN);
    } else {
        double den = 0.5 * len + 0.5 * len;
        const double *hR = CH(o, iR, jR), *hL = CH(o, iL, jL);
        const double *bR = CB(o, iR, jR), *bL = CB(o, iL, jL);
        for (size_t k = 0; k < nn; ++k) {
            shN[k] = (hR[k] - hL[k]) / den;
            sbN[k] = (bR[k] - bL[k]) / den;
        }
    }
}

/* interpolation, 2d/src/solver.f90:64-122 */
static void interpolation(oracle2d *o)
{
    int nx = o->nx, ny = o->ny;
    int xper = o->d.bc[0] == HUGKS_BC_PERIODIC, yper = o->d.bc[2] == HUGKS_BC_PERIODIC;
    if (xper) {
        periodic_ghosts(o, 1);
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) interp_cell(o, i - 1, j, i, j, i + 1, j, 1, 1);
    } else {
        for (int j = 1; j <= ny; ++j) {
            interp_cell(o, 1, j, 1, j, 2, j, 1, 0);
            interp_cell(o, nx - 1, j, nx, j, nx, j, 1, 0);
        }
        for (int j = 1; j <= ny; ++j)
            for (int i = 2; i <= nx - 1; ++i) interp_cell(o, i - 1, j, i, j, i + 1, j, 1, 1);
    }
    if (yper) {
        periodic_ghosts(o, 2);
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) interp_cell(o, i, j - 1, i, j, i, j + 1, 2, 1);
    } else {
        for (int i = 1; i <= nx; ++i) {
            interp_cell(o, i, 1, i, 1, i, 2, 2, 0);
            interp_cell(o, i, ny - 1, i, ny, i, ny, 2, 0);
        }
        for (int j = 2; j <= ny - 1; ++j)
            for (int i = 1; i <= nx; ++i) interp_cell(o, i, j - 1, i, j, i, j + 1, 2, 1);
    }
}

static void face_interior(oracle2d *o, int dir, int iL, int jL, int iR, int jR, const double *acc,
                          double face_len, double *flux, double *fh, double *fb)
{
    double len[2] = {o->d.dx, o->d.dy};
    flux_interior(&o->d, o->nn, o->weight, o->uspace, o->vspace, dir, o->dt, acc[0], acc[1],
                  face_len, CH(o, iL, jL), CB(o, iL, jL), CSH(o, iL, jL), CSB(o, iL, jL),
                  CW(o, iL, jL), len, CH(o, iR, jR), CB(o, iR, jR), CSH(o, iR, jR),
                  CSB(o, iR, jR), CW(o, iR, jR), len, flux, fh, fb, o->t0);
}

static void face_wall(oracle2d *o, int side, int dir, int rot, int i, int j, double face_len,
                      double *flux, double *fh, double *fb)
{
    double len = dir == 1 ? o->d.dx : o->d.dy;
    flux_wall(&o->d, o->nn, o->nv, o->weight, o->uspace, o->vspace, o->d.bc[side], dir, rot,
              o->d.tw[side], o->dt, face_len, CH(o, i, j), CB(o, i, j), CSH(o, i, j),
              CSB(o, i, j), len, flux, fh, fb, o->t0);
}

/* evolution, 2d/src/solver.f90:161-243 */
static void evolution(oracle2d *o)
{
    int nx = o->nx, ny = o->ny;
    size_t nn = o->nn;
    int xper = o->d.bc[0] == HUGKS_BC_PERIODIC, yper = o->d.bc[2] == HUGKS_BC_PERIODIC;
#define VF(i, j) o->vflux + 4 * VI(o, i, j), o->vfh + nn * VI(o, i, j), o->vfb + nn * VI(o, i, j)
#define HF(i, j) o->hflux + 4 * HI(o, i, j), o->hfh + nn * HI(o, i, j), o->hfb + nn * HI(o, i, j)
    if (xper) periodic_ghosts(o, 1);
    if (yper) periodic_ghosts(o, 2);
    /* vertical faces */
    for (int j = 1; j <= ny; ++j) {
        if (!xper) {
            face_wall(o, 0, 1, 1, 1, j, o->d.dy, VF(1, j));
            face_wall(o, 1, 1, -1, nx, j, o->d.dy, VF(nx + 1, j));
        }
        int i0 = xper ? 1 : 2, i1 = xper ? nx + 1 : nx;
        for (int i = i0; i <= i1; ++i)
            face_interior(o, 1, i - 1, j, i, j, o->accel_v + 2 * VI(o, i, j), o->d.dy, VF(i, j));
    }
    /* horizontal faces */
    for (int i = 1; i <= nx; ++i) {
        if (!yper) {
            face_wall(o, 2, 2, 1, i, 1, o->d.dx, HF(i, 1));
            face_wall(o, 3, 2, -1, i, ny, o->d.dx, HF(i, ny + 1));
        }
    }
    {
        int j0 = yper ? 1 : 2, j1 = yper ? ny + 1 : ny;
        for (int j = j0; j <= j1; ++j)
            for (int i = 1; i <= nx; ++i)
                face_interior(o, 2, i, j - 1, i, j, o->accel_h + 2 * HI(o, i, j), o->d.dx, HF(i, j));
    }
#undef VF
#undef HF
}

/* One velocity-axis difference of cal_fsource1, 2d/src/gravity.f90:108-187,
 * at node i (0-based) of a line of n nodes (stride st) of f, step du: the
 * one-sided second-order forms at both ends, the central one next to them
 * and the fourth-order central one inside, in the reference's expression
 * order.  As the reference writes them: the interior form takes its three
 * inner terms from g (the u-axis difference of b reads them from h), and
 * the last node's form has the sign of a forward difference when last_fwd
 * (the v axis). */
static double fd_axis(const double *f, const double *g, size_t st, int n, int i, double du,
                      int last_fwd)
{
#define F(q) f[(size_t)(q) * st]
#define G(q) g[(size_t)(q) * st]
    if (i == 0) return (-3.0 * F(0) + 4.0 * F(1) - F(2)) / du / 2.0;
    if (i == n - 1)
        return last_fwd ? (-3.0 * F(n - 1) + 4.0 * F(n - 2) - F(n - 3)) / du / 2.0
                        : (3.0 * F(n - 1) - 4.0 * F(n - 2) + F(n - 3)) / du / 2.0;
    if (i == 1 || i == n - 2) return (F(i + 1) - F(i - 1)) / du / 2.0;
    return (-F(i + 2) + 8.0 * G(i + 1) - 8.0 * G(i - 1) + G(i - 2)) / du / 12.0;
#undef F
#undef G
}

/* force_type 'finite_difference' (2d/src/solver.f90:353-356): hsource =
 * duh dt fx + dvh dt fy, bsource likewise, with du = dv = lengthU/unum */
static void force_fd(const hugks_desc *d, int nv, const double *h, const double *b, double dt,
                     double fx, double fy, double *hs, double *bs)
{
    const double du = d->fd_span / nv, dv = d->fd_span / nv;
    for (int j = 0; j < nv; ++j)
        for (int i = 0; i < nv; ++i) {
            const double duh = fd_axis(h + (size_t)nv * j, h + (size_t)nv * j, 1, nv, i, du, 0);
            const double dub = fd_axis(b + (size_t)nv * j, h + (size_t)nv * j, 1, nv, i, du, 0);
            const double dvh = fd_axis(h + i, h + i, (size_t)nv, nv, j, dv, 1);
            const double dvb = fd_axis(b + i, b + i, (size_t)nv, nv, j, dv, 1);
            hs[i + (size_t)nv * j] = duh * dt * fx + dvh * dt * fy;
            bs[i + (size_t)nv * j] = dub * dt * fx + dvb * dt * fy;
        }
}

/* update, 2d/src/solver.f90:245-395 */
static void update(oracle2d *o)
{
    const hugks_desc *d = &o->d;
    size_t nn = o->nn;
    double area = d->dx * d->dy;
    double sum_res[4] = {0, 0, 0, 0}, mass_scale = 0.0;
    double *Ho = o->t0, *Bo = o->t1, *H = o->t2, *B = o->t3, *Hp = o->t4, *Bp = o->t5;
    double *fth = o->t6, *ftb = o->t7;
    for (int j = 1; j <= o->ny; ++j)
        for (int i = 1; i <= o->nx; ++i) {
            size_t c = (size_t)(i - 1) + (size_t)o->nx * (j - 1);
            double fx = o->accel_cell[2 * c], fy = o->accel_cell[2 * c + 1];
            double *w = CW(o, i, j), *h = CH(o, i, j), *b = CB(o, i, j);
            double w_old[4], prim_old[4], prim[4], qf[2];
            memcpy(w_old, w, sizeof(w_old));
            get_primary(d, w_old, prim_old);
            reduced_maxwell(d, nn, Ho, Bo, o->uspace, o->vspace, prim_old);
            double tau_old = get_tau(d, prim_old);
            const double *fl = o->vflux + 4 * VI(o, i, j), *fr = o->vflux + 4 * VI(o, i + 1, j);
            const double *fd = o->hflux + 4 * HI(o, i, j), *fu = o->hflux + 4 * HI(o, i, j + 1);
            for (int q = 0; q < 4; ++q) w[q] = w[q] + (fl[q] - fr[q] + fd[q] - fu[q]) / area;
            get_primary(d, w, prim);
            prim[1] = prim[1] + fx * o->dt;
            prim[2] = prim[2] + fy * o->dt;
            get_conserved(d, prim, w);
            reduced_maxwell(d, nn, H, B, o->uspace, o->vspace, prim);
            double tau = get_tau(d, prim);
            for (int q = 0; q < 4; ++q) sum_res[q] += (w_old[q] - w[q]) * (w_old[q] - w[q]);
            mass_scale += fabs(w[0]);
            get_heat_flux((int)nn, o->weight, h, b, o->uspace, o->vspace, prim_old, qf);
            shakhov_part(d, (int)nn, Ho, Bo, o->uspace, o->vspace, qf, prim_old, Hp, Bp);
            for (size_t k = 0; k < nn; ++k) { Ho[k] += Hp[k]; Bo[k] += Bp[k]; }
            shakhov_part(d, (int)nn, H, B, o->uspace, o->vspace, qf, prim, Hp, Bp);
            for (size_t k = 0; k < nn; ++k) { H[k] += Hp[k]; B[k] += Bp[k]; }
            /* fth, ftb: the Hermite force terms ft1, ft2 (hsource = ft1 dt),
               or the finite-difference hsource, bsource themselves */
            const int use_fd = d->force_variant == HUGKS_FORCE_FINITE_DIFFERENCE;
            if (use_fd)
                force_fd(d, o->nv, h, b, o->dt, fx, fy, fth, ftb);
            else if (d->force_variant == HUGKS_FORCE_HERMITE_LEGACY)
                force_legacy(d, (int)nn, o->weight, o->uspace, o->vspace, d->nm, h, b, prim_old,
                             fx, fy, fth, ftb);
            else
                force_symmetric(d, (int)nn, o->weight, o->uspace, o->vspace, d->nm,
                                d->force_variant == HUGKS_FORCE_HERMITE_SYMMETRIC_B, h, b, prim_old,
                                fx, fy, fth, ftb, NULL, NULL);
            const double *hl = o->vfh + nn * VI(o, i, j), *hr = o->vfh + nn * VI(o, i + 1, j);
            const double *hd = o->hfh + nn * HI(o, i, j), *hu = o->hfh + nn * HI(o, i, j + 1);
            const double *bl = o->vfb + nn * VI(o, i, j), *br = o->vfb + nn * VI(o, i + 1, j);
            const double *bd = o->hfb + nn * HI(o, i, j), *bu = o->hfb + nn * HI(o, i, j + 1);
            double den = 1.0 + 0.5 * o->dt / tau;
            for (size_t k = 0; k < nn; ++k) {
                double hs = use_fd ? fth[k] : fth[k] * o->dt;
                double bs = use_fd ? ftb[k] : ftb[k] * o->dt;
                h[k] = (h[k] + (hl[k] - hr[k] + hd[k] - hu[k]) / area +
                        0.5 * o->dt * (H[k] / tau + (Ho[k] - h[k]) / tau_old) - hs) / den;
                b[k] = (b[k] + (bl[k] - br[k] + bd[k] - bu[k]) / area +
                        0.5 * o->dt * (B[k] / tau + (Bo[k] - b[k]) / tau_old) - bs) / den;
            }
        }
    double ncell = (double)o->nx * (double)o->ny;
    for (int q = 0; q < 4; ++q) o->res[q] = sqrt(ncell * sum_res[q]) / (mass_scale + d->smv);
}

/* ------------------------------------------------------------------ */
/* public API                                                           */
/* ------------------------------------------------------------------ */
oracle2d *oracle2d_create(const hugks_desc *d)
{
    if (!d || d->dim != 2 || d->nx < 2 || d->ny < 2 || d->nv < 2) return NULL;
    oracle2d *o = calloc(1, sizeof(*o));
    o->d = *d;
    o->nx = d->nx; o->ny = d->ny; o->nv = d->nv; o->nn = d->nv * d->nv;
    size_t nn = o->nn, ncg = (size_t)(o->nx + 2) * (o->ny + 2);
    size_t nvf = (size_t)(o->nx + 1) * o->ny, nhf = (size_t)o->nx * (o->ny + 1);
    size_t ncell = (size_t)o->nx * o->ny;
    o->uspace = malloc(nn * sizeof(double));
    o->vspace = malloc(nn * sizeof(double));
    o->weight = malloc(nn * sizeof(double));
    for (int jv = 0; jv < o->nv; ++jv)
        for (int iu = 0; iu < o->nv; ++iu) {
            o->uspace[iu + o->nv * jv] = d->u[iu];
            o->vspace[iu + o->nv * jv] = d->v[jv];
        }
    memcpy(o->weight, d->weight, nn * sizeof(double));
    o->w = calloc(4 * ncg, sizeof(double));
    o->h = calloc(nn * ncg, sizeof(double));
    o->b = calloc(nn * ncg, sizeof(double));
    o->sh = calloc(2 * nn * ncg, sizeof(double));
    o->sb = calloc(2 * nn * ncg, sizeof(double));
    o->vflux = calloc(4 * nvf, sizeof(double));
    o->vfh = calloc(nn * nvf, sizeof(double));
    o->vfb = calloc(nn * nvf, sizeof(double));
    o->hflux = calloc(4 * nhf, sizeof(double));
    o->hfh = calloc(nn * nhf, sizeof(double));
    o->hfb = calloc(nn * nhf, sizeof(double));
    o->accel_cell = malloc(2 * ncell * sizeof(double));
    o->accel_v = malloc(2 * nvf * sizeof(double));
    o->accel_h = malloc(2 * nhf * sizeof(double));
    memcpy(o->accel_cell, d->accel_cell, 2 * ncell * sizeof(double));
    memcpy(o->accel_v, d->accel_vface, 2 * nvf * sizeof(double));
    memcpy(o->accel_h, d->accel_hface, 2 * nhf * sizeof(double));
    double *scr = malloc(16 * nn * sizeof(double));
    o->t0 = scr;
    o->t1 = malloc(nn * sizeof(double)); o->t2 = malloc(nn * sizeof(double));
    o->t3 = malloc(nn * sizeof(double)); o->t4 = malloc(nn * sizeof(double));
    o->t5 = malloc(nn * sizeof(double)); o->t6 = malloc(nn * sizeof(double));
    o->t7 = malloc(nn * sizeof(double));
    /* update() uses t0 as H_old while face routines use it as a 13*nn
     * scratch; they never overlap in time */
    for (int j = 1; j <= o->ny; ++j)
        for (int i = 1; i <= o->nx; ++i) {
            /* init_distribution_function, 2d/src/initialization.f90:227-235;
             * W from init_flowField (get_conserved of prim) */
            const double *p = d->prim0 + 4 * ((size_t)(i - 1) + (size_t)o->nx * (j - 1));
            get_conserved(d, p, CW(o, i, j));
            reduced_maxwell(d, (int)nn, CH(o, i, j), CB(o, i, j), o->uspace, o->vspace, p);
        }
    o->simtime = 0.0;
    o->iter = 1;
    return o;
}

void oracle2d_destroy(oracle2d *o)
{
    if (!o) return;
    free(o->uspace); free(o->vspace); free(o->weight); free(o->w); free(o->h); free(o->b);
    free(o->sh); free(o->sb); free(o->vflux); free(o->vfh); free(o->vfb); free(o->hflux);
    free(o->hfh); free(o->hfb); free(o->accel_cell); free(o->accel_v); free(o->accel_h);
    free(o->t0); free(o->t1); free(o->t2); free(o->t3); free(o->t4); free(o->t5);
    free(o->t6); free(o->t7);
    free(o);
}

int oracle2d_step(oracle2d *o, hugks_step_info *info)
{
    timestep(o);
    interpolation(o);
    evolution(o);
    update(o);
    o->simtime = o->simtime + o->dt;
    o->iter += 1;
    if (info) {
        memset(info, 0, sizeof(*info));
        info->steps = 1;
        info->iter = o->iter;
        info->simtime = o->simtime;
        info->dt = o->dt;
        memcpy(info->res, o->res, sizeof(o->res));
        if (o->simtime >= o->d.maxtime) info->stop_reason = HUGKS_STOP_MAXTIME;
        else if (o->d.stop_on_residual && o->res[0] < o->d.eps_res && o->res[1] < o->d.eps_res &&
                 o->res[2] < o->d.eps_res && o->res[3] < o->d.eps_res)
            info->stop_reason = HUGKS_STOP_RESIDUAL;
    }
    return 0;
}

void oracle2d_get(const oracle2d *o, double *w, double *h, double *b)
{
    size_t nn = o->nn;
    for (int j = 1; j <= o->ny; ++j)
        for (int i = 1; i <= o->nx; ++i) {
            size_t c = (size_t)(i - 1) + (size_t)o->nx * (j - 1);
            size_t g = CI(o, i, j);
            if (w) memcpy(w + 4 * c, o->w + 4 * g, 4 * sizeof(double));
            if (h) memcpy(h + nn * c, o->h + nn * g, nn * sizeof(double));
            if (b) memcpy(b + nn * c, o->b + nn * g, nn * sizeof(double));
        }
}

void oracle2d_set(oracle2d *o, const double *w, const double *h, const double *b)
{
    size_t nn = o->nn;
    for (int j = 1; j <= o->ny; ++j)
        for (int i = 1; i <= o->nx; ++i) {
            size_t c = (size_t)(i - 1) + (size_t)o->nx * (j - 1);
            size_t g = CI(o, i, j);
            if (w) memcpy(o->w + 4 * g, w + 4 * c, 4 * sizeof(double));
            if (h) memcpy(o->h + nn * g, h + nn * c, nn * sizeof(double));
            if (b) memcpy(o->b + nn * g, b + nn * c, nn * sizeof(double));
        }
}

void oracle2d_time(const oracle2d *o, double *simtime, int32_t *iter, double *dt)
{
    if (simtime) *simtime = o->simtime;
    if (iter) *iter = o->iter;
    if (dt) *dt = o->dt;
}

void oracle2d_heat_flux(const oracle2d *o, double *q)
{
    /* write_macro, 2d/src/output2d.f90:156-161 */
    for (int j = 1; j <= o->ny; ++j)
        for (int i = 1; i <= o->nx; ++i) {
            size_t c = (size_t)(i - 1) + (size_t)o->nx * (j - 1);
            size_t g = CI(o, i, j);
            double prim[4];
            get_primary(&o->d, o->w + 4 * g, prim);
            get_heat_flux(o->nn, o->weight, o->h + o->nn * g, o->b + o->nn * g, o->uspace,
                          o->vspace, prim, q + 2 * c);
        }
}

static void grids(const hugks_desc *d, double **us, double **vs)
{
    int nv = d->nv, nn = nv * nv;
    *us = malloc(nn * sizeof(double));
    *vs = malloc(nn * sizeof(double));
    for (int jv = 0; jv < nv; ++jv)
        for (int iu = 0; iu < nv; ++iu) {
            (*us)[iu + nv * jv] = d->u[iu];
            (*vs)[iu + nv * jv] = d->v[jv];
        }
}

void oracle2d_flux_interior(const hugks_desc *d, int dir, double dt, double fn, double ft,
                            double face_len, const double *hL, const double *bL,
                            const double *shL, const double *sbL, const double *wL,
                            const double *lenL, const double *hR, const double *bR,
                            const double *shR, const double *sbR, const double *wR,
                            const double *lenR, double *flux, double *fh, double *fb)
{
    int nn = d->nv * d->nv;
    double *us, *vs, *scr = malloc(16 * (size_t)nn * sizeof(double));
    grids(d, &us, &vs);
    flux_interior(d, nn, d->weight, us, vs, dir, dt, fn, ft, face_len, hL, bL, shL, sbL, wL, lenL,
                  hR, bR, shR, sbR, wR, lenR, flux, fh, fb, scr);
    free(us); free(vs); free(scr);
}

void oracle2d_hermite_force(const hugks_desc *d, int variant, int order, const double *h,
                            const double *b, const double *prim, double gx, double gy,
                            double *force_h, double *force_b, double *err_h, double *err_b)
{
    int nn = d->nv * d->nv;
    double *us, *vs;
    grids(d, &us, &vs);
    if (variant == HUGKS_FORCE_HERMITE_LEGACY) {
        force_legacy(d, nn, d->weight, us, vs, order, h, b, prim, gx, gy, force_h, force_b);
        if (err_h) *err_h = 0;
        if (err_b) *err_b = 0;
    } else {
        force_symmetric(d, nn, d->weight, us, vs, order, variant == HUGKS_FORCE_HERMITE_SYMMETRIC_B,
                        h, b, prim, gx, gy, force_h, force_b, err_h, err_b);
    }
    free(us); free(vs);
}

void oracle2d_flux_wall(const hugks_desc *d, int kind, int dir, int rot, double tw, double dt,
                        double face_len, const double *h, const double *b, const double *sh,
                        const double *sb, double len, double *flux, double *fh, double *fb)
{
    int nn = d->nv * d->nv;
    double *us, *vs, *scr = malloc(16 * (size_t)nn * sizeof(double));
    grids(d, &us, &vs);
    flux_wall(d, nn, d->nv, d->weight, us, vs, kind, dir, rot, tw, dt, face_len, h, b, sh, sb, len,
              flux, fh, fb, scr);
    free(us); free(vs); free(scr);
}
