/*
 * ugks_oracle1d.c -- FP64 CPU restatement of the reference 1D Hermite-UGKS
 * step (1d/src Fortran sources).  TEST INFRASTRUCTURE ONLY (see ugks_oracle.h).
 * Each routine cites the reference lines it follows.  Boundary kinds come
 * from desc.bc[0..1] (the reference derives them from case_id,
 * 1d/src/initialization.f90:50-60).
 */
#include "ugks_oracle.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MNUM 6

struct oracle1d {
    hugks_desc d;
    int nx, nv;
    double *u, *weight;
    double *w;            /* [3*nx] */
    double *h, *b;        /* [nv*nx] */
    double *sh, *sb;      /* [nv*nx] */
    double *flux;         /* [3*(nx+1)], face i = 1..nx+1 stored at i-1 */
    double *fh, *fb;      /* [nv*(nx+1)] */
    double dt, simtime;
    int iter;
    double res[3], steady_res[3];
    double *t[12];
};

/* reduced_maxwell, 1d/src/gastheory.f90:20-27 */
static void maxwell1(const hugks_desc *d, int nv, const double *u, const double *prim,
                     double *H, double *B)
{
    for (int k = 0; k < nv; ++k) {
        double c = u[k] - prim[1];
        H[k] = prim[0] * pow(prim[2] / d->pi, 0.5) * exp(-prim[2] * c * c);
        B[k] = H[k] * d->innerk / (2.0 * prim[2]);
    }
}

/* shakhov_part, 1d/src/gastheory.f90:29-39 */
static void shakhov1(const hugks_desc *d, int nv, const double *u, const double *H,
                     const double *B, double qf, const double *prim, double *Hp, double *Bp)
{
    double c = d->shakhov_coef * (1.0 - d->prandtl) * prim[2] * prim[2] / prim[0];
    for (int k = 0; k < nv; ++k) {
        double cu = u[k] - prim[1];
        double c2 = 2.0 * prim[2] * cu * cu;
        Hp[k] = c * cu * qf * (c2 + d->innerk - 5.0) * H[k];
        Bp[k] = c * cu * qf * (c2 + d->innerk - 3.0) * B[k];
    }
}

/* get_conserved, 1d/src/gastheory.f90:41-49 */
static void conserved1(const hugks_desc *d, const double *prim, double *w)
{
    w[0] = prim[0];
    w[1] = prim[0] * prim[1];
    w[2] = 0.5 * prim[0] / prim[2] / (d->gamma - 1.0) + 0.5 * prim[0] * prim[1] * prim[1];
}

/* get_primary, 1d/src/gastheory.f90:51-58 */
static void primary1(const hugks_desc *d, const double *w, double *prim)
{
    prim[0] = w[0];
    prim[1] = w[1] / w[0];
    prim[2] = 0.5 * w[0] / (d->gamma - 1.0) / (w[2] - 0.5 * w[1] * w[1] / w[0]);
}

/* get_tau, 1d/src/gastheory.f90:77-83 */
static double tau1(const hugks_desc *d, const double *prim)
{
    return d->muref * 2.0 * pow(prim[2], 1.0 - d->omega) / prim[0];
}

/* get_heat_flux, 1d/src/gastheory.f90:98-104 */
static double heat1(int nv, const double *weight, const double *u, const double *h,
                    const double *b, const double *prim)
{
    double s1 = 0, s2 = 0;
    for (int k = 0; k < nv; ++k) {
        double c = u[k] - prim[1];
        s1 += weight[k] * c * (c * c) * h[k];
        s2 += weight[k] * c * b[k];
    }
    return 0.5 * (s1 + s2);
}

/* micro_slope, 1d/src/flux.f90:362-370 */
static void slope1(const hugks_desc *d, const double *prim, const double *sw, double *a)
{
    double K = d->innerk;
    a[2] = 4.0 * prim[2] * prim[2] / (K + 1.0) / prim[0] *
           (2.0 * sw[2] - 2.0 * prim[1] * sw[1] + sw[0] * (prim[1] * prim[1] - 0.5 * (K + 1.0) / prim[2]));
    a[1] = 2.0 * prim[2] / prim[0] * (sw[1] - prim[1] * sw[0]) - prim[1] * a[2];
    a[0] = sw[0] / prim[0] - prim[1] * a[1] - 0.5 * (prim[1] * prim[1] + 0.5 * (K + 1.0) / prim[2]) * a[2];
}

/* calc_moment_u, 1d/src/flux.f90:372-395 */
static void moment1(const hugks_desc *d, const double *prim, double *Mu, double *Mxi,
                    double *MuL, double *MuR)
{
    double U = prim[1], lam = prim[2], K = d->innerk;
    MuL[0] = 0.5 * erfc(-sqrt(lam) * U);
    MuL[1] = U * MuL[0] + 0.5 * exp(-lam * U * U) / sqrt(d->pi * lam);
    MuR[0] = 0.5 * erfc(sqrt(lam) * U);
    MuR[1] = U * MuR[0] - 0.5 * exp(-lam * U * U) / sqrt(d->pi * lam);
    for (int i = 2; i <= MNUM; ++i) {
        MuL[i] = U * MuL[i - 1] + 0.5 * (i - 1) * MuL[i - 2] / lam;
        MuR[i] = U * MuR[i - 1] + 0.5 * (i - 1) * MuR[i - 2] / lam;
    }
    for (int i = 0; i <= MNUM; ++i) Mu[i] = MuL[i] + MuR[i];
    Mxi[0] = 1.0;
    Mxi[1] = 0.5 * K / lam;
    Mxi[2] = (K * K + 2.0 * K) / (4.0 * lam * lam);
}

/* moment_uv, 1d/src/flux.f90:397-405 */
static void uv1(const double *Mu, const double *Mxi, int alpha, int delta, double *m)
{
    m[0] = Mu[alpha] * Mxi[delta / 2];
    m[1] = Mu[alpha + 1] * Mxi[delta / 2];
    m[2] = 0.5 * (Mu[alpha + 2] * Mxi[delta / 2] + Mu[alpha] * Mxi[(delta + 2) / 2]);
}

/* moment_au, 1d/src/flux.f90:407-417 */
static void au1(const double *a, const double *Mu, const double *Mxi, int alpha, double *m)
{
    double t1[3], t2[3], t3[3], t4[3];
    uv1(Mu, Mxi, alpha + 0, 0, t1);
    uv1(Mu, Mxi, alpha + 1, 0, t2);
    uv1(Mu, Mxi, alpha + 2, 0, t3);
    uv1(Mu, Mxi, alpha + 0, 2, t4);
    for (int c = 0; c < 3; ++c)
        m[c] = a[0] * t1[c] + a[1] * t2[c] + 0.5 * a[2] * t3[c] + 0.5 * a[2] * t4[c];
}

/* calc_flux, 1d/src/flux.f90:13-186 (the flux is time-integrated; no
 * face length) */
static void flux1(oracle1d *o, int iL, int iR, double *flux, double *fh, double *fb)
{
    const hugks_desc *d = &o->d;
    int nv = o->nv;
    const double *u = o->u, *wt = o->weight;
    double *h = o->t[0], *b = o->t[1], *sh = o->t[2], *sb = o->t[3];
    double *H0 = o->t[4], *B0 = o->t[5], *Hp = o->t[6], *Bp = o->t[7];
    int *delta = (int *)o->t[8];
    double phi = d->phi, dt = o->dt;
    double lenL = d->dx, lenR = d->dx;
    const double *hL = o->h + (size_t)nv * iL, *hR = o->h + (size_t)nv * iR;
    const double *bL = o->b + (size_t)nv * iL, *bR = o->b + (size_t)nv * iR;
    const double *shL = o->sh + (size_t)nv * iL, *shR = o->sh + (size_t)nv * iR;
    const double *sbL = o->sb + (size_t)nv * iL, *sbR = o->sb + (size_t)nv * iR;
    for (int k = 0; k < nv; ++k) {
        double vf = u[k] - 0.5 * phi * dt;                       /* 56-57 */
        delta[k] = (int)((copysign(1.0, vf) + 1.0) / 2.0);       /* 60 */
        double dl = delta[k], dr = 1 - delta[k];
        h[k] = (hL[k] + 0.5 * lenL * shL[k]) * dl + (hR[k] - 0.5 * lenR * shR[k]) * dr;
        b[k] = (bL[k] + 0.5 * lenL * sbL[k]) * dl + (bR[k] - 0.5 * lenR * sbR[k]) * dr;
        sh[k] = shL[k] * dl + shR[k] * dr;
        sb[k] = sbL[k] * dl + sbR[k] * dr;
    }
    double w[3] = {0, 0, 0}, e1 = 0, e2 = 0;                     /* 76-78 */
    for (int k = 0; k < nv; ++k) {
        w[0] += wt[k] * h[k];
        w[1] += wt[k] * u[k] * h[k];
        e1 += wt[k] * u[k] * u[k] * h[k];
        e2 += wt[k] * b[k];
    }
    w[2] = 0.5 * (e1 + e2);
    double prim[3], sw[3], aL[3], aR[3], aT[3];
    primary1(d, w, prim);
    double qf = heat1(nv, wt, u, h, b, prim);
    const double *wLc = o->w + 3 * iL, *wRc = o->w + 3 * iR;
    for (int c = 0; c < 3; ++c) sw[c] = (w[c] - wLc[c]) / (0.5 * lenL);   /* 89-93 */
    slope1(d, prim, sw, aL);
    for (int c = 0; c < 3; ++c) sw[c] = (wRc[c] - w[c]) / (0.5 * lenR);
    slope1(d, prim, sw, aR);
    double Mu[MNUM + 1], MuL[MNUM + 1], MuR[MNUM + 1], Mxi[3];
    double MauL[3], MauR[3], MauT[3], Mau0[3];
    moment1(d, prim, Mu, Mxi, MuL, MuR);                          /* 99-105 */
    au1(aL, MuL, Mxi, 1, MauL);
    au1(aR, MuR, Mxi, 1, MauR);
    for (int c = 0; c < 3; ++c) sw[c] = -prim[0] * (MauL[c] + MauR[c]);
    slope1(d, prim, sw, aT);
    double tau = tau1(d, prim), Mt[7], ex = exp(-dt / tau);      /* 110-117 */
    Mt[4] = tau * (1.0 - ex);
    Mt[5] = -tau * dt * ex + tau * Mt[4];
    Mt[1] = dt - Mt[4];
    Mt[2] = -tau * Mt[1] + Mt[5];
    Mt[3] = dt * dt / 2.0 - tau * Mt[1];
    Mt[6] = tau * tau * tau - 0.5 * ex * tau * (dt * dt + 2.0 * dt * tau + 2.0 * tau * tau);
    uv1(Mu, Mxi, 1, 0, Mau0);                                     /* 122-127 */
    au1(aL, MuL, Mxi, 2, MauL);
    au1(aR, MuR, Mxi, 2, MauR);
    au1(aT, Mu, Mxi, 1, MauT);
    for (int c = 0; c < 3; ++c)
        flux[c] = Mt[1] * prim[0] * Mau0[c] + Mt[2] * prim[0] * (MauL[c] + MauR[c]) +
                  Mt[3] * prim[0] * MauT[c];
    au1(aL, MuL, Mxi, 1, MauL);                                   /* 132-137 */
    au1(aR, MuR, Mxi, 1, MauR);
    for (int c = 0; c < 3; ++c)
        flux[c] = flux[c] - tau * Mt[2] * prim[0] * phi * (MauL[c] + MauR[c]) -
                  Mt[6] * prim[0] * phi * (MauL[c] + MauR[c]);
    maxwell1(d, nv, u, prim, H0, B0);                             /* 143-146 */
    shakhov1(d, nv, u, H0, B0, qf, prim, Hp, Bp);
    double S[16];
    memset(S, 0, sizeof(S));
    for (int k = 0; k < nv; ++k) {
        double wk = wt[k], uk = u[k], u2 = uk * uk;
        S[0] += wk * uk * Hp[k];
        S[1] += wk * uk * h[k];
        S[2] += wk * u2 * sh[k];
        S[3] += wk * u2 * Hp[k];
        S[4] += wk * u2 * h[k];
        S[5] += wk * uk * u2 * sh[k];
        S[6] += wk * uk * u2 * Hp[k];
        S[7] += wk * uk * Bp[k];
        S[8] += wk * uk * u2 * h[k];
        S[9] += wk * uk * b[k];
        S[10] += wk * u2 * u2 * sh[k];
        S[11] += wk * u2 * sb[k];
        S[12] += wk * uk * sh[k];
        S[13] += wk * u2 * sh[k];
        S[14] += wk * u2 * uk * sh[k];
        S[15] += wk * uk * sb[k];
    }
    flux[0] = flux[0] + Mt[1] * S[0] + Mt[4] * S[1] - Mt[5] * S[2];   /* 149-154 */
    flux[1] = flux[1] + Mt[1] * S[3] + Mt[4] * S[4] - Mt[5] * S[5];
    flux[2] = flux[2] + Mt[1] * 0.5 * (S[6] + S[7]) + Mt[4] * 0.5 * (S[8] + S[9]) -
              Mt[5] * 0.5 * (S[10] + S[11]);
    flux[0] = flux[0] + phi * Mt[6] * S[12];                           /* 157-159 */
    flux[1] = flux[1] + phi * Mt[6] * S[13];
    flux[2] = flux[2] + phi * Mt[6] * 0.5 * (S[14] + S[15]);
    for (int k = 0; k < nv; ++k) {                                     /* 164-184 */
        double uk = u[k], u2 = uk * uk, H = H0[k], Bm = B0[k];
        double dl = delta[k], dr = 1 - delta[k];
        double gLh = aL[0] * H + aL[1] * uk * H + 0.5 * aL[2] * (u2 * H + Bm);
        double gRh = aR[0] * H + aR[1] * uk * H + 0.5 * aR[2] * (u2 * H + Bm);
        double gTh = aT[0] * H + aT[1] * uk * H + 0.5 * aT[2] * (u2 * H + Bm);
        double gLb = aL[0] * Bm + aL[1] * uk * Bm + 0.5 * aL[2] * (u2 * Bm + Mxi[2] * H);
        double gRb = aR[0] * Bm + aR[1] * uk * Bm + 0.5 * aR[2] * (u2 * Bm + Mxi[2] * H);
        double gTb = aT[0] * Bm + aT[1] * uk * Bm + 0.5 * aT[2] * (u2 * Bm + Mxi[2] * H);
        double xh = Mt[1] * uk * (H + Hp[k]) + Mt[2] * u2 * gLh * dl + Mt[2] * u2 * gRh * dr +
                    Mt[3] * uk * gTh + Mt[4] * uk * h[k] - Mt[5] * u2 * sh[k];
        double xb = Mt[1] * uk * (Bm + Bp[k]) + Mt[2] * u2 * gLb * dl + Mt[2] * u2 * gRb * dr +
                    Mt[3] * uk * gTb + Mt[4] * uk * b[k] - Mt[5] * u2 * sb[k];
        fh[k] = xh - tau * Mt[2] * phi * uk * gLh * dl - tau * Mt[2] * phi * uk * gRh * dr +
                Mt[6] * phi * uk * sh[k];
        fb[k] = xb - tau * Mt[2] * phi * uk * gLb * dl - tau * Mt[2] * phi * uk * gRb * dr +
                Mt[6] * phi * uk * sb[k];
    }
}

/* calc_flux_specular 1d/src/flux.f90:188-239 and calc_flux_boundary
 * 1d/src/flux.f90:299-360 */
static void wall1(oracle1d *o, int kind, int rot, double tw, int ic, double *flux, double *fh,
                  double *fb)
{
    const hugks_desc *d = &o->d;
    int nv = o->nv;
    const double *u = o->u, *wt = o->weight;
    double *h = o->t[0], *b = o->t[1], *H0 = o->t[4], *B0 = o->t[5];
    double *th = o->t[6], *tb = o->t[7];
    int *delta = (int *)o->t[8];
    const double *hc = o->h + (size_t)nv * ic, *bc = o->b + (size_t)nv * ic;
    const double *shc = o->sh + (size_t)nv * ic, *sbc = o->sb + (size_t)nv * ic;
    for (int k = 0; k < nv; ++k) {
        delta[k] = (int)((copysign(1.0, u[k]) * rot + 1.0) / 2.0);
        h[k] = hc[k] - rot * 0.5 * d->dx * shc[k];
        b[k] = bc[k] - rot * 0.5 * d->dx * sbc[k];
    }
    if (kind == HUGKS_BC_DIFFUSE) {
        double prim[3] = {1.0, 0.0, 1.0 / tw};
        double R1 = 0, R2 = 0;
        for (int k = 0; k < nv; ++k) R1 += wt[k] * u[k] * h[k] * (double)(1 - delta[k]);
        maxwell1(d, nv, u, prim, H0, B0);
        for (int k = 0; k < nv; ++k) R2 += wt[k] * u[k] * H0[k] * (double)delta[k];
        prim[0] = -R1 / R2;
        maxwell1(d, nv, u, prim, H0, B0);
        for (int k = 0; k < nv; ++k) {
            h[k] = H0[k] * delta[k] + h[k] * (1 - delta[k]);
            b[k] = B0[k] * delta[k] + b[k] * (1 - delta[k]);
        }
    } else {
        for (int k = 0; k < nv; ++k) { th[k] = h[nv - 1 - k]; tb[k] = b[nv - 1 - k]; }
        for (int k = 0; k < nv; ++k) {
            h[k] = th[k] * delta[k] + h[k] * (1 - delta[k]);
            b[k] = tb[k] * delta[k] + b[k] * (1 - delta[k]);
        }
    }
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = 0; k < nv; ++k) {
        s0 += wt[k] * u[k] * h[k];
        s1 += wt[k] * u[k] * u[k] * h[k];
        s2 += wt[k] * u[k] * ((u[k] * u[k]) * h[k] + b[k]);
        fh[k] = o->dt * (u[k] * h[k]);
        fb[k] = o->dt * (u[k] * b[k]);
    }
    flux[0] = o->dt * s0;
    flux[1] = o->dt * s1;
    flux[2] = o->dt * (0.5 * s2);
}

/* cal_fv1, 1d/src/gravity.f90:127-154: velocity differences on the step
 * du = lengthU/unum, one-sided second order at both ends, central next to
 * them, fourth-order central inside, in the reference's expression order
 * (for nv = 3 the second node's form is written twice, the same value) */
static void calfv1(const hugks_desc *d, int nv, const double *h, const double *b, double *dvh,
                   double *dvb)
{
    const double du = d->fd_span / nv;
    const double *f[2] = {h, b};
    double *o[2] = {dvh, dvb};
    for (int a = 0; a < 2; ++a) {
        const double *x = f[a];
        double *y = o[a];
        int n = nv;
        y[0] = (-3.0 * x[0] + 4.0 * x[1] - x[2]) / 2.0 / du;
        y[1] = (x[2] - x[0]) / 2.0 / du;
        y[n - 2] = (x[n - 1] - x[n - 3]) / 2.0 / du;
        y[n - 1] = (3.0 * x[n - 1] - 4.0 * x[n - 2] + x[n - 3]) / 2.0 / du;
        for (int i = 2; i <= n - 3; ++i)
            y[i] = (-x[i + 2] + 8.0 * x[i + 1] - 8.0 * x[i - 1] + x[i - 2]) / 12.0 / du;
    }
}

/* cal_fv, 1d/src/gravity.f90:18-123 */
static void calfv(const hugks_desc *d, int nv, const double *u, const double *weight, int N,
                  const double *h, const double *b, const double *prim, double *dvh, double *dvb)
{
    double *factor = malloc(sizeof(double) * (N + 1));
    double *df = malloc(sizeof(double) * (N + 1)), *dg = malloc(sizeof(double) * (N + 1));
    double *hp = malloc(sizeof(double) * (size_t)(N + 2) * nv);
    double *v = malloc(sizeof(double) * nv), *w = malloc(sizeof(double) * nv);
    factor[0] = 1.0;
    for (int i = 1; i <= N; ++i) {
        factor[i] = 1.0;
        for (int j = 1; j <= i; ++j) factor[i] = factor[i] * j;
    }
    double t = 1.0 / prim[2];
    for (int k = 0; k < nv; ++k) {
        v[k] = (u[k] - prim[1]) / sqrt(t) * sqrt(2.0);
        w[k] = exp(-v[k] * v[k] / 2.0) / sqrt(2.0 * d->pi);
        hp[k] = 1.0;
        hp[nv + k] = v[k];
        if (N > 0) hp[2 * nv + k] = v[k] * v[k] - 1.0;
    }
    if (N > 1)
        for (int j = 3; j <= N + 1; ++j)
            for (int k = 0; k < nv; ++k)
                hp[j * nv + k] = v[k] * hp[(j - 1) * nv + k] - (j - 1.0) * hp[(j - 2) * nv + k];
    for (int j = 0; j <= N; ++j) {
        double s1 = 0, s2 = 0;
        for (int k = 0; k < nv; ++k) {
            s1 += h[k] * weight[k] * hp[j * nv + k];
            s2 += b[k] * weight[k] * hp[j * nv + k];
        }
        df[j] = s1 * sqrt(2.0 / t);
        dg[j] = s2 * sqrt(2.0 / t);
    }
    for (int k = 0; k < nv; ++k) {
        double F1 = 0.0, F2 = 0.0;
        for (int j = 0; j <= N; ++j) {
            F1 = F1 + df[j] * hp[(j + 1) * nv + k] / factor[j];
            F2 = F2 + dg[j] * hp[(j + 1) * nv + k] / factor[j];
        }
        dvh[k] = -sqrt(2.0 / t) * w[k] * F1;
        dvb[k] = -sqrt(2.0 / t) * w[k] * F2;
    }
    free(factor); free(df); free(dg); free(hp); free(v); free(w);
}

oracle1d *oracle1d_create(const hugks_desc *d)
{
    if (!d || d->dim != 1 || d->nx < 2 || d->nv < 2) return NULL;
    oracle1d *o = calloc(1, sizeof(*o));
    o->d = *d;
    o->nx = d->nx; o->nv = d->nv;
    int nx = o->nx, nv = o->nv;
    o->u = malloc(nv * sizeof(double));
    o->weight = malloc(nv * sizeof(double));
    memcpy(o->u, d->u, nv * sizeof(double));
    memcpy(o->weight, d->weight, nv * sizeof(double));
    o->w = calloc(3 * nx, sizeof(double));
    o->h = calloc((size_t)nv * nx, sizeof(double));
    o->b = calloc((size_t)nv * nx, sizeof(double));
    o->sh = calloc((size_t)nv * nx, sizeof(double));
    o->sb = calloc((size_t)nv * nx, sizeof(double));
    o->flux = calloc(3 * (nx + 1), sizeof(double));
    o->fh = calloc((size_t)nv * (nx + 1), sizeof(double));
    o->fb = calloc((size_t)nv * (nx + 1), sizeof(double));
    for (int q = 0; q < 12; ++q) o->t[q] = calloc(nv + 8, sizeof(double));
    for (int i = 0; i < nx; ++i) {
        /* init_flowField / init_distribution_function,
         * 1d/src/initialization.f90:115-193 */
        const double *p = d->prim0 + 3 * i;
        conserved1(d, p, o->w + 3 * i);
        maxwell1(d, nv, o->u, p, o->h + (size_t)nv * i, o->b + (size_t)nv * i);
    }
    o->iter = 1;
    return o;
}

void oracle1d_destroy(oracle1d *o)
{
    if (!o) return;
    free(o->u); free(o->weight); free(o->w); free(o->h); free(o->b); free(o->sh);
    free(o->sb); free(o->flux); free(o->fh); free(o->fb);
    for (int q = 0; q < 12; ++q) free(o->t[q]);
    free(o);
}

int oracle1d_step(oracle1d *o, hugks_step_info *info)
{
    const hugks_desc *d = &o->d;
    int nx = o->nx, nv = o->nv;
    /* timestep, 1d/src/solver.f90:16-47 */
    double temp = 0.0;
    for (int i = 0; i < nx; ++i) {
        double prim[3];
        primary1(d, o->w + 3 * i, prim);
        double sos = sqrt(0.5 * d->gamma / prim[2]);
        double pu = fmax(d->umax_dt, fabs(prim[1])) + sos;
        temp = fmax(temp, fmax(pu / d->dx, d->force_rate));
    }
    o->dt = d->cfl * (1.0 / temp) / 1.1;
    /* interpolation, 1d/src/solver.f90:49-104 */
    double den = 0.5 * d->dx + 0.5 * d->dx;
    for (int k = 0; k < nv; ++k) {
        o->sh[k] = (o->h[nv + k] - o->h[k]) / den;
        o->sb[k] = (o->b[nv + k] - o->b[k]) / den;
        size_t a = (size_t)nv * (nx - 1), m = (size_t)nv * (nx - 2);
        o->sh[a + k] = (o->h[a + k] - o->h[m + k]) / den;
        o->sb[a + k] = (o->b[a + k] - o->b[m + k]) / den;
    }
    for (int i = 1; i < nx - 1; ++i)
        for (int k = 0; k < nv; ++k)
            for (int q = 0; q < 2; ++q) {
                const double *f = q ? o->b : o->h;
                double *s = q ? o->sb : o->sh;
                double sL = (f[(size_t)nv * i + k] - f[(size_t)nv * (i - 1) + k]) / den;
                double sR = (f[(size_t)nv * (i + 1) + k] - f[(size_t)nv * i + k]) / den;
                s[(size_t)nv * i + k] = (copysign(1.0, sR) + copysign(1.0, sL)) * fabs(sR) * fabs(sL) /
                                        (fabs(sR) + fabs(sL) + d->smv);
            }
    /* evolution, 1d/src/solver.f90:106-139 */
    wall1(o, d->bc[0], 1, d->tw[0], 0, o->flux, o->fh, o->fb);
    wall1(o, d->bc[1], -1, d->tw[1], nx - 1, o->flux + 3 * nx, o->fh + (size_t)nv * nx,
          o->fb + (size_t)nv * nx);
    for (int i = 1; i < nx; ++i)
        flux1(o, i - 1, i, o->flux + 3 * i, o->fh + (size_t)nv * i, o->fb + (size_t)nv * i);
    /* update, 1d/src/solver.f90:141-264 */
    double sum_res[3] = {0, 0, 0}, sum_avg[3] = {0, 0, 0};
    double *Ho = o->t[0], *Bo = o->t[1], *H = o->t[2], *B = o->t[3], *Hp = o->t[4];
    double *Bp = o->t[5], *dvh = o->t[6], *dvb = o->t[7];
    o->steady_res[0] = o->steady_res[1] = o->steady_res[2] = 0.0;
    for (int i = 0; i < nx; ++i) {
        double *w = o->w + 3 * i, *h = o->h + (size_t)nv * i, *b = o->b + (size_t)nv * i;
        double w_old[3], prim_old[3], prim[3];
        memcpy(w_old, w, sizeof(w_old));
        primary1(d, w_old, prim_old);
        maxwell1(d, nv, o->u, prim_old, Ho, Bo);
        for (int c = 0; c < 3; ++c) w[c] = w[c] + (o->flux[3 * i + c] - o->flux[3 * (i + 1) + c]) / d->dx;
        primary1(d, w, prim);
        prim[1] = prim[1] + d->phi * o->dt;
        conserved1(d, prim, w);
        /* primitive_change_residual, 1d/src/solver.f90:276-291 */
        double ds = fmax(1.0, fmax(fabs(prim_old[0]), fabs(prim[0])));
        double loc[3] = {fabs(prim[0] - prim_old[0]) / ds,
                         fabs(prim[1] - prim_old[1]) / sqrt(d->tref),
                         fabs(1.0 / prim[2] - 1.0 / prim_old[2]) / d->tref};
        for (int c = 0; c < 3; ++c) o->steady_res[c] = fmax(o->steady_res[c], loc[c]);
        maxwell1(d, nv, o->u, prim, H, B);
        double tau = tau1(d, prim);
        for (int c = 0; c < 3; ++c) {
            sum_res[c] += (w_old[c] - w[c]) * (w_old[c] - w[c]);
            sum_avg[c] += fabs(w[c]);
        }
        double qf = heat1(nv, o->weight, o->u, h, b, prim_old);
        shakhov1(d, nv, o->u, Ho, Bo, qf, prim_old, Hp, Bp);
        shakhov1(d, nv, o->u, H, B, qf, prim, Hp, Bp);
        for (int k = 0; k < nv; ++k) { H[k] += Hp[k]; B[k] += Bp[k]; }
        if (d->force_variant == HUGKS_FORCE_FINITE_DIFFERENCE)
            calfv1(d, nv, h, b, dvh, dvb);   /* 1d/src/solver.f90:234-235 */
        else
            calfv(d, nv, o->u, o->weight, d->nm, h, b, prim_old, dvh, dvb);
        const double *fhl = o->fh + (size_t)nv * i, *fhr = o->fh + (size_t)nv * (i + 1);
        const double *fbl = o->fb + (size_t)nv * i, *fbr = o->fb + (size_t)nv * (i + 1);
        for (int k = 0; k < nv; ++k) {
            double hs = dvh[k] * o->dt * d->phi, bs = dvb[k] * o->dt * d->phi;
            h[k] = (h[k] + (fhl[k] - fhr[k]) / d->dx + o->dt * (H[k] / tau) - hs) / (1.0 + o->dt / tau);
            b[k] = (b[k] + (fbl[k] - fbr[k]) / d->dx + o->dt * (B[k] / tau) - bs) / (1.0 + o->dt / tau);
        }
    }
    for (int c = 0; c < 3; ++c) o->res[c] = sqrt((double)nx * sum_res[c]) / (sum_avg[c] + d->smv);
    o->iter += 1;
    o->simtime = o->simtime + o->dt;
    if (info) {
        memset(info, 0, sizeof(*info));
        info->steps = 1;
        info->iter = o->iter;
        info->simtime = o->simtime;
        info->dt = o->dt;
        memcpy(info->res, o->res, sizeof(o->res));
        memcpy(info->steady_res, o->steady_res, sizeof(o->steady_res));
        if (o->simtime >= d->maxtime) info->stop_reason = HUGKS_STOP_MAXTIME;
    }
    return 0;
}

void oracle1d_get(const oracle1d *o, double *w, double *h, double *b)
{
    if (w) memcpy(w, o->w, 3 * o->nx * sizeof(double));
    if (h) memcpy(h, o->h, (size_t)o->nv * o->nx * sizeof(double));
    if (b) memcpy(b, o->b, (size_t)o->nv * o->nx * sizeof(double));
}

void oracle1d_set(oracle1d *o, const double *w, const double *h, const double *b)
{
    if (w) memcpy(o->w, w, 3 * o->nx * sizeof(double));
    if (h) memcpy(o->h, h, (size_t)o->nv * o->nx * sizeof(double));
    if (b) memcpy(o->b, b, (size_t)o->nv * o->nx * sizeof(double));
}

void oracle1d_time(const oracle1d *o, double *simtime, int32_t *iter, double *dt)
{
    if (simtime) *simtime = o->simtime;
    if (iter) *iter = o->iter;
    if (dt) *dt = o->dt;
}

void oracle1d_heat_flux(const oracle1d *o, double *q)
{
    for (int i = 0; i < o->nx; ++i) {
        double prim[3];
        primary1(&o->d, o->w + 3 * i, prim);
        q[i] = heat1(o->nv, o->weight, o->u, o->h + (size_t)o->nv * i, o->b + (size_t)o->nv * i, prim);
    }
}

void oracle1d_hermite_force(const hugks_desc *d, int order, const double *h, const double *b,
                            const double *prim, double *dvh, double *dvb)
{
    calfv(d, d->nv, d->u, d->weight, order, h, b, prim, dvh, dvb);
}
