// Interior-face UGKS flux (gfx950): calc_flux_mtforcev / calc_flux_mtforceh,
// 2d/src/flux.f90:28-472.
//
// One wavefront per face (four faces per 256-thread workgroup, no
// workgroup barriers).  Lanes stride over the face's nv*nv velocity nodes
// in three passes; each pass re-reads the upwind cell's six inputs (h, b
// and their normal/tangential slopes) from memory, where the second and
// third reads hit the caches, instead of holding them in registers across
// the reductions:
//   pass 1  reconstruct f and its slopes at the face (flux.f90:68-87) and
//           form the 31 velocity sums that do not depend on the face state:
//           W at the face, the tangential-slope moments and every
//           Mt(4)/Mt(5)/Mt(6) moment of flux.f90:184-202;
//   pass 2  with the face state known: heat flux (get_heat_flux), force
//           compatibility (flux.f90:691-705) and the Shakhov moments,
//           split into the two terms linear in the heat flux so they are
//           reduced in the same pass;
//   pass 3  the distribution fluxes of h and b per node (flux.f90:207-236).
// Velocity sums are accumulated in FP64; the wavefront reduction is a fixed
// butterfly whose lane-0 result is broadcast, so every lane holds the same
// bits and reruns are bitwise reproducible.
#pragma once
#include "hugks_device.h"
#include "hugks_kernels.h"

namespace hugks {

template <int NV>
__device__ __forceinline__ void wave_allsum(double (&v)[NV]) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double x = v[i];
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
        v[i] = __shfl(x, 0, WAVE);
    }
}

constexpr int FACES_PER_WG = 4;

// DIR 0: vertical faces (x-normal); DIR 1: horizontal faces (y-normal).
// vn/vt are the face-local velocities: (u, v) on vertical faces, (v, -u)
// on horizontal ones, which is what the reference's rotation by
// (cosa, sina) = (1,0) / (0,1) produces.
template <int DIR, typename T>
__global__ __launch_bounds__(WAVE * FACES_PER_WG) void k_flux_interior(
    Consts c, const DevState* __restrict__ st, const double* __restrict__ W,
    const T* __restrict__ h, const T* __restrict__ b, const T* __restrict__ shx,
    const T* __restrict__ shy, const T* __restrict__ sbx, const T* __restrict__ sby,
    const typename Arith<T>::A* __restrict__ unode, const typename Arith<T>::A* __restrict__ vnode,
    const typename Arith<T>::A* __restrict__ wnode, const double* __restrict__ accel,
    T* __restrict__ fh, T* __restrict__ fb, double* __restrict__ flux, int nfaces) {
    using A = typename Arith<T>::A;
    if (st->stopped) return;
    const int lane = threadIdx.x & (WAVE - 1);
    const int f = blockIdx.x * FACES_PER_WG + (int)(threadIdx.x / WAVE);
    if (f >= nfaces) return;   // the whole wavefront leaves together
    int cL, cR, face;
    if (DIR == 0) {
        const bool per = c.bc[0] == HUGKS_BC_PERIODIC;
        const int nf = per ? c.nx + 1 : c.nx - 1;
        const int i = (per ? 0 : 1) + f % nf, j = f / nf;
        face = i + (c.nx + 1) * j;
        cL = (i - 1 + c.nx) % c.nx + c.nx * j;
        cR = i % c.nx + c.nx * j;
    } else {
        const bool per = c.bc[2] == HUGKS_BC_PERIODIC;
        const int i = f % c.nx, j = (per ? 0 : 1) + f / c.nx;
        face = i + c.nx * j;
        cL = i + c.nx * ((j - 1 + c.ny) % c.ny);
        cR = i + c.nx * (j % c.ny);
    }
    const T* sN = DIR == 0 ? shx : shy;
    const T* sT = DIR == 0 ? shy : shx;
    const T* bN = DIR == 0 ? sbx : sby;
    const T* bT = DIR == 0 ? sby : sbx;
    const double dt = st->dt;
    const double fn = accel[2 * face], ft = accel[2 * face + 1];
    const double len = DIR == 0 ? c.dx : c.dy;    // cell length normal to the face
    const double flen = DIR == 0 ? c.dy : c.dx;   // face length
    const A tsign = DIR == 0 ? (A)1 : (A)-1;
    const A hlen = (A)(0.5 * len);
    const A vshift = (A)(0.5 * fn * dt);
    const size_t baseL = (size_t)cL * c.ldv, baseR = (size_t)cR * c.ldv;
    const int nn = c.nn;

    // upwind reconstruction of node k (flux.f90:74-87)
    struct Rec {
        A vn, vt, wt, h, b, sn, st, bn, bt;
        int d;
    };
    auto rec = [&](int k, bool full) {
        Rec r;
        const A u = unode[k], v = vnode[k];
        r.wt = wnode[k];
        r.vn = DIR == 0 ? u : v;
        r.vt = DIR == 0 ? v : -u;
        r.d = upwind_delta(r.vn - vshift);
        const size_t o = (r.d ? baseL : baseR) + k;
        const A sgn = r.d ? hlen : -hlen;
        r.sn = (A)sN[o];
        r.bn = (A)bN[o];
        r.h = (A)h[o] + sgn * r.sn;
        r.b = (A)b[o] + sgn * r.bn;
        if (full) {
            r.st = tsign * (A)sT[o];
            r.bt = tsign * (A)bT[o];
        } else {
            r.st = r.bt = (A)0;
        }
        return r;
    };

    // ---- pass 1
    double s1[31];
#pragma unroll
    for (int q = 0; q < 31; ++q) s1[q] = 0.0;
#pragma unroll 1
    for (int k = lane; k < nn; k += WAVE) {
        const Rec r = rec(k, true);
        const A vn = r.vn, vt = r.vt, wt = r.wt;
        const A c2 = vn * vn + vt * vt;
        const A wh = wt * r.h, wsn = wt * r.sn, wst = wt * r.st;
        s1[0] += (double)wh;
        s1[1] += (double)(vn * wh);
        s1[2] += (double)(vt * wh);
        s1[3] += (double)(c2 * wh);
        s1[4] += (double)(wt * r.b);
        s1[5] += (double)wst;
        s1[6] += (double)(vn * wst);
        s1[7] += (double)(vt * wst);
        s1[8] += (double)(c2 * wst);
        s1[9] += (double)(wt * r.bt);
        s1[10] += (double)(vn * vn * wsn);
        s1[11] += (double)(vt * vn * wst);
        s1[12] += (double)(vn * vn * wh);
        s1[13] += (double)(vn * vn * vn * wsn);
        s1[14] += (double)(vn * vn * vt * wst);
        s1[15] += (double)(vt * vn * wh);
        s1[16] += (double)(vt * vn * vn * wsn);
        s1[17] += (double)(vt * vt * vn * wst);
        s1[18] += (double)(vn * c2 * wh);
        s1[19] += (double)(vn * wt * r.b);
        s1[20] += (double)(vn * vn * c2 * wsn);
        s1[21] += (double)(vn * vn * wt * r.bn);
        s1[22] += (double)(vt * vn * c2 * wst);
        s1[23] += (double)(vt * vn * wt * r.bt);
        s1[24] += (double)(vn * wsn);
        s1[25] += (double)(vt * vn * wsn);
        s1[26] += (double)(c2 * vn * wsn);
        s1[27] += (double)(vn * wt * r.bn);
        s1[28] += (double)(vn * vn * wst);
        s1[29] += (double)(c2 * vn * wst);
        s1[30] += (double)(vn * wt * r.bt);
    }
    wave_allsum<31>(s1);

    // ---- state at the face (flux.f90:93-99) and everything that depends
    // only on it and the pass-1 sums, folded now so the sums die here
    double p[4];
    {
        const double w[4] = {s1[0], s1[1], s1[2], 0.5 * (s1[3] + s1[4])};
        prim_from_w(c, w, p);
    }
    const double tau = tau_of(c, p[0], p[3]);
    double Mt[7];
    {
        // time-integration coefficients (flux.f90:140-147)
        const double ex = exp(-dt / tau);
        Mt[4] = tau * (1.0 - ex);
        Mt[5] = -tau * dt * ex + tau * Mt[4];
        Mt[1] = dt - Mt[4];
        Mt[2] = -tau * Mt[1] + Mt[5];
        Mt[3] = dt * dt / 2.0 - tau * Mt[1];
        Mt[6] = tau * tau * tau - 0.5 * ex * tau * (dt * dt + 2.0 * dt * tau + 2.0 * tau * tau);
    }
    double aL[4], aR[4], ab[4], fl[4];
    {
        // aL, aR (flux.f90:107-111), cell states in the face-local frame
        const double w[4] = {s1[0], s1[1], s1[2], 0.5 * (s1[3] + s1[4])};
        const double* wl = W + 4 * (size_t)cL;
        const double* wr = W + 4 * (size_t)cR;
        double locL[4], locR[4], sw[4];
        if (DIR == 0) {
            for (int q = 0; q < 4; ++q) { locL[q] = wl[q]; locR[q] = wr[q]; }
        } else {
            locL[0] = wl[0]; locL[1] = wl[2]; locL[2] = -wl[1]; locL[3] = wl[3];
            locR[0] = wr[0]; locR[1] = wr[2]; locR[2] = -wr[1]; locR[3] = wr[3];
        }
        for (int q = 0; q < 4; ++q) sw[q] = (w[q] - locL[q]) / (0.5 * len);
        micro_slope(c, p, sw, aL);
        for (int q = 0; q < 4; ++q) sw[q] = (locR[q] - w[q]) / (0.5 * len);
        micro_slope(c, p, sw, aR);
        // tangential slope (flux.f90:116-120)
        sw[0] = s1[5]; sw[1] = s1[6]; sw[2] = s1[7]; sw[3] = 0.5 * (s1[8] + s1[9]);
        micro_slope(c, p, sw, ab);
        // f0 moments of the macro flux, flux.f90:184-202 (the Shakhov
        // moments are added after pass 2)
        fl[0] = Mt[4] * s1[1] - Mt[5] * s1[10] - Mt[5] * s1[11];
        fl[1] = Mt[4] * s1[12] - Mt[5] * s1[13] - Mt[5] * s1[14];
        fl[2] = Mt[4] * s1[15] - Mt[5] * s1[16] - Mt[5] * s1[17];
        fl[3] = Mt[4] * 0.5 * (s1[18] + s1[19]) - Mt[5] * 0.5 * (s1[20] + s1[21]) -
                Mt[5] * 0.5 * (s1[22] + s1[23]);
        const double g[4] = {fn * Mt[6] * s1[24] + ft * Mt[6] * s1[6],
                             fn * Mt[6] * s1[10] + ft * Mt[6] * s1[28],
                             fn * Mt[6] * s1[25] + ft * Mt[6] * s1[11],
                             fn * Mt[6] * 0.5 * (s1[26] + s1[27]) + ft * Mt[6] * 0.5 * (s1[29] + s1[30])};
        for (int q = 0; q < 4; ++q) fl[q] += g[q];
    }
    const A U = (A)p[1], V = (A)p[2], lam = (A)p[3];
    const A rho_l_pi = (A)(p[0] * (p[3] / c.pi));
    const A kfac = (A)(c.K / (2.0 * p[3]));
    const A fnA = (A)fn, ftA = (A)ft;
    const A K = (A)c.K;

    // ---- pass 2
    double s2[17];
#pragma unroll
    for (int q = 0; q < 17; ++q) s2[q] = 0.0;
#pragma unroll 1
    for (int k = lane; k < nn; k += WAVE) {
        const Rec r = rec(k, false);
        const A vn = r.vn, vt = r.vt, wt = r.wt;
        const A cu = vn - U, cv = vt - V;
        const A cp2 = cu * cu + cv * cv;
        const A c2 = vn * vn + vt * vt;
        const A H0 = rho_l_pi * fexp(-lam * cp2);
        const A B0 = H0 * kfac;
        const A wg = wt * lam * (fnA * cu + ftA * cv);
        s2[0] += (double)(wt * cu * cp2 * r.h);
        s2[1] += (double)(wt * cu * r.b);
        s2[2] += (double)(wt * cv * cp2 * r.h);
        s2[3] += (double)(wt * cv * r.b);
        s2[4] += (double)(vn * wg * H0);
        s2[5] += (double)(vt * wg * H0);
        s2[6] += (double)(wg * (c2 * H0 + B0));
        const A lam2 = (A)2 * lam * cp2;
        const A P5H = (lam2 + K - (A)5) * H0;
        const A P3B = (lam2 + K - (A)3) * B0;
        const A wvn = wt * vn;
        s2[7] += (double)(wvn * cu * P5H);
        s2[8] += (double)(wvn * cv * P5H);
        s2[9] += (double)(wvn * vn * cu * P5H);
        s2[10] += (double)(wvn * vn * cv * P5H);
        s2[11] += (double)(wvn * vt * cu * P5H);
        s2[12] += (double)(wvn * vt * cv * P5H);
        s2[13] += (double)(wvn * c2 * cu * P5H);
        s2[14] += (double)(wvn * c2 * cv * P5H);
        s2[15] += (double)(wvn * cu * P3B);
        s2[16] += (double)(wvn * cv * P3B);
    }
    wave_allsum<17>(s2);

    // ---- time slope and the remaining macro flux (every lane computes the
    // same values; lane 0 stores the flux)
    const double qf0 = 0.5 * (s2[0] + s2[1]), qf1 = 0.5 * (s2[2] + s2[3]);
    Moments m;
    calc_moment(c, p, m);
    double aT[4];
    {
        // flux.f90:126-135
        double MauL[4], MauR[4], Mbu[4], sw[4];
        moment_au(aL, m.MuL, m.Mv, m.Mxi, 1, 0, MauL);
        moment_au(aR, m.MuR, m.Mv, m.Mxi, 1, 0, MauR);
        moment_au(ab, m.Mu, m.Mv, m.Mxi, 0, 1, Mbu);
        for (int q = 0; q < 4; ++q) sw[q] = -p[0] * (MauL[q] + MauR[q] + Mbu[q]);
        sw[1] = sw[1] + 2.0 * s2[4];
        sw[2] = sw[2] + 2.0 * s2[5];
        sw[3] = sw[3] + s2[6];
        micro_slope(c, p, sw, aT);
    }
    // The conserved-flux assembly needs only scalars; its inputs wait in a
    // per-wavefront LDS stash until pass 3 is done, so the moment algebra
    // does not overlap the per-node registers of pass 3.
    __shared__ double stash_all[FACES_PER_WG * 48];
    volatile double* stash = stash_all + 48 * (threadIdx.x / WAVE);
    if (lane == 0) {
        for (int q = 0; q < 4; ++q) {
            stash[q] = p[q];
            stash[4 + q] = aL[q];
            stash[8 + q] = aR[q];
            stash[12 + q] = ab[q];
            stash[16 + q] = aT[q];
            stash[20 + q] = fl[q];
        }
        for (int q = 1; q <= 6; ++q) stash[23 + q] = Mt[q];
        stash[30] = tau;
        stash[31] = qf0;
        stash[32] = qf1;
        for (int q = 7; q <= 16; ++q) stash[26 + q] = s2[q];
    }

    // ---- pass 3: distribution fluxes (flux.f90:207-236)
    const A mt1 = (A)Mt[1], mt2 = (A)Mt[2], mt3 = (A)Mt[3], mt4 = (A)Mt[4], mt5 = (A)Mt[5],
            mt6 = (A)Mt[6];
    const A tm2fn = (A)(tau * Mt[2] * fn), tm2ft = (A)(tau * Mt[2] * ft);
    const A mxi2 = (A)m.Mxi[2];
    const A csA = (A)(c.shk * (1.0 - c.prandtl) * p[3] * p[3] / p[0]);
    const A q0 = (A)qf0, q1 = (A)qf1;
    const A flenA = (A)flen;
    A a_L[4], a_R[4], a_b[4], a_T[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        a_L[q] = (A)aL[q]; a_R[q] = (A)aR[q]; a_b[q] = (A)ab[q]; a_T[q] = (A)aT[q];
    }
    T* oh = fh + (size_t)face * c.ldv;
    T* ob = fb + (size_t)face * c.ldv;
#pragma unroll 1
    for (int k = lane; k < nn; k += WAVE) {
        const Rec r = rec(k, true);
        const A vn = r.vn, vt = r.vt;
        const A cu = vn - U, cv = vt - V;
        const A cp2 = cu * cu + cv * cv;
        const A c2 = vn * vn + vt * vt;
        const A H = rho_l_pi * fexp(-lam * cp2);
        const A Bm = H * kfac;
        const A cq = csA * (cu * q0 + cv * q1);
        const A lam2 = (A)2 * lam * cp2;
        const A Hp = cq * (lam2 + K - (A)5) * H;
        const A Bp = cq * (lam2 + K - (A)3) * Bm;
        const A eh = c2 * H + Bm, eb = c2 * Bm + mxi2 * H;
        const A u0 = r.d ? a_L[0] : a_R[0], u1 = r.d ? a_L[1] : a_R[1];
        const A u2 = r.d ? a_L[2] : a_R[2], u3 = r.d ? a_L[3] : a_R[3];
        const A gh = u0 * H + u1 * vn * H + u2 * vt * H + (A)0.5 * u3 * eh;
        const A gb = u0 * Bm + u1 * vn * Bm + u2 * vt * Bm + (A)0.5 * u3 * eb;
        const A gbh = a_b[0] * H + a_b[1] * vn * H + a_b[2] * vt * H + (A)0.5 * a_b[3] * eh;
        const A gbb = a_b[0] * Bm + a_b[1] * vn * Bm + a_b[2] * vt * Bm + (A)0.5 * a_b[3] * eb;
        const A gTh = a_T[0] * H + a_T[1] * vn * H + a_T[2] * vt * H + (A)0.5 * a_T[3] * eh;
        const A gTb = a_T[0] * Bm + a_T[1] * vn * Bm + a_T[2] * vt * Bm + (A)0.5 * a_T[3] * eb;
        A xh = mt1 * vn * (H + Hp) + mt2 * vn * vn * gh + mt2 * vn * vt * gbh + mt3 * vn * gTh +
               mt4 * vn * r.h - mt5 * vn * vn * r.sn - mt5 * vn * vt * r.st;
        A xb = mt1 * vn * (Bm + Bp) + mt2 * vn * vn * gb + mt2 * vn * vt * gbb + mt3 * vn * gTb +
               mt4 * vn * r.b - mt5 * vn * vn * r.bn - mt5 * vn * vt * r.bt;
        xh = xh - tm2fn * vn * gh + mt6 * fnA * vn * r.sn;
        xb = xb - tm2fn * vn * gb + mt6 * fnA * vn * r.bn;
        xh = xh - tm2ft * vn * gbh + mt6 * ftA * vn * r.st;
        xb = xb - tm2ft * vn * gbb + mt6 * ftA * vn * r.bt;
        oh[k] = (T)(flenA * xh);
        ob[k] = (T)(flenA * xb);
    }
    if (lane == 0) {
        double p[4], aL[4], aR[4], ab[4], aT[4], fl[4], Mt[7], s2[17];
        for (int q = 0; q < 4; ++q) {
            p[q] = stash[q];
            aL[q] = stash[4 + q];
            aR[q] = stash[8 + q];
            ab[q] = stash[12 + q];
            aT[q] = stash[16 + q];
            fl[q] = stash[20 + q];
        }
        for (int q = 1; q <= 6; ++q) Mt[q] = stash[23 + q];
        const double tau = stash[30], qf0 = stash[31], qf1 = stash[32];
        for (int q = 7; q <= 16; ++q) s2[q] = stash[26 + q];
        Moments m;
        calc_moment(c, p, m);
        double MauL[4], MauR[4], Mbu[4], MauT[4], Mau0[4], g0[4];
        // g0 part (flux.f90:152-172)
        moment_uv(m.Mu, m.Mv, m.Mxi, 1, 0, 0, Mau0);
        moment_au(aL, m.MuL, m.Mv, m.Mxi, 2, 0, MauL);
        moment_au(aR, m.MuR, m.Mv, m.Mxi, 2, 0, MauR);
        moment_au(ab, m.Mu, m.Mv, m.Mxi, 1, 1, Mbu);
        moment_au(aT, m.Mu, m.Mv, m.Mxi, 1, 0, MauT);
        for (int q = 0; q < 4; ++q)
            g0[q] = Mt[1] * p[0] * Mau0[q] + Mt[2] * p[0] * (MauL[q] + MauR[q] + Mbu[q]) +
                    Mt[3] * p[0] * MauT[q];
        moment_au(aL, m.MuL, m.Mv, m.Mxi, 1, 0, MauL);
        moment_au(aR, m.MuR, m.Mv, m.Mxi, 1, 0, MauR);
        for (int q = 0; q < 4; ++q)
            g0[q] = g0[q] - tau * Mt[2] * p[0] * fn * (MauL[q] + MauR[q]) -
                    Mt[6] * p[0] * fn * (MauL[q] + MauR[q]);
        moment_au(ab, m.Mu, m.Mv, m.Mxi, 0, 1, Mbu);
        for (int q = 0; q < 4; ++q)
            g0[q] = g0[q] - tau * Mt[2] * p[0] * ft * Mbu[q] - Mt[6] * p[0] * ft * Mbu[q];
        // Shakhov moments from their heat-flux split (flux.f90:184-191)
        const double cs = c.shk * (1.0 - c.prandtl) * p[3] * p[3] / p[0];
        const double X0 = cs * (qf0 * s2[7] + qf1 * s2[8]);
        const double X1 = cs * (qf0 * s2[9] + qf1 * s2[10]);
        const double X2 = cs * (qf0 * s2[11] + qf1 * s2[12]);
        const double X3 = cs * (qf0 * s2[13] + qf1 * s2[14]);
        const double X4 = cs * (qf0 * s2[15] + qf1 * s2[16]);
        const double out4[4] = {g0[0] + Mt[1] * X0 + fl[0], g0[1] + Mt[1] * X1 + fl[1],
                                g0[2] + Mt[1] * X2 + fl[2], g0[3] + Mt[1] * 0.5 * (X3 + X4) + fl[3]};
        // global frame (flux.f90:242) and face length
        double g1 = out4[1], g2 = out4[2];
        if (DIR == 1) { g1 = -out4[2]; g2 = out4[1]; }
        double* out = flux + 4 * (size_t)face;
        out[0] = flen * out4[0];
        out[1] = flen * g1;
        out[2] = flen * g2;
        out[3] = flen * out4[3];
    }

}

}  // namespace hugks
