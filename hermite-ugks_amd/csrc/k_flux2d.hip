// 2D slope, interior-face flux and wall-face flux kernels (gfx950).
//
// Interior flux: one workgroup of NT threads per face, EPT velocity nodes
// per thread held in registers through three passes:
//   pass 1  reconstruct f, its normal and tangential slopes at the face and
//           form the 31 prim-independent velocity sums (W at the face, the
//           tangential slope moments and every Mt(4)/Mt(5)/Mt(6) moment);
//   pass 2  with prim known: heat flux, force compatibility and the
//           Shakhov moments, the latter split into the two terms linear in
//           the heat flux so they are reduced in the same pass;
//   pass 3  elementwise face fluxes of h and b.
// Reductions accumulate in FP64 with a fixed order.  Reference:
// calc_flux_mtforcev / calc_flux_mtforceh, 2d/src/flux.f90:28-472.
#include "hugks_device.h"
#include "hugks_kernels.h"

namespace hugks {

// ------------------------------------------------------------------ slopes
// interp_inner (van Leer) / interp_boundary (one-sided), 2d/src/solver.f90:
// 64-159.  Periodic directions use wrapped neighbours everywhere (the ghost
// copy of 2d/src/boundary.f90:16-53); wall directions use one-sided slopes
// in the two boundary cells.
template <typename T>
__global__ __launch_bounds__(256) void k_slope2d(Consts c, const DevState* __restrict__ st,
                                                 const T* __restrict__ h, const T* __restrict__ b,
                                                 T* __restrict__ shx, T* __restrict__ shy,
                                                 T* __restrict__ sbx, T* __restrict__ sby) {
    using A = typename Arith<T>::A;
    if (st->stopped) return;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t ncell = (size_t)c.nx * c.ny;
    if (idx >= ncell * c.ldv) return;
    const int cell = (int)(idx / c.ldv);
    const int k = (int)(idx - (size_t)cell * c.ldv);
    const int i = cell % c.nx, j = cell / c.nx;
    const A smv = (A)c.smv;
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {
        const bool per = c.bc[2 * dir] == HUGKS_BC_PERIODIC;
        const int n = dir == 0 ? c.nx : c.ny;
        const int p = dir == 0 ? i : j;
        const A len = (A)(dir == 0 ? c.dx : c.dy);
        int cl, cr;
        bool inner = true;
        if (per) {
            const int pl = (p - 1 + n) % n, pr = (p + 1) % n;
            cl = dir == 0 ? pl + c.nx * j : i + c.nx * pl;
            cr = dir == 0 ? pr + c.nx * j : i + c.nx * pr;
        } else if (p == 0) {
            cl = cell;
            cr = dir == 0 ? cell + 1 : cell + c.nx;
            inner = false;
        } else if (p == n - 1) {
            cl = dir == 0 ? cell - 1 : cell - c.nx;
            cr = cell;
            inner = false;
        } else {
            cl = dir == 0 ? cell - 1 : cell - c.nx;
            cr = dir == 0 ? cell + 1 : cell + c.nx;
        }
        const A den = (A)0.5 * len + (A)0.5 * len;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const T* f = q == 0 ? h : b;
            const A fl = (A)f[(size_t)cl * c.ldv + k], fr = (A)f[(size_t)cr * c.ldv + k];
            A s;
            if (inner) {
                const A fn = (A)f[(size_t)cell * c.ldv + k];
                const A sL = (fn - fl) / den, sR = (fr - fn) / den;
                s = (copysign((A)1, sR) + copysign((A)1, sL)) * fabs(sR) * fabs(sL) /
                    (fabs(sR) + fabs(sL) + smv);
            } else {
                s = (fr - fl) / den;
            }
            T* out = dir == 0 ? (q == 0 ? shx : sbx) : (q == 0 ? shy : sby);
            out[(size_t)cell * c.ldv + k] = (T)s;
        }
    }
}

// --------------------------------------------------------- interior faces
// DIR 0: vertical faces (x-normal, calc_flux_mtforcev); DIR 1: horizontal
// faces (y-normal, calc_flux_mtforceh).  vn/vt are the face-local
// velocities: (u, v) on vertical faces and (v, -u) on horizontal ones,
// which is what the reference's rotation by (cosa, sina) = (1,0) / (0,1)
// produces.
template <int DIR, typename T, int EPT, int NT>
__global__ __launch_bounds__(NT) void k_flux_interior(
    Consts c, const DevState* __restrict__ st, const double* __restrict__ W,
    const T* __restrict__ h, const T* __restrict__ b, const T* __restrict__ shx,
    const T* __restrict__ shy, const T* __restrict__ sbx, const T* __restrict__ sby,
    const typename Arith<T>::A* __restrict__ unode, const typename Arith<T>::A* __restrict__ vnode,
    const typename Arith<T>::A* __restrict__ wnode, const double* __restrict__ accel,
    T* __restrict__ fh, T* __restrict__ fb, double* __restrict__ flux) {
    using A = typename Arith<T>::A;
    if (st->stopped) return;
    __shared__ double lds[31 * (NT / WAVE)];
    const int tid = threadIdx.x;
    const int f = blockIdx.x;
    // face decode
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
    const double len = DIR == 0 ? c.dx : c.dy;     // cell length normal to the face
    const double flen = DIR == 0 ? c.dy : c.dx;    // face length
    const A tsign = DIR == 0 ? (A)1 : (A)-1;
    const A hlen = (A)(0.5 * len);
    const A vshift = (A)(0.5 * fn * dt);

    A rh[EPT], rb[EPT], rsn[EPT], rst[EPT], rbn[EPT], rbt[EPT];
    int dl[EPT];
    double s1[31];
#pragma unroll
    for (int q = 0; q < 31; ++q) s1[q] = 0.0;

    // ---- pass 1: reconstruction (flux.f90:68-87) and prim-free sums
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = tid + e * NT;
        rh[e] = rb[e] = rsn[e] = rst[e] = rbn[e] = rbt[e] = (A)0;
        dl[e] = 0;
        if (k < c.nn) {
            const A u = unode[k], v = vnode[k], wt = wnode[k];
            const A vn = DIR == 0 ? u : v;
            const A vt = DIR == 0 ? v : -u;
            const int d = upwind_delta(vn - vshift);
            dl[e] = d;
            const size_t oL = (size_t)cL * c.ldv + k, oR = (size_t)cR * c.ldv + k;
            if (d) {
                rsn[e] = (A)sN[oL];
                rbn[e] = (A)bN[oL];
                rh[e] = (A)h[oL] + hlen * rsn[e];
                rb[e] = (A)b[oL] + hlen * rbn[e];
                rst[e] = tsign * (A)sT[oL];
                rbt[e] = tsign * (A)bT[oL];
            } else {
                rsn[e] = (A)sN[oR];
                rbn[e] = (A)bN[oR];
                rh[e] = (A)h[oR] - hlen * rsn[e];
                rb[e] = (A)b[oR] - hlen * rbn[e];
                rst[e] = tsign * (A)sT[oR];
                rbt[e] = tsign * (A)bT[oR];
            }
            const A c2 = vn * vn + vt * vt;
            const A wh = wt * rh[e], wsn = wt * rsn[e], wst = wt * rst[e];
            s1[0] += (double)wh;
            s1[1] += (double)(vn * wh);
            s1[2] += (double)(vt * wh);
            s1[3] += (double)(c2 * wh);
            s1[4] += (double)(wt * rb[e]);
            s1[5] += (double)wst;
            s1[6] += (double)(vn * wst);
            s1[7] += (double)(vt * wst);
            s1[8] += (double)(c2 * wst);
            s1[9] += (double)(wt * rbt[e]);
            s1[10] += (double)(vn * vn * wsn);
            s1[11] += (double)(vt * vn * wst);
            s1[12] += (double)(vn * vn * wh);
            s1[13] += (double)(vn * vn * vn * wsn);
            s1[14] += (double)(vn * vn * vt * wst);
            s1[15] += (double)(vt * vn * wh);
            s1[16] += (double)(vt * vn * vn * wsn);
            s1[17] += (double)(vt * vt * vn * wst);
            s1[18] += (double)(vn * c2 * wh);
            s1[19] += (double)(vn * wt * rb[e]);
            s1[20] += (double)(vn * vn * c2 * wsn);
            s1[21] += (double)(vn * vn * wt * rbn[e]);
            s1[22] += (double)(vt * vn * c2 * wst);
            s1[23] += (double)(vt * vn * wt * rbt[e]);
            s1[24] += (double)(vn * wsn);
            s1[25] += (double)(vt * vn * wsn);
            s1[26] += (double)(c2 * vn * wsn);
            s1[27] += (double)(vn * wt * rbn[e]);
            s1[28] += (double)(vn * vn * wst);
            s1[29] += (double)(c2 * vn * wst);
            s1[30] += (double)(vn * wt * rbt[e]);
        }
    }
    block_sum<31, NT>(s1, lds);

    // ---- macroscopic state at the face (flux.f90:93-99)
    double w[4] = {s1[0], s1[1], s1[2], 0.5 * (s1[3] + s1[4])};
    double p[4];
    prim_from_w(c, w, p);
    const A U = (A)p[1], V = (A)p[2], lam = (A)p[3];
    const A rho_l_pi = (A)(p[0] * (p[3] / c.pi));
    const A kfac = (A)(c.K / (2.0 * p[3]));
    const A fnA = (A)fn, ftA = (A)ft;

    // ---- pass 2: heat flux (get_heat_flux), force compatibility
    // (flux.f90:691-705) and the Shakhov moments split by heat-flux
    // component.
    double s2[17];
#pragma unroll
    for (int q = 0; q < 17; ++q) s2[q] = 0.0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = tid + e * NT;
        if (k < c.nn) {
            const A u = unode[k], v = vnode[k], wt = wnode[k];
            const A vn = DIR == 0 ? u : v;
            const A vt = DIR == 0 ? v : -u;
            const A cu = vn - U, cv = vt - V;
            const A cp2 = cu * cu + cv * cv;
            const A c2 = vn * vn + vt * vt;
            const A H0 = rho_l_pi * fexp(-lam * cp2);
            const A B0 = H0 * kfac;
            const A g = fnA * cu + ftA * cv;
            const A wg = wt * lam * g;
            s2[0] += (double)(wt * cu * cp2 * rh[e]);
            s2[1] += (double)(wt * cu * rb[e]);
            s2[2] += (double)(wt * cv * cp2 * rh[e]);
            s2[3] += (double)(wt * cv * rb[e]);
            s2[4] += (double)(vn * wg * H0);
            s2[5] += (double)(vt * wg * H0);
            s2[6] += (double)(wg * (c2 * H0 + B0));
            const A lam2 = (A)2 * lam * cp2;
            const A P5H = (lam2 + (A)c.K - (A)5) * H0;
            const A P3B = (lam2 + (A)c.K - (A)3) * B0;
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
    }
    block_sum<17, NT>(s2, lds);

    // ---- scalar part (every thread, identical values)
    const double qf0 = 0.5 * (s2[0] + s2[1]), qf1 = 0.5 * (s2[2] + s2[3]);
    double sw[4], aL[4], aR[4], ab[4], aT[4];
    {
        // aL, aR (flux.f90:107-111); the cell states in face-local frame
        const double* wl = W + 4 * (size_t)cL;
        const double* wr = W + 4 * (size_t)cR;
        double locL[4], locR[4];
        if (DIR == 0) {
            locL[0] = wl[0]; locL[1] = wl[1]; locL[2] = wl[2]; locL[3] = wl[3];
            locR[0] = wr[0]; locR[1] = wr[1]; locR[2] = wr[2]; locR[3] = wr[3];
        } else {
            locL[0] = wl[0]; locL[1] = wl[2]; locL[2] = -wl[1]; locL[3] = wl[3];
            locR[0] = wr[0]; locR[1] = wr[2]; locR[2] = -wr[1]; locR[3] = wr[3];
        }
        for (int q = 0; q < 4; ++q) sw[q] = (w[q] - locL[q]) / (0.5 * len);
        micro_slope(c, p, sw, aL);
        for (int q = 0; q < 4; ++q) sw[q] = (locR[q] - w[q]) / (0.5 * len);
        micro_slope(c, p, sw, aR);
    }
    // tangential slope b (flux.f90:116-120)
    sw[0] = s1[5]; sw[1] = s1[6]; sw[2] = s1[7]; sw[3] = 0.5 * (s1[8] + s1[9]);
    micro_slope(c, p, sw, ab);
    // time slope (flux.f90:126-135)
    Moments m;
    calc_moment(c, p, m);
    double MauL[4], MauR[4], Mbu[4], MauT[4], Mau0[4];
    moment_au(aL, m.MuL, m.Mv, m.Mxi, 1, 0, MauL);
    moment_au(aR, m.MuR, m.Mv, m.Mxi, 1, 0, MauR);
    moment_au(ab, m.Mu, m.Mv, m.Mxi, 0, 1, Mbu);
    for (int q = 0; q < 4; ++q) sw[q] = -p[0] * (MauL[q] + MauR[q] + Mbu[q]);
    sw[1] = sw[1] + 2.0 * s2[4];
    sw[2] = sw[2] + 2.0 * s2[5];
    sw[3] = sw[3] + s2[6];
    micro_slope(c, p, sw, aT);
    // collision time and time-integration coefficients (flux.f90:140-147)
    const double tau = tau_of(c, p[0], p[3]);
    const double ex = exp(-dt / tau);
    double Mt[7];
    Mt[4] = tau * (1.0 - ex);
    Mt[5] = -tau * dt * ex + tau * Mt[4];
    Mt[1] = dt - Mt[4];
    Mt[2] = -tau * Mt[1] + Mt[5];
    Mt[3] = dt * dt / 2.0 - tau * Mt[1];
    Mt[6] = tau * tau * tau - 0.5 * ex * tau * (dt * dt + 2.0 * dt * tau + 2.0 * tau * tau);

    if (tid == 0) {
        double fl[4];
        // g0 part (flux.f90:152-172)
        moment_uv(m.Mu, m.Mv, m.Mxi, 1, 0, 0, Mau0);
        moment_au(aL, m.MuL, m.Mv, m.Mxi, 2, 0, MauL);
        moment_au(aR, m.MuR, m.Mv, m.Mxi, 2, 0, MauR);
        moment_au(ab, m.Mu, m.Mv, m.Mxi, 1, 1, Mbu);
        moment_au(aT, m.Mu, m.Mv, m.Mxi, 1, 0, MauT);
        for (int q = 0; q < 4; ++q)
            fl[q] = Mt[1] * p[0] * Mau0[q] + Mt[2] * p[0] * (MauL[q] + MauR[q] + Mbu[q]) +
                    Mt[3] * p[0] * MauT[q];
        moment_au(aL, m.MuL, m.Mv, m.Mxi, 1, 0, MauL);
        moment_au(aR, m.MuR, m.Mv, m.Mxi, 1, 0, MauR);
        for (int q = 0; q < 4; ++q)
            fl[q] = fl[q] - tau * Mt[2] * p[0] * fn * (MauL[q] + MauR[q]) -
                    Mt[6] * p[0] * fn * (MauL[q] + MauR[q]);
        moment_au(ab, m.Mu, m.Mv, m.Mxi, 0, 1, Mbu);
        for (int q = 0; q < 4; ++q)
            fl[q] = fl[q] - tau * Mt[2] * p[0] * ft * Mbu[q] - Mt[6] * p[0] * ft * Mbu[q];
        // g+ and f0 parts (flux.f90:184-202); Shakhov moments from their
        // heat-flux split
        const double cs = c.shk * (1.0 - c.prandtl) * p[3] * p[3] / p[0];
        const double X0 = cs * (qf0 * s2[7] + qf1 * s2[8]);
        const double X1 = cs * (qf0 * s2[9] + qf1 * s2[10]);
        const double X2 = cs * (qf0 * s2[11] + qf1 * s2[12]);
        const double X3 = cs * (qf0 * s2[13] + qf1 * s2[14]);
        const double X4 = cs * (qf0 * s2[15] + qf1 * s2[16]);
        fl[0] = fl[0] + Mt[1] * X0 + Mt[4] * s1[1] - Mt[5] * s1[10] - Mt[5] * s1[11];
        fl[1] = fl[1] + Mt[1] * X1 + Mt[4] * s1[12] - Mt[5] * s1[13] - Mt[5] * s1[14];
        fl[2] = fl[2] + Mt[1] * X2 + Mt[4] * s1[15] - Mt[5] * s1[16] - Mt[5] * s1[17];
        fl[3] = fl[3] + Mt[1] * 0.5 * (X3 + X4) + Mt[4] * 0.5 * (s1[18] + s1[19]) -
                Mt[5] * 0.5 * (s1[20] + s1[21]) - Mt[5] * 0.5 * (s1[22] + s1[23]);
        fl[0] = fl[0] + fn * Mt[6] * s1[24];
        fl[1] = fl[1] + fn * Mt[6] * s1[10];
        fl[2] = fl[2] + fn * Mt[6] * s1[25];
        fl[3] = fl[3] + fn * Mt[6] * 0.5 * (s1[26] + s1[27]);
        fl[0] = fl[0] + ft * Mt[6] * s1[6];
        fl[1] = fl[1] + ft * Mt[6] * s1[28];
        fl[2] = fl[2] + ft * Mt[6] * s1[11];
        fl[3] = fl[3] + ft * Mt[6] * 0.5 * (s1[29] + s1[30]);
        // global frame (flux.f90:242) and face length
        double g1 = fl[1], g2 = fl[2];
        if (DIR == 1) { g1 = -fl[2]; g2 = fl[1]; }
        double* out = flux + 4 * (size_t)face;
        out[0] = flen * fl[0];
        out[1] = flen * g1;
        out[2] = flen * g2;
        out[3] = flen * fl[3];
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
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = tid + e * NT;
        if (k < c.nn) {
            const A u = unode[k], v = vnode[k];
            const A vn = DIR == 0 ? u : v;
            const A vt = DIR == 0 ? v : -u;
            const A cu = vn - U, cv = vt - V;
            const A cp2 = cu * cu + cv * cv;
            const A c2 = vn * vn + vt * vt;
            const A H = rho_l_pi * fexp(-lam * cp2);
            const A Bm = H * kfac;
            const A cq = csA * (cu * q0 + cv * q1);
            const A lam2 = (A)2 * lam * cp2;
            const A Hp = cq * (lam2 + (A)c.K - (A)5) * H;
            const A Bp = cq * (lam2 + (A)c.K - (A)3) * Bm;
            const A eh = c2 * H + Bm, eb = c2 * Bm + mxi2 * H;
            const A gLh = a_L[0] * H + a_L[1] * vn * H + a_L[2] * vt * H + (A)0.5 * a_L[3] * eh;
            const A gRh = a_R[0] * H + a_R[1] * vn * H + a_R[2] * vt * H + (A)0.5 * a_R[3] * eh;
            const A gbh = a_b[0] * H + a_b[1] * vn * H + a_b[2] * vt * H + (A)0.5 * a_b[3] * eh;
            const A gTh = a_T[0] * H + a_T[1] * vn * H + a_T[2] * vt * H + (A)0.5 * a_T[3] * eh;
            const A gLb = a_L[0] * Bm + a_L[1] * vn * Bm + a_L[2] * vt * Bm + (A)0.5 * a_L[3] * eb;
            const A gRb = a_R[0] * Bm + a_R[1] * vn * Bm + a_R[2] * vt * Bm + (A)0.5 * a_R[3] * eb;
            const A gbb = a_b[0] * Bm + a_b[1] * vn * Bm + a_b[2] * vt * Bm + (A)0.5 * a_b[3] * eb;
            const A gTb = a_T[0] * Bm + a_T[1] * vn * Bm + a_T[2] * vt * Bm + (A)0.5 * a_T[3] * eb;
            const A gh = dl[e] ? gLh : gRh;
            const A gb = dl[e] ? gLb : gRb;
            A xh = mt1 * vn * (H + Hp) + mt2 * vn * vn * gh + mt2 * vn * vt * gbh + mt3 * vn * gTh +
                   mt4 * vn * rh[e] - mt5 * vn * vn * rsn[e] - mt5 * vn * vt * rst[e];
            A xb = mt1 * vn * (Bm + Bp) + mt2 * vn * vn * gb + mt2 * vn * vt * gbb + mt3 * vn * gTb +
                   mt4 * vn * rb[e] - mt5 * vn * vn * rbn[e] - mt5 * vn * vt * rbt[e];
            xh = xh - tm2fn * vn * gh + mt6 * fnA * vn * rsn[e];
            xb = xb - tm2fn * vn * gb + mt6 * fnA * vn * rbn[e];
            xh = xh - tm2ft * vn * gbh + mt6 * ftA * vn * rst[e];
            xb = xb - tm2ft * vn * gbb + mt6 * ftA * vn * rbt[e];
            oh[k] = (T)(flenA * xh);
            ob[k] = (T)(flenA * xb);
        }
    }
}

// -------------------------------------------------------------- wall faces
// One workgroup per wall face.  blockIdx enumerates the faces of the
// non-periodic sides in the order x-lo (ny), x-hi (ny), y-lo (nx),
// y-hi (nx), skipping periodic sides.  Diffuse isothermal wall:
// calc_flux_boundary, 2d/src/flux.f90:474-549; specular wall:
// calc_flux_specular_vertical / _horizon, 2d/src/flux.f90:551-689.
template <typename T, int EPT, int NT>
__global__ __launch_bounds__(NT) void k_flux_wall(
    Consts c, DevState* __restrict__ st, const T* __restrict__ h, const T* __restrict__ b,
    const T* __restrict__ shx, const T* __restrict__ shy, const T* __restrict__ sbx,
    const T* __restrict__ sby, const typename Arith<T>::A* __restrict__ unode,
    const typename Arith<T>::A* __restrict__ vnode, const typename Arith<T>::A* __restrict__ wnode,
    T* __restrict__ vfh, T* __restrict__ vfb, double* __restrict__ vflux, T* __restrict__ hfh,
    T* __restrict__ hfb, double* __restrict__ hflux) {
    using A = typename Arith<T>::A;
    if (st->stopped) return;
    __shared__ double lds[4 * (NT / WAVE)];
    const int tid = threadIdx.x;
    int f = blockIdx.x, side = -1, idx = 0;
    for (int s = 0; s < 4; ++s) {
        if (c.bc[s] == HUGKS_BC_PERIODIC) continue;
        const int n = s < 2 ? c.ny : c.nx;
        if (f < n) { side = s; idx = f; break; }
        f -= n;
    }
    if (side < 0) return;
    const int dir = side / 2;           // 0: vertical face, 1: horizontal face
    const int rot = (side & 1) ? -1 : 1;
    int cell, face;
    if (dir == 0) {
        const int i = (side & 1) ? c.nx - 1 : 0;
        cell = i + c.nx * idx;
        face = ((side & 1) ? c.nx : 0) + (c.nx + 1) * idx;
    } else {
        const int j = (side & 1) ? c.ny - 1 : 0;
        cell = idx + c.nx * j;
        face = idx + c.nx * ((side & 1) ? c.ny : 0);
    }
    const T* sN = dir == 0 ? shx : shy;
    const T* bN = dir == 0 ? sbx : sby;
    const double dt = st->dt;
    const double len = dir == 0 ? c.dx : c.dy;
    const double flen = dir == 0 ? c.dy : c.dx;
    const A hrl = (A)rot * (A)0.5 * (A)len;
    const bool diffuse = c.bc[side] == HUGKS_BC_DIFFUSE;
    const int nv = c.nv;

    A fh_[EPT], fb_[EPT];
    int dl[EPT];
    double r[2] = {0.0, 0.0};
    const double lamw = 1.0 / c.tw[side];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = tid + e * NT;
        fh_[e] = fb_[e] = (A)0;
        dl[e] = 0;
        if (k < c.nn) {
            const A u = unode[k], v = vnode[k], wt = wnode[k];
            const A vn = dir == 0 ? u : v;
            const A vt = dir == 0 ? v : -u;
            const int d = (copysign((A)1, vn) * (A)rot) > (A)0 ? 1 : 0;
            dl[e] = d;
            const size_t o = (size_t)cell * c.ldv + k;
            const A hin = (A)h[o] - hrl * (A)sN[o];
            const A bin = (A)b[o] - hrl * (A)bN[o];
            if (diffuse) {
                fh_[e] = hin;
                fb_[e] = bin;
                if (!d) r[0] += (double)(wt * vn * hin);
                else r[1] += (double)(wt * vn) * exp(-lamw * ((double)vn * vn + (double)vt * vt));
            } else if (d) {
                // mirrored node: u index flipped on vertical walls, v on horizontal
                const int iu = k % nv, jv = k / nv;
                const int km = dir == 0 ? (nv - 1 - iu) + nv * jv : iu + nv * (nv - 1 - jv);
                const size_t om = (size_t)cell * c.ldv + km;
                fh_[e] = (A)h[om] - hrl * (A)sN[om];
                fb_[e] = (A)b[om] - hrl * (A)bN[om];
            } else {
                fh_[e] = hin;
                fb_[e] = bin;
            }
        }
    }
    double rhow = 0.0;
    if (diffuse) {
        block_sum<2, NT>(r, lds);
        const double R2 = (lamw / c.pi) * r[1];
        if (R2 == 0.0) {
            if (tid == 0) atomicOr(&st->flags, 2);
        }
        rhow = -r[0] / R2;
    }
    const A Hc = (A)(rhow * (lamw / c.pi)), kf = (A)(c.K / (2.0 * lamw));
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    T* oh = (dir == 0 ? vfh : hfh) + (size_t)face * c.ldv;
    T* ob = (dir == 0 ? vfb : hfb) + (size_t)face * c.ldv;
    const A dtl = (A)(dt * flen);
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = tid + e * NT;
        if (k < c.nn) {
            const A u = unode[k], v = vnode[k], wt = wnode[k];
            const A vn = dir == 0 ? u : v;
            const A vt = dir == 0 ? v : -u;
            A hf = fh_[e], bf = fb_[e];
            if (diffuse && dl[e]) {
                const A H0 = Hc * fexp(-(A)lamw * (vn * vn + vt * vt));
                hf = H0;
                bf = H0 * kf;
            }
            const A wvh = wt * vn * hf;
            s[0] += (double)wvh;
            s[1] += (double)(vn * wvh);
            s[2] += (double)(vt * wvh);
            s[3] += (double)(wt * vn * ((vn * vn + vt * vt) * hf + bf));
            oh[k] = (T)(dtl * (vn * hf));
            ob[k] = (T)(dtl * (vn * bf));
        }
    }
    block_sum<4, NT>(s, lds);
    if (tid == 0) {
        double g1 = s[1], g2 = s[2];
        if (dir == 1) { g1 = -s[2]; g2 = s[1]; }
        double* out = (dir == 0 ? vflux : hflux) + 4 * (size_t)face;
        const double sc = dt * flen;
        out[0] = sc * s[0];
        out[1] = sc * g1;
        out[2] = sc * g2;
        out[3] = sc * (0.5 * s[3]);
    }
}

// ------------------------------------------------------------ launchers
template <typename T>
static hipError_t launch_flux2d_t(const Launch2D& L, hipStream_t s) {
    const Consts& c = L.c;
    const int NT = 256;
    const int ept = (c.ldv + NT - 1) / NT;
    const size_t ncell = (size_t)c.nx * c.ny;
    {
        const size_t n = ncell * c.ldv;
        k_slope2d<T><<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(
            c, L.st, (const T*)L.h, (const T*)L.b, (T*)L.shx, (T*)L.shy, (T*)L.sbx, (T*)L.sby);
    }
    const bool xper = c.bc[0] == HUGKS_BC_PERIODIC, yper = c.bc[2] == HUGKS_BC_PERIODIC;
    const unsigned nvf = (unsigned)((xper ? c.nx + 1 : c.nx - 1) * c.ny);
    const unsigned nhf = (unsigned)(c.nx * (yper ? c.ny + 1 : c.ny - 1));
    using A = typename Arith<T>::A;
    const A* un = (const A*)L.unode;
    const A* vn = (const A*)L.vnode;
    const A* wn = (const A*)L.wnode;
#define HUGKS_FLUX_CASE(E)                                                                     \
    case E:                                                                                    \
        if (nvf)                                                                               \
            k_flux_interior<0, T, E, 256><<<nvf, 256, 0, s>>>(                                 \
                c, L.st, L.W, (const T*)L.h, (const T*)L.b, (const T*)L.shx, (const T*)L.shy,  \
                (const T*)L.sbx, (const T*)L.sby, un, vn, wn, L.acc_v, (T*)L.vfh, (T*)L.vfb,   \
                L.vflux);                                                                      \
        if (nhf)                                                                               \
            k_flux_interior<1, T, E, 256><<<nhf, 256, 0, s>>>(                                 \
                c, L.st, L.W, (const T*)L.h, (const T*)L.b, (const T*)L.shx, (const T*)L.shy,  \
                (const T*)L.sbx, (const T*)L.sby, un, vn, wn, L.acc_h, (T*)L.hfh, (T*)L.hfb,   \
                L.hflux);                                                                      \
        if (L.nwall)                                                                           \
            k_flux_wall<T, E, 256><<<L.nwall, 256, 0, s>>>(                                    \
                c, L.st, (const T*)L.h, (const T*)L.b, (const T*)L.shx, (const T*)L.shy,       \
                (const T*)L.sbx, (const T*)L.sby, un, vn, wn, (T*)L.vfh, (T*)L.vfb, L.vflux,   \
                (T*)L.hfh, (T*)L.hfb, L.hflux);                                                \
        break;
    switch (ept) {
        HUGKS_FLUX_CASE(1)
        HUGKS_FLUX_CASE(2)
        HUGKS_FLUX_CASE(3)
        HUGKS_FLUX_CASE(4)
        HUGKS_FLUX_CASE(5)
        HUGKS_FLUX_CASE(6)
        HUGKS_FLUX_CASE(7)
        HUGKS_FLUX_CASE(8)
        default: return hipErrorInvalidValue;
    }
#undef HUGKS_FLUX_CASE
    return hipGetLastError();
}

hipError_t launch_flux2d(const Launch2D& L, hipStream_t s) {
    return L.fp64 ? launch_flux2d_t<double>(L, s) : launch_flux2d_t<float>(L, s);
}

}  // namespace hugks
