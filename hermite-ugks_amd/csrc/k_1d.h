// 1D Hermite-UGKS step kernels (gfx950): slopes, interior and wall face
// fluxes, cell update with the 1D Hermite force.  One wavefront per face
// or cell (1D velocity grids are short: nv <= 64*EPT).
// References: 1d/src/solver.f90:16-264, 1d/src/flux.f90:13-360,
// 1d/src/gravity.f90:18-123.
#pragma once
#include "hugks_device.h"
#include "hugks_kernels.h"

namespace hugks {

template <int NV>
__device__ __forceinline__ void wave_sum(double (&v)[NV]) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double x = v[i];
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
        v[i] = __shfl(x, 0, WAVE);
    }
}

// micro_slope 1D, 1d/src/flux.f90:362-370
__device__ __forceinline__ void micro_slope1(double K, const double* p, const double* sw, double* a) {
    a[2] = 4.0 * p[2] * p[2] / (K + 1.0) / p[0] *
           (2.0 * sw[2] - 2.0 * p[1] * sw[1] + sw[0] * (p[1] * p[1] - 0.5 * (K + 1.0) / p[2]));
    a[1] = 2.0 * p[2] / p[0] * (sw[1] - p[1] * sw[0]) - p[1] * a[2];
    a[0] = sw[0] / p[0] - p[1] * a[1] - 0.5 * (p[1] * p[1] + 0.5 * (K + 1.0) / p[2]) * a[2];
}
__device__ __forceinline__ void prim1(const Consts& c, const double* w, double* p) {
    p[0] = w[0];
    p[1] = w[1] / w[0];
    p[2] = 0.5 * w[0] / (c.gamma - 1.0) / (w[2] - 0.5 * w[1] * w[1] / w[0]);
}
__device__ __forceinline__ void cons1(const Consts& c, const double* p, double* w) {
    w[0] = p[0];
    w[1] = p[0] * p[1];
    w[2] = 0.5 * p[0] / p[2] / (c.gamma - 1.0) + 0.5 * p[0] * p[1] * p[1];
}
// moment_uv / moment_au 1D, 1d/src/flux.f90:397-417
__device__ __forceinline__ void uv1(const double* Mu, const double* Mxi, int al, int de, double* m) {
    m[0] = Mu[al] * Mxi[de / 2];
    m[1] = Mu[al + 1] * Mxi[de / 2];
    m[2] = 0.5 * (Mu[al + 2] * Mxi[de / 2] + Mu[al] * Mxi[(de + 2) / 2]);
}
__device__ __forceinline__ void au1(const double* a, const double* Mu, const double* Mxi, int al,
                                    double* m) {
    double t1[3], t2[3], t3[3], t4[3];
    uv1(Mu, Mxi, al, 0, t1);
    uv1(Mu, Mxi, al + 1, 0, t2);
    uv1(Mu, Mxi, al + 2, 0, t3);
    uv1(Mu, Mxi, al, 2, t4);
    for (int q = 0; q < 3; ++q) m[q] = a[0] * t1[q] + a[1] * t2[q] + 0.5 * a[2] * t3[q] + 0.5 * a[2] * t4[q];
}

// interpolation, 1d/src/solver.f90:49-104
template <typename T>
__global__ void k_slope1d(Consts c, const DevState* __restrict__ st, const T* __restrict__ h,
                          const T* __restrict__ b, T* __restrict__ sh, T* __restrict__ sb) {
    using A = typename Arith<T>::A;
    if (st->stopped) return;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)c.nx * c.ldv) return;
    const int cell = (int)(idx / c.ldv), k = (int)(idx - (size_t)cell * c.ldv);
    const A den = (A)0.5 * (A)c.dx + (A)0.5 * (A)c.dx;
    // tiny(1d0) underflows in FP32; the smallest normal float keeps 0/0 out
    const A smv = sizeof(A) == 8 ? (A)c.smv : (A)1.17549435e-38f;
    for (int q = 0; q < 2; ++q) {
        const T* f = q ? b : h;
        T* s = q ? sb : sh;
        A v;
        if (cell == 0) {
            v = ((A)f[(size_t)c.ldv + k] - (A)f[k]) / den;
        } else if (cell == c.nx - 1) {
            v = ((A)f[(size_t)cell * c.ldv + k] - (A)f[(size_t)(cell - 1) * c.ldv + k]) / den;
        } else {
            const A fl = f[(size_t)(cell - 1) * c.ldv + k], fn = f[(size_t)cell * c.ldv + k],
                    fr = f[(size_t)(cell + 1) * c.ldv + k];
            const A sL = (fn - fl) / den, sR = (fr - fn) / den;
            v = (copysign((A)1, sR) + copysign((A)1, sL)) * fabs(sR) * fabs(sL) /
                (fabs(sR) + fabs(sL) + smv);
        }
        s[(size_t)cell * c.ldv + k] = (T)v;
    }
}

// calc_flux, 1d/src/flux.f90:13-186; one wavefront per interior face
template <typename T, int EPT>
__global__ __launch_bounds__(64) void k_flux1d(Consts c, const DevState* __restrict__ st,
                                                const double* __restrict__ W, const T* __restrict__ h,
                                                const T* __restrict__ b, const T* __restrict__ sh,
                                                const T* __restrict__ sb,
                                                const typename Arith<T>::A* __restrict__ unode,
                                                const typename Arith<T>::A* __restrict__ wnode,
                                                T* __restrict__ fh, T* __restrict__ fb,
                                                double* __restrict__ flux) {
    using A = typename Arith<T>::A;
    if (st->stopped) return;
    const int face = blockIdx.x + 1;        // faces 1..nx-1 between cells face-1, face
    const int cL = face - 1, cR = face;
    const int lane = threadIdx.x;
    const double dt = st->dt, phi = c.phi;
    const A hl = (A)(0.5 * c.dx);
    const A vsh = (A)(0.5 * phi * dt);
    A rh[EPT], rb[EPT], rs[EPT], rsb[EPT];
    int dl[EPT];
    double s1[12];
    for (int q = 0; q < 12; ++q) s1[q] = 0.0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = lane + 64 * e;
        rh[e] = rb[e] = rs[e] = rsb[e] = (A)0;
        dl[e] = 0;
        if (k < c.nn) {
            const A u = unode[k], wt = wnode[k];
            const int d = upwind_delta(u - vsh);
            dl[e] = d;
            const size_t o = (size_t)(d ? cL : cR) * c.ldv + k;
            rs[e] = (A)sh[o];
            rsb[e] = (A)sb[o];
            rh[e] = d ? (A)h[o] + hl * rs[e] : (A)h[o] - hl * rs[e];
            rb[e] = d ? (A)b[o] + hl * rsb[e] : (A)b[o] - hl * rsb[e];
            const A u2 = u * u, wh = wt * rh[e], ws = wt * rs[e];
            s1[0] += (double)wh;
            s1[1] += (double)(u * wh);
            s1[2] += (double)(u2 * wh);
            s1[3] += (double)(wt * rb[e]);
            s1[4] += (double)(u2 * ws);
            s1[5] += (double)(u * u2 * ws);
            s1[6] += (double)(u * u2 * wh);
            s1[7] += (double)(u * wt * rb[e]);
            s1[8] += (double)(u2 * u2 * ws);
            s1[9] += (double)(u2 * wt * rsb[e]);
            s1[10] += (double)(u * ws);
            s1[11] += (double)(u * wt * rsb[e]);
        }
    }
    wave_sum<12>(s1);
    double w[3] = {s1[0], s1[1], 0.5 * (s1[2] + s1[3])};
    double p[3];
    prim1(c, w, p);
    const A U = (A)p[1], lam = (A)p[2];
    const A Hc = (A)(p[0] * pow(p[2] / c.pi, 0.5)), kf = (A)(c.K / (2.0 * p[2]));
    double s2[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = lane + 64 * e;
        if (k < c.nn) {
            const A u = unode[k], wt = wnode[k], cu = u - U;
            const A H0 = Hc * fexp(-lam * cu * cu), B0 = H0 * kf;
            const A l2 = (A)2 * lam * cu * cu;
            const A P5 = (l2 + (A)c.K - (A)5) * H0, P3 = (l2 + (A)c.K - (A)3) * B0;
            s2[0] += (double)(wt * cu * (cu * cu) * rh[e]);
            s2[1] += (double)(wt * cu * rb[e]);
            s2[2] += (double)(wt * u * cu * P5);
            s2[3] += (double)(wt * u * u * cu * P5);
            s2[4] += (double)(wt * u * u * u * cu * P5);
            s2[5] += (double)(wt * u * cu * P3);
        }
    }
    wave_sum<6>(s2);
    const double qf = 0.5 * (s2[0] + s2[1]);
    double sw[3], aL[3], aR[3], aT[3];
    const double* wl = W + 3 * (size_t)cL;
    const double* wr = W + 3 * (size_t)cR;
    for (int q = 0; q < 3; ++q) sw[q] = (w[q] - wl[q]) / (0.5 * c.dx);
    micro_slope1(c.K, p, sw, aL);
    for (int q = 0; q < 3; ++q) sw[q] = (wr[q] - w[q]) / (0.5 * c.dx);
    micro_slope1(c.K, p, sw, aR);
    // calc_moment_u, 1d/src/flux.f90:372-395
    double Mu[7], MuL[7], MuR[7], Mxi[3];
    {
        const double Uq = p[1], lq = p[2];
        const double ee = 0.5 * exp(-lq * Uq * Uq) / sqrt(c.pi * lq);
        MuL[0] = 0.5 * erfc(-sqrt(lq) * Uq);
        MuL[1] = Uq * MuL[0] + ee;
        MuR[0] = 0.5 * erfc(sqrt(lq) * Uq);
        MuR[1] = Uq * MuR[0] - ee;
        for (int q = 2; q <= 6; ++q) {
            MuL[q] = Uq * MuL[q - 1] + 0.5 * (q - 1) * MuL[q - 2] / lq;
            MuR[q] = Uq * MuR[q - 1] + 0.5 * (q - 1) * MuR[q - 2] / lq;
        }
        for (int q = 0; q <= 6; ++q) Mu[q] = MuL[q] + MuR[q];
        Mxi[0] = 1.0;
        Mxi[1] = 0.5 * c.K / lq;
        Mxi[2] = (c.K * c.K + 2.0 * c.K) / (4.0 * lq * lq);
    }
    double MauL[3], MauR[3], MauT[3], Mau0[3];
    au1(aL, MuL, Mxi, 1, MauL);
    au1(aR, MuR, Mxi, 1, MauR);
    for (int q = 0; q < 3; ++q) sw[q] = -p[0] * (MauL[q] + MauR[q]);
    micro_slope1(c.K, p, sw, aT);
    const double tau = c.muref * 2.0 * pow(p[2], 1.0 - c.omega) / p[0];
    const double ex = exp(-dt / tau);
    double Mt[7];
    Mt[4] = tau * (1.0 - ex);
    Mt[5] = -tau * dt * ex + tau * Mt[4];
    Mt[1] = dt - Mt[4];
    Mt[2] = -tau * Mt[1] + Mt[5];
    Mt[3] = dt * dt / 2.0 - tau * Mt[1];
    Mt[6] = tau * tau * tau - 0.5 * ex * tau * (dt * dt + 2.0 * dt * tau + 2.0 * tau * tau);
    if (lane == 0) {
        double fl[3];
        uv1(Mu, Mxi, 1, 0, Mau0);
        au1(aL, MuL, Mxi, 2, MauL);
        au1(aR, MuR, Mxi, 2, MauR);
        au1(aT, Mu, Mxi, 1, MauT);
        for (int q = 0; q < 3; ++q)
            fl[q] = Mt[1] * p[0] * Mau0[q] + Mt[2] * p[0] * (MauL[q] + MauR[q]) + Mt[3] * p[0] * MauT[q];
        au1(aL, MuL, Mxi, 1, MauL);
        au1(aR, MuR, Mxi, 1, MauR);
        for (int q = 0; q < 3; ++q)
            fl[q] = fl[q] - tau * Mt[2] * p[0] * phi * (MauL[q] + MauR[q]) -
                    Mt[6] * p[0] * phi * (MauL[q] + MauR[q]);
        const double cs = c.shk * (1.0 - c.prandtl) * p[2] * p[2] / p[0] * qf;
        const double X0 = cs * s2[2], X3 = cs * s2[3], X6 = cs * s2[4], X7 = cs * s2[5];
        fl[0] = fl[0] + Mt[1] * X0 + Mt[4] * s1[1] - Mt[5] * s1[4];
        fl[1] = fl[1] + Mt[1] * X3 + Mt[4] * s1[2] - Mt[5] * s1[5];
        fl[2] = fl[2] + Mt[1] * 0.5 * (X6 + X7) + Mt[4] * 0.5 * (s1[6] + s1[7]) -
                Mt[5] * 0.5 * (s1[8] + s1[9]);
        fl[0] = fl[0] + phi * Mt[6] * s1[10];
        fl[1] = fl[1] + phi * Mt[6] * s1[4];
        fl[2] = fl[2] + phi * Mt[6] * 0.5 * (s1[5] + s1[11]);
        for (int q = 0; q < 3; ++q) flux[4 * (size_t)face + q] = fl[q];
    }
    const A mt1 = (A)Mt[1], mt2 = (A)Mt[2], mt3 = (A)Mt[3], mt4 = (A)Mt[4], mt5 = (A)Mt[5],
            mt6 = (A)Mt[6];
    const A tm2p = (A)(tau * Mt[2] * phi), phA = (A)phi, mxi2 = (A)Mxi[2];
    const A csq = (A)(c.shk * (1.0 - c.prandtl) * p[2] * p[2] / p[0] * qf);
    A a_L[3], a_R[3], a_T[3];
    for (int q = 0; q < 3; ++q) { a_L[q] = (A)aL[q]; a_R[q] = (A)aR[q]; a_T[q] = (A)aT[q]; }
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = lane + 64 * e;
        if (k < c.nn) {
            const A u = unode[k], u2 = u * u, cu = u - U;
            const A H = Hc * fexp(-lam * cu * cu), Bm = H * kf;
            const A l2 = (A)2 * lam * cu * cu;
            const A Hp = csq * cu * (l2 + (A)c.K - (A)5) * H;
            const A Bp = csq * cu * (l2 + (A)c.K - (A)3) * Bm;
            const A eh = u2 * H + Bm, eb = u2 * Bm + mxi2 * H;
            const A gLh = a_L[0] * H + a_L[1] * u * H + (A)0.5 * a_L[2] * eh;
            const A gRh = a_R[0] * H + a_R[1] * u * H + (A)0.5 * a_R[2] * eh;
            const A gTh = a_T[0] * H + a_T[1] * u * H + (A)0.5 * a_T[2] * eh;
            const A gLb = a_L[0] * Bm + a_L[1] * u * Bm + (A)0.5 * a_L[2] * eb;
            const A gRb = a_R[0] * Bm + a_R[1] * u * Bm + (A)0.5 * a_R[2] * eb;
            const A gTb = a_T[0] * Bm + a_T[1] * u * Bm + (A)0.5 * a_T[2] * eb;
            const A gh = dl[e] ? gLh : gRh, gb = dl[e] ? gLb : gRb;
            A xh = mt1 * u * (H + Hp) + mt2 * u2 * gh + mt3 * u * gTh + mt4 * u * rh[e] - mt5 * u2 * rs[e];
            A xb = mt1 * u * (Bm + Bp) + mt2 * u2 * gb + mt3 * u * gTb + mt4 * u * rb[e] - mt5 * u2 * rsb[e];
            xh = xh - tm2p * u * gh + mt6 * phA * u * rs[e];
            xb = xb - tm2p * u * gb + mt6 * phA * u * rsb[e];
            fh[(size_t)face * c.ldv + k] = (T)xh;
            fb[(size_t)face * c.ldv + k] = (T)xb;
        }
    }
}

// calc_flux_specular (1d/src/flux.f90:188-239) / calc_flux_boundary
// (1d/src/flux.f90:299-360); block 0: left wall, block 1: right wall
template <typename T, int EPT>
__global__ __launch_bounds__(64) void k_wall1d(Consts c, DevState* __restrict__ st,
                                               const T* __restrict__ h, const T* __restrict__ b,
                                               const T* __restrict__ sh, const T* __restrict__ sb,
                                               const typename Arith<T>::A* __restrict__ unode,
                                               const typename Arith<T>::A* __restrict__ wnode,
                                               T* __restrict__ fh, T* __restrict__ fb,
                                               double* __restrict__ flux) {
    using A = typename Arith<T>::A;
    if (st->stopped) return;
    const int side = blockIdx.x;
    const int rot = side ? -1 : 1;
    const int cell = side ? c.nx - 1 : 0, face = side ? c.nx : 0;
    const int lane = threadIdx.x;
    const double dt = st->dt;
    const bool diffuse = bc_of(c, side) == HUGKS_BC_DIFFUSE;
    const double lamw = 1.0 / tw_of(c, side);
    const A hr = (A)rot * (A)0.5 * (A)c.dx;
    A fh_[EPT], fb_[EPT];
    int dl[EPT];
    double r[2] = {0, 0};
    const double Hud = pow(lamw / c.pi, 0.5), kfd = c.K / (2.0 * lamw);
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = lane + 64 * e;
        fh_[e] = fb_[e] = (A)0;
        dl[e] = 0;
        if (k < c.nn) {
            const A u = unode[k], wt = wnode[k];
            const int d = copysign((A)1, u) * (A)rot > (A)0 ? 1 : 0;
            dl[e] = d;
            const int ks = (!diffuse && d) ? c.nn - 1 - k : k;
            const size_t o = (size_t)cell * c.ldv + ks;
            fh_[e] = (A)h[o] - hr * (A)sh[o];
            fb_[e] = (A)b[o] - hr * (A)sb[o];
            if (diffuse) {
                if (!d) r[0] += (double)wt * (double)u * (double)fh_[e];
                else r[1] += (double)wt * (double)u * Hud * exp(-lamw * (double)u * (double)u);
            }
        }
    }
    double rhow = 0.0;
    if (diffuse) {
        wave_sum<2>(r);
        if (fabs(r[1]) <= 2.2250738585072014e-308) {
            if (lane == 0) atomicOr(&st->flags, 2);
        }
        rhow = -r[0] / r[1];
    }
    const double Hcd = rhow * pow(lamw / c.pi, 0.5);
    double s[3] = {0, 0, 0};
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = lane + 64 * e;
        if (k < c.nn) {
            const A u = unode[k], wt = wnode[k];
            A hf = fh_[e], bf = fb_[e];
            // conserved-flux sums in FP64 from FP64 per-node values (the
            // re-emitted mass cancels the incoming mass to FP64 round-off on
            // the FP32-storage path too); stored fluxes in the storage type
            const double ud = (double)u, wtd = (double)wt;
            double hfd = (double)hf, bfd = (double)bf;
            if (diffuse && dl[e]) {
                hfd = Hcd * exp(-lamw * ud * ud);
                bfd = hfd * kfd;
                hf = (A)hfd;
                bf = (A)bfd;
            }
            s[0] += wtd * ud * hfd;
            s[1] += wtd * ud * ud * hfd;
            s[2] += wtd * ud * ((ud * ud) * hfd + bfd);
            fh[(size_t)face * c.ldv + k] = (T)((A)dt * (u * hf));
            fb[(size_t)face * c.ldv + k] = (T)((A)dt * (u * bf));
        }
    }
    wave_sum<3>(s);
    if (lane == 0) {
        flux[4 * (size_t)face + 0] = dt * s[0];
        flux[4 * (size_t)face + 1] = dt * s[1];
        flux[4 * (size_t)face + 2] = dt * (0.5 * s[2]);
    }
}

// cal_fv1, 1d/src/gravity.f90:127-154: the velocity difference of f (nodes
// f[0..n-1], step du) at node i, one-sided second order at both ends,
// central next to them, fourth-order central inside, in the reference's
// expression order
template <typename A>
__device__ __forceinline__ double fd1(const A* f, int n, int i, double du) {
    auto F = [&](int q) { return (double)f[q]; };
    if (i == 0) return (-3.0 * F(0) + 4.0 * F(1) - F(2)) / 2.0 / du;
    if (i == n - 1) return (3.0 * F(n - 1) - 4.0 * F(n - 2) + F(n - 3)) / 2.0 / du;
    if (i == 1 || i == n - 2) return (F(i + 1) - F(i - 1)) / 2.0 / du;
    return (-F(i + 2) + 8.0 * F(i + 1) - 8.0 * F(i - 1) + F(i - 2)) / 12.0 / du;
}

// update, 1d/src/solver.f90:141-264, with cal_fv (1d/src/gravity.f90:18-123)
// or, for the finite-difference force, cal_fv1
template <typename T, int EPT>
__global__ __launch_bounds__(64) void k_update1d(Consts c, DevState* __restrict__ st,
                                                 double* __restrict__ W, double* __restrict__ part,
                                                 T* __restrict__ h, T* __restrict__ b,
                                                 const T* __restrict__ fh, const T* __restrict__ fb,
                                                 const double* __restrict__ flux,
                                                 const typename Arith<T>::A* __restrict__ unode,
                                                 const typename Arith<T>::A* __restrict__ wnode,
                                                 const double* __restrict__ unode_d,
                                                 const double* __restrict__ wnode_d) {
    using A = typename Arith<T>::A;
    if (st->stopped) return;
    const int cell = blockIdx.x, lane = threadIdx.x;
    const double dt = st->dt, phi = c.phi;
    double wo[3], po[3], wn[3], pn[3];
    for (int q = 0; q < 3; ++q) wo[q] = W[3 * (size_t)cell + q];
    prim1(c, wo, po);
    for (int q = 0; q < 3; ++q)
        wn[q] = wo[q] + (flux[4 * (size_t)cell + q] - flux[4 * (size_t)(cell + 1) + q]) / c.dx;
    prim1(c, wn, pn);
    pn[1] = pn[1] + phi * dt;
    cons1(c, pn, wn);
    const double tau = c.muref * 2.0 * pow(pn[2], 1.0 - c.omega) / pn[0];
    // heat flux at t^n and the Hermite coefficients of h^n, b^n.  The heat
    // flux, the Maxwellian and the implicit update are FP64 whatever the
    // storage type (an FP32 exponent would put a systematic relative error
    // of order lambda c^2 * 2^-24 into the moments of the next step); the
    // Hermite force uses the storage arithmetic.
    const double t = 1.0 / po[2];
    const A sc = (A)(sqrt(t)), s2 = (A)sqrt(2.0), Uo = (A)po[1];
    const int N = c.nm;
    A rh[EPT], rb[EPT], vv[EPT];
    double q2[2] = {0, 0};
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = lane + 64 * e;
        rh[e] = rb[e] = vv[e] = (A)0;
        if (k < c.nn) {
            const size_t o = (size_t)cell * c.ldv + k;
            rh[e] = (A)h[o];
            rb[e] = (A)b[o];
            const A u = unode[k];
            const double wd = wnode_d[k], cd = unode_d[k] - po[1];
            q2[0] += wd * cd * (cd * cd) * (double)rh[e];
            q2[1] += wd * cd * (double)rb[e];
            vv[e] = (u - Uo) / sc * s2;
        }
    }
    wave_sum<2>(q2);
    const double qf = 0.5 * (q2[0] + q2[1]);
    // the finite-difference force reads the neighbouring nodes of h and b
    // (other lanes) from LDS
    const bool fd = c.force_fd != 0;
    __shared__ A fdf[2 * 64 * EPT];
    if (fd) {
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            fdf[lane + 64 * e] = rh[e];
            fdf[64 * EPT + lane + 64 * e] = rb[e];
        }
        __syncthreads();
    }
    // d_j = sqrt(2/t) sum_i h_i w_i He_j(v_i); the force accumulates
    // d_j He_{j+1}(v)/j! in the same sweep over j
    const A rt2 = (A)sqrt(2.0 / t);
    A Fh[EPT], Fb[EPT], p0[EPT], p1[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        Fh[e] = Fb[e] = (A)0;
        p0[e] = (A)1;
        p1[e] = vv[e];
    }
    double fact = 1.0;
    const int nsweep = fd ? 0 : N + 1;   // no Hermite sweep for the difference force
    for (int jj = 0; jj < nsweep; ++jj) {
        if (jj > 0) fact *= (double)jj;
        double d[2] = {0, 0};
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int k = lane + 64 * e;
            if (k < c.nn) {
                const A wt = wnode[k];
                d[0] += (double)(rh[e] * wt * p0[e]);
                d[1] += (double)(rb[e] * wt * p0[e]);
            }
        }
        wave_sum<2>(d);
        const A dh = (A)(d[0] * (double)rt2 / fact), db = (A)(d[1] * (double)rt2 / fact);
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            Fh[e] += dh * p1[e];
            Fb[e] += db * p1[e];
            const A p2 = vv[e] * p1[e] - (A)(jj + 1) * p0[e];
            p0[e] = p1[e];
            p1[e] = p2;
        }
    }
    const double Hcn = pn[0] * pow(pn[2] / c.pi, 0.5), kfn = c.K / (2.0 * pn[2]);
    const double csn = c.shk * (1.0 - c.prandtl) * pn[2] * pn[2] / pn[0] * qf;
    const A inv_sqrt_2pi = (A)(1.0 / sqrt(2.0 * c.pi));
    const double itau = 1.0 / tau, den = 1.0 + dt / tau;
    const double idx = 1.0 / c.dx;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = lane + 64 * e;
        if (k < c.nn) {
            const double cun = unode_d[k] - pn[1];
            const double H = Hcn * exp(-pn[2] * cun * cun), B = H * kfn;
            const double l2 = 2.0 * pn[2] * cun * cun;
            const double Hp = H + csn * cun * (l2 + c.K - 5.0) * H;
            const double Bp = B + csn * cun * (l2 + c.K - 3.0) * B;
            double dvh, dvb;
            if (fd) {
                dvh = fd1(fdf, c.nn, k, c.fd_du);
                dvb = fd1(fdf + 64 * EPT, c.nn, k, c.fd_du);
            } else {
                const A wv = fexp(-vv[e] * vv[e] / (A)2) * inv_sqrt_2pi;
                dvh = (double)(-rt2 * wv * Fh[e]);
                dvb = (double)(-rt2 * wv * Fb[e]);
            }
            const double dfh = ((double)fh[(size_t)cell * c.ldv + k] - (double)fh[(size_t)(cell + 1) * c.ldv + k]) * idx;
            const double dfb = ((double)fb[(size_t)cell * c.ldv + k] - (double)fb[(size_t)(cell + 1) * c.ldv + k]) * idx;
            const size_t o = (size_t)cell * c.ldv + k;
            h[o] = (T)(((double)rh[e] + dfh + dt * (Hp * itau) - dvh * dt * phi) / den);
            b[o] = (T)(((double)rb[e] + dfb + dt * (Bp * itau) - dvb * dt * phi) / den);
        }
    }
    if (lane == 0) {
        double* pp = part + 9 * (size_t)cell;
        for (int q = 0; q < 3; ++q) {
            W[3 * (size_t)cell + q] = wn[q];
            pp[q] = (wo[q] - wn[q]) * (wo[q] - wn[q]);
            pp[3 + q] = fabs(wn[q]);
        }
        // primitive_change_residual, 1d/src/solver.f90:276-291
        const double ds = fmax(1.0, fmax(fabs(po[0]), fabs(pn[0])));
        pp[6] = fabs(pn[0] - po[0]) / ds;
        pp[7] = fabs(pn[1] - po[1]) / sqrt(c.tref);
        pp[8] = fabs(1.0 / pn[2] - 1.0 / po[2]) / c.tref;
        if (!(pn[2] > 0.0) || !(pn[0] > 0.0) || !isfinite(pn[0]) || !isfinite(pn[1]) || !isfinite(pn[2]))
            atomicOr(&st->flags, 1);
    }
}

}  // namespace hugks
