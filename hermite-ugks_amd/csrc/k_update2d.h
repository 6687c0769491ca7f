// 2D cell update with the Hermite spectral force term or the
// finite-difference one (gfx950).
//
// One workgroup per cell.  Reference: update(), 2d/src/solver.f90:245-395,
// with the force of symmetric_hermite_force_term_2d,
// 2d/src/hermite_symmetric2d.f90:23-104, or of cal_fsource1,
// 2d/src/gravity.f90:108-187.
//
// The Hermite force is evaluated in separable form.  With g = weight*f
// (nv x nv, u index fastest), Hx(r,iu) = He_r(xi_iu), Hy(s,jv) = He_s(eta_jv):
//   Tm(r,jv) = sum_iu g(iu,jv) Hx(r,iu)                      (Nm+1) x nv
//   D(r,s)   = (1/rt) sum_jv Tm(r,jv) Hy(s,jv) / (r! s!),     r+s <= Nm
//   R(r,jv)  = a_x sum_s D(r-1,s) Hy(s,jv) + a_y sum_s D(r,s) Hy(s+1,jv)
//   F(iu,jv) = -(1/sqrt(rt)) omega(iu,jv) sum_{r=0}^{Nm+1} Hx(r,iu) R(r,jv)
// which is the reference's triangular double sum regrouped; it costs
// O((Nm+1) nv^2) per cell instead of O((Nm+1)^2 nv^2 / 2).
#pragma once
#include "hugks_device.h"
#include "hugks_kernels.h"

namespace hugks {

// Rc of the solver's gas constant, lambda = 0.5/(Rc T) (2d/src/data2d.f90:54)
constexpr double RC = 0.5;

// One velocity-axis difference of the finite-difference force
// (cal_fsource1, 2d/src/gravity.f90:108-187) at node i of a line of n
// values f[q*st], step du: one-sided second order at both ends, central
// next to them, fourth-order central inside, in the reference's expression
// order.  As the reference writes them, the interior form takes its three
// inner terms from g (the u-axis difference of b reads them from h), and
// the last node's one-sided form has the sign of a forward difference when
// last_fwd (the v axis).
template <typename A>
__device__ __forceinline__ double fd_axis(const A* f, const A* g, int st, int n, int i,
                                          double du, bool last_fwd) {
    auto F = [&](int q) { return (double)f[q * st]; };
    auto G = [&](int q) { return (double)g[q * st]; };
    if (i == 0) return (-3.0 * F(0) + 4.0 * F(1) - F(2)) / du / 2.0;
    if (i == n - 1)
        return last_fwd ? (-3.0 * F(n - 1) + 4.0 * F(n - 2) - F(n - 3)) / du / 2.0
                        : (3.0 * F(n - 1) - 4.0 * F(n - 2) + F(n - 3)) / du / 2.0;
    if (i == 1 || i == n - 2) return (F(i + 1) - F(i - 1)) / du / 2.0;
    return (-F(i + 2) + 8.0 * G(i + 1) - 8.0 * G(i - 1) + G(i - 2)) / du / 12.0;
}


template <typename T, int EPT, int NT>
__global__ __launch_bounds__(NT) void k_update2d(
    Consts c, DevState* __restrict__ st, double* __restrict__ W, double* __restrict__ part,
    T* __restrict__ h, T* __restrict__ b, const T* __restrict__ vfh, const T* __restrict__ vfb,
    const T* __restrict__ hfh, const T* __restrict__ hfb, const double* __restrict__ vflux,
    const double* __restrict__ hflux, const typename Arith<T>::A* __restrict__ unode,
    const typename Arith<T>::A* __restrict__ vnode, const typename Arith<T>::A* __restrict__ wnode,
    const double* __restrict__ wnode_d, const double* __restrict__ u1,
    const double* __restrict__ v1, const double* __restrict__ acc_cell) {
    using A = typename Arith<T>::A;
    HIP_DYNAMIC_SHARED(double, smem_raw)
    __shared__ double red[4 * (NT / WAVE)];
    if (st->stopped) return;
    const int tid = threadIdx.x;
    const int cell = tiled_cell((int)xcd_block(blockIdx.x, gridDim.x), c.nx, c.ny);
    const int nx = c.nx, nv = c.nv, nm = c.nm, nn = c.nn;
    const int i = cell % nx, j = cell / nx;
    const double dt = st->dt;
    const double area = c.dx * c.dy;
    const int vl = i + (nx + 1) * j, vr = vl + 1;
    const int hd = i + nx * j, hu = i + nx * (j + 1);
    const double ax = acc_cell[2 * cell], ay = acc_cell[2 * cell + 1];

    // ---- macroscopic predictor (solver.f90:302-321)
    double wo[4], po[4], wn[4], pn[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) wo[q] = W[4 * (size_t)cell + q];
    prim_from_w(c, wo, po);
    const double tau_old = tau_of(c, po[0], po[3]);
#pragma unroll
    for (int q = 0; q < 4; ++q)
        wn[q] = wo[q] + (vflux[4 * (size_t)vl + q] - vflux[4 * (size_t)vr + q] +
                         hflux[4 * (size_t)hd + q] - hflux[4 * (size_t)hu + q]) / area;
    prim_from_w(c, wn, pn);
    pn[1] = pn[1] + ax * dt;
    pn[2] = pn[2] + ay * dt;
    w_from_prim(c, pn, wn);
    const double tau = tau_of(c, pn[0], pn[3]);

    // ---- LDS carve-up (the finite-difference force uses G only, holding
    // h and b themselves)
    const bool fd = c.force_fd != 0;
    A* G = reinterpret_cast<A*>(smem_raw);        // [2][nn]
    A* Hx = G + 2 * nn;                           // [nm+2][nv]
    A* Hy = Hx + (nm + 2) * nv;                   // [nm+2][nv]
    A* Tm = Hy + (nm + 2) * nv;                   // [2][nm+1][nv]
    A* D = Tm + 2 * (nm + 1) * nv;                // [2][nm+1][nm+1]
    A* R = D + 2 * (nm + 1) * (nm + 1);           // [2][nm+2][nv]

    // ---- load f^n, weighted copies, heat flux at t^n (solver.f90:337).
    // The heat flux and, below, the Maxwellians and the implicit update are
    // evaluated in FP64 whatever the storage type: the new f is close to
    // the Maxwellian of the new state whenever dt >> tau, and an FP32
    // exponent would carry a systematic relative error of order
    // lambda c^2 * 2^-24 into every moment of the next step; the Hermite
    // force (scaled by dt * a) is evaluated in the storage arithmetic.
    const double rt = RC / po[3];
    const double sqrt_rt = sqrt(rt);
    T rh[EPT], rb[EPT];
    double q4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = tid + e * NT;
        rh[e] = rb[e] = (T)0;
        if (k < nn) {
            const size_t o = (size_t)cell * c.ldv + k;
            rh[e] = h[o];
            rb[e] = b[o];
            const double wt = wnode_d[k];
            const double cu = u1[k % nv] - po[1], cv = v1[k / nv] - po[2];
            const double c2 = cu * cu + cv * cv;
            q4[0] += wt * cu * c2 * (double)rh[e];
            q4[1] += wt * cu * (double)rb[e];
            q4[2] += wt * cv * c2 * (double)rh[e];
            q4[3] += wt * cv * (double)rb[e];
            const A wa = fd ? (A)1 : wnode[k];
            G[k] = wa * (A)rh[e];
            G[nn + k] = wa * (A)rb[e];
        }
    }
    // Hermite bases He_0..He_{Nm+1} at xi = (u-U)/sqrt(rt), eta = (v-V)/sqrt(rt)
    // (build_probabilists_hermite, hermite_symmetric2d.f90:220-233)
    if (!fd && tid < 2 * nv) {
        const bool isx = tid < nv;
        const int t = isx ? tid : tid - nv;
        const A x = (A)((isx ? (u1[t] - po[1]) : (v1[t] - po[2])) / sqrt_rt);
        A* Hb = isx ? Hx : Hy;
        A p0 = (A)1, p1 = x;
        Hb[t] = p0;
        Hb[nv + t] = p1;
        for (int n = 2; n <= nm + 1; ++n) {
            const A p2 = x * p1 - (A)(n - 1) * p0;
            Hb[n * nv + t] = p2;
            p0 = p1;
            p1 = p2;
        }
    }
    block_sum<4, NT>(q4, red);   // also orders the LDS writes above
    const double qf0 = 0.5 * (q4[0] + q4[1]), qf1 = 0.5 * (q4[2] + q4[3]);

    const int m1 = nm + 1, m2 = nm + 2;
    if (!fd) {
        // ---- Tm = Hx^T g  (contract the u index)
        const int ntm = 2 * (nm + 1) * nv;
        for (int t = tid; t < ntm; t += NT) {
            const int f = t / ((nm + 1) * nv);
            const int rem = t - f * (nm + 1) * nv;
            const int r = rem / nv, jv = rem - r * nv;
            const A* g = G + f * nn + nv * jv;
            const A* hx = Hx + r * nv;
            A acc = (A)0;
            for (int iu = 0; iu < nv; ++iu) acc += g[iu] * hx[iu];
            Tm[t] = acc;
        }
        __syncthreads();
        // ---- D = jac * Tm Hy^T / (r! s!), triangular r+s <= Nm
        const double jac = 1.0 / rt;
        for (int t = tid; t < 2 * m1 * m1; t += NT) {
            const int f = t / (m1 * m1);
            const int rem = t - f * m1 * m1;
            const int r = rem / m1, s = rem - r * m1;
            A val = (A)0;
            if (r + s <= nm) {
                const A* tm = Tm + (f * m1 + r) * nv;
                const A* hy = Hy + s * nv;
                A acc = (A)0;
                for (int jv = 0; jv < nv; ++jv) acc += tm[jv] * hy[jv];
                double fr = 1.0, fs = 1.0;
                for (int q = 2; q <= r; ++q) fr *= (double)q;
                for (int q = 2; q <= s; ++q) fs *= (double)q;
                val = (A)(jac / (fr * fs)) * acc;
            }
            D[t] = val;
        }
        __syncthreads();
        // ---- R(r,jv) = a_x sum_s D(r-1,s) Hy(s,jv) + a_y sum_s D(r,s) Hy(s+1,jv)
        const A axA = (A)ax, ayA = (A)ay;
        for (int t = tid; t < 2 * m2 * nv; t += NT) {
            const int f = t / (m2 * nv);
            const int rem = t - f * m2 * nv;
            const int r = rem / nv, jv = rem - r * nv;
            const A* Df = D + f * m1 * m1;
            A acc = (A)0;
            if (r >= 1) {
                A sx = (A)0;
                for (int s = 0; s <= nm - (r - 1); ++s) sx += Df[(r - 1) * m1 + s] * Hy[s * nv + jv];
                acc += axA * sx;
            }
            if (r <= nm) {
                A sy = (A)0;
                for (int s = 0; s <= nm - r; ++s) sy += Df[r * m1 + s] * Hy[(s + 1) * nv + jv];
                acc += ayA * sy;
            }
            R[t] = acc;
        }
        __syncthreads();
    }

    // ---- force and implicit BGK update (solver.f90:337-388)
    const A inv_sqrt_rt = (A)(1.0 / sqrt_rt);
    const A two_pi = (A)(2.0 * c.pi);
    const double Hco = po[0] * (po[3] / c.pi), Hcn = pn[0] * (pn[3] / c.pi);
    const double kfo = c.K / (2.0 * po[3]), kfn = c.K / (2.0 * pn[3]);
    const double cso = c.shk * (1.0 - c.prandtl) * po[3] * po[3] / po[0];
    const double csn = c.shk * (1.0 - c.prandtl) * pn[3] * pn[3] / pn[0];
    const double half_dt = 0.5 * dt;
    const double itau = 1.0 / tau, itau_o = 1.0 / tau_old;
    const double den = 1.0 + 0.5 * dt / tau;
    const double iarea = 1.0 / area;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = tid + e * NT;
        if (k < nn) {
            const int iu = k % nv, jv = k / nv;
            // the force source hsource, bsource (solver.f90:352-375)
            double hs, bs;
            if (fd) {
                const A *gh = G, *gb = G + nn;
                const double du = c.fd_du;
                const double duh = fd_axis(gh + nv * jv, gh + nv * jv, 1, nv, iu, du, false);
                const double dub = fd_axis(gb + nv * jv, gh + nv * jv, 1, nv, iu, du, false);
                const double dvh = fd_axis(gh + iu, gh + iu, nv, nv, jv, du, true);
                const double dvb = fd_axis(gb + iu, gb + iu, nv, nv, jv, du, true);
                hs = duh * dt * ax + dvh * dt * ay;
                bs = dub * dt * ax + dvb * dt * ay;
            } else {
                const A xi = Hx[nv + iu], eta = Hy[nv + jv];
                const A om = fexp((A)-0.5 * (xi * xi + eta * eta)) / two_pi;
                A sh = (A)0, sb = (A)0;
                for (int r = 0; r <= nm + 1; ++r) {
                    const A hx = Hx[r * nv + iu];
                    sh += hx * R[r * nv + jv];
                    sb += hx * R[(m2 + r) * nv + jv];
                }
                hs = (double)(-inv_sqrt_rt * om * sh) * dt;
                bs = (double)(-inv_sqrt_rt * om * sb) * dt;
            }
            // Maxwellians and Shakhov parts at t^n and t^{n+1}, both with the
            // heat flux of h^n (solver.f90:337-347)
            const double u = u1[iu], v = v1[jv];
            const double cuo = u - po[1], cvo = v - po[2], c2o = cuo * cuo + cvo * cvo;
            const double Ho = Hco * exp(-po[3] * c2o), Bo = Ho * kfo;
            const double cqo = cso * (cuo * qf0 + cvo * qf1);
            const double Hpo = Ho + cqo * (2.0 * po[3] * c2o + c.K - 5.0) * Ho;
            const double Bpo = Bo + cqo * (2.0 * po[3] * c2o + c.K - 3.0) * Bo;
            const double cun = u - pn[1], cvn = v - pn[2], c2n = cun * cun + cvn * cvn;
            const double Hn = Hcn * exp(-pn[3] * c2n), Bn = Hn * kfn;
            const double cqn = csn * (cun * qf0 + cvn * qf1);
            const double Hpn = Hn + cqn * (2.0 * pn[3] * c2n + c.K - 5.0) * Hn;
            const double Bpn = Bn + cqn * (2.0 * pn[3] * c2n + c.K - 3.0) * Bn;
            const double dfh = ((double)vfh[(size_t)vl * c.ldv + k] - (double)vfh[(size_t)vr * c.ldv + k] +
                                (double)hfh[(size_t)hd * c.ldv + k] - (double)hfh[(size_t)hu * c.ldv + k]) * iarea;
            const double dfb = ((double)vfb[(size_t)vl * c.ldv + k] - (double)vfb[(size_t)vr * c.ldv + k] +
                                (double)hfb[(size_t)hd * c.ldv + k] - (double)hfb[(size_t)hu * c.ldv + k]) * iarea;
            const double ho = (double)rh[e], bo = (double)rb[e];
            const double hn = (ho + dfh + half_dt * (Hpn * itau + (Hpo - ho) * itau_o) - hs) / den;
            const double bn = (bo + dfb + half_dt * (Bpn * itau + (Bpo - bo) * itau_o) - bs) / den;
            const size_t o = (size_t)cell * c.ldv + k;
            h[o] = (T)hn;
            b[o] = (T)bn;
        }
    }
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            W[4 * (size_t)cell + q] = wn[q];
            part[5 * (size_t)cell + q] = (wo[q] - wn[q]) * (wo[q] - wn[q]);
        }
        part[5 * (size_t)cell + 4] = fabs(wn[0]);
        if (!(pn[3] > 0.0) || !isfinite(pn[0]) || !isfinite(pn[1]) || !isfinite(pn[2]) ||
            !isfinite(pn[3]) || !(pn[0] > 0.0))
            atomicOr(&st->flags, 1);
    }
}

}  // namespace hugks
