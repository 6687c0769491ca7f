// 2D cell update with the Hermite force on MFMA (gfx950).
//
// Same step as k_update2d.h - update(), 2d/src/solver.f90:245-395, with
// symmetric_hermite_force_term_2d, 2d/src/hermite_symmetric2d.f90:23-104 -
// but the four separable contractions of the force run as 16x16x4 matrix
// tiles (v_mfma_f32_16x16x4f32 on the mixed path, v_mfma_f64_16x16x4f64 on
// the FP64 path) instead of LDS-operand dot products:
//   Tm = Hx G            (16 x NP) = (16 x NP)(NP x NP),  G = weight * f
//   D  = Tm Hy^T / r!s!  (16 x 16), zero where r + s > Nm
//   R  = ax D1 Hy + ay D Hy1          D1 = D shifted down one row,
//                                     Hy1 = Hy shifted up one row
//   F  = Hx^T R          (NP x NP)
// with NP = nv rounded up to 16, Hx(r, iu) = He_r(xi_iu), Hy(s, jv) =
// He_s(eta_jv) for r, s = 0..15 (so Nm <= 14).  One workgroup (four
// wavefronts) per cell; the tiles of each contraction are shared out over
// the wavefronts, with a workgroup barrier between contractions.  Operands
// and results pass through LDS; the node-wise predictor, Maxwellians and
// implicit update are those of k_update2d.h.
#pragma once
#include "hugks_device.h"
#include "hugks_kernels.h"
#include "k_update2d.h"

namespace hugks {

typedef float mf_f4 __attribute__((vector_size(4 * sizeof(float))));
typedef double mf_d4 __attribute__((vector_size(4 * sizeof(double))));

template <typename T> struct Mfma;
template <> struct Mfma<float> {
    using V = mf_f4;
    // accumulator element v of lane l holds row 4*(l/16)+v, column l%16
    static __device__ __forceinline__ int row(int l, int v) { return 4 * (l >> 4) + v; }
    static __device__ __forceinline__ V mma(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<double> {
    using V = mf_d4;
    // accumulator element v of lane l holds row (l/16)+4*v, column l%16
    static __device__ __forceinline__ int row(int l, int v) { return (l >> 4) + 4 * v; }
    static __device__ __forceinline__ V mma(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

constexpr int MF_NT = 256;   // threads per workgroup (one cell)
constexpr int MF_NB = 16;    // Hermite basis rows He_0..He_15

__host__ __device__ constexpr int mf_np(int nv) { return (nv + 15) / 16 * 16; }

// LDS elements (of T) of one cell: G/F [2][nn], Hx, Hy [16][NP],
// Tm [2][16][NP], D [2][16][16], R [2][16][NP]; then, as doubles, the
// per-axis factors of the two Maxwellians and of the force weight,
// [6][NP] (counted here in T elements).
__host__ __device__ inline size_t mf_lds_elems(int nv) {
    const size_t np = mf_np(nv), nn = (size_t)nv * nv;
    return 2 * nn + 2 * MF_NB * np + 2 * MF_NB * np + 2 * MF_NB * MF_NB + 2 * MF_NB * np + 1 +
           6 * np * 2;   // doubles take at most two T elements; +1 aligns them
}

// Per-cell scalars of the update, one lane per cell, computed once instead
// of by every thread of the cell's workgroup: the macroscopic predictor
// (solver.f90:302-321), both collision times and the coefficients of the
// node-wise update; also the new W, the residual partials and the
// NaN/T<=0 flag.  Record of CP_LD doubles per cell (slots CP_*).
constexpr int CP_LD = 24;
enum : int {
    CP_PO = 0, CP_PN = 4, CP_TAUO = 8, CP_TAU = 9, CP_AX = 10, CP_AY = 11, CP_RT = 12,
    CP_HCO = 13, CP_HCN = 14, CP_KFO = 15, CP_KFN = 16, CP_CSO = 17, CP_CSN = 18,
    CP_IDEN = 19, CP_ITAU = 20, CP_ITAUO = 21,
};

__global__ __launch_bounds__(256) void k_cell_pre(Consts c, DevState* __restrict__ st,
                                                  double* __restrict__ W,
                                                  double* __restrict__ part,
                                                  const double* __restrict__ vflux,
                                                  const double* __restrict__ hflux,
                                                  const double* __restrict__ acc_cell,
                                                  double* __restrict__ cp) {
    if (st->stopped) return;
    const int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= c.nx * c.ny) return;
    const int nx = c.nx;
    const int i = cell % nx, j = cell / nx;
    const double dt = st->dt;
    const double area = c.dx * c.dy;
    const int vl = i + (nx + 1) * j, vr = vl + 1;
    const int hd = i + nx * j, hu = i + nx * (j + 1);
    const double ax = acc_cell[2 * cell], ay = acc_cell[2 * cell + 1];

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

    double* P = cp + (size_t)CP_LD * cell;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        P[CP_PO + q] = po[q];
        P[CP_PN + q] = pn[q];
    }
    P[CP_TAUO] = tau_old;
    P[CP_TAU] = tau;
    P[CP_AX] = ax;
    P[CP_AY] = ay;
    P[CP_RT] = RC / po[3];
    P[CP_HCO] = po[0] * (po[3] / c.pi);
    P[CP_HCN] = pn[0] * (pn[3] / c.pi);
    P[CP_KFO] = c.K / (2.0 * po[3]);
    P[CP_KFN] = c.K / (2.0 * pn[3]);
    P[CP_CSO] = c.shk * (1.0 - c.prandtl) * po[3] * po[3] / po[0];
    P[CP_CSN] = c.shk * (1.0 - c.prandtl) * pn[3] * pn[3] / pn[0];
    P[CP_IDEN] = 1.0 / (1.0 + 0.5 * dt / tau);
    P[CP_ITAU] = 1.0 / tau;
    P[CP_ITAUO] = 1.0 / tau_old;
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

// Shakhov-BGK update of one cell per workgroup with the Hermite force on
// MFMA; the cell's scalars come from k_cell_pre.
template <typename T, int EPT>
__global__ __launch_bounds__(MF_NT) void k_update2d_mfma(
    Consts c, const DevState* __restrict__ st, const double* __restrict__ cp,
    T* __restrict__ h, T* __restrict__ b, const T* __restrict__ vfh, const T* __restrict__ vfb,
    const T* __restrict__ hfh, const T* __restrict__ hfb,
    const typename Arith<T>::A* __restrict__ wnode, const double* __restrict__ wnode_d,
    const double* __restrict__ und, const double* __restrict__ vnd,
    const double* __restrict__ u1, const double* __restrict__ v1) {
    using A = typename Arith<T>::A;
    using M = Mfma<T>;
    using V = typename M::V;
    HIP_DYNAMIC_SHARED(double, smem_raw)
    __shared__ double red[4 * (MF_NT / WAVE)];
    if (st->stopped) return;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
    const int cell = tiled_cell((int)xcd_block(blockIdx.x, gridDim.x), c.nx, c.ny);
    const int nx = c.nx, nv = c.nv, nm = c.nm, nn = c.nn;
    const int np = mf_np(nv);
    const int i = cell % nx, j = cell / nx;
    const double dt = st->dt;
    const double area = c.dx * c.dy;
    const int vl = i + (nx + 1) * j, vr = vl + 1;
    const int hd = i + nx * j, hu = i + nx * (j + 1);
    const double* P = cp + (size_t)CP_LD * cell;
    double po[4], pn[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        po[q] = P[CP_PO + q];
        pn[q] = P[CP_PN + q];
    }
    const double ax = P[CP_AX], ay = P[CP_AY];

    // ---- LDS carve-up
    T* G = reinterpret_cast<T*>(smem_raw);        // [2][nn]; later F
    T* Hx = G + 2 * nn;                           // [16][np]
    T* Hy = Hx + MF_NB * np;                      // [16][np]
    T* Tm = Hy + MF_NB * np;                      // [2][16][np]
    T* D = Tm + 2 * MF_NB * np;                   // [2][16][16]
    T* R = D + 2 * MF_NB * MF_NB;                 // [2][16][np]
    // per-axis factors (FP64): the Maxwellians at t^n and t^{n+1} and the
    // force weight are products of a u factor and a v factor,
    // exp(-lam (u-U)^2 - lam (v-V)^2) = ex(u) ey(v)
    double* Ex = reinterpret_cast<double*>(
        (reinterpret_cast<uintptr_t>(R + 2 * MF_NB * np) + 7) & ~(uintptr_t)7);
    double* Eo_x = Ex;                            // [np] old Maxwellian, u
    double* Eo_y = Ex + np;                       //      v
    double* En_x = Ex + 2 * np;                   // new Maxwellian
    double* En_y = Ex + 3 * np;
    double* Om_x = Ex + 4 * np;                   // force weight exp(-xi^2/2)
    double* Om_y = Ex + 5 * np;

    // ---- load f^n, weighted copies, heat flux at t^n (solver.f90:337),
    // FP64 whatever the storage type
    const double rt = P[CP_RT];
    const double sqrt_rt = sqrt(rt);
    T rh[EPT], rb[EPT];
    double q4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = tid + e * MF_NT;
        rh[e] = rb[e] = (T)0;
        if (k < nn) {
            const size_t o = (size_t)cell * c.ldv + k;
            rh[e] = h[o];
            rb[e] = b[o];
            const double wt = wnode_d[k];
            const double cu = und[k] - po[1], cv = vnd[k] - po[2];
            const double c2 = cu * cu + cv * cv;
            q4[0] += wt * cu * c2 * (double)rh[e];
            q4[1] += wt * cu * (double)rb[e];
            q4[2] += wt * cv * c2 * (double)rh[e];
            q4[3] += wt * cv * (double)rb[e];
            const A wa = wnode[k];
            G[k] = (T)(wa * (A)rh[e]);
            G[nn + k] = (T)(wa * (A)rb[e]);
        }
    }
    // Hermite bases He_0..He_15 at xi = (u-U)/sqrt(rt), eta = (v-V)/sqrt(rt)
    // (build_probabilists_hermite, hermite_symmetric2d.f90:220-233); zero
    // columns beyond nv
    for (int t = tid; t < 2 * np; t += MF_NT) {
        const bool isx = t < np;
        const int idx = isx ? t : t - np;
        T* Hb = isx ? Hx : Hy;
        if (idx < nv) {
            const double xv = isx ? u1[idx] : v1[idx];
            const double co = xv - (isx ? po[1] : po[2]), cn = xv - (isx ? pn[1] : pn[2]);
            (isx ? Eo_x : Eo_y)[idx] = exp(-po[3] * co * co);
            (isx ? En_x : En_y)[idx] = exp(-pn[3] * cn * cn);
            const A x = (A)(co / sqrt_rt);
            (isx ? Om_x : Om_y)[idx] = (double)fexp((A)-0.5 * x * x);
            A p0 = (A)1, p1 = x;
            Hb[idx] = (T)p0;
            Hb[np + idx] = (T)p1;
            for (int n = 2; n < MF_NB; ++n) {
                const A p2 = x * p1 - (A)(n - 1) * p0;
                Hb[n * np + idx] = (T)p2;
                p0 = p1;
                p1 = p2;
            }
        } else {
            for (int n = 0; n < MF_NB; ++n) Hb[n * np + idx] = (T)0;
        }
    }
    block_sum<4, MF_NT>(q4, red);   // also orders the LDS writes above
    const double qf0 = 0.5 * (q4[0] + q4[1]), qf1 = 0.5 * (q4[2] + q4[3]);

    const int nt = np / 16;             // tiles along a velocity index
    const int ksn = (nv + 3) / 4;       // K steps over a velocity index
    const int lr = lane & 15, lk = lane >> 4;

    // ---- Tm = Hx G: tile (f, jt), rows r, columns jv
    for (int job = wv; job < 2 * nt; job += MF_NT / WAVE) {
        const int f = job / nt, jt = job % nt;
        const int jv = 16 * jt + lr;
        V acc = {0, 0, 0, 0};
        for (int ks = 0; ks < ksn; ++ks) {
            const int iu = 4 * ks + lk;
            const bool in = iu < nv;
            const T a = in ? Hx[lr * np + iu] : (T)0;
            const T bb = (in && jv < nv) ? G[f * nn + jv * nv + iu] : (T)0;
            acc = M::mma(a, bb, acc);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) Tm[(f * MF_NB + M::row(lane, v)) * np + jv] = acc[v];
    }
    __syncthreads();
    // ---- D = Tm Hy^T / (r! s!), r + s <= Nm
    const double jac = 1.0 / rt;
    for (int f = wv; f < 2; f += MF_NT / WAVE) {
        V acc = {0, 0, 0, 0};
        for (int ks = 0; ks < ksn; ++ks) {
            const int jv = 4 * ks + lk;
            const T a = Tm[(f * MF_NB + lr) * np + jv];     // zero beyond nv
            const T bb = Hy[lr * np + jv];
            acc = M::mma(a, bb, acc);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = M::row(lane, v), s = lr;
            T val = (T)0;
            if (r + s <= nm) {
                double fr = 1.0, fs = 1.0;
                for (int q = 2; q <= r; ++q) fr *= (double)q;
                for (int q = 2; q <= s; ++q) fs *= (double)q;
                val = (T)((A)(jac / (fr * fs)) * (A)acc[v]);
            }
            D[(f * MF_NB + r) * MF_NB + s] = val;
        }
    }
    __syncthreads();
    // ---- R(r, jv) = ax sum_s D(r-1, s) Hy(s, jv) + ay sum_s D(r, s) Hy(s+1, jv)
    const A axA = (A)ax, ayA = (A)ay;
    for (int job = wv; job < 2 * nt; job += MF_NT / WAVE) {
        const int f = job / nt, jt = job % nt;
        const int jv = 16 * jt + lr;
        V acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < MF_NB / 4; ++ks) {
            const int s = 4 * ks + lk;
            const T a1 = lr >= 1 ? (T)(axA * (A)D[(f * MF_NB + lr - 1) * MF_NB + s]) : (T)0;
            acc = M::mma(a1, Hy[s * np + jv], acc);
            const T a2 = (T)(ayA * (A)D[(f * MF_NB + lr) * MF_NB + s]);
            const T b2 = s + 1 < MF_NB ? Hy[(s + 1) * np + jv] : (T)0;
            acc = M::mma(a2, b2, acc);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) R[(f * MF_NB + M::row(lane, v)) * np + jv] = acc[v];
    }
    __syncthreads();
    // ---- F = Hx^T R, into the G region (G is dead)
    T* F = G;
    for (int job = wv; job < 2 * nt * nt; job += MF_NT / WAVE) {
        const int f = job / (nt * nt), rem = job % (nt * nt);
        const int it = rem / nt, jt = rem % nt;
        const int iu_a = 16 * it + lr, jv = 16 * jt + lr;
        V acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < MF_NB / 4; ++ks) {
            const int r = 4 * ks + lk;
            acc = M::mma(Hx[r * np + iu_a], R[(f * MF_NB + r) * np + jv], acc);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int iu = 16 * it + M::row(lane, v);
            if (iu < nv && jv < nv) F[f * nn + iu + nv * jv] = acc[v];
        }
    }
    __syncthreads();

    // ---- force and implicit BGK update (solver.f90:337-388)
    const A inv_sqrt_rt = (A)(1.0 / sqrt_rt);
    const A two_pi = (A)(2.0 * c.pi);
    const double Hco = P[CP_HCO], Hcn = P[CP_HCN];
    const double kfo = P[CP_KFO], kfn = P[CP_KFN];
    const double cso = P[CP_CSO], csn = P[CP_CSN];
    const double half_dt = 0.5 * dt;
    const double itau = P[CP_ITAU], itau_o = P[CP_ITAUO];
    const double iden = P[CP_IDEN];
    const double iarea = 1.0 / area;
    const float rnv = 1.0f / (float)nv;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = tid + e * MF_NT;
        if (k < nn) {
            // node (iu, jv) of k = iu + nv*jv without an integer division:
            // (k + 1/2)/nv stays at least 1/(2 nv) away from an integer
            const int jv = (int)(((float)k + 0.5f) * rnv), iu = k - jv * nv;
            const double u = und[k], v = vnd[k];
            const A om = (A)(Om_x[iu] * Om_y[jv]) / two_pi;
            const double Fh = (double)(-inv_sqrt_rt * om * (A)F[k]);
            const double Fb = (double)(-inv_sqrt_rt * om * (A)F[nn + k]);
            const double cuo = u - po[1], cvo = v - po[2], c2o = cuo * cuo + cvo * cvo;
            const double Ho = Hco * (Eo_x[iu] * Eo_y[jv]), Bo = Ho * kfo;
            const double cqo = cso * (cuo * qf0 + cvo * qf1);
            const double Hpo = Ho + cqo * (2.0 * po[3] * c2o + c.K - 5.0) * Ho;
            const double Bpo = Bo + cqo * (2.0 * po[3] * c2o + c.K - 3.0) * Bo;
            const double cun = u - pn[1], cvn = v - pn[2], c2n = cun * cun + cvn * cvn;
            const double Hn = Hcn * (En_x[iu] * En_y[jv]), Bn = Hn * kfn;
            const double cqn = csn * (cun * qf0 + cvn * qf1);
            const double Hpn = Hn + cqn * (2.0 * pn[3] * c2n + c.K - 5.0) * Hn;
            const double Bpn = Bn + cqn * (2.0 * pn[3] * c2n + c.K - 3.0) * Bn;
            const double dfh = ((double)vfh[(size_t)vl * c.ldv + k] - (double)vfh[(size_t)vr * c.ldv + k] +
                                (double)hfh[(size_t)hd * c.ldv + k] - (double)hfh[(size_t)hu * c.ldv + k]) * iarea;
            const double dfb = ((double)vfb[(size_t)vl * c.ldv + k] - (double)vfb[(size_t)vr * c.ldv + k] +
                                (double)hfb[(size_t)hd * c.ldv + k] - (double)hfb[(size_t)hu * c.ldv + k]) * iarea;
            const double ho = (double)rh[e], bo = (double)rb[e];
            const double hn = (ho + dfh + half_dt * (Hpn * itau + (Hpo - ho) * itau_o) - Fh * dt) * iden;
            const double bn = (bo + dfb + half_dt * (Bpn * itau + (Bpo - bo) * itau_o) - Fb * dt) * iden;
            const size_t o = (size_t)cell * c.ldv + k;
            h[o] = (T)hn;
            b[o] = (T)bn;
        }
    }
}

}  // namespace hugks
