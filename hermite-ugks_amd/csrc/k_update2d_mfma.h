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
// and results pass through LDS.  The per-cell predictor and coefficients
// come from k_cell_pre; the node-wise Maxwellians, Shakhov terms and force
// weight are products and sums of per-axis records (MR_*).
#pragma once
#include "hugks_device.h"
#include "hugks_kernels.h"
#include "k_update2d.h"

namespace hugks {

// The single-case kernels of this header are defined only where
// HUGKS_BODIES_ONLY is not set: the grouped-launch file (launch_group.hip)
// takes the device bodies alone, so no kernel is defined in two files.

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

// Threads per workgroup (one cell).  FP64: 256 or 320, whichever leaves
// fewer of the EPT x threads node slots idle (at nv = 40, 320 x 5 nodes:
// 14.0 -> 12.7 ms at 512^2 cells).  FP32 storage: 256 (320 measured slower,
// 5.9 -> 6.8 ms).
__host__ __device__ inline int mf_threads(int nn, bool fp64) {
    const int w256 = (nn + 255) / 256 * 256 - nn, w320 = (nn + 319) / 320 * 320 - nn;
    return fp64 && w320 < w256 ? 320 : 256;
}

// Compiled instances (nodes per thread, threads, nv or 0 for any nv): the
// common Gauss-Hermite grids per storage type, then the run-time kernels.
#define HUGKS_UPDM_F32(X) X(1, 256, 16) X(2, 256, 20) X(3, 256, 24) X(4, 256, 28) \
    X(4, 256, 32) X(7, 256, 40)
#define HUGKS_UPDM_F64(X) X(1, 256, 16) X(2, 256, 20) X(2, 320, 24) X(3, 320, 28) \
    X(4, 256, 32) X(5, 320, 40)
#define HUGKS_UPDM_ANY(X) X(1, 256, 0) X(2, 256, 0) X(3, 256, 0) X(4, 256, 0) X(5, 256, 0) \
    X(6, 256, 0) X(7, 256, 0) X(8, 256, 0) X(1, 320, 0) X(2, 320, 0) X(3, 320, 0)        \
    X(4, 320, 0) X(5, 320, 0) X(6, 320, 0) X(7, 320, 0)

// The instance a case runs: (nodes per thread, threads, compiled-in nv).
struct UpdKey {
    int ept, nt, nvc;
};
__host__ inline UpdKey update_key(const Consts& c, bool fp64) {
    UpdKey k;
    k.nt = mf_threads(c.nn, fp64);
    k.ept = (c.nn + k.nt - 1) / k.nt;
    k.nvc = (c.nv == 16 || c.nv == 20 || c.nv == 24 || c.nv == 28 || c.nv == 32 || c.nv == 40)
                ? c.nv : 0;
    bool listed = false;
#define HUGKS_UPDM_HAS(E, N, V) listed = listed || (k.ept == E && k.nt == N && k.nvc == V);
    if (fp64) {
        HUGKS_UPDM_F64(HUGKS_UPDM_HAS)
    } else {
        HUGKS_UPDM_F32(HUGKS_UPDM_HAS)
    }
#undef HUGKS_UPDM_HAS
    if (!listed) k.nvc = 0;   // the run-time kernel
    return k;
}
constexpr int MF_NB = 16;    // Hermite basis rows He_0..He_15

// r! for r = 0..15 (exact in double)
static __constant__ const double kFact[MF_NB] = {
    1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0, 40320.0, 362880.0, 3628800.0, 39916800.0,
    479001600.0, 6227020800.0, 87178291200.0, 1307674368000.0};

__host__ __device__ constexpr int mf_np(int nv) { return (nv + 15) / 16 * 16; }

// Per-axis records of the node-wise update (doubles, one array of nv per
// slot): for the u axis (TX, by iu) and the v axis (TY, by jv) every factor
// that depends on that velocity component alone, so that a node's
// Maxwellians, Shakhov terms and force weight are a few products and sums
// of one TX and one TY entry:
//   MR_E0, MR_E1  exp(-lam c^2) at t^n, t^{n+1}; TX also carries rho lam / pi
//   MR_X0, MR_X1  2 lam c^2, TX also K - 5: their sum is 2 lam |c|^2 + K - 5
//   MR_CQ0        peculiar velocity at t^n (heat flux), then the Shakhov
//                 projection cs q_x c_u (TX), cs q_y c_v (TY) at t^n
//   MR_OM         exp(-xi^2 / 2), TX also dt / (2 pi sqrt(RT)): the force weight
//   MR_Q1         Shakhov projection at t^{n+1}
constexpr int MF_REC = 7;
enum : int { MR_E0 = 0, MR_X0 = 1, MR_E1 = 2, MR_X1 = 3, MR_CQ0 = 4, MR_OM = 5, MR_Q1 = 6 };

// LDS layout of one cell (offsets in elements of T, records in doubles).
// Row strides are chosen for conflict-free operand reads: a tile operand
// read as 16 rows x 4 consecutive columns needs a row stride of 4 x odd
// (Hx, Hy, Tm, D, G), one read as 4 rows x 16 consecutive columns an odd
// multiple of 16 (R).
//   G [2][nv][ldG]  weight * f, later F [2][nn] (unpadded)
//   Hx, Hy [16][sH]   Hermite bases
//   Tm [2][16][sH]  first contraction; R [2][16][sR] reuses it
//   D [2][16][sD]    in the G region when it fits there
//   TX, TY [MF_REC][nv] (doubles, 16-byte aligned)
struct MfLayout {
    int np, sH, sR, sD, ldG;
    int oHx, oHy, oTm, oD, endT;
};
__host__ __device__ inline MfLayout mf_layout(int nv, int tsize) {
    MfLayout L;
    L.np = mf_np(nv);
    L.sH = L.np + 4;
    L.sR = ((L.np / 16) & 1) ? L.np : L.np + 16;
    L.sD = MF_NB + 4;
    L.ldG = (nv + 3) / 4 * 4;
    if (((L.ldG / 4) & 1) == 0) L.ldG += 4;
    L.oHx = 2 * nv * L.ldG;
    L.oHy = L.oHx + MF_NB * L.sH;
    L.oTm = L.oHy + MF_NB * L.sH;
    const int endTm = L.oTm + 2 * MF_NB * (L.sH > L.sR ? L.sH : L.sR);
    // D lives in the G region when it fits: G is dead once Tm is formed, F
    // is written there only after the last read of D
    const bool d_in_g = 2 * nv * L.ldG >= 2 * MF_NB * L.sD;
    L.oD = d_in_g ? 0 : endTm;
    const int end = d_in_g ? endTm : endTm + 2 * MF_NB * L.sD;
    L.endT = (end * tsize + 15) / 16 * 16 / tsize;   // records start 16-byte aligned
    return L;
}
__host__ __device__ inline size_t mf_lds_bytes(int nv, size_t tsize) {
    const MfLayout L = mf_layout(nv, (int)tsize);
    return (size_t)L.endT * tsize + 2 * MF_REC * (size_t)nv * sizeof(double);
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

__device__ __forceinline__ void k_cell_pre_body(Consts c, DevState* __restrict__ st,
                                                  double* __restrict__ W,
                                                  double* __restrict__ part,
                                                  const double* __restrict__ vflux,
                                                  const double* __restrict__ hflux,
                                                  const double* __restrict__ acc_cell,
                                                  double* __restrict__ cp, int bid, int nblk) {
    if (st->stopped) return;
    const int cell = bid * 256 + threadIdx.x;
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
#ifndef HUGKS_BODIES_ONLY
__global__ __launch_bounds__(256) void k_cell_pre(Consts c, DevState* __restrict__ st,
                                                  double* __restrict__ W,
                                                  double* __restrict__ part,
                                                  const double* __restrict__ vflux,
                                                  const double* __restrict__ hflux,
                                                  const double* __restrict__ acc_cell,
                                                  double* __restrict__ cp) {
    k_cell_pre_body(c, st, W, part, vflux, hflux, acc_cell, cp, (int)blockIdx.x, (int)gridDim.x);
}
#endif

// Shakhov-BGK update of one cell per workgroup with the Hermite force on
// MFMA; the cell's scalars come from k_cell_pre.
// NVC > 0: the velocity grid size as a compile-time constant (common
// Gauss-Hermite sizes), so that the LDS layout and the tile loops fold;
// NVC = 0: any nv, read at run time.
template <typename T, int EPT, int NT, int NVC>
__device__ __forceinline__ void k_update2d_mfma_body(
    const Consts& c, const DevState* __restrict__ st, const double* __restrict__ cp,
    T* __restrict__ h, T* __restrict__ b, const T* __restrict__ vfh, const T* __restrict__ vfb,
    const T* __restrict__ hfh, const T* __restrict__ hfb,
    const typename Arith<T>::A* __restrict__ wnode, const double* __restrict__ wnode_d,
    const double* __restrict__ u1, const double* __restrict__ v1, int bid, int nblk) {
    using A = typename Arith<T>::A;
    using M = Mfma<T>;
    using V = typename M::V;
    HIP_DYNAMIC_SHARED(double, smem_raw)
    __shared__ double red[4 * (NT / WAVE)];
    if (st->stopped) return;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
    const int cell = tiled_cell((int)xcd_block(bid, nblk), c.nx, c.ny);
    const int nx = c.nx, nm = c.nm;
    const int nv = NVC ? NVC : c.nv, nn = NVC ? NVC * NVC : c.nn;
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
    const MfLayout ly = mf_layout(nv, (int)sizeof(T));
    const int sH = ly.sH, sR = ly.sR, sD = ly.sD, ldG = ly.ldG, GS = nv * ly.ldG;
    T* const G = reinterpret_cast<T*>(smem_raw);
    T* const Hx = G + ly.oHx;
    T* const Hy = G + ly.oHy;
    T* const Tm = G + ly.oTm;
    T* const D = G + ly.oD;
    T* const R = Tm;       // Tm is dead once D is formed
    T* const F = G;        // G is dead once Tm is formed
    double* const TX = smem_raw + ly.endT * (int)sizeof(T) / 8;
    double* const TY = TX + MF_REC * nv;

    // ---- per-axis records and the Hermite bases He_0..He_15 at
    // xi = (u-U)/sqrt(rt), eta = (v-V)/sqrt(rt) (build_probabilists_hermite,
    // hermite_symmetric2d.f90:220-233); zero basis columns beyond nv
    const double rt = P[CP_RT];
    const double sqrt_rt = sqrt(rt);
    const double om_scale = dt / (2.0 * c.pi * sqrt_rt);
    for (int t = tid; t < 2 * np; t += NT) {
        const bool isx = t < np;
        const int idx = isx ? t : t - np;
        T* Hb = isx ? Hx : Hy;
        if (idx < nv) {
            const double xv = isx ? u1[idx] : v1[idx];
            const double co = xv - (isx ? po[1] : po[2]), cn = xv - (isx ? pn[1] : pn[2]);
            double* rec = (isx ? TX : TY) + idx;
            const double co2 = co * co, cn2 = cn * cn;
            rec[MR_E0 * nv] = (isx ? P[CP_HCO] : 1.0) * exp(-po[3] * co2);
            rec[MR_X0 * nv] = 2.0 * po[3] * co2 + (isx ? c.K - 5.0 : 0.0);
            rec[MR_E1 * nv] = (isx ? P[CP_HCN] : 1.0) * exp(-pn[3] * cn2);
            rec[MR_X1 * nv] = 2.0 * pn[3] * cn2 + (isx ? c.K - 5.0 : 0.0);
            rec[MR_CQ0 * nv] = co;
            const A x = (A)(co / sqrt_rt);
            rec[MR_OM * nv] = (isx ? om_scale : 1.0) * (double)fexp((A)-0.5 * x * x);
            A p0 = (A)1, p1 = x;
            Hb[idx] = (T)p0;
            Hb[sH + idx] = (T)p1;
            for (int n = 2; n < MF_NB; ++n) {
                const A p2 = x * p1 - (A)(n - 1) * p0;
                Hb[n * sH + idx] = (T)p2;
                p0 = p1;
                p1 = p2;
            }
        } else {
            for (int n = 0; n < MF_NB; ++n) Hb[n * sH + idx] = (T)0;
        }
    }
    __syncthreads();

    // ---- load f^n, weighted copies, heat flux at t^n (solver.f90:337),
    // FP64 whatever the storage type.  Node (iu, jv) of k = iu + nv*jv
    // without an integer division: (k + 1/2)/nv stays at least 1/(2 nv)
    // away from an integer.
    const float rnv = 1.0f / (float)nv;
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
            const int jv = (int)(((float)k + 0.5f) * rnv), iu = k - jv * nv;
            const double wt = wnode_d[k];
            const double cu = TX[MR_CQ0 * nv + iu], cv = TY[MR_CQ0 * nv + jv];
            const double c2 = cu * cu + cv * cv;
            const double wh = wt * (double)rh[e], wb = wt * (double)rb[e];
            q4[0] += cu * c2 * wh;
            q4[1] += cu * wb;
            q4[2] += cv * c2 * wh;
            q4[3] += cv * wb;
            const A wa = (A)wt;   // the A-type weight table holds (A)wnode_d
            G[jv * ldG + iu] = (T)(wa * (A)rh[e]);
            G[GS + jv * ldG + iu] = (T)(wa * (A)rb[e]);
        }
    }
    block_sum<4, NT>(q4, red);   // also orders the LDS writes above
    const double qf0 = 0.5 * (q4[0] + q4[1]), qf1 = 0.5 * (q4[2] + q4[3]);
    // Shakhov projections (read after the barriers of the contractions)
    for (int t = tid; t < 2 * nv; t += NT) {
        const bool isx = t < nv;
        const int idx = isx ? t : t - nv;
        double* rec = (isx ? TX : TY) + idx;
        const double qa = isx ? qf0 : qf1;
        const double cn = (isx ? u1[idx] : v1[idx]) - (isx ? pn[1] : pn[2]);
        rec[MR_CQ0 * nv] = P[CP_CSO] * qa * rec[MR_CQ0 * nv];
        rec[MR_Q1 * nv] = P[CP_CSN] * qa * cn;
    }

    const int nt = np / 16;             // tiles along a velocity index
    const int ksn = (nv + 3) / 4;       // K steps over a velocity index
    const int lr = lane & 15, lk = lane >> 4;

    // ---- Tm = Hx G: tile (f, jt), rows r, columns jv
    for (int job = wv; job < 2 * nt; job += NT / WAVE) {
        const int f = job / nt, jt = job % nt;
        const int jv = 16 * jt + lr;
        V acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < ksn; ++ks) {
            const int iu = 4 * ks + lk;
            const bool in = iu < nv;
            const T a = in ? Hx[lr * sH + iu] : (T)0;
            const T bb = (in && jv < nv) ? G[f * GS + jv * ldG + iu] : (T)0;
            acc = M::mma(a, bb, acc);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) Tm[(f * MF_NB + M::row(lane, v)) * sH + jv] = acc[v];
    }
    __syncthreads();
    // ---- D = Tm Hy^T / (r! s!), r + s <= Nm
    const double jac = 1.0 / rt;
    for (int f = wv; f < 2; f += NT / WAVE) {
        V acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < ksn; ++ks) {
            const int jv = 4 * ks + lk;
            const T a = Tm[(f * MF_NB + lr) * sH + jv];     // zero beyond nv
            const T bb = Hy[lr * sH + jv];
            acc = M::mma(a, bb, acc);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = M::row(lane, v), s = lr;
            T val = (T)0;
            if (r + s <= nm) val = (T)((A)(jac / (kFact[r] * kFact[s])) * (A)acc[v]);
            D[(f * MF_NB + r) * sD + s] = val;
        }
    }
    __syncthreads();
    // ---- R(r, jv) = ax sum_s D(r-1, s) Hy(s, jv) + ay sum_s D(r, s) Hy(s+1, jv)
    const A axA = (A)ax, ayA = (A)ay;
    for (int job = wv; job < 2 * nt; job += NT / WAVE) {
        const int f = job / nt, jt = job % nt;
        const int jv = 16 * jt + lr;
        V acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < MF_NB / 4; ++ks) {
            const int s = 4 * ks + lk;
            const T a1 = lr >= 1 ? (T)(axA * (A)D[(f * MF_NB + lr - 1) * sD + s]) : (T)0;
            acc = M::mma(a1, Hy[s * sH + jv], acc);
            const T a2 = (T)(ayA * (A)D[(f * MF_NB + lr) * sD + s]);
            const T b2 = s + 1 < MF_NB ? Hy[(s + 1) * sH + jv] : (T)0;
            acc = M::mma(a2, b2, acc);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) R[(f * MF_NB + M::row(lane, v)) * sR + jv] = acc[v];
    }
    __syncthreads();
    // ---- F = Hx^T R, into the G region (G is dead)
    for (int job = wv; job < 2 * nt * nt; job += NT / WAVE) {
        const int f = job / (nt * nt), rem = job % (nt * nt);
        const int it = rem / nt, jt = rem % nt;
        const int iu_a = 16 * it + lr, jv = 16 * jt + lr;
        V acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < MF_NB / 4; ++ks) {
            const int r = 4 * ks + lk;
            acc = M::mma(Hx[r * sH + iu_a], R[(f * MF_NB + r) * sR + jv], acc);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int iu = 16 * it + M::row(lane, v);
            if (iu < nv && jv < nv) F[f * nn + iu + nv * jv] = acc[v];
        }
    }
    __syncthreads();

    // ---- force and implicit BGK update (solver.f90:337-388):
    //   H' = H (1 + cq (2 lam |c|^2 + K - 5)),  B' = kf (H' + 2 cq H),
    //   h^{n+1} = (h + div + dt/2 (H'^{n+1}/tau + (H'^n - h)/tau^n) + dt F) / (1 + dt/(2 tau))
    const double kfo = P[CP_KFO], kfn = P[CP_KFN];
    const double a_n = 0.5 * dt * P[CP_ITAU], a_o = 0.5 * dt * P[CP_ITAUO];
    const double iden = P[CP_IDEN];
    const A iarea = (A)(1.0 / area);
    const T* fh_l = vfh + (size_t)vl * c.ldv;
    const T* fh_r = vfh + (size_t)vr * c.ldv;
    const T* fh_d = hfh + (size_t)hd * c.ldv;
    const T* fh_u = hfh + (size_t)hu * c.ldv;
    const T* fb_l = vfb + (size_t)vl * c.ldv;
    const T* fb_r = vfb + (size_t)vr * c.ldv;
    const T* fb_d = hfb + (size_t)hd * c.ldv;
    const T* fb_u = hfb + (size_t)hu * c.ldv;
    T* ho_p = h + (size_t)cell * c.ldv;
    T* bo_p = b + (size_t)cell * c.ldv;
    // the face-flux divergence of every node first: all its loads are in
    // flight before the node arithmetic waits on the first of them
    A dvh[EPT], dvb[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = tid + e * NT;
        dvh[e] = dvb[e] = (A)0;
        if (k < nn) {
            dvh[e] = ((A)fh_l[k] - (A)fh_r[k] + (A)fh_d[k] - (A)fh_u[k]) * iarea;
            dvb[e] = ((A)fb_l[k] - (A)fb_r[k] + (A)fb_d[k] - (A)fb_u[k]) * iarea;
        }
    }
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int k = tid + e * NT;
        if (k < nn) {
            const int jv = (int)(((float)k + 0.5f) * rnv), iu = k - jv * nv;
            const double* rx = TX + iu;
            const double* ry = TY + jv;
            const double Ho = rx[MR_E0 * nv] * ry[MR_E0 * nv];
            const double cqo = rx[MR_CQ0 * nv] + ry[MR_CQ0 * nv];
            const double Hpo = fma(Ho, cqo * (rx[MR_X0 * nv] + ry[MR_X0 * nv]), Ho);
            const double Bpo = kfo * fma(2.0 * cqo, Ho, Hpo);
            const double Hn = rx[MR_E1 * nv] * ry[MR_E1 * nv];
            const double cqn = rx[MR_Q1 * nv] + ry[MR_Q1 * nv];
            const double Hpn = fma(Hn, cqn * (rx[MR_X1 * nv] + ry[MR_X1 * nv]), Hn);
            const double Bpn = kfn * fma(2.0 * cqn, Hn, Hpn);
            const A om = (A)(rx[MR_OM * nv] * ry[MR_OM * nv]);
            const double gh = (double)(om * (A)F[k]), gb = (double)(om * (A)F[nn + k]);
            const double dfh = (double)dvh[e], dfb = (double)dvb[e];
            const double ho = (double)rh[e], bo = (double)rb[e];
            const double hn = (ho + dfh + a_n * Hpn + a_o * (Hpo - ho) + gh) * iden;
            const double bn = (bo + dfb + a_n * Bpn + a_o * (Bpo - bo) + gb) * iden;
            // streamed out: the next step reads them long after
            __builtin_nontemporal_store((T)hn, ho_p + k);
            __builtin_nontemporal_store((T)bn, bo_p + k);
        }
    }
}
// FP32 storage: at least five wavefronts per SIMD (the LDS of a cell allows
// five workgroups per CU at nv = 40).
template <typename T, int EPT, int NT, int NVC>
__global__ __launch_bounds__(NT) HUGKS_WAVES_PER_EU(sizeof(T) == 4 ? 5 : 1)
void k_update2d_mfma(
    Consts c, const DevState* __restrict__ st, const double* __restrict__ cp,
    T* __restrict__ h, T* __restrict__ b, const T* __restrict__ vfh, const T* __restrict__ vfb,
    const T* __restrict__ hfh, const T* __restrict__ hfb,
    const typename Arith<T>::A* __restrict__ wnode, const double* __restrict__ wnode_d,
    const double* __restrict__ u1, const double* __restrict__ v1) {
    k_update2d_mfma_body<T, EPT, NT, NVC>(c, st, cp, h, b, vfh, vfb, hfh, hfb, wnode, wnode_d, u1,
        v1, (int)blockIdx.x, (int)gridDim.x);
}

}  // namespace hugks
