// 2D cell update with the whole node update on MFMA (gfx950).
//
// Same step as k_update2d.h - update(), 2d/src/solver.f90:245-395, with
// symmetric_hermite_force_term_2d, 2d/src/hermite_symmetric2d.f90:23-104,
// and Shakhov-BGK with the trapezoidal collision (solver.f90:337-388) - with
// everything node-wise except the face-flux divergence moved onto 16x16x4
// matrix tiles (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64):
//
//   C  = Hx G Hy^T          Hermite coefficients, G = weight * f
//   D  = C / r!s! / rt      zero where r + s > Nm
//   R  = (ax D1 Hy + ay D Hy1) o OMy
//   F^T(jv, iu) = sum_r R(r, jv) Ho(r, iu) + sum_t Rc(t, jv) L(t, iu)
//
// with Ho(r, iu) = He_r(xi_iu) OMx(iu) dt/(2 pi sqrt(rt)) / (1 + dt/2tau):
// the force weight exp(-(xi^2 + eta^2)/2) factors over the two velocity
// axes, and so do the collision terms: with the per-axis Maxwellian factors
// ex(u) ey(v), the Shakhov projections cq = qx(u) + qy(v) and the
// polynomial 2 lam |c|^2 + K - 5 = x(u) + y(v),
//   H' = H (1 + cq (x + y))
//      = ex(1 + qx x) (x) ey + ex qx (x) ey y + ex x (x) ey qy + ex (x) ey qy y,
// four outer products per time level (and B' = kf (H' + 2 cq H) the same
// with x + 2 in place of x).  So the eight rows t of L (per distribution)
// and Rc (shared) carry a_n H'^{n+1} + a_o H'^n, scaled by 1/(1 + dt/2tau)
// like Ho, and the accumulators of the F^T tiles hold
//   (a_n H'^{n+1} + a_o H'^n + dt F) / (1 + dt/2tau)
// for the lane's nodes; the update is then
//   h^{n+1} = (1 - a_o)/(1 + a_n) h^n + div/(area (1 + a_n)) + acc.
// The heat flux at t^n (solver.f90:337) is a combination of the Hermite
// coefficients of order <= 3 (xi^3 = He3 + 3 He1, xi eta^2 = He1 He2 + He1):
//   qx = (rt^{3/2} (C30 + C12 + 4 C10) + rt^{1/2} C10[b]) / 2.
//
// Node (iu, jv) of an F^T tile sits in the accumulator element of the lane
// that holds it.  With 4-byte operands f^n and b^n are read once, at the
// start, in that layout and kept in registers (G = w f staged in LDS from
// them); with 8-byte ones G is formed in registers for each K step of Tm
// and f^n, b^n of a tile position are read again (from L2) with its face
// fluxes.  The eight face fluxes of a tile position are read before its
// tiles (the first position's during the R contraction), and h^{n+1},
// b^{n+1} are written from the accumulators.  One workgroup of NW
// wavefronts per cell; wavefront w owns the tile positions w, w + NW, ...
// (MAXP at most).
//
// The face-forming instances (FF; k_update2d_mxf) take the four faces'
// equilibrium node fluxes as 16 more K rows of the F^T tiles, built after
// the D contraction from each face's coefficients, and read the rest of
// the face divergence per node from the per-cell arrays of k_fdiv (two
// values instead of eight face fluxes).
#pragma once
#include "k_update2d_mfma.h"

namespace hugks {

// 1/r! for r = 0..15
static __constant__ const double kInvFact[MF_NB] = {
    1.0, 1.0, 1.0 / 2.0, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0, 1.0 / 720.0, 1.0 / 5040.0,
    1.0 / 40320.0, 1.0 / 362880.0, 1.0 / 3628800.0, 1.0 / 39916800.0, 1.0 / 479001600.0,
    1.0 / 6227020800.0, 1.0 / 87178291200.0, 1.0 / 1307674368000.0};

// LDS layout of one cell (elements of the tile operand type).  Row strides
// as in k_update2d_mfma.h: a tile operand read as 16 rows x 4 columns needs
// 4 x odd (Hx, Hy, Tm, D, G), one read as 4 rows x 16 columns an odd
// multiple of 16 (R, Rc).
//   region 0: G [2][nv][ldG] (weight * f; staged only for 4-byte operands),
//             later D [2][16][sD], L [2][8][sH] and Rc [8][sR]
//   Hx, Ho, Hy [16][sH]
//   region T: Tm [2][16][sH], later R [2][16][sR]
//   OMy [nv]
// G, the other operand of Tm = Hx G, is staged for 4-byte operands (FP32
// tiles): five cells of 31 KB per CU at nv = 40.  With 8-byte operands
// (FP64, bf16x3 pieces) LDS bounds the cells per CU, so G is formed in
// registers from global memory instead: 48 instead of 62 KB per cell at
// nv = 40, three cells per CU instead of two (FP64 9.40 -> 8.55 ms, bf16x3
// 8.27 -> 6.93 ms at 512^2 cells, GH40; FP32 that way: 4.41 -> 5.85 ms,
// its twenty scattered loads per tile job outweigh the occupancy).
struct MxLayout {
    int np, sH, sR, sD, ldG;
    int oD, oL, oRc, oFA, oFB, oHx, oHo, oHy, oT, oOM, end;
};
// With the interior face fluxes on the tiles (ff), the face rows (k_update2d_mx_body)
// take 24 rows of stride sR and 8 of stride sH after Rc, inside G's region where
// G is staged (G is dead once Tm is formed), and 16 rows of stride sH in Hx's
// place (Hx is dead then too).
__host__ __device__ inline MxLayout mx_layout(int nv, bool stage_g, bool ff = false) {
    MxLayout L;
    L.np = mf_np(nv);
    L.sH = L.np + 4;
    L.sR = ((L.np / 16) & 1) ? L.np : L.np + 16;
    L.sD = MF_NB + 4;
    L.ldG = (nv + 3) / 4 * 4;
    if (((L.ldG / 4) & 1) == 0) L.ldG += 4;
    L.oD = 0;
    L.oL = L.oD + 2 * MF_NB * L.sD;
    L.oRc = L.oL + 2 * 8 * L.sH;
    int r0 = L.oRc + 8 * L.sR;
    L.oFA = L.oFB = r0;
    if (ff) {
        L.oFB = r0 + 24 * L.sR;
        r0 = L.oFB + 8 * L.sH;
    }
    const int g = stage_g ? 2 * nv * L.ldG : 0;
    L.oHx = g > r0 ? g : r0;
    L.oHo = L.oHx + MF_NB * L.sH;
    L.oHy = L.oHo + MF_NB * L.sH;
    L.oT = L.oHy + MF_NB * L.sH;
    L.oOM = L.oT + 2 * MF_NB * (L.sH > L.sR ? L.sH : L.sR);
    L.end = L.oOM + nv;
    return L;
}
// G is staged exactly when the operand element is 4 bytes
__host__ __device__ inline size_t mx_lds_bytes(int nv, size_t tsize, bool ff = false) {
    return (size_t)mx_layout(nv, tsize == 4, ff).end * tsize;
}

// Compiled instances (wavefronts per cell, tile positions per wavefront,
// compiled-in nv or 0 for any nv <= 48); the bf16x3 contraction has the
// run-time-nv ones and those of the 28- and 40-node grids.
#define HUGKS_UPDX(X) X(1, 1, 16) X(4, 1, 20) X(2, 2, 20) X(4, 1, 24) X(2, 2, 24) \
    X(4, 1, 28) X(2, 2, 28) X(4, 1, 32) X(2, 2, 32) X(3, 3, 40) X(5, 2, 40) X(9, 1, 40) \
    X(1, 1, 0) X(4, 1, 0) X(2, 2, 0) X(3, 3, 0) X(5, 2, 0)
#define HUGKS_UPDX_B3(X) X(4, 1, 28) X(2, 2, 28) X(3, 3, 40) X(5, 2, 40) \
    X(1, 1, 0) X(4, 1, 0) X(2, 2, 0) X(3, 3, 0) X(5, 2, 0)

// The wavefronts per cell a case runs with by default: 3 for nv in
// 33..48 (nine tile positions, three per wavefront), 4 for 17..32, 1 up to
// 16.
__host__ __device__ inline int mx_default_nw(int nv) {
    const int nt = mf_np(nv) / 16;
    return nt >= 3 ? 3 : (nt == 2 ? 4 : 1);
}
__host__ __device__ inline int mx_maxp(int nv, int nw) {
    const int nt = mf_np(nv) / 16;
    return (nt * nt + nw - 1) / nw;
}
// The instance (nw, maxp, nvc) for nv and nw; nvc = 0 when nv has no
// compiled-in instance.  Returns false when no instance runs nw at nv.
__host__ inline bool mx_key(int nv, int nw, int* maxp, int* nvc, bool b3 = false) {
    const int mp = mx_maxp(nv, nw);
    bool in = false, any = false;
#define HUGKS_UPDX_HAS(W, P, V)                  \
    if (W == nw && P == mp && V == nv) in = true; \
    if (W == nw && P == mp && V == 0) any = true;
    if (b3) {
        HUGKS_UPDX_B3(HUGKS_UPDX_HAS)
    } else {
        HUGKS_UPDX(HUGKS_UPDX_HAS)
    }
#undef HUGKS_UPDX_HAS
    if (!in && !any) return false;
    *maxp = mp;
    *nvc = in ? nv : 0;
    return true;
}

// bf16x3 split-product operands.  An FP32 value is split exactly into three
// bf16 pieces, a = hi + mid + lo (hi = bf16(a), mid = bf16(a - hi), lo =
// a - hi - mid, which has at most 8 significant bits; all rounded to
// nearest), kept as two dwords [hi | mid << 16, lo].  One
// v_mfma_f32_16x16x32_bf16 then does one 16x16x4 step: the eight K slots
// of a lane hold the six cross products of its K index, A slots
// [h m | h m | l 0 | h m] against B slots [h h | m m | h m | l 0], i.e.
// hi hi, mid hi, hi mid, mid mid, lo hi, hi lo; the three products dropped
// (mid lo, lo mid, lo lo) are below 2^-23 |a b|, and the products are exact
// in the FP32 accumulation.  The A operand is the stored pair as it is, the
// B operand two byte permutes of it.
struct B3 {
    uint32_t x, y;
};
__device__ __forceinline__ B3 b3_split(float a) {
    const __bf16 hi = (__bf16)a;
    const float r1 = a - (float)hi;
    const __bf16 mid = (__bf16)r1;
    const __bf16 lo = (__bf16)(r1 - (float)mid);
    return B3{(uint32_t)__builtin_bit_cast(uint16_t, hi) |
                  ((uint32_t)__builtin_bit_cast(uint16_t, mid) << 16),
              (uint32_t)__builtin_bit_cast(uint16_t, lo)};
}
struct MfmaB3 {
    using V = mf_f4;
    static __device__ __forceinline__ int row(int l, int v) { return Mfma<float>::row(l, v); }
    static __device__ __forceinline__ V mma(B3 a, B3 b, V c) {
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
        const u4 av = {a.x, a.x, a.y, a.x};
        const u4 bv = {__builtin_amdgcn_perm(b.x, b.x, 0x05040100u),
                       __builtin_amdgcn_perm(b.x, b.x, 0x07060302u), b.x, b.y};
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, av),
                                                       __builtin_bit_cast(bf8, bv), c, 0, 0, 0);
    }
};
// The LDS element and the tile product of a storage type: T itself on
// 16x16x4 tiles, or (FP32 storage, SB3) the bf16x3 pieces.
template <typename T, bool SB3> struct MxOps {
    using E = T;
    using M = Mfma<T>;
    static __device__ __forceinline__ E put(typename Arith<T>::A x) { return (T)x; }
    static __device__ __forceinline__ E zero() { return (T)0; }
};
template <> struct MxOps<float, true> {
    using E = B3;
    using M = MfmaB3;
    static __device__ __forceinline__ E put(float x) { return b3_split(x); }
    static __device__ __forceinline__ E zero() { return B3{0u, 0u}; }
};

// The value, opaque to the optimizer: keeps a conversion of it where it is
// written instead of next to the load that produced the value (where its
// wider result would stay live across the kernel's phases).
template <typename X>
__device__ __forceinline__ X mx_opaque(X x) {
    __asm__ volatile("" : "+v"(x));
    return x;
}

// Inputs of the face fluxes when the update forms their equilibrium part
// itself (FF): the per-cell divergence of the rest (free transport of the
// interior faces, whole node fluxes of the wall faces; k_fdiv), the
// per-face coefficients of the flux pipeline (FC_*, indexed like its
// interior faces) and the face accelerations.
struct MxFace {
    const void *dvh, *dvb;   // [ncell][ldv]
    const double *fcoef, *acc_v, *acc_h;
};

// Slots of a face's factors in the FF table (node arithmetic): force shift,
// face state, Mt(1..3), K/(2 lam), (K^2 + 2K)/(4 lam^2), the Shakhov factor,
// heat flux, tau Mt(2) fn and ft, the normal rows' prefactor (divergence
// sign x 1/(area (1 + dt/2tau)) x face length x rho lam/pi), then the
// micro-slope coefficients a0, a1, a2, a3/2 of the tangential slope, the
// time slope, and the left and right spatial slopes.
enum : int {
    FR_VS, FR_U, FR_V, FR_LAM, FR_MT1, FR_MT2, FR_MT3, FR_KFAC, FR_MXI2, FR_CSA, FR_Q0, FR_Q1,
    FR_TM2FN, FR_TM2FT, FR_PRE, FR_PB, FR_PT = FR_PB + 4, FR_SL = FR_PT + 4, FR_SR = FR_SL + 4,
    FR_N = FR_SR + 4
};

template <typename T, int NW, int MAXP, int NVC, bool SB3 = false, bool FF = false>
__device__ __forceinline__ void k_update2d_mx_body(
    const Consts& c, const DevState* __restrict__ st, const double* __restrict__ cp,
    T* __restrict__ h, T* __restrict__ b, const T* __restrict__ vfh, const T* __restrict__ vfb,
    const T* __restrict__ hfh, const T* __restrict__ hfb,
    const typename Arith<T>::A* __restrict__ wnode, const double* __restrict__ u1,
    const double* __restrict__ v1, int bid, int nblk, MxFace fx = {}) {
    using A = typename Arith<T>::A;
    using O = MxOps<T, SB3>;
    using E = typename O::E;
    using M = typename O::M;
    using V = typename M::V;
    constexpr int NT = NW * WAVE;
    HIP_DYNAMIC_SHARED(double, smem_raw)
    // unscaled coefficients C10, C30, C12, C01, C03, C21 of h, C10, C01 of b
    __shared__ double qm[8];
    // FF: the node-independent factors of each face's rows (left, right,
    // down, up; FR_* slots), in the node arithmetic
    __shared__ typename Arith<T>::A fr[4][FR_N];
    if (st->stopped) return;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
    const int cell = tiled_cell((int)xcd_block(bid, nblk), c.nx, c.ny);
    const int nx = c.nx, nm = c.nm;
    const int nv = NVC ? NVC : c.nv;
    const int np = mf_np(nv), nt = np / 16, npos = nt * nt;
    const int i = cell % nx, j = cell / nx;
    const double dt = st->dt;
    const double* P = cp + (size_t)CP_LD * cell;
    double po[4], pn[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        po[q] = P[CP_PO + q];
        pn[q] = P[CP_PN + q];
    }
    const double iden = P[CP_IDEN];
    const double rt = P[CP_RT];
    const double sqrt_rt = sqrt(rt);

    constexpr bool STAGE_G = sizeof(E) == 4;
    const MxLayout ly = mx_layout(nv, STAGE_G, FF);
    const int sH = ly.sH, sR = ly.sR, sD = ly.sD, ldG = ly.ldG, GS = nv * ly.ldG;
    E* const sm = reinterpret_cast<E*>(smem_raw);
    E* const G = sm;               // staged G (4-byte operands)
    E* const D = sm + ly.oD;       // G is dead once Tm is formed
    E* const Lr = sm + ly.oL;
    E* const Rc = sm + ly.oRc;
    E* const Hx = sm + ly.oHx;
    E* const Ho = sm + ly.oHo;
    E* const Hy = sm + ly.oHy;
    E* const Tm = sm + ly.oT;
    E* const R = Tm;               // Tm is dead once D is formed
    A* const OMy = reinterpret_cast<A*>(sm + ly.oOM);

    // ---- node offsets.  Loads and stores are raw buffer accesses on one
    // resource per array row with one 32-bit byte offset per node, shared
    // by all the arrays: a node outside the grid gets an offset past the
    // row, which reads zero and drops a store, so no access sits under a
    // branch (a load under a branch is waited for inside it, which
    // serialises the loads).  The lane's nodes of tile position pi are
    // those of the F^T accumulators: iu = 16 it + lane % 16, jv = 16 jt +
    // row(lane, v).
    const int vl = i + (nx + 1) * j, vr = vl + 1;
    const int hd = i + nx * j, hu = i + nx * (j + 1);
    const size_t ldv = (size_t)c.ldv;
    const uint32_t rowb = (uint32_t)(nv * nv * (int)sizeof(T));
    constexpr uint32_t OUT = 0x80000000u;   // past any row
    const __amdgpu_buffer_rsrc_t rsh = buf_rsrc(h + (size_t)cell * ldv, rowb);
    const __amdgpu_buffer_rsrc_t rsb = buf_rsrc(b + (size_t)cell * ldv, rowb);
    const __amdgpu_buffer_rsrc_t rsw = buf_rsrc(wnode, rowb);
    const int lc = lane & 15;
    auto node_off = [&](int pi, int v) -> uint32_t {
        const int p = wv + NW * pi;
        const int iu = 16 * (p % nt) + lc, jv = 16 * (p / nt) + M::row(lane, v);
        return (p < npos && iu < nv && jv < nv) ? (uint32_t)((iu + nv * jv) * (int)sizeof(T))
                                                : OUT;
    };

    // ---- FF: the cell's four faces (0 left, 1 right, 2 down, 3 up).  An
    // interior face (not a wall) has the slot g of the flux pipeline's
    // per-face scratch (its walk order, face_geo).
    const int ny = c.ny;
    const bool xper = c.bc[0] == HUGKS_BC_PERIODIC, yper = c.bc[2] == HUGKS_BC_PERIODIC;
    const bool inL = xper || i > 0, inR = xper || i < nx - 1;
    const bool inD = yper || j > 0, inU = yper || j < ny - 1;
    auto face_pick = [](int fa, auto a0, auto a1, auto a2, auto a3) {
        return fa == 0 ? a0 : fa == 1 ? a1 : fa == 2 ? a2 : a3;
    };
    if constexpr (FF) {
        if (tid < 4) {
            // the face's state, time coefficients, slopes and force, with the
            // divergence sign, 1/(area (1 + dt/2tau)) and the face length
            // folded into the prefactor of its normal rows
            const int fa = tid;
            const bool isx = fa < 2;
            const bool in = face_pick(fa, inL, inR, inD, inU);
            A k[FR_N];
#pragma unroll
            for (int q = 0; q < FR_N; ++q) k[q] = (A)0;
            if (in) {
                const int nf = xper ? nx + 1 : nx - 1, nr = yper ? ny + 1 : ny - 1;
                const int ox = xper ? 0 : 1, oy = yper ? 0 : 1;
                const int g = face_pick(fa, tiled_index(i - ox, j, nf, ny),
                                        tiled_index(i + 1 - ox, j, nf, ny),
                                        nf * ny + tiled_index(i, j - oy, nx, nr),
                                        nf * ny + tiled_index(i, j + 1 - oy, nx, nr));
                const int slot = face_pick(fa, vl, vr, hd, hu);
                const double* fc = fx.fcoef + (size_t)g * FACE_COEF_LD;
                const double* ac = isx ? fx.acc_v : fx.acc_h;
                const double fn = ac[2 * slot], ft = ac[2 * slot + 1];
                const double lam = fc[FC_P + 3], rho = fc[FC_P];
                const double cdd = iden / (c.dx * c.dy);
                k[FR_VS] = (A)(0.5 * fn * dt);
                k[FR_U] = (A)fc[FC_P + 1];
                k[FR_V] = (A)fc[FC_P + 2];
                k[FR_LAM] = (A)lam;
                k[FR_MT1] = (A)fc[FC_MT];
                k[FR_MT2] = (A)fc[FC_MT + 1];
                k[FR_MT3] = (A)fc[FC_MT + 2];
                k[FR_KFAC] = (A)(c.K / (2.0 * lam));
                k[FR_MXI2] = (A)((c.K * c.K + 2.0 * c.K) / (4.0 * lam * lam));
                k[FR_CSA] = (A)(c.shk * (1.0 - c.prandtl) * lam * lam / rho);
                k[FR_Q0] = (A)fc[FC_Q];
                k[FR_Q1] = (A)fc[FC_Q + 1];
                k[FR_TM2FN] = (A)(fc[FC_TAU] * fc[FC_MT + 1] * fn);
                k[FR_TM2FT] = (A)(fc[FC_TAU] * fc[FC_MT + 1] * ft);
                k[FR_PRE] = (A)((fa & 1) ? -cdd : cdd) * (A)(isx ? c.dy : c.dx) *
                            (A)(rho * (lam / c.pi));
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    k[FR_PB + q] = (A)fc[FC_AB + q];
                    k[FR_PT + q] = (A)fc[FC_AT + q];
                    k[FR_SL + q] = (A)fc[FC_AL + q];
                    k[FR_SR + q] = (A)fc[FC_AR + q];
                }
                k[FR_PB + 3] = (A)(0.5 * fc[FC_AB + 3]);
                k[FR_PT + 3] = (A)(0.5 * fc[FC_AT + 3]);
                k[FR_SL + 3] = (A)(0.5 * fc[FC_AL + 3]);
                k[FR_SR + 3] = (A)(0.5 * fc[FC_AR + 3]);
            }
#pragma unroll
            for (int q = 0; q < FR_N; ++q) fr[fa][q] = k[q];
        }
    }

    // ---- 8-byte operands: the B operand of Tm = Hx G, G = w f, for job
    // (f, jt): node (iu = 4 ks + lk, jv = 16 jt + lr) of K step ks, loaded
    // and weighted in registers (zero outside the grid).  The wavefront's
    // first job is requested before the bases are built, so its loads are
    // in flight meanwhile.
    const int ksn = (nv + 3) / 4;       // K steps over a velocity index
    constexpr int KSN = STAGE_G ? 1 : NVC ? (NVC + 3) / 4 : (48 + 3) / 4;
    const int lr = lane & 15, lk = lane >> 4;
    A gf[KSN], gw[KSN];
    auto load_g = [&](int job) {
        const int f = job / nt, jt = job % nt;
        const int jv = 16 * jt + lr;
        const __amdgpu_buffer_rsrc_t rs = f ? rsb : rsh;
#pragma unroll
        for (int ks = 0; ks < KSN; ++ks) {
            const int iu = 4 * ks + lk;
            const bool in = ks < ksn && iu < nv && jv < nv;
            const uint32_t o = in ? (uint32_t)((iu + nv * jv) * (int)sizeof(T)) : OUT;
            gf[ks] = (A)buf_ld<T>(rs, o);
            gw[ks] = buf_ld<A>(rsw, in ? o / (uint32_t)sizeof(T) * (uint32_t)sizeof(A) : OUT);
        }
    };
    if constexpr (!STAGE_G)
        if (wv < 2 * nt) load_g(wv);

    // ---- Hermite bases He_0..He_15 at xi = (u-U)/sqrt(rt), eta = (v-V)/sqrt(rt)
    // (build_probabilists_hermite, hermite_symmetric2d.f90:220-233); Ho
    // carries the force weight of the u axis, dt/(2 pi sqrt(rt)) and
    // 1/(1 + dt/2tau); zero columns beyond nv
    {
        const A oscale = (A)(dt / (2.0 * c.pi * sqrt_rt) * iden);
        for (int t = tid; t < 2 * np; t += NT) {
            const bool isx = t < np;
            const int idx = isx ? t : t - np;
            if (idx < nv) {
                const double xv = isx ? u1[idx] : v1[idx];
                const A x = (A)((xv - (isx ? po[1] : po[2])) / sqrt_rt);
                const A om = fexp((A)-0.5 * x * x);
                const A ho = isx ? oscale * om : (A)0;
                E* Hb = isx ? Hx : Hy;
                A p0 = (A)1, p1 = x;
                Hb[idx] = O::put(p0);
                Hb[sH + idx] = O::put(p1);
                if (isx) {
                    Ho[idx] = O::put(ho * p0);
                    Ho[sH + idx] = O::put(ho * p1);
                } else {
                    OMy[idx] = om;
                }
                for (int n = 2; n < MF_NB; ++n) {
                    const A p2 = x * p1 - (A)(n - 1) * p0;
                    Hb[n * sH + idx] = O::put(p2);
                    if (isx) Ho[n * sH + idx] = O::put(ho * p2);
                    p0 = p1;
                    p1 = p2;
                }
            } else {
                E* Hb = isx ? Hx : Hy;
                for (int n = 0; n < MF_NB; ++n) {
                    Hb[n * sH + idx] = O::zero();
                    if (isx) Ho[n * sH + idx] = O::zero();
                }
            }
        }
    }

    // ---- 4-byte operands: f^n and b^n of the lane's nodes, in the layout
    // of the F^T accumulators, kept in registers for the update; weight * f
    // to LDS for the coefficients
    A rhs[STAGE_G ? MAXP : 1][4], rbs[STAGE_G ? MAXP : 1][4];
    if constexpr (STAGE_G) {
#pragma unroll
        for (int pi = 0; pi < MAXP; ++pi) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const uint32_t o = node_off(pi, v);
                rhs[pi][v] = (A)buf_ld<T>(rsh, o);
                rbs[pi][v] = (A)buf_ld<T>(rsb, o);
            }
        }
#pragma unroll
        for (int pi = 0; pi < MAXP; ++pi) {
            const int p = wv + NW * pi;
            const int iu = 16 * (p % nt) + lc;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const uint32_t o = node_off(pi, v);
                const int jv = 16 * (p / nt) + M::row(lane, v);
                if (o != OUT) {
                    const A wa = buf_ld<A>(rsw, o / (uint32_t)sizeof(T) * (uint32_t)sizeof(A));
                    G[jv * ldG + iu] = O::put(wa * rhs[pi][v]);
                    G[GS + jv * ldG + iu] = O::put(wa * rbs[pi][v]);
                }
            }
        }
    }

    // the eight face fluxes of one tile position's nodes (left/right,
    // down/up faces of h, then of b)
    const __amdgpu_buffer_rsrc_t rsf[8] = {
        buf_rsrc(vfh + (size_t)vl * ldv, rowb), buf_rsrc(vfh + (size_t)vr * ldv, rowb),
        buf_rsrc(hfh + (size_t)hd * ldv, rowb), buf_rsrc(hfh + (size_t)hu * ldv, rowb),
        buf_rsrc(vfb + (size_t)vl * ldv, rowb), buf_rsrc(vfb + (size_t)vr * ldv, rowb),
        buf_rsrc(hfb + (size_t)hd * ldv, rowb), buf_rsrc(hfb + (size_t)hu * ldv, rowb)};
    // FF: the cell's rows of the divergence arrays
    const __amdgpu_buffer_rsrc_t rsdh = buf_rsrc((const T*)fx.dvh + (size_t)cell * ldv, rowb);
    const __amdgpu_buffer_rsrc_t rsdb = buf_rsrc((const T*)fx.dvb + (size_t)cell * ldv, rowb);
    // one tile position's face fluxes and, for 8-byte operands, f^n, b^n
    // of its nodes (read a second time: the cell's rows are in L2 from the
    // Tm contraction)
    auto load_faces = [&](int pi, A (&fvo)[8][4], A (&rho)[4], A (&rbo)[4]) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const uint32_t o = node_off(pi, v);
            if constexpr (FF) {
                fvo[0][v] = (A)buf_ld<T>(rsdh, o);
                fvo[4][v] = (A)buf_ld<T>(rsdb, o);
            } else {
#pragma unroll
                for (int a = 0; a < 8; ++a) fvo[a][v] = (A)buf_ld<T>(rsf[a], o);
            }
            if constexpr (!STAGE_G) {
                rho[v] = (A)buf_ld<T>(rsh, o);
                rbo[v] = (A)buf_ld<T>(rsb, o);
            }
        }
    };
    A fv[8][4], rh[4], rb[4];
    if constexpr (STAGE_G && MAXP == 1) load_faces(0, fv, rh, rb);
    // FF with staged G (4-byte operands) and at most three tile positions
    // per wavefront: every position's divergence is requested in the
    // collision-rows phase, a whole contraction ahead of its use (4.44 ->
    // 4.14 ms at 512^2 cells, GH40; with 8-byte operands, and before the
    // face rows' factors moved to LDS, the registers it takes cost more
    // occupancy than it hides)
    constexpr bool PRE = FF && MAXP <= 3 && STAGE_G;
    A dvp[PRE ? MAXP : 1][2][4];
    __syncthreads();

    // ---- Tm = Hx G: tile (f, jt), rows r, columns jv
    for (int job = wv; job < 2 * nt; job += NW) {
        const int f = job / nt, jt = job % nt;
        const int jv = 16 * jt + lr;
        V acc = {0, 0, 0, 0};
        if constexpr (STAGE_G) {
#pragma unroll
            for (int ks = 0; ks < ksn; ++ks) {
                const int iu = 4 * ks + lk;
                const bool in = iu < nv;
                const E a = in ? Hx[lr * sH + iu] : O::zero();
                const E bb = (in && jv < nv) ? G[f * GS + jv * ldG + iu] : O::zero();
                acc = M::mma(a, bb, acc);
            }
        } else {
            if (job != wv) load_g(job);
#pragma unroll
            for (int ks = 0; ks < KSN; ++ks) {
                if (ks < ksn) {
                    const int iu = 4 * ks + lk;
                    const E a = iu < nv ? Hx[lr * sH + iu] : O::zero();
                    acc = M::mma(a, O::put(gw[ks] * gf[ks]), acc);
                }
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) Tm[(f * MF_NB + M::row(lane, v)) * sH + jv] = O::put(acc[v]);
    }
    __syncthreads();

    // ---- D = Tm Hy^T / (r! s! rt), r + s <= Nm; the coefficients of order
    // <= 3 that give the heat flux kept unscaled
    const double jac = 1.0 / rt;
    for (int f = wv; f < 2; f += NW) {
        V acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < ksn; ++ks) {
            const int jv = 4 * ks + lk;
            const E a = Tm[(f * MF_NB + lr) * sH + jv];     // zero beyond nv
            const E bb = Hy[lr * sH + jv];
            acc = M::mma(a, bb, acc);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = M::row(lane, v), s = lr;
            E val = O::zero();
            if (r + s <= nm) val = O::put((A)(jac * kInvFact[r] * kInvFact[s]) * (A)acc[v]);
            D[(f * MF_NB + r) * sD + s] = val;
            const int rs = 16 * r + s;
            int slot = -1;
            if (f == 0)
                slot = rs == 16 ? 0 : rs == 48 ? 1 : rs == 18 ? 2 : rs == 1 ? 3 : rs == 3 ? 4
                     : rs == 33 ? 5 : -1;
            else
                slot = rs == 16 ? 6 : rs == 1 ? 7 : -1;
            if (slot >= 0) qm[slot] = (double)acc[v];
        }
    }
    __syncthreads();

    // ---- heat flux at t^n from the coefficients, then the collision rows:
    // L (u axis, per distribution) and Rc (v axis, shared)
    {
        const double s1 = sqrt_rt, s3 = rt * sqrt_rt;
        const double qx = 0.5 * (s3 * (qm[1] + qm[2] + 4.0 * qm[0]) + s1 * qm[6]);
        const double qy = 0.5 * (s3 * (qm[4] + qm[5] + 4.0 * qm[3]) + s1 * qm[7]);
        const double an = 0.5 * dt * P[CP_ITAU] * iden, ao = 0.5 * dt * P[CP_ITAUO] * iden;
        const double kfo = P[CP_KFO], kfn = P[CP_KFN];
        const double cso = P[CP_CSO], csn = P[CP_CSN];
        for (int t = tid; t < 2 * np; t += NT) {
            const bool isx = t < np;
            const int idx = isx ? t : t - np;
            double r8[8] = {0, 0, 0, 0, 0, 0, 0, 0}, b8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (idx < nv) {
                const double xv = isx ? u1[idx] : v1[idx];
                const double co = xv - (isx ? po[1] : po[2]), cn = xv - (isx ? pn[1] : pn[2]);
                const double e0 = (isx ? P[CP_HCO] : 1.0) * exp(-po[3] * co * co);
                const double e1 = (isx ? P[CP_HCN] : 1.0) * exp(-pn[3] * cn * cn);
                const double k5 = isx ? c.K - 5.0 : 0.0;
                const double x0 = 2.0 * po[3] * co * co + k5, x1 = 2.0 * pn[3] * cn * cn + k5;
                const double q0 = cso * (isx ? qx : qy) * co, q1 = csn * (isx ? qx : qy) * cn;
                if (isx) {
                    r8[0] = an * e1 * (1.0 + q1 * x1);
                    r8[1] = an * e1 * q1;
                    r8[2] = an * e1 * x1;
                    r8[3] = an * e1;
                    r8[4] = ao * e0 * (1.0 + q0 * x0);
                    r8[5] = ao * e0 * q0;
                    r8[6] = ao * e0 * x0;
                    r8[7] = ao * e0;
                    b8[0] = kfn * an * e1 * (1.0 + q1 * (x1 + 2.0));
                    b8[1] = kfn * an * e1 * q1;
                    b8[2] = kfn * an * e1 * (x1 + 2.0);
                    b8[3] = kfn * an * e1;
                    b8[4] = kfo * ao * e0 * (1.0 + q0 * (x0 + 2.0));
                    b8[5] = kfo * ao * e0 * q0;
                    b8[6] = kfo * ao * e0 * (x0 + 2.0);
                    b8[7] = kfo * ao * e0;
                } else {
                    r8[0] = e1;
                    r8[1] = e1 * x1;
                    r8[2] = e1 * q1;
                    r8[3] = e1 * q1 * x1;
                    r8[4] = e0;
                    r8[5] = e0 * x0;
                    r8[6] = e0 * q0;
                    r8[7] = e0 * q0 * x0;
                }
            }
            if (isx) {
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    Lr[q * sH + idx] = O::put((A)r8[q]);
                    Lr[8 * sH + q * sH + idx] = O::put((A)b8[q]);
                }
            } else {
#pragma unroll
                for (int q = 0; q < 8; ++q) Rc[q * sR + idx] = O::put((A)r8[q]);
            }
        }
    }
    // ---- FF: the equilibrium part of the interior faces' node fluxes as
    // K rows of the F^T tiles.  On a face with normal velocity vn and
    // tangential vt (x faces: u, v; y faces: v, -u) that part of the
    // flux of h is l vn (rho lam/pi) exp(-lam (cu^2 + cv^2)) Q_h(vn, vt)
    // (flux.f90:207-236; cu = vn - U, cv = vt - V in the face's frame), and
    // Q_h is a cubic in cv whose coefficients are functions of vn alone
    // (its upwind side included):
    //   flux = sum_k N_k(vn) T_k(vt),  T_k = exp(-lam cv^2) cv^k, k = 0..3,
    // and the same for b with its own N_k.  With the divergence sign and
    // 1/(area (1 + dt/2tau)) folded into N_k, the x faces add 8 rows T_k(v)
    // (A operand, shared) x N_k(u) (B operand, per distribution) and the y
    // faces 8 rows N_k(v) (per distribution) x T_k(-u) (shared).
    E* const TxA = sm + ly.oFA;    // [8][sR]: x faces' T_k over v
    E* const NyA = TxA + 8 * sR;   // [2][8][sR]: y faces' N_k over v, h then b
    E* const TyB = sm + ly.oFB;    // [8][sH]: y faces' T_k over u
    E* const NxB = Hx;             // [2][8][sH]: x faces' N_k over u (Hx is dead)
    if constexpr (FF) {
        for (int t = tid; t < 4 * np; t += NT) {
            const int fa = t / np, idx = t - fa * np;
            const bool isx = fa < 2;
            const bool in = face_pick(fa, inL, inR, inD, inU);
            A nh[4] = {0, 0, 0, 0}, nb[4] = {0, 0, 0, 0}, tk[4] = {0, 0, 0, 0};
            if (in && idx < nv) {
                const A* k = fr[fa];
                const A U = k[FR_U], V = k[FR_V], lam = k[FR_LAM];
                const A mt1 = k[FR_MT1], mt2 = k[FR_MT2], mt3 = k[FR_MT3];
                const A K = (A)c.K;
                const A kfac = k[FR_KFAC], mxi2 = k[FR_MXI2], csA = k[FR_CSA];
                const A q0 = k[FR_Q0], q1 = k[FR_Q1];
                // normal node of this row index and its upwind side (as the
                // per-node flux decides it, in the node arithmetic)
                const A vn = (A)(isx ? u1[idx] : v1[idx]);
                const int d = upwind_delta(vn - k[FR_VS]);
                const A* as = k + (d ? FR_SL : FR_SR);
                const A s0 = as[0], s1 = as[1], s2 = as[2], s3 = as[3];
                const A s4 = s3 * mxi2;
                const A pb0 = k[FR_PB], pb1 = k[FR_PB + 1], pb2 = k[FR_PB + 2];
                const A pb3 = k[FR_PB + 3], pb4 = pb3 * mxi2;
                const A pt0 = k[FR_PT], pt1 = k[FR_PT + 1], pt2 = k[FR_PT + 2];
                const A pt3 = k[FR_PT + 3], pt4 = pt3 * mxi2;
                const A tm2fn = k[FR_TM2FN], tm2ft = k[FR_TM2FT];
                const A cu = vn - U;
                const A al = mt2 * vn - tm2fn, b0 = mt2 * V - tm2ft;
                const A a5 = (A)2 * lam * cu * cu + K - (A)5, a3 = a5 + (A)2;
                const A m2 = vn * vn + V * V;
                // h: the terms mt1 (H + H'), al g_h, be g_bh, mt3 g_Th by power of cv
                const A P0 = pb0 + pb1 * vn + pb2 * V + pb3 * (m2 + kfac);
                const A P1 = pb2 + (A)2 * pb3 * V;
                const A T0 = pt0 + pt1 * vn + pt2 * V + pt3 * (m2 + kfac);
                const A T1 = pt2 + (A)2 * pt3 * V;
                const A S0 = s0 + s1 * vn + s2 * V + s3 * (m2 + kfac);
                const A S1 = s2 + (A)2 * s3 * V;
                const A mq = mt1 * csA;
                const A qh[4] = {
                    mt1 * ((A)1 + csA * q0 * cu * a5) + al * S0 + b0 * P0 + mt3 * T0,
                    mq * q1 * a5 + al * S1 + b0 * P1 + mt2 * P0 + mt3 * T1,
                    mq * (A)2 * lam * q0 * cu + al * s3 + b0 * pb3 + mt2 * P1 + mt3 * pt3,
                    mq * (A)2 * lam * q1 + mt2 * pb3};
                // b: B = kfac H, B' with K - 3, g_b = B P + a3/2 mxi2 H
                const A P0b = kfac * (pb0 + pb1 * vn + pb2 * V + pb3 * m2) + pb4;
                const A T0b = kfac * (pt0 + pt1 * vn + pt2 * V + pt3 * m2) + pt4;
                const A S0b = kfac * (s0 + s1 * vn + s2 * V + s3 * m2) + s4;
                const A qb[4] = {
                    kfac * mt1 * ((A)1 + csA * q0 * cu * a3) + al * S0b + b0 * P0b + mt3 * T0b,
                    kfac * (mq * q1 * a3 + al * S1 + b0 * P1 + mt3 * T1) + mt2 * P0b,
                    kfac * (mq * (A)2 * lam * q0 * cu + al * s3 + b0 * pb3 + mt2 * P1 + mt3 * pt3),
                    kfac * (mq * (A)2 * lam * q1 + mt2 * pb3)};
                const A en = k[FR_PRE] * vn * fexp(-lam * cu * cu);
                // tangential node of this row index
                const A cv = (A)(isx ? v1[idx] : -u1[idx]) - V;
                const A ec = fexp(-lam * cv * cv);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    nh[k] = en * qh[k];
                    nb[k] = en * qb[k];
                }
                tk[0] = ec;
                tk[1] = ec * cv;
                tk[2] = tk[1] * cv;
                tk[3] = tk[2] * cv;
            }
            const int fr = 4 * (fa & 1);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (isx) {
                    NxB[(fr + k) * sH + idx] = O::put(nh[k]);
                    NxB[(8 + fr + k) * sH + idx] = O::put(nb[k]);
                    TxA[(fr + k) * sR + idx] = O::put(tk[k]);
                } else {
                    NyA[(fr + k) * sR + idx] = O::put(nh[k]);
                    NyA[(8 + fr + k) * sR + idx] = O::put(nb[k]);
                    TyB[(fr + k) * sH + idx] = O::put(tk[k]);
                }
            }
        }
    }
    // the first tile position's face fluxes (and f^n) in flight during the
    // R contraction; with one position per wavefront and staged G, from
    // the start
    if constexpr (PRE) {
#pragma unroll
        for (int pi = 0; pi < MAXP; ++pi) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const uint32_t o = node_off(pi, v);
                dvp[pi][0][v] = (A)buf_ld<T>(rsdh, o);
                dvp[pi][1][v] = (A)buf_ld<T>(rsdb, o);
            }
        }
    } else if constexpr (!(STAGE_G && MAXP == 1)) {
        load_faces(0, fv, rh, rb);
    }
    // ---- R(r, jv) = (ax sum_s D(r-1, s) Hy(s, jv) + ay sum_s D(r, s) Hy(s+1, jv)) OMy(jv)
    {
        const A axA = (A)P[CP_AX], ayA = (A)P[CP_AY];
        for (int job = wv; job < 2 * nt; job += NW) {
            const int f = job / nt, jt = job % nt;
            const int jv = 16 * jt + lr;
            const A om = jv < nv ? OMy[jv] : (A)0;
            if constexpr (SB3) {
                // the operands as stored; ax and ay applied to the two sums
                V ac1 = {0, 0, 0, 0}, ac2 = {0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < MF_NB / 4; ++ks) {
                    const int s = 4 * ks + lk;
                    const E a1 = lr >= 1 ? D[(f * MF_NB + lr - 1) * sD + s] : O::zero();
                    ac1 = M::mma(a1, Hy[s * sH + jv], ac1);
                    const E b2 = s + 1 < MF_NB ? Hy[(s + 1) * sH + jv] : O::zero();
                    ac2 = M::mma(D[(f * MF_NB + lr) * sD + s], b2, ac2);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    R[(f * MF_NB + M::row(lane, v)) * sR + jv] =
                        O::put((axA * ac1[v] + ayA * ac2[v]) * om);
            } else {
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
                for (int v = 0; v < 4; ++v)
                    R[(f * MF_NB + M::row(lane, v)) * sR + jv] = (T)((A)acc[v] * om);
            }
        }
    }
    __syncthreads();

    // ---- F^T tiles: force and collision terms in the accumulators, then
    // h^{n+1} = c1 h^n + cd div + acc (FP64)
    const double c1 = iden * (1.0 - 0.5 * dt * P[CP_ITAUO]);
    const double cd = iden / (c.dx * c.dy);
#pragma unroll
    for (int pi = 0; pi < MAXP; ++pi) {
        const int p = wv + NW * pi;
        if (p >= npos) continue;
        if constexpr (!PRE)
            if (pi > 0) load_faces(pi, fv, rh, rb);
        const int it = p % nt, jt = p / nt;
        const int iu = 16 * it + lc, jvb = 16 * jt;
        V acc[2];
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            V a = {0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < MF_NB / 4; ++ks) {
                const int r = 4 * ks + lk;
                a = M::mma(R[(f * MF_NB + r) * sR + jvb + lr], Ho[r * sH + iu], a);
            }
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                const int t = 4 * ks + lk;
                a = M::mma(Rc[t * sR + jvb + lr], Lr[(8 * f + t) * sH + iu], a);
            }
            if constexpr (FF) {
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    const int t = 4 * ks + lk;
                    a = M::mma(TxA[t * sR + jvb + lr], NxB[(8 * f + t) * sH + iu], a);
                    a = M::mma(NyA[(8 * f + t) * sR + jvb + lr], TyB[t * sH + iu], a);
                }
            }
            acc[f] = a;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const uint32_t o = node_off(pi, v);
            const A dh = PRE ? dvp[PRE ? pi : 0][0][v]
                       : FF  ? fv[0][v]
                             : fv[0][v] - fv[1][v] + fv[2][v] - fv[3][v];
            const A db = PRE ? dvp[PRE ? pi : 0][1][v]
                       : FF  ? fv[4][v]
                             : fv[4][v] - fv[5][v] + fv[6][v] - fv[7][v];
            const A ho = STAGE_G ? rhs[STAGE_G ? pi : 0][v] : rh[v];
            const A bo = STAGE_G ? rbs[STAGE_G ? pi : 0][v] : rb[v];
            const double hn = fma(c1, (double)mx_opaque(ho),
                                  fma(cd, (double)dh, (double)acc[0][v]));
            const double bn = fma(c1, (double)mx_opaque(bo),
                                  fma(cd, (double)db, (double)acc[1][v]));
            // streamed out (non-temporal): the next step reads them long after
            buf_st_nt<T>((T)hn, rsh, o);
            buf_st_nt<T>((T)bn, rsb, o);
        }
    }
}

#ifndef HUGKS_BODIES_ONLY
// FP32 storage: at least four wavefronts per SIMD (the LDS of a cell, 31 KB
// at nv = 40, allows five three-wavefront cells per CU).
template <typename T, int NW, int MAXP, int NVC, bool SB3 = false>
__global__ __launch_bounds__(NW * WAVE)
HUGKS_WAVES_PER_EU(sizeof(T) == 4 ? 3 : 1) void k_update2d_mx(
    Consts c, const DevState* __restrict__ st, const double* __restrict__ cp,
    T* __restrict__ h, T* __restrict__ b, const T* __restrict__ vfh, const T* __restrict__ vfb,
    const T* __restrict__ hfh, const T* __restrict__ hfb,
    const typename Arith<T>::A* __restrict__ wnode, const double* __restrict__ u1,
    const double* __restrict__ v1) {
    k_update2d_mx_body<T, NW, MAXP, NVC, SB3>(c, st, cp, h, b, vfh, vfb, hfh, hfb, wnode, u1,
                                              v1, (int)blockIdx.x, (int)gridDim.x);
}
// The same update with the equilibrium part of the interior faces' node
// fluxes formed on its tiles and the rest of the divergence read per cell
// (k_fdiv; no per-face node flux arrays).  FP32 storage: four wavefronts per
// SIMD where nv is compiled in (at most 128 VGPRs; the run-time-nv instances
// would spill there).
template <typename T, int NW, int MAXP, int NVC, bool SB3 = false>
__global__ __launch_bounds__(NW * WAVE)
HUGKS_WAVES_PER_EU(sizeof(T) == 4 ? (NVC ? 4 : 3) : 1) void k_update2d_mxf(
    Consts c, const DevState* __restrict__ st, const double* __restrict__ cp,
    T* __restrict__ h, T* __restrict__ b, const T* __restrict__ vfh, const T* __restrict__ vfb,
    const T* __restrict__ hfh, const T* __restrict__ hfb,
    const typename Arith<T>::A* __restrict__ wnode, const double* __restrict__ u1,
    const double* __restrict__ v1, MxFace fx) {
    k_update2d_mx_body<T, NW, MAXP, NVC, SB3, true>(c, st, cp, h, b, vfh, vfb, hfh, hfb, wnode,
                                                    u1, v1, (int)blockIdx.x, (int)gridDim.x,
                                                    fx);
}
#endif

}  // namespace hugks
