// Interior-face UGKS flux as a five-kernel pipeline (gfx950):
// calc_flux_mtforcev / calc_flux_mtforceh, 2d/src/flux.f90:28-472.
//
// Both face orientations go through each launch; the global interior face
// index g < nvi is a vertical face, g >= nvi a horizontal one (the branch
// is uniform per wavefront).
//
//   K1 k_fsum      one wavefront per face, one read of the upwind inputs
//                  (h, b and their normal/tangential slopes): the 37
//                  velocity sums that do not depend on the face state, in
//                  FP64 (flux.f90:68-102,116-120,184-202, and the raw
//                  moments the heat flux is expanded into)
//   K2 k_fstate    one lane per face: state at the face, tau, the time
//                  coefficients Mt, aL, aR, the tangential slope, the
//                  f0 part of the macro flux and the heat flux
//                  (flux.f90:93-147, get_heat_flux)
//   K3 k_fmaxw     one wavefront per face, no distribution reads: the
//                  Maxwellian sums of the force compatibility
//                  (flux.f90:691-705) and of the Shakhov terms;
//                  fmaxw_sep_lane computes the same sums from 1D moments
//                  when the quadrature is a tensor product
//   K4 k_fslope_t  one lane per face: time slope aT (flux.f90:126-135) and
//                  the conserved flux (flux.f90:152-246)
//   K5 k_fnode     one wavefront per face, second read of the inputs: the
//                  distribution fluxes per node (flux.f90:207-236)
//
// For a tensor-product quadrature K2, K3 and K4 run as one lane-per-face
// kernel, k_fface.
//
// The scalar moment algebra runs once per face on one lane instead of on
// every lane of the face's wavefront, and the distribution functions are
// read twice instead of three times.  Velocity sums are FP64 in a fixed
// order (bitwise reproducible), reduced over the wavefront by a
// reduce-scatter.
#pragma once
#include "hugks_device.h"
#include "hugks_kernels.h"

namespace hugks {

// The single-case kernels of this header are defined only where
// HUGKS_BODIES_ONLY is not set: the grouped-launch file (launch_group.hip)
// takes the device bodies alone, so no kernel is defined in two files.

constexpr int FSUM_N = 37;      // K1 sums per face
constexpr int FSUM_LD = FACE_SUM_LD;
constexpr int FCOEF_LD = FACE_COEF_LD;  // K2/K4 coefficients per face
constexpr int FMAXW_N = 13;     // K3 sums per face
constexpr int FMAXW_LD = FACE_MAXW_LD;
constexpr int SPLIT_FPW = 4;    // faces (wavefronts) per workgroup of K1/K3/K5

struct FaceGeo {
    int dir, face, cL, cR;
};

// Interior face g -> orientation, face slot and its two cells.  Each
// orientation's faces are walked in strips of TILE_W face columns, row by
// row inside a strip (tiled_cell on the face grid), so that with the
// XCD-contiguous workgroup order the cells around consecutive faces stay
// in one L2.
__device__ __forceinline__ FaceGeo face_geo(const Consts& c, int g, int nvi) {
    FaceGeo r;
    if (g < nvi) {
        const bool per = c.bc[0] == HUGKS_BC_PERIODIC;
        const int nf = per ? c.nx + 1 : c.nx - 1;
        const int t = tiled_cell(g, nf, c.ny);
        const int i = (per ? 0 : 1) + t % nf, j = t / nf;
        r.dir = 0;
        r.face = i + (c.nx + 1) * j;
        r.cL = (i - 1 + c.nx) % c.nx + c.nx * j;
        r.cR = i % c.nx + c.nx * j;
    } else {
        g -= nvi;
        const bool per = c.bc[2] == HUGKS_BC_PERIODIC;
        const int nr = per ? c.ny + 1 : c.ny - 1;
        const int t = tiled_cell(g, c.nx, nr);
        const int i = t % c.nx, j = (per ? 0 : 1) + t / c.nx;
        r.dir = 1;
        r.face = i + c.nx * j;
        r.cL = i + c.nx * ((j - 1 + c.ny) % c.ny);
        r.cR = i + c.nx * (j % c.ny);
    }
    return r;
}

// Slope stencils of the two cells of a face, as byte offsets of cell rows
// (side 1: the left/lower cell, upwind for d = 1; side 0: the right/upper
// cell): the cell itself, its two neighbours along the face normal and
// the two along the face, with the inner/one-sided choice of slope_nb.
struct Stencil {
    uint32_t u[2], nl[2], nr[2], tl[2], tr[2];
    bool inN[2], inT[2];
};
template <typename T>
__device__ __forceinline__ Stencil face_stencil(const Consts& c, const FaceGeo& fg) {
    Stencil S;
    const uint32_t rowB = (uint32_t)c.ldv * (uint32_t)sizeof(T);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int cell = side ? fg.cL : fg.cR;
        const SlopeNb n = slope_nb(c, cell, fg.dir), t = slope_nb(c, cell, 1 - fg.dir);
        S.u[side] = (uint32_t)cell * rowB;
        S.nl[side] = (uint32_t)n.cl * rowB;
        S.nr[side] = (uint32_t)n.cr * rowB;
        S.tl[side] = (uint32_t)t.cl * rowB;
        S.tr[side] = (uint32_t)t.cr * rowB;
        S.inN[side] = n.inner;
        S.inT[side] = t.inner;
    }
    return S;
}
template <typename T>
__device__ __forceinline__ T ld_b(const T* __restrict__ base, uint32_t boff) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}

// Upwind values of one node: f and its normal and tangential slopes for h
// and b, either read from the slope arrays or, with the stencil, formed
// from the five stencil cells exactly as k_slope2d forms them.
template <typename T, bool ST>
struct Upwind {
    using A = typename Arith<T>::A;
    const T *h, *b, *sN, *sT, *bN, *bT;
    size_t baseL, baseR;
    uint32_t rowL, rowR;   // byte offsets of the two cells' rows (arrays under 4 GiB)
    bool small;            // every distribution array is under 4 GiB
    __amdgpu_buffer_rsrc_t rh, rb, rsN, rsT, rbN, rbT;   // the six arrays, when small
    Stencil S;
    A idN, idT, smv;
    __device__ __forceinline__ void get(int d, int k, A& hu, A& bu, A& sn, A& st, A& bn,
                                        A& bt) const {
        if constexpr (!ST) {
            if (small) {
                // one 32-bit byte offset for the six arrays: loads from a
                // uniform base with no per-lane 64-bit address arithmetic
                const uint32_t ob = (d ? rowL : rowR) + (uint32_t)k * (uint32_t)sizeof(T);
                hu = (A)buf_ld<T>(rh, ob);
                bu = (A)buf_ld<T>(rb, ob);
                sn = (A)buf_ld<T>(rsN, ob);
                bn = (A)buf_ld<T>(rbN, ob);
                st = (A)buf_ld<T>(rsT, ob);
                bt = (A)buf_ld<T>(rbT, ob);
            } else {
                const size_t o = (d ? baseL : baseR) + k;
                hu = (A)h[o];
                bu = (A)b[o];
                sn = (A)sN[o];
                bn = (A)bN[o];
                st = (A)sT[o];
                bt = (A)bT[o];
            }
        } else {
            const uint32_t kb = (uint32_t)k * (uint32_t)sizeof(T);
            const uint32_t ou = (d ? S.u[1] : S.u[0]) + kb;
            const uint32_t onl = (d ? S.nl[1] : S.nl[0]) + kb;
            const uint32_t onr = (d ? S.nr[1] : S.nr[0]) + kb;
            const uint32_t otl = (d ? S.tl[1] : S.tl[0]) + kb;
            const uint32_t otr = (d ? S.tr[1] : S.tr[0]) + kb;
            const bool inN = d ? S.inN[1] : S.inN[0];
            const bool inT = d ? S.inT[1] : S.inT[0];
            hu = (A)ld_b(h, ou);
            bu = (A)ld_b(b, ou);
            sn = (A)(T)vl_slope<A, true>((A)ld_b(h, onl), hu, (A)ld_b(h, onr), inN, idN, smv);
            bn = (A)(T)vl_slope<A, true>((A)ld_b(b, onl), bu, (A)ld_b(b, onr), inN, idN, smv);
            st = (A)(T)vl_slope<A, true>((A)ld_b(h, otl), hu, (A)ld_b(h, otr), inT, idT, smv);
            bt = (A)(T)vl_slope<A, true>((A)ld_b(b, otl), bu, (A)ld_b(b, otr), inT, idT, smv);
        }
    }
};
template <typename T, bool ST>
__device__ __forceinline__ Upwind<T, ST> make_upwind(const Consts& c, const FaceGeo& fg,
                                                     const T* h, const T* b, const T* shx,
                                                     const T* shy, const T* sbx, const T* sby) {
    using A = typename Arith<T>::A;
    Upwind<T, ST> U;
    U.h = h;
    U.b = b;
    U.sN = fg.dir == 0 ? shx : shy;
    U.sT = fg.dir == 0 ? shy : shx;
    U.bN = fg.dir == 0 ? sbx : sby;
    U.bT = fg.dir == 0 ? sby : sbx;
    U.baseL = (size_t)fg.cL * c.ldv;
    U.baseR = (size_t)fg.cR * c.ldv;
    U.small = (size_t)c.nx * c.ny * c.ldv * sizeof(T) < ((size_t)1 << 32);
    U.rowL = (uint32_t)(U.baseL * sizeof(T));
    U.rowR = (uint32_t)(U.baseR * sizeof(T));
    if constexpr (!ST) {
        if (U.small) {
            const uint32_t bytes = (uint32_t)((size_t)c.nx * c.ny * c.ldv * sizeof(T));
            U.rh = buf_rsrc(U.h, bytes);
            U.rb = buf_rsrc(U.b, bytes);
            U.rsN = buf_rsrc(U.sN, bytes);
            U.rsT = buf_rsrc(U.sT, bytes);
            U.rbN = buf_rsrc(U.bN, bytes);
            U.rbT = buf_rsrc(U.bT, bytes);
        }
    }
    if constexpr (ST) U.S = face_stencil<T>(c, fg);
    const A lenN = (A)(fg.dir == 0 ? c.dx : c.dy), lenT = (A)(fg.dir == 0 ? c.dy : c.dx);
    U.idN = (A)1 / ((A)0.5 * lenN + (A)0.5 * lenN);
    U.idT = (A)1 / ((A)0.5 * lenT + (A)0.5 * lenT);
    U.smv = (A)c.smv;
    return U;
}

// ------------------------------------------------------------------ K1
// FP64 velocity tables of K1: node k's u, v and weight, from 32-byte
// records {u, v, w, 0} on the FP32-storage path (two loads per node
// instead of three: 4.37 -> 4.20 ms at 512^2 cells, GH40) and from three
// arrays on the FP64 path (the records measured 6.5 -> 7.5 ms there).
struct alignas(32) NodeRec {
    double u, v, w, pad;
};
struct NodeTabs {
    const double *u, *v, *w;   // per node
    const double* rec;         // NodeRec per node
    template <typename T>
    __device__ __forceinline__ void get(int k, double& uk, double& vk, double& wk) const {
        if constexpr (sizeof(T) == 4) {
            const NodeRec r = *reinterpret_cast<const NodeRec*>(rec + 4 * (size_t)k);
            uk = r.u;
            vk = r.v;
            wk = r.w;
        } else {
            uk = u[k];
            vk = v[k];
            wk = w[k];
        }
    }
};
template <int DIR, typename T, bool ST>
__device__ __forceinline__ void fsum_face(const Consts& c, double vshift, const Upwind<T, ST>& up,
                                          const NodeTabs& nt, double* __restrict__ out) {
    using A = typename Arith<T>::A;
    const int lane = threadIdx.x & (WAVE - 1);
    const double len = DIR == 0 ? c.dx : c.dy;
    const double hl = 0.5 * len;
    const A vs = (A)vshift;
    double s[FSUM_N];
#pragma unroll
    for (int q = 0; q < FSUM_N; ++q) s[q] = 0.0;
    // nodes per iteration: one for FP32 storage (118 VGPRs, four waves per
    // SIMD: 4.55 -> 4.38 ms at 512^2 cells, GH40) and for the stencil (ten
    // loads per node), two for FP64 (6.52 ms against 6.68 with one)
    constexpr int UNR = (ST || sizeof(T) == 4) ? 1 : 2;
#pragma unroll UNR
    for (int k = lane; k < c.nn; k += WAVE) {
        double u, v, w;
        nt.get<T>(k, u, v, w);
        // upwind choice in the node arithmetic, as the fused kernel makes it
        const int d = upwind_delta((A)(DIR == 0 ? u : v) - vs);
        A hu, bu, snA, stA, bnA, btA;
        up.get(d, k, hu, bu, snA, stA, bnA, btA);
        const double sgn = d ? hl : -hl;
        const double sn = (double)snA, bn = (double)bnA;
        const double hf = (double)hu + sgn * sn;
        const double bf = (double)bu + sgn * bn;
        // tangential slopes: the orientation sign is applied to their sums
        const double st = (double)stA, bt = (double)btA;
        const double vn = DIR == 0 ? u : v, vt = DIR == 0 ? v : -u;
        const double a = w * hf, e = w * sn, f = w * st, g = w * bf, p = w * bn, r = w * bt;
        const double vn2 = vn * vn, vt2 = vt * vt, c2 = vn2 + vt2, vnvt = vn * vt;
        const double vn3 = vn2 * vn, vn2vt = vn2 * vt, vt2vn = vt2 * vn, vnc2 = vn * c2;
        s[0] += a;
        s[1] += vn * a;
        s[2] += vt * a;
        s[4] += g;
        s[5] += f;
        s[6] += vn * f;
        s[7] += vt * f;
        s[8] += c2 * f;
        s[9] += r;
        s[10] += vn2 * e;
        s[11] += vnvt * f;
        s[12] += vn2 * a;
        s[13] += vn3 * e;
        s[14] += vn2vt * f;
        s[15] += vnvt * a;
        s[16] += vn2vt * e;
        s[17] += vt2vn * f;
        s[19] += vn * g;
        s[20] += (vn2 * c2) * e;
        s[21] += vn2 * p;
        s[22] += (vnvt * c2) * f;
        s[23] += vnvt * r;
        s[24] += vn * e;
        s[25] += vnvt * e;
        s[26] += vnc2 * e;
        s[27] += vn * p;
        s[28] += vn2 * f;
        s[29] += vnc2 * f;
        s[30] += vn * r;
        // raw moments of the heat flux expansion
        s[31] += vt2 * a;
        s[32] += vn3 * a;
        s[33] += vn2vt * a;
        s[34] += vt2vn * a;
        s[35] += (vt2 * vt) * a;
        s[36] += vt * g;
    }
    // c^2 = vn^2 + vt^2: sum w c^2 hf and sum w vn c^2 hf from their parts
    s[3] = s[12] + s[31];
    s[18] = s[32] + s[34];
    if (DIR == 1) {
        // horizontal faces take the tangential slopes with the opposite sign
        constexpr int tq[13] = {5, 6, 7, 8, 9, 11, 14, 17, 22, 23, 28, 29, 30};
#pragma unroll
        for (int t = 0; t < 13; ++t) s[tq[t]] = -s[tq[t]];
    }
    int q;
    const double x = wave_reduce_scatter<FSUM_N>(s, q);
    if (q >= 0) out[q] = x;
}

template <typename T, bool ST>
__device__ __forceinline__ void k_fsum_body(
    Consts c, const DevState* __restrict__ st, const T* __restrict__ h, const T* __restrict__ b,
    const T* __restrict__ shx, const T* __restrict__ shy, const T* __restrict__ sbx,
    const T* __restrict__ sby, const double* __restrict__ und, const double* __restrict__ vnd,
    const double* __restrict__ wnd, const double* __restrict__ node4,
    const double* __restrict__ acc_v, const double* __restrict__ acc_h,
    double* __restrict__ fsum, int nvi, int nint, int bid, int nblk) {
    if (st->stopped) return;
    const int g = __builtin_amdgcn_readfirstlane(
        (int)xcd_block(bid, nblk) * SPLIT_FPW + (int)(threadIdx.x / WAVE));
    if (g >= nint) return;
    const FaceGeo fg = face_geo(c, g, nvi);
    const double dt = st->dt;
    const Upwind<T, ST> up = make_upwind<T, ST>(c, fg, h, b, shx, shy, sbx, sby);
    const NodeTabs nt{und, vnd, wnd, node4};
    if (fg.dir == 0) {
        const double vshift = 0.5 * acc_v[2 * fg.face] * dt;
        fsum_face<0, T, ST>(c, vshift, up, nt, fsum + (size_t)g * FSUM_LD);
    } else {
        const double vshift = 0.5 * acc_h[2 * fg.face] * dt;
        fsum_face<1, T, ST>(c, vshift, up, nt, fsum + (size_t)g * FSUM_LD);
    }
}
template <typename T, bool ST>
__global__ __launch_bounds__(WAVE * SPLIT_FPW) void k_fsum(
    Consts c, const DevState* __restrict__ st, const T* __restrict__ h, const T* __restrict__ b,
    const T* __restrict__ shx, const T* __restrict__ shy, const T* __restrict__ sbx,
    const T* __restrict__ sby, const double* __restrict__ und, const double* __restrict__ vnd,
    const double* __restrict__ wnd, const double* __restrict__ node4,
    const double* __restrict__ acc_v, const double* __restrict__ acc_h,
    double* __restrict__ fsum, int nvi, int nint) {
    k_fsum_body<T, ST>(c, st, h, b, shx, shy, sbx, sby, und, vnd, wnd, node4, acc_v, acc_h, fsum,
        nvi, nint, (int)blockIdx.x, (int)gridDim.x);
}

// ------------------------------------------------------------------ K2
// Coefficient slots of one face: FC_* (hugks_kernels.h).

__device__ __forceinline__ void fstate_lane(const Consts& c, double dt,
                                            const double* __restrict__ W,
                                            const double* __restrict__ acc_v,
                                            const double* __restrict__ acc_h,
                                            const double* __restrict__ fsum, double* fcoef,
                                            const FaceGeo& fg, int g) {
    const double* s1 = fsum + (size_t)g * FSUM_LD;
    double* fc = fcoef + (size_t)g * FCOEF_LD;
    const double* acc = fg.dir == 0 ? acc_v : acc_h;
    const double fn = acc[2 * fg.face], ft = acc[2 * fg.face + 1];
    const double len = fg.dir == 0 ? c.dx : c.dy;
    double p[4];
    const double w[4] = {s1[0], s1[1], s1[2], 0.5 * (s1[3] + s1[4])};
    prim_from_w(c, w, p);
    const double tau = tau_of(c, p[0], p[3]);
    double Mt[7];
    {
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
        const double* wl = W + 4 * (size_t)fg.cL;
        const double* wr = W + 4 * (size_t)fg.cR;
        double locL[4], locR[4], sw[4];
        if (fg.dir == 0) {
            for (int q = 0; q < 4; ++q) { locL[q] = wl[q]; locR[q] = wr[q]; }
        } else {
            locL[0] = wl[0]; locL[1] = wl[2]; locL[2] = -wl[1]; locL[3] = wl[3];
            locR[0] = wr[0]; locR[1] = wr[2]; locR[2] = -wr[1]; locR[3] = wr[3];
        }
        for (int q = 0; q < 4; ++q) sw[q] = (w[q] - locL[q]) / (0.5 * len);
        micro_slope(c, p, sw, aL);
        for (int q = 0; q < 4; ++q) sw[q] = (locR[q] - w[q]) / (0.5 * len);
        micro_slope(c, p, sw, aR);
        sw[0] = s1[5]; sw[1] = s1[6]; sw[2] = s1[7]; sw[3] = 0.5 * (s1[8] + s1[9]);
        micro_slope(c, p, sw, ab);
        fl[0] = Mt[4] * s1[1] - Mt[5] * s1[10] - Mt[5] * s1[11];
        fl[1] = Mt[4] * s1[12] - Mt[5] * s1[13] - Mt[5] * s1[14];
        fl[2] = Mt[4] * s1[15] - Mt[5] * s1[16] - Mt[5] * s1[17];
        fl[3] = Mt[4] * 0.5 * (s1[18] + s1[19]) - Mt[5] * 0.5 * (s1[20] + s1[21]) -
                Mt[5] * 0.5 * (s1[22] + s1[23]);
        const double gg[4] = {fn * Mt[6] * s1[24] + ft * Mt[6] * s1[6],
                              fn * Mt[6] * s1[10] + ft * Mt[6] * s1[28],
                              fn * Mt[6] * s1[25] + ft * Mt[6] * s1[11],
                              fn * Mt[6] * 0.5 * (s1[26] + s1[27]) +
                                  ft * Mt[6] * 0.5 * (s1[29] + s1[30])};
        for (int q = 0; q < 4; ++q) fl[q] += gg[q];
    }
    // heat flux (get_heat_flux, 2d/src/gastheory2d.f90:115-128) from the raw
    // moments M_ab = sum w vn^a vt^b h and B_ab of b, expanded about (U, V)
    {
        const double U = p[1], V = p[2];
        const double M00 = s1[0], M10 = s1[1], M01 = s1[2], M20 = s1[12], M11 = s1[15];
        const double M02 = s1[31], M30 = s1[32], M21 = s1[33], M12 = s1[34], M03 = s1[35];
        const double B00 = s1[4], B10 = s1[19], B01 = s1[36];
        const double hx = M30 - 3.0 * U * M20 + 3.0 * U * U * M10 - U * U * U * M00 + M12 -
                          2.0 * V * M11 + V * V * M10 - U * M02 + 2.0 * U * V * M01 -
                          U * V * V * M00;
        const double hy = M03 - 3.0 * V * M02 + 3.0 * V * V * M01 - V * V * V * M00 + M21 -
                          2.0 * U * M11 + U * U * M01 - V * M20 + 2.0 * U * V * M10 -
                          U * U * V * M00;
        fc[FC_Q] = 0.5 * (hx + (B10 - U * B00));
        fc[FC_Q + 1] = 0.5 * (hy + (B01 - V * B00));
    }
    for (int q = 0; q < 4; ++q) {
        fc[FC_P + q] = p[q];
        fc[FC_AL + q] = aL[q];
        fc[FC_AR + q] = aR[q];
        fc[FC_AB + q] = ab[q];
        fc[FC_FL + q] = fl[q];
    }
    fc[FC_TAU] = tau;
    for (int q = 1; q <= 6; ++q) fc[FC_MT + q - 1] = Mt[q];
}

#ifndef HUGKS_BODIES_ONLY
__global__ __launch_bounds__(256) void k_fstate(Consts c, const DevState* __restrict__ st,
                                                const double* __restrict__ W,
                                                const double* __restrict__ acc_v,
                                                const double* __restrict__ acc_h,
                                                const double* __restrict__ fsum,
                                                double* __restrict__ fcoef, int nvi, int nint) {
    if (st->stopped) return;
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= nint) return;
    fstate_lane(c, st->dt, W, acc_v, acc_h, fsum, fcoef, face_geo(c, g, nvi), g);
}
#endif

// ------------------------------------------------------------------ K3
template <int DIR, typename T>
__device__ __forceinline__ void fmaxw_face(const Consts& c, const double* __restrict__ fc,
                                           double fn, double ft,
                                           const double* __restrict__ und,
                                           const double* __restrict__ vnd,
                                           const double* __restrict__ wnd,
                                           double* __restrict__ out) {
    using A = typename Arith<T>::A;
    const int lane = threadIdx.x & (WAVE - 1);
    const double U = fc[FC_P + 1], V = fc[FC_P + 2], lam = fc[FC_P + 3];
    const A lamA = (A)lam;
    const A rho_l_pi = (A)(fc[FC_P] * (lam / c.pi));
    const double kfac = c.K / (2.0 * lam);
    double s[FMAXW_N];
#pragma unroll
    for (int q = 0; q < FMAXW_N; ++q) s[q] = 0.0;
#pragma unroll 2
    for (int k = lane; k < c.nn; k += WAVE) {
        const double u = und[k], v = vnd[k], w = wnd[k];
        const double vn = DIR == 0 ? u : v, vt = DIR == 0 ? v : -u;
        const double cu = vn - U, cv = vt - V;
        const double cp2 = cu * cu + cv * cv;
        const double c2 = vn * vn + vt * vt;
        // the Maxwellian in the node arithmetic (FP32 on the mixed path)
        const double H0 = (double)(rho_l_pi * fexp(-lamA * (A)cp2));
        const double B0 = H0 * kfac;
        const double wg = w * lam * (fn * cu + ft * cv);
        s[0] += vn * wg * H0;
        s[1] += vt * wg * H0;
        s[2] += wg * (c2 * H0 + B0);
        const double lam2 = 2.0 * lam * cp2;
        const double wvn = w * vn;
        const double P5H = (lam2 + c.K - 5.0) * H0 * wvn;
        const double P3B = (lam2 + c.K - 3.0) * B0 * wvn;
        s[3] += cu * P5H;
        s[4] += cv * P5H;
        s[5] += vn * cu * P5H;
        s[6] += vn * cv * P5H;
        s[7] += vt * cu * P5H;
        s[8] += vt * cv * P5H;
        s[9] += c2 * cu * P5H;
        s[10] += c2 * cv * P5H;
        s[11] += cu * P3B;
        s[12] += cv * P3B;
    }
    int q;
    const double x = wave_reduce_scatter<FMAXW_N>(s, q);
    if (q >= 0) out[q] = x;
}

template <typename T>
__global__ __launch_bounds__(WAVE * SPLIT_FPW) void k_fmaxw(
    Consts c, const DevState* __restrict__ st, const double* __restrict__ und,
    const double* __restrict__ vnd, const double* __restrict__ wnd,
    const double* __restrict__ acc_v, const double* __restrict__ acc_h,
    const double* __restrict__ fcoef, double* __restrict__ fmaxw, int nvi, int nint) {
    if (st->stopped) return;
    const int g = __builtin_amdgcn_readfirstlane(blockIdx.x * SPLIT_FPW + (int)(threadIdx.x / WAVE));
    if (g >= nint) return;
    const FaceGeo fg = face_geo(c, g, nvi);
    const double* acc = fg.dir == 0 ? acc_v : acc_h;
    const double fn = acc[2 * fg.face], ft = acc[2 * fg.face + 1];
    const double* fc = fcoef + (size_t)g * FCOEF_LD;
    double* out = fmaxw + (size_t)g * FMAXW_LD;
    if (fg.dir == 0)
        fmaxw_face<0, T>(c, fc, fn, ft, und, vnd, wnd, out);
    else
        fmaxw_face<1, T>(c, fc, fn, ft, und, vnd, wnd, out);
}

// K3 for a tensor-product quadrature, on the face's lane.  With nodes
// (u_i, v_j), weights wx_i wy_j and the Maxwellian
// exp(-lam (cu^2 + cv^2)) = exp(-lam cu^2) exp(-lam cv^2), every sum of
// fmaxw_face is a short sum of products <f>_n <g>_t of 1D moments along
// the face normal (cu = vn - U) and tangent (cv = vt - V):
//   n[]: <1> <cu> <vn> <cu vn> <cu vn^2> <vn^2> <cu vn q5> <vn q5>
//        <cu vn^2 q5> <vn^2 q5> <cu vn^3 q5> <cu vn^3> <vn^3 q5> <vn^3>
//        <cu vn q3> <vn q3>,   q5 = 2 lam cu^2 + K - 5, q3 = 2 lam cu^2 + K - 3
//   t[]: <1> <cv> <vt> <cv vt> <vt^2> <cv vt^2> <r> <cv r> <vt r> <cv vt r>
//        <vt^2 r> <cv vt^2 r>,  r = 2 lam cv^2
// (2 lam cp^2 + K - 5 = q5 + r, c^2 = vn^2 + vt^2).  2 nv exponentials and
// about 30 nv multiply-adds per face instead of nv^2 of each.
__device__ __forceinline__ void fmaxw_sep_lane(const Consts& c, const double* __restrict__ u1,
                                               const double* __restrict__ v1,
                                               const double* __restrict__ wx,
                                               const double* __restrict__ wy,
                                               const double* __restrict__ acc_v,
                                               const double* __restrict__ acc_h,
                                               const double* fcoef, const FaceGeo& fg, int g,
                                               double* out) {
    const double* acc = fg.dir == 0 ? acc_v : acc_h;
    const double fn = acc[2 * fg.face], ft = acc[2 * fg.face + 1];
    const double* fc = fcoef + (size_t)g * FCOEF_LD;
    const double rho = fc[FC_P], U = fc[FC_P + 1], V = fc[FC_P + 2], lam = fc[FC_P + 3];
    // vertical face: normal u, tangent v; horizontal face: normal v, tangent -u
    const double* xn = fg.dir == 0 ? u1 : v1;
    const double* wn = fg.dir == 0 ? wx : wy;
    const double* xt = fg.dir == 0 ? v1 : u1;
    const double* wt = fg.dir == 0 ? wy : wx;
    const double tsgn = fg.dir == 0 ? 1.0 : -1.0;
    const double tl = 2.0 * lam, a5 = c.K - 5.0, a3 = c.K - 3.0;
    double n[16], t[12];
#pragma unroll
    for (int q = 0; q < 16; ++q) n[q] = 0.0;
#pragma unroll
    for (int q = 0; q < 12; ++q) t[q] = 0.0;
    for (int i = 0; i < c.nv; ++i) {
        const double vn = xn[i], cu = vn - U, cu2 = cu * cu;
        const double e = wn[i] * exp(-lam * cu2);
        const double vn2 = vn * vn, vn3 = vn2 * vn;
        const double q5 = tl * cu2 + a5, q3 = tl * cu2 + a3;
        const double ecu = e * cu;
        n[0] += e;
        n[1] += ecu;
        n[2] += e * vn;
        n[3] += ecu * vn;
        n[4] += ecu * vn2;
        n[5] += e * vn2;
        n[6] += ecu * vn * q5;
        n[7] += e * vn * q5;
        n[8] += ecu * vn2 * q5;
        n[9] += e * vn2 * q5;
        n[10] += ecu * vn3 * q5;
        n[11] += ecu * vn3;
        n[12] += e * vn3 * q5;
        n[13] += e * vn3;
        n[14] += ecu * vn * q3;
        n[15] += e * vn * q3;
    }
    for (int j = 0; j < c.nv; ++j) {
        const double vt = tsgn * xt[j], cv = vt - V, cv2 = cv * cv;
        const double e = wt[j] * exp(-lam * cv2);
        const double r = tl * cv2, vt2 = vt * vt, ecv = e * cv;
        t[0] += e;
        t[1] += ecv;
        t[2] += e * vt;
        t[3] += ecv * vt;
        t[4] += e * vt2;
        t[5] += ecv * vt2;
        t[6] += e * r;
        t[7] += ecv * r;
        t[8] += e * vt * r;
        t[9] += ecv * vt * r;
        t[10] += e * vt2 * r;
        t[11] += ecv * vt2 * r;
    }
    const double R = rho * (lam / c.pi), lR = lam * R, kfac = c.K / (2.0 * lam);
    out[0] = lR * (fn * n[3] * t[0] + ft * n[2] * t[1]);
    out[1] = lR * (fn * n[1] * t[2] + ft * n[0] * t[3]);
    out[2] = lR * (fn * (n[4] * t[0] + n[1] * t[4] + kfac * n[1] * t[0]) +
                   ft * (n[5] * t[1] + n[0] * t[5] + kfac * n[0] * t[1]));
    out[3] = R * (n[6] * t[0] + n[3] * t[6]);
    out[4] = R * (n[7] * t[1] + n[2] * t[7]);
    out[5] = R * (n[8] * t[0] + n[4] * t[6]);
    out[6] = R * (n[9] * t[1] + n[5] * t[7]);
    out[7] = R * (n[6] * t[2] + n[3] * t[8]);
    out[8] = R * (n[7] * t[3] + n[2] * t[9]);
    out[9] = R * (n[10] * t[0] + n[11] * t[6] + n[6] * t[4] + n[3] * t[10]);
    out[10] = R * (n[12] * t[1] + n[13] * t[7] + n[7] * t[5] + n[2] * t[11]);
    out[11] = kfac * R * (n[14] * t[0] + n[3] * t[6]);
    out[12] = kfac * R * (n[15] * t[1] + n[2] * t[7]);
}

// ------------------------------------------------------------------ K4
__device__ __forceinline__ void fslope_t_lane(const Consts& c, const double* __restrict__ acc_v,
                                              const double* __restrict__ acc_h,
                                              const double* s2, double* fcoef,
                                              double* __restrict__ vflux,
                                              double* __restrict__ hflux, const FaceGeo& fg,
                                              int g) {
    double* fc = fcoef + (size_t)g * FCOEF_LD;          // s2[q] = K3 sum q
    const double* acc = fg.dir == 0 ? acc_v : acc_h;
    const double fn = acc[2 * fg.face], ft = acc[2 * fg.face + 1];
    double p[4], aL[4], aR[4], ab[4], fl[4], Mt[7];
    for (int q = 0; q < 4; ++q) {
        p[q] = fc[FC_P + q];
        aL[q] = fc[FC_AL + q];
        aR[q] = fc[FC_AR + q];
        ab[q] = fc[FC_AB + q];
        fl[q] = fc[FC_FL + q];
    }
    for (int q = 1; q <= 6; ++q) Mt[q] = fc[FC_MT + q - 1];
    const double tau = fc[FC_TAU], qf0 = fc[FC_Q], qf1 = fc[FC_Q + 1];
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
        sw[1] = sw[1] + 2.0 * s2[0];
        sw[2] = sw[2] + 2.0 * s2[1];
        sw[3] = sw[3] + s2[2];
        micro_slope(c, p, sw, aT);
    }
    for (int q = 0; q < 4; ++q) fc[FC_AT + q] = aT[q];
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
    const double X0 = cs * (qf0 * s2[3] + qf1 * s2[4]);
    const double X1 = cs * (qf0 * s2[5] + qf1 * s2[6]);
    const double X2 = cs * (qf0 * s2[7] + qf1 * s2[8]);
    const double X3 = cs * (qf0 * s2[9] + qf1 * s2[10]);
    const double X4 = cs * (qf0 * s2[11] + qf1 * s2[12]);
    const double out4[4] = {g0[0] + Mt[1] * X0 + fl[0], g0[1] + Mt[1] * X1 + fl[1],
                            g0[2] + Mt[1] * X2 + fl[2], g0[3] + Mt[1] * 0.5 * (X3 + X4) + fl[3]};
    // global frame (flux.f90:242) and face length
    const double flen = fg.dir == 0 ? c.dy : c.dx;
    double g1 = out4[1], g2 = out4[2];
    if (fg.dir == 1) { g1 = -out4[2]; g2 = out4[1]; }
    double* out = (fg.dir == 0 ? vflux : hflux) + 4 * (size_t)fg.face;
    out[0] = flen * out4[0];
    out[1] = flen * g1;
    out[2] = flen * g2;
    out[3] = flen * out4[3];
}

#ifndef HUGKS_BODIES_ONLY
__global__ __launch_bounds__(256) void k_fslope_t(Consts c, const DevState* __restrict__ st,
                                                  const double* __restrict__ acc_v,
                                                  const double* __restrict__ acc_h,
                                                  const double* __restrict__ fmaxw,
                                                  double* __restrict__ fcoef,
                                                  double* __restrict__ vflux,
                                                  double* __restrict__ hflux, int nvi, int nint) {
    if (st->stopped) return;
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= nint) return;
    fslope_t_lane(c, acc_v, acc_h, fmaxw + (size_t)g * FMAXW_LD, fcoef, vflux, hflux,
                  face_geo(c, g, nvi), g);
}
#endif

// K2, K3 (tensor-product quadrature) and K4 of one face on one lane, the
// Maxwellian sums in registers: one launch instead of three.
__device__ __forceinline__ void k_fface_body(Consts c, const DevState* __restrict__ st,
                                               const double* __restrict__ W,
                                               const double* __restrict__ u1,
                                               const double* __restrict__ v1,
                                               const double* __restrict__ wx,
                                               const double* __restrict__ wy,
                                               const double* __restrict__ acc_v,
                                               const double* __restrict__ acc_h,
                                               const double* __restrict__ fsum, double* fcoef,
                                               double* __restrict__ vflux,
                                               double* __restrict__ hflux, int nvi, int nint,
                                               int bid, int nblk) {
    if (st->stopped) return;
    const int g = bid * 256 + threadIdx.x;
    if (g >= nint) return;
    const FaceGeo fg = face_geo(c, g, nvi);
    fstate_lane(c, st->dt, W, acc_v, acc_h, fsum, fcoef, fg, g);
    double s2[FMAXW_N];
    fmaxw_sep_lane(c, u1, v1, wx, wy, acc_v, acc_h, fcoef, fg, g, s2);
    fslope_t_lane(c, acc_v, acc_h, s2, fcoef, vflux, hflux, fg, g);
}
#ifndef HUGKS_BODIES_ONLY
__global__ __launch_bounds__(256) void k_fface(Consts c, const DevState* __restrict__ st,
                                               const double* __restrict__ W,
                                               const double* __restrict__ u1,
                                               const double* __restrict__ v1,
                                               const double* __restrict__ wx,
                                               const double* __restrict__ wy,
                                               const double* __restrict__ acc_v,
                                               const double* __restrict__ acc_h,
                                               const double* __restrict__ fsum, double* fcoef,
                                               double* __restrict__ vflux,
                                               double* __restrict__ hflux, int nvi, int nint) {
    k_fface_body(c, st, W, u1, v1, wx, wy, acc_v, acc_h, fsum, fcoef, vflux, hflux, nvi, nint,
        (int)blockIdx.x, (int)gridDim.x);
}
#endif

// ------------------------------------------------------------------ K5
template <int DIR, typename T, bool ST>
__device__ __forceinline__ void fnode_face(const Consts& c, const double* __restrict__ fc,
                                           double fn, double ft, const Upwind<T, ST>& up,
                                           const typename Arith<T>::A* __restrict__ uvnode,
                                           double dt, T* __restrict__ ofh, T* __restrict__ ofb) {
    using A = typename Arith<T>::A;
    const int lane = threadIdx.x & (WAVE - 1);
    const double len = DIR == 0 ? c.dx : c.dy;
    const double flen = DIR == 0 ? c.dy : c.dx;
    const A tsign = DIR == 0 ? (A)1 : (A)-1;
    const A hlen = (A)(0.5 * len);
    const A vshift = (A)(0.5 * fn * dt);
    const double p0 = fc[FC_P], p1 = fc[FC_P + 1], p2 = fc[FC_P + 2], p3 = fc[FC_P + 3];
    const double tau = fc[FC_TAU];
    const A U = (A)p1, V = (A)p2, lam = (A)p3;
    const A rho_l_pi = (A)(p0 * (p3 / c.pi));
    const A kfac = (A)(c.K / (2.0 * p3));
    const A fnA = (A)fn, ftA = (A)ft;
    const A K = (A)c.K;
    const A mt1 = (A)fc[FC_MT + 0], mt2 = (A)fc[FC_MT + 1], mt3 = (A)fc[FC_MT + 2],
            mt4 = (A)fc[FC_MT + 3], mt5 = (A)fc[FC_MT + 4], mt6 = (A)fc[FC_MT + 5];
    const A tm2fn = (A)(tau * fc[FC_MT + 1] * fn), tm2ft = (A)(tau * fc[FC_MT + 1] * ft);
    const A mxi2 = (A)((c.K * c.K + 2.0 * c.K) / (4.0 * p3 * p3));
    const A csA = (A)(c.shk * (1.0 - c.prandtl) * p3 * p3 / p0);
    const A q0 = (A)fc[FC_Q], q1 = (A)fc[FC_Q + 1];
    // per-face factors: with P(a) = a0 + a1 vn + a2 vt + (a3/2) c^2, the
    // Maxwellian-slope terms are g_h = H P(a) + (a3/2) B and
    // g_b = B P(a) + (a3/2) mxi2 H; the node flux is vn times a sum of
    // products of per-node coefficients and these terms
    const A flenA = (A)flen;
    A sL[5], sR[5], pb[5], pT[5];   // a0, a1, a2, a3/2, a3 mxi2/2
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        sL[q] = (A)fc[FC_AL + q];
        sR[q] = (A)fc[FC_AR + q];
        pb[q] = (A)fc[FC_AB + q];
        pT[q] = (A)fc[FC_AT + q];
    }
    sL[3] = (A)(0.5 * fc[FC_AL + 3]);
    sR[3] = (A)(0.5 * fc[FC_AR + 3]);
    pb[3] = (A)(0.5 * fc[FC_AB + 3]);
    pT[3] = (A)(0.5 * fc[FC_AT + 3]);
    sL[4] = sL[3] * mxi2;
    sR[4] = sR[3] * mxi2;
    pb[4] = pb[3] * mxi2;
    pT[4] = pT[3] * mxi2;
    const A tl = (A)2 * lam, K5 = K - (A)5, K3 = K - (A)3;
    const A m6fn = mt6 * fnA, m6ft = tsign * (mt6 * ftA), m5t = tsign * mt5;
    // one node per iteration (62 VGPRs for FP32 storage: 4.55 -> 4.32 ms at
    // 512^2 cells, GH40; FP64 unchanged)
#pragma unroll 1
    for (int k = lane; k < c.nn; k += WAVE) {
        // node k's (u, v) in the node arithmetic, one 2-element load
        struct alignas(2 * sizeof(A)) UV { A u, v; };
        const UV uv = *reinterpret_cast<const UV*>(uvnode + 2 * k);
        const A u = uv.u, v = uv.v;
        const A vn = DIR == 0 ? u : v, vt = DIR == 0 ? v : -u;
        const int d = upwind_delta(vn - vshift);
        A hu, bu, sn, stR, bn, btR;
        up.get(d, k, hu, bu, sn, stR, bn, btR);
        const A sgn = d ? hlen : -hlen;
        const A hr = hu + sgn * sn;
        const A br = bu + sgn * bn;
        const A cu = vn - U, cv = vt - V;
        const A cp2 = cu * cu + cv * cv;
        const A c2 = vn * vn + vt * vt;
        const A H = rho_l_pi * fexp(-lam * cp2);
        const A Bm = H * kfac;
        const A cq = csA * (cu * q0 + cv * q1);
        const A l2 = tl * cp2;
        const A Hp = cq * (l2 + K5) * H;
        const A Bp = cq * (l2 + K3) * Bm;
        const A s0 = d ? sL[0] : sR[0], s1 = d ? sL[1] : sR[1], s2 = d ? sL[2] : sR[2];
        const A s3 = d ? sL[3] : sR[3], s4 = d ? sL[4] : sR[4];
        const A Pu = s0 + s1 * vn + s2 * vt + s3 * c2;
        const A Pb = pb[0] + pb[1] * vn + pb[2] * vt + pb[3] * c2;
        const A PT = pT[0] + pT[1] * vn + pT[2] * vt + pT[3] * c2;
        const A gh = H * Pu + s3 * Bm, gb = Bm * Pu + s4 * H;
        const A gbh = H * Pb + pb[3] * Bm, gbb = Bm * Pb + pb[4] * H;
        const A gTh = H * PT + pT[3] * Bm, gTb = Bm * PT + pT[4] * H;
        // coefficients of g, g_b, the normal and the tangential slope
        // (flux.f90:207-236 collected by term)
        const A al = mt2 * vn - tm2fn, be = mt2 * vt - tm2ft;
        const A ga = m6fn - mt5 * vn, de = m6ft - m5t * vt;
        const A fv = flenA * vn;
        const A xh = mt1 * (H + Hp) + al * gh + be * gbh + mt3 * gTh + mt4 * hr + ga * sn +
                     de * stR;
        const A xb = mt1 * (Bm + Bp) + al * gb + be * gbb + mt3 * gTb + mt4 * br + ga * bn +
                     de * btR;
        // streamed out: read back only by the cell update
        __builtin_nontemporal_store((T)(fv * xh), ofh + k);
        __builtin_nontemporal_store((T)(fv * xb), ofb + k);
    }
}

template <typename T, bool ST>
__device__ __forceinline__ void k_fnode_body(
    Consts c, const DevState* __restrict__ st, const T* __restrict__ h, const T* __restrict__ b,
    const T* __restrict__ shx, const T* __restrict__ shy, const T* __restrict__ sbx,
    const T* __restrict__ sby, const typename Arith<T>::A* __restrict__ uvnode,
    const double* __restrict__ acc_v,
    const double* __restrict__ acc_h, const double* __restrict__ fcoef, T* __restrict__ vfh,
    T* __restrict__ vfb, T* __restrict__ hfh, T* __restrict__ hfb, int nvi, int nint, int bid,
    int nblk) {
    if (st->stopped) return;
    const int g = __builtin_amdgcn_readfirstlane(
        (int)xcd_block(bid, nblk) * SPLIT_FPW + (int)(threadIdx.x / WAVE));
    if (g >= nint) return;
    const FaceGeo fg = face_geo(c, g, nvi);
    const double dt = st->dt;
    const double* fc = fcoef + (size_t)g * FCOEF_LD;
    const Upwind<T, ST> up = make_upwind<T, ST>(c, fg, h, b, shx, shy, sbx, sby);
    if (fg.dir == 0) {
        const double fn = acc_v[2 * fg.face], ft = acc_v[2 * fg.face + 1];
        fnode_face<0, T, ST>(c, fc, fn, ft, up, uvnode, dt,
                             vfh + (size_t)fg.face * c.ldv, vfb + (size_t)fg.face * c.ldv);
    } else {
        const double fn = acc_h[2 * fg.face], ft = acc_h[2 * fg.face + 1];
        fnode_face<1, T, ST>(c, fc, fn, ft, up, uvnode, dt,
                             hfh + (size_t)fg.face * c.ldv, hfb + (size_t)fg.face * c.ldv);
    }
}
template <typename T, bool ST>
__global__ __launch_bounds__(WAVE * SPLIT_FPW) void k_fnode(
    Consts c, const DevState* __restrict__ st, const T* __restrict__ h, const T* __restrict__ b,
    const T* __restrict__ shx, const T* __restrict__ shy, const T* __restrict__ sbx,
    const T* __restrict__ sby, const typename Arith<T>::A* __restrict__ uvnode,
    const double* __restrict__ acc_v,
    const double* __restrict__ acc_h, const double* __restrict__ fcoef, T* __restrict__ vfh,
    T* __restrict__ vfb, T* __restrict__ hfh, T* __restrict__ hfb, int nvi, int nint) {
    k_fnode_body<T, ST>(c, st, h, b, shx, shy, sbx, sby, uvnode, acc_v, acc_h, fcoef, vfh,
        vfb, hfh, hfb, nvi, nint, (int)blockIdx.x, (int)gridDim.x);
}

// ------------------------------------------------------- K5 as a divergence
// With the node update forming the equilibrium part of the interior faces'
// node fluxes on its matrix tiles (update mode 3), this kernel takes K5's
// place: per cell and node the rest of the face divergence,
//   d = F_L - F_R + F_D - F_U,
// where F is, on an interior face, the free-transport part of the node flux
// (flux.f90:207-236; the terms of the upwind cell's f and slopes):
//   l vn (mt4 (f + s len/2 f_n) + (mt6 fn - mt5 vn) f_n + t (mt6 ft - mt5 vt) f_t)
// (s = +1 when the left/lower cell is upwind, else -1; t the frame sign of the
// face, +1 on x faces, -1 on y faces; vt = v on x faces, -u on y faces), and on
// a wall face the whole node flux of k_flux_wall.  The upwind cell of a node is
// chosen as K5 chooses it (the sign of vn - fn dt/2 in the node arithmetic).
// One wavefront per cell; a lane takes VW consecutive nodes of a row (16-byte
// loads and stores where a row is a multiple of 16 bytes, else VW = 1).  Raw
// buffer loads on whole arrays with 32-bit byte offsets: the host runs this
// mode only for distribution arrays under 4 GiB - 16 bytes.
constexpr int DIV_CPW = 4;   // cells (wavefronts) per workgroup of k_fdiv

template <typename T, int VW>
struct DivVec {
    using V = T __attribute__((ext_vector_type(VW)));
    static __device__ __forceinline__ V ld(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        if constexpr (sizeof(V) == 16)
            return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
        else if constexpr (sizeof(V) == 8)
            return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
        else
            return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    }
};

template <typename T, int VW>
__device__ __forceinline__ void k_fdiv_body(
    Consts c, const DevState* __restrict__ st, const T* __restrict__ h, const T* __restrict__ b,
    const T* __restrict__ shx, const T* __restrict__ shy, const T* __restrict__ sbx,
    const T* __restrict__ sby, const double* __restrict__ u1, const double* __restrict__ v1,
    const double* __restrict__ acc_v, const double* __restrict__ acc_h,
    const double* __restrict__ fcoef, const T* __restrict__ vfh, const T* __restrict__ vfb,
    const T* __restrict__ hfh, const T* __restrict__ hfb, T* __restrict__ dvh,
    T* __restrict__ dvb, int bid, int nblk) {
    using A = typename Arith<T>::A;
    using DV = DivVec<T, VW>;
    using V = typename DV::V;
    if (st->stopped) return;
    const int lane = threadIdx.x & (WAVE - 1);
    const int nx = c.nx, ny = c.ny, nv = c.nv, nn = c.nn;
    const int L = __builtin_amdgcn_readfirstlane(
        (int)xcd_block(bid, nblk) * DIV_CPW + (int)(threadIdx.x / WAVE));
    if (L >= nx * ny) return;
    const int cell = tiled_cell(L, nx, ny);
    const int i = cell % nx, j = cell / nx;
    const double dt = st->dt;
    const bool xper = c.bc[0] == HUGKS_BC_PERIODIC, yper = c.bc[2] == HUGKS_BC_PERIODIC;
    const bool inL = xper || i > 0, inR = xper || i < nx - 1;
    const bool inD = yper || j > 0, inU = yper || j < ny - 1;
    const int nf = xper ? nx + 1 : nx - 1, nr = yper ? ny + 1 : ny - 1;
    const int ox = xper ? 0 : 1, oy = yper ? 0 : 1;
    const int vl = i + (nx + 1) * j, vr = vl + 1, hd = i + nx * j, hu = i + nx * (j + 1);
    // node-independent factors of a face's free-transport flux:
    // vs = fn dt/2, f4 = l mt4, h4 = l mt4 len/2, f6n = l mt6 fn, f5 = l mt5,
    // f6t = t l mt6 ft, f5t = t l mt5 (zero on a wall face)
    struct FK {
        A vs, f4, h4, f6n, f5, f6t, f5t;
    };
    auto face_k = [&](bool in, bool isx, int g, int slot) -> FK {
        FK k{0, 0, 0, 0, 0, 0, 0};
        if (!in) return k;
        const double* fc = fcoef + (size_t)g * FACE_COEF_LD;
        const double* ac = isx ? acc_v : acc_h;
        const double fn = ac[2 * slot], ft = ac[2 * slot + 1];
        const double fl = isx ? c.dy : c.dx, len = isx ? c.dx : c.dy, ts = isx ? 1.0 : -1.0;
        const double mt4 = fc[FC_MT + 3], mt5 = fc[FC_MT + 4], mt6 = fc[FC_MT + 5];
        k.vs = (A)(0.5 * fn * dt);
        k.f4 = (A)(fl * mt4);
        k.h4 = (A)(fl * mt4 * (0.5 * len));
        k.f6n = (A)(fl * mt6 * fn);
        k.f5 = (A)(fl * mt5);
        k.f6t = (A)(ts * fl * mt6 * ft);
        k.f5t = (A)(ts * fl * mt5);
        return k;
    };
    const FK kL = face_k(inL, true, tiled_index(i - ox, j, nf, ny), vl);
    const FK kR = face_k(inR, true, tiled_index(i + 1 - ox, j, nf, ny), vr);
    const FK kD = face_k(inD, false, nf * ny + tiled_index(i, j - oy, nx, nr), hd);
    const FK kU = face_k(inU, false, nf * ny + tiled_index(i, j + 1 - oy, nx, nr), hu);
    // one face's flux of h and b at a node: normal and tangential velocity,
    // upwind side, the upwind f and slopes (normal first).  Written as
    // explicit fused multiply-adds, so every instance (one or VW nodes per
    // lane, alone or in a grouped launch) rounds the same way.
    auto flux = [](const FK& k, A vn, A vt, int d, A hu_, A sn, A st_, A bu, A bn, A bt, A& fh,
                   A& fb) {
        const A ch = vn * k.f4;
        const A cn = vn * fma(-k.f5, vn, (d ? k.h4 : -k.h4) + k.f6n);
        const A ct = vn * fma(-k.f5t, vt, k.f6t);
        fh = fma(ch, hu_, fma(cn, sn, ct * st_));
        fb = fma(ch, bu, fma(cn, bn, ct * bt));
    };
    const size_t ldv = (size_t)c.ldv;
    const uint32_t rowB = (uint32_t)(ldv * sizeof(T));
    const uint32_t nbytes = (uint32_t)((size_t)nx * ny * ldv * sizeof(T));
    constexpr uint32_t OUTF = 0xFFFFFFF0u;   // past every array (host check)
    const __amdgpu_buffer_rsrc_t rH = buf_rsrc(h, nbytes), rB = buf_rsrc(b, nbytes);
    const __amdgpu_buffer_rsrc_t rSX = buf_rsrc(shx, nbytes), rSY = buf_rsrc(shy, nbytes);
    const __amdgpu_buffer_rsrc_t rBX = buf_rsrc(sbx, nbytes), rBY = buf_rsrc(sby, nbytes);
    const uint32_t bO = (uint32_t)cell * rowB;
    const uint32_t bL = (uint32_t)((i - 1 + nx) % nx + nx * j) * rowB;
    const uint32_t bR = (uint32_t)((i + 1) % nx + nx * j) * rowB;
    const uint32_t bD = (uint32_t)(i + nx * ((j - 1 + ny) % ny)) * rowB;
    const uint32_t bU = (uint32_t)(i + nx * ((j + 1) % ny)) * rowB;
    const uint32_t rowb = (uint32_t)(nn * (int)sizeof(T));
    // the wall faces' node-flux rows (h, b) of the cell's four faces
    const __amdgpu_buffer_rsrc_t rW[8] = {
        buf_rsrc(vfh + (size_t)vl * ldv, rowb), buf_rsrc(vfh + (size_t)vr * ldv, rowb),
        buf_rsrc(hfh + (size_t)hd * ldv, rowb), buf_rsrc(hfh + (size_t)hu * ldv, rowb),
        buf_rsrc(vfb + (size_t)vl * ldv, rowb), buf_rsrc(vfb + (size_t)vr * ldv, rowb),
        buf_rsrc(hfb + (size_t)hd * ldv, rowb), buf_rsrc(hfb + (size_t)hu * ldv, rowb)};
    // f and the x and y slopes of h, then of b, at VW nodes
    struct Up {
        V h, sx, sy, b, bx, by;
    };
    auto ld = [&](uint32_t off) -> Up {
        return Up{DV::ld(rH, off), DV::ld(rSX, off), DV::ld(rSY, off),
                  DV::ld(rB, off), DV::ld(rBX, off), DV::ld(rBY, off)};
    };
    T* const oh = dvh + (size_t)cell * ldv;
    T* const ob = dvb + (size_t)cell * ldv;
#pragma unroll 1
    for (int e = lane * VW; e < nn; e += WAVE * VW) {
        const int jv = e / nv, iu = e - jv * nv;   // VW divides nv: one row
        const uint32_t oe = (uint32_t)e * (uint32_t)sizeof(T);
        const A vv = (A)v1[jv];
        A uu[VW];
#pragma unroll
        for (int q = 0; q < VW; ++q) uu[q] = (A)u1[iu + q];
        const Up me = ld(bO + oe);
        A th[VW], tb[VW];
#pragma unroll
        for (int q = 0; q < VW; ++q) th[q] = tb[q] = (A)0;
        // one face at a time: the upwind side of each node, the neighbour's
        // values where the neighbour is upwind (not loaded where no node of
        // the wavefront needs them), the flux added with the divergence
        // sign; a wall face's node fluxes as the wall kernel stored them
#pragma unroll
        for (int fa = 0; fa < 4; ++fa) {
            const bool isx = fa < 2, lo = (fa & 1) == 0;   // left/lower face of the cell
            const bool in = fa == 0 ? inL : fa == 1 ? inR : fa == 2 ? inD : inU;
            const FK& k = fa == 0 ? kL : fa == 1 ? kR : fa == 2 ? kD : kU;
            const uint32_t bn = fa == 0 ? bL : fa == 1 ? bR : fa == 2 ? bD : bU;
            if (in) {
                int d[VW];
                bool nb[VW];
                bool any = false;
#pragma unroll
                for (int q = 0; q < VW; ++q) {
                    d[q] = upwind_delta((isx ? uu[q] : vv) - k.vs);
                    // the neighbour is upwind of the left/lower face when the
                    // left/lower cell is (d = 1), of the right/upper one when
                    // the right/upper cell is (d = 0)
                    nb[q] = lo ? d[q] != 0 : d[q] == 0;
                    any = any || nb[q];
                }
                Up w = me;
                if (__any(any)) w = ld(any ? bn + oe : OUTF);
#pragma unroll
                for (int q = 0; q < VW; ++q) {
                    const Up& x = nb[q] ? w : me;
                    A fh, fb;
                    if (isx)
                        flux(k, uu[q], vv, d[q], (A)x.h[q], (A)x.sx[q], (A)x.sy[q], (A)x.b[q],
                             (A)x.bx[q], (A)x.by[q], fh, fb);
                    else
                        flux(k, vv, -uu[q], d[q], (A)x.h[q], (A)x.sy[q], (A)x.sx[q], (A)x.b[q],
                             (A)x.by[q], (A)x.bx[q], fh, fb);
                    th[q] = lo ? th[q] + fh : th[q] - fh;
                    tb[q] = lo ? tb[q] + fb : tb[q] - fb;
                }
            } else {
                const V wh = DV::ld(rW[fa], oe), wb = DV::ld(rW[4 + fa], oe);
#pragma unroll
                for (int q = 0; q < VW; ++q) {
                    th[q] = lo ? th[q] + (A)wh[q] : th[q] - (A)wh[q];
                    tb[q] = lo ? tb[q] + (A)wb[q] : tb[q] - (A)wb[q];
                }
            }
        }
        V sh, sb;
#pragma unroll
        for (int q = 0; q < VW; ++q) {
            sh[q] = (T)th[q];
            sb[q] = (T)tb[q];
        }
        // streamed out: read back only by the cell update
        __builtin_nontemporal_store(sh, reinterpret_cast<V*>(oh + e));
        __builtin_nontemporal_store(sb, reinterpret_cast<V*>(ob + e));
    }
}
template <typename T, int VW>
__global__ __launch_bounds__(WAVE * DIV_CPW) HUGKS_WAVES_PER_EU(sizeof(T) == 4 ? 4 : 3) void k_fdiv(
    Consts c, const DevState* __restrict__ st, const T* __restrict__ h, const T* __restrict__ b,
    const T* __restrict__ shx, const T* __restrict__ shy, const T* __restrict__ sbx,
    const T* __restrict__ sby, const double* __restrict__ u1, const double* __restrict__ v1,
    const double* __restrict__ acc_v, const double* __restrict__ acc_h,
    const double* __restrict__ fcoef, const T* __restrict__ vfh, const T* __restrict__ vfb,
    const T* __restrict__ hfh, const T* __restrict__ hfb, T* __restrict__ dvh,
    T* __restrict__ dvb) {
    k_fdiv_body<T, VW>(c, st, h, b, shx, shy, sbx, sby, u1, v1, acc_v, acc_h, fcoef, vfh, vfb,
                       hfh, hfb, dvh, dvb, (int)blockIdx.x, (int)gridDim.x);
}

}  // namespace hugks
