// 2D slope and wall-face flux kernels (gfx950); the interior-face flux
// kernel is in k_flux_face.h.
#pragma once
#include "hugks_device.h"
#include "hugks_kernels.h"

namespace hugks {

// Cells next to a wall, each once: the x-wall columns (i = 0, nx-1) when x
// is not periodic, then the y-wall rows (j = 0, ny-1) without the cells
// already counted.
__host__ __device__ inline int wall_cell_count(const Consts& c) {
    const bool xw = c.bc[0] != HUGKS_BC_PERIODIC, yw = c.bc[2] != HUGKS_BC_PERIODIC;
    const int nxc = xw ? (c.nx > 1 ? 2 : 1) : 0;
    const int nyc = yw ? (c.ny > 1 ? 2 : 1) : 0;
    return nxc * c.ny + nyc * (c.nx - nxc);
}
__device__ __forceinline__ int wall_cell(const Consts& c, int b) {
    const bool xw = c.bc[0] != HUGKS_BC_PERIODIC;
    const int nxc = xw ? (c.nx > 1 ? 2 : 1) : 0;
    if (b < nxc * c.ny) {
        const int j = b / nxc;
        return (b % nxc ? c.nx - 1 : 0) + c.nx * j;
    }
    b -= nxc * c.ny;
    const int w = c.nx - nxc;            // cells per wall row
    const int row = b / w;
    const int i = (nxc ? 1 : 0) + b % w;
    return i + c.nx * (row ? c.ny - 1 : 0);
}

// ------------------------------------------------------------------ slopes
constexpr int SLOPE_CPW = 4;   // cells (wavefronts) per workgroup of k_slope2d
// interp_inner (van Leer) / interp_boundary (one-sided), 2d/src/solver.f90:
// 64-159.  Periodic directions use wrapped neighbours everywhere (the ghost
// copy of 2d/src/boundary.f90:16-53); wall directions use one-sided slopes
// in the two boundary cells.
template <typename T>
__device__ __forceinline__ void k_slope2d_body(Consts c, const DevState* __restrict__ st,
                                                 const T* __restrict__ h, const T* __restrict__ b,
                                                 T* __restrict__ shx, T* __restrict__ shy,
                                                 T* __restrict__ sbx, T* __restrict__ sby,
                                                 int wall_only, int bid, int nblk) {
    using A = typename Arith<T>::A;
    if (st->stopped) return;
    // one wavefront per cell, four cells per workgroup: every cell in
    // XCD-aware strip order, or only the cells next to a wall (their normal
    // slopes feed the wall fluxes)
    const int lane = threadIdx.x & (WAVE - 1);
    const int L = __builtin_amdgcn_readfirstlane(
        (int)(wall_only ? bid : xcd_block(bid, nblk)) * SLOPE_CPW +
        (int)(threadIdx.x / WAVE));
    if (L >= (wall_only ? wall_cell_count(c) : c.nx * c.ny)) return;
    const int cell = wall_only ? wall_cell(c, L) : tiled_cell(L, c.nx, c.ny);
    const A smv = (A)c.smv;
    int cl[2], cr[2];
    bool inner[2];
    A inv_den[2];
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {
        const SlopeNb nb = slope_nb(c, cell, dir);
        const A len = (A)(dir == 0 ? c.dx : c.dy);
        cl[dir] = nb.cl;
        cr[dir] = nb.cr;
        inner[dir] = nb.inner;
        inv_den[dir] = (A)1 / ((A)0.5 * len + (A)0.5 * len);
    }
    // 16-byte vectors of VW nodes: the padded row length (a multiple of 64)
    // keeps every vector aligned and inside the row; the pad nodes are
    // computed too and never read
    constexpr int VW = 16 / (int)sizeof(T);
    struct alignas(16) Vec { T v[VW]; };
    const int ng = (c.nn + VW - 1) / VW;
    const size_t o = (size_t)cell * c.ldv;
    const size_t oL[2] = {(size_t)cl[0] * c.ldv, (size_t)cl[1] * c.ldv};
    const size_t oR[2] = {(size_t)cr[0] * c.ldv, (size_t)cr[1] * c.ldv};
    for (int g = lane; g < ng; g += WAVE) {
        const size_t k = (size_t)g * VW;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const T* f = q == 0 ? h : b;
            const Vec fc = *reinterpret_cast<const Vec*>(f + o + k);
            Vec vl[2], vr[2];
#pragma unroll
            for (int dir = 0; dir < 2; ++dir) {
                vl[dir] = *reinterpret_cast<const Vec*>(f + oL[dir] + k);
                vr[dir] = *reinterpret_cast<const Vec*>(f + oR[dir] + k);
            }
#pragma unroll
            for (int dir = 0; dir < 2; ++dir) {
                Vec so;
#pragma unroll
                for (int e = 0; e < VW; ++e) {
                    so.v[e] = (T)vl_slope((A)vl[dir].v[e], (A)fc.v[e], (A)vr[dir].v[e], inner[dir],
                                          inv_den[dir], smv);
                }
                T* out = dir == 0 ? (q == 0 ? shx : sbx) : (q == 0 ? shy : sby);
                // streamed out: read back only by the face kernels
#pragma unroll
                for (int e = 0; e < VW; ++e) __builtin_nontemporal_store(so.v[e], out + o + k + e);
            }
        }
    }
}
template <typename T>
__global__ __launch_bounds__(256) void k_slope2d(Consts c, const DevState* __restrict__ st,
                                                 const T* __restrict__ h, const T* __restrict__ b,
                                                 T* __restrict__ shx, T* __restrict__ shy,
                                                 T* __restrict__ sbx, T* __restrict__ sby,
                                                 int wall_only) {
    k_slope2d_body<T>(c, st, h, b, shx, shy, sbx, sby, wall_only, (int)blockIdx.x, (int)gridDim.x);
}

// -------------------------------------------------------------- wall faces
// One wavefront per wall face (four per workgroup).  Faces are numbered over
// the non-periodic sides in the order x-lo (ny), x-hi (ny), y-lo (nx),
// y-hi (nx).  Diffuse isothermal wall: calc_flux_boundary,
// 2d/src/flux.f90:474-549; specular wall: calc_flux_specular_vertical /
// _horizon, 2d/src/flux.f90:551-689.  The incoming half-range values are
// re-read in each pass rather than held across the reductions.
template <typename T>
__device__ __forceinline__ void k_flux_wall_body(
    Consts c, DevState* __restrict__ st, const T* __restrict__ h, const T* __restrict__ b,
    const T* __restrict__ shx, const T* __restrict__ shy, const T* __restrict__ sbx,
    const T* __restrict__ sby, const typename Arith<T>::A* __restrict__ unode,
    const typename Arith<T>::A* __restrict__ vnode, const typename Arith<T>::A* __restrict__ wnode,
    T* __restrict__ vfh, T* __restrict__ vfb, double* __restrict__ vflux, T* __restrict__ hfh,
    T* __restrict__ hfb, double* __restrict__ hflux, int nwall, int bid, int nblk) {
    using A = typename Arith<T>::A;
    if (st->stopped) return;
    const int lane = threadIdx.x & (WAVE - 1);
    int f = bid * 4 + (int)(threadIdx.x / WAVE);
    if (f >= nwall) return;
    int side = -1, idx = 0;
    for (int s = 0; s < 4; ++s) {
        if (bc_of(c, s) == HUGKS_BC_PERIODIC) continue;
        const int n = s < 2 ? c.ny : c.nx;
        if (f < n) { side = s; idx = f; break; }
        f -= n;
    }
    if (side < 0) return;
    const int dir = side / 2;           // 0: vertical face, 1: horizontal face
    const int rot = (side & 1) ? -1 : 1;
    int cell, face;
    if (dir == 0) {
        cell = ((side & 1) ? c.nx - 1 : 0) + c.nx * idx;
        face = ((side & 1) ? c.nx : 0) + (c.nx + 1) * idx;
    } else {
        cell = idx + c.nx * ((side & 1) ? c.ny - 1 : 0);
        face = idx + c.nx * ((side & 1) ? c.ny : 0);
    }
    const T* sN = dir == 0 ? shx : shy;
    const T* bN = dir == 0 ? sbx : sby;
    const double dt = st->dt;
    const double len = dir == 0 ? c.dx : c.dy;
    const double flen = dir == 0 ? c.dy : c.dx;
    const A hrl = (A)rot * (A)0.5 * (A)len;
    const bool diffuse = bc_of(c, side) == HUGKS_BC_DIFFUSE;
    const int nv = c.nv, nn = c.nn;
    const size_t base = (size_t)cell * c.ldv;
    const double lamw = 1.0 / tw_of(c, side);

    // incoming value at node k (cell value extrapolated to the face,
    // flux.f90:509-510), and for a specular wall its mirror node
    // (flux.f90:589-596 / 659-666)
    auto node = [&](int k, A& vn, A& vt, A& wt, int& d, A& hf, A& bf) {
        const A u = unode[k], v = vnode[k];
        wt = wnode[k];
        vn = dir == 0 ? u : v;
        vt = dir == 0 ? v : -u;
        d = (copysign((A)1, vn) * (A)rot) > (A)0 ? 1 : 0;
        int ks = k;
        if (!diffuse && d) {
            const int iu = k % nv, jv = k / nv;
            ks = dir == 0 ? (nv - 1 - iu) + nv * jv : iu + nv * (nv - 1 - jv);
        }
        hf = (A)h[base + ks] - hrl * (A)sN[base + ks];
        bf = (A)b[base + ks] - hrl * (A)bN[base + ks];
    };

    double rhow = 0.0;
    if (diffuse) {
        double r[2] = {0.0, 0.0};
        for (int k = lane; k < nn; k += WAVE) {
            A vn, vt, wt, hf, bf;
            int d;
            node(k, vn, vt, wt, d, hf, bf);
            if (!d) r[0] += (double)wt * (double)vn * (double)hf;
            else r[1] += (double)wt * (double)vn * exp(-lamw * ((double)vn * vn + (double)vt * vt));
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            double x = r[i];
#pragma unroll
            for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
            r[i] = __shfl(x, 0, WAVE);
        }
        const double R2 = (lamw / c.pi) * r[1];
        if (R2 == 0.0 && lane == 0) atomicOr(&st->flags, 2);
        rhow = -r[0] / R2;
    }
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    T* oh = (dir == 0 ? vfh : hfh) + (size_t)face * c.ldv;
    T* ob = (dir == 0 ? vfb : hfb) + (size_t)face * c.ldv;
    const A dtl = (A)(dt * flen);
    const double Hcd = rhow * (lamw / c.pi), kfd = c.K / (2.0 * lamw);
    for (int k = lane; k < nn; k += WAVE) {
        A vn, vt, wt, hf, bf;
        int d;
        node(k, vn, vt, wt, d, hf, bf);
        // the conserved-flux sums in FP64 from FP64 per-node values, so that
        // the re-emitted mass cancels the incoming mass to FP64 round-off
        // (R1 + rho_w R2 = 0) on the FP32-storage path too; the stored
        // per-node fluxes are rounded to the storage type
        const double vnd = (double)vn, vtd = (double)vt, wtd = (double)wt;
        double hfd = (double)hf, bfd = (double)bf;
        if (diffuse && d) {
            hfd = Hcd * exp(-lamw * (vnd * vnd + vtd * vtd));
            bfd = hfd * kfd;
            hf = (A)hfd;
            bf = (A)bfd;
        }
        const double wvh = wtd * vnd * hfd;
        s[0] += wvh;
        s[1] += vnd * wvh;
        s[2] += vtd * wvh;
        s[3] += wtd * vnd * ((vnd * vnd + vtd * vtd) * hfd + bfd);
        oh[k] = (T)(dtl * (vn * hf));
        ob[k] = (T)(dtl * (vn * bf));
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double x = s[i];
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
        s[i] = __shfl(x, 0, WAVE);
    }
    if (lane == 0) {
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
template <typename T>
__global__ __launch_bounds__(256) void k_flux_wall(
    Consts c, DevState* __restrict__ st, const T* __restrict__ h, const T* __restrict__ b,
    const T* __restrict__ shx, const T* __restrict__ shy, const T* __restrict__ sbx,
    const T* __restrict__ sby, const typename Arith<T>::A* __restrict__ unode,
    const typename Arith<T>::A* __restrict__ vnode, const typename Arith<T>::A* __restrict__ wnode,
    T* __restrict__ vfh, T* __restrict__ vfb, double* __restrict__ vflux, T* __restrict__ hfh,
    T* __restrict__ hfb, double* __restrict__ hflux, int nwall) {
    k_flux_wall_body<T>(c, st, h, b, shx, shy, sbx, sby, unode, vnode, wnode, vfh, vfb, vflux, hfh,
        hfb, hflux, nwall, (int)blockIdx.x, (int)gridDim.x);
}

}  // namespace hugks
