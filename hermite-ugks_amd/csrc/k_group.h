// Grouped launches for batches of small cases (gfx950): one launch runs one
// kernel type for every case of a batch, so a batch step is about a dozen
// launches instead of a dozen per case.  Workgroup b of a grouped launch
// belongs to the case e whose range [start[e], start[e+1]) holds it and runs
// that case's kernel body as workgroup b - start[e] of start[e+1] - start[e];
// each case's arguments are its Launch2D, copied to device memory.  The
// bodies are the ones the single-case kernels run, so a case computes the
// same bits in a batch as alone.
#pragma once
#include "k_flux2d.h"
#include "k_flux_split.h"
#include "k_flux_cellsum.h"
#include "k_misc.h"
#include "k_update2d_mx.h"

namespace hugks {

// Table of one grouped launch (ints): [n, start[0..n], case[0..n-1]].
struct GroupIdx {
    int cs, bid, nblk;
};
__device__ __forceinline__ GroupIdx group_index(const int* __restrict__ tab) {
    const int n = tab[0];
    const int* start = tab + 1;
    const int b = (int)blockIdx.x;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    GroupIdx g;
    g.cs = tab[2 + n + lo];
    g.bid = b - start[lo];
    g.nblk = start[lo + 1] - start[lo];
    return g;
}

// interior faces of a case: vertical nvf, all nint
__host__ __device__ inline void face_counts(const Consts& c, int& nvf, int& nint) {
    const bool xper = c.bc[0] == HUGKS_BC_PERIODIC, yper = c.bc[2] == HUGKS_BC_PERIODIC;
    nvf = (xper ? c.nx + 1 : c.nx - 1) * c.ny;
    nint = nvf + c.nx * (yper ? c.ny + 1 : c.ny - 1);
}

__global__ __launch_bounds__(RED_NT) void kg_timestep(const int* __restrict__ tab,
                                                      const Launch2D* __restrict__ Ls) {
    const GroupIdx g = group_index(tab);
    const Launch2D& L = Ls[g.cs];
    k_timestep_body(L.c, 2, L.st, L.W, nullptr, 0, g.bid, g.nblk);
}

template <typename T>
__global__ __launch_bounds__(WAVE * SLOPE_CPW) void kg_slope(const int* __restrict__ tab,
                                                             const Launch2D* __restrict__ Ls) {
    const GroupIdx g = group_index(tab);
    const Launch2D& L = Ls[g.cs];
    k_slope2d_body<T>(L.c, L.st, (const T*)L.h, (const T*)L.b, (T*)L.shx, (T*)L.shy, (T*)L.sbx,
                      (T*)L.sby, 0, g.bid, g.nblk);
}

template <typename T>
__global__ __launch_bounds__(WAVE * SPLIT_FPW) void kg_fsum(const int* __restrict__ tab,
                                                            const Launch2D* __restrict__ Ls) {
    using A = typename Arith<T>::A;
    const GroupIdx g = group_index(tab);
    const Launch2D& L = Ls[g.cs];
    int nvf, nint;
    face_counts(L.c, nvf, nint);
    k_fsum_body<T, false>(L.c, L.st, (const T*)L.h, (const T*)L.b, (const T*)L.shx,
                          (const T*)L.shy, (const T*)L.sbx, (const T*)L.sby, L.unode_d,
                          L.vnode_d, L.wnode_d, L.node4, L.acc_v, L.acc_h, L.fsum, nvf, nint,
                          g.bid, g.nblk);
}

// the cell-sum pipeline (every case of the plan supports it): slopes and
// per-cell sums, then each face's sums from two bins
template <typename T, int VW>
__global__ __launch_bounds__(WAVE * SLOPE_CPW) HUGKS_WAVES_PER_EU(3) void kg_cell_sum(const int* __restrict__ tab,
                                                                const Launch2D* __restrict__ Ls) {
    __shared__ __align__(16) unsigned char lds[SLOPE_CPW * cs_lds_per_wave<T>()];
    const GroupIdx g = group_index(tab);
    const Launch2D& L = Ls[g.cs];
    k_cell_sum_body<T, VW>(L.c, L.st, (const T*)L.h, (const T*)L.b, (T*)L.shx, (T*)L.shy,
                           (T*)L.sbx, (T*)L.sby, L.u1, L.v1, L.wx, L.wy, L.csum, L.cs_js, lds,
                           g.bid, g.nblk);
}

template <typename T>
__global__ __launch_bounds__(256) void kg_fface_cs(const int* __restrict__ tab,
                                                   const Launch2D* __restrict__ Ls) {
    const GroupIdx g = group_index(tab);
    const Launch2D& L = Ls[g.cs];
    int nvf, nint;
    face_counts(L.c, nvf, nint);
    k_fface_cs_body<T>(L.c, L.st, L.W, (const T*)L.h, (const T*)L.b, (const T*)L.shx,
                       (const T*)L.shy, (const T*)L.sbx, (const T*)L.sby, L.u1, L.v1, L.wx, L.wy,
                       L.unode_d, L.vnode_d, L.wnode_d, L.node4, L.acc_v, L.acc_h, L.csum,
                       L.fsum, L.fcoef, L.vflux, L.hflux, nvf, nint, g.bid);
}

__global__ __launch_bounds__(256) void kg_fface(const int* __restrict__ tab,
                                                const Launch2D* __restrict__ Ls) {
    const GroupIdx g = group_index(tab);
    const Launch2D& L = Ls[g.cs];
    int nvf, nint;
    face_counts(L.c, nvf, nint);
    k_fface_body(L.c, L.st, L.W, L.u1, L.v1, L.wx, L.wy, L.acc_v, L.acc_h, L.fsum, L.fcoef,
                 L.vflux, L.hflux, nvf, nint, g.bid, g.nblk);
}

template <typename T>
__global__ __launch_bounds__(WAVE * SPLIT_FPW) void kg_fnode(const int* __restrict__ tab,
                                                             const Launch2D* __restrict__ Ls) {
    using A = typename Arith<T>::A;
    const GroupIdx g = group_index(tab);
    const Launch2D& L = Ls[g.cs];
    int nvf, nint;
    face_counts(L.c, nvf, nint);
    k_fnode_body<T, false>(L.c, L.st, (const T*)L.h, (const T*)L.b, (const T*)L.shx,
                           (const T*)L.shy, (const T*)L.sbx, (const T*)L.sby,
                           (const A*)L.uvnode, L.acc_v, L.acc_h, L.fcoef,
                           (T*)L.vfh, (T*)L.vfb, (T*)L.hfh, (T*)L.hfb, nvf, nint, g.bid, g.nblk);
}

template <typename T>
__global__ __launch_bounds__(256) void kg_wall(const int* __restrict__ tab,
                                               const Launch2D* __restrict__ Ls) {
    using A = typename Arith<T>::A;
    const GroupIdx g = group_index(tab);
    const Launch2D& L = Ls[g.cs];
    k_flux_wall_body<T>(L.c, L.st, (const T*)L.h, (const T*)L.b, (const T*)L.shx,
                        (const T*)L.shy, (const T*)L.sbx, (const T*)L.sby, (const A*)L.unode,
                        (const A*)L.vnode, (const A*)L.wnode, (T*)L.vfh, (T*)L.vfb, L.vflux,
                        (T*)L.hfh, (T*)L.hfb, L.hflux, (int)L.nwall, g.bid, g.nblk);
}

__global__ __launch_bounds__(256) void kg_cell_pre(const int* __restrict__ tab,
                                                   const Launch2D* __restrict__ Ls) {
    const GroupIdx g = group_index(tab);
    const Launch2D& L = Ls[g.cs];
    k_cell_pre_body(L.c, L.st, L.W, L.part, L.vflux, L.hflux, L.acc_cell, L.cellpre, g.bid,
                    g.nblk);
}

// FF: the node update forming the interior faces' equilibrium flux (update
// mode 3), with the waves per SIMD of its single-case kernel
template <typename T, int NW, int MAXP, int NVC, bool FF>
__global__ __launch_bounds__(NW * WAVE)
HUGKS_WAVES_PER_EU(sizeof(T) == 4 ? (FF && NVC ? 4 : 3) : 1) void kg_update(
    const int* __restrict__ tab, const Launch2D* __restrict__ Ls) {
    const GroupIdx g = group_index(tab);
    const Launch2D& L = Ls[g.cs];
    const MxFace fx{L.dvh, L.dvb, L.fcoef, L.acc_v, L.acc_h};
    k_update2d_mx_body<T, NW, MAXP, NVC, false, FF>(
        L.c, L.st, L.cellpre, (T*)L.h, (T*)L.b, (const T*)L.vfh, (const T*)L.vfb,
        (const T*)L.hfh, (const T*)L.hfb, (const typename Arith<T>::A*)L.wnode, L.u1, L.v1,
        g.bid, g.nblk, fx);
}

// the per-cell face divergence of update mode 3 (k_fdiv)
template <typename T, int VW>
__global__ __launch_bounds__(WAVE * DIV_CPW)
HUGKS_WAVES_PER_EU(sizeof(T) == 4 ? 4 : 3) void kg_fdiv(const int* __restrict__ tab,
                                                       const Launch2D* __restrict__ Ls) {
    const GroupIdx g = group_index(tab);
    const Launch2D& L = Ls[g.cs];
    k_fdiv_body<T, VW>(L.c, L.st, (const T*)L.h, (const T*)L.b, (const T*)L.shx,
                       (const T*)L.shy, (const T*)L.sbx, (const T*)L.sby, L.u1, L.v1, L.acc_v,
                       L.acc_h, L.fcoef, (const T*)L.vfh, (const T*)L.vfb, (const T*)L.hfh,
                       (const T*)L.hfb, (T*)L.dvh, (T*)L.dvb, g.bid, g.nblk);
}

__global__ __launch_bounds__(RED_NT) void kg_finalize(const int* __restrict__ tab,
                                                      const Launch2D* __restrict__ Ls) {
    const GroupIdx g = group_index(tab);
    const Launch2D& L = Ls[g.cs];
    k_finalize_body(L.c, 2, L.st, L.part, nullptr, 0, g.bid, g.nblk);
}

}  // namespace hugks
