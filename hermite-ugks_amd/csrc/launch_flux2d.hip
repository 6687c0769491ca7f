// Host launchers of the 2D slope, interior-face and wall-face kernels.
#include "k_flux2d.h"
#include "k_flux_face.h"
#include "k_flux_split.h"
#include "k_flux_cellsum.h"

namespace hugks {

template <typename T>
static hipError_t launch_flux2d_t(const Launch2D& L, hipStream_t s) {
    const Consts& c = L.c;
    const size_t ncell = (size_t)c.nx * c.ny;
    const bool cs = L.flux_mode == HUGKS_FLUX_CELLSUM;
    if (L.flux_mode != HUGKS_FLUX_STENCIL && !cs) {
        HookScope hs(L, "slope2d");
        hipLaunchKernelGGL((k_slope2d<T>), dim3((unsigned)((ncell + SLOPE_CPW - 1) / SLOPE_CPW)),
                           dim3(WAVE * SLOPE_CPW), 0, s, c,
                           L.st, (const T*)L.h, (const T*)L.b, (T*)L.shx, (T*)L.shy, (T*)L.sbx,
                           (T*)L.sby, 0);
    }
    if (cs) {
        // 16-byte pieces where a row is a multiple of 16 bytes
        constexpr int VW = 16 / (int)sizeof(T);
        // the slopes and the per-cell sums in one pass
        HookScope hs(L, "slope_sum");
        auto kc = (c.nv * (int)sizeof(T)) % 16 == 0 ? k_cell_sum<T, VW> : k_cell_sum<T, 1>;
        hipLaunchKernelGGL(kc, dim3((unsigned)((ncell + SLOPE_CPW - 1) / SLOPE_CPW)),
                           dim3(WAVE * SLOPE_CPW), 0, s, c, L.st, (const T*)L.h, (const T*)L.b,
                           (T*)L.shx, (T*)L.shy, (T*)L.sbx, (T*)L.sby, L.u1, L.v1, L.wx, L.wy,
                           L.csum, L.cs_js);
    } else if (L.flux_mode == HUGKS_FLUX_STENCIL && L.nwall) {
        // the wall fluxes read the slope arrays of the wall cells
        HookScope hs(L, "slope2d");
        hipLaunchKernelGGL((k_slope2d<T>),
                           dim3((unsigned)((wall_cell_count(c) + SLOPE_CPW - 1) / SLOPE_CPW)),
                           dim3(WAVE * SLOPE_CPW), 0, s,
                           c, L.st, (const T*)L.h, (const T*)L.b, (T*)L.shx, (T*)L.shy,
                           (T*)L.sbx, (T*)L.sby, 1);
    }
    const bool xper = c.bc[0] == HUGKS_BC_PERIODIC, yper = c.bc[2] == HUGKS_BC_PERIODIC;
    const int nvf = (xper ? c.nx + 1 : c.nx - 1) * c.ny;
    const int nhf = c.nx * (yper ? c.ny + 1 : c.ny - 1);
    using A = typename Arith<T>::A;
    const A* un = (const A*)L.unode;
    const A* vn = (const A*)L.vnode;
    const A* wn = (const A*)L.wnode;
    const int fpw = FACES_PER_WG;
    const int nint = nvf + nhf;
    if (L.flux_mode != HUGKS_FLUX_FUSED && nint > 0) {
        const bool stn = L.flux_mode == HUGKS_FLUX_STENCIL;
        const unsigned gw = (unsigned)((nint + SPLIT_FPW - 1) / SPLIT_FPW);
        const unsigned gl = (unsigned)((nint + 255) / 256);
        if (cs) {
            // the face sums from the cell bins inside the lane-per-face
            // kernel (the cell-sum flux needs a tensor-product quadrature)
            HookScope hs(L, "flux_face");
            hipLaunchKernelGGL((k_fface_cs<T>), dim3(gl), dim3(256), 0, s, c, L.st, L.W,
                               (const T*)L.h, (const T*)L.b, (const T*)L.shx, (const T*)L.shy,
                               (const T*)L.sbx, (const T*)L.sby, L.u1, L.v1, L.wx, L.wy,
                               L.unode_d, L.vnode_d, L.wnode_d, L.node4, L.acc_v, L.acc_h,
                               L.csum, L.fsum, L.fcoef, L.vflux, L.hflux, nvf, nint);
        } else {
            HookScope hs(L, "flux_sum");
            auto kf = stn ? k_fsum<T, true> : k_fsum<T, false>;
            hipLaunchKernelGGL(kf, dim3(gw), dim3(WAVE * SPLIT_FPW), 0, s, c, L.st,
                               (const T*)L.h, (const T*)L.b, (const T*)L.shx, (const T*)L.shy,
                               (const T*)L.sbx, (const T*)L.sby, L.unode_d, L.vnode_d,
                               L.wnode_d, L.node4, L.acc_v, L.acc_h, L.fsum, nvf, nint);
        }
        if (!cs && L.wx) {
            // tensor-product quadrature: face state, Maxwellian sums and
            // conserved flux in one lane-per-face kernel
            HookScope hs(L, "flux_face");
            hipLaunchKernelGGL((k_fface), dim3(gl), dim3(256), 0, s, c, L.st, L.W, L.u1, L.v1,
                               L.wx, L.wy, L.acc_v, L.acc_h, L.fsum, L.fcoef, L.vflux, L.hflux,
                               nvf, nint);
        } else if (!cs) {
            {
                HookScope hs(L, "flux_state");
                hipLaunchKernelGGL((k_fstate), dim3(gl), dim3(256), 0, s, c, L.st, L.W, L.acc_v,
                                   L.acc_h, L.fsum, L.fcoef, nvf, nint);
            }
            {
                HookScope hs(L, "flux_maxw");
                hipLaunchKernelGGL((k_fmaxw<T>), dim3(gw), dim3(WAVE * SPLIT_FPW), 0, s, c, L.st,
                                   L.unode_d, L.vnode_d, L.wnode_d, L.acc_v, L.acc_h, L.fcoef,
                                   L.fmaxw, nvf, nint);
            }
            {
                HookScope hs(L, "flux_slope_t");
                hipLaunchKernelGGL((k_fslope_t), dim3(gl), dim3(256), 0, s, c, L.st, L.acc_v,
                                   L.acc_h, L.fmaxw, L.fcoef, L.vflux, L.hflux, nvf, nint);
            }
        }
        // (update mode 3: k_fdiv below instead)
        if (L.update_mode != 3) {
            HookScope hs(L, "flux_node");
            auto kn = stn ? k_fnode<T, true> : k_fnode<T, false>;
            hipLaunchKernelGGL(kn, dim3(gw), dim3(WAVE * SPLIT_FPW), 0, s, c, L.st,
                               (const T*)L.h, (const T*)L.b, (const T*)L.shx, (const T*)L.shy,
                               (const T*)L.sbx, (const T*)L.sby, (const A*)L.uvnode, L.acc_v,
                               L.acc_h, L.fcoef, (T*)L.vfh, (T*)L.vfb, (T*)L.hfh, (T*)L.hfb, nvf,
                               nint);
        }
    } else if (nvf > 0) {
        HookScope hs(L, "flux_vface");
        hipLaunchKernelGGL((k_flux_interior<0, T>), dim3((unsigned)((nvf + fpw - 1) / fpw)),
                           dim3(WAVE * fpw), 0, s, c, L.st, L.W, (const T*)L.h, (const T*)L.b,
                           (const T*)L.shx, (const T*)L.shy, (const T*)L.sbx, (const T*)L.sby, un,
                           vn, wn, L.acc_v, (T*)L.vfh, (T*)L.vfb, L.vflux, nvf);
    }
    if (L.flux_mode == HUGKS_FLUX_FUSED && nhf > 0) {
        HookScope hs(L, "flux_hface");
        hipLaunchKernelGGL((k_flux_interior<1, T>), dim3((unsigned)((nhf + fpw - 1) / fpw)),
                           dim3(WAVE * fpw), 0, s, c, L.st, L.W, (const T*)L.h, (const T*)L.b,
                           (const T*)L.shx, (const T*)L.shy, (const T*)L.sbx, (const T*)L.sby, un,
                           vn, wn, L.acc_h, (T*)L.hfh, (T*)L.hfb, L.hflux, nhf);
    }
    if (L.nwall) {
        HookScope hs(L, "flux_wall");
        hipLaunchKernelGGL((k_flux_wall<T>), dim3((L.nwall + 3) / 4), dim3(256), 0, s, c, L.st,
                           (const T*)L.h, (const T*)L.b, (const T*)L.shx, (const T*)L.shy,
                           (const T*)L.sbx, (const T*)L.sby, un, vn, wn, (T*)L.vfh, (T*)L.vfb,
                           L.vflux, (T*)L.hfh, (T*)L.hfb, L.hflux, (int)L.nwall);
    }
    if (L.update_mode == 3) {
        // the face divergence but the equilibrium part of the interior
        // faces, which the update forms on its tiles (after the walls: it
        // reads their node fluxes)
        HookScope hs(L, "flux_div");
        constexpr int VW = 16 / (int)sizeof(T);
        auto kd = (c.nv * (int)sizeof(T)) % 16 == 0 ? k_fdiv<T, VW> : k_fdiv<T, 1>;
        hipLaunchKernelGGL(kd, dim3((unsigned)((ncell + DIV_CPW - 1) / DIV_CPW)),
                           dim3(WAVE * DIV_CPW), 0, s, c, L.st, (const T*)L.h, (const T*)L.b,
                           (const T*)L.shx, (const T*)L.shy, (const T*)L.sbx, (const T*)L.sby,
                           L.u1, L.v1, L.acc_v, L.acc_h, L.fcoef, (const T*)L.vfh,
                           (const T*)L.vfb, (const T*)L.hfh, (const T*)L.hfb, (T*)L.dvh,
                           (T*)L.dvb);
    }
    return hipGetLastError();
}

hipError_t launch_flux2d(const Launch2D& L, hipStream_t s) {
    return L.fp64 ? launch_flux2d_t<double>(L, s) : launch_flux2d_t<float>(L, s);
}

}  // namespace hugks
