// Host launchers of the 2D slope, interior-face and wall-face kernels.
#include "k_flux2d.h"
#include "k_flux_face.h"

namespace hugks {

template <typename T>
static hipError_t launch_flux2d_t(const Launch2D& L, hipStream_t s) {
    const Consts& c = L.c;
    const size_t ncell = (size_t)c.nx * c.ny;
    {
        HookScope hs(L, "slope2d");
        const size_t n = ncell * c.ldv;
        hipLaunchKernelGGL((k_slope2d<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c,
                           L.st, (const T*)L.h, (const T*)L.b, (T*)L.shx, (T*)L.shy, (T*)L.sbx,
                           (T*)L.sby);
    }
    const bool xper = c.bc[0] == HUGKS_BC_PERIODIC, yper = c.bc[2] == HUGKS_BC_PERIODIC;
    const int nvf = (xper ? c.nx + 1 : c.nx - 1) * c.ny;
    const int nhf = c.nx * (yper ? c.ny + 1 : c.ny - 1);
    using A = typename Arith<T>::A;
    const A* un = (const A*)L.unode;
    const A* vn = (const A*)L.vnode;
    const A* wn = (const A*)L.wnode;
    const int fpw = FACES_PER_WG;
    if (nvf > 0) {
        HookScope hs(L, "flux_vface");
        hipLaunchKernelGGL((k_flux_interior<0, T>), dim3((unsigned)((nvf + fpw - 1) / fpw)),
                           dim3(WAVE * fpw), 0, s, c, L.st, L.W, (const T*)L.h, (const T*)L.b,
                           (const T*)L.shx, (const T*)L.shy, (const T*)L.sbx, (const T*)L.sby, un,
                           vn, wn, L.acc_v, (T*)L.vfh, (T*)L.vfb, L.vflux, nvf);
    }
    if (nhf > 0) {
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
    return hipGetLastError();
}

hipError_t launch_flux2d(const Launch2D& L, hipStream_t s) {
    return L.fp64 ? launch_flux2d_t<double>(L, s) : launch_flux2d_t<float>(L, s);
}

}  // namespace hugks
