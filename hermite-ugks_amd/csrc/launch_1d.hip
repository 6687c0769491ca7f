// Host launchers for the kernels of k_1d.h.
#include "k_1d.h"

namespace hugks {

template <typename T>
static hipError_t step1d_t(const Launch2D& L, hipStream_t s) {
    const Consts& c = L.c;
    using A = typename Arith<T>::A;
    const size_t n = (size_t)c.nx * c.ldv;
    HookScope hs(L, "step1d");
    hipLaunchKernelGGL((k_slope1d<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c, L.st, (const T*)L.h, (const T*)L.b,
                                                              (T*)L.shx, (T*)L.sbx);
    const int ept = (c.ldv + 63) / 64;
    const A* un = (const A*)L.unode;
    const A* wn = (const A*)L.wnode;
#define HUGKS_1D_CASE(E)                                                                       \
    case E:                                                                                    \
        hipLaunchKernelGGL((k_wall1d<T, E>), dim3(2), dim3(64), 0, s, c, L.st, (const T*)L.h, (const T*)L.b, (const T*)L.shx, \
                                        (const T*)L.sbx, un, wn, (T*)L.vfh, (T*)L.vfb, L.vflux); \
        if (c.nx > 1)                                                                          \
            hipLaunchKernelGGL((k_flux1d<T, E>), dim3(c.nx - 1), dim3(64), 0, s, c, L.st, L.W, (const T*)L.h, (const T*)L.b,  \
                                                   (const T*)L.shx, (const T*)L.sbx, un, wn,   \
                                                   (T*)L.vfh, (T*)L.vfb, L.vflux);             \
        hipLaunchKernelGGL((k_update1d<T, E>), dim3(c.nx), dim3(64), 0, s, c, L.st, L.W, L.part, (T*)L.h, (T*)L.b,            \
                                             (const T*)L.vfh, (const T*)L.vfb, L.vflux, un, wn, L.unode_d, L.wnode_d); \
        break;
    switch (ept) {
        HUGKS_1D_CASE(1)
        HUGKS_1D_CASE(2)
        HUGKS_1D_CASE(3)
        HUGKS_1D_CASE(4)
        default: return hipErrorInvalidValue;
    }
#undef HUGKS_1D_CASE
    return hipGetLastError();
}

hipError_t launch_step1d(const Launch2D& L, hipStream_t s) {
    return L.fp64 ? step1d_t<double>(L, s) : step1d_t<float>(L, s);
}

}  // namespace hugks
