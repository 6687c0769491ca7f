// Host launchers for the kernels of k_update2d.h.
#include "k_update2d.h"
#include "k_update2d_mfma.h"
#include "k_update2d_mx.h"

namespace hugks {

size_t update2d_lds_bytes(const Consts& c, bool fp64) {
    const size_t nv = c.nv, nm = c.nm, nn = c.nn;
    const size_t n = 2 * nn + 2 * (nm + 2) * nv + 2 * (nm + 1) * nv + 2 * (nm + 1) * (nm + 1) +
                     2 * (nm + 2) * nv;
    return n * (fp64 ? sizeof(double) : sizeof(float));
}

template <typename T>
static hipError_t launch_update_t(const Launch2D& L, hipStream_t s, bool prepare) {
    const Consts& c = L.c;
    constexpr int NT = 256;
    const int ept = (c.ldv + NT - 1) / NT;
    const unsigned ncell = (unsigned)(c.nx * c.ny);
    using A = typename Arith<T>::A;
#define HUGKS_UPD_CASE(E)                                                                       \
    case E: {                                                                                   \
        if (prepare)                                                                            \
            return hipFuncSetAttribute((const void*)k_update2d<T, E, NT>,                       \
                                       hipFuncAttributeMaxDynamicSharedMemorySize,              \
                                       (int)L.update_lds);                                      \
        HookScope hs(L, "update2d");                                                            \
        hipLaunchKernelGGL((k_update2d<T, E, NT>), dim3(ncell), dim3(NT), L.update_lds, s, \
            c, L.st, L.W, L.part, (T*)L.h, (T*)L.b, (const T*)L.vfh, (const T*)L.vfb,           \
            (const T*)L.hfh, (const T*)L.hfb, L.vflux, L.hflux, (const A*)L.unode,              \
            (const A*)L.vnode, (const A*)L.wnode, L.wnode_d, L.u1, L.v1, L.acc_cell);          \
        break;                                                                                  \
    }
    switch (ept) {
        HUGKS_UPD_CASE(1)
        HUGKS_UPD_CASE(2)
        HUGKS_UPD_CASE(3)
        HUGKS_UPD_CASE(4)
        HUGKS_UPD_CASE(5)
        HUGKS_UPD_CASE(6)
        HUGKS_UPD_CASE(7)
        HUGKS_UPD_CASE(8)
        default: return hipErrorInvalidValue;
    }
#undef HUGKS_UPD_CASE
    return hipGetLastError();
}

// (the finite-difference force has no Hermite contraction: LDS kernel only)
bool update2d_mfma_supported(const Consts& c) {
    return c.nv <= 48 && c.nm <= MF_NB - 2 && !c.force_fd;
}

size_t update2d_mfma_lds_bytes(const Consts& c, bool fp64) {
    return mf_lds_bytes(c.nv, fp64 ? sizeof(double) : sizeof(float));
}

template <typename T>
static hipError_t launch_update_mfma_t(const Launch2D& L, hipStream_t s, bool prepare) {
    const Consts& c = L.c;
    const UpdKey key = update_key(c, sizeof(T) == 8);
    const unsigned ncell = (unsigned)(c.nx * c.ny);
    const size_t lds = update2d_mfma_lds_bytes(c, sizeof(T) == 8);
    using A = typename Arith<T>::A;
    if (!prepare) {
        HookScope hs(L, "cell_pre");
        hipLaunchKernelGGL((k_cell_pre), dim3((ncell + 255) / 256), dim3(256), 0, s, c, L.st, L.W,
                           L.part, L.vflux, L.hflux, L.acc_cell, L.cellpre);
    }
#define HUGKS_UPDM_CASE(E, N, V)                                                                \
    if (key.ept == E && key.nt == N && key.nvc == V) {                                          \
        if (prepare)                                                                            \
            return hipFuncSetAttribute((const void*)k_update2d_mfma<T, E, N, V>,                \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);   \
        HookScope hs(L, "update2d");                                                            \
        hipLaunchKernelGGL((k_update2d_mfma<T, E, N, V>), dim3(ncell), dim3(N), lds, s, c,      \
            L.st, L.cellpre, (T*)L.h, (T*)L.b, (const T*)L.vfh, (const T*)L.vfb,                \
            (const T*)L.hfh, (const T*)L.hfb, (const A*)L.wnode, L.wnode_d, L.u1, L.v1);        \
        return hipGetLastError();                                                               \
    }
    if (sizeof(T) == 4) {
        HUGKS_UPDM_F32(HUGKS_UPDM_CASE)
    } else {
        HUGKS_UPDM_F64(HUGKS_UPDM_CASE)
    }
    HUGKS_UPDM_ANY(HUGKS_UPDM_CASE)
#undef HUGKS_UPDM_CASE
    return hipErrorInvalidValue;
}

// LDS element size: the storage type, or the two dwords of a bf16x3 operand
static size_t mx_elem(bool fp64, bool b3) {
    return fp64 ? sizeof(double) : b3 ? sizeof(B3) : sizeof(float);
}

bool update2d_mx_supported(const Consts& c, bool fp64, bool b3, size_t lds_limit, int nw) {
    int mp, nvc;
    return update2d_mfma_supported(c) && !(fp64 && b3) && mx_key(c.nv, nw, &mp, &nvc, b3) &&
           mx_lds_bytes(c.nv, mx_elem(fp64, b3)) <= lds_limit;
}

int update2d_mx_default_waves(const Consts& c) { return mx_default_nw(c.nv); }

// The face-forming update in FP64 at nv 33..48 runs faster with five
// wavefronts per cell than with three (9.09 against 9.76 ms at 512^2 cells,
// GH40; the plain node update the other way round, 10.29 against 8.35).
int update2d_default_waves(const Launch2D& L, int mode) {
    const int nw = mx_default_nw(L.c.nv);
    if (mode == 3 && L.fp64 && nw == 3 && update2d_mxf_supported(L, 5)) return 5;
    return nw;
}

bool update2d_mxf_supported(const Launch2D& L, int nw) {
    const Consts& c = L.c;
    int mp, nvc;
    const size_t fbytes = (size_t)c.nx * c.ny * c.ldv * (L.fp64 ? sizeof(double) : sizeof(float));
    return !L.b3 && update2d_mfma_supported(c) && mx_key(c.nv, nw, &mp, &nvc, false) &&
           mx_lds_bytes(c.nv, mx_elem(L.fp64, false), true) <= L.lds_limit && L.dvh && L.dvb &&
           fbytes < (size_t)0xFFFFFFF0u;
}

template <typename T, bool SB3>
static hipError_t launch_update_mx_t(const Launch2D& L, hipStream_t s, int nw, bool prepare,
                                     bool ff) {
    const Consts& c = L.c;
    int maxp = 0, nvc = 0;
    if (!mx_key(c.nv, nw, &maxp, &nvc, SB3)) return hipErrorInvalidValue;
    if (ff && SB3) return hipErrorInvalidValue;
    const unsigned ncell = (unsigned)(c.nx * c.ny);
    const size_t lds = mx_lds_bytes(c.nv, mx_elem(sizeof(T) == 8, SB3), ff);
    const MxFace fx{L.dvh, L.dvb, L.fcoef, L.acc_v, L.acc_h};
    if (!prepare) {
        HookScope hs(L, "cell_pre");
        hipLaunchKernelGGL((k_cell_pre), dim3((ncell + 255) / 256), dim3(256), 0, s, c, L.st, L.W,
                           L.part, L.vflux, L.hflux, L.acc_cell, L.cellpre);
    }
#define HUGKS_UPDX_CASE(W, P, V)                                                              \
    if (nw == W && maxp == P && nvc == V && ff) {                                             \
        if constexpr (!SB3) {                                                                 \
            if (prepare)                                                                      \
                return hipFuncSetAttribute((const void*)k_update2d_mxf<T, W, P, V>,           \
                                           hipFuncAttributeMaxDynamicSharedMemorySize,        \
                                           (int)lds);                                         \
            HookScope hs(L, "update2d");                                                      \
            hipLaunchKernelGGL((k_update2d_mxf<T, W, P, V>), dim3(ncell), dim3(W * WAVE),     \
                lds, s, c, L.st, L.cellpre, (T*)L.h, (T*)L.b, (const T*)L.vfh,                \
                (const T*)L.vfb, (const T*)L.hfh, (const T*)L.hfb,                            \
                (const typename Arith<T>::A*)L.wnode, L.u1, L.v1, fx);                        \
            return hipGetLastError();                                                         \
        }                                                                                     \
    }                                                                                         \
    if (nw == W && maxp == P && nvc == V && !ff) {                                            \
        if (prepare)                                                                          \
            return hipFuncSetAttribute((const void*)k_update2d_mx<T, W, P, V, SB3>,           \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
        HookScope hs(L, "update2d");                                                          \
        hipLaunchKernelGGL((k_update2d_mx<T, W, P, V, SB3>), dim3(ncell), dim3(W * WAVE), lds, \
            s, c, L.st, L.cellpre, (T*)L.h, (T*)L.b, (const T*)L.vfh, (const T*)L.vfb,        \
            (const T*)L.hfh, (const T*)L.hfb, (const typename Arith<T>::A*)L.wnode, L.u1,     \
            L.v1);                                                                            \
        return hipGetLastError();                                                             \
    }
    if constexpr (SB3) {
        HUGKS_UPDX_B3(HUGKS_UPDX_CASE)
    } else {
        HUGKS_UPDX(HUGKS_UPDX_CASE)
    }
#undef HUGKS_UPDX_CASE
    return hipErrorInvalidValue;
}

static hipError_t launch_update_mx(const Launch2D& L, hipStream_t s, int nw, bool prepare,
                                   bool ff = false) {
    if (L.fp64) return launch_update_mx_t<double, false>(L, s, nw, prepare, ff);
    return L.b3 ? launch_update_mx_t<float, true>(L, s, nw, prepare, ff)
                : launch_update_mx_t<float, false>(L, s, nw, prepare, ff);
}

hipError_t launch_update2d(const Launch2D& L, hipStream_t s) {
    if (L.update_mode == 3) return launch_update_mx(L, s, L.update_nw, false, true);
    if (L.update_mode == 2) return launch_update_mx(L, s, L.update_nw, false);
    if (L.update_mode == 1)
        return L.fp64 ? launch_update_mfma_t<double>(L, s, false)
                      : launch_update_mfma_t<float>(L, s, false);
    return L.fp64 ? launch_update_t<double>(L, s, false) : launch_update_t<float>(L, s, false);
}

// Dynamic-LDS limits of both update kernels of this case, set once before
// any launch (so that no launch, captured into a graph or not, needs to).
hipError_t prepare_update2d(const Launch2D& L) {
    hipError_t e = L.fp64 ? launch_update_t<double>(L, nullptr, true)
                          : launch_update_t<float>(L, nullptr, true);
    if (e == hipSuccess && update2d_mfma_supported(L.c))
        e = L.fp64 ? launch_update_mfma_t<double>(L, nullptr, true)
                   : launch_update_mfma_t<float>(L, nullptr, true);
    for (int nw = 1; nw <= 16 && e == hipSuccess; ++nw) {
        if (!update2d_mx_supported(L.c, L.fp64, L.b3, L.lds_limit, nw)) continue;
        e = launch_update_mx(L, nullptr, nw, true);
        if (e == hipSuccess && update2d_mxf_supported(L, nw))
            e = launch_update_mx(L, nullptr, nw, true, true);
    }
    return e;
}

}  // namespace hugks
