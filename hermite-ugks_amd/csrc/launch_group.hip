// Host launchers of the grouped batch kernels (k_group.h).
// the kernel bodies only: the single-case kernels belong to the other
// launch files
#define HUGKS_BODIES_ONLY
#include "k_group.h"

namespace hugks {

void group_blocks(const Launch2D& L, int* slope, int* fsum, int* fface, int* fnode, int* wall,
                  int* cpre) {
    const Consts& c = L.c;
    int nvf, nint;
    face_counts(c, nvf, nint);
    const int ncell = c.nx * c.ny;
    *slope = (ncell + SLOPE_CPW - 1) / SLOPE_CPW;
    *fsum = *fnode = (nint + SPLIT_FPW - 1) / SPLIT_FPW;
    *fface = (nint + 255) / 256;
    *wall = ((int)L.nwall + 3) / 4;
    *cpre = (ncell + 255) / 256;
}

void group_update_key(const Launch2D& L, int* nw, int* maxp, int* nvc) {
    *nw = L.update_nw;
    if (!mx_key(L.c.nv, L.update_nw, maxp, nvc)) *maxp = *nvc = -1;
}

size_t group_update_lds(const Launch2D& L) {
    return mx_lds_bytes(L.c.nv, L.fp64 ? sizeof(double) : sizeof(float), L.update_mode == 3);
}

int group_fdiv_blocks(const Launch2D& L) { return (L.c.nx * L.c.ny + DIV_CPW - 1) / DIV_CPW; }

template <typename T>
static hipError_t group_update_t(const GroupUpd& u, const Launch2D* Ls, hipStream_t s,
                                 bool prepare) {
#define HUGKS_GUPD_CASE(W, P, V)                                                            \
    if (u.nw == W && u.maxp == P && u.nvc == V) {                                           \
        const void* kf = u.ff ? (const void*)kg_update<T, W, P, V, true>                    \
                              : (const void*)kg_update<T, W, P, V, false>;                  \
        if (prepare)                                                                        \
            return hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize,      \
                                       (int)u.lds);                                         \
        if (u.ff)                                                                           \
            hipLaunchKernelGGL((kg_update<T, W, P, V, true>), dim3(u.blocks),               \
                               dim3(W * WAVE), u.lds, s, u.tab, Ls);                        \
        else                                                                                \
            hipLaunchKernelGGL((kg_update<T, W, P, V, false>), dim3(u.blocks),              \
                               dim3(W * WAVE), u.lds, s, u.tab, Ls);                        \
        return hipGetLastError();                                                           \
    }
    HUGKS_UPDX(HUGKS_GUPD_CASE)
#undef HUGKS_GUPD_CASE
    return hipErrorInvalidValue;
}

hipError_t prepare_group_update(bool fp64, const GroupUpd& u) {
    return fp64 ? group_update_t<double>(u, nullptr, nullptr, true)
                : group_update_t<float>(u, nullptr, nullptr, true);
}

template <typename T>
static hipError_t group_step_t(const GroupPlan& P, hipStream_t s) {
    {
        HookCall hc(P.hook, P.hook_ctx, "kg_timestep");
        hipLaunchKernelGGL((kg_timestep), dim3(P.g_ts), dim3(RED_NT), 0, s, P.t_ts, P.Ls);
    }
    if (P.cellsum) {
        {
            constexpr int VW = 16 / (int)sizeof(T);
            HookCall hc(P.hook, P.hook_ctx, "kg_cell_sum");
            auto kc = P.cs_vec ? kg_cell_sum<T, VW> : kg_cell_sum<T, 1>;
            hipLaunchKernelGGL(kc, dim3(P.g_slope), dim3(WAVE * SLOPE_CPW), 0, s, P.t_slope,
                               P.Ls);
        }
        {
            HookCall hc(P.hook, P.hook_ctx, "kg_fface");
            hipLaunchKernelGGL((kg_fface_cs<T>), dim3(P.g_fface), dim3(256), 0, s, P.t_fface,
                               P.Ls);
        }
    } else {
        {
            HookCall hc(P.hook, P.hook_ctx, "kg_slope");
            hipLaunchKernelGGL((kg_slope<T>), dim3(P.g_slope), dim3(WAVE * SLOPE_CPW), 0, s,
                               P.t_slope, P.Ls);
        }
        {
            HookCall hc(P.hook, P.hook_ctx, "kg_fsum");
            hipLaunchKernelGGL((kg_fsum<T>), dim3(P.g_fsum), dim3(WAVE * SPLIT_FPW), 0, s,
                               P.t_fsum, P.Ls);
        }
        {
            HookCall hc(P.hook, P.hook_ctx, "kg_fface");
            hipLaunchKernelGGL((kg_fface), dim3(P.g_fface), dim3(256), 0, s, P.t_fface, P.Ls);
        }
    }
    if (!P.face) {
        HookCall hc(P.hook, P.hook_ctx, "kg_fnode");
        hipLaunchKernelGGL((kg_fnode<T>), dim3(P.g_fnode), dim3(WAVE * SPLIT_FPW), 0, s,
                           P.t_fnode, P.Ls);
    }
    if (P.g_wall > 0) {
        HookCall hc(P.hook, P.hook_ctx, "kg_wall");
        hipLaunchKernelGGL((kg_wall<T>), dim3(P.g_wall), dim3(256), 0, s, P.t_wall, P.Ls);
    }
    if (P.face) {
        // after the walls: the divergence reads their node fluxes
        constexpr int VW = 16 / (int)sizeof(T);
        HookCall hc(P.hook, P.hook_ctx, "kg_fdiv");
        auto kd = P.cs_vec ? kg_fdiv<T, VW> : kg_fdiv<T, 1>;
        hipLaunchKernelGGL(kd, dim3(P.g_fdiv), dim3(WAVE * DIV_CPW), 0, s, P.t_fdiv, P.Ls);
    }
    {
        HookCall hc(P.hook, P.hook_ctx, "kg_cell_pre");
        hipLaunchKernelGGL((kg_cell_pre), dim3(P.g_cpre), dim3(256), 0, s, P.t_cpre, P.Ls);
    }
    for (int q = 0; q < P.nupd; ++q) {
        HookCall hc(P.hook, P.hook_ctx, "kg_update");
        hipError_t e = group_update_t<T>(P.upd[q], P.Ls, s, false);
        if (e != hipSuccess) return e;
    }
    {
        HookCall hc(P.hook, P.hook_ctx, "kg_finalize");
        hipLaunchKernelGGL((kg_finalize), dim3(P.g_fin), dim3(RED_NT), 0, s, P.t_fin, P.Ls);
    }
    return hipGetLastError();
}

hipError_t launch_group_step(const GroupPlan& P, hipStream_t s) {
    return P.fp64 ? group_step_t<double>(P, s) : group_step_t<float>(P, s);
}

}  // namespace hugks
