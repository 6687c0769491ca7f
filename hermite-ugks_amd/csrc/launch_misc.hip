// Host launchers for the kernels of k_misc.h.
#include "k_misc.h"

namespace hugks {

// Workgroups of the first reduction pass; 0 (one single-workgroup pass)
// for grids a single workgroup reduces quickly, where a second launch
// would cost more than it saves.
static int pred_blocks(const Consts& c) {
    if (c.nx * c.ny < PRED_MIN_CELLS) return 0;
    const int n = (c.nx * c.ny + PRED_NT - 1) / PRED_NT;
    return n < PRED_NB ? n : PRED_NB;
}

hipError_t launch_finalize(const Launch2D& L, int dim, hipStream_t s) {
    HookScope hs(L, "finalize");
    const int nb = pred_blocks(L.c);
    double* red = L.red + PRED_NB;
    if (nb > 0)
        hipLaunchKernelGGL((k_part_sum), dim3(nb), dim3(PRED_NT), 0, s, L.c, dim, L.st, L.part, red);
    hipLaunchKernelGGL((k_finalize), dim3(1), dim3(RED_NT), 0, s, L.c, dim, L.st, L.part, red, nb);
    return hipGetLastError();
}

hipError_t launch_begin(DevState* st, double t_stop, hipStream_t s) {
    hipLaunchKernelGGL((k_begin), dim3(1), dim3(64), 0, s, st, t_stop);
    return hipGetLastError();
}
hipError_t launch_begin_all(DevState* const* sts, int n, double t_stop, hipStream_t s) {
    hipLaunchKernelGGL((k_begin_all), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sts, n,
                       t_stop);
    return hipGetLastError();
}
hipError_t launch_gather_states(DevState* const* sts, int n, DevState* out, hipStream_t s) {
    hipLaunchKernelGGL((k_gather_states), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sts,
                       n, out);
    return hipGetLastError();
}

hipError_t launch_timestep(const Launch2D& L, int dim, hipStream_t s) {
    HookScope hs(L, "timestep");
    const int nb = pred_blocks(L.c);
    if (nb > 0)
        hipLaunchKernelGGL((k_rate_part), dim3(nb), dim3(PRED_NT), 0, s, L.c, dim, L.st, L.W, L.red);
    hipLaunchKernelGGL((k_timestep), dim3(1), dim3(RED_NT), 0, s, L.c, dim, L.st, L.W, L.red, nb);
    return hipGetLastError();
}

hipError_t launch_init_f(const Launch2D& L, int dim, const double* prim0, hipStream_t s) {
    const size_t n = (size_t)L.c.nx * L.c.ny * L.c.ldv;
    hipLaunchKernelGGL((k_init_f), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, L.c, dim, prim0, L.u1, L.v1, L.W, L.h,
                                                         L.b, L.fp64 ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_macro(const Launch2D& L, int dim, double* q, hipStream_t s) {
    const unsigned ncell = (unsigned)(L.c.nx * L.c.ny);
    if (L.fp64)
        hipLaunchKernelGGL((k_macro<double>), dim3(ncell), dim3(256), 0, s, L.c, dim, L.W, (const double*)L.h,
                                              (const double*)L.b, L.u1, L.v1, L.wnode_d, q);
    else
        hipLaunchKernelGGL((k_macro<float>), dim3(ncell), dim3(256), 0, s, L.c, dim, L.W, (const float*)L.h, (const float*)L.b,
                                             L.u1, L.v1, L.wnode_d, q);
    return hipGetLastError();
}

hipError_t launch_convert(const void* src, bool src_fp64, double* dst, size_t ncell, int nn,
                          int ldv, bool, hipStream_t s) {
    const size_t n = ncell * nn;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (src_fp64) hipLaunchKernelGGL((k_to_dense<double>), dim3(g), dim3(256), 0, s, (const double*)src, dst, ncell, nn, ldv);
    else hipLaunchKernelGGL((k_to_dense<float>), dim3(g), dim3(256), 0, s, (const float*)src, dst, ncell, nn, ldv);
    return hipGetLastError();
}

hipError_t launch_convert_in(const double* src, void* dst, bool dst_fp64, size_t ncell, int nn,
                             int ldv, hipStream_t s) {
    const size_t n = ncell * ldv;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (dst_fp64) hipLaunchKernelGGL((k_from_dense<double>), dim3(g), dim3(256), 0, s, src, (double*)dst, ncell, nn, ldv);
    else hipLaunchKernelGGL((k_from_dense<float>), dim3(g), dim3(256), 0, s, src, (float*)dst, ncell, nn, ldv);
    return hipGetLastError();
}

}  // namespace hugks
