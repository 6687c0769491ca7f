// Host-side launch interface of the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include "hugks.h"
#include "hugks_device.h"

namespace hugks {

// Optional per-launch hook (timing): called with phase 0 right before and
// phase 1 right after each kernel launch on the launch stream.
typedef void (*LaunchHook)(void* ctx, int phase, const char* name);

// Everything one 2D/1D step touches on the device (pointers are device
// pointers; T-typed arrays are void* and interpreted by precision).
struct Launch2D {
    Consts c;
    bool fp64;
    bool b3;             // bf16x3 split-product Hermite contractions (FP32 storage)
    DevState* st;
    double* W;
    double* part;        // per-cell residual partials [5*ncell]
    void *h, *b, *shx, *shy, *sbx, *sby;
    void *vfh, *vfb, *hfh, *hfb;
    void *dvh, *dvb;     // update mode 3: per-cell face divergence of the free transport
                         // and the walls (k_fdiv; [ncell][ldv], or null)
    double *vflux, *hflux;
    const void *unode, *vnode, *wnode;       // per-node tables in the arithmetic type
    const double *unode_d, *vnode_d, *wnode_d;  // FP64 tables (moments)
    const double* node4;     // per node {u, v, w, 0} (FP64), 32-byte records
    const void* uvnode;      // per node {u, v} in the arithmetic type
    const double *u1, *v1;                   // 1D node lists [nv] (double)
    const double *wx, *wy;   // weight factors w[i + nv j] = wx[i] wy[j]; null if the
                             // weights do not factor
    const double *acc_cell, *acc_v, *acc_h;
    unsigned nwall;      // number of wall faces (non-periodic sides)
    int flux_mode;       // interior faces: 0 one fused kernel, 1 five-kernel pipeline,
                         // 2 the pipeline with slopes from the cell stencils, 3 the
                         // pipeline with the face sums formed per cell in the slope pass
    int cs_js;           // flux mode 3: first v row with v >= 0 (rows below: v < 0);
                         // -1 where the mode is unavailable (nv > 64, unsorted v)
    int stencil_ok;      // distribution arrays under 4 GiB (32-bit byte offsets)
    int update_mode;     // cell update: 0 LDS dot products, 1 MFMA Hermite tiles,
                         // 2 MFMA node update (force and collision on tiles), 3 mode 2
                         // forming the equilibrium part of the interior faces' node
                         // fluxes too (flux modes 1 and 3; k_fdiv replaces K5)
    int update_nw;       // wavefronts per cell of update mode 2
    size_t lds_limit;    // largest dynamic LDS of one workgroup on the device
    double *fsum, *fcoef, *fmaxw;  // per-face scratch of the pipeline
    double* csum;        // flux mode 3: per-cell half-space sums [ncell][4][CELL_SUM_LD]
    double* cellpre;     // per-cell scalars of the MFMA update [ncell][24]
    double* red;         // partials of the cell reductions [RED_SCRATCH]
    int update_nt;       // threads per update workgroup
    size_t update_lds;   // dynamic LDS bytes of the update kernel
    LaunchHook hook;     // may be null
    void* hook_ctx;
};

// Leading dimensions of the per-face scratch of the interior-face pipeline
// (k_flux_split.h).
constexpr int FACE_SUM_LD = 40;
constexpr int FACE_COEF_LD = 40;
constexpr int FACE_MAXW_LD = 16;
// Coefficient slots of one face in the per-face scratch fcoef (FACE_COEF_LD
// doubles per face), written by the face kernels of k_flux_split.h and
// k_flux_cellsum.h.
enum : int {
    FC_P = 0,      // prim at the face, 4
    FC_TAU = 4,
    FC_MT = 5,     // Mt(1..6) at 5..10
    FC_AL = 11,    // 4
    FC_AR = 15,    // 4
    FC_AB = 19,    // 4
    FC_FL = 23,    // 4: f0 and force part of the macro flux
    FC_Q = 27,     // heat flux, 2
    FC_AT = 29,    // 4 (K4)
};
// Leading dimension of one half-space bin of the per-cell sums (flux mode 3).
constexpr int CELL_SUM_LD = 40;
// Grids below this many cells reduce over their cells in one workgroup (one
// launch); larger ones in two passes.
constexpr int PRED_MIN_CELLS = 16384;
// Doubles of the cell-reduction partials (k_misc.h: PRED_NB x (1 + PRED_LD)).
constexpr int RED_SCRATCH = 1024 * 10;

// Brackets one launch with a hook given by function and context.
struct HookCall {
    LaunchHook f;
    void* ctx;
    const char* name;
    HookCall(LaunchHook fn, void* c, const char* n) : f(fn), ctx(c), name(n) {
        if (f) f(ctx, 0, name);
    }
    ~HookCall() {
        if (f) f(ctx, 1, name);
    }
};

// Brackets one launch with the hook.
struct HookScope {
    const Launch2D& L;
    const char* name;
    HookScope(const Launch2D& l, const char* n) : L(l), name(n) {
        if (L.hook) L.hook(L.hook_ctx, 0, name);
    }
    ~HookScope() {
        if (L.hook) L.hook(L.hook_ctx, 1, name);
    }
};

hipError_t launch_begin(DevState* st, double t_stop, hipStream_t s);
// batch forms: the states of n cases (device array of pointers)
hipError_t launch_begin_all(DevState* const* sts, int n, double t_stop, hipStream_t s);
hipError_t launch_gather_states(DevState* const* sts, int n, DevState* out, hipStream_t s);
hipError_t launch_timestep(const Launch2D& L, int dim, hipStream_t s);
hipError_t launch_flux2d(const Launch2D& L, hipStream_t s);
hipError_t launch_update2d(const Launch2D& L, hipStream_t s);
hipError_t prepare_update2d(const Launch2D& L);
hipError_t launch_finalize(const Launch2D& L, int dim, hipStream_t s);
hipError_t launch_init_f(const Launch2D& L, int dim, const double* prim0, hipStream_t s);
hipError_t launch_macro(const Launch2D& L, int dim, double* q, hipStream_t s);
hipError_t launch_convert(const void* src, bool src_fp64, double* dst, size_t ncell, int nn,
                          int ldv, bool to_dense, hipStream_t s);
hipError_t launch_convert_in(const double* src, void* dst, bool dst_fp64, size_t ncell, int nn,
                             int ldv, hipStream_t s);

// Grouped launches of a batch (k_group.h): device tables and grid sizes of
// one batch step; updates are grouped by kernel instance.
struct GroupUpd {
    int nw, maxp, nvc;   // the node-update instance (mx_key)
    bool ff;             // the instance forming the interior faces' equilibrium flux
    int blocks;
    size_t lds;          // largest LDS of the instance's cases
    const int* tab;
};
struct GroupPlan {
    bool fp64;
    bool cellsum;        // the cell-sum pipeline (every case supports it), else the split one
    bool cs_vec;         // its 16-byte loads (every case's velocity row a multiple of 16 bytes)
    bool face;           // update mode 3 (every case supports it): k_fdiv instead of K5
    const Launch2D* Ls;  // per-case arguments, device memory
    const int *t_ts, *t_slope, *t_fsum, *t_fface, *t_fnode, *t_wall, *t_cpre, *t_fin, *t_fdiv;
    int g_ts, g_slope, g_fsum, g_fface, g_fnode, g_wall, g_cpre, g_fin, g_fdiv;
    GroupUpd upd[16];
    int nupd;
    LaunchHook hook;     // may be null (timing of eager, not captured, launches)
    void* hook_ctx;
};
hipError_t launch_group_step(const GroupPlan& P, hipStream_t s);
hipError_t prepare_group_update(bool fp64, const GroupUpd& u);
// per-case workgroup counts of the grouped kernels (slope, fsum, fface,
// fnode, wall, cell_pre; time step and finalisation take one each)
void group_blocks(const Launch2D& L, int* slope, int* fsum, int* fface, int* fnode, int* wall,
                  int* cpre);
// node-update instance of a case as (wavefronts per cell, tile positions
// per wavefront, compiled-in nv or 0) and its dynamic LDS bytes
void group_update_key(const Launch2D& L, int* nw, int* maxp, int* nvc);
size_t group_update_lds(const Launch2D& L);
int group_fdiv_blocks(const Launch2D& L);   // workgroups of k_fdiv for a case

// 1D path
hipError_t launch_step1d(const Launch2D& L, hipStream_t s);

size_t update2d_lds_bytes(const Consts& c, bool fp64);
bool update2d_mfma_supported(const Consts& c);
// update mode 2 with nw wavefronts per cell: a compiled instance, nv <= 48,
// Nm <= 14 and its LDS within lds_limit bytes (b3: the bf16x3 contraction,
// FP32 storage only)
bool update2d_mx_supported(const Consts& c, bool fp64, bool b3, size_t lds_limit, int nw);
int update2d_mx_default_waves(const Consts& c);
// the default wavefronts per cell of update mode 2 or 3 for the case
int update2d_default_waves(const Launch2D& L, int mode);
// update mode 3 with nw wavefronts per cell: as mode 2 (not b3), the LDS of its
// face rows within the limit, the divergence arrays allocated and each
// distribution array under 4 GiB - 16 bytes (32-bit offsets)
bool update2d_mxf_supported(const Launch2D& L, int nw);
size_t update2d_mfma_lds_bytes(const Consts& c, bool fp64);

}  // namespace hugks
