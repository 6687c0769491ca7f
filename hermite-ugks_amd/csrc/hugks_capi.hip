// C-ABI of libhugks.so (declared in include/hugks.h): handle lifetime,
// device allocation, step sequencing and data exchange.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <tuple>
#include <mutex>
#include <map>

#include "hugks.h"
#include "block_writer.h"
#include "hugks_device.h"
#include "hugks_kernels.h"
#include <memory>

using namespace hugks;

namespace {

thread_local std::string g_create_error;

struct KernelTimer {
    const char* name;
    double ms;
    int calls;
};

// Per-kernel timing: every launch is bracketed by two HIP events recorded
// on the launch stream (no synchronisation); elapsed times are collected
// when the host asks for them or when many records are pending.
struct KTimes {
    hipStream_t stream = nullptr;
    std::vector<KernelTimer> timers;
    struct Rec {
        const char* name;
        hipEvent_t e0, e1;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    hipEvent_t open_ev = nullptr;
};

void kt_harvest(KTimes& t) {
    if (t.pending.empty()) return;
    (void)hipStreamSynchronize(t.stream);
    for (auto& r : t.pending) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r.e0, r.e1);
        bool found = false;
        for (auto& k : t.timers)
            if (k.name == r.name) {
                k.ms += ms;
                k.calls += 1;
                found = true;
                break;
            }
        if (!found) t.timers.push_back({r.name, (double)ms, 1});
        t.pool.push_back(r.e0);
        t.pool.push_back(r.e1);
    }
    t.pending.clear();
}

hipEvent_t kt_event(KTimes& t) {
    if (t.pool.empty()) {
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    hipEvent_t e = t.pool.back();
    t.pool.pop_back();
    return e;
}

void kt_hook(void* ctx, int phase, const char* name) {
    KTimes& t = *static_cast<KTimes*>(ctx);
    if (phase == 0) {
        if (t.pending.size() >= 4096) kt_harvest(t);
        t.open_ev = kt_event(t);
        (void)hipEventRecord(t.open_ev, t.stream);
    } else {
        hipEvent_t e1 = kt_event(t);
        (void)hipEventRecord(e1, t.stream);
        t.pending.push_back({name, t.open_ev, e1});
    }
}

void kt_release(KTimes& t) {
    for (auto& r : t.pending) {
        (void)hipEventDestroy(r.e0);
        (void)hipEventDestroy(r.e1);
    }
    t.pending.clear();
    for (hipEvent_t e : t.pool) (void)hipEventDestroy(e);
    t.pool.clear();
}

int kt_report(KTimes& t, int32_t cap, const char** names, double* ms, int32_t* calls,
              int32_t* count) {
    kt_harvest(t);
    int n = 0;
    for (auto& k : t.timers) {
        if (n < cap) {
            if (names) names[n] = k.name;
            if (ms) ms[n] = k.ms;
            if (calls) calls[n] = k.calls;
        }
        ++n;
    }
    if (count) *count = n;
    return HUGKS_OK;
}

}  // namespace

struct hugks_handle {
    hugks_desc d{};
    int dim = 2;
    bool fp64 = true;
    bool b3 = false;    // bf16x3 Hermite contractions (FP32 storage otherwise)
    Launch2D L{};
    hipStream_t stream = nullptr;
    int device = 0;
    std::vector<void*> allocs;
    std::string err;
    bool timing = false;
    KTimes kt;
    std::unique_ptr<BlockWriter> writer;   // macro-file output thread, made on first use
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    size_t ncell = 0;
    uint64_t run_hash = 0;   // of the inputs that define the run (checkpoints)
    bool in_batch = false;   // a batch's graphs hold this case's schedule
};

// FNV-1a, 64-bit, over byte ranges.
static uint64_t fnv1a(uint64_t hsh, const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
        hsh ^= c[i];
        hsh *= 1099511628211ull;
    }
    return hsh;
}

// Hash of what defines a run besides its shape: velocity nodes, weights,
// accelerations, Hermite order and the physical constants and boundary
// conditions.  Loop controls (cfl, maxtime, residual and steady-state
// stops) are left out: a restored run may continue under other ones.
static uint64_t run_hash_of(const hugks_desc* d, size_t ncell) {
    uint64_t r = 1469598103934665603ull;
    const int nn = d->dim == 2 ? d->nv * d->nv : d->nv;
    r = fnv1a(r, d->u, d->nv * sizeof(double));
    if (d->dim == 2) r = fnv1a(r, d->v, d->nv * sizeof(double));
    r = fnv1a(r, d->weight, nn * sizeof(double));
    if (d->dim == 2) {
        r = fnv1a(r, d->accel_cell, 2 * ncell * sizeof(double));
        r = fnv1a(r, d->accel_vface, 2 * (size_t)(d->nx + 1) * d->ny * sizeof(double));
        r = fnv1a(r, d->accel_hface, 2 * (size_t)d->nx * (d->ny + 1) * sizeof(double));
    }
    const int32_t iv[] = {d->nm, d->tau_model, d->dt_rule, d->bc[0], d->bc[1], d->bc[2],
                          d->bc[3], d->force_variant};
    const double dv[] = {d->dx, d->dy, d->pi, d->smv, d->innerk, d->gamma, d->prandtl,
                         d->shakhov_coef, d->muref, d->omega, d->tw[0], d->tw[1], d->tw[2],
                         d->tw[3], d->phi, d->umax_dt, d->vmax_dt, d->force_rate,
                         d->fd_span};
    r = fnv1a(r, iv, sizeof iv);
    return fnv1a(r, dv, sizeof dv);
}

static int fail(hugks_handle* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    else g_create_error = msg;
    return code;
}

static int hip_fail(hugks_handle* h, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(h, e == hipErrorOutOfMemory ? HUGKS_E_NOMEM : HUGKS_E_DEVICE, m);
}

template <typename P>
static hipError_t dalloc(hugks_handle* h, P** p, size_t bytes) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes < 16 ? 16 : bytes);
    if (e != hipSuccess) return e;
    h->allocs.push_back(q);
    *p = (P*)q;
    return hipSuccess;
}

static void release(hugks_handle* h) {
    if (!h) return;
    h->writer.reset();   // writes the files still queued
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void* p : h->allocs) (void)hipFree(p);
    h->allocs.clear();
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    kt_release(h->kt);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    h->stream = nullptr;
}

// One reference step of h on stream s (the handle's own stream, or a
// stream being captured into a batch graph: then without timing events).
static hipError_t enqueue_step_on(hugks_handle* h, hipStream_t s, bool timed) {
    hipError_t e;
    Launch2D& L = h->L;
    L.hook = timed && h->timing ? kt_hook : nullptr;
    L.hook_ctx = &h->kt;
    e = launch_timestep(L, h->dim, s);
    if (e != hipSuccess) return e;
    if (h->dim == 2) {
        e = launch_flux2d(L, s);
        if (e != hipSuccess) return e;
        e = launch_update2d(L, s);
        if (e != hipSuccess) return e;
    } else {
        e = launch_step1d(L, s);
        if (e != hipSuccess) return e;
    }
    return launch_finalize(L, h->dim, s);
}

static hipError_t enqueue_step(hugks_handle* h) { return enqueue_step_on(h, h->stream, true); }

// A case's step info and status from a host copy of its device state.
static int state_info(hugks_handle* h, const DevState& s, hugks_step_info* info) {
    if (info) {
        memset(info, 0, sizeof(*info));
        info->steps = s.steps;
        info->iter = s.iter;
        info->stop_reason = s.stop_reason;
        info->device_flags = s.flags;
        info->simtime = s.simtime;
        info->dt = s.dt;
        for (int q = 0; q < 4; ++q) info->res[q] = s.res[q];
        for (int q = 0; q < 3; ++q) info->steady_res[q] = s.steady[q];
        info->mass_rel_change = s.mass_rel;
    }
    // the wall flux runs before the cell update in a step (evolution before
    // update, 2d/src/main.f90:80-81), so a zero wall coefficient is the
    // first failure of its step; the non-finite state it leads to is not
    if (s.flags & 2) return fail(h, HUGKS_E_WALL, "zero outgoing wall mass coefficient");
    if (s.flags & 1) return fail(h, HUGKS_E_NUMERIC, "non-finite state or T <= 0 on the device");
    return HUGKS_OK;
}
static int read_state(hugks_handle* h, hugks_step_info* info) {
    DevState s;
    hipError_t e = hipMemcpyAsync(&s, h->L.st, sizeof(s), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "reading device state");
    return state_info(h, s, info);
}

extern "C" {

const char* hugks_last_error(const hugks_handle* h) {
    return h ? h->err.c_str() : g_create_error.c_str();
}

const char* hugks_device_info(void) {
    static std::string info;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        info = "no HIP device";
        return info.c_str();
    }
    hipDeviceProp_t p;
    (void)hipGetDeviceProperties(&p, 0);
    char buf[256];
    snprintf(buf, sizeof buf, "%s %s CUs=%d LDS/CU=%zu KiB HBM=%.1f GiB", p.name, p.gcnArchName,
             p.multiProcessorCount, (size_t)p.maxSharedMemoryPerMultiProcessor / 1024,
             p.totalGlobalMem / 1073741824.0);
    info = buf;
    return info.c_str();
}

int hugks_create(const hugks_desc* d, hugks_handle** out) {
    if (!d || !out) return fail(nullptr, HUGKS_E_ARG, "null descriptor or output pointer");
    *out = nullptr;
    if (d->abi_version != HUGKS_ABI_VERSION)
        return fail(nullptr, HUGKS_E_ARG, "descriptor abi_version mismatch");
    if (d->dim != 1 && d->dim != 2) return fail(nullptr, HUGKS_E_ARG, "dim must be 1 or 2");
    const int dim = d->dim;
    if (d->nx < 2 || (dim == 2 && d->ny < 2) || (dim == 1 && d->ny != 1))
        return fail(nullptr, HUGKS_E_ARG, "need nx >= 2 (and ny >= 2 in 2D, ny == 1 in 1D)");
    if (d->nv < 2) return fail(nullptr, HUGKS_E_ARG, "nv must be >= 2");
    if (d->nm < 0) return fail(nullptr, HUGKS_E_ARG, "Hermite order Nm must be non-negative");
    if (d->precision < HUGKS_FP64 || d->precision > HUGKS_MIXED_BF16X3)
        return fail(nullptr, HUGKS_E_ARG, "unknown precision");
    if (d->precision == HUGKS_MIXED_BF16X3 && d->dim != 2)
        return fail(nullptr, HUGKS_E_UNSUPPORTED,
                    "HUGKS_MIXED_BF16X3 applies to the 2D Hermite contraction only");
    if (!d->u || !d->weight || !d->prim0 || (dim == 2 && (!d->v || !d->accel_cell ||
                                                          !d->accel_vface || !d->accel_hface)))
        return fail(nullptr, HUGKS_E_ARG, "missing node, weight, prim0 or acceleration array");
    if (!(d->dx > 0) || (dim == 2 && !(d->dy > 0)))
        return fail(nullptr, HUGKS_E_ARG, "cell sizes must be positive");
    if (d->tau_model != HUGKS_TAU_MU_OVER_RHO && d->tau_model != HUGKS_TAU_VHS)
        return fail(nullptr, HUGKS_E_ARG, "unknown collision-time model");
    if (dim == 2 ? (d->dt_rule != HUGKS_DT_2D_GENERAL && d->dt_rule != HUGKS_DT_2D_WALL_NORMAL)
                 : d->dt_rule != HUGKS_DT_1D)
        return fail(nullptr, HUGKS_E_ARG, "time-step rule does not match the dimension");
    if (d->force_variant < HUGKS_FORCE_HERMITE_SYMMETRIC ||
        d->force_variant > HUGKS_FORCE_FINITE_DIFFERENCE)
        return fail(nullptr, HUGKS_E_ARG, "unknown force variant");
    if (d->force_variant == HUGKS_FORCE_FINITE_DIFFERENCE && (!(d->fd_span > 0) || d->nv < 3))
        return fail(nullptr, HUGKS_E_ARG,
                    "the finite-difference force needs fd_span > 0 and nv >= 3");
    if (d->force_variant == HUGKS_FORCE_FINITE_DIFFERENCE && d->precision == HUGKS_MIXED_BF16X3)
        return fail(nullptr, HUGKS_E_UNSUPPORTED,
                    "HUGKS_MIXED_BF16X3 names the Hermite contraction; the finite-difference "
                    "force has none");
    if (d->steady_steps < 0 || (d->steady_steps > 0 && (dim != 1 || !(d->steady_tol >= 0) ||
                                                         !(d->mass_tol >= 0))))
        return fail(nullptr, HUGKS_E_ARG, "steady test needs dim 1 and tolerances >= 0");
    if (!(d->cfl > 0) || !(d->prandtl > 0) || !(d->muref > 0) || !(d->gamma > 1) ||
        !(d->innerk >= 0) || !(d->pi > 0) || !(d->smv > 0))
        return fail(nullptr, HUGKS_E_ARG,
                    "cfl, prandtl, muref, pi, smv must be positive, gamma > 1, innerk >= 0");
    for (int s = 0; s < (dim == 2 ? 4 : 2); ++s) {
        if (d->bc[s] < HUGKS_BC_PERIODIC || d->bc[s] > HUGKS_BC_DIFFUSE)
            return fail(nullptr, HUGKS_E_ARG, "unknown boundary kind");
        if (d->bc[s] == HUGKS_BC_DIFFUSE && !(d->tw[s] > 0))
            return fail(nullptr, HUGKS_E_ARG, "diffuse wall temperature must be positive");
    }
    if (dim == 2 && ((d->bc[0] == HUGKS_BC_PERIODIC) != (d->bc[1] == HUGKS_BC_PERIODIC) ||
                     (d->bc[2] == HUGKS_BC_PERIODIC) != (d->bc[3] == HUGKS_BC_PERIODIC)))
        return fail(nullptr, HUGKS_E_ARG, "periodic sides must come in pairs");
    if (dim == 1 && (d->bc[0] == HUGKS_BC_PERIODIC || d->bc[1] == HUGKS_BC_PERIODIC))
        return fail(nullptr, HUGKS_E_UNSUPPORTED, "1D supports specular and diffuse walls");
    const int nn = dim == 2 ? d->nv * d->nv : d->nv;
    const int ldv = (nn + 63) / 64 * 64;
    if (dim == 2 && ldv > 2048)
        return fail(nullptr, HUGKS_E_UNSUPPORTED, "nv*nv above 2048 has no 2D kernel");
    if (dim == 1 && ldv > 256) return fail(nullptr, HUGKS_E_UNSUPPORTED, "1D supports nv <= 256");
    if (dim == 2 && d->nm + 2 > 2 * d->nv)
        return fail(nullptr, HUGKS_E_UNSUPPORTED, "Hermite order above 2*nv-2 is not supported");
    for (int k = 0; k < nn; ++k)
        if (!(d->weight[k] >= 0) || !std::isfinite(d->weight[k]))
            return fail(nullptr, HUGKS_E_ARG, "weights must be finite and non-negative");
    const size_t ncell = (size_t)d->nx * d->ny;
    for (size_t c = 0; c < ncell; ++c) {
        const double* p = d->prim0 + (size_t)(dim + 2) * c;
        if (!(p[0] > 0) || !(p[dim + 1] > 0))
            return fail(nullptr, HUGKS_E_ARG, "initial density and lambda must be positive");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, HUGKS_E_NOHIP, "no HIP device visible");

    hugks_handle* h = new hugks_handle();
    h->run_hash = run_hash_of(d, ncell);
    h->d = *d;
    h->d.u = h->d.v = h->d.weight = h->d.prim0 = nullptr;
    h->d.accel_cell = h->d.accel_vface = h->d.accel_hface = nullptr;
    h->dim = dim;
    h->fp64 = d->precision == HUGKS_FP64;
    h->b3 = d->precision == HUGKS_MIXED_BF16X3;
    h->ncell = ncell;
    (void)hipGetDevice(&h->device);
    Consts& c = h->L.c;
    c.nx = d->nx; c.ny = d->ny; c.nv = d->nv; c.nn = nn; c.ldv = ldv; c.nm = d->nm;
    for (int s = 0; s < 4; ++s) { c.bc[s] = d->bc[s]; c.tw[s] = d->tw[s]; }
    c.tau_model = d->tau_model; c.dt_rule = d->dt_rule; c.clamp_dt = d->clamp_dt;
    c.stop_on_residual = d->stop_on_residual;
    c.dx = d->dx; c.dy = dim == 2 ? d->dy : 1.0; c.pi = d->pi; c.smv = d->smv; c.K = d->innerk;
    c.gamma = d->gamma; c.prandtl = d->prandtl; c.shk = d->shakhov_coef; c.muref = d->muref;
    c.omega = d->omega; c.cfl = d->cfl; c.maxtime = d->maxtime; c.eps_res = d->eps_res;
    c.umax_dt = d->umax_dt; c.vmax_dt = d->vmax_dt; c.force_rate = d->force_rate;
    c.tref = d->tref; c.phi = d->phi;
    c.force_fd = d->force_variant == HUGKS_FORCE_FINITE_DIFFERENCE;
    c.fd_du = d->fd_span / d->nv;
    c.steady_steps = dim == 1 ? d->steady_steps : 0;
    c.steady_tol = d->steady_tol; c.mass_tol = d->mass_tol;
    Launch2D& L = h->L;
    L.fp64 = h->fp64;
    L.b3 = h->b3;
    const size_t es = h->fp64 ? sizeof(double) : sizeof(float);
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { release(h); delete h; return hip_fail(nullptr, e, "stream"); }
    h->kt.stream = h->stream;
    // every runtime call of the set-up is checked: a failed copy of the
    // node tables, the initial state or the accelerations must not leave a
    // handle with unset inputs behind HUGKS_OK
#define HCHK(call)                                                                 \
    do {                                                                           \
        e = (call);                                                                \
        if (e != hipSuccess) {                                                     \
            int rc = hip_fail(nullptr, e, "hugks_create: " #call);                 \
            release(h);                                                            \
            delete h;                                                              \
            return rc;                                                             \
        }                                                                          \
    } while (0)
    HCHK(hipEventCreate(&h->ev0));
    HCHK(hipEventCreate(&h->ev1));
    const size_t fsz = ncell * ldv * es;
    const size_t nvf = dim == 2 ? (size_t)(d->nx + 1) * d->ny : (size_t)d->nx + 1;
    const size_t nhf = dim == 2 ? (size_t)d->nx * (d->ny + 1) : 1;
    const int np = dim == 2 ? 5 : 9;
    double* prim0_d = nullptr;
#define ALLOC(ptr, bytes)                                                          \
    do {                                                                           \
        e = dalloc(h, &(ptr), (bytes));                                            \
        if (e != hipSuccess) {                                                     \
            int rc = hip_fail(nullptr, e, "device allocation");                    \
            release(h);                                                            \
            delete h;                                                              \
            return rc;                                                             \
        }                                                                          \
    } while (0)
    ALLOC(L.st, sizeof(DevState));
    ALLOC(L.W, ncell * (dim + 2) * sizeof(double));
    ALLOC(L.part, ncell * np * sizeof(double));
    ALLOC(L.red, RED_SCRATCH * sizeof(double));
    ALLOC(L.h, fsz);
    ALLOC(L.b, fsz);
    ALLOC(L.shx, fsz);
    ALLOC(L.sbx, fsz);
    if (dim == 2) {
        ALLOC(L.shy, fsz);
        ALLOC(L.sby, fsz);
        ALLOC(L.hfh, nhf * ldv * es);
        ALLOC(L.hfb, nhf * ldv * es);
        ALLOC(L.hflux, nhf * 4 * sizeof(double));
    }
    ALLOC(L.vfh, nvf * ldv * es);
    ALLOC(L.vfb, nvf * ldv * es);
    ALLOC(L.vflux, nvf * 4 * sizeof(double));
    double *un_d, *vn_d, *wn_d, *u1, *v1;
    ALLOC(un_d, ldv * sizeof(double));
    ALLOC(vn_d, ldv * sizeof(double));
    ALLOC(wn_d, ldv * sizeof(double));
    double* n4_d;
    ALLOC(n4_d, ldv * 4 * sizeof(double));
    ALLOC(u1, d->nv * sizeof(double));
    ALLOC(v1, d->nv * sizeof(double));
    ALLOC(prim0_d, ncell * (dim + 2) * sizeof(double));
    // per-node tables: u/v of node k and its weight, zero in the pad
    std::vector<double> un(ldv, 0.0), vn(ldv, 0.0), wn(ldv, 0.0);
    for (int k = 0; k < nn; ++k) {
        if (dim == 2) {
            un[k] = d->u[k % d->nv];
            vn[k] = d->v[k / d->nv];
        } else {
            un[k] = d->u[k];
        }
        wn[k] = d->weight[k];
    }
    std::vector<double> vtab(d->nv, 0.0);
    if (dim == 2) for (int k = 0; k < d->nv; ++k) vtab[k] = d->v[k];
    HCHK(hipMemcpy(un_d, un.data(), ldv * sizeof(double), hipMemcpyHostToDevice));
    HCHK(hipMemcpy(vn_d, vn.data(), ldv * sizeof(double), hipMemcpyHostToDevice));
    HCHK(hipMemcpy(wn_d, wn.data(), ldv * sizeof(double), hipMemcpyHostToDevice));
    {
        std::vector<double> n4(4 * (size_t)ldv, 0.0);
        for (int k = 0; k < ldv; ++k) {
            n4[4 * (size_t)k] = un[k];
            n4[4 * (size_t)k + 1] = vn[k];
            n4[4 * (size_t)k + 2] = wn[k];
        }
        HCHK(hipMemcpy(n4_d, n4.data(), 4 * (size_t)ldv * sizeof(double), hipMemcpyHostToDevice));
    }
    HCHK(hipMemcpy(u1, d->u, d->nv * sizeof(double), hipMemcpyHostToDevice));
    HCHK(hipMemcpy(v1, vtab.data(), d->nv * sizeof(double), hipMemcpyHostToDevice));
    HCHK(hipMemcpy(prim0_d, d->prim0, ncell * (dim + 2) * sizeof(double), hipMemcpyHostToDevice));
    L.unode_d = un_d; L.vnode_d = vn_d; L.wnode_d = wn_d; L.u1 = u1; L.v1 = v1; L.node4 = n4_d;
    {
        void* uv;
        const size_t es_a = h->fp64 ? sizeof(double) : sizeof(float);
        ALLOC(uv, 2 * (size_t)ldv * es_a);
        if (h->fp64) {
            std::vector<double> t(2 * (size_t)ldv);
            for (int k = 0; k < ldv; ++k) { t[2 * k] = un[k]; t[2 * k + 1] = vn[k]; }
            HCHK(hipMemcpy(uv, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
        } else {
            std::vector<float> t(2 * (size_t)ldv);
            for (int k = 0; k < ldv; ++k) { t[2 * k] = (float)un[k]; t[2 * k + 1] = (float)vn[k]; }
            HCHK(hipMemcpy(uv, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
        }
        L.uvnode = uv;
    }
    if (h->fp64) {
        L.unode = un_d; L.vnode = vn_d; L.wnode = wn_d;
    } else {
        float *uf, *vf, *wf;
        ALLOC(uf, ldv * sizeof(float));
        ALLOC(vf, ldv * sizeof(float));
        ALLOC(wf, ldv * sizeof(float));
        std::vector<float> a(ldv), bb(ldv), cc(ldv);
        for (int k = 0; k < ldv; ++k) { a[k] = (float)un[k]; bb[k] = (float)vn[k]; cc[k] = (float)wn[k]; }
        HCHK(hipMemcpy(uf, a.data(), ldv * sizeof(float), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(vf, bb.data(), ldv * sizeof(float), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(wf, cc.data(), ldv * sizeof(float), hipMemcpyHostToDevice));
        L.unode = uf; L.vnode = vf; L.wnode = wf;
    }
    if (dim == 2) {
        double *ac, *av, *ah;
        ALLOC(ac, ncell * 2 * sizeof(double));
        ALLOC(av, nvf * 2 * sizeof(double));
        ALLOC(ah, nhf * 2 * sizeof(double));
        HCHK(hipMemcpy(ac, d->accel_cell, ncell * 2 * sizeof(double), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(av, d->accel_vface, nvf * 2 * sizeof(double), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(ah, d->accel_hface, nhf * 2 * sizeof(double), hipMemcpyHostToDevice));
        L.acc_cell = ac; L.acc_v = av; L.acc_h = ah;
        unsigned nwall = 0;
        for (int s = 0; s < 4; ++s)
            if (d->bc[s] != HUGKS_BC_PERIODIC) nwall += s < 2 ? d->ny : d->nx;
        L.nwall = nwall;
        // interior-face pipeline scratch (per-face sums and coefficients)
        const bool xper = d->bc[0] == HUGKS_BC_PERIODIC, yper = d->bc[2] == HUGKS_BC_PERIODIC;
        const size_t nint = (size_t)(xper ? d->nx + 1 : d->nx - 1) * d->ny +
                            (size_t)d->nx * (yper ? d->ny + 1 : d->ny - 1);
        L.stencil_ok = fsz < ((size_t)1 << 32);
        // default: the cell-sum pipeline where it applies (set below), else
        // the split one
        L.flux_mode = HUGKS_FLUX_CELLSUM;
        bool maxw_nodes = false;
        if (const char* fm = getenv("HUGKS_FLUX_KERNEL")) {
            if (strcmp(fm, "fused") == 0) L.flux_mode = HUGKS_FLUX_FUSED;
            if (strcmp(fm, "split") == 0) L.flux_mode = HUGKS_FLUX_SPLIT;
            if (strcmp(fm, "stencil") == 0 && L.stencil_ok) L.flux_mode = HUGKS_FLUX_STENCIL;
            if (strcmp(fm, "cellsum") == 0) L.flux_mode = HUGKS_FLUX_CELLSUM;
            if (strcmp(fm, "split-nodes") == 0) {
                L.flux_mode = HUGKS_FLUX_SPLIT;
                maxw_nodes = true;
            }
        }
        // Factors of a tensor-product quadrature, w[i + nv j] = wx[i] wy[j]
        // (both quadratures of velocity.f90:43,128 are such products); the
        // Maxwellian sums of the flux then reduce to 1D moments.  Taken
        // from the row and column of the largest weight and accepted when
        // they reproduce every weight to 1e-13.
        {
            const int nv = d->nv;
            int km = 0;
            for (int k = 1; k < nn; ++k)
                if (d->weight[k] > d->weight[km]) km = k;
            const int im = km % nv, jm = km / nv;
            const double wmax = d->weight[km];
            std::vector<double> fx(nv), fy(nv);
            bool sep = wmax > 0 && !maxw_nodes;
            for (int i = 0; i < nv; ++i) fx[i] = d->weight[i + nv * jm];
            for (int j = 0; j < nv; ++j) fy[j] = d->weight[im + nv * j] / wmax;
            for (int k = 0; sep && k < nn; ++k) {
                const double w = d->weight[k], pr = fx[k % nv] * fy[k / nv];
                if (!(fabs(pr - w) <= 1e-13 * w)) sep = false;
            }
            if (sep) {
                double *wxd, *wyd;
                ALLOC(wxd, nv * sizeof(double));
                ALLOC(wyd, nv * sizeof(double));
                HCHK(hipMemcpy(wxd, fx.data(), nv * sizeof(double), hipMemcpyHostToDevice));
                HCHK(hipMemcpy(wyd, fy.data(), nv * sizeof(double), hipMemcpyHostToDevice));
                L.wx = wxd;
                L.wy = wyd;
            }
        }
        ALLOC(L.fsum, (nint ? nint : 1) * FACE_SUM_LD * sizeof(double));
        // per-cell half-space sums (HUGKS_FLUX_CELLSUM): one lane per u
        // column (nv <= 64), the v nodes split as rows [0, js) negative,
        // [js, nv) non-negative (upwind_delta's sign test), and weights
        // that factor (w = wx wy)
        L.cs_js = -1;
        if (d->nv <= 64 && L.wx) {
            int js = 0;
            while (js < d->nv && std::signbit(d->v[js])) ++js;
            bool split = true;
            for (int j = js; j < d->nv; ++j) split = split && !std::signbit(d->v[j]);
            if (split) {
                L.cs_js = js;
                ALLOC(L.csum, ncell * 4 * CELL_SUM_LD * sizeof(double));
            }
        }
        if (L.flux_mode == HUGKS_FLUX_CELLSUM && L.cs_js < 0) L.flux_mode = HUGKS_FLUX_SPLIT;
        // the per-cell face divergence of HUGKS_UPDATE_MFMA_FACE (k_fdiv)
        if (!h->b3 && update2d_mfma_supported(c) && fsz < (size_t)0xFFFFFFF0u) {
            ALLOC(L.dvh, fsz);
            ALLOC(L.dvb, fsz);
        }
        ALLOC(L.fcoef, (nint ? nint : 1) * FACE_COEF_LD * sizeof(double));
        ALLOC(L.fmaxw, (nint ? nint : 1) * FACE_MAXW_LD * sizeof(double));
        ALLOC(L.cellpre, ncell * 24 * sizeof(double));
        L.update_lds = update2d_lds_bytes(c, h->fp64);
        hipDeviceProp_t prop;
        HCHK(hipGetDeviceProperties(&prop, h->device));
        L.lds_limit = prop.sharedMemPerBlockOptin > 0 ? (size_t)prop.sharedMemPerBlockOptin
                                                      : (size_t)prop.sharedMemPerBlock;
        L.update_nw = update2d_mx_default_waves(c);
        const bool mx = update2d_mx_supported(c, h->fp64, h->b3, L.lds_limit, L.update_nw);
        if (h->b3 && !mx) {
            release(h);
            delete h;
            return fail(nullptr, HUGKS_E_UNSUPPORTED,
                        "HUGKS_MIXED_BF16X3 needs the MFMA node update (nv <= 48, Nm <= 14)");
        }
        // default: the node update forming the interior faces' equilibrium
        // flux where it applies, else the MFMA node update, else the MFMA
        // Hermite tiles, else the LDS kernel
        const bool mxf =
            (L.flux_mode == HUGKS_FLUX_CELLSUM || L.flux_mode == HUGKS_FLUX_SPLIT) &&
            update2d_mxf_supported(L, L.update_nw);
        L.update_mode = mxf ? 3 : mx ? 2 : update2d_mfma_supported(c) ? 1 : 0;
        if (const char* um = getenv("HUGKS_UPDATE_KERNEL"); um && !h->b3) {
            if (strcmp(um, "lds") == 0) L.update_mode = 0;
            if (strcmp(um, "mfma") == 0 && update2d_mfma_supported(c)) L.update_mode = 1;
            if (strcmp(um, "mfma-node") == 0 && mx) L.update_mode = 2;
            if (strcmp(um, "mfma-face") == 0 && mxf) L.update_mode = 3;
        }
        if (L.update_mode >= 2) L.update_nw = update2d_default_waves(L, L.update_mode);
        if (L.update_lds > (size_t)prop.sharedMemPerBlockOptin && prop.sharedMemPerBlockOptin > 0) {
            release(h);
            delete h;
            return fail(nullptr, HUGKS_E_UNSUPPORTED, "Hermite working set exceeds LDS");
        }
        e = prepare_update2d(L);
        if (e != hipSuccess) {
            int rc = hip_fail(nullptr, e, "update kernel attributes");
            release(h);
            delete h;
            return rc;
        }
    }
#undef ALLOC
    // zero everything that pads or is read before written, then the
    // initial state
    HCHK(hipMemsetAsync(L.shx, 0, fsz, h->stream));
    HCHK(hipMemsetAsync(L.sbx, 0, fsz, h->stream));
    if (dim == 2) {
        HCHK(hipMemsetAsync(L.shy, 0, fsz, h->stream));
        HCHK(hipMemsetAsync(L.sby, 0, fsz, h->stream));
        HCHK(hipMemsetAsync(L.hfh, 0, nhf * ldv * es, h->stream));
        HCHK(hipMemsetAsync(L.hfb, 0, nhf * ldv * es, h->stream));
        HCHK(hipMemsetAsync(L.hflux, 0, nhf * 4 * sizeof(double), h->stream));
    }
    HCHK(hipMemsetAsync(L.vfh, 0, nvf * ldv * es, h->stream));
    HCHK(hipMemsetAsync(L.vfb, 0, nvf * ldv * es, h->stream));
    HCHK(hipMemsetAsync(L.vflux, 0, nvf * 4 * sizeof(double), h->stream));
    DevState s0{};
    s0.iter = 1;
    s0.t_stop = INFINITY;
    // initial mass: sum of rho * cell size in cell order (2d/src/main.f90:63-68)
    for (size_t q = 0; q < ncell; ++q) s0.mass0 += d->prim0[(size_t)(dim + 2) * q] * (c.dx * c.dy);
    HCHK(hipMemcpyAsync(L.st, &s0, sizeof(s0), hipMemcpyHostToDevice, h->stream));
    e = launch_init_f(L, dim, prim0_d, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        int rc = hip_fail(nullptr, e, "initial state");
        release(h);
        delete h;
        return rc;
    }
#undef HCHK
    *out = h;
    return HUGKS_OK;
}

int hugks_destroy(hugks_handle* h) {
    if (!h) return HUGKS_OK;
    release(h);
    delete h;
    return HUGKS_OK;
}

// ----------------------------------------------------------- output

int hugks_write_blocks_async(hugks_handle* h, const char* path, const char* header,
                             int32_t per_line, int32_t nblock, const int64_t* block_len,
                             const double* values) {
    if (!path || !header || per_line < 1 || nblock < 0 || (nblock > 0 && !block_len))
        return fail(h, HUGKS_E_ARG, "write_blocks: null path/header/lengths or bad counts");
    BlockFile f;
    f.path = path;
    f.header = header;
    f.per_line = per_line;
    int64_t total = 0;
    for (int32_t k = 0; k < nblock; ++k) {
        if (block_len[k] < 0) return fail(h, HUGKS_E_ARG, "write_blocks: negative block length");
        total += block_len[k];
    }
    if (total > 0 && !values) return fail(h, HUGKS_E_ARG, "write_blocks: null values");
    f.block_len.assign(block_len, block_len + nblock);
    f.values.assign(values, values + total);
    if (!h) {
        const std::string e = write_block_file(f);
        return e.empty() ? HUGKS_OK : fail(nullptr, HUGKS_E_IO, e);
    }
    if (!h->writer) h->writer = std::make_unique<BlockWriter>();
    h->writer->push(std::move(f));
    return HUGKS_OK;
}

int hugks_write_wait(hugks_handle* h) {
    if (!h) return fail(nullptr, HUGKS_E_ARG, "null handle");
    if (!h->writer) return HUGKS_OK;
    const std::string e = h->writer->wait();
    return e.empty() ? HUGKS_OK : fail(h, HUGKS_E_IO, e);
}

int hugks_step(hugks_handle* h, hugks_step_info* info) {
    if (!h) return fail(nullptr, HUGKS_E_ARG, "null handle");
    hipError_t e = launch_begin(h->L.st, INFINITY, h->stream);
    if (e == hipSuccess) e = enqueue_step(h);
    if (e != hipSuccess) return hip_fail(h, e, "step launch");
    return read_state(h, info);
}

int hugks_advance(hugks_handle* h, int32_t max_steps, hugks_step_info* info) {
    return hugks_advance_until(h, max_steps, INFINITY, info);
}

int hugks_advance_until(hugks_handle* h, int32_t max_steps, double t_stop,
                        hugks_step_info* info) {
    if (!h) return fail(nullptr, HUGKS_E_ARG, "null handle");
    if (std::isnan(t_stop)) return fail(h, HUGKS_E_ARG, "t_stop is NaN");
    hipError_t e = launch_begin(h->L.st, t_stop, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "advance launch");
    if (max_steps > 0) {
        for (int32_t s = 0; s < max_steps; ++s) {
            e = enqueue_step(h);
            if (e != hipSuccess) return hip_fail(h, e, "advance launch");
        }
        return read_state(h, info);
    }
    // no cap: run in chunks until an exit condition holds
    const int chunk = 64;
    for (;;) {
        for (int s = 0; s < chunk; ++s) {
            e = enqueue_step(h);
            if (e != hipSuccess) return hip_fail(h, e, "advance launch");
        }
        DevState s;
        e = hipMemcpyAsync(&s, h->L.st, sizeof(s), hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "advance state");
        if (s.stopped || s.flags) break;
    }
    return read_state(h, info);
}

int hugks_get_macro(hugks_handle* h, double* w, double* q) {
    if (!h) return fail(nullptr, HUGKS_E_ARG, "null handle");
    const size_t ncell = h->ncell;
    hipError_t e = hipSuccess;
    if (w) e = hipMemcpyAsync(w, h->L.W, ncell * (h->dim + 2) * sizeof(double),
                              hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && q) {
        double* qd = nullptr;
        e = hipMallocAsync((void**)&qd, ncell * h->dim * sizeof(double), h->stream);
        if (e == hipSuccess) e = launch_macro(h->L, h->dim, qd, h->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(q, qd, ncell * h->dim * sizeof(double), hipMemcpyDeviceToHost,
                               h->stream);
        if (qd) (void)hipFreeAsync(qd, h->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "get_macro");
    return HUGKS_OK;
}

}  // extern "C"

// Staging of hugks_get_macro_noalloc: a file-scope device array instead of
// a per-call allocation, so that callers that cannot allocate device
// memory (code under graph capture, tight serving loops) can read the
// output fields.  Every GPU holds one copy, at its own address, for the
// life of the process; host code uses the per-device address from
// hipGetSymbolAddress, never the array's name.  Every k_macro work-group
// writes all dim entries of its cell, so nothing has to start at zero.
// Calls that overlap on one GPU share the array and race.
namespace {
constexpr size_t kMacroScratchDoubles = HUGKS_MACRO_NOALLOC_MAX;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace
static __device__ __align__(256) double scratch[kMacroScratchDoubles];

// The device address of scratch on device dev, looked up once per device
// with that device current.
static hipError_t scratch_on(int dev, double** out) {
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) {
            const hipError_t e2 = hipSetDevice(cur);
            if (e == hipSuccess) e = e2;
        }
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

extern "C" {

int hugks_get_macro_noalloc(hugks_handle* h, double* w, double* q) {
    if (!h) return fail(nullptr, HUGKS_E_ARG, "null handle");
    const size_t ncell = h->ncell;
    if (q && (size_t)h->dim * ncell > kMacroScratchDoubles)
        return fail(h, HUGKS_E_UNSUPPORTED,
                    "hugks_get_macro_noalloc: dim*ncell = " + std::to_string(h->dim * ncell) +
                        " heat-flux values exceed its static staging of " +
                        std::to_string(kMacroScratchDoubles) + "; use hugks_get_macro");
    hipError_t e = hipSuccess;
    if (w) e = hipMemcpyAsync(w, h->L.W, ncell * (h->dim + 2) * sizeof(double),
                              hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && q) {
        double* qd = nullptr;
        e = scratch_on(h->device, &qd);
        if (e == hipSuccess) e = launch_macro(h->L, h->dim, qd, h->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(q, qd, ncell * h->dim * sizeof(double), hipMemcpyDeviceToHost,
                               h->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "get_macro_noalloc");
    return HUGKS_OK;
}

int hugks_get_f(hugks_handle* h, double* hf, double* bf) {
    if (!h) return fail(nullptr, HUGKS_E_ARG, "null handle");
    const Consts& c = h->L.c;
    const size_t n = h->ncell * c.nn;
    double* tmp = nullptr;
    hipError_t e = hipMallocAsync((void**)&tmp, n * sizeof(double), h->stream);
    for (int q = 0; q < 2 && e == hipSuccess; ++q) {
        double* dst = q ? bf : hf;
        if (!dst) continue;
        e = launch_convert(q ? h->L.b : h->L.h, h->fp64, tmp, h->ncell, c.nn, c.ldv, true, h->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dst, tmp, n * sizeof(double), hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    }
    if (tmp) (void)hipFreeAsync(tmp, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "get_f");
    return HUGKS_OK;
}

int hugks_set_f(hugks_handle* h, const double* hf, const double* bf, const double* w) {
    if (!h) return fail(nullptr, HUGKS_E_ARG, "null handle");
    const Consts& c = h->L.c;
    const size_t n = h->ncell * c.nn;
    double* tmp = nullptr;
    hipError_t e = hipMallocAsync((void**)&tmp, n * sizeof(double), h->stream);
    for (int q = 0; q < 2 && e == hipSuccess; ++q) {
        const double* src = q ? bf : hf;
        if (!src) continue;
        e = hipMemcpyAsync(tmp, src, n * sizeof(double), hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess)
            e = launch_convert_in(tmp, q ? h->L.b : h->L.h, h->fp64, h->ncell, c.nn, c.ldv, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    }
    if (e == hipSuccess && w) {
        e = hipMemcpyAsync(h->L.W, w, h->ncell * (h->dim + 2) * sizeof(double),
                           hipMemcpyHostToDevice, h->stream);
        // the new W is the reference of the relative mass change
        double m0 = 0.0;
        for (size_t q = 0; q < h->ncell; ++q) m0 += w[(size_t)(h->dim + 2) * q] * (c.dx * c.dy);
        if (e == hipSuccess)
            e = hipMemcpyAsync(&h->L.st->mass0, &m0, sizeof(double), hipMemcpyHostToDevice,
                               h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    }
    // a new state: a device flag raised by an earlier step no longer holds
    if (e == hipSuccess) e = hipMemsetAsync(&h->L.st->flags, 0, sizeof(int), h->stream);
    if (tmp) (void)hipFreeAsync(tmp, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "set_f");
    return HUGKS_OK;
}

int hugks_debug_fetch(hugks_handle* h, int32_t which, double* out) {
    if (!h || !out) return fail(h, HUGKS_E_ARG, "null handle or output");
    const Consts& c = h->L.c;
    const bool two = h->dim == 2;
    const size_t nvf = two ? (size_t)(c.nx + 1) * c.ny : (size_t)c.nx + 1;
    const size_t nhf = two ? (size_t)c.nx * (c.ny + 1) : 0;
    const void* src = nullptr;
    size_t items = 0;
    switch (which) {
        case HUGKS_ARR_SHX: src = h->L.shx; items = h->ncell; break;
        case HUGKS_ARR_SHY: src = two ? h->L.shy : nullptr; items = h->ncell; break;
        case HUGKS_ARR_SBX: src = h->L.sbx; items = h->ncell; break;
        case HUGKS_ARR_SBY: src = two ? h->L.sby : nullptr; items = h->ncell; break;
        case HUGKS_ARR_VFH: src = h->L.vfh; items = nvf; break;
        case HUGKS_ARR_VFB: src = h->L.vfb; items = nvf; break;
        case HUGKS_ARR_HFH: src = two ? h->L.hfh : nullptr; items = nhf; break;
        case HUGKS_ARR_HFB: src = two ? h->L.hfb : nullptr; items = nhf; break;
        case HUGKS_ARR_VFLUX: {
            hipError_t e = hipMemcpyAsync(out, h->L.vflux, nvf * 4 * sizeof(double),
                                          hipMemcpyDeviceToHost, h->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
            return e == hipSuccess ? HUGKS_OK : hip_fail(h, e, "debug_fetch");
        }
        case HUGKS_ARR_HFLUX: {
            if (!two) return fail(h, HUGKS_E_ARG, "no horizontal faces in 1D");
            hipError_t e = hipMemcpyAsync(out, h->L.hflux, nhf * 4 * sizeof(double),
                                          hipMemcpyDeviceToHost, h->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
            return e == hipSuccess ? HUGKS_OK : hip_fail(h, e, "debug_fetch");
        }
        default: return fail(h, HUGKS_E_ARG, "unknown array id");
    }
    if (!src) return fail(h, HUGKS_E_ARG, "array does not exist for this dimension");
    const size_t n = items * c.nn;
    double* tmp = nullptr;
    hipError_t e = hipMallocAsync((void**)&tmp, n * sizeof(double), h->stream);
    if (e == hipSuccess) e = launch_convert(src, h->fp64, tmp, items, c.nn, c.ldv, true, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, n * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (tmp) (void)hipFreeAsync(tmp, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return e == hipSuccess ? HUGKS_OK : hip_fail(h, e, "debug_fetch");
}

int hugks_get_time(hugks_handle* h, double* simtime, int32_t* iter) {
    if (!h) return fail(nullptr, HUGKS_E_ARG, "null handle");
    DevState s;
    hipError_t e = hipMemcpyAsync(&s, h->L.st, sizeof(s), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "get_time");
    if (simtime) *simtime = s.simtime;
    if (iter) *iter = s.iter;
    return HUGKS_OK;
}

// ----------------------------------------------------------- checkpoint
namespace {
constexpr char kCkMagic[8] = {'H', 'U', 'G', 'K', 'S', 'C', 'K', '2'};
struct CkHeader {
    char magic[8];
    int32_t abi, dim, nx, ny, nv, nm, precision, ldv;
    int32_t iter, steady_count, flags, pad;
    double dt, simtime, mass0, mass_rel, res[4], steady[3];
    uint64_t run_hash;
};
struct FileCloser {
    FILE* f;
    ~FileCloser() {
        if (f) fclose(f);
    }
};
}  // namespace

int hugks_checkpoint_write(hugks_handle* h, const char* path) {
    if (!h || !path) return fail(h, HUGKS_E_ARG, "null handle or path");
    const Consts& c = h->L.c;
    const size_t es = h->fp64 ? sizeof(double) : sizeof(float);
    const size_t fbytes = h->ncell * c.ldv * es, wbytes = h->ncell * (h->dim + 2) * sizeof(double);
    DevState st;
    hipError_t e = hipMemcpyAsync(&st, h->L.st, sizeof(st), hipMemcpyDeviceToHost, h->stream);
    std::vector<char> buf(wbytes > fbytes ? wbytes : fbytes);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "checkpoint_write");
    CkHeader hd;
    memset(&hd, 0, sizeof(hd));
    memcpy(hd.magic, kCkMagic, 8);
    hd.abi = HUGKS_ABI_VERSION;
    hd.dim = h->dim; hd.nx = c.nx; hd.ny = c.ny; hd.nv = c.nv; hd.nm = c.nm;
    hd.precision = h->d.precision; hd.ldv = c.ldv;
    hd.run_hash = h->run_hash;
    hd.iter = st.iter; hd.steady_count = st.steady_count; hd.flags = st.flags;
    hd.dt = st.dt; hd.simtime = st.simtime; hd.mass0 = st.mass0; hd.mass_rel = st.mass_rel;
    for (int q = 0; q < 4; ++q) hd.res[q] = st.res[q];
    for (int q = 0; q < 3; ++q) hd.steady[q] = st.steady[q];
    FileCloser fc{fopen(path, "wb")};
    if (!fc.f) return fail(h, HUGKS_E_ARG, std::string("cannot open for writing: ") + path);
    bool ok = fwrite(&hd, sizeof(hd), 1, fc.f) == 1;
    const void* srcs[3] = {h->L.W, h->L.h, h->L.b};
    const size_t sizes[3] = {wbytes, fbytes, fbytes};
    for (int a = 0; a < 3 && ok; ++a) {
        e = hipMemcpy(buf.data(), srcs[a], sizes[a], hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(h, e, "checkpoint_write");
        ok = fwrite(buf.data(), 1, sizes[a], fc.f) == sizes[a];
    }
    if (!ok || fflush(fc.f) != 0) return fail(h, HUGKS_E_ARG, std::string("write failed: ") + path);
    return HUGKS_OK;
}

int hugks_checkpoint_read(hugks_handle* h, const char* path) {
    if (!h || !path) return fail(h, HUGKS_E_ARG, "null handle or path");
    const Consts& c = h->L.c;
    const size_t es = h->fp64 ? sizeof(double) : sizeof(float);
    const size_t fbytes = h->ncell * c.ldv * es, wbytes = h->ncell * (h->dim + 2) * sizeof(double);
    FileCloser fc{fopen(path, "rb")};
    if (!fc.f) return fail(h, HUGKS_E_ARG, std::string("cannot open: ") + path);
    CkHeader hd;
    if (fread(&hd, sizeof(hd), 1, fc.f) != 1 || memcmp(hd.magic, kCkMagic, 8) != 0)
        return fail(h, HUGKS_E_ARG, std::string("not a checkpoint file: ") + path);
    if (hd.abi != HUGKS_ABI_VERSION || hd.dim != h->dim || hd.nx != c.nx || hd.ny != c.ny ||
        hd.nv != c.nv || hd.nm != c.nm || hd.precision != h->d.precision || hd.ldv != c.ldv)
        return fail(h, HUGKS_E_ARG,
                    "checkpoint shape, Hermite order or precision differs from the handle");
    if (hd.run_hash != h->run_hash)
        return fail(h, HUGKS_E_ARG,
                    "checkpoint was written by a run with other velocity nodes, weights, "
                    "accelerations or physical constants");
    std::vector<char> buf(wbytes > fbytes ? wbytes : fbytes);
    void* dsts[3] = {h->L.W, h->L.h, h->L.b};
    const size_t sizes[3] = {wbytes, fbytes, fbytes};
    for (int a = 0; a < 3; ++a) {
        if (fread(buf.data(), 1, sizes[a], fc.f) != sizes[a])
            return fail(h, HUGKS_E_ARG, std::string("checkpoint file is truncated: ") + path);
        hipError_t e = hipMemcpy(dsts[a], buf.data(), sizes[a], hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(h, e, "checkpoint_read");
    }
    DevState st;
    hipError_t e = hipMemcpy(&st, h->L.st, sizeof(st), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(h, e, "checkpoint_read");
    st.iter = hd.iter; st.steady_count = hd.steady_count; st.flags = hd.flags;
    st.dt = hd.dt; st.simtime = hd.simtime; st.mass0 = hd.mass0; st.mass_rel = hd.mass_rel;
    for (int q = 0; q < 4; ++q) st.res[q] = hd.res[q];
    for (int q = 0; q < 3; ++q) st.steady[q] = hd.steady[q];
    st.stopped = 0;
    st.stop_reason = HUGKS_STOP_NONE;
    st.steps = 0;
    e = hipMemcpy(h->L.st, &st, sizeof(st), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(h, e, "checkpoint_read");
    return HUGKS_OK;
}

int hugks_kernel_times(hugks_handle* h, int32_t cap, const char** names, double* ms,
                       int32_t* calls, int32_t* count) {
    if (!h) return fail(nullptr, HUGKS_E_ARG, "null handle");
    return kt_report(h->kt, cap, names, ms, calls, count);
}

int hugks_reset_kernel_times(hugks_handle* h) {
    if (!h) return fail(nullptr, HUGKS_E_ARG, "null handle");
    kt_harvest(h->kt);
    h->kt.timers.clear();
    return HUGKS_OK;
}

int hugks_set_timing(hugks_handle* h, int32_t on) {
    if (!h) return fail(nullptr, HUGKS_E_ARG, "null handle");
    h->timing = on != 0;
    return HUGKS_OK;
}


// ----------------------------------------------------------- schedules

int hugks_set_schedule(hugks_handle* h, int32_t flux_kernel, int32_t update_kernel) {
    if (!h) return fail(nullptr, HUGKS_E_ARG, "null handle");
    if (h->in_batch)
        return fail(h, HUGKS_E_ARG, "the case belongs to a batch, whose graphs fix its schedule");
    if (flux_kernel < HUGKS_FLUX_FUSED || flux_kernel > HUGKS_FLUX_CELLSUM ||
        update_kernel < HUGKS_UPDATE_LDS || update_kernel > HUGKS_UPDATE_MFMA_FACE)
        return fail(h, HUGKS_E_ARG, "unknown kernel choice");
    if (h->dim != 2) return HUGKS_OK;   // 1D has one kernel of each kind
    if (update_kernel == HUGKS_UPDATE_MFMA && !update2d_mfma_supported(h->L.c))
        return fail(h, HUGKS_E_UNSUPPORTED, "MFMA update needs nv <= 48 and Nm <= 14");
    if (h->L.c.force_fd && update_kernel != HUGKS_UPDATE_LDS)
        return fail(h, HUGKS_E_UNSUPPORTED,
                    "the finite-difference force runs in HUGKS_UPDATE_LDS only");
    if (h->b3 && update_kernel != HUGKS_UPDATE_MFMA_NODE)
        return fail(h, HUGKS_E_UNSUPPORTED,
                    "HUGKS_MIXED_BF16X3 runs the MFMA node update only");
    if (update_kernel == HUGKS_UPDATE_MFMA_NODE &&
        !update2d_mx_supported(h->L.c, h->fp64, h->b3, h->L.lds_limit, h->L.update_nw))
        return fail(h, HUGKS_E_UNSUPPORTED,
                    "MFMA node update needs nv <= 48, Nm <= 14 and its LDS on the device");
    if (update_kernel == HUGKS_UPDATE_MFMA_FACE &&
        ((flux_kernel != HUGKS_FLUX_CELLSUM && flux_kernel != HUGKS_FLUX_SPLIT) ||
         !update2d_mxf_supported(h->L, h->L.update_nw)))
        return fail(h, HUGKS_E_UNSUPPORTED,
                    "HUGKS_UPDATE_MFMA_FACE needs HUGKS_FLUX_CELLSUM or HUGKS_FLUX_SPLIT, FP32 "
                    "or FP64 storage, the node update's conditions and distribution arrays "
                    "under 4 GiB");
    if (flux_kernel == HUGKS_FLUX_STENCIL && !h->L.stencil_ok)
        return fail(h, HUGKS_E_UNSUPPORTED, "stencil flux needs distribution arrays under 4 GiB");
    if (flux_kernel == HUGKS_FLUX_CELLSUM && h->L.cs_js < 0)
        return fail(h, HUGKS_E_UNSUPPORTED,
                    "cell-sum flux needs nv <= 64 and the v nodes negative before non-negative");
    h->L.flux_mode = flux_kernel;
    h->L.update_mode = update_kernel;
    return HUGKS_OK;
}

int hugks_set_update_waves(hugks_handle* h, int32_t waves) {
    if (!h) return fail(nullptr, HUGKS_E_ARG, "null handle");
    if (h->in_batch)
        return fail(h, HUGKS_E_ARG, "the case belongs to a batch, whose graphs fix its schedule");
    if (h->dim != 2) return HUGKS_OK;
    const int nw = waves > 0 ? waves : update2d_default_waves(h->L, h->L.update_mode);
    if (nw > 16 || !update2d_mx_supported(h->L.c, h->fp64, h->b3, h->L.lds_limit, nw) ||
        (h->L.update_mode == HUGKS_UPDATE_MFMA_FACE && !update2d_mxf_supported(h->L, nw)))
        return fail(h, HUGKS_E_UNSUPPORTED,
                    "no MFMA node-update kernel with " + std::to_string(nw) +
                        " wavefronts per cell for this velocity grid");
    h->L.update_nw = nw;
    return HUGKS_OK;
}

int hugks_get_update_waves(hugks_handle* h, int32_t* waves) {
    if (!h || !waves) return fail(h, HUGKS_E_ARG, "null handle or output");
    *waves = h->dim == 2 ? h->L.update_nw : 0;
    return HUGKS_OK;
}

int hugks_get_schedule(hugks_handle* h, int32_t* flux_kernel, int32_t* update_kernel) {
    if (!h) return fail(nullptr, HUGKS_E_ARG, "null handle");
    if (flux_kernel) *flux_kernel = h->dim == 2 ? h->L.flux_mode : HUGKS_FLUX_FUSED;
    if (update_kernel) *update_kernel = h->dim == 2 ? h->L.update_mode : HUGKS_UPDATE_LDS;
    return HUGKS_OK;
}

}  // extern "C"

// Autotuner: for one case shape, time every kernel schedule and, for the
// MFMA node update, every compiled wavefronts-per-cell tile, on a scratch
// handle of that shape (2 warm-up and 4 timed steps, HIP events on the
// handle's stream), and keep the fastest.  For a case stepped in grouped
// launches only the tile of the grouped schedule (split flux, MFMA node
// update) is searched.  Results are cached per shape for the life of the
// process.
namespace {
struct TuneKey {
    int dim, nx, ny, nv, nm, precision, bc0, bc2, force, grouped;
    bool operator<(const TuneKey& o) const {
        return std::tie(dim, nx, ny, nv, nm, precision, bc0, bc2, force, grouped) <
               std::tie(o.dim, o.nx, o.ny, o.nv, o.nm, o.precision, o.bc0, o.bc2, o.force,
                        o.grouped);
    }
};
struct TuneResult {
    int flux, update, waves;
};
std::map<TuneKey, TuneResult> g_tuned;
std::mutex g_tune_mu;

// ms per step of the handle's current schedule over 4 steps after 2
hipError_t time_schedule(hugks_handle* h, float* ms) {
    hipError_t e = launch_begin(h->L.st, INFINITY, h->stream);
    for (int s = 0; s < 2 && e == hipSuccess; ++s) e = enqueue_step_on(h, h->stream, false);
    if (e == hipSuccess) e = hipEventRecord(h->ev0, h->stream);
    for (int s = 0; s < 4 && e == hipSuccess; ++s) e = enqueue_step_on(h, h->stream, false);
    if (e == hipSuccess) e = hipEventRecord(h->ev1, h->stream);
    if (e == hipSuccess) e = hipEventSynchronize(h->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, h->ev0, h->ev1);
    if (e == hipSuccess) *ms /= 4.f;
    return e;
}
}  // namespace

static int tune_case(const hugks_desc* d, bool grouped, int* flux_kernel, int* update_kernel,
                     int* waves, double* best_ms) {
    // grouped: 1, or 2 with the grouped launches' face-forming update asked
    // off (HUGKS_UPDATE_KERNEL=mfma-node)
    const char* um = getenv("HUGKS_UPDATE_KERNEL");
    const bool plain = um && strcmp(um, "mfma-node") == 0;
    const TuneKey key{d->dim, d->nx, d->ny, d->nv, d->nm, d->precision, d->bc[0], d->bc[2],
                      d->force_variant, grouped ? (plain ? 2 : 1) : 0};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            *flux_kernel = it->second.flux;
            *update_kernel = it->second.update;
            *waves = it->second.waves;
            if (best_ms) *best_ms = -1.0;
            return HUGKS_OK;
        }
    }
    hugks_handle* h = nullptr;
    int rc = hugks_create(d, &h);
    if (rc != HUGKS_OK) return rc;
    TuneResult r{h->L.flux_mode, h->L.update_mode, h->dim == 2 ? h->L.update_nw : 0};
    double best = 1e300;
    if (d->dim == 2) {
        // a case of the grouped launches: the tile for the schedule its
        // plan runs (the cell-sum flux and the face-forming update where the
        // case supports them, as batch_group_plan picks them)
        const bool gface = !plain && update2d_mxf_supported(h->L, h->L.update_nw);
        const int gfk = h->L.cs_js >= 0 ? HUGKS_FLUX_CELLSUM : HUGKS_FLUX_SPLIT;
        const int guk = gface ? HUGKS_UPDATE_MFMA_FACE : HUGKS_UPDATE_MFMA_NODE;
        const int f0 = grouped ? gfk : HUGKS_FLUX_FUSED;
        const int f1 = grouped ? gfk : HUGKS_FLUX_CELLSUM;
        const int u0 = grouped ? guk : HUGKS_UPDATE_LDS;
        const int u1 = grouped ? guk : HUGKS_UPDATE_MFMA_FACE;
        for (int fk = f0; fk <= f1; ++fk)
            for (int uk = u0; uk <= u1; ++uk) {
                if (hugks_set_schedule(h, fk, uk) != HUGKS_OK) continue;
                const bool node = uk >= HUGKS_UPDATE_MFMA_NODE;
                for (int nw = node ? 1 : 0; nw <= (node ? 16 : 0); ++nw) {
                    if (node && hugks_set_update_waves(h, nw) != HUGKS_OK) continue;
                    float ms = 0.f;
                    hipError_t e = time_schedule(h, &ms);
                    if (e != hipSuccess) {
                        rc = hip_fail(nullptr, e, "autotune");
                        hugks_destroy(h);
                        return rc;
                    }
                    if (ms < best) {
                        best = ms;
                        r = TuneResult{fk, uk, node ? nw : h->L.update_nw};
                    }
                }
            }
    }
    hugks_destroy(h);
    *flux_kernel = r.flux;
    *update_kernel = r.update;
    *waves = r.waves;
    if (best_ms) *best_ms = best < 1e300 ? best : -1.0;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[key] = r;
    return HUGKS_OK;
}

extern "C" {

int hugks_tune(const hugks_desc* d, int32_t* flux_kernel, int32_t* update_kernel,
               double* ms_per_step) {
    int32_t nw = 0;
    return hugks_tune_tiles(d, flux_kernel, update_kernel, &nw, ms_per_step);
}

int hugks_tune_tiles(const hugks_desc* d, int32_t* flux_kernel, int32_t* update_kernel,
                     int32_t* update_waves, double* ms_per_step) {
    if (!d || !flux_kernel || !update_kernel || !update_waves)
        return fail(nullptr, HUGKS_E_ARG, "null descriptor or output");
    int fk = 0, uk = 0, nw = 0;
    int rc = tune_case(d, false, &fk, &uk, &nw, ms_per_step);
    if (rc != HUGKS_OK) return rc;
    *flux_kernel = fk;
    *update_kernel = uk;
    *update_waves = nw;
    return HUGKS_OK;
}

// ------------------------------------------------------------ batches

}  // extern "C"

struct hugks_batch {
    std::vector<hugks_handle*> cases;
    bool grouped = false;       // some cases run in grouped launches (k_group.h)
    std::vector<char> in_group; // per case: stepped by the grouped launches
    std::vector<GroupPlan> plans;   // one per storage type present
    std::vector<void*> gmem;    // per-case arguments and the plans' tables
    DevState** d_sts = nullptr;     // the cases' device states (device array)
    DevState* d_gather = nullptr;   // their copies side by side (device)
    DevState* h_gather = nullptr;   // and on the host (pinned)
    hipStream_t main = nullptr;
    std::vector<hipStream_t> lanes;
    std::vector<hipEvent_t> joins;
    hipEvent_t fork = nullptr;
    hipGraph_t graph = nullptr;     // one step of every case
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph_k = nullptr;   // graph_steps steps, one after the other
    hipGraphExec_t exec_k = nullptr;
    int graph_steps = 1;
    std::string err;
    bool timing = false;   // eager steps with timed grouped launches
    KTimes kt;
};

static void batch_release(hugks_batch* b) {
    if (!b) return;
    if (b->main) (void)hipStreamSynchronize(b->main);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->exec_k) (void)hipGraphExecDestroy(b->exec_k);
    if (b->graph_k) (void)hipGraphDestroy(b->graph_k);
    for (hipEvent_t e : b->joins) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (hipStream_t s : b->lanes) (void)hipStreamDestroy(s);
    if (b->main) (void)hipStreamDestroy(b->main);
    kt_release(b->kt);
    for (void* p : b->gmem) (void)hipFree(p);
    if (b->d_sts) (void)hipFree(b->d_sts);
    if (b->d_gather) (void)hipFree(b->d_gather);
    if (b->h_gather) (void)hipHostFree(b->h_gather);
    for (hugks_handle* h : b->cases) hugks_destroy(h);
    delete b;
}

static int batch_fail(hugks_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    else g_create_error = msg;
    return code;
}

extern "C" {

// The small 2D cases whose step can be the split flux with factoring
// weights and the MFMA update run in grouped launches (one launch per
// kernel type and storage type for all of them) when there are at least two;
// the other cases (and those of the bf16x3 contraction) keep their own
// launches on the lane streams.
static bool case_groupable(const hugks_handle* h) {
    const Launch2D& L = h->L;
    return h->dim == 2 && L.wx && !h->b3 &&
           update2d_mx_supported(L.c, h->fp64, false, L.lds_limit, L.update_nw) &&
           L.c.nx * L.c.ny < PRED_MIN_CELLS;
}
static std::vector<char> batch_groupable(const hugks_batch* b) {
    std::vector<char> g(b->cases.size(), 0);
    if (const char* e = getenv("HUGKS_BATCH_GROUPED"))
        if (strcmp(e, "0") == 0) return g;
    int count = 0;
    for (size_t i = 0; i < b->cases.size(); ++i) {
        g[i] = case_groupable(b->cases[i]) ? 1 : 0;
        count += g[i];
    }
    if (count < 2) std::fill(g.begin(), g.end(), 0);
    return g;
}

// Plan of the grouped launches of the cases ids (one storage type): tables
// ([n, start[0..n], case[0..n-1]] each, case = index into Ls) in one device
// allocation.
static hipError_t batch_group_plan_of(hugks_batch* b, const std::vector<int>& ids,
                                      const Launch2D* Ls) {
    std::vector<std::vector<int>> tabs;
    std::vector<int> blocks;
    auto table = [&](const std::vector<std::pair<int, int>>& cs) {   // (case, workgroups)
        std::vector<int> t(1 + cs.size() + 1 + cs.size());
        t[0] = (int)cs.size();
        int acc = 0;
        for (size_t e = 0; e < cs.size(); ++e) {
            t[1 + e] = acc;
            acc += cs[e].second;
            t[2 + cs.size() + e] = cs[e].first;
        }
        t[1 + cs.size()] = acc;
        tabs.push_back(t);
        blocks.push_back(acc);
        return (int)tabs.size() - 1;
    };
    std::vector<std::pair<int, int>> one, sl, fs, ff, fn, wl, cp, fd;
    std::map<std::tuple<int, int, int>, std::vector<std::pair<int, int>>> upd;
    std::map<std::tuple<int, int, int>, size_t> upd_lds;
    for (int i : ids) {
        hugks_handle* h = b->cases[i];
        int s1, s2, s3, s4, s5, s6, e, t, v;
        group_blocks(h->L, &s1, &s2, &s3, &s4, &s5, &s6);
        one.push_back({i, 1});
        sl.push_back({i, s1});
        fs.push_back({i, s2});
        ff.push_back({i, s3});
        fn.push_back({i, s4});
        if (s5 > 0) wl.push_back({i, s5});
        cp.push_back({i, s6});
        fd.push_back({i, group_fdiv_blocks(h->L)});
        group_update_key(h->L, &e, &t, &v);
        const auto key = std::make_tuple(e, t, v);
        upd[key].push_back({i, h->L.c.nx * h->L.c.ny});
        const size_t lds = group_update_lds(h->L);
        upd_lds[key] = std::max(upd_lds[key], lds);
    }
    GroupPlan P{};
    if (upd.size() > sizeof(P.upd) / sizeof(P.upd[0])) return hipErrorInvalidValue;
    P.fp64 = b->cases[ids[0]]->fp64;
    P.cellsum = b->cases[ids[0]]->L.flux_mode == HUGKS_FLUX_CELLSUM;
    P.face = b->cases[ids[0]]->L.update_mode == HUGKS_UPDATE_MFMA_FACE;
    P.cs_vec = true;
    for (int i : ids) {
        const int es = b->cases[i]->fp64 ? 8 : 4;
        P.cs_vec = P.cs_vec && (b->cases[i]->L.c.nv * es) % 16 == 0;
    }
    const int k_ts = table(one), k_sl = table(sl), k_fs = table(fs), k_ff = table(ff),
              k_fn = table(fn), k_wl = wl.empty() ? -1 : table(wl), k_cp = table(cp),
              k_fd = table(fd);
    std::vector<int> k_up;
    for (auto& kv : upd) k_up.push_back(table(kv.second));
    size_t ints = 0;
    std::vector<size_t> off;
    for (auto& t : tabs) {
        off.push_back(ints);
        ints += t.size();
    }
    void* tmem = nullptr;
    hipError_t e = hipMalloc(&tmem, ints * sizeof(int));
    if (e != hipSuccess) return e;
    b->gmem.push_back(tmem);
    std::vector<int> flat;
    for (auto& t : tabs) flat.insert(flat.end(), t.begin(), t.end());
    e = hipMemcpy(tmem, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    auto tp = [&](int k) { return static_cast<const int*>(tmem) + off[k]; };
    P.Ls = Ls;
    P.t_ts = P.t_fin = tp(k_ts);
    P.g_ts = P.g_fin = blocks[k_ts];
    P.t_slope = tp(k_sl); P.g_slope = blocks[k_sl];
    P.t_fsum = tp(k_fs); P.g_fsum = blocks[k_fs];
    P.t_fface = tp(k_ff); P.g_fface = blocks[k_ff];
    P.t_fnode = tp(k_fn); P.g_fnode = blocks[k_fn];
    P.t_wall = k_wl >= 0 ? tp(k_wl) : nullptr;
    P.g_wall = k_wl >= 0 ? blocks[k_wl] : 0;
    P.t_cpre = tp(k_cp); P.g_cpre = blocks[k_cp];
    P.t_fdiv = tp(k_fd); P.g_fdiv = blocks[k_fd];
    P.nupd = 0;
    int q = 0;
    for (auto& kv : upd) {
        GroupUpd& u = P.upd[P.nupd++];
        u.nw = std::get<0>(kv.first);
        u.maxp = std::get<1>(kv.first);
        u.nvc = std::get<2>(kv.first);
        u.blocks = blocks[k_up[q]];
        u.tab = tp(k_up[q]);
        u.lds = upd_lds[kv.first];
        u.ff = P.face;
        ++q;
        e = prepare_group_update(P.fp64, u);
        if (e != hipSuccess) return e;
    }
    b->plans.push_back(P);
    return hipSuccess;
}

// Per-case arguments to device memory, then one plan per storage type.
static hipError_t batch_group_plan(hugks_batch* b) {
    const int n = (int)b->cases.size();
    void* lmem = nullptr;
    hipError_t e = hipMalloc(&lmem, n * sizeof(Launch2D));
    if (e != hipSuccess) return e;
    b->gmem.push_back(lmem);
    std::vector<Launch2D> Ls(n);
    std::vector<int> ids[2];
    for (int i = 0; i < n; ++i)
        if (b->in_group[i]) ids[b->cases[i]->fp64 ? 1 : 0].push_back(i);
    // the grouped schedule: the cell-sum flux where every case of the
    // storage type supports it, else the split flux; the node update forming
    // the interior faces' equilibrium flux where every case supports it
    // (HUGKS_UPDATE_KERNEL=mfma-node: the plain node update), else the MFMA
    // node update
    const char* um = getenv("HUGKS_UPDATE_KERNEL");
    const bool face_ok = !(um && strcmp(um, "mfma-node") == 0);
    for (int t = 0; t < 2; ++t) {
        bool cs = true, face = face_ok;
        for (int i : ids[t]) {
            cs = cs && b->cases[i]->L.cs_js >= 0;
            face = face && update2d_mxf_supported(b->cases[i]->L, b->cases[i]->L.update_nw);
        }
        for (int i : ids[t])
            hugks_set_schedule(b->cases[i], cs ? HUGKS_FLUX_CELLSUM : HUGKS_FLUX_SPLIT,
                               face ? HUGKS_UPDATE_MFMA_FACE : HUGKS_UPDATE_MFMA_NODE);
    }
    for (int i = 0; i < n; ++i) Ls[i] = b->cases[i]->L;
    e = hipMemcpy(lmem, Ls.data(), n * sizeof(Launch2D), hipMemcpyHostToDevice);
    for (int t = 0; t < 2 && e == hipSuccess; ++t)
        if (!ids[t].empty()) e = batch_group_plan_of(b, ids[t], static_cast<const Launch2D*>(lmem));
    return e;
}

int hugks_batch_create(int32_t n, const hugks_desc* descs, int32_t flags, hugks_batch** out) {
    if (n <= 0 || !descs || !out) return batch_fail(nullptr, HUGKS_E_ARG, "empty batch or null pointer");
    *out = nullptr;
    hugks_batch* b = new hugks_batch();
    for (int32_t i = 0; i < n; ++i) {
        hugks_handle* h = nullptr;
        int rc = hugks_create(&descs[i], &h);
        if (rc != HUGKS_OK) {
            std::string m = "case " + std::to_string(i) + ": " + g_create_error;
            batch_release(b);
            return batch_fail(nullptr, rc, m);
        }
        b->cases.push_back(h);
    }
    // grouped cases run the grouped schedule (split flux, MFMA node
    // update), for which the tuner picks the wavefronts per cell; it picks
    // schedule and tile of the cases stepped on their own
    b->in_group = batch_groupable(b);
    for (int32_t i = 0; i < n && (flags & HUGKS_BATCH_AUTOTUNE); ++i) {
        int fk = -1, uk = -1, nw = 0;
        int rc = tune_case(&descs[i], b->in_group[i] != 0, &fk, &uk, &nw, nullptr);
        if (rc == HUGKS_OK && b->cases[i]->dim == 2) {
            rc = hugks_set_schedule(b->cases[i], fk, uk);
            if (rc == HUGKS_OK && uk >= HUGKS_UPDATE_MFMA_NODE)
                rc = hugks_set_update_waves(b->cases[i], nw);
            if (rc != HUGKS_OK) g_create_error = b->cases[i]->err;
        }
        if (rc != HUGKS_OK) {
            std::string m = "case " + std::to_string(i) + ": " + g_create_error;
            batch_release(b);
            return batch_fail(nullptr, rc, m);
        }
    }
    // one step of every case, fanned out over a few streams and captured
    // into one graph: a replay advances the whole batch by one step
    hipError_t e = hipStreamCreateWithFlags(&b->main, hipStreamNonBlocking);
    b->kt.stream = b->main;
    // the cases' states: reset and read back for all cases at once
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_sts, n * sizeof(DevState*));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_gather, n * sizeof(DevState));
    if (e == hipSuccess) e = hipHostMalloc((void**)&b->h_gather, n * sizeof(DevState), 0);
    if (e == hipSuccess) {
        std::vector<DevState*> sts(n);
        for (int32_t i = 0; i < n; ++i) sts[i] = b->cases[i]->L.st;
        e = hipMemcpy(b->d_sts, sts.data(), n * sizeof(DevState*), hipMemcpyHostToDevice);
    }
    int lanes = 8;
    if (const char* ls = getenv("HUGKS_BATCH_LANES")) lanes = atoi(ls) > 0 ? atoi(ls) : lanes;
    const int nl = n < lanes ? n : lanes;
    for (int l = 0; l < nl && e == hipSuccess; ++l) {
        hipStream_t s = nullptr;
        e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e == hipSuccess) b->lanes.push_back(s);
        hipEvent_t ev = nullptr;
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) b->joins.push_back(ev);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    // cases finish their own set-up on their streams first
    for (hugks_handle* h : b->cases)
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    int nsep = 0;   // cases with their own launches
    for (int32_t i = 0; i < n; ++i) nsep += b->in_group[i] ? 0 : 1;
    b->grouped = e == hipSuccess && nsep < n;
    if (b->grouped) e = batch_group_plan(b);
    // k steps of every case captured into one graph: per step, the cases
    // with their own launches on the lanes, forked from the batch stream,
    // and the grouped launches on the batch stream beside them
    const int nlu = nsep < nl ? nsep : nl;
    auto capture = [&](int k, hipGraph_t* graph, hipGraphExec_t* exec) {
        hipError_t r = hipStreamBeginCapture(b->main, hipStreamCaptureModeThreadLocal);
        for (int step = 0; step < k && r == hipSuccess; ++step) {
            if (nlu > 0) {
                r = hipEventRecord(b->fork, b->main);
                for (int l = 0; l < nlu && r == hipSuccess; ++l)
                    r = hipStreamWaitEvent(b->lanes[l], b->fork, 0);
                int q = 0;
                for (int32_t i = 0; i < n && r == hipSuccess; ++i)
                    if (!b->in_group[i]) r = enqueue_step_on(b->cases[i], b->lanes[q++ % nlu], false);
            }
            for (size_t q = 0; q < b->plans.size() && r == hipSuccess; ++q)
                r = launch_group_step(b->plans[q], b->main);
            for (int l = 0; l < nlu && r == hipSuccess; ++l) {
                r = hipEventRecord(b->joins[l], b->lanes[l]);
                if (r == hipSuccess) r = hipStreamWaitEvent(b->main, b->joins[l], 0);
            }
        }
        if (r == hipSuccess) {
            r = hipStreamEndCapture(b->main, graph);
        } else {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(b->main, &g);
            if (g) (void)hipGraphDestroy(g);
        }
        if (r == hipSuccess) r = hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0);
        return r;
    };
    // the multi-step graph: a case that stops partway through turns the
    // rest of its launches into no-ops (its device state's stop flag)
    b->graph_steps = 8;
    if (const char* gs = getenv("HUGKS_BATCH_GRAPH_STEPS"))
        b->graph_steps = atoi(gs) > 0 ? atoi(gs) : b->graph_steps;
    if (e == hipSuccess) e = capture(1, &b->graph, &b->exec);
    if (e == hipSuccess && b->graph_steps > 1) e = capture(b->graph_steps, &b->graph_k, &b->exec_k);
    if (e != hipSuccess) {
        std::string m = std::string("batch graph: ") + hipGetErrorString(e);
        batch_release(b);
        return batch_fail(nullptr, e == hipErrorOutOfMemory ? HUGKS_E_NOMEM : HUGKS_E_DEVICE, m);
    }
    // the graphs hold each case's launches: its schedule is fixed from here
    for (hugks_handle* h : b->cases) h->in_batch = true;
    *out = b;
    return HUGKS_OK;
}

int hugks_batch_destroy(hugks_batch* b) {
    batch_release(b);
    return HUGKS_OK;
}

const char* hugks_batch_last_error(const hugks_batch* b) {
    return b ? b->err.c_str() : g_create_error.c_str();
}

int32_t hugks_batch_size(const hugks_batch* b) { return b ? (int32_t)b->cases.size() : 0; }

int32_t hugks_batch_grouped(const hugks_batch* b) {
    if (!b) return 0;
    int32_t k = 0;
    for (char g : b->in_group) k += g ? 1 : 0;
    return k;
}

int hugks_batch_handle(hugks_batch* b, int32_t i, hugks_handle** h) {
    if (!b || !h || i < 0 || i >= (int32_t)b->cases.size())
        return batch_fail(b, HUGKS_E_ARG, "case index out of range");
    *h = b->cases[i];
    return HUGKS_OK;
}

// One step of every case: the captured graph, or, with timing on, the same
// launches issued eagerly on the batch stream, the grouped ones bracketed
// by timing events (the cases with their own launches run untimed).
static hipError_t batch_step(hugks_batch* b) {
    if (!b->timing) return hipGraphLaunch(b->exec, b->main);
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < b->cases.size() && e == hipSuccess; ++i)
        if (!b->in_group[i]) e = enqueue_step_on(b->cases[i], b->main, false);
    for (size_t q = 0; q < b->plans.size() && e == hipSuccess; ++q) {
        GroupPlan P = b->plans[q];
        P.hook = kt_hook;
        P.hook_ctx = &b->kt;
        e = launch_group_step(P, b->main);
    }
    return e;
}

// n steps of every case: replays of the multi-step graph, then of the
// one-step graph for the remainder
static hipError_t batch_steps(hugks_batch* b, int n) {
    hipError_t e = hipSuccess;
    if (!b->timing && b->exec_k)
        for (; n >= b->graph_steps && e == hipSuccess; n -= b->graph_steps)
            e = hipGraphLaunch(b->exec_k, b->main);
    for (; n > 0 && e == hipSuccess; --n) e = batch_step(b);
    return e;
}

int hugks_batch_set_timing(hugks_batch* b, int32_t on) {
    if (!b) return batch_fail(nullptr, HUGKS_E_ARG, "null batch");
    b->timing = on != 0;
    return HUGKS_OK;
}

int hugks_batch_kernel_times(hugks_batch* b, int32_t cap, const char** names, double* ms,
                             int32_t* calls, int32_t* count) {
    if (!b) return batch_fail(nullptr, HUGKS_E_ARG, "null batch");
    return kt_report(b->kt, cap, names, ms, calls, count);
}

int hugks_batch_reset_kernel_times(hugks_batch* b) {
    if (!b) return batch_fail(nullptr, HUGKS_E_ARG, "null batch");
    kt_harvest(b->kt);
    b->kt.timers.clear();
    return HUGKS_OK;
}

int hugks_batch_advance(hugks_batch* b, int32_t max_steps, hugks_step_info* infos) {
    if (!b) return batch_fail(nullptr, HUGKS_E_ARG, "null batch");
    const int n = (int)b->cases.size();
    // every case's state to the host: one gather kernel and one copy
    auto fetch = [&]() {
        hipError_t r = launch_gather_states(b->d_sts, n, b->d_gather, b->main);
        if (r == hipSuccess)
            r = hipMemcpyAsync(b->h_gather, b->d_gather, n * sizeof(DevState),
                               hipMemcpyDeviceToHost, b->main);
        if (r == hipSuccess) r = hipStreamSynchronize(b->main);
        return r;
    };
    hipError_t e = launch_begin_all(b->d_sts, n, INFINITY, b->main);
    if (max_steps > 0) {
        e = batch_steps(b, max_steps);
    } else {
        // no cap: replay until every case met an exit condition
        for (;;) {
            e = batch_steps(b, 32);
            if (e == hipSuccess) e = fetch();
            if (e != hipSuccess) break;
            bool all = true;
            for (int i = 0; i < n; ++i) all = all && (b->h_gather[i].stopped || b->h_gather[i].flags);
            if (all) break;
        }
    }
    if (e == hipSuccess) e = fetch();
    if (e != hipSuccess) return batch_fail(b, HUGKS_E_DEVICE, std::string("batch advance: ") + hipGetErrorString(e));
    int first = HUGKS_OK;
    for (int i = 0; i < n; ++i) {
        int rc = state_info(b->cases[i], b->h_gather[i], infos ? &infos[i] : nullptr);
        if (rc != HUGKS_OK && first == HUGKS_OK) {
            first = rc;
            b->err = "case " + std::to_string(i) + ": " + b->cases[i]->err;
        }
    }
    return first;
}

}  // extern "C"
