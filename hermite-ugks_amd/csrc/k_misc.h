// Time-step size, step finalisation (residual, loop state, exit tests),
// initialisation of f, output moments and layout conversion (gfx950).
#pragma once
#include "hugks_device.h"
#include "hugks_kernels.h"

namespace hugks {

// The single-case kernels of this header are defined only where
// HUGKS_BODIES_ONLY is not set: the grouped-launch file (launch_group.hip)
// takes the device bodies alone, so no kernel is defined in two files.

constexpr int RED_NT = 1024;

template <int NV, int NT = RED_NT>
__device__ __forceinline__ void block_max(double (&v)[NV], double* lds) {
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double x = v[i];
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, WAVE));
        v[i] = x;
    }
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) lds[wid * NV + i] = v[i];
    __syncthreads();
    // one wave's partials in registers at a time (as block_sum)
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = lds[i];
#pragma unroll 1
    for (int w = 1; w < NT / WAVE; ++w) {
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] = fmax(v[i], lds[w * NV + i]);
    }
    __syncthreads();
}

// Marks the start of a host call: later kernels run again even if a
// previous call ended on an exit condition, unless a device flag is raised
// (the state of a failed step is not stepped further; hugks_set_f and a
// checkpoint restore clear the flags).
__device__ __forceinline__ void begin_state(DevState* st, double t_stop) {
    st->stopped = st->flags ? 1 : 0;
    st->stop_reason = st->flags ? HUGKS_STOP_ERROR : 0;
    st->steps = 0;
    st->t_stop = t_stop;
}
#ifndef HUGKS_BODIES_ONLY
__global__ void k_begin(DevState* st, double t_stop) {
    if (threadIdx.x == 0) begin_state(st, t_stop);
}
// k_begin for every case of a batch, one thread per case.
__global__ void k_begin_all(DevState* const* __restrict__ sts, int n, double t_stop) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) begin_state(sts[i], t_stop);
}
// The states of a batch's cases side by side, for one copy to the host.
__global__ void k_gather_states(DevState* const* __restrict__ sts, int n,
                                DevState* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = *sts[i];
}
#endif

// Signal rate of one cell for the time step (timestep, 2d/src/solver.f90:
// 13-61 and 1d/src/solver.f90:16-47).
__device__ __forceinline__ double cell_rate(const Consts& c, int dim, const double* __restrict__ W,
                                            int cell) {
    if (dim == 2) {
        double p[4];
        prim_from_w(c, W + 4 * (size_t)cell, p);
        const double sos = sqrt(0.5 * c.gamma / p[3]);
        const double pu = fmax(c.umax_dt, fabs(p[1])) + sos;
        const double pv = fmax(c.vmax_dt, fabs(p[2])) + sos;
        return c.dt_rule == HUGKS_DT_2D_WALL_NORMAL ? pv / c.dy
                                                    : (c.dy * pu + c.dx * pv) / (c.dx * c.dy);
    }
    const double* w = W + 3 * (size_t)cell;
    const double U = w[1] / w[0];
    const double lam = 0.5 * w[0] / (c.gamma - 1.0) / (w[2] - 0.5 * w[1] * w[1] / w[0]);
    const double sos = sqrt(0.5 * c.gamma / lam);
    return (fmax(c.umax_dt, fabs(U)) + sos) / c.dx;
}

// Reductions over the cells run in two passes: PRED_NT-thread workgroups
// reduce strided cell ranges to one partial each (at most PRED_NB), then
// one workgroup reduces the partials in index order.
constexpr int PRED_NT = 256;
constexpr int PRED_NB = 1024;
constexpr int PRED_LD = 9;     // per-partial slots of the finalize pass
static_assert(PRED_NB * (1 + PRED_LD) <= RED_SCRATCH, "reduction scratch");

// First pass of the time step: the largest signal rate per workgroup.
#ifndef HUGKS_BODIES_ONLY
__global__ __launch_bounds__(PRED_NT) void k_rate_part(Consts c, int dim,
                                                       const DevState* __restrict__ st,
                                                       const double* __restrict__ W,
                                                       double* __restrict__ red) {
    __shared__ double lds[PRED_NT / WAVE];
    if (st->stopped) return;
    const int ncell = c.nx * c.ny;
    double t[1] = {0.0};
    for (int cell = blockIdx.x * PRED_NT + threadIdx.x; cell < ncell; cell += gridDim.x * PRED_NT)
        t[0] = fmax(t[0], cell_rate(c, dim, W, cell));
    block_max<1, PRED_NT>(t, lds);
    if (threadIdx.x == 0) red[blockIdx.x] = t[0];
}
#endif

// timestep: dt from the largest rate, over the nred partials of
// k_rate_part or, with nred = 0, straight over the cells (small grids: one
// launch instead of two).  The 1D loop tests simtime >= maxtime before the
// step (1d/src/main.f90:84), so that test is made here.
__device__ __forceinline__ void k_timestep_body(Consts c, int dim, DevState* st,
                                                     const double* __restrict__ W,
                                                     const double* __restrict__ red, int nred,
                                                     int bid, int nblk) {
    __shared__ double lds[RED_NT / WAVE];
    if (st->stopped) return;
    if (dim == 1 && st->simtime >= c.maxtime) {
        if (threadIdx.x == 0) {
            st->stopped = 1;
            st->stop_reason = HUGKS_STOP_MAXTIME;
        }
        return;
    }
    double t[1] = {0.0};
    if (nred > 0) {
        for (int q = threadIdx.x; q < nred; q += RED_NT) t[0] = fmax(t[0], red[q]);
    } else {
        for (int cell = threadIdx.x; cell < c.nx * c.ny; cell += RED_NT)
            t[0] = fmax(t[0], cell_rate(c, dim, W, cell));
    }
    block_max<1>(t, lds);
    if (threadIdx.x == 0) {
        const double temp = fmax(t[0], c.force_rate);
        const double tmin = 1.0 / temp;
        double dt = c.cfl * tmin / 1.1;
        if (c.clamp_dt) dt = fmin(dt, c.maxtime - st->simtime);
        st->dt = dt;
    }
}
#ifndef HUGKS_BODIES_ONLY
__global__ __launch_bounds__(RED_NT) void k_timestep(Consts c, int dim, DevState* st,
                                                     const double* __restrict__ W,
                                                     const double* __restrict__ red, int nred) {
    k_timestep_body(c, dim, st, W, red, nred, (int)blockIdx.x, (int)gridDim.x);
}
#endif

// First pass of the finalisation: per workgroup, the sums of the per-cell
// residual partials (fixed order) and, in 1D, the maxima of the steady
// residuals.
#ifndef HUGKS_BODIES_ONLY
__global__ __launch_bounds__(PRED_NT) void k_part_sum(Consts c, int dim,
                                                      const DevState* __restrict__ st,
                                                      const double* __restrict__ part,
                                                      double* __restrict__ red) {
    __shared__ double lds[6 * (PRED_NT / WAVE)];
    if (st->stopped) return;
    const int ncell = c.nx * c.ny;
    const int np = dim == 2 ? 5 : 9;
    double s[6] = {0, 0, 0, 0, 0, 0};
    double mx[3] = {0, 0, 0};
    for (int cell = blockIdx.x * PRED_NT + threadIdx.x; cell < ncell; cell += gridDim.x * PRED_NT) {
        const double* p = part + (size_t)np * cell;
        if (dim == 2) {
            for (int q = 0; q < 5; ++q) s[q] += p[q];
        } else {
            for (int q = 0; q < 6; ++q) s[q] += p[q];
            for (int q = 0; q < 3; ++q) mx[q] = fmax(mx[q], p[6 + q]);
        }
    }
    block_sum<6, PRED_NT>(s, lds);
    if (dim == 1) block_max<3, PRED_NT>(mx, lds);
    if (threadIdx.x == 0) {
        double* r = red + (size_t)PRED_LD * blockIdx.x;
        for (int q = 0; q < 6; ++q) r[q] = s[q];
        for (int q = 0; q < 3; ++q) r[6 + q] = mx[q];
    }
}
#endif

// Residual and loop state after update(): 2d/src/solver.f90:393 with
// 2d/src/main.f90:82,102-104; 1d/src/solver.f90:262 with 1d/src/main.f90:
// 106-107.  Partials per cell: 2D (dw^2 x4, |rho|); 1D (dw^2 x3, |w| x3,
// steady residual x3), reduced by k_part_sum; one workgroup sums its
// partials in index order.
__device__ __forceinline__ void k_finalize_body(Consts c, int dim, DevState* st,
                                                     const double* __restrict__ part,
                                                     const double* __restrict__ red, int nred,
                                                     int bid, int nblk) {
    __shared__ double lds[9 * (RED_NT / WAVE)];
    if (st->stopped) return;
    const int ncell = c.nx * c.ny;
    double s[6] = {0, 0, 0, 0, 0, 0};
    double mx[3] = {0, 0, 0};
    if (nred > 0) {
        for (int b = threadIdx.x; b < nred; b += RED_NT) {
            const double* r = red + (size_t)PRED_LD * b;
            for (int q = 0; q < 6; ++q) s[q] += r[q];
            for (int q = 0; q < 3; ++q) mx[q] = fmax(mx[q], r[6 + q]);
        }
    } else {
        // small grids: straight over the per-cell partials
        const int np = dim == 2 ? 5 : 9;
        for (int cell = threadIdx.x; cell < ncell; cell += RED_NT) {
            const double* p = part + (size_t)np * cell;
            if (dim == 2) {
                for (int q = 0; q < 5; ++q) s[q] += p[q];
            } else {
                for (int q = 0; q < 6; ++q) s[q] += p[q];
                for (int q = 0; q < 3; ++q) mx[q] = fmax(mx[q], p[6 + q]);
            }
        }
    }
    block_sum<6, RED_NT>(s, lds);
    if (dim == 1) block_max<3>(mx, lds);
    if (threadIdx.x == 0) {
        const double n = (double)ncell;
        bool all_small = true;
        if (dim == 2) {
            for (int q = 0; q < 4; ++q) {
                st->res[q] = sqrt(n * s[q]) / (s[4] + c.smv);
                all_small = all_small && st->res[q] < c.eps_res;
            }
            st->simtime = st->simtime + st->dt;
            st->iter += 1;
        } else {
            for (int q = 0; q < 3; ++q) {
                st->res[q] = sqrt(n * s[q]) / (s[3 + q] + c.smv);
                st->steady[q] = mx[q];
            }
            st->iter += 1;
            st->simtime = st->simtime + st->dt;
            all_small = false;
        }
        st->steps += 1;
        // mass = sum of rho * cell size (rho > 0 is enforced by the update)
        const double mass = (dim == 2 ? s[4] * (c.dx * c.dy) : s[3] * c.dx);
        st->mass_rel = (mass - st->mass0) / (fabs(st->mass0) + c.smv);
        if (st->flags) {
            // the reference stops the run at the error (error stop); later
            // steps of this call do nothing and the host returns the status
            st->stopped = 1;
            st->stop_reason = HUGKS_STOP_ERROR;
        } else if (dim == 2 && st->simtime >= c.maxtime) {
            st->stopped = 1;
            st->stop_reason = HUGKS_STOP_MAXTIME;
        } else if (dim == 2 && c.stop_on_residual && all_small) {
            st->stopped = 1;
            st->stop_reason = HUGKS_STOP_RESIDUAL;
        }
        // Fourier steady test (fourier_is_steady, 1d/src/solver.f90:266-274;
        // counting and exit, 1d/src/main.f90:108-126)
        if (dim == 1 && c.steady_steps > 0) {
            const double m = fmax(fmax(fabs(st->steady[0]), fabs(st->steady[1])), fabs(st->steady[2]));
            if (m <= c.steady_tol && fabs(st->mass_rel) <= c.mass_tol)
                st->steady_count += 1;
            else
                st->steady_count = 0;
            if (st->steady_count >= c.steady_steps && !st->stopped) {
                st->stopped = 1;
                st->stop_reason = HUGKS_STOP_STEADY;
            }
        }
        if (!st->stopped && st->simtime >= st->t_stop) {
            st->stopped = 1;
            st->stop_reason = HUGKS_STOP_TIME;
        }
    }
}
#ifndef HUGKS_BODIES_ONLY
__global__ __launch_bounds__(RED_NT) void k_finalize(Consts c, int dim, DevState* st,
                                                     const double* __restrict__ part,
                                                     const double* __restrict__ red, int nred) {
    k_finalize_body(c, dim, st, part, red, nred, (int)blockIdx.x, (int)gridDim.x);
}
#endif

// Reduced Maxwellian initial state, 2d/src/initialization.f90:227-235
// (1d/src/initialization.f90:185-192) with W = get_conserved(prim0).
#ifndef HUGKS_BODIES_ONLY
__global__ void k_init_f(Consts c, int dim, const double* __restrict__ prim0,
                         const double* __restrict__ u1, const double* __restrict__ v1,
                         double* __restrict__ W, void* h, void* b, int fp64) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t ncell = (size_t)c.nx * c.ny;
    if (idx >= ncell * c.ldv) return;
    const size_t cell = idx / c.ldv;
    const int k = (int)(idx - cell * c.ldv);
    double H = 0.0, B = 0.0;
    if (dim == 2) {
        const double* p = prim0 + 4 * cell;
        if (k < c.nn) {
            const int iu = k % c.nv, jv = k / c.nv;
            const double cu = u1[iu] - p[1], cv = v1[jv] - p[2];
            H = p[0] * (p[3] / c.pi) * exp(-p[3] * (cu * cu + cv * cv));
            B = H * c.K / (2.0 * p[3]);
        }
        if (k == 0) w_from_prim(c, p, W + 4 * cell);
    } else {
        const double* p = prim0 + 3 * cell;
        if (k < c.nn) {
            const double cu = u1[k] - p[1];
            H = p[0] * pow(p[2] / c.pi, 0.5) * exp(-p[2] * cu * cu);
            B = H * c.K / (2.0 * p[2]);
        }
        if (k == 0) {
            double* w = W + 3 * cell;
            w[0] = p[0];
            w[1] = p[0] * p[1];
            w[2] = 0.5 * p[0] / p[2] / (c.gamma - 1.0) + 0.5 * p[0] * p[1] * p[1];
        }
    }
    if (fp64) {
        ((double*)h)[idx] = H;
        ((double*)b)[idx] = B;
    } else {
        ((float*)h)[idx] = (float)H;
        ((float*)b)[idx] = (float)B;
    }
}
#endif

// Heat flux per cell for output: get_heat_flux(h, b, uspace, vspace, prim)
// (2d/src/gastheory2d.f90:121-129; 1d/src/gastheory.f90:98-104).
template <typename T>
__global__ __launch_bounds__(256) void k_macro(Consts c, int dim, const double* __restrict__ W,
                                               const T* __restrict__ h, const T* __restrict__ b,
                                               const double* __restrict__ u1,
                                               const double* __restrict__ v1,
                                               const double* __restrict__ wnode,
                                               double* __restrict__ q) {
    __shared__ double lds[4 * 4];
    const int cell = blockIdx.x;
    double s[4] = {0, 0, 0, 0};
    if (dim == 2) {
        double p[4];
        prim_from_w(c, W + 4 * (size_t)cell, p);
        for (int k = threadIdx.x; k < c.nn; k += 256) {
            const int iu = k % c.nv, jv = k / c.nv;
            const double cu = u1[iu] - p[1], cv = v1[jv] - p[2];
            const double c2 = cu * cu + cv * cv;
            const double hh = (double)h[(size_t)cell * c.ldv + k];
            const double bb = (double)b[(size_t)cell * c.ldv + k];
            const double wt = wnode[k];
            s[0] += wt * cu * c2 * hh;
            s[1] += wt * cu * bb;
            s[2] += wt * cv * c2 * hh;
            s[3] += wt * cv * bb;
        }
    } else {
        const double* w = W + 3 * (size_t)cell;
        const double U = w[1] / w[0];
        for (int k = threadIdx.x; k < c.nn; k += 256) {
            const double cu = u1[k] - U;
            const double hh = (double)h[(size_t)cell * c.ldv + k];
            const double bb = (double)b[(size_t)cell * c.ldv + k];
            s[0] += wnode[k] * cu * (cu * cu) * hh;
            s[1] += wnode[k] * cu * bb;
        }
    }
    block_sum<4, 256>(s, lds);
    if (threadIdx.x == 0) {
        if (dim == 2) {
            q[2 * (size_t)cell] = 0.5 * (s[0] + s[1]);
            q[2 * (size_t)cell + 1] = 0.5 * (s[2] + s[3]);
        } else {
            q[cell] = 0.5 * (s[0] + s[1]);
        }
    }
}

template <typename T>
__global__ void k_to_dense(const T* __restrict__ src, double* __restrict__ dst, size_t ncell,
                           int nn, int ldv) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ncell * nn) return;
    const size_t cell = idx / nn;
    const int k = (int)(idx - cell * nn);
    dst[idx] = (double)src[cell * ldv + k];
}

template <typename T>
__global__ void k_from_dense(const double* __restrict__ src, T* __restrict__ dst, size_t ncell,
                             int nn, int ldv) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ncell * ldv) return;
    const size_t cell = idx / ldv;
    const int k = (int)(idx - cell * ldv);
    dst[idx] = k < nn ? (T)src[cell * nn + k] : (T)0;
}

}  // namespace hugks
