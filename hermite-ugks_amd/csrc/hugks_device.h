// Device-side building blocks shared by the Hermite-UGKS kernels (gfx950).
//
// Layout in HBM (one handle = one case):
//   W        double[ncell][4]          conserved variables, cell c = i + nx*j
//   h, b     T[ncell][ldv]             reduced distributions, node k = iu + nv*jv,
//                                      ldv = nv*nv rounded up to 64 (pad is zero)
//   sh?, sb? T[ncell][ldv]             van Leer slopes in x and y
//   vfh,vfb  T[(nx+1)*ny][ldv]         vertical-face distribution fluxes
//   hfh,hfb  T[nx*(ny+1)][ldv]         horizontal-face distribution fluxes
//   vflux    double[(nx+1)*ny][4]      face conserved fluxes (FP64 always)
//   hflux    double[nx*(ny+1)][4]
// T is double (HUGKS_FP64) or float (mixed modes).  Velocity-space sums
// are always accumulated in double.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "hugks.h"

// Occupancy hint of a kernel (waves per SIMD); a host build of the sources
// for testing defines it away beforehand.
#ifndef HUGKS_WAVES_PER_EU
#define HUGKS_WAVES_PER_EU(n) __attribute__((amdgpu_waves_per_eu(n)))
#endif

namespace hugks {

constexpr int WAVE = 64;

// Per-launch constants (by value).
struct Consts {
    int nx, ny, nv, nn, ldv, nm;
    int bc[4];
    int tau_model, dt_rule, clamp_dt, stop_on_residual, steady_steps;
    int force_fd;     // the finite-difference force instead of the Hermite one
    double fd_du;     // its velocity step, lengthU / nv
    double dx, dy, pi, smv, K, gamma, prandtl, shk, muref, omega;
    double steady_tol, mass_tol;
    double cfl, maxtime, eps_res, umax_dt, vmax_dt, force_rate, tref, phi;
    double tw[4];
};

// Loop state kept on the device so a sequence of steps needs no host
// round trip.
struct DevState {
    double dt;
    double simtime;
    double res[4];
    double steady[3];
    double t_stop;    // stop time of the current host call
    double mass0;     // reference mass of the relative mass change
    double mass_rel;  // (mass - mass0) / (|mass0| + smv) after the last step
    int iter;
    int stopped;      // set once an exit condition held; later steps no-op
    int stop_reason;
    int flags;        // bit 0: NaN/T<=0, bit 1: zero wall coefficient
    int steps;        // steps executed since the last host reset
    int steady_count; // consecutive steps the 1D steady test has held
};

// ---------------------------------------------------------------- scalars
// Boundary kind and wall temperature of side s (0..3) by selects, not by
// a run-time index into the arrays of c: a Consts copied from memory (the
// grouped kernels read each case's from a table) would otherwise be kept
// in scratch memory to be indexed.
__host__ __device__ __forceinline__ int bc_of(const Consts& c, int s) {
    return s == 0 ? c.bc[0] : s == 1 ? c.bc[1] : s == 2 ? c.bc[2] : c.bc[3];
}
__host__ __device__ __forceinline__ double tw_of(const Consts& c, int s) {
    return s == 0 ? c.tw[0] : s == 1 ? c.tw[1] : s == 2 ? c.tw[2] : c.tw[3];
}

// get_primary / get_conserved, 2d/src/gastheory2d.f90:28-46
__device__ __forceinline__ void prim_from_w(const Consts& c, const double* w, double* p) {
    p[0] = w[0];
    p[1] = w[1] / w[0];
    p[2] = w[2] / w[0];
    p[3] = 0.5 * w[0] / (c.gamma - 1.0) / (w[3] - 0.5 * (w[1] * w[1] + w[2] * w[2]) / w[0]);
}
__device__ __forceinline__ void w_from_prim(const Consts& c, const double* p, double* w) {
    w[0] = p[0];
    w[1] = p[0] * p[1];
    w[2] = p[0] * p[2];
    w[3] = 0.5 * p[0] / p[3] / (c.gamma - 1.0) + 0.5 * p[0] * (p[1] * p[1] + p[2] * p[2]);
}
// get_tau, 2d/src/gastheory2d.f90:87-97
__device__ __forceinline__ double tau_of(const Consts& c, double rho, double lam) {
    if (c.tau_model == 0) return c.muref / rho;
    return c.muref * 2.0 * pow(lam, 1.0 - c.omega) / rho;
}
// micro_slope, 2d/src/flux.f90:711-722
__device__ __forceinline__ void micro_slope(const Consts& c, const double* p, const double* sw,
                                            double* a) {
    const double K = c.K;
    a[3] = 4.0 * p[3] * p[3] / (K + 2.0) / p[0] *
           (2.0 * sw[3] - 2.0 * p[1] * sw[1] - 2.0 * p[2] * sw[2] +
            sw[0] * (p[1] * p[1] + p[2] * p[2] - 0.5 * (K + 2.0) / p[3]));
    a[2] = 2.0 * p[3] / p[0] * (sw[2] - p[2] * sw[0]) - p[2] * a[3];
    a[1] = 2.0 * p[3] / p[0] * (sw[1] - p[1] * sw[0]) - p[1] * a[3];
    a[0] = sw[0] / p[0] - p[1] * a[1] - p[2] * a[2] -
           0.5 * (p[1] * p[1] + p[2] * p[2] + 0.5 * (K + 2.0) / p[3]) * a[3];
}

// Half-range and full-range Maxwellian moments, 2d/src/flux.f90:730-762
struct Moments {
    double Mu[7], MuL[7], MuR[7], Mv[7], Mxi[3];
};
__device__ __forceinline__ void calc_moment(const Consts& c, const double* p, Moments& m) {
    const double U = p[1], V = p[2], lam = p[3];
    const double e = 0.5 * exp(-lam * U * U) / sqrt(c.pi * lam);
    m.MuL[0] = 0.5 * erfc(-sqrt(lam) * U);
    m.MuL[1] = U * m.MuL[0] + e;
    m.MuR[0] = 0.5 * erfc(sqrt(lam) * U);
    m.MuR[1] = U * m.MuR[0] - e;
#pragma unroll
    for (int i = 2; i <= 6; ++i) {
        m.MuL[i] = U * m.MuL[i - 1] + 0.5 * (i - 1) * m.MuL[i - 2] / lam;
        m.MuR[i] = U * m.MuR[i - 1] + 0.5 * (i - 1) * m.MuR[i - 2] / lam;
    }
#pragma unroll
    for (int i = 0; i <= 6; ++i) m.Mu[i] = m.MuL[i] + m.MuR[i];
    m.Mv[0] = 1.0;
    m.Mv[1] = V;
#pragma unroll
    for (int i = 2; i <= 6; ++i) m.Mv[i] = V * m.Mv[i - 1] + 0.5 * (i - 1) * m.Mv[i - 2] / lam;
    m.Mxi[0] = 1.0;
    m.Mxi[1] = 0.5 * c.K / lam;
    m.Mxi[2] = (c.K * c.K + 2.0 * c.K) / (4.0 * lam * lam);
}
// moment_uv, 2d/src/flux.f90:772-782
__device__ __forceinline__ void moment_uv(const double* Mu, const double* Mv, const double* Mxi,
                                          int al, int be, int th, double* o) {
    o[0] = Mu[al] * Mv[be] * Mxi[th / 2];
    o[1] = Mu[al + 1] * Mv[be] * Mxi[th / 2];
    o[2] = Mu[al] * Mv[be + 1] * Mxi[th / 2];
    o[3] = 0.5 * (Mu[al + 2] * Mv[be] * Mxi[th / 2] + Mu[al] * Mv[be + 2] * Mxi[th / 2] +
                  Mu[al] * Mv[be] * Mxi[(th + 2) / 2]);
}
// moment_au, 2d/src/flux.f90:792-804
__device__ __forceinline__ void moment_au(const double* a, const double* Mu, const double* Mv,
                                          const double* Mxi, int al, int be, double* o) {
    double t1[4], t2[4], t3[4], t4[4], t5[4], t6[4];
    moment_uv(Mu, Mv, Mxi, al, be, 0, t1);
    moment_uv(Mu, Mv, Mxi, al + 1, be, 0, t2);
    moment_uv(Mu, Mv, Mxi, al, be + 1, 0, t3);
    moment_uv(Mu, Mv, Mxi, al + 2, be, 0, t4);
    moment_uv(Mu, Mv, Mxi, al, be + 2, 0, t5);
    moment_uv(Mu, Mv, Mxi, al, be, 2, t6);
#pragma unroll
    for (int q = 0; q < 4; ++q)
        o[q] = a[0] * t1[q] + a[1] * t2[q] + a[2] * t3[q] + 0.5 * a[3] * t4[q] +
               0.5 * a[3] * t5[q] + 0.5 * a[3] * t6[q];
}

// Fortran sign(1d0, x): -1 for -0.0 as well (copysign semantics).
__device__ __forceinline__ int upwind_delta(double x) { return copysign(1.0, x) > 0.0 ? 1 : 0; }
__device__ __forceinline__ int upwind_delta(float x) { return copysignf(1.0f, x) > 0.0f ? 1 : 0; }

// ----------------------------------------------------------- reductions
// Sum of NV doubles over the block; every thread receives the totals.
// Fixed order (wave butterfly, then waves in index order), so results
// are bitwise reproducible run to run.  lds must hold NV*(NT/64) doubles.
template <int NV, int NT>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double x = v[i];
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
        v[i] = x;
    }
    if (NT > WAVE) {
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < NV; ++i) lds[wid * NV + i] = v[i];
        }
        __syncthreads();
        // waves in index order for every value; for large blocks one
        // wave's partials in registers at a time (the same additions): with
        // the loop over waves unrolled the compiler keeps all NV x NT/64
        // partials live (a 1024-thread block with NV = 6 spilled to scratch)
        if constexpr (NV * (NT / WAVE) <= 32) {
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                double s = lds[i];
#pragma unroll
                for (int w = 1; w < NT / WAVE; ++w) s += lds[w * NV + i];
                v[i] = s;
            }
        } else {
#pragma unroll
            for (int i = 0; i < NV; ++i) v[i] = lds[i];
#pragma unroll 1
            for (int w = 1; w < NT / WAVE; ++w) {
#pragma unroll
                for (int i = 0; i < NV; ++i) v[i] += lds[w * NV + i];
            }
        }
        __syncthreads();
    } else {
        // one wave: broadcast lane 0's value so all lanes agree bitwise
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] = __shfl(v[i], 0, WAVE);
    }
}

// Stage of wave_reduce_scatter: the M values of each lane are halved to
// ceil(M/2) by exchanging with the lane O apart; the lower lane of the pair
// keeps the first half, the upper lane the second (zero-padded when M is
// odd).  On the way back pos becomes the value index the lane ends up
// holding, or -1 where its path ran through padding.
// The exchanges across 32 and 16 lanes use the gfx950 lane swaps:
// v_permlane32_swap(a, b) leaves a = [a lanes 0-31, b lanes 0-31] and
// b = [a lanes 32-63, b lanes 32-63], v_permlane16_swap does the same per
// pair of 16-lane rows, so a + b is the stage's sum on every lane (a
// partner's lo value on the lower lane of a pair, hi on the upper) with no
// selects and no LDS round trip; the same additions as the shuffle form.
template <int O>
__device__ __forceinline__ double swap_sum(double lo_v, double hi_v) {
    static_assert(O == 32 || O == 16, "lane swaps exist for 32 and 16");
    const unsigned long long a = __builtin_bit_cast(unsigned long long, lo_v);
    const unsigned long long b = __builtin_bit_cast(unsigned long long, hi_v);
    const unsigned a0 = (unsigned)a, a1 = (unsigned)(a >> 32);
    const unsigned b0 = (unsigned)b, b1 = (unsigned)(b >> 32);
    unsigned r00, r01, r10, r11;
    if constexpr (O == 32) {
        const auto w0 = __builtin_amdgcn_permlane32_swap(a0, b0, false, false);
        const auto w1 = __builtin_amdgcn_permlane32_swap(a1, b1, false, false);
        r00 = w0[0]; r01 = w0[1]; r10 = w1[0]; r11 = w1[1];
    } else {
        const auto w0 = __builtin_amdgcn_permlane16_swap(a0, b0, false, false);
        const auto w1 = __builtin_amdgcn_permlane16_swap(a1, b1, false, false);
        r00 = w0[0]; r01 = w0[1]; r10 = w1[0]; r11 = w1[1];
    }
    const double x = __builtin_bit_cast(double, (unsigned long long)r00 | ((unsigned long long)r10 << 32));
    const double y = __builtin_bit_cast(double, (unsigned long long)r01 | ((unsigned long long)r11 << 32));
    return x + y;
}

template <int M, int O>
__device__ __forceinline__ double wave_rs_stage(double* x, int lane, int& pos) {
    constexpr int M2 = (M + 1) / 2;
    const bool hi = (lane & O) != 0;
#pragma unroll
    for (int t = 0; t < M2; ++t) {
        const double lo_v = x[t];
        const double hi_v = (t + M2 < M) ? x[t + M2] : 0.0;
        if constexpr (O == 32 || O == 16) {
            x[t] = swap_sum<O>(lo_v, hi_v);
        } else {
            const double keep = hi ? hi_v : lo_v;
            const double send = hi ? lo_v : hi_v;
            x[t] = keep + __shfl_xor(send, O, WAVE);
        }
    }
    double r;
    if constexpr (O > 1) {
        r = wave_rs_stage<M2, O / 2>(x, lane, pos);
    } else {
        r = x[0];
        pos = 0;
    }
    if (pos >= 0) {
        pos += hi ? M2 : 0;
        if (pos >= M) pos = -1;
    }
    return r;
}

// Sum of N doubles over the wavefront, scattered over the lanes: returns
// the total of value q on exactly one lane, with q in *q, and q = -1 on
// the other lanes.  About N shuffles instead of the 6N of a butterfly per
// value; fixed order, so bitwise reproducible.  v is clobbered.
template <int N>
__device__ __forceinline__ double wave_reduce_scatter(double (&v)[N], int& q) {
    static_assert(N >= 1 && N <= WAVE, "one value per lane at most");
    q = 0;
    return wave_rs_stage<N, WAVE / 2>(v, (int)(threadIdx.x & (WAVE - 1)), q);
}

// ------------------------------------------------- XCD-aware cell order
// The workgroups of one launch are dealt to the eight XCDs round-robin
// (workgroup b to XCD b % 8), and each XCD has its own L2.  xcd_block
// renumbers workgroups so that every XCD runs one contiguous range of the
// logical order (a bijection on [0, nb)); tiled_cell walks the cells in
// strips of TILE_W columns, row by row inside a strip, so the neighbour
// cells a kernel reads (i +- 1 and j +- 1) were touched a few dozen cells
// earlier on the same XCD and are still in its L2.
constexpr int NXCD = 8;
constexpr int TILE_W = 32;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
    const unsigned q = nb / NXCD, r = nb % NXCD, x = b % NXCD;
    return x * q + (x < r ? x : r) + b / NXCD;
}
__device__ __forceinline__ int tiled_cell(int L, int nx, int ny) {
    const int s = L / (TILE_W * ny);
    const int rem = L - s * TILE_W * ny;
    const int ws = nx - s * TILE_W < TILE_W ? nx - s * TILE_W : TILE_W;
    const int jj = rem / ws, ii = rem - jj * ws;
    return s * TILE_W + ii + nx * jj;
}
// The inverse of tiled_cell: the walk position of cell (i, j) of an nx x ny grid.
__device__ __forceinline__ int tiled_index(int i, int j, int nx, int ny) {
    const int s = i / TILE_W, ii = i - s * TILE_W;
    const int ws = nx - s * TILE_W < TILE_W ? nx - s * TILE_W : TILE_W;
    return s * TILE_W * ny + j * ws + ii;
}

// ----------------------------------------------------------- slope stencil
// Cells of the slope of `cell` along dir (0: x, 1: y), as interp_inner /
// interp_boundary choose them (2d/src/solver.f90:64-159): periodic
// directions wrap (the ghost copy of 2d/src/boundary.f90:16-53); along a
// wall the two boundary cells take one-sided differences (inner = false).
struct SlopeNb {
    int cl, cr;
    bool inner;
};
__host__ __device__ __forceinline__ SlopeNb slope_nb(const Consts& c, int cell, int dir) {
    const int i = cell % c.nx, j = cell / c.nx;
    const bool per = bc_of(c, 2 * dir) == HUGKS_BC_PERIODIC;
    const int n = dir == 0 ? c.nx : c.ny;
    const int p = dir == 0 ? i : j;
    const int step = dir == 0 ? 1 : c.nx;
    SlopeNb r;
    r.inner = true;
    if (per) {
        const int pl = (p - 1 + n) % n, pr = (p + 1) % n;
        r.cl = dir == 0 ? pl + c.nx * j : i + c.nx * pl;
        r.cr = dir == 0 ? pr + c.nx * j : i + c.nx * pr;
    } else if (p == 0) {
        r.cl = cell;
        r.cr = cell + step;
        r.inner = false;
    } else if (p == n - 1) {
        r.cl = cell - step;
        r.cr = cell;
        r.inner = false;
    } else {
        r.cl = cell - step;
        r.cr = cell + step;
    }
    return r;
}

// Slope at a node from its stencil values: van Leer (solver.f90:146-151)
// for an inner cell, the one-sided difference at a wall.  FAST divides by
// the hardware reciprocal (FP32: within 1 ulp) instead of an IEEE division.
__device__ __forceinline__ float fast_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double fast_div(double a, double b) { return a / b; }
template <typename A, bool FAST = false>
__device__ __forceinline__ A vl_slope(A fl, A fn, A fr, bool inner, A inv_den, A smv) {
    if (inner) {
        const A sL = (fn - fl) * inv_den, sR = (fr - fn) * inv_den;
        const A num = (copysign((A)1, sR) + copysign((A)1, sL)) * fabs(sR) * fabs(sL);
        const A den = fabs(sR) + fabs(sL) + smv;
        return FAST ? fast_div(num, den) : num / den;
    }
    return (fr - fl) * inv_den;
}

// Orders one wavefront's LDS accesses across its lanes: what any lane wrote
// before is visible to every lane after (the compiler keeps its LDS
// accesses on their side; the hardware completes one wavefront's LDS
// operations in order).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------ buffer loads
// Raw buffer loads: a scalar resource (base, byte size) and a 32-bit byte
// offset per lane, so a load needs no per-lane 64-bit address arithmetic;
// the hardware returns zero for offsets past the size.  Word 3 of the
// resource is the gfx9-family raw-buffer format (32-bit data).
constexpr int BUF_RSRC_WORD3 = 0x00020000;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                             BUF_RSRC_WORD3);
}
template <typename T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}

// Raw buffer store with the non-temporal cache policy (the data is read
// back only by a later kernel); an offset at or past the resource's size
// is dropped by the hardware range check.
template <typename T>
__device__ __forceinline__ void buf_st_nt(T v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    constexpr int kNonTemporal = 2;
    if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)off, 0,
                                              kNonTemporal);
    else
        __builtin_amdgcn_raw_buffer_store_b64(
            __builtin_bit_cast(unsigned __attribute__((vector_size(8))), v), r, (int)off, 0,
            kNonTemporal);
}

template <typename T> struct Arith;           // per-node arithmetic type
template <> struct Arith<double> { using A = double; };
template <> struct Arith<float> { using A = float; };

__device__ __forceinline__ double fexp(double x) { return exp(x); }
__device__ __forceinline__ float fexp(float x) { return expf(x); }

}  // namespace hugks
