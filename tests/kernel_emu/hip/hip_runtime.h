// CPU emulation of the subset of the HIP runtime and device intrinsics the
// library uses.  TEST INFRASTRUCTURE ONLY: it lets the unmodified kernel
// and launcher sources (hermite-ugks_amd/csrc) be compiled by g++ and run
// on the host, one workgroup at a time with one fiber per work-item,
// under AddressSanitizer.  __syncthreads is a barrier over the workgroup,
// shuffles exchange through a workgroup buffer, __shared__ arrays are
// static (one workgroup runs at a time), dynamic LDS is a heap block of the
// exact launch size, and fresh allocations are filled with NaN bytes.
#pragma once
#include <math.h>

#include <ucontext.h>

#include <tuple>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <functional>
#include <vector>

#define __global__
#define __device__
#define __constant__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define HUGKS_WAVES_PER_EU(n)
#define __shared__ static
#define HIP_DYNAMIC_SHARED(type, var) type* var = reinterpret_cast<type*>(::emu::dyn_smem);
// wave-uniform values only: the emulation has no scalar register file
#define __builtin_amdgcn_readfirstlane(x) (x)
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }
// cache hints have no meaning on the CPU
#define __builtin_nontemporal_store(v, p) (*(p) = (v))
// raw buffer loads: base and byte size; offsets past the size read zero
struct __amdgpu_buffer_rsrc_t {
    const char* base;
    uint32_t bytes;
};
inline __amdgpu_buffer_rsrc_t __builtin_amdgcn_make_buffer_rsrc(void* p, short, int bytes, int) {
    return {static_cast<const char*>(p), (uint32_t)bytes};
}
inline uint32_t __builtin_amdgcn_raw_buffer_load_b32(__amdgpu_buffer_rsrc_t r, int off, int, int) {
    uint32_t v = 0;
    if ((uint64_t)(uint32_t)off + 4 <= r.bytes) memcpy(&v, r.base + (uint32_t)off, 4);
    return v;
}
inline uint64_t __builtin_amdgcn_raw_buffer_load_b64(__amdgpu_buffer_rsrc_t r, int off, int, int) {
    uint64_t v = 0;
    if ((uint64_t)(uint32_t)off + 8 <= r.bytes) memcpy(&v, r.base + (uint32_t)off, 8);
    return v;
}

inline void __builtin_amdgcn_raw_buffer_store_b32(uint32_t v, __amdgpu_buffer_rsrc_t r, int off, int,
                                                  int) {
    if ((uint64_t)(uint32_t)off + 4 <= r.bytes) memcpy(const_cast<char*>(r.base) + (uint32_t)off, &v, 4);
}
inline void __builtin_amdgcn_raw_buffer_store_b64(unsigned __attribute__((vector_size(8))) v,
                                                  __amdgpu_buffer_rsrc_t r, int off, int, int) {
    if ((uint64_t)(uint32_t)off + 8 <= r.bytes) memcpy(const_cast<char*>(r.base) + (uint32_t)off, &v, 8);
}
// v_perm_b32: byte i of the result is byte sel[i] of {b (bytes 0-3), a (4-7)};
// selector 0x0c gives 0x00, larger ones 0xff (the cases the kernels use: 0-7)
inline uint32_t __builtin_amdgcn_perm(uint32_t a, uint32_t b, uint32_t sel) {
    const uint64_t ab = ((uint64_t)a << 32) | b;
    uint32_t r = 0;
    for (int i = 0; i < 4; ++i) {
        const uint32_t s = (sel >> (8 * i)) & 0xffu;
        const uint32_t byte = s < 8 ? (uint32_t)((ab >> (8 * s)) & 0xffu) : s == 0x0c ? 0u : 0xffu;
        r |= byte << (8 * i);
    }
    return r;
}
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (reinterpret_cast<const void*>(&(x)))

struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace emu {
// Work-items of one workgroup are cooperative fibers on one OS thread; a
// barrier switches back to the scheduler, which resumes every live fiber
// in turn (a work-item that returned no longer takes part, as an exited
// wave no longer takes part in s_barrier).
struct Fiber {
    ucontext_t ctx;
    char* stack;
    size_t size;
    bool done;
    dim3 t;
};
extern dim3 tid, bid, bdim, gdim;
extern Fiber* fibers_base;
extern unsigned nthreads;
// Counting barrier over the live fibers [first, first + count): a fiber
// that arrived waits (yielding) until every live fiber of the range has
// arrived; fibers that returned no longer count.
struct Bar {
    unsigned arrived;
    unsigned gen;
};
extern Bar block_bar;
extern Bar wave_bar[16];
void barrier(Bar& b, unsigned first, unsigned count);
void wave_barrier();
extern unsigned char* dyn_smem;
extern unsigned char xbuf[1024 * 8];
extern long launches;
extern ucontext_t sched_ctx;
extern Fiber* cur;
extern void (*body)(void*);
extern void* body_arg;
void yield_to_sched();
void fiber_entry();
void run_block(Fiber* fibers, unsigned nt);

template <class K, class... A>
struct Call {
    K kernel;
    std::tuple<A...> args;
    static void run(void* p) {
        Call* c = (Call*)p;
        std::apply(c->kernel, c->args);
    }
};

// Stream capture: while a capture is open, launches are recorded (by
// value) instead of run; a graph replays them in recording order, which is
// one valid order of the captured dependencies.
extern bool capturing;
extern std::vector<std::function<void()>>* capture_list;

template <class K, class... A>
void launch_now(K kernel, dim3 g, dim3 b, size_t shmem, A... args);

template <class K, class... A>
void launch(K kernel, dim3 g, dim3 b, size_t shmem, A... args) {
    if (capturing) {
        capture_list->push_back([=]() { launch_now(kernel, g, b, shmem, args...); });
        return;
    }
    launch_now(kernel, g, b, shmem, args...);
}

template <class K, class... A>
void launch_now(K kernel, dim3 g, dim3 b, size_t shmem, A... args) {
    const unsigned nt = b.x * b.y * b.z;
    const unsigned long nblk = (unsigned long)g.x * g.y * g.z;
    ++launches;
    Call<K, A...> call{kernel, std::make_tuple(args...)};
    body = &Call<K, A...>::run;
    body_arg = &call;
    unsigned char* d = (unsigned char*)malloc(shmem ? shmem : 1);
    dyn_smem = d;
    bdim = b;
    gdim = g;
    std::vector<Fiber> fibers(nt);
    const size_t ss = 256 * 1024;
    for (auto& f : fibers) {
        f.size = ss;
        f.stack = (char*)malloc(ss);
    }
    for (unsigned long blk = 0; blk < nblk; ++blk) {
        memset(d, 0xff, shmem ? shmem : 1);
        bid = dim3((unsigned)(blk % g.x), (unsigned)((blk / g.x) % g.y),
                   (unsigned)(blk / ((unsigned long)g.x * g.y)));
        for (unsigned t = 0; t < nt; ++t)
            fibers[t].t = dim3(t % b.x, (t / b.x) % b.y, t / (b.x * b.y));
        fibers_base = fibers.data();
        nthreads = nt;
        run_block(fibers.data(), nt);
    }
    for (auto& f : fibers) free(f.stack);
    free(d);
}
}  // namespace emu

#define threadIdx (::emu::tid)
#define blockIdx (::emu::bid)
#define blockDim (::emu::bdim)
#define gridDim (::emu::gdim)

inline void __syncthreads() { emu::barrier(emu::block_bar, 0, emu::nthreads); }
// wavefront-scope ordering: the lanes of a wavefront run as separate fibers
// here, so the wave barrier is a real one and the fences are no-ops
#define __builtin_amdgcn_fence(order, scope) ((void)0)
// the lanes' stores are in host memory before any later load: a wave
// barrier orders them across the lanes (fibers) of the wavefront
inline void __threadfence() { emu::wave_barrier(); }
inline void __builtin_amdgcn_wave_barrier() { emu::wave_barrier(); }

// Wave-level operations exchange through xbuf between two wave barriers
// (every lane of the wavefront takes part, as on the hardware).
template <class T>
inline T __shfl_xor(T v, int mask, int width = 64) {
    static_assert(sizeof(T) <= 8, "shuffle of at most 8 bytes");
    const unsigned t = threadIdx.x;
    std::memcpy(emu::xbuf + 8 * t, &v, sizeof(T));
    emu::wave_barrier();
    const unsigned src = (t & ~(unsigned)(width - 1)) | ((t & (unsigned)(width - 1)) ^ (unsigned)mask);
    T r;
    std::memcpy(&r, emu::xbuf + 8 * src, sizeof(T));
    emu::wave_barrier();
    return r;
}
template <class T>
inline T __shfl(T v, int lane, int width = 64) {
    const unsigned t = threadIdx.x;
    std::memcpy(emu::xbuf + 8 * t, &v, sizeof(T));
    emu::wave_barrier();
    const unsigned src = (t & ~(unsigned)(width - 1)) | ((unsigned)lane & (unsigned)(width - 1));
    T r;
    std::memcpy(&r, emu::xbuf + 8 * src, sizeof(T));
    emu::wave_barrier();
    return r;
}

// Bit l set where lane l of the wavefront passes a true predicate.
inline unsigned long long __ballot(int pred) {
    const unsigned t = threadIdx.x, w0 = t & ~63u;
    emu::xbuf[8 * t] = pred ? 1 : 0;
    emu::wave_barrier();
    unsigned long long m = 0;
    for (unsigned l = 0; l < 64; ++l)
        if (emu::xbuf[8 * (w0 + l)]) m |= 1ull << l;
    emu::wave_barrier();
    return m;
}

// v_permlane32_swap / v_permlane16_swap (gfx950): returns {new vdst, new
// src0}.  32: lanes 32-63 of vdst trade places with lanes 0-31 of src0;
// 16: the odd 16-lane rows of vdst with the even rows of src0.
struct PlSwap {
    unsigned v[2];
    unsigned operator[](int i) const { return v[i]; }
};
namespace emu {
inline PlSwap permlane_swap(unsigned vd, unsigned s0, unsigned w) {
    const unsigned t = threadIdx.x, l = t & 63u;
    std::memcpy(emu::xbuf + 8 * t, &vd, 4);
    std::memcpy(emu::xbuf + 8 * t + 4, &s0, 4);
    emu::wave_barrier();
    const bool upper = (l & w) != 0;
    PlSwap r;
    if (upper) {
        std::memcpy(&r.v[0], emu::xbuf + 8 * (t - w) + 4, 4);   // src0 of the lower row
        r.v[1] = s0;
    } else {
        r.v[0] = vd;
        std::memcpy(&r.v[1], emu::xbuf + 8 * (t + w), 4);       // vdst of the upper row
    }
    emu::wave_barrier();
    return r;
}
}  // namespace emu
inline PlSwap __builtin_amdgcn_permlane32_swap(unsigned vd, unsigned s0, bool, bool) {
    return emu::permlane_swap(vd, s0, 32);
}
inline PlSwap __builtin_amdgcn_permlane16_swap(unsigned vd, unsigned s0, bool, bool) {
    return emu::permlane_swap(vd, s0, 16);
}

// 16x16x4 matrix tiles, lane layouts as on gfx950: A(i,k) from lane
// 16k+i, B(k,j) from lane 16k+j; result element v of lane l is
// C(4*(l/16)+v, l%16) for f32 and C(l/16+4*v, l%16) for f64.
namespace emu {
extern double mfa[1024], mfb[1024];
template <class V, class S>
inline V mfma16x16x4(S a, S b, V c, bool f64) {
    const unsigned t = threadIdx.x, w0 = t & ~63u, l = t & 63u;
    mfa[t] = (double)a;
    mfb[t] = (double)b;
    wave_barrier();
    V r = c;
    for (int v = 0; v < 4; ++v) {
        const unsigned i = f64 ? (l >> 4) + 4 * v : 4 * (l >> 4) + v, j = l & 15;
        S acc = c[v];
        for (unsigned k = 0; k < 4; ++k)
            acc = (S)std::fma((S)mfa[w0 + 16 * k + i], (S)mfb[w0 + 16 * k + j], acc);
        r[v] = acc;
    }
    wave_barrier();
    return r;
}
}  // namespace emu
typedef float emu_f4 __attribute__((vector_size(4 * sizeof(float))));
typedef double emu_d4 __attribute__((vector_size(4 * sizeof(double))));
inline emu_f4 __builtin_amdgcn_mfma_f32_16x16x4f32(float a, float b, emu_f4 c, int, int, int) {
    return emu::mfma16x16x4<emu_f4, float>(a, b, c, false);
}
inline emu_d4 __builtin_amdgcn_mfma_f64_16x16x4f64(double a, double b, emu_d4 c, int, int, int) {
    return emu::mfma16x16x4<emu_d4, double>(a, b, c, true);
}
// 16x16x32 bf16 tiles: lane l holds A(l%16, 8(l/16)+s) and B(8(l/16)+s, l%16)
// in slot s; the f32 result layout as 16x16x4f32.
typedef __bf16 emu_bf8 __attribute__((ext_vector_type(8)));
namespace emu {
extern float mbfa[1024 * 8], mbfb[1024 * 8];
}
inline emu_f4 __builtin_amdgcn_mfma_f32_16x16x32_bf16(emu_bf8 a, emu_bf8 b, emu_f4 c, int, int, int) {
    const unsigned t = threadIdx.x, w0 = t & ~63u, l = t & 63u;
    for (int s = 0; s < 8; ++s) {
        emu::mbfa[8 * t + s] = (float)a[s];
        emu::mbfb[8 * t + s] = (float)b[s];
    }
    emu::wave_barrier();
    emu_f4 r = c;
    for (int v = 0; v < 4; ++v) {
        const unsigned i = 4 * (l >> 4) + v, j = l & 15;
        float acc = c[v];
        for (unsigned k = 0; k < 32; ++k) {
            const unsigned la = w0 + 16 * (k / 8) + i, lb = w0 + 16 * (k / 8) + j;
            acc = std::fma(emu::mbfa[8 * la + k % 8], emu::mbfb[8 * lb + k % 8], acc);
        }
        r[v] = acc;
    }
    emu::wave_barrier();
    return r;
}
inline int atomicOr(int* p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }

// ---- runtime API
typedef enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNoDevice = 100,
    hipErrorInvalidDevice = 101,
} hipError_t;
typedef struct emu_stream* hipStream_t;
typedef struct emu_event* hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost,
                     hipMemcpyDeviceToDevice, hipMemcpyDefault };
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 8 };
#define hipStreamNonBlocking 1
struct hipDeviceProp_t {
    char name[256];
    char gcnArchName[256];
    int multiProcessorCount;
    size_t maxSharedMemoryPerMultiProcessor;
    size_t totalGlobalMem;
    size_t sharedMemPerBlockOptin;
    size_t sharedMemPerBlock;
};
#define hipLaunchKernelGGL(kernel, g, b, sh, st, ...) ::emu::launch(kernel, g, b, sh, __VA_ARGS__)

inline const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "success" : "emu error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidDevice; }
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) {
    *p = const_cast<void*>(sym);
    return hipSuccess;
}
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
    memset(p, 0, sizeof(*p));
    strcpy(p->name, "CPU emulation");
    strcpy(p->gcnArchName, "emu");
    p->multiProcessorCount = 1;
    p->maxSharedMemoryPerMultiProcessor = 160 * 1024;
    p->totalGlobalMem = (size_t)1 << 34;
    p->sharedMemPerBlockOptin = 160 * 1024;
    p->sharedMemPerBlock = 64 * 1024;
    return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    *s = (hipStream_t)malloc(1);
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t s) { free(s); return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = (hipEvent_t)malloc(1); return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { free(e); return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
inline hipError_t hipMalloc(void** p, size_t n) {
    *p = malloc(n);
    if (!*p) return hipErrorOutOfMemory;
    memset(*p, 0xff, n);
    return hipSuccess;
}
inline hipError_t hipFree(void* p) { free(p); return hipSuccess; }
inline hipError_t hipHostMalloc(void** p, size_t n, unsigned = 0) { return hipMalloc(p, n); }
inline hipError_t hipHostFree(void* p) { free(p); return hipSuccess; }
inline hipError_t hipMallocAsync(void** p, size_t n, hipStream_t) { return hipMalloc(p, n); }
inline hipError_t hipFreeAsync(void* p, hipStream_t) { free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
    memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { memset(d, v, n); return hipSuccess; }
inline hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return hipSuccess; }
#define hipEventDisableTiming 2
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal,
                            hipStreamCaptureModeRelaxed };
struct emu_graph { std::vector<std::function<void()>> ops; };
typedef emu_graph* hipGraph_t;
typedef emu_graph* hipGraphExec_t;
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
    emu::capture_list = new std::vector<std::function<void()>>();
    emu::capturing = true;
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
    emu::capturing = false;
    *g = new emu_graph{std::move(*emu::capture_list)};
    delete emu::capture_list;
    emu::capture_list = nullptr;
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, void*, void*, size_t) {
    *x = new emu_graph{g->ops};
    return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t) {
    for (auto& op : x->ops) op();
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t x) { delete x; return hipSuccess; }
