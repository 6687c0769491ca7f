// State and fiber scheduler of the CPU emulation of the HIP runtime (see
// hip/hip_runtime.h).  TEST INFRASTRUCTURE ONLY.
#include <hip/hip_runtime.h>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define EMU_ASAN 1
#endif

namespace emu {
dim3 tid, bid, bdim, gdim;
Fiber* fibers_base = nullptr;
unsigned nthreads = 0;
Bar block_bar = {0, 0};
Bar wave_bar[16] = {};
double mfa[1024], mfb[1024];
float mbfa[1024 * 8], mbfb[1024 * 8];
bool capturing = false;
std::vector<std::function<void()>>* capture_list = nullptr;
unsigned char* dyn_smem = nullptr;
unsigned char xbuf[1024 * 8];
long launches = 0;
ucontext_t sched_ctx;
Fiber* cur = nullptr;
void (*body)(void*) = nullptr;
void* body_arg = nullptr;
static void* sched_fake_stack = nullptr;

static void switch_to(Fiber* f) {
#ifdef EMU_ASAN
    void* fs = nullptr;
    __sanitizer_start_switch_fiber(&fs, f->stack, f->size);
#endif
    cur = f;
    tid = f->t;
    swapcontext(&sched_ctx, &f->ctx);
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(fs, nullptr, nullptr);
#endif
}

void yield_to_sched() {
    Fiber* f = cur;
#ifdef EMU_ASAN
    void* fs = nullptr;
    __sanitizer_start_switch_fiber(&fs, nullptr, 0);
#endif
    swapcontext(&f->ctx, &sched_ctx);
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(fs, nullptr, nullptr);
#endif
}

static unsigned live_in(unsigned first, unsigned count) {
    unsigned n = 0;
    for (unsigned t = first; t < first + count && t < nthreads; ++t) n += !fibers_base[t].done;
    return n;
}

void barrier(Bar& b, unsigned first, unsigned count) {
    const unsigned g = b.gen;
    b.arrived += 1;
    for (;;) {
        if (b.gen != g) return;
        if (b.arrived >= live_in(first, count)) {
            b.arrived = 0;
            b.gen += 1;
            return;
        }
        yield_to_sched();
    }
}

void wave_barrier() {
    const unsigned w = tid.x / 64;
    barrier(wave_bar[w], 64 * w, 64);
}

void fiber_entry() {
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(nullptr, nullptr, nullptr);
#endif
    body(body_arg);
    cur->done = true;
#ifdef EMU_ASAN
    __sanitizer_start_switch_fiber(nullptr, nullptr, 0);
#endif
    setcontext(&sched_ctx);
}

void run_block(Fiber* fibers, unsigned nt) {
    block_bar = {0, 0};
    for (auto& wb : wave_bar) wb = {0, 0};
    for (unsigned t = 0; t < nt; ++t) {
        Fiber& f = fibers[t];
        f.done = false;
        getcontext(&f.ctx);
        f.ctx.uc_stack.ss_sp = f.stack;
        f.ctx.uc_stack.ss_size = f.size;
        f.ctx.uc_link = nullptr;
        makecontext(&f.ctx, fiber_entry, 0);
    }
    (void)sched_fake_stack;
    bool live = true;
    while (live) {
        live = false;
        for (unsigned t = 0; t < nt; ++t) {
            if (fibers[t].done) continue;
            switch_to(&fibers[t]);
            if (!fibers[t].done) live = true;
        }
    }
}
}  // namespace emu
