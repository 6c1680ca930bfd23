// Runtime of the CPU emulation declared in include/hip/hip_runtime.h (tests
// only). Each block runs its threads as fibers on the launching OS thread;
// barriers and 64-lane wave collectives are released once every live thread
// (every live lane of the wave) has arrived, and a state in which nothing can
// run or be released is reported as a deadlock and aborts.
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define EMU_ASAN 1
#else
#define EMU_ASAN 0
#endif

struct ihipStream_t {
    bool capturing = false;
    ihipGraph* graph = nullptr;
};
struct ihipEvent_t {
    std::chrono::steady_clock::time_point t;
};
struct ihipGraph {
    std::vector<std::function<void()>> nodes;
};
struct hipGraphExec {
    std::vector<std::function<void()>> nodes;
};

// Minimal x86-64 context switch: pushes the callee-saved registers, MXCSR and
// the x87 control word on the current stack, stores the stack pointer in
// *from and continues on `to` (a stack saved the same way, or a fresh one
// prepared by make_fiber_stack). No signal-mask system call, unlike
// swapcontext.
extern "C" void rpd_emu_switch(void** from, void* to);
asm(R"(
    .text
    .globl rpd_emu_switch
    .type rpd_emu_switch,@function
rpd_emu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    subq $8, %rsp
    stmxcsr (%rsp)
    fnstcw 4(%rsp)
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    ldmxcsr (%rsp)
    fldcw 4(%rsp)
    addq $8, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size rpd_emu_switch, .-rpd_emu_switch
)");

namespace rpd_emu {
namespace {

constexpr size_t kStack = 512 * 1024;
constexpr int kWave = 64;
constexpr size_t kMaxDynLds = 160 * 1024;  // LDS of one CU on gfx950

enum FiberState { RUNNABLE, AT_BARRIER, AT_WAVE, DONE };

struct Fiber {
    void* sp = nullptr;       // saved stack pointer while switched out
    char* stack = nullptr;
    FiberState state = RUNNABLE;
    unsigned tid = 0;
};

struct Block {
    dim3 grid, dim;
    uint3 idx{0, 0, 0};
    std::vector<Fiber> fibers;
    void* sched_sp = nullptr;
    int cur = -1;
    void* dyn = nullptr;
    std::vector<uint64_t> slot;     // per thread: wave-collective input
    std::vector<uint64_t> result;   // per thread: wave-collective output
    std::vector<int> mode, arg;
    std::vector<void*> site;        // per thread: return address of the collective's caller
    std::vector<long> count;        // per thread: wave collectives completed
    const std::function<void()>* body = nullptr;
    const void* sched_stack_bottom = nullptr;
    size_t sched_stack_size = 0;
};

thread_local Block* g_block = nullptr;
thread_local hipError_t g_last = hipSuccess;

void fail_launch(const char* what) {
    std::fprintf(stderr, "[cpu-emu] %s\n", what);
    std::fflush(stderr);
    std::abort();
}

Fiber& me() { return g_block->fibers[size_t(g_block->cur)]; }

// fiber -> scheduler
void yield_to_scheduler() {
    Block* b = g_block;
    Fiber& f = me();
#if EMU_ASAN
    void* fake = nullptr;
    __sanitizer_start_switch_fiber(&fake, b->sched_stack_bottom, b->sched_stack_size);
#endif
    rpd_emu_switch(&f.sp, b->sched_sp);
#if EMU_ASAN
    __sanitizer_finish_switch_fiber(fake, &b->sched_stack_bottom, &b->sched_stack_size);
#endif
}

void fiber_entry() {
    Block* b = g_block;
#if EMU_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &b->sched_stack_bottom, &b->sched_stack_size);
#endif
    (*b->body)();
    me().state = DONE;
#if EMU_ASAN
    __sanitizer_start_switch_fiber(nullptr, b->sched_stack_bottom, b->sched_stack_size);
#endif
    void* dead = nullptr;
    rpd_emu_switch(&dead, b->sched_sp);  // never resumed
    __builtin_unreachable();
}

// A fresh stack laid out like one switched out by rpd_emu_switch: control
// words, six zeroed callee-saved registers, then fiber_entry as the return
// address, placed so that fiber_entry starts with the stack alignment of a
// called function.
void* make_fiber_stack(char* base, size_t size) {
    uintptr_t top = (reinterpret_cast<uintptr_t>(base) + size) & ~uintptr_t(15);
    uint64_t* sp = reinterpret_cast<uint64_t*>(top);
    *--sp = 0;                                              // padding: entry sees rsp = 8 (mod 16)
    *--sp = reinterpret_cast<uint64_t>(&fiber_entry);       // return address
    for (int r = 0; r < 6; ++r) *--sp = 0;                  // rbp rbx r12 r13 r14 r15
    uint32_t ctl[2];
    asm volatile("stmxcsr %0" : "=m"(ctl[0]));
    asm volatile("fnstcw %0" : "=m"(ctl[1]));
    --sp;
    std::memcpy(sp, ctl, 8);
    return sp;
}

// scheduler -> fiber k
void resume(Block& b, int k) {
    b.cur = k;
    Fiber& f = b.fibers[size_t(k)];
#if EMU_ASAN
    void* fake = nullptr;
    __sanitizer_start_switch_fiber(&fake, f.stack, kStack);
#endif
    rpd_emu_switch(&b.sched_sp, f.sp);
#if EMU_ASAN
    __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
    b.cur = -1;
}

// the collective of wave w, once all its live lanes have deposited their input
void complete_wave(Block& b, int w, int n) {
    const int lo = w * kWave, hi = std::min(n, lo + kWave);
    uint64_t ballot = 0;
    int md = -1;
    void* site0 = nullptr;
    for (int t = lo; t < hi; ++t) {
        const Fiber& f = b.fibers[size_t(t)];
        if (f.state != AT_WAVE) continue;
        if (md < 0) { md = b.mode[size_t(t)]; site0 = b.site[size_t(t)]; }
        else if (md != b.mode[size_t(t)] || site0 != b.site[size_t(t)]) {
            for (int u = lo; u < hi; ++u)
                if (b.fibers[size_t(u)].state == AT_WAVE)
                {
                    Dl_info di{};
                    dladdr(b.site[size_t(u)], &di);
                    std::fprintf(stderr, "[cpu-emu] lane %d (collective #%ld): operation %d called from %s+0x%lx\n",
                                 u - lo, b.count[size_t(u)], b.mode[size_t(u)], di.dli_fname ? di.dli_fname : "?",
                                 (unsigned long)((char*)b.site[size_t(u)] - (char*)di.dli_fbase));
                }
            fail_launch("divergent wave collectives (lanes of one wave in different collectives)");
        }
        if (b.slot[size_t(t)]) ballot |= 1ull << (t - lo);
    }
    for (int t = lo; t < hi; ++t) {
        const Fiber& f = b.fibers[size_t(t)];
        if (f.state != AT_WAVE) continue;
        const int lane = t - lo, a = b.arg[size_t(t)];
        if (md == 0) {
            b.result[size_t(t)] = ballot;
        } else {
            const int src = md == 1 ? a : (lane ^ a);
            const int st = lo + ((src % kWave) + kWave) % kWave;
            // a source lane outside the block or already finished reads as its own value
            b.result[size_t(t)] = (st < hi && b.fibers[size_t(st)].state == AT_WAVE) ? b.slot[size_t(st)] : b.slot[size_t(t)];
        }
    }
    // release only after every result is computed: sources are read while still waiting
    for (int t = lo; t < hi; ++t) {
        Fiber& f = b.fibers[size_t(t)];
        if (f.state != AT_WAVE) continue;
        f.state = RUNNABLE;
        ++b.count[size_t(t)];
    }
}

void run_block(Block& b) {
    const int n = int(b.fibers.size());
    for (int t = 0; t < n; ++t) {
        Fiber& f = b.fibers[size_t(t)];
        f.state = RUNNABLE;
        f.tid = unsigned(t);
        b.count[size_t(t)] = 0;
        f.sp = make_fiber_stack(f.stack, kStack);
    }
    while (true) {
        for (int t = 0; t < n; ++t)
            while (b.fibers[size_t(t)].state == RUNNABLE) resume(b, t);
        int live = 0, at_barrier = 0;
        bool released = false;
        for (int w = 0; w * kWave < n; ++w) {
            int wl = 0, ww = 0;
            for (int t = w * kWave; t < std::min(n, (w + 1) * kWave); ++t) {
                const FiberState s = b.fibers[size_t(t)].state;
                if (s != DONE) ++wl;
                if (s == AT_WAVE) ++ww;
            }
            if (ww > 0 && ww == wl) { complete_wave(b, w, n); released = true; }
        }
        if (released) continue;
        for (const Fiber& f : b.fibers) {
            if (f.state != DONE) ++live;
            if (f.state == AT_BARRIER) ++at_barrier;
        }
        if (live == 0) break;
        if (at_barrier == live) {
            for (Fiber& f : b.fibers) if (f.state == AT_BARRIER) f.state = RUNNABLE;
            continue;
        }
        char msg[160];
        std::snprintf(msg, sizeof(msg), "deadlock in block (%u,%u,%u): %d live threads, %d at __syncthreads",
                      b.idx.x, b.idx.y, b.idx.z, live, at_barrier);
        fail_launch(msg);
    }
}

void execute(dim3 grid, dim3 block, size_t dyn_lds, const std::function<void()>& body) {
    const unsigned nthreads = block.x * block.y * block.z;
    Block b;
    b.grid = grid;
    b.dim = block;
    b.body = &body;
    b.fibers.resize(nthreads);
    for (Fiber& f : b.fibers) f.stack = static_cast<char*>(std::malloc(kStack));
    b.slot.assign(nthreads, 0);
    b.result.assign(nthreads, 0);
    b.mode.assign(nthreads, 0);
    b.arg.assign(nthreads, 0);
    b.site.assign(nthreads, nullptr);
    b.count.assign(nthreads, 0);
    Block* outer = g_block;
    g_block = &b;
    for (unsigned z = 0; z < grid.z; ++z)
        for (unsigned y = 0; y < grid.y; ++y)
            for (unsigned x = 0; x < grid.x; ++x) {
                b.idx = {x, y, z};
                // fresh per block, so overruns of the declared size are caught; filled
                // with a pattern because LDS starts with undefined contents
                b.dyn = std::malloc(dyn_lds ? dyn_lds : 1);
                std::memset(b.dyn, 0xA5, dyn_lds ? dyn_lds : 1);
                run_block(b);
                std::free(b.dyn);
                b.dyn = nullptr;
            }
    g_block = outer;
    for (Fiber& f : b.fibers) std::free(f.stack);
}

}  // namespace

hipError_t launch(dim3 grid, dim3 block, size_t dyn_lds, hipStream_t s, std::function<void()> body) {
    const unsigned long long nthreads = 1ull * block.x * block.y * block.z;
    if (nthreads == 0 || nthreads > 1024 || grid.x == 0 || grid.y == 0 || grid.z == 0 || grid.y > 65535 ||
        grid.z > 65535 || dyn_lds > kMaxDynLds) {
        g_last = hipErrorInvalidConfiguration;
        return g_last;
    }
    if (s && s->capturing) {
        s->graph->nodes.push_back([grid, block, dyn_lds, body]() { execute(grid, block, dyn_lds, body); });
        return hipSuccess;
    }
    execute(grid, block, dyn_lds, body);
    return hipSuccess;
}

uint3 thread_idx() {
    const Block& b = *g_block;
    const unsigned t = b.fibers[size_t(b.cur)].tid;
    return {t % b.dim.x, (t / b.dim.x) % b.dim.y, t / (b.dim.x * b.dim.y)};
}
uint3 block_idx() { return g_block->idx; }
dim3 block_dim() { return g_block->dim; }
dim3 grid_dim() { return g_block->grid; }
void* dynamic_lds() { return g_block->dyn; }

void block_barrier() {
    me().state = AT_BARRIER;
    yield_to_scheduler();
}

uint64_t wave_exchange_u64(uint64_t v, int mode, int arg) {
    Block& b = *g_block;
    const size_t t = size_t(b.cur);
    b.site[t] = __builtin_return_address(0);
    b.slot[t] = v;
    b.mode[t] = mode;
    b.arg[t] = arg;
    me().state = AT_WAVE;
    yield_to_scheduler();
    return b.result[size_t(b.cur)];
}

}  // namespace rpd_emu

// ---- runtime API --------------------------------------------------------------------------
using rpd_emu::g_last;

// While a stream captures, the device runtime rejects synchronous copies,
// allocator calls and stream/event creation from other threads; so does the
// emulation (the capturing thread itself only enqueues work).
static std::atomic<int> g_capturing{0};
static thread_local int t_capturing = 0;
static bool capture_conflict() { return g_capturing.load() > t_capturing; }
#define EMU_NO_CAPTURE() \
    do { if (capture_conflict()) return hipErrorStreamCaptureImplicit; } while (0)

hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
    std::memset(p, 0, sizeof(*p));
    std::strcpy(p->name, "cpu-emulation");
    p->totalGlobalMem = size_t(16) << 30;
    p->multiProcessorCount = 4;   // few blocks per persistent launch keeps emulation quick
    p->warpSize = 64;
    p->maxThreadsPerBlock = 1024;
    p->sharedMemPerBlock = 64 * 1024;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
    if (a != hipDeviceAttributeMultiprocessorCount) return hipErrorInvalidValue;
    *v = 4;
    return hipSuccess;
}
hipError_t hipMemGetInfo(size_t* f, size_t* t) { *f = size_t(12) << 30; *t = size_t(16) << 30; return hipSuccess; }
hipError_t hipMalloc(void** p, size_t bytes) {
    EMU_NO_CAPTURE();
    *p = std::malloc(bytes ? bytes : 1);
    if (!*p) return hipErrorOutOfMemory;
    // Device memory is not cleared by the allocator, and what it holds depends
    // on earlier allocations: each allocation gets a different fill byte, so
    // results that read memory nobody wrote differ between otherwise equal runs.
    static std::atomic<unsigned> n_alloc{0};
    const unsigned char fill = (unsigned char)(0x81 + 2 * (n_alloc.fetch_add(1) % 61));
    std::memset(*p, fill, bytes ? bytes : 1);
    return hipSuccess;
}
hipError_t hipFree(void* p) { EMU_NO_CAPTURE(); std::free(p); return hipSuccess; }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
    EMU_NO_CAPTURE();
    if (n) std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st) {
    if (st && st->capturing) {
        st->graph->nodes.push_back([d, s, n]() { if (n) std::memmove(d, s, n); });
        return hipSuccess;
    }
    if (n) std::memmove(d, s, n);   // asynchronous on a stream: no capture conflict
    (void)k;
    return hipSuccess;
}
hipError_t hipMemset(void* d, int v, size_t n) { if (n) std::memset(d, v, n); return hipSuccess; }
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st) {
    if (st && st->capturing) {
        st->graph->nodes.push_back([d, v, n]() { if (n) std::memset(d, v, n); });
        return hipSuccess;
    }
    return hipMemset(d, v, n);
}
hipError_t hipStreamCreate(hipStream_t* s) { EMU_NO_CAPTURE(); *s = new ihipStream_t(); return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { return hipStreamCreate(s); }
hipError_t hipStreamDestroy(hipStream_t s) { EMU_NO_CAPTURE(); delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t s) {
    return (s && s->capturing) ? hipErrorStreamCaptureUnsupported : hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) { EMU_NO_CAPTURE(); *e = new ihipEvent_t(); (*e)->t = std::chrono::steady_clock::now(); return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { EMU_NO_CAPTURE(); delete e; return hipSuccess; }
// Every launch runs when it is called, so work queued before the event is done.
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t, unsigned) {
    return (s && s->capturing) ? hipErrorStreamCaptureUnsupported : hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    if (s && s->capturing) return hipErrorStreamCaptureUnsupported;
    e->t = std::chrono::steady_clock::now();
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    if (!s || s->capturing) return hipErrorInvalidValue;
    s->capturing = true;
    s->graph = new ihipGraph();
    ++g_capturing;
    ++t_capturing;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
    if (!s || !s->capturing) return hipErrorInvalidValue;
    s->capturing = false;
    --g_capturing;
    --t_capturing;
    *g = s->graph;
    s->graph = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    *e = new hipGraphExec();
    (*e)->nodes = g->nodes;
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    for (const auto& n : e->nodes) n();
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipGetLastError() { const hipError_t e = g_last; g_last = hipSuccess; return e; }
hipError_t hipPeekAtLastError() { return g_last; }
const char* hipGetErrorString(hipError_t e) {
    switch (e) {
        case hipSuccess: return "hipSuccess";
        case hipErrorInvalidValue: return "hipErrorInvalidValue";
        case hipErrorOutOfMemory: return "hipErrorOutOfMemory";
        case hipErrorInvalidConfiguration: return "hipErrorInvalidConfiguration";
        case hipErrorStreamCaptureUnsupported: return "hipErrorStreamCaptureUnsupported";
        case hipErrorStreamCaptureImplicit: return "hipErrorStreamCaptureImplicit";
        default: return "hipErrorUnknown";
    }
}
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void*, int, size_t lds) {
    *n = lds > 160 * 1024 ? 0 : 1;
    return hipSuccess;
}
