#include "hip/hip_runtime.h"
#include <csetjmp>
#include <csignal>
#include <map>
#include <sys/mman.h>
#include <ucontext.h>
#include <unistd.h>
#include <vector>

namespace emu {
struct Fiber {
    jmp_buf jb;
    ucontext_t uc;
    uint3e t;
    int state = 3;  // 0 ready, 1 barrier, 2 wave op, 3 done
    bool started = false;
    int kind = 0, arg = 0, width = 64;
    long long val = 0, res = 0;
    char* stack = nullptr;
};
static const size_t STK = 128 * 1024;
static std::vector<Fiber*> F;
static jmp_buf sched_jb;
Fiber* g_cur = nullptr;
static uint3e g_b, g_bd, g_gd;
static const char* g_name = "";
static size_t g_smem = 0;
static unsigned char* g_smem_ptr = nullptr;
static void (*g_entry)(void*) = nullptr;
static void* g_arg = nullptr;
static long long g_launches = 0;
static std::map<void*, std::pair<void*, size_t>> g_maps;

const uint3e& tidx() { return g_cur->t; }
const uint3e& bidx() { return g_b; }
const uint3e& bdim() { return g_bd; }
const uint3e& gdim() { return g_gd; }

static void yield_to_sched() {
    if (!_setjmp(g_cur->jb)) _longjmp(sched_jb, 1);
}
void syncthreads() {
    g_cur->state = 1;
    yield_to_sched();
}
long long wave_op(int kind, long long v, int arg, int width) {
    g_cur->state = 2;
    g_cur->kind = kind;
    g_cur->val = v;
    g_cur->arg = arg;
    g_cur->width = width;
    yield_to_sched();
    return g_cur->res;
}
static void trampoline() {
    g_entry(g_arg);
    g_cur->state = 3;
    _longjmp(sched_jb, 1);
}

// guarded memory: the buffer ends at a PROT_NONE page, another one precedes it
static void* galloc(size_t bytes) {
    const size_t pg = (size_t)sysconf(_SC_PAGESIZE);
    const size_t r = (std::max<size_t>(bytes, 8) + 7) & ~(size_t)7;
    const size_t dp = (r + pg - 1) / pg;
    const size_t total = (dp + 2) * pg;
    char* base = (char*)mmap(nullptr, total, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (base == MAP_FAILED) return nullptr;
    mprotect(base, pg, PROT_NONE);
    mprotect(base + (1 + dp) * pg, pg, PROT_NONE);
    char* data = base + (1 + dp) * pg - r;
    g_maps[data] = {base, total};
    return data;
}
static void gfree(void* p) {
    auto it = g_maps.find(p);
    if (it == g_maps.end()) return;
    munmap(it->second.first, it->second.second);
    g_maps.erase(it);
}
static bool in_guard(void* a, void** owner) {
    const size_t pg = (size_t)sysconf(_SC_PAGESIZE);
    for (auto& kv : g_maps) {
        char* base = (char*)kv.second.first;
        const size_t total = kv.second.second;
        char* x = (char*)a;
        if (x >= base && x < base + pg) { *owner = kv.first; return true; }
        if (x >= base + total - pg && x < base + total) { *owner = kv.first; return true; }
    }
    return false;
}
static void on_segv(int, siginfo_t* si, void*) {
    void* owner = nullptr;
    const bool g = in_guard(si->si_addr, &owner);
    char buf[512];
    int n = snprintf(buf, sizeof buf,
                     "EMU FAULT: address %p (%s, buffer %p) in kernel %s launch %lld block (%u,%u,%u) thread (%u,%u,%u)\n",
                     si->si_addr, g ? "guard page" : "not a guard page", owner, g_name, g_launches, g_b.x, g_b.y,
                     g_b.z, g_cur ? g_cur->t.x : 0, g_cur ? g_cur->t.y : 0, g_cur ? g_cur->t.z : 0);
    (void)!write(2, buf, n);
    _exit(86);
}
static struct Init {
    Init() {
        static char alt[1 << 16];
        stack_t ss{};
        ss.ss_sp = alt;
        ss.ss_size = sizeof alt;
        sigaltstack(&ss, nullptr);
        struct sigaction sa{};
        sa.sa_sigaction = on_segv;
        sa.sa_flags = SA_SIGINFO | SA_ONSTACK;
        sigaction(SIGSEGV, &sa, nullptr);
        sigaction(SIGBUS, &sa, nullptr);
    }
} g_init;

void* dyn_smem() { return g_smem_ptr; }

void begin_kernel(const char* name, dim3 g, dim3 b, size_t smem) {
    g_name = name;
    g_gd = {g.x, g.y, g.z};
    g_bd = {b.x, b.y, b.z};
    g_smem = smem;
    g_smem_ptr = (unsigned char*)galloc(smem ? smem : 8);
    memset(g_smem_ptr, 0xff, smem ? smem : 8);
    ++g_launches;
    if (getenv("EMU_TRACE")) fprintf(stderr, "launch %lld %s grid (%u,%u,%u) block (%u,%u,%u) smem %zu\n",
                                     g_launches, name, g.x, g.y, g.z, b.x, b.y, b.z, smem);
}
void end_kernel() {
    gfree(g_smem_ptr);
    g_smem_ptr = nullptr;
}

void run_block(void (*entry)(void*), void* arg, unsigned bx, unsigned by, unsigned bz) {
    g_entry = entry;
    g_arg = arg;
    g_b = {bx, by, bz};
    const unsigned n = g_bd.x * g_bd.y * g_bd.z;
    while (F.size() < n) {
        Fiber* f = new Fiber();
        f->stack = (char*)malloc(STK);
        F.push_back(f);
    }
    for (unsigned t = 0; t < n; ++t) {
        Fiber* f = F[t];
        f->t = {t % g_bd.x, (t / g_bd.x) % g_bd.y, t / (g_bd.x * g_bd.y)};
        f->state = 0;
        f->started = false;
    }
    for (;;) {
        for (unsigned t = 0; t < n; ++t) {
            Fiber* f = F[t];
            if (f->state != 0) continue;
            g_cur = f;
            if (!_setjmp(sched_jb)) {
                if (!f->started) {
                    f->started = true;
                    getcontext(&f->uc);
                    f->uc.uc_stack.ss_sp = f->stack;
                    f->uc.uc_stack.ss_size = STK;
                    f->uc.uc_link = nullptr;
                    makecontext(&f->uc, trampoline, 0);
                    ucontext_t tmp;
                    swapcontext(&tmp, &f->uc);
                } else {
                    _longjmp(f->jb, 1);
                }
            }
        }
        // every fibre is at a barrier, a wave operation or done
        unsigned done = 0, bar = 0, wop = 0;
        for (unsigned t = 0; t < n; ++t) {
            const int s = F[t]->state;
            done += s == 3; bar += s == 1; wop += s == 2;
        }
        if (done == n) break;
        if (wop) {
            for (unsigned w0 = 0; w0 < n; w0 += 64) {
                const unsigned w1 = std::min(n, w0 + 64);
                int kind = -1;
                bool any = false, mixed = false;
                for (unsigned t = w0; t < w1; ++t) {
                    Fiber* f = F[t];
                    if (f->state == 2) {
                        if (kind >= 0 && f->kind != kind) mixed = true;
                        kind = f->kind;
                        any = true;
                    } else if (f->state == 1) {
                        mixed = true;
                    }
                }
                if (!any) continue;
                if (mixed) {
                    fprintf(stderr, "EMU ERROR: divergent wave operation in kernel %s block (%u,%u,%u)\n", g_name, bx,
                            by, bz);
                    _exit(87);
                }
                unsigned long long ball = 0;
                for (unsigned t = w0; t < w1; ++t)
                    if (F[t]->state == 2 && F[t]->val) ball |= 1ull << (t - w0);
                for (unsigned t = w0; t < w1; ++t) {
                    Fiber* f = F[t];
                    if (f->state != 2) continue;
                    const unsigned l = t - w0;
                    const int wd = f->width;
                    const unsigned base = l & ~(unsigned)(wd - 1);
                    long long r = f->val;
                    auto get = [&](unsigned src, long long own) -> long long {
                        const unsigned ts = w0 + src;
                        if (ts < w1 && F[ts]->state == 2) return F[ts]->val;
                        return own;
                    };
                    if (kind == 0) r = get(base + ((unsigned)f->arg & (unsigned)(wd - 1)), f->val);
                    else if (kind == 1) {
                        const unsigned src = l ^ (unsigned)f->arg;
                        r = ((src & ~(unsigned)(wd - 1)) == base) ? get(src, f->val) : f->val;
                    } else if (kind == 2) r = (l - base) >= (unsigned)f->arg ? get(l - f->arg, f->val) : f->val;
                    else if (kind == 3) r = (long long)ball;
                    else if (kind == 4) r = get((unsigned)f->arg, 0);
                    f->res = r;
                }
                for (unsigned t = w0; t < w1; ++t)
                    if (F[t]->state == 2) F[t]->state = 0;
            }
            continue;
        }
        for (unsigned t = 0; t < n; ++t)
            if (F[t]->state == 1) F[t]->state = 0;
    }
    g_cur = nullptr;
}
}  // namespace emu

// ------------------------------------------------------------------ runtime
struct emu_stream { int x; };
struct emu_event { int x; };
hipError_t hipMallocAsync(void** p, size_t bytes, hipStream_t) { *p = emu::galloc(bytes); return *p ? hipSuccess : 2; }
hipError_t hipMalloc(void** p, size_t bytes) { return hipMallocAsync(p, bytes, nullptr); }
hipError_t hipFreeAsync(void* p, hipStream_t) { emu::gfree(p); return hipSuccess; }
hipError_t hipFree(void* p) { emu::gfree(p); return hipSuccess; }
hipError_t hipMemsetAsync(void* p, int v, size_t n, hipStream_t) { memset(p, v, n); return hipSuccess; }
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) { memmove(d, s, n); return hipSuccess; }
hipError_t hipMemcpy2DAsync(void* d, size_t dp, const void* s, size_t sp, size_t w, size_t h, hipMemcpyKind, hipStream_t) {
    for (size_t r = 0; r < h; ++r) memmove((char*)d + r * dp, (const char*)s + r * sp, w);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = new emu_stream(); return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = new emu_event(); return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) { *e = new emu_event(); return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }  // launches run to completion at once
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0; return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) { *p = calloc(1, bytes); return hipSuccess; }
hipError_t hipHostFree(void* p) { free(p); return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "emu error"; }
hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) { *v = 163840; return hipSuccess; }
hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) { *g = nullptr; return hipErrorNotSupported; }
hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) { return hipErrorNotSupported; }
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
