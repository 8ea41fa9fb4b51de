// Host emulation of the HIP execution model: fibers per work-item, one host thread per resident
// work-group, in-order dispatch of the work-groups of a launch (test infrastructure only; the
// interface is declared in hip/hip_runtime.h next to this file).
#include "hip/hip_runtime.h"

#include <sys/mman.h>
#include <unistd.h>

#include <array>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#if defined(__has_feature)
#if __has_feature(address_sanitizer)
#define EMU_ASAN 1
#include <sanitizer/asan_interface.h>
#include <sanitizer/common_interface_defs.h>
#endif
#endif

namespace emu {

namespace {

constexpr size_t kStack = 1 << 20;   // per work-item stack (kernels keep large register arrays)
enum { RUN = 0, BAR = 1, WAVE = 2, DONE = 3 };

struct LaneData {
    uint64_t w[4];
};

// Switches stacks: saves the callee-saved registers on the current stack, stores the stack
// pointer in *save_sp and resumes the context saved at new_sp (System V x86-64; the kernels never
// change the floating-point control words, so those are not switched).
extern "C" void emu_switch(void **save_sp, void *new_sp);
asm(R"(
    .text
    .globl emu_switch
    .type emu_switch, @function
emu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size emu_switch, .-emu_switch
)");

struct Fiber {
    void *sp = nullptr;
    char *stack = nullptr;
    dim3 tid;
    int linear = 0, wave = 0;
    int st = RUN;
    uint64_t site = 0;
    unsigned bar_gen[2] = {~0u, ~0u};
    int pred[2] = {0, 0};
    unsigned naps = 0;
    void *fake = nullptr;
};

struct Wave {
    LaneData slot[kWave];   // each lane's operand at its pending operation
    LaneData snap[kWave];   // operands of the lanes of the last released operation
    uint64_t snap_mask = 0;
    int run = 0, wait = 0;  // lanes that can run, lanes waiting at a wavefront operation
};

// One work-group on one host thread.  Work-items switch directly to the next runnable one when
// they block; barriers and wavefront operations are released by counters as the last lane that
// could still run blocks.
struct BlockCtx {
    std::vector<Fiber> f;
    std::vector<Wave> waves;
    void *sched_sp = nullptr;
    const void *sched_bottom = nullptr;
    size_t sched_size = 0;
    int n = 0, live = 0, at_bar = 0;
    Fiber *cur = nullptr;
    dim3 bidx, bdim, gdim;
    const std::function<void()> *body = nullptr;
    unsigned bar_gen = 0;
    const char *what = "";
    long lin = 0;
    long events = 0, idle_mark = 0, idle = 0;   // blocking events; naps since the last one
    std::chrono::steady_clock::time_point deadline;
};

thread_local BlockCtx *tl = nullptr;
thread_local hipError_t last_error = hipSuccess;

[[noreturn]] void die(const char *fmt, const char *what, long a = 0, long b = 0) {
    std::fprintf(stderr, "[emu] ");
    std::fprintf(stderr, fmt, what, a, b);
    std::fprintf(stderr, "\n");
    std::fflush(stderr);
    std::abort();
}

// switch from the running fiber (or the host stack, from == nullptr) to `to` (or back to the
// host stack, to == nullptr); `final`: `from` never resumes
void do_switch(BlockCtx &b, Fiber *from, Fiber *to, bool final) {
    void **save = from ? &from->sp : &b.sched_sp;
    void *dest = to ? to->sp : b.sched_sp;
    b.cur = to;
#ifdef EMU_ASAN
    void **fake = final ? nullptr : (from ? &from->fake : nullptr);
    void *sched_fake = nullptr;
    if (!from) fake = &sched_fake;
    __sanitizer_start_switch_fiber(fake, to ? to->stack : b.sched_bottom, to ? kStack : b.sched_size);
    emu_switch(save, dest);
    __sanitizer_finish_switch_fiber(from ? from->fake : sched_fake, nullptr, nullptr);
#else
    (void)final;
    emu_switch(save, dest);
#endif
}

void release_wave(BlockCtx &b, int w) {
    Wave &wv = b.waves[w];
    if (wv.run > 0 || wv.wait == 0) return;
    const int l0 = w * kWave, l1 = std::min(b.n, l0 + kWave);
    uint64_t site = 0;
    for (int l = l0; l < l1; ++l)
        if (b.f[l].st == WAVE) {
            site = b.f[l].site;
            break;
        }
    wv.snap_mask = 0;
    for (int i = 0; i < kWave; ++i) {
        Fiber *g = l0 + i < l1 ? &b.f[l0 + i] : nullptr;
        if (g && g->st == WAVE && g->site == site) {
            wv.snap[i] = wv.slot[i];
            wv.snap_mask |= 1ull << i;
            g->st = RUN;
            ++wv.run;
            --wv.wait;
        } else {
            wv.snap[i] = LaneData{};
        }
    }
}

void release_barrier(BlockCtx &b) {
    if (b.at_bar == 0 || b.at_bar != b.live) return;
    ++b.bar_gen;
    b.at_bar = 0;
    for (int i = 0; i < b.n; ++i)
        if (b.f[i].st == BAR) {
            b.f[i].st = RUN;
            ++b.waves[b.f[i].wave].run;
        }
}

// The running fiber f gives up the host thread (f->st already set): to the next runnable fiber
// after it, to itself if it is the only one, or back to the host stack when the work-group ended.
void switch_next(BlockCtx &b, Fiber *f) {
    Fiber *next = nullptr;
    for (int k = 1; k <= b.n; ++k) {
        Fiber &g = b.f[(f->linear + k) % b.n];
        if (g.st == RUN) {
            next = &g;
            break;
        }
    }
    if (!next) {
        if (b.live == 0) {
            do_switch(b, f, nullptr, true);
            die("%s: finished work-item resumed", b.what);
        }
        int bar = 0, wave = 0;
        for (int i = 0; i < b.n; ++i) {
            bar += b.f[i].st == BAR;
            wave += b.f[i].st == WAVE;
        }
        std::fprintf(stderr, "[emu] %s: work-group %ld deadlocked: %d at the work-group barrier, %d at a wavefront "
                     "operation, %d exited (divergent barrier or wavefront operation)\n", b.what, b.lin, bar, wave,
                     b.n - b.live);
        std::fflush(stderr);
        std::abort();
    }
    if (next == f) return;
    do_switch(b, f, next, f->st == DONE);
}

void block_on(Fiber *f, int st) {
    BlockCtx &b = *tl;
    Wave &wv = b.waves[f->wave];
    --wv.run;
    if (st == BAR) ++b.at_bar;
    else if (st == WAVE) ++wv.wait;
    else --b.live;
    f->st = st;
    ++b.events;
    release_wave(b, f->wave);
    release_barrier(b);
    switch_next(b, f);
}

[[noreturn]] void fiber_entry() {
    BlockCtx *b = tl;
    Fiber *f = b->cur;
#ifdef EMU_ASAN
    const void *bottom = nullptr;
    size_t size = 0;
    __sanitizer_finish_switch_fiber(nullptr, &bottom, &size);
    if (!b->sched_bottom) {   // the first work-item is entered from the host stack
        b->sched_bottom = bottom;
        b->sched_size = size;
    }
#endif
    (*b->body)();
    block_on(f, DONE);
    die("%s: finished work-item resumed", b->what);
}

char *new_stack() {
    void *p = mmap(nullptr, kStack, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (p == MAP_FAILED) die("%s: stack allocation failed", "mmap");
    mprotect(p, 4096, PROT_NONE);   // guard page: a stack overflow faults instead of corrupting
    return static_cast<char *>(p);
}

// Runs one work-group to completion on the calling host thread.
void run_block(BlockCtx &b, long lin, dim3 grid, dim3 block, const char *what, const std::function<void()> &body,
               std::chrono::steady_clock::time_point deadline) {
    const int n = (int)(block.x * block.y * block.z);
    if ((int)b.f.size() < n) {
        const size_t old = b.f.size();
        b.f.resize(n);
        for (size_t i = old; i < (size_t)n; ++i) b.f[i].stack = new_stack();
    }
    const int nw = (n + kWave - 1) / kWave;
    b.waves.assign(nw, Wave{});
    b.n = n;
    b.live = n;
    b.at_bar = 0;
    b.bdim = block;
    b.gdim = grid;
    b.bidx = dim3((unsigned)(lin % grid.x), (unsigned)((lin / grid.x) % grid.y), (unsigned)(lin / ((long)grid.x * grid.y)));
    b.body = &body;
    b.bar_gen = 0;
    b.what = what;
    b.lin = lin;
    b.deadline = deadline;
    b.sched_bottom = nullptr;
    for (int i = 0; i < n; ++i) {
        Fiber &f = b.f[i];
        f.linear = i;
        f.wave = i / kWave;
        f.tid = dim3((unsigned)(i % block.x), (unsigned)((i / block.x) % block.y), (unsigned)(i / (block.x * block.y)));
        f.st = RUN;
        f.bar_gen[0] = f.bar_gen[1] = ~0u;
        f.naps = 0;
        f.fake = nullptr;
        ++b.waves[f.wave].run;
        // initial frame, as emu_switch leaves one: six registers and the return address
        // fiber_entry (entered with rsp = 8 mod 16, as after a call)
#ifdef EMU_ASAN
        // the previous work-group's frames left their redzones poisoned on this stack
        __asan_unpoison_memory_region(f.stack + 4096, kStack - 4096);
#endif
        uint64_t *top = reinterpret_cast<uint64_t *>(f.stack + kStack - 64 - 8);
        uint64_t *sp = top - 7;
        for (int r = 0; r < 6; ++r) sp[r] = 0;
        sp[6] = reinterpret_cast<uint64_t>(&fiber_entry);
        f.sp = sp;
    }
    tl = &b;
    do_switch(b, nullptr, &b.f[0], false);
    tl = nullptr;
}

// ---- worker pool: kResident host threads, each runs one work-group at a time -------------------
// A launch publishes its parameters and a ticket (launch number << 32 | next work-group) under the
// mutex; a worker copies the parameters of the launch it joins and claims work-groups only while
// the ticket still carries that launch, so a worker that wakes late can neither run nor count a
// work-group of another launch.  The launch returns once every work-group has finished.
struct Pool {
    struct Launch {
        const std::function<void()> *body = nullptr;
        dim3 grid, block;
        long nblocks = 0;
        const char *what = "";
        std::chrono::steady_clock::time_point deadline;
    };
    std::mutex m, launch_m;
    std::condition_variable cv, done_cv;
    std::vector<std::thread> th;
    std::vector<BlockCtx *> ctxs;   // each worker's work-group (for the stall report)
    uint32_t job = 0;
    Launch cur;
    std::atomic<uint64_t> ticket{0};
    long finished = 0;
    std::chrono::steady_clock::time_point started;

    Pool() {
        ctxs.assign(kResident, nullptr);
        for (int i = 0; i < kResident; ++i) th.emplace_back([this, i] { worker(i); });
    }
    // A launch that makes no progress for a minute: print where its work-groups wait.
    void report_stall() {
        std::fprintf(stderr, "[emu] %s: no completion after %ld s: %ld of %ld work-groups finished, next %ld\n", cur.what,
                     (long)std::chrono::duration_cast<std::chrono::seconds>(std::chrono::steady_clock::now() - started).count(),
                     finished, cur.nblocks, (long)(ticket.load() & 0xffffffffu));
        for (int i = 0; i < kResident; ++i) {
            const BlockCtx *c = ctxs[i];
            if (!c || c->live == 0) continue;
            int run = 0, napping = 0;
            for (int j = 0; j < c->n; ++j) {
                run += c->f[j].st == RUN;
                napping += c->f[j].st == RUN && c->f[j].naps > 0;
            }
            std::fprintf(stderr, "[emu]   worker %d: work-group %ld, %d live, %d at the barrier, %d runnable (%d polled)\n", i,
                         c->lin, c->live, c->at_bar, run, napping);
        }
        std::fflush(stderr);
    }
    void worker(int id) {
        BlockCtx ctx;
        ctxs[id] = &ctx;
        uint32_t seen = 0;
        for (;;) {
            Launch l;
            uint32_t mine;
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return job != seen; });
                seen = mine = job;
                l = cur;
            }
            for (;;) {
                uint64_t t = ticket.load();
                bool claimed = false;
                while ((uint32_t)(t >> 32) == mine && (long)(t & 0xffffffffu) < l.nblocks)
                    if (ticket.compare_exchange_weak(t, t + 1)) {
                        claimed = true;
                        break;
                    }
                if (!claimed) break;
                run_block(ctx, (long)(t & 0xffffffffu), l.grid, l.block, l.what, *l.body, l.deadline);
                std::lock_guard<std::mutex> lk(m);
                if (++finished == l.nblocks) done_cv.notify_all();
            }
        }
    }
    void run(const char *w, dim3 g, dim3 b, const std::function<void()> &f) {
        std::lock_guard<std::mutex> one(launch_m);
        std::unique_lock<std::mutex> lk(m);
        cur.what = w;
        cur.grid = g;
        cur.block = b;
        cur.body = &f;
        cur.nblocks = (long)g.x * g.y * g.z;
        started = std::chrono::steady_clock::now();
        cur.deadline = started + std::chrono::seconds(900);
        finished = 0;
        ++job;
        ticket.store((uint64_t)job << 32);
        cv.notify_all();
        while (!done_cv.wait_for(lk, std::chrono::seconds(60), [&] { return finished == cur.nblocks; })) report_stall();
    }
};

Pool &pool() {
    static Pool *p = new Pool();   // never destroyed: its threads live until the process ends
    return *p;
}

std::unordered_map<const void *, Invoker> &registry() {
    static std::unordered_map<const void *, Invoker> r;
    return r;
}

}  // namespace

const dim3 &thread_idx() { return tl->cur->tid; }
const dim3 &block_idx() { return tl->bidx; }
const dim3 &block_dim() { return tl->bdim; }
const dim3 &grid_dim() { return tl->gdim; }

void block_barrier() {
    Fiber *f = tl->cur;
    const unsigned g = tl->bar_gen;
    f->bar_gen[g & 1] = g;
    block_on(f, BAR);
}

int block_reduce(int pred, int op) {
    Fiber *f = tl->cur;
    const unsigned g = tl->bar_gen;
    f->pred[g & 1] = pred;
    f->bar_gen[g & 1] = g;
    block_on(f, BAR);
    int acc = op == 0 ? 1 : 0;
    for (int i = 0; i < tl->n; ++i) {
        const Fiber &o = tl->f[i];
        if (o.bar_gen[g & 1] != g) continue;
        const int p = o.pred[g & 1] != 0;
        if (op == 0) acc = acc && p;
        else if (op == 1) acc = acc || p;
        else acc += p;
    }
    return acc;
}

void nap() {
    BlockCtx &b = *tl;
    Fiber *f = b.cur;
    ++f->naps;
    if (b.events != b.idle_mark) {
        b.idle_mark = b.events;
        b.idle = 0;
    }
    // a work-group whose runnable work-items only poll waits for other work-groups: give their
    // host threads the cores
    if (++b.idle > 4L * b.live) {   // every runnable work-item polled four times without progress
        b.idle = 0;
        std::this_thread::sleep_for(std::chrono::microseconds(5));
        if (std::chrono::steady_clock::now() > b.deadline)
            die("%s: work-group %ld exceeded the emulator's time limit", b.what, b.lin);
    }
    switch_next(b, f);
}

uint64_t wave_gather(uint64_t site, const void *mine, size_t bytes, const void **all, int *lane) {
    Fiber *f = tl->cur;
    const int l = f->linear % kWave;
    Wave &wv = tl->waves[f->wave];
    std::memset(&wv.slot[l], 0, sizeof(LaneData));
    std::memcpy(&wv.slot[l], mine, bytes);
    f->site = site;
    block_on(f, WAVE);
    // released: the snapshot stays valid until every lane of this operation blocked again
    *all = wv.snap;
    *lane = l;
    return wv.snap_mask;
}

int readfirstlane(uint64_t site, int v) {
    const LaneData *all;
    int lane;
    const uint64_t m = wave_gather(site, &v, sizeof(v), reinterpret_cast<const void **>(&all), &lane);
    return (int)all[__builtin_ctzll(m)].w[0];
}

int readlane(uint64_t site, int v, int l) {
    const LaneData *all;
    int lane;
    wave_gather(site, &v, sizeof(v), reinterpret_cast<const void **>(&all), &lane);
    return (int)all[l & (kWave - 1)].w[0];
}

uint64_t ballot(uint64_t site, int pred) {
    const LaneData *all;
    int lane;
    const int p = pred != 0;
    const uint64_t m = wave_gather(site, &p, sizeof(p), reinterpret_cast<const void **>(&all), &lane);
    uint64_t r = 0;
    for (int i = 0; i < kWave; ++i)
        if ((m >> i & 1) && all[i].w[0]) r |= 1ull << i;
    return r;
}

// DPP source lane of `lane` for the controls the kernels use (-1: no valid source)
static int dpp_source(int lane, int ctrl) {
    const int row = lane & ~15, in = lane & 15;
    if (ctrl >= 0 && ctrl <= 0xFF) return (lane & ~3) + ((ctrl >> (2 * (lane & 3))) & 3);   // quad_perm
    if (ctrl >= 0x101 && ctrl <= 0x10F) { const int s = in + (ctrl & 15); return s < 16 ? row + s : -1; }   // row_shl
    if (ctrl >= 0x111 && ctrl <= 0x11F) { const int s = in - (ctrl & 15); return s >= 0 ? row + s : -1; }   // row_shr
    if (ctrl >= 0x121 && ctrl <= 0x12F) return row + ((in - (ctrl & 15)) & 15);                            // row_ror
    if (ctrl == 0x140) return row + (15 - in);                                                             // row_mirror
    if (ctrl == 0x141) return row + (in & 8) + (7 - (in & 7));                                            // row_half_mirror
    if (ctrl == 0x142) return row >= 16 ? row - 1 : -1;                                                    // row_bcast15
    if (ctrl == 0x143) return lane >= 32 ? 31 : -1;                                                        // row_bcast31
    std::fprintf(stderr, "[emu] unsupported DPP control 0x%x\n", ctrl);
    std::abort();
}

int dpp(uint64_t site, int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl, bool keep_old_if_masked) {
    const LaneData *all;
    int lane;
    const uint64_t m = wave_gather(site, &src, sizeof(src), reinterpret_cast<const void **>(&all), &lane);
    const int row = lane >> 4, bank = (lane >> 2) & 3;
    const int own = keep_old_if_masked ? old : src;   // v_mov_b32_dpp keeps the destination: here the source
    if (!((row_mask >> row) & 1) || !((bank_mask >> bank) & 1)) return own;
    const int s = dpp_source(lane, ctrl);
    if (s < 0 || !((m >> s) & 1)) return bound_ctrl ? 0 : own;
    return (int)all[s].w[0];
}

int ds_swizzle(uint64_t site, int v, int pattern) {
    const LaneData *all;
    int lane;
    wave_gather(site, &v, sizeof(v), reinterpret_cast<const void **>(&all), &lane);
    if (pattern & 0x8000) {
        std::fprintf(stderr, "[emu] unsupported ds_swizzle pattern 0x%x\n", pattern);
        std::abort();
    }
    const int and_m = pattern & 31, or_m = (pattern >> 5) & 31, xor_m = (pattern >> 10) & 31;
    const int s = (lane & ~31) | ((((lane & 31) & and_m) | or_m) ^ xor_m);
    return (int)all[s].w[0];
}

// MFMA: gather A and B of the 64 lanes, then the lane's accumulator elements.
f32x4 mfma_f32_16x16x4(uint64_t site, float a, float b, f32x4 c) {
    const LaneData *all;
    int lane;
    float ab[2] = {a, b};
    wave_gather(site, ab, sizeof(ab), reinterpret_cast<const void **>(&all), &lane);
    auto A = [&](int i, int k) { float x[2]; std::memcpy(x, &all[k * 16 + i], 8); return x[0]; };
    auto B = [&](int k, int j) { float x[2]; std::memcpy(x, &all[k * 16 + j], 8); return x[1]; };
    const int li = lane & 15, lk = lane >> 4;
    f32x4 d;
    for (int e = 0; e < 4; ++e) {
        const int r = 4 * lk + e;
        float s = c[e];
        for (int k = 0; k < 4; ++k) s = std::fma(A(r, k), B(k, li), s);
        d[e] = s;
    }
    return d;
}

f64x4 mfma_f64_16x16x4(uint64_t site, double a, double b, f64x4 c) {
    const LaneData *all;
    int lane;
    double ab[2] = {a, b};
    wave_gather(site, ab, sizeof(ab), reinterpret_cast<const void **>(&all), &lane);
    auto A = [&](int i, int k) { double x[2]; std::memcpy(x, &all[k * 16 + i], 16); return x[0]; };
    auto B = [&](int k, int j) { double x[2]; std::memcpy(x, &all[k * 16 + j], 16); return x[1]; };
    const int li = lane & 15, lk = lane >> 4;
    f64x4 d;
    for (int e = 0; e < 4; ++e) {
        const int r = lk + 4 * e;
        double s = c[e];
        for (int k = 0; k < 4; ++k) s = std::fma(A(r, k), B(k, li), s);
        d[e] = s;
    }
    return d;
}

f32x16 mfma_f32_32x32x2(uint64_t site, float a, float b, f32x16 c) {
    const LaneData *all;
    int lane;
    float ab[2] = {a, b};
    wave_gather(site, ab, sizeof(ab), reinterpret_cast<const void **>(&all), &lane);
    auto A = [&](int i, int k) { float x[2]; std::memcpy(x, &all[k * 32 + i], 8); return x[0]; };
    auto B = [&](int k, int j) { float x[2]; std::memcpy(x, &all[k * 32 + j], 8); return x[1]; };
    const int col = lane & 31;
    f32x16 d;
    for (int e = 0; e < 16; ++e) {
        const int r = 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        float s = c[e];
        for (int k = 0; k < 2; ++k) s = std::fma(A(r, k), B(k, col), s);
        d[e] = s;
    }
    return d;
}

void run_grid(const char *what, dim3 grid, dim3 block, bool cooperative, const std::function<void()> &body) {
    const long nthreads = (long)block.x * block.y * block.z;
    if (grid.x == 0 || grid.y == 0 || grid.z == 0 || nthreads == 0 || nthreads > 1024 || grid.y > 65535 || grid.z > 65535) {
        last_error = hipErrorInvalidConfiguration;
        return;
    }
    if ((long)grid.x * grid.y * grid.z > 0xffffffffL) {   // the pool's ticket holds 32 bits of work-group index
        last_error = hipErrorInvalidConfiguration;
        return;
    }
    if (cooperative && (long)grid.x * grid.y * grid.z > kResident) {
        last_error = hipErrorCooperativeLaunchTooLarge;
        return;
    }
    static const bool trace = std::getenv("EMU_TRACE") != nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    if (trace) {
        std::fprintf(stderr, "[emu] launch %s grid (%u,%u,%u) block (%u,%u,%u)%s\n", what, grid.x, grid.y, grid.z, block.x,
                     block.y, block.z, cooperative ? " cooperative" : "");
        std::fflush(stderr);
    }
    pool().run(what, grid, block, body);
    if (trace) {
        std::fprintf(stderr, "[emu]   done in %.3f s\n",
                     std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        std::fflush(stderr);
    }
}

void set_launch_error(hipError_t e) { last_error = e; }

void register_kernel(const void *f, Invoker inv) { registry()[f] = std::move(inv); }
const Invoker *find_kernel(const void *f) {
    auto it = registry().find(f);
    return it == registry().end() ? nullptr : &it->second;
}

}  // namespace emu

// ---- runtime API ---------------------------------------------------------------------------------
struct emu_stream {
    unsigned flags = 0;
};
struct emu_event {
    std::chrono::steady_clock::time_point t;
};

hipError_t hipMalloc(void **p, size_t bytes) {
    if (!p) return hipErrorInvalidValue;
    if (bytes == 0) { *p = nullptr; return hipSuccess; }
    void *q = nullptr;
    if (posix_memalign(&q, 256, bytes) != 0) return hipErrorOutOfMemory;
    std::memset(q, 0xFF, bytes);   // not zero: fresh device memory holds whatever it held
    *p = q;
    return hipSuccess;
}
hipError_t hipFree(void *p) { std::free(p); return hipSuccess; }
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned) { return hipMalloc(p, bytes); }
hipError_t hipHostFree(void *p) { std::free(p); return hipSuccess; }
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind) {
    if (bytes) std::memmove(dst, src, bytes);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind k, hipStream_t) { return hipMemcpy(dst, src, bytes, k); }
hipError_t hipMemcpy2D(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                       hipMemcpyKind) {
    if (width > dpitch || width > spitch) return hipErrorInvalidValue;
    for (size_t r = 0; r < height; ++r)
        std::memmove(static_cast<char *>(dst) + r * dpitch, static_cast<const char *>(src) + r * spitch, width);
    return hipSuccess;
}
hipError_t hipMemcpy2DAsync(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                            hipMemcpyKind k, hipStream_t) {
    return hipMemcpy2D(dst, dpitch, src, spitch, width, height, k);
}
hipError_t hipMemset(void *p, int v, size_t bytes) {
    if (bytes) std::memset(p, v, bytes);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void *p, int v, size_t bytes, hipStream_t) { return hipMemset(p, v, bytes); }
hipError_t hipMemGetInfo(size_t *f, size_t *t) {
    *f = (size_t)16 << 30;
    *t = (size_t)288 << 30;
    return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t *s) { *s = new emu_stream(); return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags) { *s = new emu_stream{flags}; return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e) { *e = new emu_event(); (*e)->t = std::chrono::steady_clock::now(); return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { e->t = std::chrono::steady_clock::now(); return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
    *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
    return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidDevice; }
hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d) {
    if (d != 0) return hipErrorInvalidDevice;
    std::memset(p, 0, sizeof(*p));
    std::snprintf(p->name, sizeof(p->name), "host emulation of gfx950");
    std::snprintf(p->gcnArchName, sizeof(p->gcnArchName), "gfx950");
    p->totalGlobalMem = (size_t)288 << 30;
    p->sharedMemPerBlock = 160 * 1024;
    p->warpSize = 64;
    p->maxThreadsPerBlock = 1024;
    p->maxThreadsDim[0] = p->maxThreadsDim[1] = p->maxThreadsDim[2] = 1024;
    p->maxGridSize[0] = 0x7fffffff;
    p->maxGridSize[1] = p->maxGridSize[2] = 65535;
    p->multiProcessorCount = emu::kCUs;
    p->cooperativeLaunch = 1;
    p->maxThreadsPerMultiProcessor = 2048;
    p->major = 9;
    p->minor = 5;
    return hipSuccess;
}
hipError_t hipGetLastError() {
    const hipError_t e = emu::last_error;
    emu::last_error = hipSuccess;
    return e;
}
hipError_t hipPeekAtLastError() { return emu::last_error; }
const char *hipGetErrorString(hipError_t e) {
    switch (e) {
        case hipSuccess: return "no error";
        case hipErrorInvalidValue: return "invalid argument";
        case hipErrorOutOfMemory: return "out of memory";
        case hipErrorInvalidConfiguration: return "invalid configuration argument";
        case hipErrorCooperativeLaunchTooLarge: return "cooperative launch too large";
        case hipErrorInvalidDeviceFunction: return "invalid device function";
        default: return "emulated HIP error";
    }
}
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, const void *, int, size_t) {
    *n = emu::kBlocksPerCU;
    return hipSuccess;
}
hipError_t hipLaunchCooperativeKernel(const void *f, dim3 grid, dim3 block, void **args, unsigned, hipStream_t) {
    const emu::Invoker *inv = emu::find_kernel(f);
    if (!inv) return hipErrorInvalidDeviceFunction;
    if ((long)grid.x * grid.y * grid.z > emu::kResident) return hipErrorCooperativeLaunchTooLarge;
    emu::last_error = hipSuccess;
    emu::run_grid("cooperative kernel", grid, block, true, [&]() { (*inv)(args); });
    const hipError_t e = emu::last_error;
    emu::last_error = hipSuccess;
    return e;
}
