// CPU emulation of gfx950 wave64 execution, for testing the engine's device
// code without a GPU -- TEST INFRASTRUCTURE ONLY (tests/test_emu.py builds
// and runs it; libmbe.so never contains it).
//
// One workgroup is one wave of 64 lanes. Each lane runs as a ucontext fiber
// on the wave's OS thread; a lane runs until it reaches a cross-lane
// operation (ballot, readlane, readfirstlane, DPP, shuffle, barrier), and
// the operation completes once every lane has reached it, as in lockstep
// execution. A wave whose lanes reach different cross-lane operations, or
// finish while others wait, aborts with a message: on the GPU that would be
// a divergent cross-lane operation. Waves run on separate OS threads, so
// global atomics and the work-stealing deque see real concurrency. LDS is
// a thread_local array per wave.
#pragma once
#include <sched.h>
#include <ucontext.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#define __device__
#define __global__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __shared__ __thread

#define __HIP_MEMORY_SCOPE_WAVEFRONT 2
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __HIP_MEMORY_SCOPE_SYSTEM 5

namespace emu {

enum Op { kBallot = 1, kFirst, kReadLane, kDpp, kShfl, kBarrier };

struct Wave {
    ucontext_t sched;
    ucontext_t ctx[64];
    int state[64];  // 0 runnable, 1 waiting at a cross-lane op, 2 done
    int op[64];
    uint64_t val[64], old[64], arg[64], res[64];
    void (*fn)(void*);
    void* fnarg;
};

struct Dim {
    uint32_t x, y, z;
};

inline thread_local Wave* g_wave = nullptr;
inline thread_local int g_lane = 0;
inline thread_local uint32_t g_block = 0, g_grid = 1;

[[noreturn]] inline void die(const char* what) {
    fprintf(stderr, "wave emulator: %s (block %u)\n", what, g_block);
    abort();
}

inline uint64_t xlane(int op, uint64_t val, uint64_t old = 0, uint64_t arg = 0) {
    Wave* w = g_wave;
    const int l = g_lane;
    w->op[l] = op;
    w->val[l] = val;
    w->old[l] = old;
    w->arg[l] = arg;
    w->state[l] = 1;
    swapcontext(&w->ctx[l], &w->sched);
    return w->res[l];
}

inline void lane_main(int l) {
    Wave* w = g_wave;
    w->fn(w->fnarg);
    w->state[l] = 2;
    swapcontext(&w->ctx[l], &w->sched);
}

inline uint32_t dpp_lane(int l, uint32_t ctrl, uint32_t row_mask, bool& ok) {
    const int row = l >> 4;
    ok = false;
    if (!((row_mask >> row) & 1)) return 0;
    if (ctrl >= 0x111 && ctrl <= 0x11F) {  // row_shr:n
        const int n = ctrl & 0xF;
        if ((l & 15) < n) return 0;
        ok = true;
        return l - n;
    }
    if (ctrl == 0x142) {  // row_bcast:15 -> lane 15 of the previous row
        if (row == 0) return 0;
        ok = true;
        return row * 16 - 1;
    }
    if (ctrl == 0x143) {  // row_bcast:31 -> lane 31, rows 2 and 3
        if (row < 2) return 0;
        ok = true;
        return 31;
    }
    die("unsupported DPP control");
}

// Run fn(arg) as one wave of 64 lanes on the calling thread.
inline void run_wave(uint32_t block, uint32_t grid, void (*fn)(void*), void* arg) {
    constexpr size_t kStack = 512 * 1024;
    Wave* w = new Wave();
    char* stacks = (char*)malloc(kStack * 64);
    if (!stacks) die("out of memory for lane stacks");
    // poison the lane stacks: a local read before it is written holds
    // nonzero garbage here, as registers do on the GPU
    memset(stacks, 0xA5, kStack * 64);
    w->fn = fn;
    w->fnarg = arg;
    g_wave = w;
    g_block = block;
    g_grid = grid;
    for (int l = 0; l < 64; l++) {
        getcontext(&w->ctx[l]);
        w->ctx[l].uc_stack.ss_sp = stacks + (size_t)l * kStack;
        w->ctx[l].uc_stack.ss_size = kStack;
        w->ctx[l].uc_link = nullptr;
        makecontext(&w->ctx[l], (void (*)())lane_main, 1, l);
        w->state[l] = 0;
    }
    for (;;) {
        for (int l = 0; l < 64; l++)
            if (w->state[l] == 0) {
                g_lane = l;
                swapcontext(&w->sched, &w->ctx[l]);
            }
        int nwait = 0, ndone = 0, op = 0;
        for (int l = 0; l < 64; l++) {
            if (w->state[l] == 2) ndone++;
            if (w->state[l] == 1) {
                if (nwait && w->op[l] != op) die("lanes reached different cross-lane operations");
                op = w->op[l];
                nwait++;
            }
        }
        if (!nwait) break;
        if (ndone) die("lanes finished while others wait at a cross-lane operation");
        switch (op) {
            case kBallot: {
                uint64_t m = 0;
                for (int l = 0; l < 64; l++) m |= (uint64_t)(w->val[l] != 0) << l;
                for (int l = 0; l < 64; l++) w->res[l] = m;
                break;
            }
            case kFirst:
                for (int l = 0; l < 64; l++) w->res[l] = w->val[0];
                break;
            case kReadLane:
                for (int l = 0; l < 64; l++) {
                    if (w->arg[l] != w->arg[0] || w->arg[l] > 63) die("readlane: lane not uniform");
                    w->res[l] = w->val[w->arg[l]];
                }
                break;
            case kShfl:
                for (int l = 0; l < 64; l++) w->res[l] = w->val[(l ^ w->arg[l]) & 63];
                break;
            case kDpp:
                for (int l = 0; l < 64; l++) {
                    bool ok;
                    const uint32_t s = dpp_lane(l, w->arg[l] & 0xFFFF, (w->arg[l] >> 16) & 0xF, ok);
                    w->res[l] = ok ? (uint32_t)w->val[s] : (uint32_t)w->old[l];
                }
                break;
            case kBarrier:
                break;
            default:
                die("unknown cross-lane operation");
        }
        for (int l = 0; l < 64; l++) w->state[l] = 0;
    }
    g_wave = nullptr;
    free(stacks);
    delete w;
}

}  // namespace emu

#define threadIdx (emu::Dim{(uint32_t)emu::g_lane, 0, 0})
#define blockIdx (emu::Dim{emu::g_block, 0, 0})
#define blockDim (emu::Dim{64, 1, 1})
#define gridDim (emu::Dim{emu::g_grid, 1, 1})

inline unsigned long long __ballot(int p) { return emu::xlane(emu::kBallot, p != 0); }
inline int __builtin_amdgcn_readfirstlane(int v) { return (int)emu::xlane(emu::kFirst, (uint32_t)v); }
inline int __builtin_amdgcn_readlane(int v, int l) {
    return (int)emu::xlane(emu::kReadLane, (uint32_t)v, 0, (uint64_t)l);
}
inline int __shfl_xor(int v, int m) { return (int)emu::xlane(emu::kShfl, (uint32_t)v, 0, (uint64_t)m); }
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int, bool) {
    return (int)emu::xlane(emu::kDpp, (uint32_t)src, (uint32_t)old,
                           (uint64_t)(ctrl & 0xFFFF) | ((uint64_t)row_mask << 16));
}
inline void __syncthreads() { emu::xlane(emu::kBarrier, 0); }
inline void __threadfence() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
inline void __builtin_amdgcn_fence(int, const char*) { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
inline void __builtin_amdgcn_s_sleep(int) { sched_yield(); }
inline void __builtin_amdgcn_s_waitcnt(int) {}
inline unsigned long long wall_clock64() {
    return (unsigned long long)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
               .count() / 10;
}
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }

template <class T, class U>
inline T atomicAdd(T* p, U v) { return __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U>
inline T atomicOr(T* p, U v) { return __atomic_fetch_or(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U>
inline T atomicAnd(T* p, U v) { return __atomic_fetch_and(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U>
inline T atomicMax(T* p, U v) {
    T c = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (c < (T)v &&
           !__atomic_compare_exchange_n(p, &c, (T)v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return c;
}
template <class T, class U, class V>
inline T atomicCAS(T* p, U cmp, V val) {
    T c = (T)cmp;
    __atomic_compare_exchange_n(p, &c, (T)val, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return c;
}
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), __ATOMIC_SEQ_CST)
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), __ATOMIC_SEQ_CST)

template <class A, class B>
inline typename std::common_type<A, B>::type min(A a, B b) {
    return a < b ? a : b;
}
template <class A, class B>
inline typename std::common_type<A, B>::type max(A a, B b) {
    return a > b ? a : b;
}
