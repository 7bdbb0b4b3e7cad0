// hip_emu.hpp -- a small SIMT emulation of the HIP intrinsics the search code uses,
// so that uastar_amd/csrc/search.hpp runs unmodified on the CPU in tests.
//
// One host thread per lane. Cross-lane operations (shuffles, readlane) go through a
// per-wave exchange buffer between two waits on a 64-thread barrier, which is valid
// because the device code calls them with all 64 lanes active. __syncthreads is a
// barrier over the block's threads. Atomics map to sequentially consistent host
// atomics (stronger than the device's scoped relaxed atomics).
#pragma once
#include <atomic>
#include <barrier>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <thread>
#include <type_traits>

#define UA_EMU 1
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __shared__ static thread_local_block_shared
#define __constant__
#define __launch_bounds__(x)

#define __HIP_MEMORY_SCOPE_SINGLETHREAD 1
#define __HIP_MEMORY_SCOPE_WAVEFRONT 2
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __HIP_MEMORY_SCOPE_SYSTEM 5

// Function-scope __shared__ variables are shared by every thread of the (single)
// emulated block: plain statics.
#define thread_local_block_shared

struct EmuDim {
    unsigned x = 0, y = 0, z = 0;
};

namespace emu {
struct Wave {
    std::barrier<> bar{64};
    uint64_t buf[64];
};
struct Block {
    std::barrier<> *bar;
    Wave *waves;
    unsigned nthreads;
};
inline thread_local unsigned tid = 0;
inline Block *block = nullptr;

inline uint64_t exchange(uint64_t v, unsigned src) {
    Wave &w = block->waves[tid >> 6];
    w.buf[tid & 63] = v;
    w.bar.arrive_and_wait();
    uint64_t r = w.buf[src & 63];
    w.bar.arrive_and_wait();
    return r;
}
}  // namespace emu

struct EmuThreadIdx {
    unsigned y = 0, z = 0;
    struct X {
        operator unsigned() const { return emu::tid; }
    } x;
};
inline EmuThreadIdx threadIdx;
struct EmuBlockDim {
    unsigned y = 1, z = 1;
    struct X {
        operator unsigned() const { return emu::block->nthreads; }
    } x;
};
inline EmuBlockDim blockDim;
inline EmuDim blockIdx{0, 0, 0};
inline EmuDim gridDim{1, 1, 1};

inline void __syncthreads() { emu::block->bar->arrive_and_wait(); }
inline unsigned __lane_id() { return emu::tid & 63; }

template <class T>
inline T __shfl(T v, int src, int width = 64) {
    (void)width;
    uint64_t r = emu::exchange((uint64_t)(std::make_unsigned_t<T>)v, (unsigned)src);
    return (T)r;
}
template <class T>
inline T __shfl_xor(T v, int m, int width = 64) {
    (void)width;
    uint64_t r = emu::exchange((uint64_t)(std::make_unsigned_t<T>)v, (emu::tid & 63) ^ (unsigned)m);
    return (T)r;
}
template <class T>
inline T __shfl_up(T v, unsigned d, int width = 64) {
    (void)width;
    unsigned l = emu::tid & 63;
    uint64_t r = emu::exchange((uint64_t)(std::make_unsigned_t<T>)v, l >= d ? l - d : l);
    return (T)r;
}
// Mask of the lanes whose predicate holds (every lane reads the whole exchange buffer).
inline uint64_t __ballot(int pred) {
    emu::Wave &w = emu::block->waves[emu::tid >> 6];
    w.buf[emu::tid & 63] = pred != 0;
    w.bar.arrive_and_wait();
    uint64_t m = 0;
    for (int i = 0; i < 64; ++i) m |= (w.buf[i] & 1u) << i;
    w.bar.arrive_and_wait();
    return m;
}
inline int __builtin_amdgcn_readlane(int v, int lane) {
    return (int)(uint32_t)emu::exchange((uint64_t)(uint32_t)v, (unsigned)lane);
}
inline int __builtin_amdgcn_readfirstlane(int v) {
    return (int)(uint32_t)emu::exchange((uint64_t)(uint32_t)v, 0u);
}
inline uint64_t __builtin_amdgcn_s_memrealtime() {
    return (uint64_t)std::chrono::steady_clock::now().time_since_epoch().count() / 10;
}
inline void __builtin_amdgcn_s_sleep(int) { std::this_thread::yield(); }
inline void __builtin_amdgcn_s_waitcnt(int) { std::atomic_thread_fence(std::memory_order_seq_cst); }
#define __builtin_amdgcn_fence(order, scope) std::atomic_thread_fence(std::memory_order_seq_cst)

template <class T, class U>
inline T __hip_atomic_load(const T *p, int, int) {
    return __atomic_load_n(p, __ATOMIC_SEQ_CST);
}
template <class T, class U = T>
inline void __hip_atomic_store(T *p, U v, int, int) {
    __atomic_store_n(p, (T)v, __ATOMIC_SEQ_CST);
}
template <class T>
inline T __hip_atomic_load(const T *p, int o, int s) {
    return __hip_atomic_load<T, T>(p, o, s);
}
template <class T, class U>
inline T __hip_atomic_fetch_add(T *p, U v, int, int) {
    return __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST);
}
template <class T, class U>
inline T __hip_atomic_fetch_or(T *p, U v, int, int) {
    return __atomic_fetch_or(p, (T)v, __ATOMIC_SEQ_CST);
}
template <class T, class U>
inline T __hip_atomic_fetch_min(T *p, U v, int, int) {
    T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while ((T)v < cur && !__atomic_compare_exchange_n(p, &cur, (T)v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return cur;
}
template <class T, class U>
inline T __hip_atomic_fetch_max(T *p, U v, int, int) {
    T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while ((T)v > cur && !__atomic_compare_exchange_n(p, &cur, (T)v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return cur;
}
template <class T, class U>
inline bool __hip_atomic_compare_exchange_strong(T *p, T *expected, U desired, int, int, int) {
    return __atomic_compare_exchange_n(p, expected, (T)desired, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
}
inline unsigned atomicAdd(unsigned *p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
