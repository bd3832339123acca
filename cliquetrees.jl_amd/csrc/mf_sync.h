// mf_sync.h — the bounded waits of the persistent kernels (device code only).
//
// A task waits on counters that tasks with smaller tickets set, so every
// wait ends in a correct launch.  The bound is a watchdog for a corrupt plan
// or an engine bug: a wait gives up only after `limit` ticks of the constant
// wall clock (s_memrealtime; the host converts seconds with the device's
// wall-clock rate) during which neither the awaited counter nor the launch's
// ticket counter moved.  The ticket counter advances whenever any workgroup
// finishes a task and takes the next ticket, so a legitimately long task
// (a wide update tile of a huge front) never trips it while the rest of the
// launch makes progress, and a long wait behind such a task restarts its
// clock whenever the awaited counter advances.  A wait that gives up sets
// the hang flag, which makes every other wait return, and the host reports
// MF_EHIP.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mf {

__device__ __forceinline__ int32_t sync_ld(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t sync_ld(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Progress watchdog state of one waiter.
struct Watch {
    unsigned long long t0 = 0;
    uint32_t tick = 0;
    int32_t seen = 0;
    bool armed = false;

    // Called every few hundred polls with the awaited counter's value v:
    // true when the wait must give up.
    __device__ __forceinline__ bool expired(int32_t v, const uint32_t* ticket, unsigned long long limit) {
        const uint32_t tk = sync_ld(ticket);
        const unsigned long long now = wall_clock64();
        if (!armed || tk != tick || v != seen) {
            armed = true;
            tick = tk;
            seen = v;
            t0 = now;
            return false;
        }
        return now - t0 > limit;
    }
};

// Records the first wait of a launch that gave up: hang[0] = 1, hang[1] =
// the target, hang[2] = the value seen, hang[3] = the waiter's ticket (-1:
// unknown), hang[4..5] = the awaited address (the host names the counter).
__device__ __forceinline__ void give_up(int32_t* hang, const int32_t* p, int32_t target, int32_t v,
                                        int32_t ticket) {
    if (atomicCAS(hang, 0, 1) == 0) {
        const unsigned long long a = (unsigned long long)p;
        __hip_atomic_store(hang + 1, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(hang + 2, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(hang + 3, ticket, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(hang + 4, (int32_t)(a & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(hang + 5, (int32_t)(a >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Wait until *p >= target; returns the value seen (>= target unless the
// wait gave up or another wait did).
__device__ __forceinline__ int32_t bounded_wait(const int32_t* p, int32_t target, int32_t* hang,
                                                const uint32_t* ticket, unsigned long long limit,
                                                int sleep) {
    int32_t v = sync_ld(p);
    Watch w;
    for (int spins = 1; v < target; ++spins) {
        if ((spins & 255) == 0) {
            // once any wait has given up, every other wait returns at once
            if (sync_ld(hang)) break;
            if (w.expired(v, ticket, limit)) {
                give_up(hang, p, target, v, -1);
                break;
            }
        }
        if (sleep > 1) __builtin_amdgcn_s_sleep(2);
        else __builtin_amdgcn_s_sleep(1);
        v = sync_ld(p);
    }
    return v;
}

}  // namespace mf
