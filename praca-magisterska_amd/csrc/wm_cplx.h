// Device-side complex helpers and the bounded cross-workgroup wait.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace wm {

template <typename T> struct CT;
template <> struct CT<double> { using V = double2; };
template <> struct CT<float> { using V = float2; };

// Spin limit for device-side waits: with the s_sleep below this is well
// over a second, far above any legitimate wait, and every waiting wave
// leaves once any wave has recorded a timeout in the error word.
constexpr long kSpinLimit = 1L << 24;

// Arrive at and wait on a counter shared by `members` workgroups.  `target`
// is the counter value once everyone has arrived at this barrier instance.
// Thread 0 of the block does the atomics; the block is fenced on both sides.
__device__ inline void cluster_barrier(int *cnt, int target, int *err)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        long it = 0;
        while (__hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
            if (++it > kSpinLimit) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Wait until *flag >= want (flag written by an earlier-dispatched workgroup
// of the same grid).  Bounded like cluster_barrier.
__device__ inline void wait_flag(const int *flag, int want, int *err)
{
    if (threadIdx.x == 0) {
        long it = 0;
        while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < want) {
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
            if (++it > kSpinLimit) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Publish *flag = value after this block's prior global stores.
__device__ inline void set_flag(int *flag, int value)
{
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(flag, value, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Smith (1962) reciprocal of a complex number.
__device__ inline double2 crecip(double re, double im)
{
    if (fabs(re) >= fabs(im)) {
        const double t = im / re, d = re + im * t;
        return make_double2(1.0 / d, -t / d);
    }
    const double t = re / im, d = im + re * t;
    return make_double2(t / d, -1.0 / d);
}

}  // namespace wm
