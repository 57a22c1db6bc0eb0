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

// Tagged 8-byte packets {32-bit tag, 32-bit payload} for data handed between
// workgroups of one grid: a reader polls a packet until its tag matches, so
// neither a counter nor a release/acquire fence (an L2 write-back and
// invalidate on gfx950) is needed.  8-byte loads and stores are single-copy
// atomic; agent-scope relaxed atomics reach the device coherence point.
__device__ inline void put_pkt(uint64_t *p, uint32_t tag, uint32_t v)
{
    __hip_atomic_store(p, ((uint64_t)tag << 32) | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint64_t ld_pkt(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded-wait bookkeeping for a polling lane: true if the lane must stop
// (the error word is set, or this lane ran out of spins and set it).
__device__ inline bool spin_give_up(long &it, int *err)
{
    if ((++it & 255) == 0) {
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return true;
        if (it > kSpinLimit) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return true;
        }
    }
    __builtin_amdgcn_s_sleep(1);
    return false;
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
