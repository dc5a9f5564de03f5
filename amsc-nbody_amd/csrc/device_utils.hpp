// device_utils.hpp -- wave64 primitives for gfx950.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nbh {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Number of set bits of `mask` below this lane (v_mbcnt_lo/hi).
__device__ __forceinline__ unsigned lanes_below(unsigned long long mask)
{
    return __builtin_amdgcn_mbcnt_hi(
        static_cast<unsigned>(mask >> 32),
        __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u));
}

// Lanes of the wave whose 8-bit digit equals this lane's (8 ballots).
__device__ __forceinline__ unsigned long long match_digit8(unsigned d,
                                                           bool valid)
{
    unsigned long long m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long v = __ballot(bit);
        m &= bit ? v : ~v;
    }
    return m;
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = fmin(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = fmax(v, __shfl_xor(v, o));
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T v)
{
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T y = __shfl_up(v, o);
        if (lane >= o)
            v += y;
    }
    return v;
}

// Block-wide exclusive sum over 256 threads (4 waves); `tmp` holds 4 words
// of LDS.  Returns this thread's exclusive prefix and the block total.
__device__ __forceinline__ unsigned block256_exclusive_sum(unsigned v,
                                                           unsigned* tmp,
                                                           unsigned& total)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const unsigned inc = wave_inclusive_sum(v);
    if (lane == 63)
        tmp[wave] = inc;
    __syncthreads();
    unsigned off = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const unsigned t = tmp[w];
        if (w < wave)
            off += t;
        total += t;
    }
    __syncthreads();
    return off + inc - v;
}

// Adds one per active lane to the LDS counter cnt[key], one atomic per
// distinct key in the wave; returns each active lane's slot (the counter's
// value before its own increment, lanes of one key in lane order).  Every
// lane of the wave must call it (inactive lanes with active = false).
__device__ __forceinline__ int wave_claim(int* cnt, int key, bool active)
{
    const int lane = static_cast<int>(threadIdx.x & 63);
    int slot = 0;
    unsigned long long todo = __ballot(active);
    while (todo != 0ull) {
        const int leader = __ffsll(static_cast<long long>(todo)) - 1;
        const int k0 = __shfl(key, leader);
        const unsigned long long same = __ballot(active && key == k0);
        int base = 0;
        if (lane == leader)
            base = atomicAdd(&cnt[k0], __popcll(same));
        base = __shfl(base, leader);
        if (active && key == k0)
            slot = base + __popcll(same & ((1ull << lane) - 1ull));
        todo &= ~same;
    }
    return slot;
}

// Agent-scope atomics for cross-workgroup protocols (look-back scans,
// arrival counters).  gfx950 has one L2 per XCD: agent scope makes these
// coherent across all eight.
template <typename T>
__device__ __forceinline__ T load_agent(const T* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ __forceinline__ void store_agent(T* p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int fetch_add_acq_rel(int* p, int v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL,
                                  __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void raise_error(int* err, int code)
{
    atomicCAS(err, 0, code);
}

} // namespace nbh
