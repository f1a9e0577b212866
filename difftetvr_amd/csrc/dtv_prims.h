// Wave64 / workgroup primitives for gfx950: ballots, prefix scans, DPP reductions, look-back status words.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dtv {

constexpr int kWave = 64;

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (lane_id() == 0) ? 0ull : (~0ull >> (64 - lane_id())); }
__device__ __forceinline__ int popc64(uint64_t m) { return __popcll(m); }

// Element count of a launch sized for n_cap: n_dev (may be NULL) holds the actual count in device memory, read
// at run time and clamped to n_cap, so that a launch sequence can be recorded before the count is known.
__device__ __forceinline__ int64_t run_count(int64_t n_cap, const uint64_t* n_dev) {
    if (n_dev == nullptr) return n_cap;
    const uint64_t d = *n_dev;
    return d < (uint64_t)n_cap ? (int64_t)d : n_cap;
}

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Sum across the wave, result valid in every lane.  DPP row operations within 16-lane rows, then
// row broadcasts; all lanes must be active.
__device__ __forceinline__ float wave_sum(float v) {
    int x = __float_as_int(v);
    // quad_perm [1,0,3,2] and [2,3,0,1]: quad sums
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, true)));
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, true)));
    // row_half_mirror: sums of 8 lanes
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, true)));
    // row_mirror: sums of 16 lanes
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_mov_dpp(x, 0x140, 0xF, 0xF, true)));
    // combine the four rows through readlane (scalar broadcast)
    float r0 = __int_as_float(__builtin_amdgcn_readlane(x, 0));
    float r1 = __int_as_float(__builtin_amdgcn_readlane(x, 16));
    float r2 = __int_as_float(__builtin_amdgcn_readlane(x, 32));
    float r3 = __int_as_float(__builtin_amdgcn_readlane(x, 48));
    return (r0 + r1) + (r2 + r3);
}

// Exclusive scan over the threads of a workgroup of BLOCK threads (BLOCK multiple of 64, <= 1024).
template <int BLOCK>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp /* >= BLOCK/64 + 1 */,
                                                         uint32_t& total) {
    constexpr int W = BLOCK / 64;
    const int lane = lane_id();
    const int wid = threadIdx.x / 64;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) s_tmp[wid] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int w = 0; w < W; w++) {
            uint32_t t = s_tmp[w];
            s_tmp[w] = run;
            run += t;
        }
        s_tmp[W] = run;
    }
    __syncthreads();
    total = s_tmp[W];
    uint32_t r = s_tmp[wid] + inc - v;
    __syncthreads();
    return r;
}

// Look-back status words: 2 flag bits + 30 value bits.
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagInc = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1;

__device__ __forceinline__ void status_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t status_load(const uint32_t* p) {
    return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store64(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t status_load64(const uint64_t* p) {
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void atomic_add_f32(float* p, float v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace dtv
