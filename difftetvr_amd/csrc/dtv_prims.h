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

namespace detail {
// One DPP level: pairs (x, y) of n values; lanes whose BIT is clear keep x, the others keep y, each adding the
// value the partner lane sends (the partner's copy of the kept value).
template <int CTRL, int BIT, int N>
__device__ __forceinline__ void dpp_halve(const float (&in)[N], float (&out)[(N + 1) / 2], int lane) {
    const bool hi = (lane >> BIT) & 1;
#pragma unroll
    for (int k = 0; k < (N + 1) / 2; k++) {
        const float x = in[2 * k], y = in[2 * k + 1 < N ? 2 * k + 1 : 2 * k];
        const float keep = hi ? y : x, send = hi ? x : y;
        out[k] = keep + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(send), CTRL, 0xF, 0xF, true));
    }
}
}  // namespace detail

// Reduce-scatter of N <= 32 per-lane values within each row of 16 lanes (each row reduces its own values):
// lane L returns in out[k] the row's sum of value 16 k + row_scatter_index(L) (meaningful when below N).  Each
// level pairs the values and halves them across one row-lane bit (3 and 2 by the DPP row mirror and half-row
// mirror, 1 and 0 by quad permutations): about 3 VALU instructions per pair and level, where a full row sum of
// every value costs about 8 each.  All lanes must be active.
__device__ __forceinline__ int row_scatter_index(int lane) {
    return ((lane >> 3) & 1) | ((lane >> 1) & 2) | ((lane << 1) & 4) | ((lane << 3) & 8);
}

template <int N>
__device__ __forceinline__ void row_reduce_scatter(const float (&v)[N], float (&out)[(N + 15) / 16]) {
    static_assert(N >= 1 && N <= 32, "row_reduce_scatter: 1 to 32 values");
    constexpr int N1 = (N + 1) / 2, N2 = (N1 + 1) / 2, N3 = (N2 + 1) / 2;
    const int lane = lane_id();
    float a[N1], b[N2], c[N3];
    detail::dpp_halve<0x140, 3>(v, a, lane);    // row_mirror: lane i <-> 15 - i
    detail::dpp_halve<0x141, 2>(a, b, lane);    // row_half_mirror: lane i <-> 7 - i
    detail::dpp_halve<0x4E, 1>(b, c, lane);     // quad_perm [2,3,0,1]: lane i <-> i ^ 2
    detail::dpp_halve<0xB1, 0>(c, out, lane);   // quad_perm [1,0,3,2]: lane i <-> i ^ 1
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
