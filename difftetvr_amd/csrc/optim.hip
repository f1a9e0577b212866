// Fused Adam step over several parameter groups, and the MSE image loss with its adjoint, for gfx950.
//
// Adam follows torch.optim.Adam's single-tensor update (the optimiser pytests/train.py:318,351 uses):
//   m = m + (1-b1)(g - m);  v = b2 v + (1-b2) g^2;  p -= (lr / (1-b1^t)) * m / (sqrt(v)/sqrt(1-b2^t) + eps)
// then the colour clamp p -= min(p, 0) (train.py:352-356) when requested, and optionally zeroes g for the next
// step (optimizer.zero_grad(set_to_none=False), train.py:342).  Elementwise and HBM-bound: 16 B read + 12 B
// written per parameter (+4 B when the gradient is zeroed); float4 vector access.
#include "dtv_kernels.h"
#include "dtv_prims.h"

namespace dtv {

constexpr int kMaxAdamGroups = 8;

struct AdamLaunch {
    AdamGroupDev g[kMaxAdamGroups];
    int64_t block_start[kMaxAdamGroups + 1];  // first block of each group (in float4 units of work)
    int ngroups;
    const int64_t* step_dev;  // non-NULL: steps taken so far, in device memory (the bias corrections use +1)
};

__device__ __forceinline__ float adam_one(float p, float gr, float& m, float& v, const AdamGroupDev& G) {
    m = m + G.w1 * (gr - m);
    v = v * G.beta2 + G.w2 * (gr * gr);
    const float denom = sqrtf(v) / G.bc2_sqrt + G.eps;
    p = p + (-G.step_size) * (m / denom);
    if (G.clamp_nonneg) p = p - fminf(p, 0.0f);
    return p;
}

__global__ __launch_bounds__(256) void k_adam(AdamLaunch A) {
    int gi = 0;
    while (gi + 1 < A.ngroups && (int64_t)blockIdx.x >= A.block_start[gi + 1]) gi++;
    AdamGroupDev G = A.g[gi];
    if (A.step_dev != nullptr) {
        // bias corrections of step t = taken + 1, in double as on the host
        const double t = (double)(*A.step_dev + 1);
        G.step_size = (float)(G.lr / (1.0 - pow(G.beta1, t)));
        G.bc2_sqrt = (float)sqrt(1.0 - pow(G.beta2d, t));
    }
    const int64_t i4 = ((int64_t)blockIdx.x - A.block_start[gi]) * 256 + threadIdx.x;
    const int64_t n4 = G.n / 4;
    if (i4 < n4) {
        float4 p = reinterpret_cast<float4*>(G.param)[i4];
        float4 g = reinterpret_cast<float4*>(G.grad)[i4];
        float4 m = reinterpret_cast<float4*>(G.m)[i4];
        float4 v = reinterpret_cast<float4*>(G.v)[i4];
        p.x = adam_one(p.x, g.x, m.x, v.x, G);
        p.y = adam_one(p.y, g.y, m.y, v.y, G);
        p.z = adam_one(p.z, g.z, m.z, v.z, G);
        p.w = adam_one(p.w, g.w, m.w, v.w, G);
        reinterpret_cast<float4*>(G.param)[i4] = p;
        reinterpret_cast<float4*>(G.m)[i4] = m;
        reinterpret_cast<float4*>(G.v)[i4] = v;
        if (G.zero_grad) reinterpret_cast<float4*>(G.grad)[i4] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    // tail (n not a multiple of 4)
    const int64_t tail = G.n - n4 * 4;
    if (i4 == n4 && tail > 0) {
        for (int64_t i = n4 * 4; i < G.n; i++) {
            float m = G.m[i], v = G.v[i];
            G.param[i] = adam_one(G.param[i], G.grad[i], m, v, G);
            G.m[i] = m;
            G.v[i] = v;
            if (G.zero_grad) G.grad[i] = 0.f;
        }
    }
}

__global__ void k_step_increment(int64_t* step) { *step += 1; }

void launch_adam(const AdamGroupDev* groups, int ngroups, int64_t* step_dev, hipStream_t s) {
    AdamLaunch A{};
    A.ngroups = ngroups;
    A.step_dev = step_dev;
    int64_t blocks = 0;
    for (int i = 0; i < ngroups; i++) {
        A.g[i] = groups[i];
        A.block_start[i] = blocks;
        const int64_t work = groups[i].n / 4 + 1;  // +1 covers the scalar tail
        blocks += (work + 255) / 256;
    }
    A.block_start[ngroups] = blocks;
    if (blocks > 0) hipLaunchKernelGGL(k_adam, dim3((unsigned)blocks), dim3(256), 0, s, A);
    if (step_dev != nullptr) hipLaunchKernelGGL(k_step_increment, dim3(1), dim3(1), 0, s, step_dev);
}

// dimg = 2 (img - tgt) / denom ; loss += sum((img - tgt)^2) / denom  (denom = n: the mean; denom = n / items:
// the sum of the items' means)
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ img, const float* __restrict__ tgt, int64_t n,
                                              int64_t denom, float* __restrict__ dimg, float* __restrict__ loss) {
    const float inv_n = 1.0f / (float)denom;
    float acc = 0.f;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float d = img[i] - tgt[i];
        dimg[i] = 2.0f * d * inv_n;
        acc += d * d;
    }
    if (loss != nullptr) {
        // one atomic per block: the loss address is shared by the whole grid
        __shared__ float s_part[4];
        const float ws = wave_sum(acc);
        if (lane_id() == 0) s_part[threadIdx.x / 64] = ws;
        __syncthreads();
        if (threadIdx.x == 0) atomic_add_f32(loss, ((s_part[0] + s_part[1]) + (s_part[2] + s_part[3])) * inv_n);
    }
}

void launch_mse(const float* img, const float* tgt, int64_t n, int64_t denom, float* dimg, float* loss,
                hipStream_t s) {
    if (loss != nullptr) fill_u32(reinterpret_cast<uint32_t*>(loss), 0u, 1, s);
    int64_t blocks = (n + 255) / 256;
    if (blocks > 512) blocks = 512;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_mse, dim3((unsigned)blocks), dim3(256), 0, s, img, tgt, n, denom, dimg, loss);
}

}  // namespace dtv
