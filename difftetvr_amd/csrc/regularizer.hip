// Tet-quality regulariser gradient on gfx950 (Data/Shaders/Optimizer/TetRegularizer.glsl:108-166,
// Renderer/TetRegularizerPass.cpp:91-95).
//
// Per tet: q = 6 sqrt(2) V / l_rms^3 (Parthasarathy et al. 1991) with V = -dot(cross(p1-p0, p2-p0), p3-p0) / 6
// and l_rms^2 = S / 6, S = sum of the 6 squared edge lengths; the loss term is the mirrored softplus
// (lambda / beta) log(1 + exp(-beta q)), whose derivative is dL/dq = -lambda / (1 + exp(beta q)).
// Written as q = -12 sqrt(3) D / S^{3/2} with the triple product D = dot(cross(p1-p0, p2-p0), p3-p0):
//   dq/dp = (12 sqrt(3) * dD/dp * S - 18 sqrt(3) * D * dS/dp) / S^{5/2} * (-1)
// and, as in the reference, S^{5/2} is clamped below at 1e-5 (this damps the gradient of small tets).
// Gradients are accumulated (+=) into grad_positions with one float atomic per value.
#include "dtv_kernels.h"
#include "dtv_prims.h"

namespace dtv {

__global__ __launch_bounds__(256) void k_tet_regularizer(const int32_t* __restrict__ cells,
                                                          const float* __restrict__ pos, int64_t T, float lambda,
                                                          float beta, float* __restrict__ grad_pos) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= T) return;
    const int4 idx = *reinterpret_cast<const int4*>(cells + 4 * t);
    const int32_t vid[4] = {idx.x, idx.y, idx.z, idx.w};
    float3_ p[4];
#pragma unroll
    for (int i = 0; i < 4; i++) p[i] = load_pos(pos, vid[i]);
    const float3_ a = sub(p[1], p[0]), b = sub(p[2], p[0]), c = sub(p[3], p[0]);
    const float D = dot(cross(a, b), c);
    float S = 0.f;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = i + 1; j < 4; j++) {
            const float3_ e = sub(p[i], p[j]);
            S += dot(e, e);
        }
    const float k12 = 20.784609690826528f;   // 12 sqrt(3)
    const float q = -k12 * D / (S * sqrtf(S));
    const float dLdq = -lambda / (1.0f + expf(beta * q));
    const float S52 = fmaxf(S * S * sqrtf(S), 1e-5f);
    // dD/dp: p1 -> cross(b, c), p2 -> cross(c, a), p3 -> cross(a, b), p0 -> minus their sum
    float3_ dD[4];
    dD[1] = cross(b, c);
    dD[2] = cross(c, a);
    dD[3] = cross(a, b);
    dD[0] = scale(add(add(dD[1], dD[2]), dD[3]), -1.0f);
    float3_ sum = add(add(p[0], p[1]), add(p[2], p[3]));
#pragma unroll
    for (int i = 0; i < 4; i++) {
        // dS/dp_i = 2 (4 p_i - sum)
        const float3_ dS = scale(sub(scale(p[i], 4.0f), sum), 2.0f);
        const float3_ g = scale(sub(scale(dD[i], k12 * S), scale(dS, 1.5f * k12 * D)), -dLdq / S52);
        const int64_t v = vid[i];
        atomic_add_f32(&grad_pos[3 * v + 0], g.x);
        atomic_add_f32(&grad_pos[3 * v + 1], g.y);
        atomic_add_f32(&grad_pos[3 * v + 2], g.z);
    }
}

void launch_tet_regularizer(const int32_t* cells, const float* pos, int64_t T, float lambda, float beta,
                            float* grad_pos, hipStream_t s) {
    if (T <= 0) return;
    hipLaunchKernelGGL(k_tet_regularizer, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, s, cells, pos, T, lambda,
                       beta, grad_pos);
}

}  // namespace dtv
