/*
 * lt_film.hip - film post-processing: resolve and tone mapping on the GPU.
 *
 * The reference resolves in the visualizer (visualizer.cpp:355-370: pixel =
 * sum / float(sampleCount)) and tone-maps in a fragment shader
 * (final_resolve_ps.glsl:24-44): color = exposure * (rgb, 1), then clamp to
 * [0, 1] or the filmic curve x = max(0, c - 0.004),
 * (x (6.2x + 0.5)) / (x (6.2x + 1.7) + 0.06) with alpha 1, clamped.
 */
#include "lt_internal.h"

namespace lt {

namespace {

__global__ void k_tonemap(const float *sum, const uint32_t *count, uint32_t n, int op, float exposure, float4 *out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t c = count[i];
        float r = 0.0f, g = 0.0f, b = 0.0f;
        if (c) {
            const float fc = (float)c;
            r = sum[3 * (size_t)i] / fc;
            g = sum[3 * (size_t)i + 1] / fc;
            b = sum[3 * (size_t)i + 2] / fc;
        }
        float4 v = make_float4(exposure * r, exposure * g, exposure * b, exposure * 1.0f);
        if (op == 1) {
            float x[3] = {fmaxf(0.0f, v.x - 0.004f), fmaxf(0.0f, v.y - 0.004f), fmaxf(0.0f, v.z - 0.004f)};
            float y[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) y[k] = (x[k] * (6.2f * x[k] + 0.5f)) / (x[k] * (6.2f * x[k] + 1.7f) + 0.06f);
            v = make_float4(y[0], y[1], y[2], 1.0f);
        }
        v.x = fminf(fmaxf(v.x, 0.0f), 1.0f);
        v.y = fminf(fmaxf(v.y, 0.0f), 1.0f);
        v.z = fminf(fmaxf(v.z, 0.0f), 1.0f);
        v.w = fminf(fmaxf(v.w, 0.0f), 1.0f);
        out[i] = v;
    }
}

} // namespace

int launch_tonemap(hipStream_t st, const float *sum, const uint32_t *count, uint32_t n, int op, float exposure,
                   float *rgba, int num_cus) {
    if (n == 0) return LT_OK;
    uint32_t blocks = std::min<uint32_t>((n + 255) / 256, (uint32_t)std::max(1, num_cus) * 16);
    hipLaunchKernelGGL(k_tonemap, dim3(blocks), dim3(256), 0, st, sum, count, n, op, exposure, (float4 *)rgba);
    LT_HIP(hipGetLastError());
    return LT_OK;
}

} // namespace lt
