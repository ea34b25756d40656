// Clip-space transform of the session's vertices (diffrp's clip_space,
// diffrp/rendering/surface_deferred.py:229-238 through transform_point,
// diffrp/utils/shader_ops.py:123-136): pos[b, v] = VP_b [x_v, y_v, z_v, 1]^T
// for n views at once, and its backward to the world vertices,
// grad_xyz[v] = sum over views b (in view order) of grad_pos[b, v] VP_b[:, 0:3].
// Each row is one fma chain, m0 x + (m1 y + (m2 z + m3)), evaluated the same
// way for every view.
#include "drp_common.h"
#include "drp_kernels.h"

namespace drp {

__global__ void __launch_bounds__(256) k_clip_fwd(const float *__restrict__ xyz, int V, const float *__restrict__ VP,
                                                  float4 *__restrict__ pos) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (v >= V) return;
    const float x = xyz[3 * (long long)v], y = xyz[3 * (long long)v + 1], z = xyz[3 * (long long)v + 2];
    const float *m = VP + 16 * b;  // row-major 4 x 4
    float r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = fma_(m[4 * i], x, fma_(m[4 * i + 1], y, fma_(m[4 * i + 2], z, m[4 * i + 3])));
    pos[(long long)b * V + v] = make_float4(r[0], r[1], r[2], r[3]);
}

// eight lanes per vertex, lane s taking views s, s + 8, ... (summed in view
// order), then the lanes' sums added in a fixed tree: a grid of 8 V lanes
// keeps the GPU's memory pipes busy where one lane per vertex was a fraction
// of a wave per SIMD
constexpr int kClipLanes = 8;
__global__ void __launch_bounds__(256) k_clip_bwd(const float4 *__restrict__ grad_pos, int V, int n,
                                                  const float *__restrict__ VP, float *__restrict__ grad_xyz) {
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int v = (int)(gt / kClipLanes), s = (int)(gt % kClipLanes);
    if (v >= V) return;  // all lanes of v together
    float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll 8
    for (int b = s; b < n; b += kClipLanes) {  // unrolled: eight views' loads in flight
        const float4 g = grad_pos[(long long)b * V + v];
        const float *m = VP + 16 * b;
        // column k of VP_b: m[k], m[4 + k], m[8 + k], m[12 + k]
        gx += fma_(g.x, m[0], fma_(g.y, m[4], fma_(g.z, m[8], g.w * m[12])));
        gy += fma_(g.x, m[1], fma_(g.y, m[5], fma_(g.z, m[9], g.w * m[13])));
        gz += fma_(g.x, m[2], fma_(g.y, m[6], fma_(g.z, m[10], g.w * m[14])));
    }
    for (int o = 1; o < kClipLanes; o <<= 1) {  // (a + b on one lane, b + a on the other: the same sum)
        gx += __shfl_xor(gx, o, 64);
        gy += __shfl_xor(gy, o, 64);
        gz += __shfl_xor(gz, o, 64);
    }
    if (s != 0) return;
    float *o = grad_xyz + 3 * (long long)v;
    o[0] = gx;
    o[1] = gy;
    o[2] = gz;
}

int launch_clip_fwd(const float *xyz, int V, const float *VP, int n, float *pos, hipStream_t st) {
    if (V <= 0 || n <= 0) return 0;
    hipLaunchKernelGGL(k_clip_fwd, dim3((V + 255) / 256, n), dim3(256), 0, st, xyz, V, VP, (float4 *)pos);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_clip_bwd(const float *grad_pos, int V, const float *VP, int n, float *grad_xyz, hipStream_t st) {
    if (V <= 0 || n <= 0) return 0;
    const long long lanes = (long long)V * kClipLanes;
    hipLaunchKernelGGL(k_clip_bwd, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, st, (const float4 *)grad_pos, V,
                       n, VP, grad_xyz);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace drp
