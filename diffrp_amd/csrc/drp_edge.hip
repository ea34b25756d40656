// Backward of the edge-gradient visibility operator ("Rasterized Edge
// Gradients", diffrp/rendering/surface_deferred.py:120-169): one pass over the
// pixels of a bottom-row-first image.
//
// For each vertical pair (a = row i, b = row i+1) and horizontal pair
// (a = column j, b = column j+1) whose triangle ids differ:
//   e = s * 0.5 * <g_a + g_b, c_a - c_b>   (s = -1 vertical, +1 horizontal)
//   a receives e when b is empty or a's fragment lies in b's screen triangle,
//   b receives e when a is empty or b's fragment lies in a's screen triangle.
// grad_frag_xy[p] = (sum of horizontal receipts * W/2, sum of vertical * H/2).
// "Lies in" is the sign test of the three 2D edge determinants (on an edge
// counts as inside).
#include "drp_common.h"
#include "drp_kernels.h"

namespace drp {

__device__ __forceinline__ float sign2d(float2 p1, float2 p2, float2 p3) {
    return (p1.x - p3.x) * (p2.y - p3.y) - (p2.x - p3.x) * (p1.y - p3.y);
}

__device__ __forceinline__ bool point_in_tri(float2 p, const float2 *t3) {
    float d1 = sign2d(p, t3[0], t3[1]);
    float d2 = sign2d(p, t3[1], t3[2]);
    float d3 = sign2d(p, t3[2], t3[0]);
    bool neg = d1 < 0.f || d2 < 0.f || d3 < 0.f;
    bool pos = d1 > 0.f || d2 > 0.f || d3 > 0.f;
    return !(neg && pos);
}

// receipt of pixel `self` from its pair with `other`; `self_is_a` orders the pair
template <int C>
__device__ __forceinline__ float edge_receipt(const EdgeLaunch &L, long long self, long long other, bool self_is_a,
                                              float s) {
    int ts = float_to_triidx(L.rast[4 * self + 3]);
    int to = float_to_triidx(L.rast[4 * other + 3]);
    if (ts == to) return 0.f;
    long long a = self_is_a ? self : other, b = self_is_a ? other : self;
    float dot = 0.f;
    const float *ga = L.grad_out + a * C, *gb = L.grad_out + b * C;
    const float *ca = L.rgb + a * C, *cb = L.rgb + b * C;
#pragma unroll
    for (int k = 0; k < C; ++k) dot += (ga[k] + gb[k]) * (ca[k] - cb[k]);
    float e = s * 0.5f * dot;
    // self receives when the other pixel is empty or self's fragment lies in the other's triangle
    if (to == 0) return e;
    if ((unsigned)(to - 1) >= (unsigned)L.F) return 0.f;
    float2 p = make_float2(L.frag_xy[2 * self], L.frag_xy[2 * self + 1]);
    const float2 *tri = reinterpret_cast<const float2 *>(L.tris_xy) + 3 * (long long)(to - 1);
    float2 t3[3] = {tri[0], tri[1], tri[2]};
    return point_in_tri(p, t3) ? e : 0.f;
}

template <int C>
__global__ void __launch_bounds__(256) k_edge_grad_bwd(EdgeLaunch L) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long npx = (long long)L.H * L.W;
    if (g >= npx) return;
    int r = (int)(g / L.W), c = (int)(g % L.W);
    float gx = 0.f, gy = 0.f;
    if (c + 1 < L.W) gx += edge_receipt<C>(L, g, g + 1, true, 1.f);
    if (c > 0) gx += edge_receipt<C>(L, g, g - 1, false, 1.f);
    if (r + 1 < L.H) gy += edge_receipt<C>(L, g, g + L.W, true, -1.f);
    if (r > 0) gy += edge_receipt<C>(L, g, g - L.W, false, -1.f);
    L.grad_frag_xy[2 * g] = gx * ((float)L.W * 0.5f);
    L.grad_frag_xy[2 * g + 1] = gy * ((float)L.H * 0.5f);
}

int launch_edge_grad_bwd(const EdgeLaunch &L, hipStream_t st) {
    long long npx = (long long)L.H * L.W;
    if (npx == 0) return 0;
    dim3 grid((unsigned)((npx + 255) / 256)), block(256);
    switch (L.C) {
        case 1: hipLaunchKernelGGL(k_edge_grad_bwd<1>, grid, block, 0, st, L); break;
        case 3: hipLaunchKernelGGL(k_edge_grad_bwd<3>, grid, block, 0, st, L); break;
        case 4: hipLaunchKernelGGL(k_edge_grad_bwd<4>, grid, block, 0, st, L); break;
        default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace drp
