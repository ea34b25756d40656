// Per-pixel rasterize gradient shared by the backward kernels.
#pragma once
#include "drp_common.h"

namespace drp {

// ----------------------------------------------------------------------------
// backward math of one pixel: d(u, v, z/w [, db]) -> d(x, y, z, w) of the corners
// ----------------------------------------------------------------------------
template <bool DB>
__device__ __forceinline__ void raster_pixel_grad(const TriVerts &v, float px, float py, float4 g, float4 gd,
                                                  float4 out[3]) {
    PixEval p;
    eval_pixel(v, px, py, p);
    float u = p.e[0] / p.s, vv = p.e[1] / p.s, zw = p.zn / p.wn;
    float gu = g.x, gv = g.y, gz = g.z;
    float gX[3] = {0.f, 0.f, 0.f}, gY[3] = {0.f, 0.f, 0.f}, gw[3] = {0.f, 0.f, 0.f}, gs = 0.f;
    if (DB) {
        float ex[3], ey[3];
        ex[0] = fma_(v.w[2], p.Y[1], -(v.w[1] * p.Y[2]));
        ex[1] = fma_(v.w[0], p.Y[2], -(v.w[2] * p.Y[0]));
        ex[2] = fma_(v.w[1], p.Y[0], -(v.w[0] * p.Y[1]));
        ey[0] = fma_(v.w[1], p.X[2], -(v.w[2] * p.X[1]));
        ey[1] = fma_(v.w[2], p.X[0], -(v.w[0] * p.X[2]));
        ey[2] = fma_(v.w[0], p.X[1], -(v.w[1] * p.X[0]));
        float sx = ex[0] + ex[1] + ex[2], sy = ey[0] + ey[1] + ey[2];
        float Aux = ex[0] - u * sx, Auy = ey[0] - u * sy, Avx = ex[1] - vv * sx, Avy = ey[1] - vv * sy;
        float a = gd.x / p.s, bb = gd.y / p.s, cc = gd.z / p.s, dd = gd.w / p.s;  // gd pre-scaled by kx / ky
        float gsx = -(a * u + cc * vv), gsy = -(bb * u + dd * vv);
        float gex[3] = {a + gsx, cc + gsx, gsx}, gey[3] = {bb + gsy, dd + gsy, gsy};
        gu += -(a * sx + bb * sy);
        gv += -(cc * sx + dd * sy);
        gs += -(a * Aux + bb * Auy + cc * Avx + dd * Avy) / p.s;
        gw[2] += gex[0] * p.Y[1]; gY[1] += gex[0] * v.w[2]; gw[1] -= gex[0] * p.Y[2]; gY[2] -= gex[0] * v.w[1];
        gw[0] += gex[1] * p.Y[2]; gY[2] += gex[1] * v.w[0]; gw[2] -= gex[1] * p.Y[0]; gY[0] -= gex[1] * v.w[2];
        gw[1] += gex[2] * p.Y[0]; gY[0] += gex[2] * v.w[1]; gw[0] -= gex[2] * p.Y[1]; gY[1] -= gex[2] * v.w[0];
        gw[1] += gey[0] * p.X[2]; gX[2] += gey[0] * v.w[1]; gw[2] -= gey[0] * p.X[1]; gX[1] -= gey[0] * v.w[2];
        gw[2] += gey[1] * p.X[0]; gX[0] += gey[1] * v.w[2]; gw[0] -= gey[1] * p.X[2]; gX[2] -= gey[1] * v.w[0];
        gw[0] += gey[2] * p.X[1]; gX[1] += gey[2] * v.w[0]; gw[1] -= gey[2] * p.X[0]; gX[0] -= gey[2] * v.w[1];
    }
    float inv_s = 1.f / p.s;
    float G0 = (gu * (1.f - u) + gv * (-vv)) * inv_s + gs + gz * (v.z[0] - zw * v.w[0]) / p.wn;
    float G1 = (gu * (-u) + gv * (1.f - vv)) * inv_s + gs + gz * (v.z[1] - zw * v.w[1]) / p.wn;
    float G2 = (gu * (-u) + gv * (-vv)) * inv_s + gs + gz * (v.z[2] - zw * v.w[2]) / p.wn;
    gX[0] += G2 * p.Y[1] - G1 * p.Y[2];
    gX[1] += G0 * p.Y[2] - G2 * p.Y[0];
    gX[2] += G1 * p.Y[0] - G0 * p.Y[1];
    gY[0] += G1 * p.X[2] - G2 * p.X[1];
    gY[1] += G2 * p.X[0] - G0 * p.X[2];
    gY[2] += G0 * p.X[1] - G1 * p.X[0];
    float gzw = gz / p.wn;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float ek = p.e[k];
        out[k] = make_float4(gX[k], gY[k], gzw * ek, gw[k] - px * gX[k] - py * gY[k] - gzw * zw * ek);
    }
}

__device__ __forceinline__ float4 db_grad_scaled(float4 gd, int H, int W) {
    float kx = 2.f / (float)W, ky = 2.f / (float)H;
    return make_float4(gd.x * kx, gd.y * ky, gd.z * kx, gd.w * ky);
}

}  // namespace drp
