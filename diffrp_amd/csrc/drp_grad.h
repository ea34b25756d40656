// Per-pixel rasterize gradient shared by the backward kernels.
#pragma once
#include "drp_common.h"

namespace drp {

// ----------------------------------------------------------------------------
// backward math of one pixel: d(u, v, z/w [, db]) -> d(x, y, z, w) of the corners
//
// Evaluated in double precision from the float inputs.  The derivatives of
// thin triangles seen edge-on (s small against the e_k) lose ~|e|/s of their
// relative precision in the e_k - u s style cancellations; in float that is
// above 1e-3 at the silhouette of a dense mesh, in double it is negligible.
// The forward stays in float (bit-exact contract); only the gradient is
// computed more accurately.  MI355X runs double FMAs at the rate of unpacked
// float ones, and this code is gather-bound.
// ----------------------------------------------------------------------------
__device__ __forceinline__ double dfma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// 1/x in double to within about an ulp: the hardware reciprocal estimate
// (~26 bits) refined by two Newton steps, the refinement that the compiler's
// correctly rounded division also uses, without its scaling and fix-up for
// exponents near the ends of the range (s and wn of a covered pixel are far
// from them; x = 0 gives a non-finite result, as the division does).
__device__ __forceinline__ double drcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = dfma(-x, r, 1.0);
    r = dfma(r, e, r);
    e = dfma(-x, r, 1.0);
    return dfma(r, e, r);
}

// NDC centre of pixel column c (row r) in double: (2c + 1 - W) / W
__device__ __forceinline__ double pix_ndc_d(int c, int W) { return (double)(2 * c + 1 - W) / (double)W; }

// px, py: the pixel's NDC centre, pix_ndc_d(c, W), pix_ndc_d(r, H)
template <bool DB>
__device__ __forceinline__ void raster_pixel_grad_at(const TriVerts &v, double px, double py, float4 g, float4 gd,
                                                     float4 out[3]) {
    double X[3], Y[3], wk[3], zk[3], e[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        wk[k] = v.w[k];
        zk[k] = v.z[k];
        X[k] = dfma(-wk[k], px, (double)v.x[k]);
        Y[k] = dfma(-wk[k], py, (double)v.y[k]);
    }
    e[0] = dfma(X[1], Y[2], -(X[2] * Y[1]));
    e[1] = dfma(X[2], Y[0], -(X[0] * Y[2]));
    e[2] = dfma(X[0], Y[1], -(X[1] * Y[0]));
    const double s = (e[0] + e[1]) + e[2];
    const double zn = dfma(e[0], zk[0], dfma(e[1], zk[1], e[2] * zk[2]));
    const double wn = dfma(e[0], wk[0], dfma(e[1], wk[1], e[2] * wk[2]));
    const double inv_s = drcp(s), inv_wn = drcp(wn);
    const double u = e[0] * inv_s, vv = e[1] * inv_s, zw = zn * inv_wn;
    double gu = g.x, gv = g.y, gz = g.z;
    double gX[3] = {0.0, 0.0, 0.0}, gY[3] = {0.0, 0.0, 0.0}, gw[3] = {0.0, 0.0, 0.0}, gs = 0.0;
    if (DB) {
        double ex[3], ey[3];
        ex[0] = dfma(wk[2], Y[1], -(wk[1] * Y[2]));
        ex[1] = dfma(wk[0], Y[2], -(wk[2] * Y[0]));
        ex[2] = dfma(wk[1], Y[0], -(wk[0] * Y[1]));
        ey[0] = dfma(wk[1], X[2], -(wk[2] * X[1]));
        ey[1] = dfma(wk[2], X[0], -(wk[0] * X[2]));
        ey[2] = dfma(wk[0], X[1], -(wk[1] * X[0]));
        double sx = ex[0] + ex[1] + ex[2], sy = ey[0] + ey[1] + ey[2];
        double Aux = ex[0] - u * sx, Auy = ey[0] - u * sy, Avx = ex[1] - vv * sx, Avy = ey[1] - vv * sy;
        // gd holds d rast_db pre-scaled by (2/W, 2/H)
        double a = gd.x * inv_s, bb = gd.y * inv_s, cc = gd.z * inv_s, dd = gd.w * inv_s;
        double gsx = -(a * u + cc * vv), gsy = -(bb * u + dd * vv);
        double gex[3] = {a + gsx, cc + gsx, gsx}, gey[3] = {bb + gsy, dd + gsy, gsy};
        gu += -(a * sx + bb * sy);
        gv += -(cc * sx + dd * sy);
        gs += -(a * Aux + bb * Auy + cc * Avx + dd * Avy) * inv_s;
        gw[2] += gex[0] * Y[1]; gY[1] += gex[0] * wk[2]; gw[1] -= gex[0] * Y[2]; gY[2] -= gex[0] * wk[1];
        gw[0] += gex[1] * Y[2]; gY[2] += gex[1] * wk[0]; gw[2] -= gex[1] * Y[0]; gY[0] -= gex[1] * wk[2];
        gw[1] += gex[2] * Y[0]; gY[0] += gex[2] * wk[1]; gw[0] -= gex[2] * Y[1]; gY[1] -= gex[2] * wk[0];
        gw[1] += gey[0] * X[2]; gX[2] += gey[0] * wk[1]; gw[2] -= gey[0] * X[1]; gX[1] -= gey[0] * wk[2];
        gw[2] += gey[1] * X[0]; gX[0] += gey[1] * wk[2]; gw[0] -= gey[1] * X[2]; gX[2] -= gey[1] * wk[0];
        gw[0] += gey[2] * X[1]; gX[1] += gey[2] * wk[0]; gw[1] -= gey[2] * X[0]; gX[0] -= gey[2] * wk[1];
    }
    double gzw = gz * inv_wn;
    double G0 = (gu * (1.0 - u) + gv * (-vv)) * inv_s + gs + gzw * (zk[0] - zw * wk[0]);
    double G1 = (gu * (-u) + gv * (1.0 - vv)) * inv_s + gs + gzw * (zk[1] - zw * wk[1]);
    double G2 = (gu * (-u) + gv * (-vv)) * inv_s + gs + gzw * (zk[2] - zw * wk[2]);
    gX[0] += G2 * Y[1] - G1 * Y[2];
    gX[1] += G0 * Y[2] - G2 * Y[0];
    gX[2] += G1 * Y[0] - G0 * Y[1];
    gY[0] += G1 * X[2] - G2 * X[1];
    gY[1] += G2 * X[0] - G0 * X[2];
    gY[2] += G0 * X[1] - G1 * X[0];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double ek = e[k];
        out[k] = make_float4((float)gX[k], (float)gY[k], (float)(gzw * ek),
                             (float)(gw[k] - px * gX[k] - py * gY[k] - gzw * zw * ek));
    }
}

template <bool DB>
__device__ __forceinline__ void raster_pixel_grad(const TriVerts &v, int c, int r, int H, int W, float4 g, float4 gd,
                                                  float4 out[3]) {
    raster_pixel_grad_at<DB>(v, pix_ndc_d(c, W), pix_ndc_d(r, H), g, gd, out);
}

__device__ __forceinline__ float4 db_grad_scaled(float4 gd, int H, int W) {
    float kx = 2.f / (float)W, ky = 2.f / (float)H;
    return make_float4(gd.x * kx, gd.y * ky, gd.z * kx, gd.w * ky);
}

}  // namespace drp
