// Texture sampling for the shading stage after interpolation: diffrp's
// sample2d / sample3d (diffrp/utils/shader_ops.py:198-310), used by
// GLTFSampler (diffrp/materials/gltf_material.py:15-22) and image-based
// lighting, which call torch's grid_sample on a permuted (C, H, W) view.
// Here the texture stays channels-last ([H, W, C] / [D, H, W, C]): a tap is
// C contiguous floats.  The coordinate mapping is diffrp's: uv (0, 0) is the
// bottom-left texel corner (last row), x = 2u - 1, y = 1 - 2v (z = 2w - 1), and
// grid_sample's unnormalisation, border / reflection padding and
// align_corners (latlong) rules; "cyclic" samples the 2 x 2 tiled texture at
// (uv mod 1) / 2 shifted into [1/4, 3/4) with reflection padding, as diffrp
// does.  Forward: one thread per sample.  Backward: the texture gradient is
// scattered with float atomics (as grid_sample's backward), the coordinate
// gradient is written per sample.
#include "drp_common.h"
#include "drp_kernels.h"

namespace drp {

enum : int { kWrapBorder = 0, kWrapReflection = 1, kWrapCyclic = 2, kWrapLatlong = 3 };
enum : int { kModeBilinear = 0, kModeNearest = 1, kModeBicubic = 2 };

// grid_sample's source index for one axis with its derivative d(index)/d(grid)
__device__ __forceinline__ float pad_coord(float x, int size, bool align, bool reflect, float &dg);

__device__ __forceinline__ float source_index(float g, int size, bool align, bool reflect, bool raw, float &dg) {
    float x;
    if (align) {
        x = (g + 1.f) * 0.5f * (float)(size - 1);
        dg = (float)(size - 1) * 0.5f;
    } else {
        x = ((g + 1.f) * (float)size - 1.f) * 0.5f;
        dg = (float)size * 0.5f;
    }
    return raw ? x : pad_coord(x, size, align, reflect, dg);  // bicubic pads each tap instead
}

// border / reflection padding of a source coordinate, with its derivative factor
__device__ __forceinline__ float pad_coord(float x, int size, bool align, bool reflect, float &dg) {
    if (reflect) {
        const float lo = align ? 0.f : -0.5f;  // reflection interval [lo, lo + span]
        const float span = align ? (float)(size - 1) : (float)size;
        if (span <= 0.f) {
            x = 0.f;
            dg = 0.f;
        } else {
            float in = x - lo;
            if (in < 0.f) {
                in = -in;
                dg = -dg;
            }
            const float extra = fmodf(in, span);
            const int flips = (int)floorf(in / span);
            if (flips % 2 == 0) {
                x = extra + lo;
            } else {
                x = span - extra + lo;
                dg = -dg;
            }
        }
    }
    // clip to [0, size - 1] (border padding, and after a reflection)
    if (x <= 0.f) {
        x = 0.f;
        dg = 0.f;
    } else if (x >= (float)(size - 1)) {
        x = (float)(size - 1);
        dg = 0.f;
    }
    return x;
}

struct TexAxis {
    float pos;    // source index
    float dpos;   // d pos / d (texture coordinate)
    int size;     // sampled size (2x for cyclic)
    int period;   // texel period (the texture's size)
};

// texture coordinate -> source index of one axis; flip: y axis (1 - 2v)
__device__ __forceinline__ TexAxis make_axis(float u, int size, int wrap, bool flip, bool raw = false) {
    TexAxis a;
    a.period = size;
    float du = 1.f;
    if (wrap == kWrapCyclic) {
        float m = u - floorf(u);  // remainder(u, 1), derivative 1
        m *= 0.5f;
        if (!(m >= 0.25f)) m += 0.5f;
        u = m;
        du = 0.5f;
        size *= 2;
    }
    float g = flip ? -(u * 2.f - 1.f) : (u * 2.f - 1.f);
    float dg;
    const bool align = wrap == kWrapLatlong;
    const bool reflect = wrap != kWrapBorder;
    a.pos = source_index(g, size, align, reflect, raw, dg);
    a.size = size;
    a.dpos = dg * (flip ? -2.f : 2.f) * du;
    return a;
}

__device__ __forceinline__ int texel(const TexAxis &a, int i) { return a.size == a.period ? i : i % a.period; }

// bicubic (Keys, A = -0.75, as grid_sample): weights of the taps x0 - 1 .. x0 + 2 at
// fraction t, the weights' derivatives with respect to -t, and each tap's padded texel
__device__ __forceinline__ void cubic_weights(float t, float w[4], float dw[4]) {
    const float A = -0.75f;
    float x = t + 1.f;
    w[0] = ((A * x - 5.f * A) * x + 8.f * A) * x - 4.f * A;
    x = t;
    w[1] = ((A + 2.f) * x - (A + 3.f)) * x * x + 1.f;
    x = 1.f - t;
    w[2] = ((A + 2.f) * x - (A + 3.f)) * x * x + 1.f;
    x = 2.f - t;
    w[3] = ((A * x - 5.f * A) * x + 8.f * A) * x - 4.f * A;
    x = -1.f - t;
    dw[0] = (-3.f * A * x - 10.f * A) * x - 8.f * A;
    x = -t;
    dw[1] = (-3.f * (A + 2.f) * x - 2.f * (A + 3.f)) * x;
    x = 1.f - t;
    dw[2] = (3.f * (A + 2.f) * x - 2.f * (A + 3.f)) * x;
    x = 2.f - t;
    dw[3] = (3.f * A * x - 10.f * A) * x + 8.f * A;
}

__device__ __forceinline__ int cubic_tap(const TexAxis &a, int i, int wrap) {
    float dg = 0.f;
    const float x = pad_coord((float)i, a.size, wrap == kWrapLatlong, wrap != kWrapBorder, dg);
    return texel(a, (int)x);
}

template <int CV>
__global__ void __launch_bounds__(256) k_sample2d_fwd(const float *__restrict__ tex, int H, int W, int C,
                                                      const float *__restrict__ uv, long long n, int wrap, int mode,
                                                      float *__restrict__ out) {
    const int Cn = CV > 0 ? CV : C;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool raw = mode == kModeBicubic;
    const TexAxis ax = make_axis(uv[2 * i], W, wrap, false, raw), ay = make_axis(uv[2 * i + 1], H, wrap, true, raw);
    float *o = out + i * Cn;
    if (mode == kModeBicubic) {
        const float fx0 = floorf(ax.pos), fy0 = floorf(ay.pos);
        float wx[4], wy[4], dwx[4], dwy[4];
        cubic_weights(ax.pos - fx0, wx, dwx);
        cubic_weights(ay.pos - fy0, wy, dwy);
        int tx[4], ty[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            tx[k] = cubic_tap(ax, (int)fx0 - 1 + k, wrap);
            ty[k] = cubic_tap(ay, (int)fy0 - 1 + k, wrap);
        }
        for (int c = 0; c < Cn; ++c) {
            float rows[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float *row = tex + (long long)ty[k] * W * Cn + c;
                rows[k] = row[tx[0] * Cn] * wx[0] + row[tx[1] * Cn] * wx[1] + row[tx[2] * Cn] * wx[2] +
                          row[tx[3] * Cn] * wx[3];
            }
            o[c] = rows[0] * wy[0] + rows[1] * wy[1] + rows[2] * wy[2] + rows[3] * wy[3];
        }
        return;
    }
    if (mode == kModeNearest) {
        const int x = (int)rintf(ax.pos), y = (int)rintf(ay.pos);
        const bool in = x >= 0 && x < ax.size && y >= 0 && y < ay.size;
        const float *t = tex + ((long long)texel(ay, in ? y : 0) * W + texel(ax, in ? x : 0)) * Cn;
        for (int c = 0; c < Cn; ++c) o[c] = in ? t[c] : 0.f;
        return;
    }
    const int x0 = (int)floorf(ax.pos), y0 = (int)floorf(ay.pos);
    const float fx = ax.pos - (float)x0, fy = ay.pos - (float)y0;
    const float w[4] = {(1.f - fx) * (1.f - fy), fx * (1.f - fy), (1.f - fx) * fy, fx * fy};
    long long off[4];
    bool in[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int x = x0 + (k & 1), y = y0 + (k >> 1);
        in[k] = x >= 0 && x < ax.size && y >= 0 && y < ay.size;
        off[k] = in[k] ? ((long long)texel(ay, y) * W + texel(ax, x)) * Cn : 0;
    }
    for (int c = 0; c < Cn; ++c) {
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (in[k]) s += tex[off[k] + c] * w[k];
        o[c] = s;
    }
}

template <int CV>
__global__ void __launch_bounds__(256) k_sample2d_bwd(const float *__restrict__ tex, int H, int W, int C,
                                                      const float *__restrict__ uv, long long n, int wrap, int mode,
                                                      const float *__restrict__ gout, float *__restrict__ gtex,
                                                      float *__restrict__ guv) {
    const int Cn = CV > 0 ? CV : C;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool raw = mode == kModeBicubic;
    const TexAxis ax = make_axis(uv[2 * i], W, wrap, false, raw), ay = make_axis(uv[2 * i + 1], H, wrap, true, raw);
    const float *g = gout + i * Cn;
    if (mode == kModeBicubic) {
        const float fx0 = floorf(ax.pos), fy0 = floorf(ay.pos);
        float wx[4], wy[4], dwx[4], dwy[4];
        cubic_weights(ax.pos - fx0, wx, dwx);
        cubic_weights(ay.pos - fy0, wy, dwy);
        int tx[4], ty[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            tx[k] = cubic_tap(ax, (int)fx0 - 1 + k, wrap);
            ty[k] = cubic_tap(ay, (int)fy0 - 1 + k, wrap);
        }
        float gx = 0.f, gy = 0.f;
        for (int c = 0; c < Cn; ++c) {
            const float gc = g[c];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const long long off = ((long long)ty[a] * W + tx[b]) * Cn + c;
                    const float v = tex[off];
                    gx -= v * dwx[b] * wy[a] * gc;  // dw holds -d w / d t
                    gy -= v * wx[b] * dwy[a] * gc;
                    if (gtex) atomicAdd(gtex + off, wy[a] * wx[b] * gc);
                }
        }
        if (guv) {
            guv[2 * i] = gx * ax.dpos;
            guv[2 * i + 1] = gy * ay.dpos;
        }
        return;
    }
    if (mode == kModeNearest) {
        const int x = (int)rintf(ax.pos), y = (int)rintf(ay.pos);
        if (gtex && x >= 0 && x < ax.size && y >= 0 && y < ay.size) {
            float *t = gtex + ((long long)texel(ay, y) * W + texel(ax, x)) * Cn;
            for (int c = 0; c < Cn; ++c) atomicAdd(t + c, g[c]);
        }
        if (guv) {
            guv[2 * i] = 0.f;
            guv[2 * i + 1] = 0.f;
        }
        return;
    }
    const int x0 = (int)floorf(ax.pos), y0 = (int)floorf(ay.pos);
    const float fx = ax.pos - (float)x0, fy = ay.pos - (float)y0;
    const float w[4] = {(1.f - fx) * (1.f - fy), fx * (1.f - fy), (1.f - fx) * fy, fx * fy};
    // d w / d fx and d w / d fy
    const float wx[4] = {-(1.f - fy), 1.f - fy, -fy, fy}, wy[4] = {-(1.f - fx), -fx, 1.f - fx, fx};
    long long off[4];
    bool in[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int x = x0 + (k & 1), y = y0 + (k >> 1);
        in[k] = x >= 0 && x < ax.size && y >= 0 && y < ay.size;
        off[k] = in[k] ? ((long long)texel(ay, y) * W + texel(ax, x)) * Cn : 0;
    }
    float gx = 0.f, gy = 0.f;
    for (int c = 0; c < Cn; ++c) {
        const float gc = g[c];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!in[k]) continue;
            const float v = tex[off[k] + c];
            gx += wx[k] * v * gc;
            gy += wy[k] * v * gc;
            if (gtex) atomicAdd(gtex + off[k] + c, w[k] * gc);
        }
    }
    if (guv) {
        guv[2 * i] = gx * ax.dpos;
        guv[2 * i + 1] = gy * ay.dpos;
    }
}

template <int CV>
__global__ void __launch_bounds__(256) k_sample3d_fwd(const float *__restrict__ tex, int D, int H, int W, int C,
                                                      const float *__restrict__ uvw, long long n, int wrap, int mode,
                                                      float *__restrict__ out) {
    const int Cn = CV > 0 ? CV : C;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const TexAxis ax = make_axis(uvw[3 * i], W, wrap, false), ay = make_axis(uvw[3 * i + 1], H, wrap, true),
                  az = make_axis(uvw[3 * i + 2], D, wrap, false);
    float *o = out + i * Cn;
    if (mode == kModeNearest) {
        const int x = (int)rintf(ax.pos), y = (int)rintf(ay.pos), z = (int)rintf(az.pos);
        const bool in = x >= 0 && x < W && y >= 0 && y < H && z >= 0 && z < D;
        const float *t = tex + (((long long)(in ? z : 0) * H + (in ? y : 0)) * W + (in ? x : 0)) * Cn;
        for (int c = 0; c < Cn; ++c) o[c] = in ? t[c] : 0.f;
        return;
    }
    const int x0 = (int)floorf(ax.pos), y0 = (int)floorf(ay.pos), z0 = (int)floorf(az.pos);
    const float fx = ax.pos - (float)x0, fy = ay.pos - (float)y0, fz = az.pos - (float)z0;
    for (int c = 0; c < Cn; ++c) o[c] = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int x = x0 + (k & 1), y = y0 + ((k >> 1) & 1), z = z0 + (k >> 2);
        if (!(x >= 0 && x < W && y >= 0 && y < H && z >= 0 && z < D)) continue;
        const float w = ((k & 1) ? fx : 1.f - fx) * (((k >> 1) & 1) ? fy : 1.f - fy) * ((k >> 2) ? fz : 1.f - fz);
        const float *t = tex + (((long long)z * H + y) * W + x) * Cn;
        for (int c = 0; c < Cn; ++c) o[c] += t[c] * w;
    }
}

template <int CV>
__global__ void __launch_bounds__(256) k_sample3d_bwd(const float *__restrict__ tex, int D, int H, int W, int C,
                                                      const float *__restrict__ uvw, long long n, int wrap, int mode,
                                                      const float *__restrict__ gout, float *__restrict__ gtex,
                                                      float *__restrict__ guvw) {
    const int Cn = CV > 0 ? CV : C;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const TexAxis ax = make_axis(uvw[3 * i], W, wrap, false), ay = make_axis(uvw[3 * i + 1], H, wrap, true),
                  az = make_axis(uvw[3 * i + 2], D, wrap, false);
    const float *g = gout + i * Cn;
    if (mode == kModeNearest) {
        const int x = (int)rintf(ax.pos), y = (int)rintf(ay.pos), z = (int)rintf(az.pos);
        if (gtex && x >= 0 && x < W && y >= 0 && y < H && z >= 0 && z < D) {
            float *t = gtex + (((long long)z * H + y) * W + x) * Cn;
            for (int c = 0; c < Cn; ++c) atomicAdd(t + c, g[c]);
        }
        if (guvw) guvw[3 * i] = guvw[3 * i + 1] = guvw[3 * i + 2] = 0.f;
        return;
    }
    const int x0 = (int)floorf(ax.pos), y0 = (int)floorf(ay.pos), z0 = (int)floorf(az.pos);
    const float fx = ax.pos - (float)x0, fy = ay.pos - (float)y0, fz = az.pos - (float)z0;
    float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int x = x0 + (k & 1), y = y0 + ((k >> 1) & 1), z = z0 + (k >> 2);
        if (!(x >= 0 && x < W && y >= 0 && y < H && z >= 0 && z < D)) continue;
        const float wxs = (k & 1) ? fx : 1.f - fx, wys = ((k >> 1) & 1) ? fy : 1.f - fy, wzs = (k >> 2) ? fz : 1.f - fz;
        const float dx = (k & 1) ? 1.f : -1.f, dy = ((k >> 1) & 1) ? 1.f : -1.f, dz = (k >> 2) ? 1.f : -1.f;
        const float *t = tex + (((long long)z * H + y) * W + x) * Cn;
        float *gt = gtex ? gtex + (((long long)z * H + y) * W + x) * Cn : nullptr;
        float dot = 0.f;
        for (int c = 0; c < Cn; ++c) {
            dot += t[c] * g[c];
            if (gt) atomicAdd(gt + c, wxs * wys * wzs * g[c]);
        }
        gx += dx * wys * wzs * dot;
        gy += wxs * dy * wzs * dot;
        gz += wxs * wys * dz * dot;
    }
    if (guvw) {
        guvw[3 * i] = gx * ax.dpos;
        guvw[3 * i + 1] = gy * ay.dpos;
        guvw[3 * i + 2] = gz * az.dpos;
    }
}

#define DRP_TEX_DISPATCH(KERNEL, ...)                                                          \
    switch (C) {                                                                               \
        case 1: hipLaunchKernelGGL(KERNEL<1>, grid, dim3(256), 0, st, __VA_ARGS__); break;     \
        case 3: hipLaunchKernelGGL(KERNEL<3>, grid, dim3(256), 0, st, __VA_ARGS__); break;     \
        case 4: hipLaunchKernelGGL(KERNEL<4>, grid, dim3(256), 0, st, __VA_ARGS__); break;     \
        default: hipLaunchKernelGGL(KERNEL<0>, grid, dim3(256), 0, st, __VA_ARGS__); break;    \
    }

int launch_sample2d(const float *tex, int H, int W, int C, const float *uv, long long n, int wrap, int mode,
                    float *out, hipStream_t st) {
    if (n == 0) return 0;
    dim3 grid((unsigned)((n + 255) / 256));
    DRP_TEX_DISPATCH(k_sample2d_fwd, tex, H, W, C, uv, n, wrap, mode, out)
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_sample2d_bwd(const float *tex, int H, int W, int C, const float *uv, long long n, int wrap, int mode,
                        const float *gout, float *gtex, float *guv, hipStream_t st) {
    if (n == 0) return 0;
    dim3 grid((unsigned)((n + 255) / 256));
    DRP_TEX_DISPATCH(k_sample2d_bwd, tex, H, W, C, uv, n, wrap, mode, gout, gtex, guv)
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_sample3d(const float *tex, int D, int H, int W, int C, const float *uvw, long long n, int wrap, int mode,
                    float *out, hipStream_t st) {
    if (n == 0) return 0;
    dim3 grid((unsigned)((n + 255) / 256));
    DRP_TEX_DISPATCH(k_sample3d_fwd, tex, D, H, W, C, uvw, n, wrap, mode, out)
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_sample3d_bwd(const float *tex, int D, int H, int W, int C, const float *uvw, long long n, int wrap,
                        int mode, const float *gout, float *gtex, float *guvw, hipStream_t st) {
    if (n == 0) return 0;
    dim3 grid((unsigned)((n + 255) / 256));
    DRP_TEX_DISPATCH(k_sample3d_bwd, tex, D, H, W, C, uvw, n, wrap, mode, gout, gtex, guvw)
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace drp
