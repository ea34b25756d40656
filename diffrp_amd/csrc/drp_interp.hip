// Attribute interpolation (forward, screen-space derivatives, backward).
// Pixel formula (diffrp/rendering/interpolator.py:43-45):
//   out = (a0 - a2) * u + ((a1 - a2) * v + a2)
#include "drp_common.h"
#include "drp_kernels.h"

namespace drp {

struct InterpPix {
    bool hit;
    int i0, i1, i2;
    float u, v;
};

__device__ __forceinline__ InterpPix interp_pix(const float4 *__restrict__ rast, const int *__restrict__ tri,
                                                long long pix, int F, int V) {
    InterpPix p;
    p.hit = false;
    float4 r = rast[pix];
    if (!(r.w > 0.f)) return p;
    int t = float_to_triidx(r.w) - 1;
    if (t < 0 || t >= F) return p;
    const int *t3 = tri + 3 * (long long)t;
    p.i0 = t3[0]; p.i1 = t3[1]; p.i2 = t3[2];
    if ((unsigned)p.i0 >= (unsigned)V || (unsigned)p.i1 >= (unsigned)V || (unsigned)p.i2 >= (unsigned)V) return p;
    p.u = r.x;
    p.v = r.y;
    p.hit = true;
    return p;
}

// one thread per pixel; C channels per thread (C is small: 1..64 typical)
template <int CV>  // CV > 0: compile-time channel count, 0: runtime
__global__ void __launch_bounds__(256) k_interp_fwd(const float *__restrict__ attr, int V, int C, int attr_batched,
                                                    const float4 *__restrict__ rast, const int *__restrict__ tri,
                                                    int F, long long total_pix, long long pix_per_img,
                                                    float *__restrict__ out, const float4 *__restrict__ rast_db,
                                                    const int *__restrict__ diff_idx, int n_diff,
                                                    float *__restrict__ out_da) {
    const int NC = CV > 0 ? CV : C;
    long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total_pix) return;
    InterpPix p = interp_pix(rast, tri, gid, F, V);
    float *o = out + gid * NC;
    typedef float v4 __attribute__((ext_vector_type(4)));
    if (!p.hit) {
        if (CV > 0 && (CV % 4) == 0) {
            const v4 z = {0.f, 0.f, 0.f, 0.f};
            for (int ch = 0; ch < NC; ch += 4) __builtin_nontemporal_store(z, reinterpret_cast<v4 *>(o + ch));
        } else {
            for (int ch = 0; ch < NC; ++ch) o[ch] = 0.f;
        }
        if (out_da)
            for (int j = 0; j < 2 * n_diff; ++j) out_da[gid * 2 * n_diff + j] = 0.f;
        return;
    }
    const float *A = attr + (attr_batched ? (gid / pix_per_img) * (long long)V * NC : 0);
    const float *a0 = A + (long long)p.i0 * NC, *a1 = A + (long long)p.i1 * NC, *a2 = A + (long long)p.i2 * NC;
    if (CV > 0 && (CV % 4) == 0) {
        for (int ch = 0; ch < NC; ch += 4) {
            float4 x0 = *reinterpret_cast<const float4 *>(a0 + ch);
            float4 x1 = *reinterpret_cast<const float4 *>(a1 + ch);
            float4 x2 = *reinterpret_cast<const float4 *>(a2 + ch);
            float4 y;
            y.x = fma_(x0.x - x2.x, p.u, fma_(x1.x - x2.x, p.v, x2.x));
            y.y = fma_(x0.y - x2.y, p.u, fma_(x1.y - x2.y, p.v, x2.y));
            y.z = fma_(x0.z - x2.z, p.u, fma_(x1.z - x2.z, p.v, x2.z));
            y.w = fma_(x0.w - x2.w, p.u, fma_(x1.w - x2.w, p.v, x2.w));
            v4 yy = {y.x, y.y, y.z, y.w};
            __builtin_nontemporal_store(yy, reinterpret_cast<v4 *>(o + ch));
        }
    } else {
        for (int ch = 0; ch < NC; ++ch) {
            float x2 = a2[ch];
            o[ch] = fma_(a0[ch] - x2, p.u, fma_(a1[ch] - x2, p.v, x2));
        }
    }
    if (out_da) {
        float4 db = rast_db[gid];
        float *od = out_da + gid * 2 * n_diff;
        for (int j = 0; j < n_diff; ++j) {
            int ch = diff_idx[j];
            if ((unsigned)ch >= (unsigned)NC) { od[2 * j] = od[2 * j + 1] = 0.f; continue; }
            float d0 = a0[ch] - a2[ch], d1 = a1[ch] - a2[ch];
            od[2 * j + 0] = fma_(d0, db.x, d1 * db.z);
            od[2 * j + 1] = fma_(d0, db.y, d1 * db.w);
        }
    }
}

// backward: grad_attr scatter (atomics), grad_rast (u, v) written per pixel
template <int CV>
__global__ void __launch_bounds__(256) k_interp_bwd(const float *__restrict__ attr, int V, int C, int attr_batched,
                                                    const float4 *__restrict__ rast, const int *__restrict__ tri,
                                                    int F, long long total_pix, long long pix_per_img,
                                                    const float *__restrict__ grad_out, float *__restrict__ grad_attr,
                                                    float4 *__restrict__ grad_rast, const float4 *__restrict__ rast_db,
                                                    const int *__restrict__ diff_idx, int n_diff,
                                                    const float *__restrict__ grad_da, float4 *__restrict__ grad_db) {
    const int NC = CV > 0 ? CV : C;
    long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total_pix) return;
    InterpPix p = interp_pix(rast, tri, gid, F, V);
    if (!p.hit) {
        grad_rast[gid] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (grad_db) grad_db[gid] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    long long aoff = attr_batched ? (gid / pix_per_img) * (long long)V * NC : 0;
    const float *A = attr + aoff;
    float *GA = grad_attr + aoff;
    const float *a0 = A + (long long)p.i0 * NC, *a1 = A + (long long)p.i1 * NC, *a2 = A + (long long)p.i2 * NC;
    float *g0 = GA + (long long)p.i0 * NC, *g1 = GA + (long long)p.i1 * NC, *g2 = GA + (long long)p.i2 * NC;
    const float *g = grad_out + gid * NC;
    float w2 = (1.f - p.u) - p.v;
    float su = 0.f, sv = 0.f;
    for (int ch = 0; ch < NC; ++ch) {
        float gc = g[ch];
        if (gc != 0.f) {
            atomicAdd(g0 + ch, p.u * gc);
            atomicAdd(g1 + ch, p.v * gc);
            atomicAdd(g2 + ch, w2 * gc);
        }
        float x2 = a2[ch];
        su += gc * (a0[ch] - x2);
        sv += gc * (a1[ch] - x2);
    }
    grad_rast[gid] = make_float4(su, sv, 0.f, 0.f);
    if (grad_da && n_diff > 0) {
        float4 db = rast_db[gid];
        const float *gd = grad_da + gid * 2 * n_diff;
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
        for (int j = 0; j < n_diff; ++j) {
            int ch = diff_idx[j];
            if ((unsigned)ch >= (unsigned)NC) continue;
            float gx = gd[2 * j], gy = gd[2 * j + 1];
            float d0 = a0[ch] - a2[ch], d1 = a1[ch] - a2[ch];
            float k0 = gx * db.x + gy * db.y, k1 = gx * db.z + gy * db.w;
            atomicAdd(g0 + ch, k0);
            atomicAdd(g1 + ch, k1);
            atomicAdd(g2 + ch, -(k0 + k1));
            s0 += gx * d0; s1 += gy * d0; s2 += gx * d1; s3 += gy * d1;
        }
        if (grad_db) grad_db[gid] = make_float4(s0, s1, s2, s3);
    } else if (grad_db) {
        grad_db[gid] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

#define DRP_INTERP_DISPATCH(KERNEL, ...)                                                          \
    switch (L.cv) {                                                                               \
        case 1: hipLaunchKernelGGL(KERNEL<1>, grid, block, 0, st, __VA_ARGS__); break;            \
        case 2: hipLaunchKernelGGL(KERNEL<2>, grid, block, 0, st, __VA_ARGS__); break;            \
        case 3: hipLaunchKernelGGL(KERNEL<3>, grid, block, 0, st, __VA_ARGS__); break;            \
        case 4: hipLaunchKernelGGL(KERNEL<4>, grid, block, 0, st, __VA_ARGS__); break;            \
        case 8: hipLaunchKernelGGL(KERNEL<8>, grid, block, 0, st, __VA_ARGS__); break;            \
        case 16: hipLaunchKernelGGL(KERNEL<16>, grid, block, 0, st, __VA_ARGS__); break;          \
        default: hipLaunchKernelGGL(KERNEL<0>, grid, block, 0, st, __VA_ARGS__); break;           \
    }

int launch_interpolate(const InterpLaunch &L, hipStream_t st) {
    if (L.total_pix == 0) return 0;
    dim3 grid((unsigned)((L.total_pix + 255) / 256)), block(256);
    DRP_INTERP_DISPATCH(k_interp_fwd, L.attr, L.V, L.C, L.attr_batched, (const float4 *)L.rast, L.tri, L.F,
                        L.total_pix, L.pix_per_img, L.out, (const float4 *)L.rast_db, L.diff_idx, L.n_diff, L.out_da)
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ----------------------------------------------------------------------------
// deterministic backward (no atomics)
//   k_tile_reduce / k_tile_gather (drp_tribwd.hip): grad_rast per pixel and,
//       per (view group, triangle), (u, v, 1-u-v) * g per corner
//   k_gather_attr: per vertex, sum the corner partials (CSR, fixed order)
// ----------------------------------------------------------------------------
// grad_attr[out][v] += sum over groups g in the output's range and incident corners
template <int NC>
__global__ void __launch_bounds__(256) k_gather_attr(const float *__restrict__ partial, int n_groups, int F, int V,
                                                     int per_output_groups, const int *__restrict__ inc_off,
                                                     const int *__restrict__ inc, float *__restrict__ grad_attr) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    int out = blockIdx.y;
    if (v >= V) return;
    float s[NC];
#pragma unroll
    for (int ch = 0; ch < NC; ++ch) s[ch] = 0.f;
    int i0 = inc_off[v], i1 = inc_off[v + 1];
    int g0 = per_output_groups ? out : 0, g1 = per_output_groups ? out + 1 : n_groups;
    for (int g = g0; g < g1; ++g) {
        const float *P = partial + (long long)g * F * 3 * NC;
        if (NC % 4 == 0 && NC <= 8) {
            // four corners' loads in flight at once (corner partials are NC * 4 bytes from an
            // aligned base), summed in corner order
            for (int i = i0; i < i1; i += 4) {
                float4 x[4][NC / 4];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (i + k < i1) {
                        const float4 *q = reinterpret_cast<const float4 *>(P + (long long)inc[i + k] * NC);
#pragma unroll
                        for (int h = 0; h < NC / 4; ++h) x[k][h] = q[h];
                    }
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (i + k < i1) {
#pragma unroll
                        for (int h = 0; h < NC / 4; ++h) {
                            s[4 * h] += x[k][h].x; s[4 * h + 1] += x[k][h].y;
                            s[4 * h + 2] += x[k][h].z; s[4 * h + 3] += x[k][h].w;
                        }
                    }
            }
        } else {
            for (int i = i0; i < i1; ++i) {
                const float *q = P + (long long)inc[i] * NC;
#pragma unroll
                for (int ch = 0; ch < NC; ++ch) s[ch] += q[ch];
            }
        }
    }
    float *o = grad_attr + ((long long)out * V + v) * NC;
#pragma unroll
    for (int ch = 0; ch < NC; ++ch) o[ch] += s[ch];
}

// The same for 8 and 16 channels with `lanes` lanes per (vertex, four
// channels) (1, 2 or 4: up to the groups each output sums): lane s takes
// groups s, s + lanes, ... (two groups' corner rows in flight together, each
// group's corners summed in order), and the lanes' sums are added in a fixed
// tree.  (One lane per (vertex, quad) left the 64-view grid at barely one wave
// per SIMD waiting on its loads.)
constexpr int kAttrLanes = 4;
template <int NC>
__global__ void __launch_bounds__(256) k_gather_attr_q(const float *__restrict__ partial, int n_groups, int F,
                                                       int V, int per_output_groups, const int *__restrict__ inc_off,
                                                       const int *__restrict__ inc, float *__restrict__ grad_attr,
                                                       int lanes) {
    constexpr int NQ = NC / 4;
    const long long tl = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = (int)(tl & (lanes - 1));
    const long long tq = tl / lanes;
    const int v = (int)(tq / NQ), q = (int)(tq % NQ);
    const int out = blockIdx.y;
    if (v >= V) return;  // all lanes of (v, q) together
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const int i0 = inc_off[v], i1 = inc_off[v + 1], deg = i1 - i0;
    const int g0 = per_output_groups ? out : 0, g1 = per_output_groups ? out + 1 : n_groups;
    const long long gstride = (long long)F * 3 * NC;
    if (deg <= 8) {
        int c[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) c[k] = k < deg ? inc[i0 + k] : 0;
        for (int g = g0 + s; g < g1; g += 2 * lanes) {
            float4 x[2][8];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int gh = g + h * lanes;
                const float *P = partial + (long long)gh * gstride + 4 * q;
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (gh < g1 && k < deg) x[h][k] = *reinterpret_cast<const float4 *>(P + (long long)c[k] * NC);
            }
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (g + h * lanes < g1 && k < deg) {
                        acc.x += x[h][k].x; acc.y += x[h][k].y; acc.z += x[h][k].z; acc.w += x[h][k].w;
                    }
        }
    } else {
        for (int g = g0 + s; g < g1; g += lanes) {
            const float *P = partial + (long long)g * gstride + 4 * q;
            for (int i = i0; i < i1; i += 4) {  // four corners' loads in flight, summed in corner order
                float4 x[4];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (i + k < i1) x[k] = *reinterpret_cast<const float4 *>(P + (long long)inc[i + k] * NC);
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (i + k < i1) {
                        acc.x += x[k].x; acc.y += x[k].y; acc.z += x[k].z; acc.w += x[k].w;
                    }
            }
        }
    }
    for (int o = 1; o < lanes; o <<= 1) {  // (a + b on one lane, b + a on the other: the same sum)
        acc.x += __shfl_xor(acc.x, o, 64);
        acc.y += __shfl_xor(acc.y, o, 64);
        acc.z += __shfl_xor(acc.z, o, 64);
        acc.w += __shfl_xor(acc.w, o, 64);
    }
    if (s != 0) return;
    float *o = grad_attr + ((long long)out * V + v) * NC + 4 * q;
    o[0] += acc.x;
    o[1] += acc.y;
    o[2] += acc.z;
    o[3] += acc.w;
}

template <int NC>
static void launch_gather_attr_t(const float *partial, int groups, int F, int V, int outs, int per_output_groups,
                                 const int *inc_off, const int *inc, float *grad_attr, hipStream_t st) {
    if constexpr (NC == 8 || NC == 16) {
        const int per_out = per_output_groups ? 1 : groups;
        const int lanes = per_out >= kAttrLanes ? kAttrLanes : (per_out >= 2 ? 2 : 1);
        const long long n = (long long)V * (NC / 4) * lanes;
        hipLaunchKernelGGL(k_gather_attr_q<NC>, dim3((unsigned)((n + 255) / 256), outs), dim3(256), 0, st, partial,
                           groups, F, V, per_output_groups, inc_off, inc, grad_attr, lanes);
    } else {
        hipLaunchKernelGGL(k_gather_attr<NC>, dim3((V + 255) / 256, outs), dim3(256), 0, st, partial, groups, F, V,
                           per_output_groups, inc_off, inc, grad_attr);
    }
}

template <int NC>
static void launch_interp_bwd_tri_t(const InterpLaunch &L, hipStream_t st) {
    int n_img = L.um.n ? 1 : (int)(L.total_pix / L.pix_per_img);  // an uneven batch sums as one group
    int vpg = L.attr_batched ? 1 : L.views_per_group;
    int groups = (n_img + vpg - 1) / vpg;
    TriBwdLaunch T{};
    T.tri = L.tri;
    T.rast = L.rast;
    T.grad_out = L.grad_out;
    T.attr = L.attr;
    T.V = L.V;
    T.attr_batched = L.attr_batched;
    T.grad_rast_out = L.grad_rast;
    T.boxes = L.boxes;
    T.partial = L.partial;
    T.n_img = n_img;
    T.F = L.F;
    T.H = L.H;
    T.W = L.W;
    T.views_per_block = vpg;
    if (L.um.n) {
        T.um = L.um;
        T.views_per_block = 1;
        tile_records_layout_tiles(T, L.records, L.um_tiles, L.F);
    } else {
        tile_records_layout(T, L.records, n_img, L.H, L.W, (long long)n_img * L.F);
    }
    launch_tri_bwd_attr(T, NC, st);
    int outs = L.attr_batched ? n_img : 1;
    launch_gather_attr_t<NC>(L.partial, groups, L.F, L.V, outs, L.attr_batched, L.inc_off, L.inc, L.grad_attr, st);
}

int launch_gather_attr(const float *partial, int groups, int F, int V, int C, const int *inc_off, const int *inc,
                       float *grad_attr, hipStream_t st) {
    if (V == 0) return 0;
    switch (C) {
        case 4: launch_gather_attr_t<4>(partial, groups, F, V, 1, 0, inc_off, inc, grad_attr, st); break;
        case 8: launch_gather_attr_t<8>(partial, groups, F, V, 1, 0, inc_off, inc, grad_attr, st); break;
        default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_interpolate_bwd(const InterpLaunch &L, hipStream_t st) {
    if (L.total_pix == 0) return 0;
    if (L.boxes && L.partial && L.records && L.inc_off && L.inc && L.n_diff == 0 && L.cv > 0) {
        switch (L.cv) {
            case 1: launch_interp_bwd_tri_t<1>(L, st); break;
            case 2: launch_interp_bwd_tri_t<2>(L, st); break;
            case 3: launch_interp_bwd_tri_t<3>(L, st); break;
            case 4: launch_interp_bwd_tri_t<4>(L, st); break;
            case 8: launch_interp_bwd_tri_t<8>(L, st); break;
            case 16: launch_interp_bwd_tri_t<16>(L, st); break;
            default: return -1;
        }
        return hipGetLastError() == hipSuccess ? 0 : -2;
    }
    dim3 grid((unsigned)((L.total_pix + 255) / 256)), block(256);
    DRP_INTERP_DISPATCH(k_interp_bwd, L.attr, L.V, L.C, L.attr_batched, (const float4 *)L.rast, L.tri, L.F,
                        L.total_pix, L.pix_per_img, L.grad_out, L.grad_attr, (float4 *)L.grad_rast,
                        (const float4 *)L.rast_db, L.diff_idx, L.n_diff, L.grad_da, (float4 *)L.grad_db)
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace drp
