// Device-side conventions of the rasterisation path (gfx950).
//
// The operation sequence of every floating-point quantity that reaches the
// output buffers (snapping, barycentrics, z/w) is fixed here and compiled with
// -ffp-contract=off, so triangle ids and depth are reproducible bit for bit by
// a scalar CPU implementation of the same conventions:
//   pixel (r, c) of a bottom-row-first W x H image has NDC centre
//   ((2c+1-W)/W, (2r+1-H)/H);
//   "direct" triangles (all w > 0, |x|,|y| <= 4096 w) are snapped to 1/16 pixel,
//   X = rint(((x/w) + 1) * (W/2) * 16), and covered by exact integer edge
//   functions with a top-left rule (y up, CCW after orientation normalisation);
//   other triangles use the 2D-homogeneous inside test;
//   barycentrics: X_k = x_k - w_k px, Y_k = y_k - w_k py, e0 = X1 Y2 - X2 Y1,
//   e1 = X2 Y0 - X0 Y2, e2 = X0 Y1 - X1 Y0, s = (e0+e1)+e2, u = e0/s, v = e1/s,
//   z/w = (sum e_k z_k)/(sum e_k w_k); fragments with s == 0 or z/w outside
//   [-1, 1] are dropped; the smallest key (orderable(z/w) << 32 | tri) wins.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace drp {

constexpr int kSubpix = 16;            // 4 sub-pixel bits
constexpr int kHalfSub = 8;            // pixel centre offset in sub-pixel units
constexpr float kGuard = 4096.0f;      // guard band in NDC units
constexpr int kMaxRes = 16384;

enum : int { kCull = 0, kDirect = 1, kHomog = 2 };

// ----------------------------------------------------------------------------
// triangle-id codec of diffrp/rendering/interpolator.py:24-29
// ----------------------------------------------------------------------------
__device__ __forceinline__ float triidx_to_float(int n) {
    return (n <= 0x01000000) ? (float)n : __uint_as_float(0x4a800000u + (uint32_t)n);
}
__device__ __forceinline__ int float_to_triidx(float x) {
    return (x <= 16777216.0f) ? (int)x : (int)(__float_as_uint(x) - 0x4a800000u);
}

__device__ __forceinline__ uint32_t ord_f32(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ uint64_t frag_key(float zw, int t) {
    return ((uint64_t)ord_f32(zw) << 32) | (uint32_t)t;
}

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

__device__ __forceinline__ float pix_ndc(int c, int W) { return (float)(2 * c + 1 - W) / (float)W; }

// ----------------------------------------------------------------------------
// image table: uniform batches compute it arithmetically, uneven batches read it
// ----------------------------------------------------------------------------
struct ImageDesc {
    long long pos_off;   // float offset of this image's vertex 0 in pos
    long long pix_off;   // pixel offset of this image in the output buffers
    int V;               // vertices visible to this image
    int f0, f1;          // triangle range [f0, f1)
    int H, W;
    int tiles_x, tiles_y;
    int tile_off;        // first global tile of this image
    int list_off;        // offset of this image's spill list
    int list_cap;        // capacity of that list
    long long bin_off;   // first bin slot of this image's first tile
    int bin_cap;         // bin capacity of this image's tiles
};

struct Batch {
    const ImageDesc *descs;    // uneven batches: device table (nullptr = uniform)
    const int *tile_to_img;    // uneven batches: global tile -> image
    int n_img;
    // uniform batches
    int V, F, H, W, pos_batched;
    const int *ranges;         // device [n_img, 3] (start, count, spill-list offset) or nullptr
    int tiles_x, tiles_y;
    int bin_cap;               // bin capacity of every tile
    int local_ids;             // rast ids count from each image's first triangle (uneven batches)
};

__device__ __forceinline__ ImageDesc image_desc(const Batch &b, int img) {
    if (b.descs) return b.descs[img];
    ImageDesc d;
    d.V = b.V;
    d.pos_off = b.pos_batched ? (long long)img * b.V * 4 : 0;
    d.H = b.H;
    d.W = b.W;
    d.pix_off = (long long)img * b.H * b.W;
    int f0 = 0, f1 = b.F;
    d.list_off = img * b.F;
    if (b.ranges) {
        f0 = b.ranges[3 * img];
        f1 = f0 + b.ranges[3 * img + 1];
        f0 = f0 < 0 ? 0 : f0;
        f1 = f1 > b.F ? b.F : f1;
        f1 = f1 < f0 ? f0 : f1;
        d.list_off = b.ranges[3 * img + 2];
    }
    d.f0 = f0;
    d.f1 = f1;
    d.tiles_x = b.tiles_x;
    d.tiles_y = b.tiles_y;
    d.tile_off = img * b.tiles_x * b.tiles_y;
    d.list_cap = f1 - f0;
    d.bin_cap = b.bin_cap;
    d.bin_off = (long long)d.tile_off * b.bin_cap;
    return d;
}

__device__ __forceinline__ int tile_image(const Batch &b, int tile) {
    if (b.descs) return b.tile_to_img[tile];
    return tile / (b.tiles_x * b.tiles_y);
}

// ----------------------------------------------------------------------------
// triangle setup
// ----------------------------------------------------------------------------
struct TriVerts {
    float x[3], y[3], z[3], w[3];
};

struct TriSetup {
    int cls;
    int X[3], Y[3];      // snapped, CCW order (direct only)
    int tl[3];           // top-left flag of edge k -> k+1
    int c0, c1, r0, r1;  // candidate pixel box, inclusive
};

__device__ __forceinline__ bool load_tri(const float *__restrict__ pos, const int *__restrict__ tri, int t,
                                         const ImageDesc &d, TriVerts &v) {
    const int *t3 = tri + 3 * (long long)t;
    int i0 = t3[0], i1 = t3[1], i2 = t3[2];
    if ((unsigned)i0 >= (unsigned)d.V || (unsigned)i1 >= (unsigned)d.V || (unsigned)i2 >= (unsigned)d.V) return false;
    const float4 *p = reinterpret_cast<const float4 *>(pos + d.pos_off);
    float4 a = p[i0], b = p[i1], c = p[i2];
    v.x[0] = a.x; v.y[0] = a.y; v.z[0] = a.z; v.w[0] = a.w;
    v.x[1] = b.x; v.y[1] = b.y; v.z[1] = b.z; v.w[1] = b.w;
    v.x[2] = c.x; v.y[2] = c.y; v.z[2] = c.z; v.w[2] = c.w;
    return true;
}

__device__ __forceinline__ long long floor_div16(long long a) {  // floor(a / 16)
    return a >> 4;
}

// pixel box (centres inside the snapped vertices' bounding box), clipped to
// the image; false when empty
__device__ __forceinline__ bool direct_box(const int X[3], const int Y[3], int H, int W, int &bc0, int &bc1,
                                           int &br0, int &br1) {
    int mnx = min(X[0], min(X[1], X[2])), mxx = max(X[0], max(X[1], X[2]));
    int mny = min(Y[0], min(Y[1], Y[2])), mxy = max(Y[0], max(Y[1], Y[2]));
    long long c0 = -floor_div16(-((long long)mnx - kHalfSub)), c1 = floor_div16((long long)mxx - kHalfSub);
    long long r0 = -floor_div16(-((long long)mny - kHalfSub)), r1 = floor_div16((long long)mxy - kHalfSub);
    c0 = c0 < 0 ? 0 : c0;
    r0 = r0 < 0 ? 0 : r0;
    c1 = c1 > W - 1 ? W - 1 : c1;
    r1 = r1 > H - 1 ? H - 1 : r1;
    if (c0 > c1 || r0 > r1) return false;
    bc0 = (int)c0; bc1 = (int)c1; br0 = (int)r0; br1 = (int)r1;
    return true;
}

__device__ __forceinline__ void setup_tri(const TriVerts &v, int H, int W, TriSetup &s) {
    s.cls = kCull;
    for (int k = 0; k < 3; ++k)
        if (!isfinite(v.x[k]) || !isfinite(v.y[k]) || !isfinite(v.z[k]) || !isfinite(v.w[k])) return;
    bool direct = true;
    for (int k = 0; k < 3; ++k)
        direct = direct && (v.w[k] > 0.f && fabsf(v.x[k]) <= kGuard * v.w[k] && fabsf(v.y[k]) <= kGuard * v.w[k]);
    if (direct) {
        float hw = (float)W * 0.5f, hh = (float)H * 0.5f;
        int X[3], Y[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float nx = v.x[k] / v.w[k], ny = v.y[k] / v.w[k];
            float fx = (nx + 1.0f) * hw, fy = (ny + 1.0f) * hh;
            X[k] = (int)rintf(fx * (float)kSubpix);
            Y[k] = (int)rintf(fy * (float)kSubpix);
        }
        long long area = (long long)(X[1] - X[0]) * (Y[2] - Y[0]) - (long long)(X[2] - X[0]) * (Y[1] - Y[0]);
        if (area == 0) return;
        s.X[0] = X[0]; s.Y[0] = Y[0];
        if (area > 0) { s.X[1] = X[1]; s.Y[1] = Y[1]; s.X[2] = X[2]; s.Y[2] = Y[2]; }
        else { s.X[1] = X[2]; s.Y[1] = Y[2]; s.X[2] = X[1]; s.Y[2] = Y[1]; }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            int j = k == 2 ? 0 : k + 1;
            int dx = s.X[j] - s.X[k], dy = s.Y[j] - s.Y[k];
            s.tl[k] = (dy < 0) || (dy == 0 && dx < 0);
        }
        if (!direct_box(X, Y, H, W, s.c0, s.c1, s.r0, s.r1)) return;
        s.cls = kDirect;
        return;
    }
    // homogeneous triangle: conservative pixel box of its w > 0 part.
    // Trivial reject: every vertex beyond one frustum side (pixel rays have |x|,|y| < w).
    if ((v.x[0] > v.w[0] && v.x[1] > v.w[1] && v.x[2] > v.w[2]) ||
        (-v.x[0] > v.w[0] && -v.x[1] > v.w[1] && -v.x[2] > v.w[2]) ||
        (v.y[0] > v.w[0] && v.y[1] > v.w[1] && v.y[2] > v.w[2]) ||
        (-v.y[0] > v.w[0] && -v.y[1] > v.w[1] && -v.y[2] > v.w[2]))
        return;
    if (!(v.w[0] > 0.f) && !(v.w[1] > 0.f) && !(v.w[2] > 0.f)) return;  // nothing in front of the eye
    // projected vertices with w > 0, plus the directions at infinity of edges
    // crossing w = 0 (their projections run off to infinity that way)
    float mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int j = k == 2 ? 0 : k + 1;
        if (v.w[k] > 0.f) {
            float nx = v.x[k] / v.w[k], ny = v.y[k] / v.w[k];
            mnx = fminf(mnx, nx); mxx = fmaxf(mxx, nx);
            mny = fminf(mny, ny); mxy = fmaxf(mxy, ny);
        }
        if ((v.w[k] > 0.f) != (v.w[j] > 0.f)) {
            float t = v.w[k] / (v.w[k] - v.w[j]);
            float dx = v.x[k] + t * (v.x[j] - v.x[k]), dy = v.y[k] + t * (v.y[j] - v.y[k]);
            if (dx > 0.f) mxx = INFINITY;
            if (dx < 0.f) mnx = -INFINITY;
            if (dy > 0.f) mxy = INFINITY;
            if (dy < 0.f) mny = -INFINITY;
            if (dx == 0.f && dy == 0.f) { mnx = mny = -INFINITY; mxx = mxy = INFINITY; }
        }
    }
    mnx = fmaxf(mnx, -2.f); mxx = fminf(mxx, 2.f);
    mny = fmaxf(mny, -2.f); mxy = fminf(mxy, 2.f);
    int c0 = (int)floorf((mnx + 1.f) * 0.5f * (float)W) - 1;
    int c1 = (int)floorf((mxx + 1.f) * 0.5f * (float)W) + 1;
    int r0 = (int)floorf((mny + 1.f) * 0.5f * (float)H) - 1;
    int r1 = (int)floorf((mxy + 1.f) * 0.5f * (float)H) + 1;
    c0 = max(c0, 0); r0 = max(r0, 0);
    c1 = min(c1, W - 1); r1 = min(r1, H - 1);
    if (c0 > c1 || r0 > r1) return;
    s.c0 = c0; s.c1 = c1; s.r0 = r0; s.r1 = r1;
    s.cls = kHomog;
}

// ----------------------------------------------------------------------------
// per-pixel barycentrics
// ----------------------------------------------------------------------------
struct PixEval {
    float X[3], Y[3], e[3], s, zn, wn;
};

__device__ __forceinline__ void eval_pixel(const TriVerts &v, float px, float py, PixEval &p) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        p.X[k] = fma_(-v.w[k], px, v.x[k]);
        p.Y[k] = fma_(-v.w[k], py, v.y[k]);
    }
    p.e[0] = fma_(p.X[1], p.Y[2], -(p.X[2] * p.Y[1]));
    p.e[1] = fma_(p.X[2], p.Y[0], -(p.X[0] * p.Y[2]));
    p.e[2] = fma_(p.X[0], p.Y[1], -(p.X[1] * p.Y[0]));
    p.s = (p.e[0] + p.e[1]) + p.e[2];
    p.zn = fma_(p.e[0], v.z[0], fma_(p.e[1], v.z[1], p.e[2] * v.z[2]));
    p.wn = fma_(p.e[0], v.w[0], fma_(p.e[1], v.w[1], p.e[2] * v.w[2]));
}

__device__ __forceinline__ bool homog_inside(const PixEval &p) {
    if (p.wn > 0.f) return p.e[0] >= 0.f && p.e[1] >= 0.f && p.e[2] >= 0.f;
    if (p.wn < 0.f) return p.e[0] <= 0.f && p.e[1] <= 0.f && p.e[2] <= 0.f;
    return false;
}

// screen-space derivatives of (u, v) in pixel units: (du/dX, du/dY, dv/dX, dv/dY)
__device__ __forceinline__ float4 bary_db(const TriVerts &v, const PixEval &p, float u, float vv, int H, int W) {
    float ex0 = fma_(v.w[2], p.Y[1], -(v.w[1] * p.Y[2]));
    float ex1 = fma_(v.w[0], p.Y[2], -(v.w[2] * p.Y[0]));
    float ex2 = fma_(v.w[1], p.Y[0], -(v.w[0] * p.Y[1]));
    float ey0 = fma_(v.w[1], p.X[2], -(v.w[2] * p.X[1]));
    float ey1 = fma_(v.w[2], p.X[0], -(v.w[0] * p.X[2]));
    float ey2 = fma_(v.w[0], p.X[1], -(v.w[1] * p.X[0]));
    float sx = (ex0 + ex1) + ex2, sy = (ey0 + ey1) + ey2;
    float kx = 2.0f / (float)W, ky = 2.0f / (float)H;
    float4 r;
    r.x = (fma_(-u, sx, ex0) / p.s) * kx;
    r.y = (fma_(-u, sy, ey0) / p.s) * ky;
    r.z = (fma_(-vv, sx, ex1) / p.s) * kx;
    r.w = (fma_(-vv, sy, ey1) / p.s) * ky;
    return r;
}

// XCD-aware block remap: consecutive logical blocks land on the same XCD
// (gfx950 dispatches workgroups round-robin over its 8 XCDs).
// floor(a / b) for 0 <= a < b * 65536, 1 <= b <= 65536: float estimate, one fix-up step
// (the estimate's error is below 1/64, so it is off by at most one).
// Bitonic sort of the wave's 64 keys across lanes (ascending by lane), no barriers.
template <typename K>
__device__ __forceinline__ K wave_sort64(K k) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            K o = __shfl_xor(k, stride, 64);
            bool keep_min = ((lane & stride) == 0) == ((lane & size) == 0);
            k = keep_min ? (o < k ? o : k) : (o > k ? o : k);
        }
    }
    return k;
}

__device__ __forceinline__ int div_box(int a, int b) {
    int q = (int)((float)a * __builtin_amdgcn_rcpf((float)b));
    int r = a - q * b;
    if (r < 0) --q;
    else if (r >= b) ++q;
    return q;
}

__device__ __forceinline__ int xcd_remap(int bid, int nblocks_padded) {
    int per = nblocks_padded >> 3;
    return (bid & 7) * per + (bid >> 3);
}

// blocks -> work items in groups of `group` consecutive items per XCD, the
// groups dealt round robin (block bid runs on XCD bid & 7): neighbouring items
// share an L2 while every XCD gets a share of each region of the grid.  The
// grid is padded to a multiple of 8 * group.
__device__ __forceinline__ int xcd_deal(int bid, int group) {
    const int i = bid >> 3, g = i / group;
    return (g * 8 + (bid & 7)) * group + (i - g * group);
}

}  // namespace drp
