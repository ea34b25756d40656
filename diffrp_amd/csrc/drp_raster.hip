// Rasterizer kernels for gfx950: triangle setup + screen-tile binning, per-tile
// coverage/depth in registers with triangle records staged in LDS, and the
// backward pass (barycentric -> clip-space position chain).
#include "drp_common.h"
#include "drp_kernels.h"

namespace drp {

// ----------------------------------------------------------------------------
// setup + binning: one thread per (image, triangle)
//   small triangles (<= kMaxBinTiles tiles) go to each tile's bin;
//   large ones, and ones whose bin is full, go to the image's spill list;
//   if that list overflows too, every tile of the image scans all triangles.
// ----------------------------------------------------------------------------
constexpr int kMaxBinTiles = 16;

__global__ void __launch_bounds__(256) k_setup_bin(Batch b, const float *__restrict__ pos,
                                                   const int *__restrict__ tri, int tile_w, int tile_h,
                                                   int bin_cap, int *__restrict__ bin_count,
                                                   int *__restrict__ bins, int *__restrict__ list_count,
                                                   int *__restrict__ lists) {
    int img = blockIdx.y;
    ImageDesc d = image_desc(b, img);
    int t = d.f0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.f1) return;
    TriVerts v;
    if (!load_tri(pos, tri, t, d, v)) return;
    TriSetup s;
    setup_tri(v, d.H, d.W, s);
    if (s.cls == kCull) return;
    int tx0 = s.c0 / tile_w, tx1 = s.c1 / tile_w, ty0 = s.r0 / tile_h, ty1 = s.r1 / tile_h;
    int ntiles = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
    bool spill = ntiles > kMaxBinTiles;
    if (!spill) {
        for (int ty = ty0; ty <= ty1; ++ty)
            for (int tx = tx0; tx <= tx1; ++tx) {
                int tile = d.tile_off + ty * d.tiles_x + tx;
                int slot = atomicAdd(&bin_count[tile], 1);
                if (slot < bin_cap) bins[(long long)tile * bin_cap + slot] = t;
                else spill = true;
            }
    }
    if (spill) {
        int slot = atomicAdd(&list_count[img], 1);
        if (slot < d.list_cap) lists[(long long)d.list_off + slot] = t;
    }
}

// ----------------------------------------------------------------------------
// fine raster: one workgroup per tile, one thread per pixel
// ----------------------------------------------------------------------------
struct TileRec {
    int t, cls, r0, r1;
    long long E0[3];
    int A[3], B[3];
    TriVerts v;
};

template <int TW, int TH, bool PEEL, bool DB>
__global__ void __launch_bounds__(TW *TH) k_fine(Batch b, const float *__restrict__ pos,
                                                 const int *__restrict__ tri, int bin_cap,
                                                 const int *__restrict__ bin_count, const int *__restrict__ bins,
                                                 const int *__restrict__ list_count,
                                                 const int *__restrict__ lists, int n_tiles, int n_tiles_pad,
                                                 const float4 *__restrict__ prev, float4 *__restrict__ rast,
                                                 float4 *__restrict__ rast_db) {
    constexpr int NT = TW * TH;
    __shared__ TileRec rec[NT];
    __shared__ int nrec;
    int tile = xcd_remap(blockIdx.x, n_tiles_pad);
    if (tile >= n_tiles) return;
    int img = tile_image(b, tile);
    ImageDesc d = image_desc(b, img);
    int lt = tile - d.tile_off;
    int tx = lt % d.tiles_x, ty = lt / d.tiles_x;
    int tid = threadIdx.x;
    int c = tx * TW + (tid % TW), r = ty * TH + (tid / TW);
    bool inimg = c < d.W && r < d.H;
    long long pix = d.pix_off + (long long)r * d.W + c;
    // wave footprint (rows) for wave-uniform rejection
    constexpr int ROWS_PER_WAVE = 64 / TW > 0 ? 64 / TW : 1;
    int wave_r0 = ty * TH + (tid / 64) * ROWS_PER_WAVE;
    int wave_r1 = wave_r0 + ROWS_PER_WAVE - 1;
    int tc0 = tx * TW, tr0 = ty * TH;

    uint64_t prev_key = 0;
    bool alive = inimg;
    if (PEEL && inimg) {
        float4 pv = prev[pix];
        if (pv.w > 0.f) prev_key = frag_key(pv.z, float_to_triidx(pv.w) - 1);
        else alive = false;
    }
    uint64_t best = ~0ull;
    float bu = 0.f, bv = 0.f, bz = 0.f;
    float4 bdb = make_float4(0.f, 0.f, 0.f, 0.f);
    float px = pix_ndc(c, d.W), py = pix_ndc(r, d.H);
    int oxf = (tid % TW) * kSubpix, oyf = (tid / TW) * kSubpix;

    int cnt = bin_count[tile];
    int lcnt = list_count[img];
    bool full_scan = lcnt > d.list_cap;
    int nbin = full_scan ? 0 : min(cnt, bin_cap);
    int nlist = full_scan ? (d.f1 - d.f0) : lcnt;
    int total = nbin + nlist;

    for (int base = 0; base < total; base += NT) {
        if (tid == 0) nrec = 0;
        __syncthreads();
        int idx = base + tid;
        if (idx < total) {
            int t = idx < nbin ? bins[(long long)tile * bin_cap + idx]
                               : (full_scan ? d.f0 + (idx - nbin) : lists[(long long)d.list_off + idx - nbin]);
            TriVerts v;
            if (load_tri(pos, tri, t, d, v)) {
                TriSetup s;
                setup_tri(v, d.H, d.W, s);
                bool hit = s.cls != kCull && s.c1 >= tc0 && s.c0 < tc0 + TW && s.r1 >= tr0 && s.r0 < tr0 + TH;
                if (hit) {
                    int k = atomicAdd(&nrec, 1);
                    TileRec &R = rec[k];
                    R.t = t;
                    R.cls = s.cls;
                    R.r0 = s.r0;
                    R.r1 = s.r1;
                    R.v = v;
                    if (s.cls == kDirect) {
                        long long Px = (long long)kSubpix * tc0 + kHalfSub, Py = (long long)kSubpix * tr0 + kHalfSub;
#pragma unroll
                        for (int e = 0; e < 3; ++e) {
                            int j = e == 2 ? 0 : e + 1;
                            int dx = s.X[j] - s.X[e], dy = s.Y[j] - s.Y[e];
                            R.E0[e] = (long long)dx * (Py - s.Y[e]) - (long long)dy * (Px - s.X[e]) - (s.tl[e] ? 0 : 1);
                            R.A[e] = -dy;
                            R.B[e] = dx;
                        }
                    }
                }
            }
        }
        __syncthreads();
        int n = nrec;
        for (int k = 0; k < n; ++k) {
            // wave-uniform row rejection
            int rr0 = __builtin_amdgcn_readfirstlane(rec[k].r0);
            int rr1 = __builtin_amdgcn_readfirstlane(rec[k].r1);
            if (rr1 < wave_r0 || rr0 > wave_r1) continue;
            if (!alive) continue;
            const TileRec &R = rec[k];
            bool cov;
            if (R.cls == kDirect) {
                long long E0 = R.E0[0] + (long long)R.A[0] * oxf + (long long)R.B[0] * oyf;
                long long E1 = R.E0[1] + (long long)R.A[1] * oxf + (long long)R.B[1] * oyf;
                long long E2 = R.E0[2] + (long long)R.A[2] * oxf + (long long)R.B[2] * oyf;
                cov = (E0 | E1 | E2) >= 0;
            } else {
                cov = true;  // decided by the homogeneous test below
            }
            if (!cov) continue;
            TriVerts v = R.v;
            PixEval p;
            eval_pixel(v, px, py, p);
            if (R.cls == kHomog && !homog_inside(p)) continue;
            if (p.s == 0.f) continue;
            float zw = p.zn / p.wn;
            if (!(zw >= -1.f && zw <= 1.f)) continue;
            uint64_t key = frag_key(zw, R.t);
            if (PEEL && key <= prev_key) continue;
            if (key < best) {
                best = key;
                bz = zw;
                bu = p.e[0] / p.s;
                bv = p.e[1] / p.s;
                if (DB) bdb = bary_db(v, p, bu, bv, d.H, d.W);
            }
        }
        __syncthreads();
    }
    if (!inimg) return;
    float4 o;
    if (best != ~0ull) {
        int t = (int)(uint32_t)(best & 0xffffffffu);
        o = make_float4(bu, bv, bz, triidx_to_float(t + 1));
    } else {
        o = make_float4(0.f, 0.f, 0.f, 0.f);
        bdb = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    rast[pix] = o;
    if (DB) rast_db[pix] = bdb;
}

// ----------------------------------------------------------------------------
// rasterize backward: one thread per pixel, gradients scattered to the three
// vertices of the pixel's triangle
// ----------------------------------------------------------------------------
template <bool DB>
__global__ void __launch_bounds__(256) k_raster_bwd(Batch b, const float *__restrict__ pos,
                                                    const int *__restrict__ tri, const float4 *__restrict__ rast,
                                                    const float4 *__restrict__ grad_rast,
                                                    const float4 *__restrict__ grad_db, float *__restrict__ grad_pos,
                                                    long long total_pix, int pix_per_img) {
    long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total_pix) return;
    int img;
    ImageDesc d;
    if (b.descs) {
        // uneven batch: locate the image by its pixel range (images are few)
        img = 0;
        while (img + 1 < b.n_img && b.descs[img + 1].pix_off <= gid) ++img;
        d = b.descs[img];
    } else {
        img = (int)(gid / pix_per_img);
        d = image_desc(b, img);
    }
    float4 rv = rast[gid];
    if (!(rv.w > 0.f)) return;
    int t = float_to_triidx(rv.w) - 1;
    if (t < 0 || t >= b.F) return;
    float4 g = grad_rast[gid];
    float4 gd = DB ? grad_db[gid] : make_float4(0.f, 0.f, 0.f, 0.f);
    if (g.x == 0.f && g.y == 0.f && g.z == 0.f && (!DB || (gd.x == 0.f && gd.y == 0.f && gd.z == 0.f && gd.w == 0.f)))
        return;
    TriVerts v;
    if (!load_tri(pos, tri, t, d, v)) return;
    long long q = gid - d.pix_off;
    int r = (int)(q / d.W), c = (int)(q % d.W);
    float px = pix_ndc(c, d.W), py = pix_ndc(r, d.H);
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
        float kx = 2.f / (float)d.W, ky = 2.f / (float)d.H;
        float Aux = ex[0] - u * sx, Auy = ey[0] - u * sy, Avx = ex[1] - vv * sx, Avy = ey[1] - vv * sy;
        float a = gd.x * kx / p.s, bb = gd.y * ky / p.s, cc = gd.z * kx / p.s, dd = gd.w * ky / p.s;
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
    const int *t3 = tri + 3 * (long long)t;
    float *G = grad_pos + d.pos_off;
    float gzw = gz / p.wn;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float *gp = G + 4 * (long long)t3[k];
        float ek = p.e[k];
        atomicAdd(gp + 0, gX[k]);
        atomicAdd(gp + 1, gY[k]);
        if (gz != 0.f) atomicAdd(gp + 2, gzw * ek);
        atomicAdd(gp + 3, gw[k] - px * gX[k] - py * gY[k] - gzw * zw * ek);
    }
}

// ----------------------------------------------------------------------------
// host launchers
// ----------------------------------------------------------------------------
__global__ void k_zero_i32(int *p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

template <int TW, int TH>
static void launch_fine_t(const RasterLaunch &L, hipStream_t st) {
    int pad = (L.n_tiles + 7) & ~7;
    dim3 grid(pad), block(TW * TH);
    if (L.prev) {
        if (L.rast_db)
            hipLaunchKernelGGL((k_fine<TW, TH, true, true>), grid, block, 0, st, L.b, L.pos, L.tri, L.bin_cap,
                               L.bin_count, L.bins, L.list_count, L.lists, L.n_tiles, pad, (const float4 *)L.prev,
                               (float4 *)L.rast, (float4 *)L.rast_db);
        else
            hipLaunchKernelGGL((k_fine<TW, TH, true, false>), grid, block, 0, st, L.b, L.pos, L.tri, L.bin_cap,
                               L.bin_count, L.bins, L.list_count, L.lists, L.n_tiles, pad, (const float4 *)L.prev,
                               (float4 *)L.rast, (float4 *)nullptr);
    } else {
        if (L.rast_db)
            hipLaunchKernelGGL((k_fine<TW, TH, false, true>), grid, block, 0, st, L.b, L.pos, L.tri, L.bin_cap,
                               L.bin_count, L.bins, L.list_count, L.lists, L.n_tiles, pad, (const float4 *)nullptr,
                               (float4 *)L.rast, (float4 *)L.rast_db);
        else
            hipLaunchKernelGGL((k_fine<TW, TH, false, false>), grid, block, 0, st, L.b, L.pos, L.tri, L.bin_cap,
                               L.bin_count, L.bins, L.list_count, L.lists, L.n_tiles, pad, (const float4 *)nullptr,
                               (float4 *)L.rast, (float4 *)nullptr);
    }
}

int launch_rasterize(const RasterLaunch &L, hipStream_t st) {
    // zero bin counters and spill-list counters (a kernel, not a memset node)
    long long nz = (long long)L.n_tiles + L.b.n_img;
    hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, st, L.bin_count, nz);
    if (L.max_tris > 0) {
        dim3 grid((L.max_tris + 255) / 256, L.b.n_img);
        hipLaunchKernelGGL(k_setup_bin, grid, dim3(256), 0, st, L.b, L.pos, L.tri, L.tile_w, L.tile_h, L.bin_cap,
                           L.bin_count, L.bins, L.list_count, L.lists);
    }
    if (L.tile_w == 8 && L.tile_h == 8) launch_fine_t<8, 8>(L, st);
    else if (L.tile_w == 16 && L.tile_h == 16) launch_fine_t<16, 16>(L, st);
    else if (L.tile_w == 32 && L.tile_h == 8) launch_fine_t<32, 8>(L, st);
    else if (L.tile_w == 16 && L.tile_h == 8) launch_fine_t<16, 8>(L, st);
    else return -1;
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_rasterize_bwd(const RasterBwdLaunch &L, hipStream_t st) {
    if (L.total_pix == 0) return 0;
    dim3 grid((unsigned)((L.total_pix + 255) / 256));
    if (L.grad_db)
        hipLaunchKernelGGL(k_raster_bwd<true>, grid, dim3(256), 0, st, L.b, L.pos, L.tri, (const float4 *)L.rast,
                           (const float4 *)L.grad_rast, (const float4 *)L.grad_db, L.grad_pos, L.total_pix,
                           L.pix_per_img);
    else
        hipLaunchKernelGGL(k_raster_bwd<false>, grid, dim3(256), 0, st, L.b, L.pos, L.tri, (const float4 *)L.rast,
                           (const float4 *)L.grad_rast, (const float4 *)nullptr, L.grad_pos, L.total_pix,
                           L.pix_per_img);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace drp
