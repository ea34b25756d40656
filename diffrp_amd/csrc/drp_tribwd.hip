// Triangle-centric backward passes, lane-balanced (gfx950).
//
// A block owns 256 items, item = (view, triangle): TB triangles x VB views.
// Items whose triangle won no pixel in the forward (hit flag clear) are
// skipped.  The pixel boxes of the others are expanded into candidates, one
// lane per candidate (block prefix sum + search), so lanes stay busy however
// uneven the boxes are.  Each round of 256 candidates:
//   A  a candidate whose pixel shows the item's triangle writes its
//      contribution to LDS (component-major rows, padded: conflict-free);
//   B  one thread per (16-candidate chunk, component) sums its chunk in
//      order, restarting at every item head, and stores each segment sum at
//      the segment's first slot (fixed trip count: no divergence);
//   C  each item adds, in order, the segment sums that start inside its range.
// The summation order depends only on the boxes, so results are reproducible
// bit for bit (no float atomics anywhere).
//
//   POS  (VB = 1): d rast (u, v, z/w [, db]) -> d clip (x, y, z, w) of the 3 corners,
//                  partial [n, F, 3] float4, reduced per vertex by k_gather_pos.
//   ATTR (VB <= 8): (u, v, 1-u-v) * g of the 3 corners, summed over the block's
//                  VB views, partial [groups, F, 3, C], reduced by k_gather_attr.
#include "drp_common.h"
#include "drp_grad.h"
#include "drp_kernels.h"

namespace drp {

constexpr int kItems = 256;
constexpr int kRound = 256;          // candidates per round (one per lane)
constexpr int kChunk = 16;           // candidates per reduction thread
constexpr int kRowStride = kRound + 1;

__device__ __forceinline__ int search_start(const int *cstart, int i) {
    int lo = 0, hi = kItems - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (cstart[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int block_scan256(int v, int *wsum, int &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        int s = wsum[w];
        if (w < wave) pre += s;
        tot += s;
    }
    total = tot;
    return pre + x - v;
}

// floor(a / b) for 0 <= a < b * 65536, 1 <= b <= 65536: float estimate, one fix-up step
// (the estimate's error is below 1/64, so it is off by at most one).
__device__ __forceinline__ int div_box(int a, int b) {
    int q = (int)((float)a * __builtin_amdgcn_rcpf((float)b));
    int r = a - q * b;
    if (r < 0) --q;
    else if (r >= b) ++q;
    return q;
}

// MODE 0: POS (NV = 12), MODE 1: ATTR (NV = 3 C)
template <int MODE, int C, bool DB>
__global__ void __launch_bounds__(256) k_tri_bwd(TriBwdLaunch L) {
    constexpr int NV = MODE == 0 ? 12 : 3 * C;
    __shared__ int cstart[kItems + 1];
    __shared__ int wsum[4];
    __shared__ int2 ibox[kItems];
    __shared__ unsigned char head[kRound];
    __shared__ TriVerts iverts[MODE == 0 ? kItems : 1];
    __shared__ float contrib[NV * kRowStride];
    const int VB = L.views_per_block, TB = kItems / VB;
    const int tid = threadIdx.x;
    const int tblk = blockIdx.x, vgrp = blockIdx.y;
    // this thread's item
    int it_t = tblk * TB + tid % TB, it_v = vgrp * VB + tid / TB;
    bool valid = it_t < L.F && it_v < L.n_img;
    int c0 = 1, c1 = 0, r0 = 0, r1 = -1;
    if (valid) {
        ItemBox b = L.boxes[(long long)it_v * L.F + it_t];
        if (b.hit) {
            c0 = b.cc & 0xffff; c1 = (unsigned)b.cc >> 16; r0 = b.rr & 0xffff; r1 = (unsigned)b.rr >> 16;
        }
    }
    int count = (c0 <= c1) ? (c1 - c0 + 1) * (r1 - r0 + 1) : 0;
    if (MODE == 0 && count > 0) {
        ImageDesc d = image_desc(L.b, it_v);
        if (!load_tri(L.pos, L.tri, it_t, d, iverts[tid])) count = 0;
    }
    ibox[tid] = make_int2(c0 | (c1 << 16), r0 | (r1 << 16));
    int total;
    int ex = block_scan256(count, wsum, total);
    cstart[tid] = ex;
    if (tid == kItems - 1) cstart[kItems] = ex + count;
    __syncthreads();
    float acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = 0.f;
    const int my0 = ex, my1 = ex + count;
    const long long npx = (long long)L.H * L.W;
    for (int base = 0; base < total; base += kRound) {
        // ---- A: contributions of this round's candidates
        int i = base + tid;
        float val[NV];
#pragma unroll
        for (int j = 0; j < NV; ++j) val[j] = 0.f;
        bool is_head = true;
        if (i < total) {
            int k = search_start(cstart, i);
            int2 b = ibox[k];
            int kc0 = b.x & 0xffff, kc1 = (unsigned)b.x >> 16, kr0 = b.y & 0xffff;
            int bw = kc1 - kc0 + 1, local = i - cstart[k];
            is_head = local == 0;
            int dr = div_box(local, bw);
            int r = kr0 + dr, c = kc0 + (local - dr * bw);
            int t = tblk * TB + k % TB, view = vgrp * VB + k / TB;
            long long pix = (long long)view * npx + (long long)r * L.W + c;
            float4 rv = reinterpret_cast<const float4 *>(L.rast)[pix];
            if (rv.w == triidx_to_float(t + 1)) {
                if (MODE == 0) {
                    float4 g = reinterpret_cast<const float4 *>(L.grad_rast)[pix];
                    float4 gd = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (DB) gd = db_grad_scaled(reinterpret_cast<const float4 *>(L.grad_db)[pix], L.H, L.W);
                    float4 o[3];
                    raster_pixel_grad<DB>(iverts[k], c, r, L.H, L.W, g, gd, o);
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        val[4 * q] = o[q].x; val[4 * q + 1] = o[q].y; val[4 * q + 2] = o[q].z; val[4 * q + 3] = o[q].w;
                    }
                } else {
                    float u = rv.x, v = rv.y, w2 = (1.f - u) - v;
                    float gv[C];
                    if (C % 4 == 0) {
                        const float4 *g4 = reinterpret_cast<const float4 *>(L.grad_out + pix * C);
#pragma unroll
                        for (int q = 0; q < C / 4; ++q) {
                            float4 x = g4[q];
                            gv[4 * q] = x.x; gv[4 * q + 1] = x.y; gv[4 * q + 2] = x.z; gv[4 * q + 3] = x.w;
                        }
                    } else {
#pragma unroll
                        for (int ch = 0; ch < C; ++ch) gv[ch] = L.grad_out[pix * C + ch];
                    }
#pragma unroll
                    for (int ch = 0; ch < C; ++ch) {
                        val[ch] = u * gv[ch];
                        val[C + ch] = v * gv[ch];
                        val[2 * C + ch] = w2 * gv[ch];
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < NV; ++j) contrib[j * kRowStride + tid] = val[j];
        head[tid] = is_head;
        __syncthreads();
        // ---- B: segment sums per (chunk, component), stored at each segment's first slot
        for (int w = tid; w < (kRound / kChunk) * NV; w += 256) {
            int j = w % NV, ch = w / NV;  // neighbouring lanes: neighbouring rows (distinct banks)
            float *row = contrib + j * kRowStride;
            int p0 = ch * kChunk, seg = p0;
            float s = row[p0];
#pragma unroll
            for (int p = p0 + 1; p < p0 + kChunk; ++p) {
                float x = row[p];
                if (head[p]) {
                    row[seg] = s;
                    seg = p;
                    s = x;
                } else {
                    s += x;
                }
            }
            row[seg] = s;
        }
        __syncthreads();
        // ---- C: the item adds the segments that start in its range, in order
        int lo = max(my0, base), hi = min(my1, base + kRound);
        if (lo < hi) {
            int p = lo - base, pe = hi - base;
#pragma unroll
            for (int j = 0; j < NV; ++j) acc[j] += contrib[j * kRowStride + p];
            for (int q = (p / kChunk + 1) * kChunk; q < pe; q += kChunk) {
#pragma unroll
                for (int j = 0; j < NV; ++j) acc[j] += contrib[j * kRowStride + q];
            }
        }
        __syncthreads();
    }
    if (MODE == 0) {
        if (!valid) return;
        float4 *P = reinterpret_cast<float4 *>(L.partial) + ((long long)it_v * L.F + it_t) * 3;
        P[0] = make_float4(acc[0], acc[1], acc[2], acc[3]);
        P[1] = make_float4(acc[4], acc[5], acc[6], acc[7]);
        P[2] = make_float4(acc[8], acc[9], acc[10], acc[11]);
    } else {
        // combine the VB views of each triangle in view order
        float *stage = contrib;  // kItems * NV floats fit in NV * kRowStride
#pragma unroll
        for (int j = 0; j < NV; ++j) stage[tid * NV + j] = acc[j];
        __syncthreads();
        if (tid < TB && tblk * TB + tid < L.F) {
            float s[NV];
#pragma unroll
            for (int j = 0; j < NV; ++j) s[j] = 0.f;
            for (int vi = 0; vi < VB; ++vi) {
                const float *q = stage + (vi * TB + tid) * NV;
#pragma unroll
                for (int j = 0; j < NV; ++j) s[j] += q[j];
            }
            float *P = L.partial + ((long long)vgrp * L.F + tblk * TB + tid) * NV;
#pragma unroll
            for (int j = 0; j < NV; ++j) P[j] = s[j];
        }
    }
}

int launch_tri_bwd_pos(const TriBwdLaunch &L, hipStream_t st) {
    dim3 grid((L.F + kItems - 1) / kItems, L.n_img);
    if (L.grad_db) hipLaunchKernelGGL((k_tri_bwd<0, 1, true>), grid, dim3(256), 0, st, L);
    else hipLaunchKernelGGL((k_tri_bwd<0, 1, false>), grid, dim3(256), 0, st, L);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_tri_bwd_attr(const TriBwdLaunch &L, int C, hipStream_t st) {
    int VB = L.views_per_block, TB = kItems / VB;
    dim3 grid((L.F + TB - 1) / TB, (L.n_img + VB - 1) / VB);
    switch (C) {
        case 1: hipLaunchKernelGGL((k_tri_bwd<1, 1, false>), grid, dim3(256), 0, st, L); break;
        case 2: hipLaunchKernelGGL((k_tri_bwd<1, 2, false>), grid, dim3(256), 0, st, L); break;
        case 3: hipLaunchKernelGGL((k_tri_bwd<1, 3, false>), grid, dim3(256), 0, st, L); break;
        case 4: hipLaunchKernelGGL((k_tri_bwd<1, 4, false>), grid, dim3(256), 0, st, L); break;
        case 8: hipLaunchKernelGGL((k_tri_bwd<1, 8, false>), grid, dim3(256), 0, st, L); break;
        case 16: hipLaunchKernelGGL((k_tri_bwd<1, 16, false>), grid, dim3(256), 0, st, L); break;
        default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace drp
