// Rasterizer kernels for gfx950.
//
// Forward: (1) k_setup_bin, one thread per (image, triangle): setup (stored as
// a 32-byte setup record: snapped vertices, class, top-left flags), per-(image,
// triangle) pixel box, and wave-aggregated insertion into fixed-capacity
// screen-tile bins (full bins and large triangles spill to a per-image list);
// (2) k_fine, one workgroup per tile: triangle records staged
// in LDS, then coverage by rows -- lanes take (record, row of its box inside
// the tile) items from a block prefix sum and find the row's exact covered
// span; the covered pixels are queued and evaluated by all lanes -- with the
// depth test resolved by 64-bit LDS atomicMin on (orderable(z/w) << 32 |
// triangle); (u, v) are recomputed for the winner only.  Tiles with many
// entries are split into chunks: (3) k_fine_split drains the chunk queue and
// merges the chunks' depth keys with a device-wide 64-bit minimum.
//
// The setup pass also records each (image, triangle)'s pixel box, and k_fine
// flags the triangles that won a pixel; the deterministic backward
// (drp_tribwd.hip) uses both to find each triangle's per-tile sums, and a CSR
// vertex gather (k_gather_pos) sums the per-corner partials.  A per-pixel
// atomic kernel handles inputs that come without pixel boxes.
#include "drp_common.h"
#include "drp_grad.h"
#include "drp_kernels.h"

namespace drp {

constexpr int kMaxBinTiles = 64;  // larger triangles (and full bins) go to the per-image list that every tile scans
constexpr int kQueuePerLane = 2;  // fine raster: covered-pixel queue entries per lane and pass
constexpr int kFineWavesPerEU = 6;  // occupancy target of k_fine (VGPR budget 512 / 6)
constexpr int kMaxFineNT = 256;     // largest tile (pixels)
constexpr int kFineXcdGroup = 16;   // fine raster: consecutive tiles per XCD turn when tiles are dealt

__device__ __forceinline__ int2 pack_box(int c0, int c1, int r0, int r1) {
    return make_int2(c0 | (c1 << 16), r0 | (r1 << 16));
}
__device__ __forceinline__ int2 empty_box() { return make_int2(0xffff, 0xffff); }  // c0 = 65535 > c1 = 0

__device__ __forceinline__ unsigned long long lanemask_lt() {
    return (1ull << (threadIdx.x & 63)) - 1ull;
}

// ----------------------------------------------------------------------------
// setup records: two int4 per (image, triangle), at the image's list offset
//   direct: {X0, X1, X2, meta}, {Y0, Y1, Y2, 0}  (snapped, counter-clockwise)
//   homogeneous: {c0 | c1 << 16, r0 | r1 << 16, 0, meta}, {0, 0, 0, 0}
//   meta = class | top-left flags << 2 (class kCull: nothing to draw)
// ----------------------------------------------------------------------------
__device__ __forceinline__ void store_setup(int4 *__restrict__ setup, long long i, const TriSetup &s) {
    int4 a = make_int4(0, 0, 0, s.cls), b = make_int4(0, 0, 0, 0);
    if (s.cls == kDirect) {
        a = make_int4(s.X[0], s.X[1], s.X[2], s.cls | (s.tl[0] << 2) | (s.tl[1] << 3) | (s.tl[2] << 4));
        b = make_int4(s.Y[0], s.Y[1], s.Y[2], 0);
    } else if (s.cls == kHomog) {
        a = make_int4(s.c0 | (s.c1 << 16), s.r0 | (s.r1 << 16), 0, s.cls);
    }
    setup[2 * i] = a;
    setup[2 * i + 1] = b;
}

// ----------------------------------------------------------------------------
// setup + binning
// ----------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_setup_bin(Batch b, const float *__restrict__ pos,
                                                   const int *__restrict__ tri, int tile_w, int tile_h,
                                                   int *__restrict__ bin_count,
                                                   int *__restrict__ bins, int *__restrict__ list_count,
                                                   int *__restrict__ lists, int4 *__restrict__ setup,
                                                   ItemBox *__restrict__ boxes) {
    int img = blockIdx.y;
    ImageDesc d = image_desc(b, img);
    int t = d.f0 + blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = t < d.f1;
    TriSetup s;
    s.cls = kCull;
    if (valid) {
        TriVerts v;
        if (load_tri(pos, tri, t, d, v)) setup_tri(v, d.H, d.W, s);
        store_setup(setup, (long long)d.list_off + (t - d.f0), s);
        if (boxes) {
            int2 pb = s.cls == kCull ? empty_box() : pack_box(s.c0, s.c1, s.r0, s.r1);
            boxes[(long long)d.list_off + (t - d.f0)] = ItemBox{pb.x, pb.y, 0};  // hit: set by k_fine
        }
    }
    bool live = valid && s.cls != kCull;
    int tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1, ntiles = 0;
    if (live) {
        tx0 = s.c0 / tile_w; tx1 = s.c1 / tile_w; ty0 = s.r0 / tile_h; ty1 = s.r1 / tile_h;
        ntiles = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
    }
    bool spill = live && ntiles > kMaxBinTiles;
    int mine = (live && !spill) ? ntiles : 0;
    int ncols = tx1 - tx0 + 1;
    const int lane = threadIdx.x & 63;
    // The wave's (triangle, tile) pairs are numbered in lane order (lane i's
    // triangle owns pairs [excl, incl)) and inserted 64 per round, so a round
    // is full whatever the spread of tile counts.  Each lane finds its pair's
    // triangle by a binary search over the lanes' inclusive counts; the wave
    // sorts the pairs' (tile, lane) keys; the first lane of each run of one
    // tile reserves the run's slots with one atomic -- all runs at once -- and
    // every sorted lane writes the triangle of the pair its key came from.  A
    // full bin hands the flag back to the pair's lane, and a triangle with a
    // full pair goes to the spill list.  (Every shuffle and ballot below runs
    // in wave-uniform control flow.)
    constexpr unsigned kNoTile = (1u << 26) - 1;  // tile field of a lane with nothing to insert
    int incl = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    const int excl = incl - mine;
    const int total = __shfl(incl, 63, 64);
    for (int r0 = 0; r0 < total; r0 += 64) {
        const int p = r0 + lane;
        const bool active = p < total;
        int sl = 0;  // the pair's lane: the number of lanes whose pairs end at or before p
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1) {
            const int v = __shfl(incl, sl + step - 1, 64);
            if (v <= p) sl += step;
        }
        const int k = p - __shfl(excl, sl, 64);
        const int nc = __shfl(ncols, sl, 64), px0 = __shfl(tx0, sl, 64), py0 = __shfl(ty0, sl, 64);
        const int tp = __shfl(t, sl, 64);
        unsigned key = (kNoTile << 6) | (unsigned)lane;  // unique: the sorted lanes are a permutation
        if (active) {
            const int ty = py0 + k / nc, tx = px0 + k % nc;
            key = ((unsigned)(ty * d.tiles_x + tx) << 6) | (unsigned)lane;
        }
        const unsigned sk = wave_sort64<unsigned>(key);
        const unsigned tl = sk >> 6;
        const int src = (int)(sk & 63);
        const bool valid = tl != kNoTile;
        const unsigned prev = __shfl_up(sk, 1, 64);
        const bool head = valid && (lane == 0 || (prev >> 6) != tl);
        const unsigned long long heads = __ballot(head);
        const int nvalid = __popcll(__ballot(valid));
        const unsigned long long upto = (2ull << lane) - 1;  // lanes <= this one
        const int h = valid ? 63 - __clzll(heads & upto) : 0;
        const unsigned long long after = heads & ~upto;
        const int e = after ? __ffsll(after) - 1 : nvalid;  // end of this lane's run
        int base = 0;
        if (head) base = atomicAdd(&bin_count[d.tile_off + tl], e - lane);
        base = __shfl(base, h, 64);
        const int tsrc = __shfl(tp, src, 64);
        bool full = false;
        if (valid) {
            const int slot = base + (lane - h);
            if (slot < d.bin_cap) bins[d.bin_off + (long long)tl * d.bin_cap + slot] = tsrc;
            else full = true;
        }
        const bool pair_full = __builtin_amdgcn_ds_permute(src << 2, full ? 1 : 0) != 0;
        const unsigned long long fm = __ballot(pair_full && active);
        // this lane's triangle's pairs of this round: lanes [lo, hi)
        const int lo = max(excl, r0) - r0, hi = min(incl, r0 + 64) - r0;
        if (hi > lo) {
            const unsigned long long m = (hi - lo == 64 ? ~0ull : ((1ull << (hi - lo)) - 1)) << lo;
            if (fm & m) spill = true;
        }
    }
    // spill list (per image): wave-aggregated as well
    unsigned long long sm = __ballot(spill);
    if (sm) {
        int leader = __ffsll(sm) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(&list_count[img], __popcll(sm));
        base = __shfl(base, leader);
        if (spill) {
            int slot = base + __popcll(sm & lanemask_lt());
            if (slot < d.list_cap) lists[(long long)d.list_off + slot] = t;
        }
    }
}

// ----------------------------------------------------------------------------
// fine raster
// ----------------------------------------------------------------------------
struct FineRec {
    int X[3], Y[3];   // snapped vertices minus the tile's first pixel centre (sub-pixels; direct class)
    TriVerts v;
    int t;
    int box;          // tile-local box columns x0, x1 and first row y0 (6 bits each), class << 18, top-left
                      // flags << 20, small-coordinate flag << 23
};
__device__ __forceinline__ int rec_x0(int box) { return box & 63; }
__device__ __forceinline__ int rec_y0(int box) { return (box >> 6) & 63; }
__device__ __forceinline__ int rec_x1(int box) { return (box >> 12) & 63; }
__device__ __forceinline__ int rec_cls(int box) { return (box >> 18) & 3; }

// Covered columns of tile row y of a direct-class record, intersected with
// [xl, xr]: every edge e -> j needs E_e(x) = C_e + a_e x >= 0 with
// C_e = dx (16 y - Y_e) + dy X_e - [not top-left], a_e = -16 dy (dx, dy the
// edge vector, vertices relative to the tile's first pixel centre; 64-bit).  The crossing -C_e / a_e is estimated in floating
// point (clamped to [xl - 1, xr + 1], where its error is far below one
// column) and settled by one exact 64-bit step either way, so the span is
// exactly the set of pixels whose three edge functions are non-negative.
__device__ __forceinline__ void row_span(const FineRec &R, int box, int y, int &xl, int &xr) {
    // branch-free per edge: the lanes of a wave take edges of either direction
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const int j = e == 2 ? 0 : e + 1;
        const int dx = R.X[j] - R.X[e], dy = R.Y[j] - R.Y[e];
        const long long a = -(long long)dy * kSubpix;
        const long long cc = (long long)dx * (kSubpix * y - R.Y[e]) + (long long)dy * R.X[e] -
                             (((box >> (20 + e)) & 1) ? 0 : 1);
        const bool flat = a == 0, rising = a > 0;  // rising: covered for x >= crossing, else x <= crossing
        const float den = flat ? 1.0f : (float)(-dy) * (float)kSubpix;
        double q = -(double)cc * (double)__builtin_amdgcn_rcpf(den);
        q = fmin(fmax(q, (double)(xl - 1)), (double)(xr + 1));
        int x0 = (int)(rising ? ceil(q) : floor(q));
        const int d = rising ? -1 : 1;  // the step towards the uncovered side
        if (cc + a * (x0 + d) >= 0) x0 += d;
        else if (cc + a * x0 < 0) x0 -= d;
        const int nxl = flat ? xl : (rising ? max(xl, x0) : xl);
        const int nxr = flat ? (cc < 0 ? xl - 1 : xr) : (rising ? xr : min(xr, x0));
        xl = nxl;
        xr = nxr;
    }
}


// The same span for records whose vertices lie within 2^14 sub-pixels of the
// tile's first pixel centre (flag bit 23 of the box): every product then fits
// in 32 bits (|C_e| < 2^30.2, |a_e x| < 2^24.1 for x in [-2, 33]) and a float
// estimate suffices -- a crossing inside the clamp window has |C_e| < 2^24.1,
// exact in float, so the estimate is off by far less than one column, and
// one exact 32-bit step either way settles it (tests/test_rowspan_math.py
// models this arithmetic against brute-force coverage: slivers and the
// coordinate limit included, the reciprocal perturbed by one ulp).
constexpr int kSmallRec = 1 << 14;
__device__ __forceinline__ void row_span32(const FineRec &R, int box, int y, int &xl, int &xr) {
    // branch-free per edge, the same decisions as row_span in 32-bit arithmetic
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const int j = e == 2 ? 0 : e + 1;
        const int dx = R.X[j] - R.X[e], dy = R.Y[j] - R.Y[e];
        const int a = -dy * kSubpix;
        const int cc = dx * (kSubpix * y - R.Y[e]) + dy * R.X[e] - (((box >> (20 + e)) & 1) ? 0 : 1);
        const bool flat = a == 0, rising = a > 0;
        const float den = flat ? 1.0f : (float)(-dy) * (float)kSubpix;
        float q = -(float)cc * __builtin_amdgcn_rcpf(den);
        q = fminf(fmaxf(q, (float)(xl - 1)), (float)(xr + 1));
        int x0 = (int)(rising ? ceilf(q) : floorf(q));
        const int d = rising ? -1 : 1;
        if (cc + a * (x0 + d) >= 0) x0 += d;
        else if (cc + a * x0 < 0) x0 -= d;
        const int nxl = flat ? xl : (rising ? max(xl, x0) : xl);
        const int nxr = flat ? (cc < 0 ? xl - 1 : xr) : (rising ? xr : min(xr, x0));
        xl = nxl;
        xr = nxr;
    }
}

template <int NT>
__device__ __forceinline__ int block_excl_scan(int v, int *wsum, int &total) {
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
    for (int w = 0; w < NT / 64; ++w) {
        int s = wsum[w];
        if (w < wave) pre += s;
        tot += s;
    }
    total = tot;
    return pre + x - v;
}

__device__ __forceinline__ float inv_ord_f32(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Depth cull of a direct-class triangle against tile rows whose every pixel
// already holds a fragment: true when no pixel centre of the box
// [px0, px1] x [py0, py1] can receive from this triangle a depth z/w at or
// below zcut, the farthest current nearest depth over the box's rows, so that
// the triangle cannot change any depth key (the test is a strict >, whatever
// the triangle index).
//
// With X_k = x_k - w_k px, Y_k = y_k - w_k py and e_k = X_i Y_j - X_j Y_i
// ((i, j) = (k+1, k+2) mod 3), wn* = sum e_k w_k is the determinant of the
// rows (x_k, y_k, w_k), the same at every pixel, and zn* = sum e_k z_k is
// affine in (px, py), so the exact depth zn* / wn* is affine over the box:
// its minimum is at a corner.  eval_pixel's float zn and wn (X, Y one
// rounding each; e_k two products and an fma: 4.02 u T_k; the nested fmas:
// 3.01 u sum |e_k z_k|) are within Ez = 7.04 u sum T_k |z_k| and
// Ew = 7.04 u sum T_k w_k of the exact ones at any pixel centre of the box,
// with u = 2^-24 and T_k = |X_i Y_j| + |X_j Y_i| bounded through the box's
// end columns and rows (X_k is affine in px, Y_k in py); 8 u in place of
// 7.04 u covers the roundings of the bound itself.  The corners are evaluated
// with eval_pixel too, and their depths taken as zn * rcp(wn) (within 4 u of
// zn / wn).  With W0 = |wn(corner 0)| - Ew <= |wn*| > 3 Ew (slivers are never
// culled), every |wn| >= W0 - Ew > 0; the exact depth over the box is at most
// ZB = 2 (QA (1 + 5 u) + Ez / (W0 - Ew)) in magnitude (QA: the corners'
// largest |depth|), a pixel's computed depth is within E + u (ZB + E) of the
// exact one, E = (Ez + ZB Ew) / (W0 - Ew), and the exact minimum is within
// E + 4 u QA (1 + 5 u) of the corners' smallest depth: the margin below
// (2 E + 8 u QA + 2 u (ZB + E)) exceeds the sum.
__device__ __forceinline__ bool depth_culled(const TriVerts &v, float px0, float px1, float py0, float py1,
                                             float zcut) {
    constexpr float u = 0x1p-24f;
    float ez = 0.f, ew = 0.f;
    {
        float MX[3], MY[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            MX[k] = fmaxf(fabsf(fmaf(-v.w[k], px0, v.x[k])), fabsf(fmaf(-v.w[k], px1, v.x[k])));
            MY[k] = fmaxf(fabsf(fmaf(-v.w[k], py0, v.y[k])), fabsf(fmaf(-v.w[k], py1, v.y[k])));
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int i = k == 2 ? 0 : k + 1, j = k == 0 ? 2 : k - 1;
            const float T = fmaf(MX[i], MY[j], MX[j] * MY[i]);
            ez = fmaf(T, fabsf(v.z[k]), ez);
            ew = fmaf(T, v.w[k], ew);
        }
        ez *= 8.f * u;
        ew *= 8.f * u;
    }
    float qmin = INFINITY, qa = 0.f, w0 = 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        PixEval p;
        eval_pixel(v, (c & 1) ? px1 : px0, (c & 2) ? py1 : py0, p);
        if (c == 0) w0 = fabsf(p.wn) - ew;
        const float q = p.zn * __builtin_amdgcn_rcpf(p.wn);
        qmin = fminf(qmin, q);
        qa = fmaxf(qa, fabsf(q));
    }
    if (!(w0 > 3.f * ew) || !(qa < INFINITY)) return false;
    const float zb = 2.f * (qa * (1.f + 5.f * u) + ez / (w0 - ew));
    const float E = (ez + zb * ew) / (w0 - ew);
    return qmin - (2.f * E + 8.f * u * qa + 2.f * u * (zb + E)) > zcut;
}

// LDS of one fine-raster block
template <int TW, int TH, bool PEEL>
struct FineLds {
    static constexpr int NT = TW * TH;
    static constexpr int kQueue = NT * kQueuePerLane;
    FineRec rec[NT];
    unsigned short rowmap[NT];          // row item of a sub-round -> record << 5 | row within the record
    int queue[kQueue];                  // covered pixels: record | x << 8 | y << 16 (tile-local)
    unsigned long long zbuf[NT];        // depth-test keys
    unsigned long long pkey[PEEL ? NT : 1];
    float pxs[TW], pys[TH];             // NDC centres of the tile's columns and rows (one division each per tile)
    int wsum[3 * (NT / 64)];            // scan partials: records; row spans (two, alternating)
    float zrow[TH];                     // per tile row: the farthest current nearest depth (depth cull)
    int bcast;                          // block-wide broadcast slot
};

// one tile's view of the work: its image, its pixel for this lane, and its
// entry list (bin entries, then the image's spill list or every triangle)
struct FineTile {
    ImageDesc d;
    int tile, tc0, tr0, tw_img, th_img;
    int nbin, total;
    bool full_scan;
};

__device__ __forceinline__ int rfl(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ long long rfl(long long x) {
    return (long long)(((unsigned long long)(unsigned)rfl((int)(x >> 32)) << 32) | (unsigned)rfl((int)x));
}
// the block's image descriptor in scalar registers (the compiler cannot tell
// that a load after the block's own stores is uniform)
__device__ __forceinline__ ImageDesc uniform_desc(const ImageDesc &s) {
    ImageDesc d;
    d.pos_off = rfl(s.pos_off);
    d.pix_off = rfl(s.pix_off);
    d.V = rfl(s.V);
    d.f0 = rfl(s.f0);
    d.f1 = rfl(s.f1);
    d.H = rfl(s.H);
    d.W = rfl(s.W);
    d.tiles_x = rfl(s.tiles_x);
    d.tiles_y = rfl(s.tiles_y);
    d.tile_off = rfl(s.tile_off);
    d.list_off = rfl(s.list_off);
    d.list_cap = rfl(s.list_cap);
    d.bin_off = rfl(s.bin_off);
    d.bin_cap = rfl(s.bin_cap);
    return d;
}

template <int TW, int TH, bool PEEL>
__device__ __forceinline__ FineTile fine_begin(FineLds<TW, TH, PEEL> &S, const Batch &b, int tile,
                                               const int *__restrict__ bin_count,
                                               const int *__restrict__ list_count, const float4 *__restrict__ prev) {
    FineTile T;
    T.tile = tile;
    const int img = rfl(tile_image(b, tile));
    T.d = uniform_desc(image_desc(b, img));
    const ImageDesc &d = T.d;
    const int lt = tile - d.tile_off;
    const int tx = lt % d.tiles_x, ty = lt / d.tiles_x;
    const int tid = threadIdx.x;
    T.tc0 = tx * TW;
    T.tr0 = ty * TH;
    S.zbuf[tid] = ~0ull;
    if (tid < TW) S.pxs[tid] = pix_ndc(T.tc0 + tid, d.W);
    if (tid < TH) S.pys[tid] = pix_ndc(T.tr0 + tid, d.H);
    if (PEEL) {
        unsigned long long pk = ~0ull;  // empty previous layer: nothing can follow
        const int c = T.tc0 + (tid % TW), r = T.tr0 + (tid / TW);
        if (c < d.W && r < d.H) {
            float4 pv = prev[d.pix_off + (long long)r * d.W + c];
            if (pv.w > 0.f) pk = frag_key(pv.z, float_to_triidx(pv.w) - 1 + (b.local_ids ? d.f0 : 0));
        }
        S.pkey[tid] = pk;
    }
    T.tw_img = min(TW, d.W - T.tc0);  // tile clipped to the image
    T.th_img = min(TH, d.H - T.tr0);
    const int cnt = rfl(bin_count[tile]);
    const int lcnt = rfl(list_count[img]);
    T.full_scan = lcnt > d.list_cap;
    T.nbin = T.full_scan ? 0 : min(cnt, d.bin_cap);
    const int nlist = T.full_scan ? (d.f1 - d.f0) : lcnt;
    T.total = T.nbin + nlist;
    return T;
}

// depth test of the tile's entries [e0, e1) into S.zbuf; with CULL, after the
// first round (from the first when S.zbuf was seeded with real keys),
// direct-class triangles are culled against the tile's depths
template <int TW, int TH, bool PEEL, bool CULL>
__device__ __forceinline__ void fine_entries(FineLds<TW, TH, PEEL> &S, const FineTile &T, int e0, int e1,
                                             const float *__restrict__ pos, const int *__restrict__ tri,
                                             const int *__restrict__ bins, const int *__restrict__ lists,
                                             const int4 *__restrict__ setup, bool seeded) {
    constexpr int NT = TW * TH;
    constexpr int kQueue = NT * kQueuePerLane;
    const ImageDesc &d = T.d;
    const int tid = threadIdx.x;
    const int tc0 = T.tc0, tr0 = T.tr0;
    int par = 0;  // scan partials of the row spans in use; alternates every sub-round
    for (int base = e0; base < e1; base += NT) {
        // After the first round: per tile row, the farthest of its current
        // nearest depths (infinite while a pixel of the row has no fragment
        // yet), against which direct-class triangles are culled over the rows
        // of their box (depth_culled).
        const bool cull_round = CULL && (base > e0 || seeded);
        if (cull_round) {
            float zt = -INFINITY;
            if ((tid % TW) < T.tw_img && (tid / TW) < T.th_img) {
                const unsigned long long key = S.zbuf[tid];
                zt = key == ~0ull ? INFINITY : inv_ord_f32((uint32_t)(key >> 32));
            }
#pragma unroll
            for (int o = TW / 2; o >= 1; o >>= 1) zt = fmaxf(zt, __shfl_xor(zt, o, 64));  // a row: TW lanes
            if (tid % TW == 0) S.zrow[tid / TW] = zt;
            __syncthreads();
        }
        int idx = base + tid;
        int count = 0;
        if (idx < e1) {
            int t;
            if (idx < T.nbin) t = bins[d.bin_off + (long long)(T.tile - d.tile_off) * d.bin_cap + idx];
            else if (T.full_scan) t = d.f0 + (idx - T.nbin);
            else t = lists[(long long)d.list_off + idx - T.nbin];
            if (t >= d.f0 && t < d.f1) {  // every entry names a triangle of this image
                // both setup words and the corner indices depend on t alone: one round of loads,
                // then the corners' positions
                const long long si = 2 * ((long long)d.list_off + (t - d.f0));
                const int4 sa = setup[si], sb = setup[si + 1];
                const int *t3 = tri + 3 * (long long)t;
                const int i0 = t3[0], i1 = t3[1], i2 = t3[2];
                const bool in_range = (unsigned)i0 < (unsigned)d.V && (unsigned)i1 < (unsigned)d.V &&
                                      (unsigned)i2 < (unsigned)d.V;
                TriVerts v;
                if (in_range) {
                    const float4 *P = reinterpret_cast<const float4 *>(pos + d.pos_off);
                    const float4 a = P[i0], b = P[i1], c = P[i2];
                    v.x[0] = a.x; v.y[0] = a.y; v.z[0] = a.z; v.w[0] = a.w;
                    v.x[1] = b.x; v.y[1] = b.y; v.z[1] = b.z; v.w[1] = b.w;
                    v.x[2] = c.x; v.y[2] = c.y; v.z[2] = c.z; v.w[2] = c.w;
                }
                const int cls = sa.w & 3;
                if (cls != kCull) {
                    int c0, c1, r0, r1;
                    if (cls == kDirect) {
                        const int X[3] = {sa.x, sa.y, sa.z}, Y[3] = {sb.x, sb.y, sb.z};
                        direct_box(X, Y, d.H, d.W, c0, c1, r0, r1);  // not empty: the setup pass kept it
                    } else {
                        c0 = sa.x & 0xffff; c1 = (unsigned)sa.x >> 16;
                        r0 = sa.y & 0xffff; r1 = (unsigned)sa.y >> 16;
                    }
                    const int x0 = max(c0 - tc0, 0), x1 = min(c1 - tc0, T.tw_img - 1);
                    const int y0 = max(r0 - tr0, 0), y1 = min(r1 - tr0, T.th_img - 1);
                    if (x0 <= x1 && y0 <= y1 && in_range) {
                        FineRec &R = S.rec[tid];
                        R.t = t;
                        R.v = v;
                        count = y1 - y0 + 1;  // row items
                        int box = x0 | (y0 << 6) | (x1 << 12) | (cls << 18);
                        if (cls == kDirect) {
                            const int Px = kSubpix * tc0 + kHalfSub, Py = kSubpix * tr0 + kHalfSub;
                            R.X[0] = sa.x - Px; R.X[1] = sa.y - Px; R.X[2] = sa.z - Px;
                            R.Y[0] = sb.x - Py; R.Y[1] = sb.y - Py; R.Y[2] = sb.z - Py;
                            box |= ((sa.w >> 2) & 7) << 20;  // top-left flags
                            const int m = max(max(max(abs(R.X[0]), abs(R.X[1])), abs(R.X[2])),
                                              max(max(abs(R.Y[0]), abs(R.Y[1])), abs(R.Y[2])));
                            if (m < kSmallRec) box |= 1 << 23;  // 32-bit row spans
                        }
                        R.box = box;
                    }
                }
            }
        }
        if (cull_round && count > 0) {  // the lane's own record, read back from LDS
            const FineRec &R = S.rec[tid];
            const int box = R.box, x0 = rec_x0(box), y0 = rec_y0(box);
            float zcut = -INFINITY;
            for (int y = y0; y < y0 + count; ++y) zcut = fmaxf(zcut, S.zrow[y]);
            if (rec_cls(box) == kDirect && zcut < INFINITY &&
                depth_culled(R.v, S.pxs[x0], S.pxs[rec_x1(box)], S.pys[y0], S.pys[y0 + count - 1], zcut))
                count = 0;
        }
        int nrows;
        const int ex = block_excl_scan<NT>(count, S.wsum, nrows);  // this lane's record owns rows [ex, ex + count)
        // Row items (record, tile row) in sub-rounds of NT: each lane finds the
        // exact covered span of its row; the covered pixels are queued (block
        // prefix sum, at most kQueue per pass) and evaluated by all lanes, so
        // the barycentric and depth work runs without divergence.  The depth
        // test is a minimum, so the queue order does not matter.
        for (int sb = 0; sb < nrows; sb += NT, par ^= 1) {
            // each record marks its rows of this sub-round's window (at most 32 rows per record)
            for (int rr = max(ex, sb), re = min(ex + count, sb + NT); rr < re; ++rr)
                S.rowmap[rr - sb] = (unsigned short)((tid << 5) | (rr - ex));
            __syncthreads();
            const int ri = sb + tid;
            int len = 0, k = 0, y = 0, xl = 0;
            if (ri < nrows) {
                const int m = S.rowmap[tid];
                k = m >> 5;
                const FineRec &R = S.rec[k];
                const int box = R.box;
                y = rec_y0(box) + (m & 31);
                xl = rec_x0(box);
                int xr = rec_x1(box);
                if (rec_cls(box) == kDirect) {
                    if ((box >> 23) & 1) row_span32(R, box, y, xl, xr);
                    else row_span(R, box, y, xl, xr);
                }
                len = max(xr - xl + 1, 0);
            }
            int npix;
            const int off = block_excl_scan<NT>(len, S.wsum + (NT / 64) * (1 + par), npix);
            for (int q0 = 0; q0 < npix; q0 += kQueue) {
                const int lo = max(off, q0), hi = min(off + len, q0 + kQueue);
                for (int i = lo; i < hi; ++i) S.queue[i - q0] = k | ((xl + i - off) << 8) | (y << 16);
                __syncthreads();
                const int nq = min(kQueue, npix - q0);
                for (int qi = tid; qi < nq; qi += NT) {
                    int en = S.queue[qi];
                    const FineRec &R = S.rec[en & 255];
                    int lx = (en >> 8) & 255, ly = en >> 16;
                    TriVerts v = R.v;
                    float px = S.pxs[lx], py = S.pys[ly];
                    PixEval p;
                    eval_pixel(v, px, py, p);
                    if ((rec_cls(R.box) != kHomog || homog_inside(p)) && p.s != 0.f) {
                        float zw = p.zn / p.wn;
                        if (zw >= -1.f && zw <= 1.f) {
                            unsigned long long key = frag_key(zw, R.t);
                            int q = ly * TW + lx;
                            if (!PEEL || key > S.pkey[q]) atomicMin(&S.zbuf[q], key);
                        }
                    }
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
}

// output of this lane's pixel from the winning key (u, v recomputed for the winner)
template <int TW, bool DB>
__device__ __forceinline__ void fine_output(const FineTile &T, unsigned long long key, const Batch &b,
                                            const float *__restrict__ pos, const int *__restrict__ tri,
                                            float4 *__restrict__ rast, float4 *__restrict__ rast_db,
                                            ItemBox *__restrict__ boxes, float px, float py) {
    const ImageDesc &d = T.d;
    const int c = T.tc0 + ((int)threadIdx.x % TW), r = T.tr0 + ((int)threadIdx.x / TW);
    if (c >= d.W || r >= d.H) return;
    const long long pix = d.pix_off + (long long)r * d.W + c;
    if (key == ~0ull) {
        rast[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (DB) rast_db[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    int t = (int)(uint32_t)(key & 0xffffffffu);
    float zw = inv_ord_f32((uint32_t)(key >> 32));
    TriVerts v;
    load_tri(pos, tri, t, d, v);
    PixEval p;
    eval_pixel(v, px, py, p);
    float u = p.e[0] / p.s, vv = p.e[1] / p.s;
    rast[pix] = make_float4(u, vv, zw, triidx_to_float(t + 1 - (b.local_ids ? d.f0 : 0)));
    if (DB) rast_db[pix] = bary_db(v, p, u, vv, d.H, d.W);
    if (boxes) boxes[(long long)d.list_off + (t - d.f0)].hit = 1;  // same value from every winning pixel
}

// Heavy tiles (more than kSplitRounds rounds of entries) are split into
// chunks of at least kChunkRounds rounds, at most kMaxChunks per tile: the
// tile's own block takes chunk 0 and queues the rest for k_fine_split, whose
// blocks merge their depth keys with device-scope 64-bit atomicMin; the last
// chunk to finish writes the tile's output.  The minimum is order-free, so
// the result is the same as one block's.
constexpr int kChunkRounds = 4;
constexpr int kSplitRounds = 8;
constexpr int kMaxChunks = 64;

// Split bookkeeping: three control words after the image's list counters
// (slots taken, queue length, queue head) and one workspace region of
// `slots` entries: int4 info (tile, chunk size, chunk count), int done
// counter, 63 queued chunks (slot << 8 | chunk), kMaxFineNT depth keys.
struct SplitWs {
    int *ctl;
    int4 *info;
    int *done;
    int *queue;
    unsigned long long *gkey;
};
__device__ __forceinline__ SplitWs split_ws(int *list_count, int n_img, char *base, int slots) {
    SplitWs w;
    w.ctl = list_count + n_img;
    w.info = reinterpret_cast<int4 *>(base);
    w.done = reinterpret_cast<int *>(base + 16 * (size_t)slots);
    w.queue = w.done + slots;
    w.gkey = reinterpret_cast<unsigned long long *>(base + 272 * (size_t)slots);  // after 16 + 4 + 252 bytes per slot
    return w;
}
static_assert(kMaxFineNT == 256, "split_region_bytes assumes 256 keys per slot");

template <int TW, int TH, bool PEEL, bool DB, bool CULL>
__global__ void __launch_bounds__(TW *TH, kFineWavesPerEU) k_fine(Batch b, const float *__restrict__ pos,
                                                 const int *__restrict__ tri,
                                                 const int *__restrict__ bin_count, const int *__restrict__ bins,
                                                 int *__restrict__ list_count, const int *__restrict__ lists,
                                                 const int4 *__restrict__ setup, int n_tiles, int xcd_group,
                                                 char *split_base, int split_slots, const float4 *__restrict__ prev,
                                                 float4 *__restrict__ rast, float4 *__restrict__ rast_db,
                                                 ItemBox *__restrict__ boxes) {
    constexpr int NT = TW * TH;
    static_assert(NT % 64 == 0, "tile must be a whole number of waves");
    __shared__ FineLds<TW, TH, PEEL> S;
    int tile = xcd_deal(blockIdx.x, xcd_group);
    if (tile >= n_tiles) return;
    const FineTile T = fine_begin<TW, TH, PEEL>(S, b, tile, bin_count, list_count, prev);
    int e1 = T.total;
    int slot = -1;
    if (split_slots > 0 && T.total > kSplitRounds * NT) {
        const int per = NT * max(kChunkRounds, T.total / (NT * kMaxChunks) + 1);
        const int nch = (T.total + per - 1) / per;  // >= 2
        if (threadIdx.x == 0) {
            const SplitWs sp = split_ws(list_count, b.n_img, split_base, split_slots);
            int s = atomicAdd(&sp.ctl[0], 1);
            if (s < split_slots) {
                sp.info[s] = make_int4(tile, per, nch, 0);
                sp.done[s] = 0;
                const int q = atomicAdd(&sp.ctl[1], nch - 1);  // at most (kMaxChunks - 1) per slot: fits
                for (int k = 1; k < nch; ++k) sp.queue[q + k - 1] = (s << 8) | k;
            } else {
                s = -1;
            }
            S.bcast = s;
        }
        __syncthreads();
        slot = __builtin_amdgcn_readfirstlane(S.bcast);
        if (slot >= 0) e1 = per;
    }
    fine_entries<TW, TH, PEEL, CULL>(S, T, 0, e1, pos, tri, bins, lists, setup, false);
    if (slot >= 0) {  // chunk 0 of a split tile: its keys seed the merge
        const SplitWs sp = split_ws(list_count, b.n_img, split_base, split_slots);
        sp.gkey[(long long)slot * kMaxFineNT + threadIdx.x] = S.zbuf[threadIdx.x];
        return;
    }
    fine_output<TW, DB>(T, S.zbuf[threadIdx.x], b, pos, tri, rast, rast_db, boxes, S.pxs[threadIdx.x % TW],
                        S.pys[threadIdx.x / TW]);
}

// chunks 1.. of split tiles: a fixed grid drains the queue k_fine filled
template <int TW, int TH, bool PEEL, bool DB, bool CULL>
__global__ void __launch_bounds__(TW *TH, kFineWavesPerEU) k_fine_split(Batch b, const float *__restrict__ pos,
                                                 const int *__restrict__ tri,
                                                 const int *__restrict__ bin_count, const int *__restrict__ bins,
                                                 int *__restrict__ list_count, const int *__restrict__ lists,
                                                 const int4 *__restrict__ setup, char *split_base,
                                                 int split_slots, const float4 *__restrict__ prev,
                                                 float4 *__restrict__ rast, float4 *__restrict__ rast_db,
                                                 ItemBox *__restrict__ boxes) {
    __shared__ FineLds<TW, TH, PEEL> S;
    const SplitWs sp = split_ws(list_count, b.n_img, split_base, split_slots);
    const int nq = sp.ctl[1];  // final: written by the previous kernel
    if (nq == 0) return;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) S.bcast = atomicAdd(&sp.ctl[2], 1);
        __syncthreads();
        // the item is the same for every lane: keep it (and what follows from it) in scalar registers
        const int it = __builtin_amdgcn_readfirstlane(S.bcast);
        if (it >= nq) return;
        const int e = __builtin_amdgcn_readfirstlane(sp.queue[it]);
        const int slot = e >> 8, chunk = e & 255;
        const int4 inf = sp.info[slot];
        const int itile = __builtin_amdgcn_readfirstlane(inf.x), per = __builtin_amdgcn_readfirstlane(inf.y),
                  nch = __builtin_amdgcn_readfirstlane(inf.z);
        const FineTile T = fine_begin<TW, TH, PEEL>(S, b, itile, bin_count, list_count, prev);
        const int e0 = chunk * per, e1 = min(e0 + per, T.total);
        unsigned long long *g = sp.gkey + (long long)slot * kMaxFineNT + tid;
        // With the cull, the chunk starts from the tile's keys merged so far
        // (chunk 0's, stored before this kernel began, and any chunk merged
        // since; an atomic load, so never half of one key and half of
        // another): each is a real fragment's key, so the depth test and the
        // cull see bounds of the final minimum from the first round.
        if (CULL) S.zbuf[tid] = __hip_atomic_load(g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        fine_entries<TW, TH, PEEL, CULL>(S, T, e0, e1, pos, tri, bins, lists, setup, CULL);
        const unsigned long long mine = S.zbuf[tid];
        if (mine != ~0ull) atomicMin(g, mine);
        __threadfence();
        __syncthreads();
        if (tid == 0) S.bcast = atomicAdd(&sp.done[slot], 1);
        __syncthreads();
        const bool last = S.bcast == nch - 2;  // chunks 1 .. nch-1 run here
        if (last) {
            __threadfence();
            const unsigned long long key = atomicMin(g, ~0ull);  // read at the point of coherence
            fine_output<TW, DB>(T, key, b, pos, tri, rast, rast_db, boxes, S.pxs[tid % TW], S.pys[tid / TW]);
        }
        __syncthreads();  // S is reused by the next item
    }
}

// per-pixel fallback: atomics into grad_pos
template <bool DB>
__global__ void __launch_bounds__(256) k_raster_bwd(Batch b, const float *__restrict__ pos,
                                                    const int *__restrict__ tri, const float4 *__restrict__ rast,
                                                    const float4 *__restrict__ grad_rast,
                                                    const float4 *__restrict__ grad_db, float *__restrict__ grad_pos,
                                                    long long total_pix, int pix_per_img) {
    long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total_pix) return;
    int img = (int)(gid / pix_per_img);
    ImageDesc d = image_desc(b, img);
    float4 rv = rast[gid];
    if (!(rv.w > 0.f)) return;
    int t = float_to_triidx(rv.w) - 1;
    if (t < 0 || t >= b.F) return;
    float4 g = grad_rast[gid];
    float4 gd = DB ? db_grad_scaled(grad_db[gid], d.H, d.W) : make_float4(0.f, 0.f, 0.f, 0.f);
    if (g.x == 0.f && g.y == 0.f && g.z == 0.f && (!DB || (gd.x == 0.f && gd.y == 0.f && gd.z == 0.f && gd.w == 0.f)))
        return;
    TriVerts v;
    if (!load_tri(pos, tri, t, d, v)) return;
    long long q = gid - d.pix_off;
    int r = (int)(q / d.W), c = (int)(q % d.W);
    float4 o[3];
    raster_pixel_grad<DB>(v, c, r, d.H, d.W, g, gd, o);
    const int *t3 = tri + 3 * (long long)t;
    float *G = grad_pos + d.pos_off;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float *gp = G + 4 * (long long)t3[k];
        atomicAdd(gp + 0, o[k].x);
        atomicAdd(gp + 1, o[k].y);
        if (o[k].z != 0.f) atomicAdd(gp + 2, o[k].z);
        atomicAdd(gp + 3, o[k].w);
    }
}

// vertex gather: grad_pos[img][v] += sum over incident corners of partial[pimg][t][corner]
// (pos_batched: one output per image; else all images sum into one buffer).
// The corners of a block's 256 vertices are one contiguous range of each
// image's partials (CSR order): it is staged in LDS with coalesced loads
// when it fits, and each vertex then sums its slice from there, in corner
// order as the direct loads do.
constexpr int kGatherPosLds = 2048;  // float4 per block (32 KiB)
__global__ void __launch_bounds__(256) k_gather_pos(const float4 *__restrict__ partial, int n_img, int F, int V,
                                                    int pos_batched, const int *__restrict__ inc_off,
                                                    float4 *__restrict__ grad_pos) {
    __shared__ float4 buf[kGatherPosLds];
    const int vb = blockIdx.x * blockDim.x, v = vb + threadIdx.x;
    const int oimg = blockIdx.y;
    const int b0 = inc_off[vb], b1 = inc_off[min(vb + (int)blockDim.x, V)];  // (vb < V: the grid covers V)
    const bool staged = b1 - b0 <= kGatherPosLds;  // uniform per block
    const int i0 = v < V ? inc_off[v] : b0, i1 = v < V ? inc_off[v + 1] : b0;
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    const int ia = pos_batched ? oimg : 0, ib = pos_batched ? oimg + 1 : n_img;
    for (int img = ia; img < ib; ++img) {
        const float4 *P = partial + (long long)img * F * 3;  // this vertex's corners are P[i0 .. i1)
        if (staged) {
            if (img > ia) __syncthreads();  // the previous image's sums have read the buffer
            for (int i = b0 + (int)threadIdx.x; i < b1; i += blockDim.x) buf[i - b0] = P[i];
            __syncthreads();
            for (int i = i0; i < i1; ++i) {
                const float4 q = buf[i - b0];
                s.x += q.x; s.y += q.y; s.z += q.z; s.w += q.w;
            }
            continue;
        }
        // eight corners' loads in flight at once, summed in corner order
        for (int i = i0; i < i1; i += 8) {
            float4 q[8];
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (i + k < i1) q[k] = P[i + k];
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (i + k < i1) { s.x += q[k].x; s.y += q[k].y; s.z += q[k].z; s.w += q[k].w; }
        }
    }
    if (v >= V) return;
    float4 *o = grad_pos + (long long)oimg * V + v;
    float4 cur = *o;
    *o = make_float4(cur.x + s.x, cur.y + s.y, cur.z + s.z, cur.w + s.w);
}

// slot[c] = position of corner c in the incidence list (inverse permutation);
// -1 for corners whose vertex index is outside [0, V) (not in the list)
__global__ void k_corner_slots_clear(long long n, int *__restrict__ slot) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) slot[i] = -1;
}
__global__ void k_corner_slots(const int *__restrict__ inc, const int *__restrict__ inc_off, int V,
                               int *__restrict__ slot) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < inc_off[V]) slot[inc[i]] = (int)i;
}

// ----------------------------------------------------------------------------
// host launchers
// ----------------------------------------------------------------------------
__global__ void k_zero_i32(int *p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// covered pixels (id channel > 0) per image: grid (blocks per image, images),
// 4 pixels per thread and pass, one atomic per wave
__global__ void __launch_bounds__(256) k_count_covered(const float4 *__restrict__ rast, long long npx,
                                                       int *__restrict__ counts) {
    const int img = blockIdx.y;
    const float4 *R = rast + (long long)img * npx;
    int n = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < npx; i += (long long)gridDim.x * blockDim.x)
        n += R[i].w > 0.f ? 1 : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o, 64);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(&counts[img], n);
}

int launch_count_covered(const float *rast, int n_img, long long npx, int *counts, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((n_img + 255) / 256)), dim3(256), 0, st, counts, (long long)n_img);
    long long blocks = (npx + 1023) / 1024;
    if (blocks > 256) blocks = 256;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_count_covered, dim3((unsigned)blocks, (unsigned)n_img), dim3(256), 0, st,
                       reinterpret_cast<const float4 *>(rast), npx, counts);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <int TW, int TH, bool PEEL, bool DB, bool CULL>
static void launch_fine_c(const RasterLaunch &L, hipStream_t st) {
    // Eight or more images of one size: each XCD takes one contiguous eighth of
    // the tiles (whole images, so its L2 holds one image's triangles and setup
    // records at a time).  Otherwise groups of kFineXcdGroup tiles are dealt
    // round robin, so no XCD is left with the heavy images or the heavy band.
    const bool eighths = L.b.descs == nullptr && L.b.n_img >= 8;
    const int group = eighths ? max(1, (L.n_tiles + 7) / 8) : kFineXcdGroup;
    const long long unit = 8ll * group;
    const int pad = (int)((L.n_tiles + unit - 1) / unit * unit);
    hipLaunchKernelGGL((k_fine<TW, TH, PEEL, DB, CULL>), dim3(pad), dim3(TW * TH), 0, st, L.b, L.pos, L.tri,
                       L.bin_count, L.bins, L.list_count, L.lists, L.setup, L.n_tiles, group, L.split_base,
                       L.split_slots, (const float4 *)L.prev, (float4 *)L.rast, (float4 *)L.rast_db, L.boxes);
    if (L.split_slots > 0) {
        // six blocks of 256 lanes per CU, whatever the tile
        const int grid = 6 * 256 * (256 / (TW * TH));
        hipLaunchKernelGGL((k_fine_split<TW, TH, PEEL, DB, CULL>), dim3(grid), dim3(TW * TH), 0, st, L.b, L.pos, L.tri,
                           L.bin_count, L.bins, L.list_count, L.lists, L.setup, L.split_base, L.split_slots,
                           (const float4 *)L.prev, (float4 *)L.rast, (float4 *)L.rast_db, L.boxes);
    }
}

// The culling kernels only where tiles hold many rounds of entries: in
// one-round tiles the cull never runs, yet its registers and code slow the
// fine raster slightly.
template <int TW, int TH, bool PEEL, bool DB>
static void launch_fine_v(const RasterLaunch &L, hipStream_t st) {
    if (L.n_tiles <= 0) return;
    if (L.depth_cull) launch_fine_c<TW, TH, PEEL, DB, true>(L, st);
    else launch_fine_c<TW, TH, PEEL, DB, false>(L, st);
}

template <int TW, int TH>
static void launch_fine_t(const RasterLaunch &L, hipStream_t st) {
    if (L.prev) {
        if (L.rast_db) launch_fine_v<TW, TH, true, true>(L, st);
        else launch_fine_v<TW, TH, true, false>(L, st);
    } else {
        if (L.rast_db) launch_fine_v<TW, TH, false, true>(L, st);
        else launch_fine_v<TW, TH, false, false>(L, st);
    }
}

int launch_rasterize(const RasterLaunch &L, hipStream_t st) {
    // zero bin counters, spill-list counters and the split controls (a kernel, not a memset node)
    long long nz = (long long)L.n_tiles + L.b.n_img + kSplitCtl;
    hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, st, L.bin_count, nz);
    if (L.max_tris > 0) {
        dim3 grid((L.max_tris + 255) / 256, L.b.n_img);
        hipLaunchKernelGGL(k_setup_bin, grid, dim3(256), 0, st, L.b, L.pos, L.tri, L.tile_w, L.tile_h,
                           L.bin_count, L.bins, L.list_count, L.lists, L.setup, L.boxes);
    }
    const bool tile_known = (L.tile_w == 8 && L.tile_h == 8) || (L.tile_w == 16 && L.tile_h == 16) ||
                            (L.tile_w == 32 && L.tile_h == 8) || (L.tile_w == 16 && L.tile_h == 8);
    if (!tile_known) return -1;
    // the antialias preparation (latency-bound) beside the fine raster (issue-bound)
    Fork f{};
    const bool par = L.aa_prepare && fork_begin(st, f);
    int rc = L.aa_prepare ? launch_antialias_prepare(*L.aa_prepare, par ? f.side : st) : 0;
    if (L.tile_w == 8 && L.tile_h == 8) launch_fine_t<8, 8>(L, st);
    else if (L.tile_w == 16 && L.tile_h == 16) launch_fine_t<16, 16>(L, st);
    else if (L.tile_w == 32 && L.tile_h == 8) launch_fine_t<32, 8>(L, st);
    else launch_fine_t<16, 8>(L, st);
    if (par && fork_end(st, f)) rc = -2;
    if (rc) return rc;
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_corner_slots(const int *inc, const int *inc_off, int V, int F, int *slot, hipStream_t st) {
    long long nc = 3ll * F;
    if (nc == 0) return 0;
    hipLaunchKernelGGL(k_corner_slots_clear, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, nc, slot);
    hipLaunchKernelGGL(k_corner_slots, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, inc, inc_off, V, slot);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_gather_pos(const float *partial, int n_img, int F, int V, int pos_batched, const int *inc_off,
                      float *grad_pos, hipStream_t st) {
    if (V == 0) return 0;
    int n_out = pos_batched ? n_img : 1;
    hipLaunchKernelGGL(k_gather_pos, dim3((V + 255) / 256, n_out), dim3(256), 0, st, (const float4 *)partial, n_img, F,
                       V, pos_batched, inc_off, (float4 *)grad_pos);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_rasterize_bwd(const RasterBwdLaunch &L, hipStream_t st) {
    if (L.total_pix == 0) return 0;
    if (L.boxes && L.partial && L.records && L.inc_off && L.inc) {
        TriBwdLaunch T{};
        T.b = L.b;
        T.pos = L.pos;
        T.tri = L.tri;
        T.rast = L.rast;
        T.grad_rast = L.grad_rast;
        T.grad_db = L.grad_db;
        T.boxes = L.boxes;
        T.partial = L.partial;
        T.n_img = L.b.n_img;
        T.F = L.b.F;
        T.H = L.b.H;
        T.W = L.b.W;
        T.views_per_block = 1;
        if (L.um.n) {  // uneven batch: the items' tiles, one partial
            T.um = L.um;
            T.local_ids = L.local_ids;
            T.n_img = 1;
            tile_records_layout_tiles(T, L.records, L.um_tiles, L.b.F);
        } else {
            tile_records_layout(T, L.records, L.b.n_img, L.b.H, L.b.W, (long long)L.b.n_img * L.b.F);
        }
        long long nc = 3ll * L.b.F;
        hipLaunchKernelGGL(k_corner_slots_clear, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, nc,
                           L.corner_slot);
        hipLaunchKernelGGL(k_corner_slots, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, L.inc, L.inc_off,
                           L.b.V, L.corner_slot);
        T.corner_slot = L.corner_slot;
        if (int rc = launch_tri_bwd_pos(T, st)) return rc;
        int n_out = L.b.pos_batched ? L.b.n_img : 1;
        dim3 gg((L.b.V + 255) / 256, n_out);
        hipLaunchKernelGGL(k_gather_pos, gg, dim3(256), 0, st, (const float4 *)L.partial, T.n_img, L.b.F, L.b.V,
                           L.b.pos_batched, L.inc_off, (float4 *)L.grad_pos);
        return hipGetLastError() == hipSuccess ? 0 : -2;
    }
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
