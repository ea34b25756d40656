// Deterministic backward passes of rasterize (positions) and interpolate
// (attributes), pixel-centric in two kernels (gfx950).
//
// k_tile_reduce, one workgroup per 16 x 16 pixel tile of one image: every
// pixel computes its contribution once (coalesced loads, no atomics) and the
// 256 (triangle, pixel) keys are sorted in LDS (bitonic); each run of equal
// triangles -- a segment -- is summed in pixel order, per component, and
// written as a record (triangle, sums).  The records of a tile come out
// sorted by triangle.
//
// k_tile_gather, one thread per (image [group], triangle) that won a pixel
// in the forward: the triangle's pixel box names the tiles it touches; the
// thread binary-searches each tile's records for its triangle and adds them in
// tile order.  All sums run in a fixed order, so results are reproducible bit
// for bit.
//
//   POS  (NV = 12): d rast (u, v, z/w [, db]) -> d clip (x, y, z, w) of the 3
//        corners (double-precision chain, drp_grad.h); partial [n, 3F] float4
//        with the corners of each vertex adjacent (incidence order), summed per
//        vertex by k_gather_pos.
//   ATTR (NV = 3 C): (u, v, 1-u-v) * g per corner; the same pass writes
//        grad_rast = (<g, a0-a2>, <g, a1-a2>, 0, 0); k_tile_gather sums the
//        views of a group, partial [groups, F, 3, C], reduced by k_gather_attr.
#include "drp_common.h"
#include "drp_grad.h"
#include "drp_kernels.h"

namespace drp {

constexpr int kBT = 16;            // backward tile edge
constexpr int kBTP = kBT * kBT;    // pixels (and records) per tile
constexpr int kRowStride = kBTP + 1;
// Sort keys: triangle << 8 | pixel, 32-bit while triangle ids fit 24 bits,
// else 64-bit; an uncovered pixel's triangle field is all ones.
template <typename K> struct SortKey;
template <> struct SortKey<unsigned> {
    static constexpr unsigned kEmptyT = (1u << 24) - 1;
};
template <> struct SortKey<unsigned long long> {
    static constexpr unsigned long long kEmptyT = (1ull << 40) - 1;
};


// number of entries of the sorted run a[0..64) below k (reads a[0..63] only)
template <typename K>
__device__ __forceinline__ int count_below64(const K *a, K k) {
    int base = 0;
#pragma unroll
    for (int half = 32; half >= 1; half >>= 1)
        if (a[base + half - 1] < k) base += half;
    return base + (a[base] < k ? 1 : 0);
}

__device__ __forceinline__ int block_excl_scan256(int v, int *wsum, int &total) {
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

// MODE 0: POS (NV = 12), MODE 1: ATTR (NV = 3 C), MODE 2: both from one pass
// (NV = 3 C + 12: the attribute sums, then the position sums of the grad_rast
// this pass writes -- what MODE 0 would read back -- so both come out of one
// sort and one record search, with the same sums as the two passes)
template <int MODE>
constexpr int tile_nv(int C) { return MODE == 0 ? 12 : (MODE == 1 ? 3 * C : 3 * C + 12); }

template <int MODE, int C, bool DB, typename K>
__global__ void __launch_bounds__(256) k_tile_reduce(TriBwdLaunch L) {
    constexpr int NV = tile_nv<MODE>(C);
    constexpr K kEmptyT = SortKey<K>::kEmptyT;
    __shared__ K key[kBTP];
    __shared__ union {
        K run[kBTP];             // the waves' sorted runs (merge)
        int seg_start[kBTP + 1]; // segment starts (after the merge)
    } U;
    // contributions: four components per float4, quad-major ([NV / 4][pixel]) when NV % 4 == 0,
    // else component rows with a padded stride
    constexpr bool QUAD = NV % 4 == 0;
    __shared__ __align__(16) float contrib[QUAD ? NV * kBTP : NV * kRowStride];
    __shared__ int wsum[4];
    __shared__ double spx[kBT], spy[kBT];  // NDC centres of the tile's columns and rows (position chain)
    const int tid = threadIdx.x;
    const int tile = xcd_remap(blockIdx.x, L.n_tiles_pad);
    if (tile >= L.n_tiles) return;  // uniform per block
    // the tile's image: arithmetic for a uniform batch, the item map for an uneven one
    int img, lt, H = L.H, W = L.W, tiles_x = L.tiles_x;
    long long pix0;
    if (L.um.n) {
        img = um_find(L.um.bt16, L.um.n, tile);
        lt = tile - L.um.bt16[img];
        H = L.um.hw[2 * img];
        W = L.um.hw[2 * img + 1];
        tiles_x = (W + kBT - 1) / kBT;
        pix0 = L.um.pix[img];
    } else {
        img = tile / L.tiles_per_img;
        lt = tile - img * L.tiles_per_img;
        pix0 = (long long)img * L.H * L.W;
    }
    const int r = (lt / tiles_x) * kBT + (tid >> 4), c = (lt % tiles_x) * kBT + (tid & 15);
    if (MODE != 1) {  // one double division per column and row of the tile, not two per pixel
        if (tid < kBT) spx[tid] = pix_ndc_d((lt % tiles_x) * kBT + tid, W);
        else if (tid < 2 * kBT) spy[tid - kBT] = pix_ndc_d((lt / tiles_x) * kBT + tid - kBT, H);
        __syncthreads();
    }
    float val[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) val[j] = 0.f;
    int t = -1;
    if (r < H && c < W) {
        const long long pix = pix0 + (long long)r * W + c;
        float4 rv = reinterpret_cast<const float4 *>(L.rast)[pix];
        if (rv.w > 0.f) {
            int id = float_to_triidx(rv.w);
            if (L.um.n && L.local_ids) id += L.um.tri[img];  // item-local id -> batch triangle + 1
            if (id >= 1 && id <= L.F) t = id - 1;
        }
        if (MODE == 0) {
            if (t >= 0) {
                ImageDesc d = image_desc(L.b, img);
                TriVerts tv;
                if (load_tri(L.pos, L.tri, t, d, tv)) {
                    float4 g = reinterpret_cast<const float4 *>(L.grad_rast)[pix];
                    float4 gd = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (DB) gd = db_grad_scaled(reinterpret_cast<const float4 *>(L.grad_db)[pix], H, W);
                    float4 o[3];
                    raster_pixel_grad_at<DB>(tv, spx[tid & 15], spy[tid >> 4], g, gd, o);
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        val[4 * q] = o[q].x; val[4 * q + 1] = o[q].y; val[4 * q + 2] = o[q].z; val[4 * q + 3] = o[q].w;
                    }
                } else {
                    t = -1;
                }
            }
        } else {
            float4 gr = make_float4(0.f, 0.f, 0.f, 0.f);
            if (t >= 0) {
                const int *t3 = L.tri + 3 * (long long)t;
                int i0 = t3[0], i1 = t3[1], i2 = t3[2];
                if ((unsigned)i0 < (unsigned)L.V && (unsigned)i1 < (unsigned)L.V && (unsigned)i2 < (unsigned)L.V) {
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
                    const float *A = L.attr + (L.attr_batched ? (long long)img * L.V * C : 0);
                    const float *a0 = A + (long long)i0 * C, *a1 = A + (long long)i1 * C, *a2 = A + (long long)i2 * C;
                    float u = rv.x, v = rv.y, w2 = (1.f - u) - v, su = 0.f, sv = 0.f;
                    float r0[C], r1[C], r2[C];  // the three attribute rows (float4 loads when C % 4 == 0)
                    if (C % 4 == 0) {
#pragma unroll
                        for (int q = 0; q < C / 4; ++q) {
                            const float4 x0 = reinterpret_cast<const float4 *>(a0)[q];
                            const float4 x1 = reinterpret_cast<const float4 *>(a1)[q];
                            const float4 x2 = reinterpret_cast<const float4 *>(a2)[q];
                            r0[4 * q] = x0.x; r0[4 * q + 1] = x0.y; r0[4 * q + 2] = x0.z; r0[4 * q + 3] = x0.w;
                            r1[4 * q] = x1.x; r1[4 * q + 1] = x1.y; r1[4 * q + 2] = x1.z; r1[4 * q + 3] = x1.w;
                            r2[4 * q] = x2.x; r2[4 * q + 1] = x2.y; r2[4 * q + 2] = x2.z; r2[4 * q + 3] = x2.w;
                        }
                    } else {
#pragma unroll
                        for (int ch = 0; ch < C; ++ch) { r0[ch] = a0[ch]; r1[ch] = a1[ch]; r2[ch] = a2[ch]; }
                    }
#pragma unroll
                    for (int ch = 0; ch < C; ++ch) {
                        float gc = gv[ch], x2 = r2[ch];
                        su += gc * (r0[ch] - x2);
                        sv += gc * (r1[ch] - x2);
                        val[ch] = u * gc;
                        val[C + ch] = v * gc;
                        val[2 * C + ch] = w2 * gc;
                    }
                    gr = make_float4(su, sv, 0.f, 0.f);
                    if (MODE == 2) {  // the position chain of this grad_rast
                        ImageDesc d = image_desc(L.b, img);
                        TriVerts tv;
                        if (load_tri(L.pos, L.tri, t, d, tv)) {
                            float4 o[3];
                            raster_pixel_grad_at<false>(tv, spx[tid & 15], spy[tid >> 4], gr,
                                                        make_float4(0.f, 0.f, 0.f, 0.f), o);
#pragma unroll
                            for (int q = 0; q < 3; ++q) {
                                val[3 * C + 4 * q] = o[q].x; val[3 * C + 4 * q + 1] = o[q].y;
                                val[3 * C + 4 * q + 2] = o[q].z; val[3 * C + 4 * q + 3] = o[q].w;
                            }
                        }
                    }
                } else {
                    t = -1;
                }
            }
            reinterpret_cast<float4 *>(L.grad_rast_out)[pix] = gr;
        }
    }
    if (QUAD) {
        float4 *c4 = reinterpret_cast<float4 *>(contrib);
#pragma unroll
        for (int j = 0; j < NV; j += 4)
            c4[(j / 4) * kBTP + tid] = make_float4(val[j], val[j + 1], val[j + 2], val[j + 3]);
    } else {
#pragma unroll
        for (int j = 0; j < NV; ++j) contrib[j * kRowStride + tid] = val[j];
    }
    // sort the 256 keys (triangle << 8 | pixel; uncovered pixels last, keys unique):
    // each wave sorts its 64 across lanes, then every key's place is its lane
    // plus the keys below it in the other three runs
    {
        K k = wave_sort64<K>(((t >= 0 ? (K)t : kEmptyT) << 8) | (K)tid);
        const int wave = tid >> 6;
        U.run[tid] = k;
        __syncthreads();
        int place = tid & 63;
#pragma unroll
        for (int w = 0; w < 4; ++w)
            if (w != wave) place += count_below64<K>(U.run + 64 * w, k);
        key[place] = k;
        __syncthreads();
    }
    // segments: runs of one triangle in sorted order
    K sk = key[tid];
    bool valid = (sk >> 8) != kEmptyT;
    bool head = valid && (tid == 0 || (key[tid - 1] >> 8) != (sk >> 8));
    int nseg;
    int seg = block_excl_scan256(head ? 1 : 0, wsum, nseg);
    int nvalid = __syncthreads_count(valid);
    if (head) U.seg_start[seg] = tid;
    if (tid == 0) U.seg_start[nseg] = nvalid;
    __syncthreads();
    const long long rbase = (long long)tile * kBTP;
    if (head) L.rec_t[rbase + seg] = (int)(sk >> 8);
    if (tid == 0) L.rec_count[tile] = nseg;
    if (QUAD) {
        // one thread per (segment, four components): each key read serves four sums (pixel order)
        constexpr int NQ = NV / 4;
        const float4 *c4 = reinterpret_cast<const float4 *>(contrib);
        for (int w = tid; w < nseg * NQ; w += kBTP) {
            const int sg = w / NQ, jq = w - sg * NQ;
            const int p0 = U.seg_start[sg], p1 = U.seg_start[sg + 1];
            const float4 *row = c4 + jq * kBTP;
            float4 sum = row[(int)(key[p0] & 255)];
            for (int p = p0 + 1; p < p1; ++p) {
                const float4 x = row[(int)(key[p] & 255)];
                sum.x += x.x; sum.y += x.y; sum.z += x.z; sum.w += x.w;
            }
            reinterpret_cast<float4 *>(L.rec_v + (rbase + sg) * NV)[jq] = sum;
        }
    } else {
        // one thread per (segment, component): consecutive threads -> consecutive record floats
        for (int w = tid; w < nseg * NV; w += kBTP) {
            int sg = w / NV, j = w - sg * NV;
            int p0 = U.seg_start[sg], p1 = U.seg_start[sg + 1];
            const float *row = contrib + j * kRowStride;
            float s = row[(int)(key[p0] & 255)];
            for (int p = p0 + 1; p < p1; ++p) s += row[(int)(key[p] & 255)];
            L.rec_v[(rbase + sg) * NV + j] = s;
        }
    }
}

// Add the records of triangle t over images [v0, v1): for each image where t
// won a pixel, the tiles of its pixel box, in (image, tile) order.  The
// searches (records sorted by triangle) run kSearch at a time in lockstep, so
// their dependent loads overlap instead of queueing one after another.
constexpr int kSearch = 4;

template <int NV>
__device__ __forceinline__ void gather_records(const TriBwdLaunch &L, int t, int v0, int v1, float *acc) {
    // uniform batch: images v0 .. v1, box of (image, t) at image * F + t, tile
    // image * tiles_per_img + ty * tiles_x + tx; uneven batch: t's item only,
    // box at t, tiles from the item's first
    long long box_stride = L.F;
    int tile_stride = L.tiles_per_img, tiles_x = L.tiles_x, tile_base = 0;
    if (L.um.n) {
        const int item = um_find(L.um.tri, L.um.n, t);
        v0 = 0;
        v1 = 1;
        box_stride = 0;
        tile_stride = 0;
        tile_base = L.um.bt16[item];
        tiles_x = (L.um.hw[2 * item + 1] + kBT - 1) / kBT;
    }
    int img = v0 - 1, tx = 0, tx0 = 0, tx1 = -1, ty = 1, ty1 = 0;  // empty range: the first fill reads a box
    for (;;) {
        long long rb[kSearch];
        int lo[kSearch], hi[kSearch], hit[kSearch];
        int n = 0;
        bool done = false;
#pragma unroll
        for (int k = 0; k < kSearch; ++k) {
            rb[k] = 0;
            lo[k] = 0;
            hi[k] = -1;
            hit[k] = -1;
            if (done) continue;
            if (ty > ty1) {  // next image where t won a pixel
                done = true;
                while (++img < v1) {
                    ItemBox b = L.boxes[(long long)img * box_stride + t];
                    int c0 = b.cc & 0xffff, c1 = (unsigned)b.cc >> 16, r0 = b.rr & 0xffff, r1 = (unsigned)b.rr >> 16;
                    if (!b.hit || c0 > c1 || r0 > r1) continue;
                    tx0 = tx = c0 / kBT;
                    tx1 = c1 / kBT;
                    ty = r0 / kBT;
                    ty1 = r1 / kBT;
                    done = false;
                    break;
                }
                if (done) continue;
            }
            const int tile = tile_base + img * tile_stride + ty * tiles_x + tx;
            rb[k] = (long long)tile * kBTP;
            hi[k] = tile;  // replaced by the record count below
            ++n;
            if (++tx > tx1) {
                tx = tx0;
                ++ty;
            }
        }
#pragma unroll
        for (int k = 0; k < kSearch; ++k)
            if (k < n) hi[k] = L.rec_count[hi[k]] - 1;
        for (bool more = true; more;) {
            more = false;
#pragma unroll
            for (int k = 0; k < kSearch; ++k) {
                if (hit[k] >= 0 || lo[k] > hi[k]) continue;
                const int mid = (lo[k] + hi[k]) >> 1;
                const int tm = L.rec_t[rb[k] + mid];
                if (tm < t) lo[k] = mid + 1;
                else if (tm > t) hi[k] = mid - 1;
                else hit[k] = mid;
                more = more || (hit[k] < 0 && lo[k] <= hi[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < kSearch; ++k) {
            if (hit[k] < 0) continue;
            const float *q = L.rec_v + (rb[k] + hit[k]) * NV;
            if (NV % 4 == 0) {  // records are 16-byte aligned: NV * 4 bytes each from an aligned base
#pragma unroll
                for (int j = 0; j < NV; j += 4) {
                    const float4 x = *reinterpret_cast<const float4 *>(q + j);
                    acc[j] += x.x; acc[j + 1] += x.y; acc[j + 2] += x.z; acc[j + 3] += x.w;
                }
            } else {
#pragma unroll
                for (int j = 0; j < NV; ++j) acc[j] += q[j];
            }
        }
        if (n < kSearch) return;
    }
}

// MODE 2: the attribute sums over the group's images (as MODE 1), and per
// image the position sums written to that image's corner slots (as MODE 0);
// records come in (image, tile) order, so an image's position sums are complete
// when the next image's first record (or the end) arrives.
template <int NA>
__device__ __forceinline__ void gather_records_fused(const TriBwdLaunch &L, int t, int v0, int v1, float *acc) {
    constexpr int NV = NA + 12;
    const int *slot = L.corner_slot + 3 * (long long)t;
    auto flush = [&](int img, const float *q12) {
        float4 *P = reinterpret_cast<float4 *>(L.partial_pos) + (long long)img * 3 * L.F;
        if (slot[0] >= 0) P[slot[0]] = make_float4(q12[0], q12[1], q12[2], q12[3]);  // -1: vertex out of range
        if (slot[1] >= 0) P[slot[1]] = make_float4(q12[4], q12[5], q12[6], q12[7]);
        if (slot[2] >= 0) P[slot[2]] = make_float4(q12[8], q12[9], q12[10], q12[11]);
    };
    const float zero12[12] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float pos[12];
    int cur = v0 - 1;  // image whose position sums are in `pos`
    int img = v0 - 1, tx = 0, tx0 = 0, tx1 = -1, ty = 1, ty1 = 0;
    for (;;) {
        long long rb[kSearch];
        int lo[kSearch], hi[kSearch], hit[kSearch], im[kSearch];
        int n = 0;
        bool done = false;
#pragma unroll
        for (int k = 0; k < kSearch; ++k) {
            rb[k] = 0;
            lo[k] = 0;
            hi[k] = -1;
            hit[k] = -1;
            im[k] = -1;
            if (done) continue;
            if (ty > ty1) {
                done = true;
                while (++img < v1) {
                    ItemBox b = L.boxes[(long long)img * L.F + t];
                    int c0 = b.cc & 0xffff, c1 = (unsigned)b.cc >> 16, r0 = b.rr & 0xffff, r1 = (unsigned)b.rr >> 16;
                    if (!b.hit || c0 > c1 || r0 > r1) continue;
                    tx0 = tx = c0 / kBT;
                    tx1 = c1 / kBT;
                    ty = r0 / kBT;
                    ty1 = r1 / kBT;
                    done = false;
                    break;
                }
                if (done) continue;
            }
            const int tile = img * L.tiles_per_img + ty * L.tiles_x + tx;
            rb[k] = (long long)tile * kBTP;
            hi[k] = tile;
            im[k] = img;
            ++n;
            if (++tx > tx1) {
                tx = tx0;
                ++ty;
            }
        }
#pragma unroll
        for (int k = 0; k < kSearch; ++k)
            if (k < n) hi[k] = L.rec_count[hi[k]] - 1;
        for (bool more = true; more;) {
            more = false;
#pragma unroll
            for (int k = 0; k < kSearch; ++k) {
                if (hit[k] >= 0 || lo[k] > hi[k]) continue;
                const int mid = (lo[k] + hi[k]) >> 1;
                const int tm = L.rec_t[rb[k] + mid];
                if (tm < t) lo[k] = mid + 1;
                else if (tm > t) hi[k] = mid - 1;
                else hit[k] = mid;
                more = more || (hit[k] < 0 && lo[k] <= hi[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < kSearch; ++k) {
            if (k >= n) continue;
            if (im[k] != cur) {  // a new image: the previous one's sums are complete
                if (cur >= v0) flush(cur, pos);
                for (int e = cur + 1; e < im[k]; ++e) flush(e, zero12);
                cur = im[k];
#pragma unroll
                for (int j = 0; j < 12; ++j) pos[j] = 0.f;
            }
            if (hit[k] < 0) continue;
            const float *q = L.rec_v + (rb[k] + hit[k]) * NV;
#pragma unroll
            for (int j = 0; j < NV; j += 4) {
                const float4 x = *reinterpret_cast<const float4 *>(q + j);
                if (j < NA) {
                    acc[j] += x.x; acc[j + 1] += x.y; acc[j + 2] += x.z; acc[j + 3] += x.w;
                } else {
                    pos[j - NA] += x.x; pos[j - NA + 1] += x.y; pos[j - NA + 2] += x.z; pos[j - NA + 3] += x.w;
                }
            }
        }
        if (n < kSearch) break;
    }
    if (cur >= v0) flush(cur, pos);
    for (int e = cur + 1; e < v1; ++e) flush(e, zero12);
}

template <int MODE, int C>
__global__ void __launch_bounds__(256, 6) k_tile_gather(TriBwdLaunch L) {  // 6 waves per SIMD: latency-bound searches
    constexpr int NV = tile_nv<MODE>(C);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int grp = blockIdx.y;
    if (t >= L.F) return;
    float acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = 0.f;
    const int VB = L.views_per_block;
    const int v0 = grp * VB, v1 = min(v0 + VB, L.n_img);
    if (MODE == 2) gather_records_fused<3 * C>(L, t, v0, v1, acc);
    else gather_records<NV>(L, t, v0, v1, acc);
    if (MODE == 0) {
        // corners go to their vertex's slots (CSR order), so the vertex gather reads contiguously
        float4 *P = reinterpret_cast<float4 *>(L.partial) + (long long)grp * 3 * L.F;
        const int *slot = L.corner_slot + 3 * (long long)t;
        if (slot[0] >= 0) P[slot[0]] = make_float4(acc[0], acc[1], acc[2], acc[3]);  // -1: vertex out of range
        if (slot[1] >= 0) P[slot[1]] = make_float4(acc[4], acc[5], acc[6], acc[7]);
        if (slot[2] >= 0) P[slot[2]] = make_float4(acc[8], acc[9], acc[10], acc[11]);
    } else {
        constexpr int NA = 3 * C;  // attribute part (MODE 1: all of the record)
        float *P = L.partial + ((long long)grp * L.F + t) * NA;
        if (NA % 4 == 0) {
#pragma unroll
            for (int j = 0; j < NA; j += 4)
                *reinterpret_cast<float4 *>(P + j) = make_float4(acc[j], acc[j + 1], acc[j + 2], acc[j + 3]);
        } else {
#pragma unroll
            for (int j = 0; j < NA; ++j) P[j] = acc[j];
        }
    }
}

static void tile_grid(const TriBwdLaunch &L, dim3 &grid) {
    grid = dim3((unsigned)L.n_tiles_pad);
}

int launch_tri_bwd_pos(const TriBwdLaunch &L, hipStream_t st) {
    dim3 grid;
    tile_grid(L, grid);
    const bool k32 = L.F < (int)SortKey<unsigned>::kEmptyT;
    if (L.grad_db) {
        if (k32) hipLaunchKernelGGL((k_tile_reduce<0, 1, true, unsigned>), grid, dim3(kBTP), 0, st, L);
        else hipLaunchKernelGGL((k_tile_reduce<0, 1, true, unsigned long long>), grid, dim3(kBTP), 0, st, L);
    } else {
        if (k32) hipLaunchKernelGGL((k_tile_reduce<0, 1, false, unsigned>), grid, dim3(kBTP), 0, st, L);
        else hipLaunchKernelGGL((k_tile_reduce<0, 1, false, unsigned long long>), grid, dim3(kBTP), 0, st, L);
    }
    hipLaunchKernelGGL((k_tile_gather<0, 1>), dim3((L.F + 255) / 256, L.n_img), dim3(256), 0, st, L);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <int C>
static void launch_attr_c(const TriBwdLaunch &L, hipStream_t st) {
    dim3 grid;
    tile_grid(L, grid);
    if (L.F < (int)SortKey<unsigned>::kEmptyT)
        hipLaunchKernelGGL((k_tile_reduce<1, C, false, unsigned>), grid, dim3(kBTP), 0, st, L);
    else
        hipLaunchKernelGGL((k_tile_reduce<1, C, false, unsigned long long>), grid, dim3(kBTP), 0, st, L);
    int groups = (L.n_img + L.views_per_block - 1) / L.views_per_block;
    hipLaunchKernelGGL((k_tile_gather<1, C>), dim3((L.F + 255) / 256, groups), dim3(256), 0, st, L);
}

template <int C>
static void launch_fused_c(const TriBwdLaunch &L, hipStream_t st) {
    dim3 grid;
    tile_grid(L, grid);
    if (L.F < (int)SortKey<unsigned>::kEmptyT)
        hipLaunchKernelGGL((k_tile_reduce<2, C, false, unsigned>), grid, dim3(kBTP), 0, st, L);
    else
        hipLaunchKernelGGL((k_tile_reduce<2, C, false, unsigned long long>), grid, dim3(kBTP), 0, st, L);
    int groups = (L.n_img + L.views_per_block - 1) / L.views_per_block;
    hipLaunchKernelGGL((k_tile_gather<2, C>), dim3((L.F + 255) / 256, groups), dim3(256), 0, st, L);
}

int launch_tri_bwd_fused(const TriBwdLaunch &L, int C, hipStream_t st) {
    switch (C) {
        case 4: launch_fused_c<4>(L, st); break;
        case 8: launch_fused_c<8>(L, st); break;
        default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_tri_bwd_attr(const TriBwdLaunch &L, int C, hipStream_t st) {
    switch (C) {
        case 1: launch_attr_c<1>(L, st); break;
        case 2: launch_attr_c<2>(L, st); break;
        case 3: launch_attr_c<3>(L, st); break;
        case 4: launch_attr_c<4>(L, st); break;
        case 8: launch_attr_c<8>(L, st); break;
        case 16: launch_attr_c<16>(L, st); break;
        default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

size_t tile_records_bytes_tiles(long long tiles, int nv) {
    auto a256 = [](size_t x) { return (x + 255) & ~(size_t)255; };
    return a256((size_t)tiles * 4) + a256((size_t)tiles * kBTP * 4) + a256((size_t)tiles * kBTP * nv * 4);
}

size_t tile_records_bytes(int n_img, int H, int W, int nv) {
    return tile_records_bytes_tiles((long long)n_img * ((W + kBT - 1) / kBT) * ((H + kBT - 1) / kBT), nv);
}

void tile_records_layout_tiles(TriBwdLaunch &T, void *ws, long long n_tiles) {
    T.n_tiles = (int)n_tiles;
    T.n_tiles_pad = (T.n_tiles + 7) & ~7;
    auto a256 = [](size_t x) { return (x + 255) & ~(size_t)255; };
    char *b = (char *)ws;
    T.rec_count = (int *)b;
    b += a256((size_t)T.n_tiles * 4);
    T.rec_t = (int *)b;
    b += a256((size_t)T.n_tiles * kBTP * 4);
    T.rec_v = (float *)b;
}

void tile_records_layout(TriBwdLaunch &T, void *ws, int n_img, int H, int W, int nv) {
    T.tiles_x = (W + kBT - 1) / kBT;
    T.tiles_per_img = T.tiles_x * ((H + kBT - 1) / kBT);
    T.n_tiles = n_img * T.tiles_per_img;
    T.n_tiles_pad = (T.n_tiles + 7) & ~7;
    auto a256 = [](size_t x) { return (x + 255) & ~(size_t)255; };
    char *b = (char *)ws;
    T.rec_count = (int *)b;
    b += a256((size_t)T.n_tiles * 4);
    T.rec_t = (int *)b;
    b += a256((size_t)T.n_tiles * kBTP * 4);
    T.rec_v = (float *)b;
    (void)nv;
}

}  // namespace drp
