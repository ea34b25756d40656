// Deterministic backward passes of rasterize (positions) and interpolate
// (attributes), pixel-centric in two kernels (gfx950).
//
// k_tile_reduce, one workgroup per 16 x 16 pixel tile of one image: every
// pixel computes its contribution once (coalesced loads, no atomics) and the
// 256 (triangle, pixel) keys are sorted in LDS (bitonic); the contributions
// are stored in sorted order, and each run of equal triangles -- a segment --
// is summed in pixel order, per component, and written as a record
// (triangle, sums).  The records of a tile come out sorted by triangle.
//
// k_tile_gather, one thread per (image [group], triangle) that won a pixel
// in the forward: the triangle's pixel box names the tiles it touches; the
// tile pass stored the record index of the box's first kBoxSlots tiles per
// box, further tiles' records are binary-searched, and the thread adds them in
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

// Quads [q0, q0 + NQP) of every segment's record: one thread per (segment,
// quad), the contributions of its quad (rows of c4, one per quad of the phase)
// summed in pixel order over the contiguous segment.
template <int NQP>
__device__ __forceinline__ void segment_sums(const float4 *c4, const int *seg_start, int nseg, float *rec, int nv,
                                             int q0) {
    for (int w = threadIdx.x; w < nseg * NQP; w += kBTP) {
        const int sg = w / NQP, jq = w - sg * NQP;
        const int p0 = seg_start[sg], p1 = seg_start[sg + 1];
        const float4 *row = c4 + jq * kBTP;
        float4 sum = row[p0];
        for (int p = p0 + 1; p < p1; ++p) {
            const float4 x = row[p];
            sum.x += x.x; sum.y += x.y; sum.z += x.z; sum.w += x.w;
        }
        reinterpret_cast<float4 *>(rec + (long long)sg * nv)[q0 + jq] = sum;
    }
}

// Phase PH of the staged contributions: once the previous phase's sums have
// read the buffer, this pixel's quads [PH QP, min(NQ, (PH + 1) QP)) go to its
// sorted place, then their segment sums.
template <int PH, int QP, int NQ, int NVAL>
__device__ __forceinline__ void reduce_phase(float4 *c4, const float (&val)[NVAL], int my, const int *seg_start,
                                             int nseg, float *rec, int nv) {
    constexpr int j0 = PH * QP, j1 = (PH + 1) * QP < NQ ? (PH + 1) * QP : NQ;
    __syncthreads();
#pragma unroll
    for (int j = j0; j < j1; ++j)
        c4[(j - j0) * kBTP + my] = make_float4(val[4 * j], val[4 * j + 1], val[4 * j + 2], val[4 * j + 3]);
    __syncthreads();
    segment_sums<j1 - j0>(c4, seg_start, nseg, rec, nv, j0);
}


// MODE 0: POS (NV = 12), MODE 1: ATTR (NV = 3 C), MODE 2: both from one pass
// (NV = 3 C + 12: the attribute sums, then the position sums of the grad_rast
// this pass writes -- what MODE 0 would read back -- so both come out of one
// sort and one record search, with the same sums as the two passes)
template <int MODE>
constexpr int tile_nv(int C) { return MODE == 0 ? 12 : (MODE == 1 ? 3 * C : 3 * C + 12); }

// Contribution phases by record width (LDS per block: QP quads x 4 KiB) and
// the waves per SIMD the register budget is cut for (1: no cut): the
// 36-float records (C = 8, both sums) in three phases of three quads with 64
// VGPRs (eight waves; a few registers spill) beat two phases at six waves
// (0.86 against 0.89 ms for the fused backward at 64 x 512^2); 24-float
// records are fastest in two.
template <int MODE, int C>
constexpr int reduce_phases() {
    return tile_nv<MODE>(C) % 4 ? 1 : (tile_nv<MODE>(C) > 32 ? 3 : (tile_nv<MODE>(C) > 16 ? 2 : 1));
}
template <int MODE, int C>
constexpr int reduce_waves() { return tile_nv<MODE>(C) > 32 && tile_nv<MODE>(C) <= 36 ? 8 : 1; }

template <int MODE, int C, bool DB, typename K>
__global__ void __launch_bounds__(256, (reduce_waves<MODE, C>())) k_tile_reduce(TriBwdLaunch L) {
    constexpr int NV = tile_nv<MODE>(C);
    constexpr K kEmptyT = SortKey<K>::kEmptyT;
    __shared__ K key[kBTP];
    __shared__ union {
        K run[kBTP];             // the waves' sorted runs (merge)
        int seg_start[kBTP + 1]; // segment starts (after the merge)
    } U;
    // contributions: four components per float4, quad-major ([quad][pixel]) when NV % 4 == 0,
    // else component rows with a padded stride.  LDS sets this kernel's
    // occupancy, so records of more than four quads go through a buffer of
    // QP quads in NPH phases (quads [0, QP), [QP, 2 QP), ...); the buffer holds
    // the NDC centres of the tile's columns and rows before the sort.
    constexpr bool QUAD = NV % 4 == 0;
    constexpr int NQ = NV / 4;
    constexpr int NPH = reduce_phases<MODE, C>();
    constexpr int QP = (NQ + NPH - 1) / NPH;  // quads per phase (the last may have fewer)
    __shared__ __align__(16) union {
        float contrib[QUAD ? 4 * QP * kBTP : NV * kRowStride];
        double sp[2 * kBT];  // column centres, then row centres (position chain)
    } B;
    __shared__ int wsum[4];
    __shared__ unsigned char rank[kBTP];  // sorted place of each pixel's key
    double *spx = B.sp, *spy = B.sp + kBT;
    float *contrib = B.contrib;
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
            if (MODE == 1 || L.grad_rast_out) reinterpret_cast<float4 *>(L.grad_rast_out)[pix] = gr;
        }
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
        rank[(int)(k & 255)] = (unsigned char)place;
        __syncthreads();
    }
    // this pixel's contributions at its sorted place, so that every segment's
    // contributions are contiguous (read by the sums after the barriers below)
    const int my = rank[tid];
    if (QUAD) {
        float4 *c4 = reinterpret_cast<float4 *>(contrib);
#pragma unroll
        for (int j = 0; j < QP; ++j)
            c4[j * kBTP + my] = make_float4(val[4 * j], val[4 * j + 1], val[4 * j + 2], val[4 * j + 3]);
    } else {
#pragma unroll
        for (int j = 0; j < NV; ++j) contrib[j * kRowStride + my] = val[j];
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
    if (head) {
        const int st = (int)(sk >> 8);
        L.rec_t[rbase + seg] = st;
        if (L.box_slot) {  // the record's place among the first tiles of st's box
            const long long bi = (L.um.n ? 0ll : (long long)img * L.F) + st;
            const ItemBox b = L.boxes[bi];
            const int tx = lt % tiles_x, ty = lt / tiles_x;
            const int tx0 = (b.cc & 0xffff) / kBT, tx1 = ((unsigned)b.cc >> 16) / kBT, ty0 = (b.rr & 0xffff) / kBT;
            const int k = (ty - ty0) * (tx1 - tx0 + 1) + (tx - tx0);
            if (tx >= tx0 && tx <= tx1 && ty >= ty0 && k < kBoxSlots) L.box_slot[kBoxSlots * bi + k] = (int)(rbase + seg);
        }
    }
    if (tid == 0) L.rec_count[tile] = nseg;
    if (QUAD) {
        // one thread per (segment, four components), summed in pixel order over the contiguous segment
        float4 *c4 = reinterpret_cast<float4 *>(contrib);
        float *rec = L.rec_v + rbase * NV;
        segment_sums<QP>(c4, U.seg_start, nseg, rec, NV, 0);
        if constexpr (NPH > 1) reduce_phase<1, QP, NQ>(c4, val, my, U.seg_start, nseg, rec, NV);
        if constexpr (NPH > 2) reduce_phase<2, QP, NQ>(c4, val, my, U.seg_start, nseg, rec, NV);
        static_assert(NPH <= 3, "at most three phases");
    } else {
        // one thread per (segment, component): consecutive threads -> consecutive record floats
        for (int w = tid; w < nseg * NV; w += kBTP) {
            int sg = w / NV, j = w - sg * NV;
            int p0 = U.seg_start[sg], p1 = U.seg_start[sg + 1];
            const float *row = contrib + j * kRowStride;
            float s = row[p0];
            for (int p = p0 + 1; p < p1; ++p) s += row[p];
            L.rec_v[(rbase + sg) * NV + j] = s;
        }
    }
}

// Records of triangle t in one image, in tile order over its pixel box (box
// index bi).  The first kBoxSlots tiles of the box are looked up through
// box_slot, which the tile pass filled with the record index of every
// (box, tile) it wrote; a slot is trusted only when it points into that tile's
// records and names t there (the workspace is not cleared between calls, so a
// slot of a tile where t won no pixel holds whatever an earlier call left).
// Further tiles, and every tile without box_slot, are binary-searched
// (records sorted by triangle), kSearch at a time in lockstep so that their
// dependent loads overlap.
constexpr int kSearch = 4;

// one image's box of the triangle and its box slots, loaded together
struct BoxLoad {
    ItemBox b;
    int4 s;
};
__device__ __forceinline__ BoxLoad load_box(const TriBwdLaunch &L, long long bi) {
    BoxLoad x;
    x.b = L.boxes[bi];
    x.s = L.box_slot ? *reinterpret_cast<const int4 *>(L.box_slot + kBoxSlots * bi) : make_int4(-1, -1, -1, -1);
    return x;
}

template <int NV, typename Add>
__device__ __forceinline__ void image_records(const TriBwdLaunch &L, int t, const BoxLoad &x, int tile_base,
                                              int tiles_x, Add &&add) {
    const ItemBox &b = x.b;
    const int c0 = b.cc & 0xffff, c1 = (unsigned)b.cc >> 16, r0 = b.rr & 0xffff, r1 = (unsigned)b.rr >> 16;
    if (!b.hit || c0 > c1 || r0 > r1) return;
    const int tx0 = c0 / kBT, tx1 = c1 / kBT, ty0 = r0 / kBT, bw = tx1 - tx0 + 1;
    const int nb = bw * (r1 / kBT - ty0 + 1);
    int k0 = 0;
    if (L.box_slot) {
        const int sv[kBoxSlots] = {x.s.x, x.s.y, x.s.z, x.s.w};
        int idx[kBoxSlots], cnt[kBoxSlots], rt[kBoxSlots], seg[kBoxSlots];
        int tx = tx0, ty = ty0;
#pragma unroll
        for (int j = 0; j < kBoxSlots; ++j) {
            idx[j] = -1;
            cnt[j] = 0;
            rt[j] = -1;
            seg[j] = -1;
            if (j < nb) {
                const int tile = tile_base + ty * tiles_x + tx;
                seg[j] = sv[j] - tile * kBTP;
                if (seg[j] >= 0 && seg[j] < kBTP) {  // inside the tile's records: both loads in bounds
                    cnt[j] = L.rec_count[tile];
                    rt[j] = L.rec_t[sv[j]];
                }
                if (++tx > tx1) {
                    tx = tx0;
                    ++ty;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kBoxSlots; ++j)
            if (seg[j] >= 0 && seg[j] < cnt[j] && rt[j] == t) idx[j] = sv[j];
#pragma unroll
        for (int j = 0; j < kBoxSlots; ++j)
            if (idx[j] >= 0) add(L.rec_v + (long long)idx[j] * NV);
        k0 = min(nb, kBoxSlots);
    }
    for (int k = k0; k < nb; k += kSearch) {
        long long rb[kSearch];
        int lo[kSearch], hi[kSearch], hit[kSearch];
#pragma unroll
        for (int j = 0; j < kSearch; ++j) {
            rb[j] = 0;
            lo[j] = 0;
            hi[j] = -1;
            hit[j] = -1;
            if (k + j < nb) {
                const int q = (k + j) / bw;
                const int tile = tile_base + (ty0 + q) * tiles_x + tx0 + (k + j - q * bw);
                rb[j] = (long long)tile * kBTP;
                hi[j] = L.rec_count[tile] - 1;
            }
        }
        for (bool more = true; more;) {
            more = false;
#pragma unroll
            for (int j = 0; j < kSearch; ++j) {
                if (hit[j] >= 0 || lo[j] > hi[j]) continue;
                const int mid = (lo[j] + hi[j]) >> 1;
                const int tm = L.rec_t[rb[j] + mid];
                if (tm < t) lo[j] = mid + 1;
                else if (tm > t) hi[j] = mid - 1;
                else hit[j] = mid;
                more = more || (hit[j] < 0 && lo[j] <= hi[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < kSearch; ++j)
            if (hit[j] >= 0) add(L.rec_v + (rb[j] + hit[j]) * NV);
    }
}

// Records of triangle t over images [v0, v1) in (image, tile) order; done(img)
// after each image.  Uniform batch: box of (image, t) at image * F + t, tile
// image * tiles_per_img + ty * tiles_x + tx.  Uneven batch: t's item only, box
// at t, tiles from the item's first.  The next image's box and slots are
// loaded before this image's records are read.
template <int NV, typename Add, typename Done>
__device__ __forceinline__ void triangle_records(const TriBwdLaunch &L, int t, int v0, int v1, Add &&add,
                                                 Done &&done) {
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
    if (v0 >= v1) return;
    BoxLoad cur = load_box(L, v0 * box_stride + t);
    for (int img = v0; img < v1; ++img) {
        BoxLoad nxt;
        if (img + 1 < v1) nxt = load_box(L, (img + 1) * box_stride + t);
        image_records<NV>(L, t, cur, tile_base + img * tile_stride, tiles_x, add);
        done(img);
        if (img + 1 < v1) cur = nxt;
    }
}

template <int NV>
__device__ __forceinline__ void add_record(float *acc, const float *q) {
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

template <int NV>
__device__ __forceinline__ void gather_records(const TriBwdLaunch &L, int t, int v0, int v1, float *acc) {
    triangle_records<NV>(L, t, v0, v1, [&](const float *q) { add_record<NV>(acc, q); }, [](int) {});
}

// MODE 2: the attribute sums over the group's images (as MODE 1), and per
// image the position sums, written to that image's corner slots (as MODE 0)
template <int NA>
__device__ __forceinline__ void gather_records_fused(const TriBwdLaunch &L, int t, int v0, int v1, float *acc) {
    const int *slot = L.corner_slot + 3 * (long long)t;
    float pos[12];
#pragma unroll
    for (int j = 0; j < 12; ++j) pos[j] = 0.f;
    triangle_records<NA + 12>(
        L, t, v0, v1,
        [&](const float *q) {
            add_record<NA>(acc, q);
            add_record<12>(pos, q + NA);
        },
        [&](int img) {
            float4 *P = reinterpret_cast<float4 *>(L.partial_pos) + (long long)img * 3 * L.F;
            if (slot[0] >= 0) P[slot[0]] = make_float4(pos[0], pos[1], pos[2], pos[3]);  // -1: vertex out of range
            if (slot[1] >= 0) P[slot[1]] = make_float4(pos[4], pos[5], pos[6], pos[7]);
            if (slot[2] >= 0) P[slot[2]] = make_float4(pos[8], pos[9], pos[10], pos[11]);
#pragma unroll
            for (int j = 0; j < 12; ++j) pos[j] = 0.f;
        });
}

// Waves per SIMD the register budget is cut for: 6 keeps the latency-bound
// searches overlapped; records wider than 16 floats need the accumulators of
// 4 waves (128 VGPRs), beyond 40 floats of 3, to stay out of scratch.
template <int MODE, int C>
constexpr int gather_waves() { return tile_nv<MODE>(C) <= 16 ? 6 : (tile_nv<MODE>(C) <= 40 ? 4 : 3); }

template <int MODE, int C>
__global__ void __launch_bounds__(256, (gather_waves<MODE, C>())) k_tile_gather(TriBwdLaunch L) {
    constexpr int NV = tile_nv<MODE>(C);
    // XCD k takes the k-th eighth of the (group, block) grid (its width a
    // multiple of 8): neighbouring triangles share their tiles' record lines
    const int lin = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int grp = lin / gridDim.x;
    const int t = (lin - grp * (int)gridDim.x) * blockDim.x + threadIdx.x;
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

// blocks per group of the tile gather: a multiple of 8 (xcd_remap)
static unsigned gather_width(int F) { return (unsigned)((((F + 255) / 256) + 7) & ~7); }

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
    hipLaunchKernelGGL((k_tile_gather<0, 1>), dim3(gather_width(L.F), L.n_img), dim3(256), 0, st, L);
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
    hipLaunchKernelGGL((k_tile_gather<1, C>), dim3(gather_width(L.F), groups), dim3(256), 0, st, L);
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
    hipLaunchKernelGGL((k_tile_gather<2, C>), dim3(gather_width(L.F), groups), dim3(256), 0, st, L);
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

// workspace: record counts [tiles], triangles [tiles, 256], box slots
// [n_boxes, kBoxSlots], sums [tiles, 256, nv]; the box slots hold int record
// indices, so they are left out (every tile searched) past 2^31 records
static bool box_slots_fit(long long n_tiles) { return n_tiles * kBTP < (1ll << 31); }

static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t tile_records_bytes_tiles(long long n_tiles, int nv, long long n_boxes) {
    const size_t slots = box_slots_fit(n_tiles) ? a256((size_t)n_boxes * kBoxSlots * 4) : 0;
    return a256((size_t)n_tiles * 4) + a256((size_t)n_tiles * kBTP * 4) + slots + a256((size_t)n_tiles * kBTP * nv * 4);
}

size_t tile_records_bytes(int n_img, int H, int W, int nv, long long n_boxes) {
    return tile_records_bytes_tiles((long long)n_img * ((W + kBT - 1) / kBT) * ((H + kBT - 1) / kBT), nv, n_boxes);
}

void tile_records_layout_tiles(TriBwdLaunch &T, void *ws, long long n_tiles, long long n_boxes) {
    T.n_tiles = (int)n_tiles;
    T.n_tiles_pad = (T.n_tiles + 7) & ~7;
    char *b = (char *)ws;
    T.rec_count = (int *)b;
    b += a256((size_t)n_tiles * 4);
    T.rec_t = (int *)b;
    b += a256((size_t)n_tiles * kBTP * 4);
    T.box_slot = nullptr;
    if (box_slots_fit(n_tiles)) {
        T.box_slot = (int *)b;
        b += a256((size_t)n_boxes * kBoxSlots * 4);
    }
    T.rec_v = (float *)b;
}

void tile_records_layout(TriBwdLaunch &T, void *ws, int n_img, int H, int W, long long n_boxes) {
    T.tiles_x = (W + kBT - 1) / kBT;
    T.tiles_per_img = T.tiles_x * ((H + kBT - 1) / kBT);
    tile_records_layout_tiles(T, ws, (long long)n_img * T.tiles_per_img, n_boxes);
}

}  // namespace drp
