// Launch descriptors shared by the kernel files and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include "drp_common.h"

namespace drp {

// Uneven batches (items of different resolutions and meshes, packed back to
// back): prefix sums over the items of their pixels, vertices, triangles and of
// the tiles each pass uses.  The item a work unit (block, tile, pixel) belongs
// to is the last item whose prefix is <= the unit's index (um_find).  n = 0:
// uniform batch, every image [H, W] of the launch.
struct UnevenMap {
    const long long *pix;   // [n + 1] pixel offsets
    const int *vert;        // [n + 1] vertex offsets
    const int *tri;         // [n + 1] triangle offsets
    const int *bt16;        // [n + 1] 16 x 16 tiles (deterministic backward)
    const int *blk;         // [n + 1] 64 x 8 blocks (antialias pairs)
    const int *t32;         // [n + 1] 32 x 32 tiles (antialias backward)
    const int *hw;          // [2 n] (H, W) per item
    int n;
};

template <typename T>
__device__ __forceinline__ int um_find(const T *pre, int n, T x) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Pixel box of one (image, triangle) from the forward setup: columns c0..c1 in
// cc (c0 | c1 << 16), rows r0..r1 in rr; c0 > c1 when empty.  hit is set by the
// forward when the triangle won at least one pixel of that image.
struct ItemBox {
    int cc, rr, hit;
};

// heavy-tile split of the fine raster: control words after the bin and list
// counters, at most this many split tiles per call
constexpr int kSplitCtl = 4;
constexpr int kSplitSlotsMax = 4096;
inline int split_slots(long long n_tiles) { return (int)(n_tiles < kSplitSlotsMax ? n_tiles : kSplitSlotsMax); }
// per slot: int4 of bookkeeping, a done counter, 63 queued chunks, 256 depth keys
inline size_t split_region_bytes(int slots) { return (size_t)slots * (16 + 4 + 4 * 63 + 8 * 256); }

struct AALaunch;

struct RasterLaunch {
    Batch b;
    const float *pos;
    const int *tri;
    int tile_w, tile_h;
    int *bin_count;   // [n_tiles], immediately followed by list_count [n_img] and split_ctl [kSplitCtl]
    int *list_count;
    char *split_base; // heavy-tile split region (split_region_bytes(split_slots))
    int split_slots;
    int *bins;        // [n_tiles * bin_cap]
    int *lists;       // spill lists
    int4 *setup;      // setup records: two int4 per (image, triangle), at the image's list offset
    ItemBox *boxes;   // optional [n_img, F] pixel box + hit flag per (image, triangle)
    int n_tiles;
    int max_tris;     // largest triangle range of any image
    const float *prev;
    float *rast;
    float *rast_db;
    const AALaunch *aa_prepare;  // optional: the antialias preparation, run beside the fine raster
    int depth_cull;   // fine raster with the depth cull (depth_cull_wanted)
};

// The fine raster culls triangles by depth where bins hold more than
// kCullRounds rounds of a tile's lanes (a bin capacity learnt or tuned from
// the exact counts is the largest count + 15 %, rounded up to a power of two).
constexpr int kCullRounds = 4;
inline int depth_cull_wanted(int bin_cap, int tile_w, int tile_h) {
    return bin_cap > kCullRounds * tile_w * tile_h ? 1 : 0;
}

struct RasterBwdLaunch {
    Batch b;
    const float *pos;
    const int *tri;
    const float *rast, *grad_rast, *grad_db;
    float *grad_pos;
    long long total_pix;
    int pix_per_img;
    // triangle-centric path (all set) or per-pixel atomics (any null)
    const ItemBox *boxes;
    float *partial;          // [n_img, 3F] float4, corners in vertex (CSR) order
    void *records;           // tile records (tile_records_bytes(n_img, H, W, 12, n_img * F))
    int *corner_slot;        // [3F]: position of corner 3t+k in the incidence list
    const int *inc_off, *inc;
    UnevenMap um;            // uneven batch (um.n > 0): packed items, global vertices
    long long um_tiles;      // uneven batch: 16 x 16 tiles of all items
    int local_ids;           // uneven batch: rast ids count from each item's first triangle (else global)
};

struct InterpLaunch {
    const float *attr;
    int V, C, attr_batched;
    int cv;           // compiled channel variant (0 = generic loop)
    const float *rast;
    const int *tri;
    int F;
    long long total_pix, pix_per_img;
    const float *rast_db;
    const int *diff_idx;
    int n_diff;
    float *out, *out_da;
    const float *grad_out, *grad_da;
    float *grad_attr, *grad_rast, *grad_db;
    // triangle-centric backward (all set) or per-pixel atomics (any null)
    int H, W;
    const ItemBox *boxes;
    float *partial;          // [groups, F, 3, C]
    void *records;           // tile records (tile_records_bytes(n_img, H, W, 3 C, n_img * F))
    const int *inc_off, *inc;
    int views_per_group;
    UnevenMap um;            // uneven batch (um.n > 0): packed items, global ids and vertices, one group
    long long um_tiles;      // uneven batch: 16 x 16 tiles of all items
};

struct TopoLaunch {
    const int *tri;
    int F;
    unsigned long long *keys;
    int *cnt;
    int2 *slots;
    unsigned mask;
    int *opp;
};

struct AALaunch {
    const float *color, *rast, *pos;
    const int *tri, *opp;
    int n_img, H, W, C, V, F, pos_batched;
    int cv;           // compiled channel variant (0 = generic loop)
    double2 *proj;               // [n_proj, V] pixel coordinates (NaN: w <= 0)
    signed char *orient;
    unsigned char *silh;
    float2 *weights;
    int2 *decisions;             // [n, H, W] pair decisions for the backward (nullptr: not kept)
    int reuse;                   // backward: workspace holds the forward's state
    int prepared;                // forward: drp_antialias_prepare filled the per-vertex / per-triangle state
    float *out;
    const float *grad_out;
    float *grad_color, *grad_pos;
    float boost;
    UnevenMap um;                // uneven batch (um.n > 0): n_img items, packed; pos / tri global
    long long total_pix;         // uneven batch: pixels of all items
    int total_blk, total_t32;    // uneven batch: pair blocks and backward tiles of all items
};

// tiles of a triangle's pixel box whose record index the tile pass stores
constexpr int kBoxSlots = 4;

struct TriBwdLaunch {
    Batch b;
    const float *pos;
    const int *tri;
    const float *rast, *grad_rast, *grad_db, *grad_out;
    const float *attr;       // attribute mode: attributes (for grad_rast)
    int V, attr_batched;
    float *grad_rast_out;    // attribute mode: grad_rast written by the tile pass (fused mode: null skips it)
    const ItemBox *boxes;
    float *partial;
    float *partial_pos;      // fused mode: the position partial [n_img, 3F] float4 (partial: attributes)
    const int *corner_slot;  // positions mode: partial index of corner 3t+k (vertex order)
    int n_img, F, H, W;
    int views_per_block;     // images summed per partial (attribute mode); 1 for positions
    // per-tile records (16 x 16 tiles): count, triangle, sums
    int *rec_count, *rec_t;
    float *rec_v;
    int *box_slot;           // [boxes, kBoxSlots] record index of the box's first tiles (nullptr: search)
    int tiles_x, tiles_per_img, n_tiles, n_tiles_pad;
    UnevenMap um;            // uneven batch (um.n > 0): tiles of all items, one group
    int local_ids;           // uneven batch: rast ids count from each item's first triangle (else global)
};

// Fork / join onto the device's auxiliary stream (drp_fork.hip).  fork_begin
// returns false (f.side null) when there is none: run everything on `main`.
struct Fork {
    hipStream_t side;
    hipEvent_t join;
};
bool fork_begin(hipStream_t main, Fork &f);
int fork_end(hipStream_t main, const Fork &f);

int launch_tri_bwd_pos(const TriBwdLaunch &L, hipStream_t st);
int launch_tri_bwd_attr(const TriBwdLaunch &L, int C, hipStream_t st);
int launch_tri_bwd_fused(const TriBwdLaunch &L, int C, hipStream_t st);
int launch_corner_slots(const int *inc, const int *inc_off, int V, int F, int *slot, hipStream_t st);
int launch_gather_pos(const float *partial, int n_img, int F, int V, int pos_batched, const int *inc_off,
                      float *grad_pos, hipStream_t st);
int launch_gather_attr(const float *partial, int groups, int F, int V, int C, const int *inc_off, const int *inc,
                       float *grad_attr, hipStream_t st);
// workspace of the tile records for nv floats per record and n_boxes pixel
// boxes (n_img * F for a uniform batch, the triangles of an uneven one), and
// its layout
size_t tile_records_bytes(int n_img, int H, int W, int nv, long long n_boxes);
void tile_records_layout(TriBwdLaunch &T, void *ws, int n_img, int H, int W, long long n_boxes);
// the same for an uneven batch of n_tiles 16 x 16 tiles in all
size_t tile_records_bytes_tiles(long long n_tiles, int nv, long long n_boxes);
void tile_records_layout_tiles(TriBwdLaunch &T, void *ws, long long n_tiles, long long n_boxes);

struct EdgeLaunch {
    const float *grad_out, *rgb, *rast, *frag_xy, *tris_xy;
    float *grad_frag_xy;
    int H, W, C, F;
};

int launch_edge_grad_bwd(const EdgeLaunch &L, hipStream_t st);
int launch_rasterize(const RasterLaunch &L, hipStream_t st);
int launch_rasterize_bwd(const RasterBwdLaunch &L, hipStream_t st);
int launch_interpolate(const InterpLaunch &L, hipStream_t st);
int launch_interpolate_bwd(const InterpLaunch &L, hipStream_t st);
int launch_topology(const TopoLaunch &L, hipStream_t st);
int launch_antialias(const AALaunch &L, hipStream_t st);
int launch_antialias_bwd(const AALaunch &L, hipStream_t st);
int launch_antialias_prepare(const AALaunch &L, hipStream_t st);
size_t incidence_ws_bytes(int V, int F);
size_t stencil_partition_ws_bytes(long long npx, int n_owner);
int launch_sample2d(const float *tex, int H, int W, int C, const float *uv, long long n, int wrap, int mode,
                    float *out, hipStream_t st);
int launch_sample2d_bwd(const float *tex, int H, int W, int C, const float *uv, long long n, int wrap, int mode,
                        const float *gout, float *gtex, float *guv, hipStream_t st);
int launch_sample3d(const float *tex, int D, int H, int W, int C, const float *uvw, long long n, int wrap, int mode,
                    float *out, hipStream_t st);
int launch_sample3d_bwd(const float *tex, int D, int H, int W, int C, const float *uvw, long long n, int wrap,
                        int mode, const float *gout, float *gtex, float *guvw, hipStream_t st);
int launch_stencil_partition(const float *rast, long long npx, const int *stencils, int F, int n_owner, int *order,
                             int *starts, int *ends, void *ws, hipStream_t st);
int launch_count_covered(const float *rast, int n_img, long long npx, int *counts, hipStream_t st);
int launch_incidence(const int *tri, int F, int V, int *inc_off, int *inc, void *ws, hipStream_t st);
size_t l2_loss_ws_bytes();
int launch_l2_loss(const float *a, const float *b, long long n, float *loss, float *grad_a, void *ws,
                   hipStream_t st);
int launch_l2_loss_scale(float *grad, long long n, const float *grad_loss, hipStream_t st);
int launch_l2_loss_bwd(const float *a, const float *b, long long n, const float *grad_loss, float *ga, float *gb,
                       hipStream_t st);
int launch_clip_fwd(const float *xyz, int V, const float *VP, int n, float *pos, hipStream_t st);
int launch_clip_bwd(const float *grad_pos, int V, const float *VP, int n, float *grad_xyz, hipStream_t st);

}  // namespace drp
