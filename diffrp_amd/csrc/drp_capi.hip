// C-ABI of the rasterisation path: plain pointers and sizes, work enqueued on
// the caller's stream, no allocation (the caller passes a workspace sized by
// the *_workspace_size queries), no implicit device synchronisation, so every
// entry point can be captured into a HIP graph.  Status: 0 ok, < 0 error with
// a message in drp_last_error() (thread-local).
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>

#include "../../include/drp.h"
#include "drp_kernels.h"

using namespace drp;

namespace {
thread_local std::string g_err;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr int kMaxDecisionTris = (1 << 27) - 1;

bool tile_ok(int tw, int th) {
    return (tw == 8 && th == 8) || (tw == 16 && th == 16) || (tw == 32 && th == 8) || (tw == 16 && th == 8);
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// views folded into one attribute partial: a power of two in [1, 8]
int views_per_block(int v) {
    int p = 1;
    while (p * 2 <= v && p < 8) p *= 2;
    return p;
}

int channel_variant(int C) { return (C == 1 || C == 2 || C == 3 || C == 4 || C == 8 || C == 16) ? C : 0; }

struct RasterWs {
    int *bin_count, *list_count, *bins, *lists;
    int2 *list_box;
    size_t bytes;
};

RasterWs raster_ws_layout(void *base, long long n_tiles, int n_img, long long list_total, int bin_cap) {
    RasterWs w;
    size_t off = 0;
    char *b = (char *)base;
    size_t lt = (size_t)(list_total > 0 ? list_total : 1);
    w.bin_count = (int *)(b + off);
    w.list_count = w.bin_count + n_tiles;
    off = align256(sizeof(int) * (size_t)(n_tiles + n_img));
    w.bins = (int *)(b + off);
    off += align256(sizeof(int) * (size_t)n_tiles * (size_t)bin_cap);
    w.lists = (int *)(b + off);
    off += align256(sizeof(int) * lt);
    w.list_box = (int2 *)(b + off);
    off += align256(sizeof(int2) * lt);
    w.bytes = off;
    return w;
}

int check_dims(int n_img, int H, int W) {
    if (n_img < 0) return fail(-1, "n_img must be >= 0 (got %d)", n_img);
    if (H < 1 || W < 1 || H > kMaxRes || W > kMaxRes)
        return fail(-1, "resolution %dx%d outside [1, %d]", H, W, kMaxRes);
    return 0;
}
}  // namespace

extern "C" {

const char *drp_last_error(void) { return g_err.c_str(); }

int drp_abi_version(void) { return DRP_ABI_VERSION; }

size_t drp_rasterize_workspace_size(int n_img, int H, int W, long long list_total, int tile_w, int tile_h,
                                    int bin_cap) {
    if (n_img <= 0 || H <= 0 || W <= 0 || tile_w <= 0 || tile_h <= 0 || bin_cap <= 0) return 0;
    long long tx = (W + tile_w - 1) / tile_w, ty = (H + tile_h - 1) / tile_h;
    return raster_ws_layout(nullptr, tx * ty * n_img, n_img, list_total, bin_cap).bytes;
}

int drp_rasterize_fwd(void *stream, const float *pos, int n_img, int V, int pos_batched, const int32_t *tri, int F,
                      const int32_t *ranges_dev, int H, int W, const float *prev, float *rast, float *rast_db,
                      void *workspace, size_t workspace_bytes, int tile_w, int tile_h, int bin_cap,
                      long long list_total, int max_range, int32_t *boxes) {
    if (int e = check_dims(n_img, H, W)) return e;
    if (n_img == 0) return 0;
    if (V < 0 || F < 0) return fail(-1, "negative V (%d) or F (%d)", V, F);
    if (!rast) return fail(-1, "rast output is null");
    if ((V > 0 && !pos) || (F > 0 && !tri)) return fail(-1, "pos/tri pointer is null");
    if (pos && !aligned16(pos)) return fail(-1, "pos must be 16-byte aligned");
    if (!aligned16(rast) || (rast_db && !aligned16(rast_db)) || (prev && !aligned16(prev)))
        return fail(-1, "rast/rast_db/prev must be 16-byte aligned");
    if (!tile_ok(tile_w, tile_h)) return fail(-1, "unsupported tile shape %dx%d", tile_w, tile_h);
    if (bin_cap < 1) return fail(-1, "bin capacity must be >= 1");
    Batch b{};
    b.n_img = n_img;
    b.V = V;
    b.F = F;
    b.H = H;
    b.W = W;
    b.pos_batched = pos_batched;
    b.ranges = ranges_dev;
    b.tiles_x = (W + tile_w - 1) / tile_w;
    b.tiles_y = (H + tile_h - 1) / tile_h;
    long long n_tiles = (long long)b.tiles_x * b.tiles_y * n_img;
    if (n_tiles > 0x7fffffffll) return fail(-1, "too many tiles");
    if (!ranges_dev) {
        list_total = (long long)n_img * F;
        max_range = F;
    }
    RasterWs ws = raster_ws_layout(workspace, n_tiles, n_img, list_total, bin_cap);
    if (!workspace || workspace_bytes < ws.bytes)
        return fail(-1, "workspace too small: need %zu bytes, got %zu", ws.bytes, workspace_bytes);
    RasterLaunch L{};
    L.b = b;
    L.pos = pos;
    L.tri = tri;
    L.tile_w = tile_w;
    L.tile_h = tile_h;
    L.bin_cap = bin_cap;
    L.bin_count = ws.bin_count;
    L.list_count = ws.list_count;
    L.bins = ws.bins;
    L.lists = ws.lists;
    L.list_box = ws.list_box;
    L.boxes = (ItemBox *)boxes;
    if (boxes && ((uintptr_t)boxes & 3)) return fail(-1, "boxes must be 4-byte aligned");
    L.n_tiles = (int)n_tiles;
    L.max_tris = max_range;
    L.prev = prev;
    L.rast = rast;
    L.rast_db = rast_db;
    int rc = launch_rasterize(L, (hipStream_t)stream);
    if (rc) return fail(rc, "rasterize launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

size_t drp_rasterize_bwd_workspace_size(int n_img, int H, int W, int F) {
    if (n_img <= 0 || F <= 0 || H <= 0 || W <= 0) return 0;
    return align256((size_t)n_img * (size_t)F * 3 * 16) + tile_records_bytes(n_img, H, W, 12);
}

int drp_rasterize_bwd(void *stream, const float *pos, int n_img, int V, int pos_batched, const int32_t *tri, int F,
                      const int32_t *ranges_dev, int H, int W, const float *rast, const float *grad_rast,
                      const float *grad_db, float *grad_pos, const int32_t *boxes, const int32_t *inc_off,
                      const int32_t *inc, void *workspace, size_t workspace_bytes) {
    if (int e = check_dims(n_img, H, W)) return e;
    if (n_img == 0) return 0;
    if (!pos || !tri || !rast || !grad_rast || !grad_pos) return fail(-1, "null pointer argument");
    if (!aligned16(pos) || !aligned16(rast) || !aligned16(grad_rast) || (grad_db && !aligned16(grad_db)) ||
        !aligned16(grad_pos))
        return fail(-1, "pos/rast/grad buffers must be 16-byte aligned");
    bool tri_path = boxes && inc_off && inc && !ranges_dev;
    if (tri_path && (!workspace || workspace_bytes < drp_rasterize_bwd_workspace_size(n_img, H, W, F)))
        return fail(-1, "workspace too small: need %zu bytes, got %zu",
                    drp_rasterize_bwd_workspace_size(n_img, H, W, F),
                    workspace_bytes);
    RasterBwdLaunch L{};
    if (tri_path) {
        L.boxes = (const ItemBox *)boxes;
        L.partial = (float *)workspace;
        L.records = (char *)workspace + align256((size_t)n_img * (size_t)F * 3 * 16);
        L.inc_off = inc_off;
        L.inc = inc;
    }
    L.b.n_img = n_img;
    L.b.V = V;
    L.b.F = F;
    L.b.H = H;
    L.b.W = W;
    L.b.pos_batched = pos_batched;
    L.b.ranges = ranges_dev;
    L.b.tiles_x = 1;
    L.b.tiles_y = 1;
    L.pos = pos;
    L.tri = tri;
    L.rast = rast;
    L.grad_rast = grad_rast;
    L.grad_db = grad_db;
    L.grad_pos = grad_pos;
    L.total_pix = (long long)n_img * H * W;
    L.pix_per_img = H * W;
    int rc = launch_rasterize_bwd(L, (hipStream_t)stream);
    if (rc) return fail(rc, "rasterize_bwd launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

int drp_interpolate_fwd(void *stream, const float *attr, int n_img, int V, int attr_batched, int C,
                        const float *rast, int H, int W, const int32_t *tri, int F, const float *rast_db,
                        const int32_t *diff_idx, int n_diff, float *out, float *out_da) {
    if (int e = check_dims(n_img, H, W)) return e;
    if (n_img == 0) return 0;
    if (C < 1) return fail(-1, "attribute channel count must be >= 1");
    if (!attr || !rast || !tri || !out) return fail(-1, "null pointer argument");
    if (n_diff > 0 && (!rast_db || !diff_idx || !out_da)) return fail(-1, "diff_attrs needs rast_db and out_da");
    if (!aligned16(rast) || (rast_db && !aligned16(rast_db))) return fail(-1, "rast/rast_db must be 16-byte aligned");
    InterpLaunch L{};
    L.attr = attr;
    L.V = V;
    L.C = C;
    L.cv = channel_variant(C);
    if (C % 4 == 0 && (!aligned16(attr) || !aligned16(out))) L.cv = 0;  // float4 path needs alignment
    L.attr_batched = attr_batched;
    L.rast = rast;
    L.tri = tri;
    L.F = F;
    L.total_pix = (long long)n_img * H * W;
    L.pix_per_img = (long long)H * W;
    L.rast_db = rast_db;
    L.diff_idx = diff_idx;
    L.n_diff = n_diff;
    L.out = out;
    L.out_da = n_diff > 0 ? out_da : nullptr;
    int rc = launch_interpolate(L, (hipStream_t)stream);
    if (rc) return fail(rc, "interpolate launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

static size_t interp_partial_bytes(int n_img, int F, int C, int attr_batched, int views_per_group) {
    int vpg = attr_batched ? 1 : views_per_block(views_per_group);
    size_t groups = ((size_t)n_img + vpg - 1) / vpg;
    return align256(groups * (size_t)F * 3 * (size_t)C * 4);
}

size_t drp_interpolate_bwd_workspace_size(int n_img, int H, int W, int F, int C, int attr_batched,
                                          int views_per_group) {
    if (n_img <= 0 || F <= 0 || C <= 0 || H <= 0 || W <= 0) return 0;
    return interp_partial_bytes(n_img, F, C, attr_batched, views_per_group) + tile_records_bytes(n_img, H, W, 3 * C);
}

int drp_interpolate_bwd(void *stream, const float *attr, int n_img, int V, int attr_batched, int C,
                        const float *rast, int H, int W, const int32_t *tri, int F, const float *rast_db,
                        const int32_t *diff_idx, int n_diff, const float *grad_out, const float *grad_da,
                        float *grad_attr, float *grad_rast, float *grad_db, const int32_t *boxes,
                        int views_per_group, const int32_t *inc_off, const int32_t *inc, void *workspace,
                        size_t workspace_bytes) {
    if (int e = check_dims(n_img, H, W)) return e;
    if (n_img == 0) return 0;
    if (C < 1) return fail(-1, "attribute channel count must be >= 1");
    if (!attr || !rast || !tri || !grad_out || !grad_attr || !grad_rast) return fail(-1, "null pointer argument");
    if (n_diff > 0 && (!rast_db || !diff_idx)) return fail(-1, "diff_attrs needs rast_db");
    if (!aligned16(rast) || !aligned16(grad_rast) || (grad_db && !aligned16(grad_db)) ||
        (rast_db && !aligned16(rast_db)))
        return fail(-1, "rast-shaped buffers must be 16-byte aligned");
    InterpLaunch L{};
    L.attr = attr;
    L.V = V;
    L.C = C;
    L.attr_batched = attr_batched;
    L.rast = rast;
    L.tri = tri;
    L.F = F;
    L.total_pix = (long long)n_img * H * W;
    L.pix_per_img = (long long)H * W;
    L.rast_db = rast_db;
    L.diff_idx = diff_idx;
    L.n_diff = n_diff;
    L.cv = channel_variant(C);
    L.grad_out = grad_out;
    L.grad_da = n_diff > 0 ? grad_da : nullptr;
    L.grad_attr = grad_attr;
    L.grad_rast = grad_rast;
    L.grad_db = grad_db;
    L.H = H;
    L.W = W;
    L.views_per_group = attr_batched ? 1 : views_per_block(views_per_group);
    bool tri_path = boxes && inc_off && inc && n_diff == 0 && L.cv > 0;
    if (tri_path) {
        size_t need = drp_interpolate_bwd_workspace_size(n_img, H, W, F, C, attr_batched, L.views_per_group);
        if (!workspace || workspace_bytes < need)
            return fail(-1, "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
        if (C % 4 == 0 && !aligned16(grad_out)) return fail(-1, "grad_out must be 16-byte aligned");
        L.boxes = (const ItemBox *)boxes;
        L.partial = (float *)workspace;
        L.records = (char *)workspace + interp_partial_bytes(n_img, F, C, attr_batched, L.views_per_group);
        L.inc_off = inc_off;
        L.inc = inc;
    }
    int rc = launch_interpolate_bwd(L, (hipStream_t)stream);
    if (rc) return fail(rc, "interpolate_bwd launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

int drp_edge_grad_bwd(void *stream, const float *grad_out, const float *rgb, int H, int W, int C, const float *rast,
                      const float *frag_xy, const float *tris_xy, int F, float *grad_frag_xy) {
    if (int e = check_dims(1, H, W)) return e;
    if (!(C == 1 || C == 3 || C == 4)) return fail(-1, "edge gradient supports 1, 3 or 4 channels (got %d)", C);
    if (!grad_out || !rgb || !rast || !frag_xy || !grad_frag_xy || (F > 0 && !tris_xy))
        return fail(-1, "null pointer argument");
    EdgeLaunch L{grad_out, rgb, rast, frag_xy, tris_xy, grad_frag_xy, H, W, C, F};
    int rc = launch_edge_grad_bwd(L, (hipStream_t)stream);
    if (rc) return fail(rc, "edge gradient launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

size_t drp_incidence_workspace_size(int V) { return V > 0 ? incidence_ws_bytes(V) : 0; }

int drp_incidence_build(void *stream, const int32_t *tri, int F, int V, int32_t *inc_off, int32_t *inc,
                        void *workspace, size_t workspace_bytes) {
    if (F < 0 || V < 0) return fail(-1, "negative F or V");
    if (F > 0x2aaaaaaa) return fail(-1, "too many triangles");
    if (V == 0) return 0;
    if ((F > 0 && !tri) || !inc_off || (F > 0 && !inc)) return fail(-1, "null pointer argument");
    if (!workspace || workspace_bytes < incidence_ws_bytes(V))
        return fail(-1, "workspace too small: need %zu bytes, got %zu", incidence_ws_bytes(V), workspace_bytes);
    int rc = launch_incidence(tri, F, V, inc_off, inc, workspace, (hipStream_t)stream);
    if (rc) return fail(rc, "incidence launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

static unsigned topo_mask(int F) {
    unsigned long long need = 3ull * (unsigned long long)F * 2ull;  // load factor <= 1/2 over edge slots
    unsigned long long n = 1024;
    while (n < need) n <<= 1;
    return (unsigned)(n - 1);
}

size_t drp_topology_workspace_size(int F) {
    if (F <= 0) return 0;
    size_t n = (size_t)topo_mask(F) + 1;
    return align256(n * 8) + align256(n * 4) + align256(n * 8);
}

int drp_topology_build(void *stream, const int32_t *tri, int F, int32_t *opp, void *workspace,
                       size_t workspace_bytes) {
    if (F < 0) return fail(-1, "negative F");
    if (F == 0) return 0;
    if (F > 0x2aaaaaaa) return fail(-1, "too many triangles for the topology table");
    if (!tri || !opp) return fail(-1, "null pointer argument");
    size_t need = drp_topology_workspace_size(F);
    if (!workspace || workspace_bytes < need)
        return fail(-1, "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    unsigned mask = topo_mask(F);
    size_t n = (size_t)mask + 1;
    char *b = (char *)workspace;
    TopoLaunch L{};
    L.tri = tri;
    L.F = F;
    L.keys = (unsigned long long *)b;
    L.cnt = (int *)(b + align256(n * 8));
    L.slots = (int2 *)(b + align256(n * 8) + align256(n * 4));
    L.mask = mask;
    L.opp = opp;
    int rc = launch_topology(L, (hipStream_t)stream);
    if (rc) return fail(rc, "topology launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

size_t drp_antialias_workspace_size(int n_img, int H, int W, int V, int F, int pos_batched) {
    if (n_img <= 0) return 0;
    size_t n_proj = pos_batched ? (size_t)n_img : 1;
    size_t npx = (size_t)n_img * H * W;
    return align256(n_proj * (size_t)V * 16) + 2 * align256(n_proj * (size_t)F) + 2 * align256(npx * 8);
}

static int aa_common(AALaunch &L, void *stream, const float *color, int n_img, int H, int W, int C,
                     const float *rast, const float *pos, int V, int pos_batched, const int32_t *tri, int F,
                     const int32_t *opp, void *workspace, size_t workspace_bytes) {
    if (int e = check_dims(n_img, H, W)) return e;
    if (C < 1) return fail(-1, "color channel count must be >= 1");
    if (!color || !rast || !pos || !tri || !opp) return fail(-1, "null pointer argument");
    if (!aligned16(pos) || !aligned16(rast)) return fail(-1, "pos/rast must be 16-byte aligned");
    if (C % 4 == 0 && !aligned16(color)) return fail(-1, "color must be 16-byte aligned when C %% 4 == 0");
    size_t need = drp_antialias_workspace_size(n_img, H, W, V, F, pos_batched);
    if (!workspace || workspace_bytes < need)
        return fail(-1, "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    size_t n_proj = pos_batched ? (size_t)n_img : 1;
    char *b = (char *)workspace;
    L.color = color;
    L.rast = rast;
    L.pos = pos;
    L.tri = tri;
    L.opp = opp;
    L.n_img = n_img;
    L.H = H;
    L.W = W;
    L.C = C;
    L.V = V;
    L.F = F;
    L.pos_batched = pos_batched;
    L.cv = (C == 1 || C == 3 || C == 4 || C == 8 || C == 16) ? C : 0;
    size_t npx = (size_t)n_img * H * W;
    size_t o_proj = 0, o_orient = align256(n_proj * (size_t)V * 16);
    size_t o_silh = o_orient + align256(n_proj * (size_t)F), o_w = o_silh + align256(n_proj * (size_t)F);
    size_t o_dec = o_w + align256(npx * 8);
    L.proj = (double2 *)(b + o_proj);
    L.orient = (signed char *)(b + o_orient);
    L.silh = (unsigned char *)(b + o_silh);
    L.weights = (float2 *)(b + o_w);
    // pair decisions pack the triangle into 27 bits; larger meshes recompute them in the backward
    L.decisions = F < kMaxDecisionTris ? (int2 *)(b + o_dec) : nullptr;
    (void)stream;
    return 0;
}

int drp_antialias_fwd(void *stream, const float *color, int n_img, int H, int W, int C, const float *rast,
                      const float *pos, int V, int pos_batched, const int32_t *tri, int F, const int32_t *opp,
                      float *out, void *workspace, size_t workspace_bytes) {
    if (n_img == 0) return 0;
    AALaunch L{};
    if (int e = aa_common(L, stream, color, n_img, H, W, C, rast, pos, V, pos_batched, tri, F, opp, workspace,
                          workspace_bytes))
        return e;
    if (!out) return fail(-1, "null output");
    if (C % 4 == 0 && !aligned16(out)) return fail(-1, "out must be 16-byte aligned when C %% 4 == 0");
    L.out = out;
    int rc = launch_antialias(L, (hipStream_t)stream);
    if (rc) return fail(rc, "antialias launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

int drp_antialias_bwd(void *stream, const float *color, int n_img, int H, int W, int C, const float *rast,
                      const float *pos, int V, int pos_batched, const int32_t *tri, int F, const int32_t *opp,
                      const float *grad_out, float *grad_color, float *grad_pos, float pos_gradient_boost,
                      int reuse_forward, void *workspace, size_t workspace_bytes) {
    if (n_img == 0) return 0;
    AALaunch L{};
    if (int e = aa_common(L, stream, color, n_img, H, W, C, rast, pos, V, pos_batched, tri, F, opp, workspace,
                          workspace_bytes))
        return e;
    if (!grad_out || !grad_color) return fail(-1, "null gradient buffer");
    if (C % 4 == 0 && (!aligned16(grad_out) || !aligned16(grad_color)))
        return fail(-1, "grad_out/grad_color must be 16-byte aligned when C %% 4 == 0");
    L.grad_out = grad_out;
    L.grad_color = grad_color;
    L.grad_pos = grad_pos;
    L.boost = pos_gradient_boost;
    L.reuse = reuse_forward && L.decisions != nullptr;
    int rc = launch_antialias_bwd(L, (hipStream_t)stream);
    if (rc) return fail(rc, "antialias_bwd launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

}  // extern "C"
