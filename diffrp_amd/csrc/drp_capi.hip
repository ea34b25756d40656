// C-ABI of the rasterisation path: plain pointers and sizes, work enqueued on
// the caller's stream, no allocation (the caller passes a workspace sized by
// the *_workspace_size queries), no implicit device synchronisation, so every
// entry point can be captured into a HIP graph.  Status: 0 ok, < 0 error with
// a message in drp_last_error() (thread-local).
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
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
    int4 *setup;
    char *split_base;
    int split_slots;
    size_t bytes;
};

RasterWs raster_ws_layout(void *base, long long n_tiles, int n_img, long long list_total, long long bin_slots) {
    RasterWs w;
    size_t off = 0;
    char *b = (char *)base;
    size_t lt = (size_t)(list_total > 0 ? list_total : 1);
    w.bin_count = (int *)(b + off);
    w.list_count = w.bin_count + n_tiles;
    off = align256(sizeof(int) * (size_t)(n_tiles + n_img + kSplitCtl));
    w.bins = (int *)(b + off);
    off += align256(sizeof(int) * (size_t)bin_slots);
    w.lists = (int *)(b + off);
    off += align256(sizeof(int) * lt);
    w.setup = (int4 *)(b + off);
    off += align256(2 * sizeof(int4) * lt);
    // heavy-tile split region (its control words follow the list counters)
    w.split_slots = split_slots(n_tiles);
    w.split_base = b + off;
    off += align256(split_region_bytes(w.split_slots));
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
    return raster_ws_layout(nullptr, tx * ty * n_img, n_img, list_total, tx * ty * n_img * (long long)bin_cap).bytes;
}

static void aa_layout(AALaunch &L, void *workspace, int n_img, int H, int W, int V, int F, int pos_batched);

static int raster_fwd(void *stream, const float *pos, int n_img, int V, int pos_batched, const int32_t *tri, int F,
                      const int32_t *ranges_dev, int H, int W, const float *prev, float *rast, float *rast_db,
                      void *workspace, size_t workspace_bytes, int tile_w, int tile_h, int bin_cap,
                      long long list_total, int max_range, int32_t *boxes, const AALaunch *aa) {
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
    b.bin_cap = bin_cap;
    RasterWs ws = raster_ws_layout(workspace, n_tiles, n_img, list_total, n_tiles * (long long)bin_cap);
    if (!workspace || workspace_bytes < ws.bytes)
        return fail(-1, "workspace too small: need %zu bytes, got %zu", ws.bytes, workspace_bytes);
    RasterLaunch L{};
    L.b = b;
    L.pos = pos;
    L.tri = tri;
    L.tile_w = tile_w;
    L.tile_h = tile_h;
    L.bin_count = ws.bin_count;
    L.list_count = ws.list_count;
    L.bins = ws.bins;
    L.lists = ws.lists;
    L.setup = ws.setup;
    L.split_base = ws.split_base;
    L.split_slots = ws.split_slots;
    L.boxes = (ItemBox *)boxes;
    if (boxes && ((uintptr_t)boxes & 3)) return fail(-1, "boxes must be 4-byte aligned");
    L.n_tiles = (int)n_tiles;
    L.max_tris = max_range;
    L.prev = prev;
    L.rast = rast;
    L.rast_db = rast_db;
    L.aa_prepare = aa;
    L.depth_cull = depth_cull_wanted(bin_cap, tile_w, tile_h);
    int rc = launch_rasterize(L, (hipStream_t)stream);
    if (rc) return fail(rc, "rasterize launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

int drp_rasterize_fwd(void *stream, const float *pos, int n_img, int V, int pos_batched, const int32_t *tri, int F,
                      const int32_t *ranges_dev, int H, int W, const float *prev, float *rast, float *rast_db,
                      void *workspace, size_t workspace_bytes, int tile_w, int tile_h, int bin_cap,
                      long long list_total, int max_range, int32_t *boxes) {
    return raster_fwd(stream, pos, n_img, V, pos_batched, tri, F, ranges_dev, H, W, prev, rast, rast_db, workspace,
                      workspace_bytes, tile_w, tile_h, bin_cap, list_total, max_range, boxes, nullptr);
}

int drp_rasterize_fwd_aa_prepare(void *stream, const float *pos, int n_img, int V, int pos_batched,
                                 const int32_t *tri, int F, int H, int W, float *rast, float *rast_db,
                                 void *workspace, size_t workspace_bytes, int tile_w, int tile_h, int bin_cap,
                                 int32_t *boxes, const int32_t *opp, void *aa_workspace, size_t aa_workspace_bytes) {
    if (int e = check_dims(n_img, H, W)) return e;
    if (n_img == 0) return 0;
    if (n_img > 65535) return fail(-1, "antialias takes at most 65535 images per call (got %d)", n_img);
    if (V <= 0 || F <= 0) return fail(-1, "the antialias preparation needs vertices and triangles");
    if (!opp) return fail(-1, "null pointer argument");
    if (pos_batched && (long long)n_img * V >= (1ll << 31))
        return fail(-1, "antialias with per-image positions takes fewer than 2^31 vertices in all");
    const size_t need = drp_antialias_workspace_size(n_img, H, W, V, F, pos_batched);
    if (!aa_workspace || aa_workspace_bytes < need)
        return fail(-1, "antialias workspace too small: need %zu bytes, got %zu", need, aa_workspace_bytes);
    AALaunch A{};
    A.pos = pos;
    A.tri = tri;
    A.opp = opp;
    A.n_img = n_img;
    A.H = H;
    A.W = W;
    A.V = V;
    A.F = F;
    A.pos_batched = pos_batched;
    aa_layout(A, aa_workspace, n_img, H, W, V, F, pos_batched);
    return raster_fwd(stream, pos, n_img, V, pos_batched, tri, F, nullptr, H, W, nullptr, rast, rast_db, workspace,
                      workspace_bytes, tile_w, tile_h, bin_cap, (long long)n_img * F, F, boxes, &A);
}

// ---- allocation-free rasterize ---------------------------------------------
// The workspace of drp_rasterize_fwd_noalloc is one file-scope device array,
// `scratch`, of kNoAllocScratch bytes: each GPU the library runs on holds its
// own copy until the process ends.  Host code never passes the array's name
// to a kernel (on the host it is only a stand-in symbol): its device address
// is looked up once per device, with that device current, and kept here.
namespace {
constexpr size_t kNoAllocScratch = DRP_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxScratchDevices = 64;
__device__ __align__(256) unsigned char scratch[kNoAllocScratch];
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices];

int scratch_address(void **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(-2, "hipGetDevice failed: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxScratchDevices) return fail(-1, "device %d beyond the %d supported", dev, kMaxScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(-2, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

int drp_rasterize_fwd_noalloc(void *stream, const float *pos, int n_img, int V, int pos_batched,
                              const int32_t *tri, int F, const int32_t *ranges_dev, int H, int W, const float *prev,
                              float *rast, float *rast_db, int tile_w, int tile_h, int bin_cap, long long list_total,
                              int max_range, int32_t *boxes) {
    if (int e = check_dims(n_img, H, W)) return e;
    if (n_img == 0) return 0;
    if (!tile_ok(tile_w, tile_h)) return fail(-1, "unsupported tile shape %dx%d", tile_w, tile_h);
    if (bin_cap < 1) return fail(-1, "bin capacity must be >= 1");
    if (F < 0) return fail(-1, "negative F (%d)", F);
    long long lt = ranges_dev ? list_total : (long long)n_img * F;
    size_t need = drp_rasterize_workspace_size(n_img, H, W, lt, tile_w, tile_h, bin_cap);
    if (need > kNoAllocScratch)
        return fail(-3, "rasterize needs %zu bytes of scratch; the allocation-free entry point has %zu "
                        "(use drp_rasterize_fwd with a workspace, or fewer images / a smaller bin capacity)",
                    need, kNoAllocScratch);
    void *ws = nullptr;
    if (int e = scratch_address(&ws)) return e;
    return drp_rasterize_fwd(stream, pos, n_img, V, pos_batched, tri, F, ranges_dev, H, W, prev, rast, rast_db, ws,
                             kNoAllocScratch, tile_w, tile_h, bin_cap, list_total, max_range, boxes);
}

// ---- uneven batches --------------------------------------------------------
// Plan (host and device copies byte-identical): header, ImageDesc[n_img]
// (16-byte aligned), tile -> image map int32[n_tiles].
struct BatchPlanHeader {
    int n_img, tile_w, tile_h, max_count;
    long long n_tiles, n_pix, n_tri, n_vert, bin_slots;
};

// bin capacity of an item when none is given: ~6 F / tiles + 32 (a closed mesh
// covers ~2 tiles per triangle, dense tiles ~3x the mean), a power of two
static int auto_bin_cap(int F, long long tiles) {
    double est = 6.0 * F / (double)(tiles > 0 ? tiles : 1) + 32.0;
    int cap = 32;
    while (cap < est && cap < (1 << 16)) cap *= 2;
    return cap;
}

static size_t plan_desc_off() { return (sizeof(BatchPlanHeader) + 15) & ~(size_t)15; }
static size_t plan_t2i_off(int n_img) { return plan_desc_off() + (size_t)n_img * sizeof(ImageDesc); }
// item map (UnevenMap) after the tile -> image table: pixel prefix (int64), the
// vertex, triangle, 16 x 16 tile, 64 x 8 block and 32 x 32 tile prefixes
// (int32, n + 1 each), then (H, W) per item
static size_t plan_map_off(int n_img, long long n_tiles) {
    return (plan_t2i_off(n_img) + sizeof(int32_t) * (size_t)n_tiles + 15) & ~(size_t)15;
}
static size_t plan_map_bytes(int n_img) { return 8 * (size_t)(n_img + 1) + 20 * (size_t)(n_img + 1) + 8 * (size_t)n_img; }
static UnevenMap plan_map(const void *plan, int n_img, long long n_tiles) {
    UnevenMap m;
    const char *b = (const char *)plan + plan_map_off(n_img, n_tiles);
    m.pix = (const long long *)b;
    const int32_t *q = (const int32_t *)(b + 8 * (size_t)(n_img + 1));
    m.vert = q;
    m.tri = q + (n_img + 1);
    m.bt16 = q + 2 * (n_img + 1);
    m.blk = q + 3 * (n_img + 1);
    m.t32 = q + 4 * (n_img + 1);
    m.hw = q + 5 * (n_img + 1);
    m.n = n_img;
    return m;
}

static long long plan_tiles(const int32_t *items, int n_img, int tile_w, int tile_h) {
    long long n = 0;
    for (int i = 0; i < n_img; ++i)
        n += (long long)((items[5 * i + 3] + tile_w - 1) / tile_w) * ((items[5 * i + 2] + tile_h - 1) / tile_h);
    return n;
}

size_t drp_batch_plan_size(const int32_t *items, int n_img, int tile_w, int tile_h) {
    if (!items || n_img <= 0 || !tile_ok(tile_w, tile_h)) return 0;
    return plan_map_off(n_img, plan_tiles(items, n_img, tile_w, tile_h)) + plan_map_bytes(n_img);
}

int drp_batch_plan(const int32_t *items, int n_img, int tile_w, int tile_h, void *plan, size_t plan_bytes) {
    if (!items || n_img <= 0) return fail(-1, "empty batch");
    if (!tile_ok(tile_w, tile_h)) return fail(-1, "unsupported tile shape %dx%d", tile_w, tile_h);
    long long vb = 0, fb = 0, pb = 0, tb = 0, sb = 0;
    int max_count = 0;
    for (int i = 0; i < n_img; ++i) {
        int V = items[5 * i], F = items[5 * i + 1], H = items[5 * i + 2], W = items[5 * i + 3];
        int cap = items[5 * i + 4];
        if (V < 0 || F < 0 || cap < 0) return fail(-1, "item %d: negative V, F or bin capacity", i);
        if (int e = check_dims(1, H, W)) return e;
        long long nt = (long long)((W + tile_w - 1) / tile_w) * ((H + tile_h - 1) / tile_h);
        vb += V;
        fb += F;
        pb += (long long)H * W;
        tb += nt;
        sb += nt * (cap > 0 ? cap : auto_bin_cap(F, nt));
        max_count = F > max_count ? F : max_count;
    }
    if (tb > 0x7fffffffll || fb > 0x7fffffffll || vb > 0x7fffffffll) return fail(-1, "batch too large");
    size_t need = plan_map_off(n_img, tb) + plan_map_bytes(n_img);
    if (!plan || plan_bytes < need) return fail(-1, "plan buffer too small: need %zu bytes", need);
    char *base = (char *)plan;
    BatchPlanHeader h{n_img, tile_w, tile_h, max_count, tb, pb, fb, vb, sb};
    memcpy(base, &h, sizeof(h));
    ImageDesc *D = (ImageDesc *)(base + plan_desc_off());
    int32_t *t2i = (int32_t *)(base + plan_t2i_off(n_img));
    {
        UnevenMap m = plan_map(plan, n_img, tb);
        long long *pix = (long long *)m.pix;
        int32_t *vert = (int32_t *)m.vert, *tri = (int32_t *)m.tri, *bt16 = (int32_t *)m.bt16;
        int32_t *blk = (int32_t *)m.blk, *t32 = (int32_t *)m.t32, *hw = (int32_t *)m.hw;
        long long p = 0, v = 0, f = 0, a = 0, k = 0, c = 0;
        for (int i = 0; i <= n_img; ++i) {
            pix[i] = p;
            vert[i] = (int32_t)v;
            tri[i] = (int32_t)f;
            bt16[i] = (int32_t)a;
            blk[i] = (int32_t)k;
            t32[i] = (int32_t)c;
            if (i == n_img) break;
            int V = items[5 * i], F = items[5 * i + 1], H = items[5 * i + 2], W = items[5 * i + 3];
            hw[2 * i] = H;
            hw[2 * i + 1] = W;
            p += (long long)H * W;
            v += V;
            f += F;
            a += (long long)((W + 15) / 16) * ((H + 15) / 16);
            k += (long long)((W + 63) / 64) * ((H + 7) / 8);
            c += (long long)((W + 31) / 32) * ((H + 31) / 32);
        }
        if (a > 0x7fffffffll || k > 0x7fffffffll || c > 0x7fffffffll) return fail(-1, "batch too large");
    }
    vb = fb = pb = tb = sb = 0;
    for (int i = 0; i < n_img; ++i) {
        int V = items[5 * i], F = items[5 * i + 1], H = items[5 * i + 2], W = items[5 * i + 3];
        ImageDesc d{};
        d.pos_off = 4 * vb;
        d.pix_off = pb;
        d.V = V;
        d.f0 = (int)fb;
        d.f1 = (int)(fb + F);
        d.H = H;
        d.W = W;
        d.tiles_x = (W + tile_w - 1) / tile_w;
        d.tiles_y = (H + tile_h - 1) / tile_h;
        d.tile_off = (int)tb;
        d.list_off = (int)fb;
        d.list_cap = F;
        long long nt = (long long)d.tiles_x * d.tiles_y;
        d.bin_cap = items[5 * i + 4] > 0 ? items[5 * i + 4] : auto_bin_cap(F, nt);
        d.bin_off = sb;
        memcpy(D + i, &d, sizeof(d));
        for (long long k = 0; k < nt; ++k) t2i[tb + k] = i;
        vb += V;
        fb += F;
        pb += (long long)H * W;
        tb += nt;
        sb += nt * d.bin_cap;
    }
    return 0;
}

int drp_batch_plan_info(const void *plan, long long *n_pix, long long *n_tiles, long long *n_tri, long long *n_vert) {
    if (!plan) return fail(-1, "null plan");
    BatchPlanHeader h;
    memcpy(&h, plan, sizeof(h));
    if (n_pix) *n_pix = h.n_pix;
    if (n_tiles) *n_tiles = h.n_tiles;
    if (n_tri) *n_tri = h.n_tri;
    if (n_vert) *n_vert = h.n_vert;
    return 0;
}

size_t drp_rasterize_batch_workspace_size(const void *plan) {
    if (!plan) return 0;
    BatchPlanHeader h;
    memcpy(&h, plan, sizeof(h));
    return raster_ws_layout(nullptr, h.n_tiles, h.n_img, h.n_tri, h.bin_slots).bytes;
}

int drp_rasterize_batch_fwd(void *stream, const float *pos, const int32_t *tri, const void *plan,
                            const void *plan_dev, const float *prev, float *rast, float *rast_db, void *workspace,
                            size_t workspace_bytes, int32_t *boxes, int global_ids) {
    if (!plan || !plan_dev) return fail(-1, "null plan");
    BatchPlanHeader h;
    memcpy(&h, plan, sizeof(h));
    if (h.n_img <= 0 || !tile_ok(h.tile_w, h.tile_h)) return fail(-1, "invalid plan");
    if (!rast) return fail(-1, "rast output is null");
    if ((h.n_vert > 0 && !pos) || (h.n_tri > 0 && !tri)) return fail(-1, "pos/tri pointer is null");
    if ((pos && !aligned16(pos)) || !aligned16(rast) || (rast_db && !aligned16(rast_db)) ||
        (prev && !aligned16(prev)) || !aligned16(plan_dev))
        return fail(-1, "pos/rast/rast_db/prev/plan must be 16-byte aligned");
    if (boxes && ((uintptr_t)boxes & 3)) return fail(-1, "boxes must be 4-byte aligned");
    RasterWs ws = raster_ws_layout(workspace, h.n_tiles, h.n_img, h.n_tri, h.bin_slots);
    if (!workspace || workspace_bytes < ws.bytes)
        return fail(-1, "workspace too small: need %zu bytes, got %zu", ws.bytes, workspace_bytes);
    Batch b{};
    b.descs = (const ImageDesc *)((const char *)plan_dev + plan_desc_off());
    b.tile_to_img = (const int *)((const char *)plan_dev + plan_t2i_off(h.n_img));
    b.n_img = h.n_img;
    b.V = (int)h.n_vert;
    b.F = (int)h.n_tri;
    b.local_ids = global_ids ? 0 : 1;
    RasterLaunch L{};
    L.b = b;
    L.pos = pos;
    L.tri = tri;
    L.tile_w = h.tile_w;
    L.tile_h = h.tile_h;
    L.bin_count = ws.bin_count;
    L.list_count = ws.list_count;
    L.bins = ws.bins;
    L.lists = ws.lists;
    L.setup = ws.setup;
    L.split_base = ws.split_base;
    L.split_slots = ws.split_slots;
    L.boxes = (ItemBox *)boxes;
    L.n_tiles = (int)h.n_tiles;
    L.max_tris = h.max_count;
    L.prev = prev;
    L.rast = rast;
    L.rast_db = rast_db;
    {
        const ImageDesc *D = (const ImageDesc *)((const char *)plan + plan_desc_off());
        int cap = 0;
        for (int i = 0; i < h.n_img; ++i) cap = D[i].bin_cap > cap ? D[i].bin_cap : cap;
        L.depth_cull = depth_cull_wanted(cap, h.tile_w, h.tile_h);
    }
    int rc = launch_rasterize(L, (hipStream_t)stream);
    if (rc) return fail(rc, "batch rasterize launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

// ---- uneven batches: interpolate, antialias and the backward passes --------
// The items' buffers packed back to back (rast [n_pix, 4] with global triangle
// ids from drp_rasterize_batch_fwd(..., global_ids = 1)); pos [n_vert, 4] and
// attr [n_vert, C] concatenated; tri [n_tri, 3] with global vertex indices.
namespace {
struct PlanView {
    BatchPlanHeader h;
    UnevenMap um;
    long long bt16, blk, t32;  // totals of the item map's tilings
};

int plan_view(const void *plan, const void *plan_dev, PlanView &v) {
    if (!plan || !plan_dev) return fail(-1, "null plan");
    if (!aligned16(plan_dev)) return fail(-1, "plan must be 16-byte aligned");
    memcpy(&v.h, plan, sizeof(v.h));
    if (v.h.n_img <= 0 || !tile_ok(v.h.tile_w, v.h.tile_h)) return fail(-1, "invalid plan");
    const UnevenMap host = plan_map(plan, v.h.n_img, v.h.n_tiles);
    v.bt16 = host.bt16[v.h.n_img];
    v.blk = host.blk[v.h.n_img];
    v.t32 = host.t32[v.h.n_img];
    v.um = plan_map(plan_dev, v.h.n_img, v.h.n_tiles);
    return 0;
}
}  // namespace

int drp_batch_interpolate_fwd(void *stream, const void *plan, const float *attr, int C, const float *rast,
                              const int32_t *tri, float *out) {
    if (!plan) return fail(-1, "null plan");
    BatchPlanHeader h;
    memcpy(&h, plan, sizeof(h));
    if (C < 1) return fail(-1, "attribute channel count must be >= 1");
    if (h.n_pix == 0) return 0;
    if (!rast || !out || (h.n_vert > 0 && !attr) || (h.n_tri > 0 && !tri)) return fail(-1, "null pointer argument");
    if (!aligned16(rast) || (C % 4 == 0 && (!aligned16(out) || (attr && !aligned16(attr)))))
        return fail(-1, "rast (and out, attr when C %% 4 == 0) must be 16-byte aligned");
    InterpLaunch L{};
    L.attr = attr;
    L.V = (int)h.n_vert;
    L.C = C;
    L.rast = rast;
    L.tri = tri;
    L.F = (int)h.n_tri;
    L.total_pix = h.n_pix;
    L.pix_per_img = h.n_pix;
    L.cv = channel_variant(C);
    L.out = out;
    int rc = launch_interpolate(L, (hipStream_t)stream);
    if (rc) return fail(rc, "batch interpolate launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

static size_t batch_interp_partial_bytes(const BatchPlanHeader &h, int C) {
    return align256((size_t)h.n_tri * 3 * (size_t)C * 4);
}

size_t drp_batch_interpolate_bwd_workspace_size(const void *plan, int C) {
    PlanView v;
    if (!plan || C <= 0) return 0;
    memcpy(&v.h, plan, sizeof(v.h));
    long long tiles = plan_map(plan, v.h.n_img, v.h.n_tiles).bt16[v.h.n_img];
    return batch_interp_partial_bytes(v.h, C) + tile_records_bytes_tiles(tiles, 3 * C, v.h.n_tri);
}

int drp_batch_interpolate_bwd(void *stream, const void *plan, const void *plan_dev, const float *attr, int C,
                              const float *rast, const int32_t *tri, const float *grad_out, float *grad_attr,
                              float *grad_rast, const int32_t *boxes, const int32_t *inc_off, const int32_t *inc,
                              void *workspace, size_t workspace_bytes) {
    PlanView v;
    if (int e = plan_view(plan, plan_dev, v)) return e;
    const int cv = channel_variant(C);
    if (cv == 0) return fail(-1, "batch interpolate backward supports C in {1, 2, 3, 4, 8, 16} (got %d)", C);
    if (v.h.n_pix == 0) return 0;
    if (!rast || !grad_out || !grad_rast || (v.h.n_vert > 0 && (!attr || !grad_attr)) ||
        (v.h.n_tri > 0 && (!tri || !boxes || !inc_off || !inc)))
        return fail(-1, "null pointer argument");
    if (!aligned16(rast) || !aligned16(grad_rast) || (C % 4 == 0 && (!aligned16(grad_out) || !aligned16(attr))))
        return fail(-1, "rast-shaped buffers (and attr, grad_out when C %% 4 == 0) must be 16-byte aligned");
    size_t need = drp_batch_interpolate_bwd_workspace_size(plan, C);
    if (!workspace || workspace_bytes < need)
        return fail(-1, "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    InterpLaunch L{};
    L.attr = attr;
    L.V = (int)v.h.n_vert;
    L.C = C;
    L.rast = rast;
    L.tri = tri;
    L.F = (int)v.h.n_tri;
    L.total_pix = v.h.n_pix;
    L.pix_per_img = v.h.n_pix;
    L.cv = cv;
    L.grad_out = grad_out;
    L.grad_attr = grad_attr;
    L.grad_rast = grad_rast;
    L.views_per_group = 1;
    L.boxes = (const ItemBox *)boxes;
    L.partial = (float *)workspace;
    L.records = (char *)workspace + batch_interp_partial_bytes(v.h, C);
    L.inc_off = inc_off;
    L.inc = inc;
    L.um = v.um;
    L.um_tiles = v.bt16;
    if (v.h.n_tri == 0) {  // nothing covered: zero grad_rast through the per-pixel path
        L.boxes = nullptr;
        L.partial = nullptr;
    }
    int rc = launch_interpolate_bwd(L, (hipStream_t)stream);
    if (rc) return fail(rc, "batch interpolate backward launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

// the attribute and position backward of an uneven batch from one pass over
// its pixels (the batched form of drp_interpolate_rasterize_bwd)
size_t drp_batch_interpolate_rasterize_bwd_workspace_size(const void *plan, int C) {
    if (!plan || C <= 0) return 0;
    BatchPlanHeader h;
    memcpy(&h, plan, sizeof(h));
    if (h.n_tri <= 0) return 0;
    long long tiles = plan_map(plan, h.n_img, h.n_tiles).bt16[h.n_img];
    return align256((size_t)h.n_tri * 3 * (size_t)C * 4) + align256((size_t)h.n_tri * 3 * 16) +
           tile_records_bytes_tiles(tiles, 3 * C + 12, h.n_tri) + align256((size_t)h.n_tri * 3 * 4);
}

int drp_batch_interpolate_rasterize_bwd(void *stream, const void *plan, const void *plan_dev, const float *attr,
                                        int C, const float *pos, const float *rast, const int32_t *tri,
                                        const float *grad_out, float *grad_attr, float *grad_pos,
                                        const int32_t *boxes, const int32_t *inc_off, const int32_t *inc,
                                        void *workspace, size_t workspace_bytes) {
    PlanView v;
    if (int e = plan_view(plan, plan_dev, v)) return e;
    if (!(C == 4 || C == 8)) return fail(-1, "the fused backward takes C in {4, 8} (got %d)", C);
    if (v.h.n_pix == 0 || v.h.n_tri == 0 || v.h.n_vert == 0) return 0;
    if (!attr || !pos || !rast || !tri || !grad_out || !grad_attr || !grad_pos || !boxes || !inc_off || !inc)
        return fail(-1, "null pointer argument");
    if (!aligned16(attr) || !aligned16(pos) || !aligned16(rast) || !aligned16(grad_out) || !aligned16(grad_pos))
        return fail(-1, "attr, pos, rast, grad_out and grad_pos must be 16-byte aligned");
    const size_t need = drp_batch_interpolate_rasterize_bwd_workspace_size(plan, C);
    if (!workspace || workspace_bytes < need)
        return fail(-1, "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    const int V = (int)v.h.n_vert, F = (int)v.h.n_tri;
    char *b = (char *)workspace;
    float *partial_attr = (float *)b;
    b += align256((size_t)F * 3 * (size_t)C * 4);
    float *partial_pos = (float *)b;
    b += align256((size_t)F * 3 * 16);
    void *records = b;
    b += tile_records_bytes_tiles(v.bt16, 3 * C + 12, F);
    int *slot = (int *)b;
    hipStream_t st = (hipStream_t)stream;
    if (int rc = launch_corner_slots(inc, inc_off, V, F, slot, st))
        return fail(rc, "corner slots launch failed: %s", hipGetErrorString(hipGetLastError()));
    TriBwdLaunch T{};
    T.b.n_img = 1;  // one image of packed items: the item map names each tile's item
    T.b.V = V;
    T.b.F = F;
    T.b.H = 1;
    T.b.W = 1;
    T.b.pos_batched = 0;
    T.pos = pos;
    T.tri = tri;
    T.rast = rast;
    T.grad_out = grad_out;
    T.attr = attr;
    T.V = V;
    T.attr_batched = 0;
    T.grad_rast_out = nullptr;
    T.boxes = (const ItemBox *)boxes;
    T.partial = partial_attr;
    T.partial_pos = partial_pos;
    T.corner_slot = slot;
    T.n_img = 1;
    T.F = F;
    T.H = 1;
    T.W = 1;
    T.views_per_block = 1;
    T.um = v.um;
    T.local_ids = 0;
    tile_records_layout_tiles(T, records, v.bt16, F);
    int rc = launch_tri_bwd_fused(T, C, st);
    Fork f{};  // the two vertex gathers are independent
    const bool par = !rc && fork_begin(st, f);
    if (!rc) rc = launch_gather_attr(partial_attr, 1, F, V, C, inc_off, inc, grad_attr, par ? f.side : st);
    if (!rc) rc = launch_gather_pos(partial_pos, 1, F, V, 0, inc_off, grad_pos, st);
    if (par && fork_end(st, f)) rc = -2;
    if (rc) return fail(rc, "batch fused backward launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

size_t drp_batch_rasterize_bwd_workspace_size(const void *plan) {
    if (!plan) return 0;
    BatchPlanHeader h;
    memcpy(&h, plan, sizeof(h));
    if (h.n_tri <= 0) return 0;
    long long tiles = plan_map(plan, h.n_img, h.n_tiles).bt16[h.n_img];
    return align256((size_t)h.n_tri * 3 * 16) + tile_records_bytes_tiles(tiles, 12, h.n_tri) +
           align256((size_t)h.n_tri * 3 * 4);
}

int drp_batch_rasterize_bwd(void *stream, const void *plan, const void *plan_dev, const float *pos, const int32_t *tri,
                            const float *rast, int local_ids, const float *grad_rast, const float *grad_db,
                            float *grad_pos, const int32_t *boxes, const int32_t *inc_off, const int32_t *inc,
                            void *workspace, size_t workspace_bytes) {
    PlanView v;
    if (int e = plan_view(plan, plan_dev, v)) return e;
    if (v.h.n_pix == 0 || v.h.n_tri == 0) return 0;
    if (!pos || !tri || !rast || !grad_rast || !grad_pos || !boxes || !inc_off || !inc)
        return fail(-1, "null pointer argument");
    if (!aligned16(pos) || !aligned16(rast) || !aligned16(grad_rast) || !aligned16(grad_pos) ||
        (grad_db && !aligned16(grad_db)))
        return fail(-1, "pos, rast, grad_rast, grad_db and grad_pos must be 16-byte aligned");
    size_t need = drp_batch_rasterize_bwd_workspace_size(plan);
    if (!workspace || workspace_bytes < need)
        return fail(-1, "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    RasterBwdLaunch L{};
    Batch b{};
    b.n_img = 1;
    b.V = (int)v.h.n_vert;
    b.F = (int)v.h.n_tri;
    b.H = 1;
    b.W = 1;
    b.pos_batched = 0;
    L.b = b;
    L.pos = pos;
    L.tri = tri;
    L.rast = rast;
    L.grad_rast = grad_rast;
    L.grad_db = grad_db;
    L.grad_pos = grad_pos;
    L.total_pix = v.h.n_pix;
    L.pix_per_img = 1;
    L.boxes = (const ItemBox *)boxes;
    L.partial = (float *)workspace;
    L.records = (char *)workspace + align256((size_t)v.h.n_tri * 3 * 16);
    L.corner_slot = (int *)((char *)L.records + tile_records_bytes_tiles(v.bt16, 12, v.h.n_tri));
    L.inc_off = inc_off;
    L.inc = inc;
    L.um = v.um;
    L.um_tiles = v.bt16;
    L.local_ids = local_ids ? 1 : 0;
    int rc = launch_rasterize_bwd(L, (hipStream_t)stream);
    if (rc) return fail(rc, "batch rasterize backward launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

size_t drp_batch_antialias_workspace_size(const void *plan) {
    if (!plan) return 0;
    BatchPlanHeader h;
    memcpy(&h, plan, sizeof(h));
    return align256((size_t)h.n_vert * 16) + 2 * align256((size_t)h.n_tri) + 2 * align256((size_t)h.n_pix * 8);
}

static int batch_aa_common(AALaunch &L, const PlanView &v, const float *color, int C, const float *rast,
                           const float *pos, const int32_t *tri, const int32_t *opp, void *workspace,
                           size_t workspace_bytes, const void *plan) {
    if (C < 1) return fail(-1, "color channel count must be >= 1");
    if (!color || !rast || (v.h.n_vert > 0 && !pos) || (v.h.n_tri > 0 && (!tri || !opp)))
        return fail(-1, "null pointer argument");
    if (!aligned16(pos) || !aligned16(rast)) return fail(-1, "pos/rast must be 16-byte aligned");
    if (C % 4 == 0 && !aligned16(color)) return fail(-1, "color must be 16-byte aligned when C %% 4 == 0");
    size_t need = drp_batch_antialias_workspace_size(plan);
    if (!workspace || workspace_bytes < need)
        return fail(-1, "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    char *b = (char *)workspace;
    L.color = color;
    L.rast = rast;
    L.pos = pos;
    L.tri = tri;
    L.opp = opp;
    L.n_img = v.h.n_img;
    L.H = 1;
    L.W = 1;
    L.C = C;
    L.V = (int)v.h.n_vert;
    L.F = (int)v.h.n_tri;
    L.pos_batched = 0;
    L.cv = (C == 1 || C == 3 || C == 4 || C == 8 || C == 16) ? C : 0;
    size_t o_orient = align256((size_t)v.h.n_vert * 16);
    size_t o_silh = o_orient + align256((size_t)v.h.n_tri), o_w = o_silh + align256((size_t)v.h.n_tri);
    size_t o_dec = o_w + align256((size_t)v.h.n_pix * 8);
    L.proj = (double2 *)b;
    L.orient = (signed char *)(b + o_orient);
    L.silh = (unsigned char *)(b + o_silh);
    L.weights = (float2 *)(b + o_w);
    L.decisions = v.h.n_tri < kMaxDecisionTris ? (int2 *)(b + o_dec) : nullptr;
    L.um = v.um;
    L.total_pix = v.h.n_pix;
    L.total_blk = (int)v.blk;
    L.total_t32 = (int)v.t32;
    return 0;
}

int drp_batch_antialias_fwd(void *stream, const void *plan, const void *plan_dev, const float *color, int C,
                            const float *rast, const float *pos, const int32_t *tri, const int32_t *opp, float *out,
                            void *workspace, size_t workspace_bytes) {
    PlanView v;
    if (int e = plan_view(plan, plan_dev, v)) return e;
    if (v.h.n_pix == 0) return 0;
    AALaunch L{};
    if (int e = batch_aa_common(L, v, color, C, rast, pos, tri, opp, workspace, workspace_bytes, plan)) return e;
    if (!out) return fail(-1, "null output");
    if (C % 4 == 0 && !aligned16(out)) return fail(-1, "out must be 16-byte aligned when C %% 4 == 0");
    L.out = out;
    int rc = launch_antialias(L, (hipStream_t)stream);
    if (rc) return fail(rc, "batch antialias launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

int drp_batch_antialias_bwd(void *stream, const void *plan, const void *plan_dev, const float *color, int C,
                            const float *rast, const float *pos, const int32_t *tri, const int32_t *opp,
                            const float *grad_out, float *grad_color, float *grad_pos, float pos_gradient_boost,
                            int reuse_forward, void *workspace, size_t workspace_bytes) {
    PlanView v;
    if (int e = plan_view(plan, plan_dev, v)) return e;
    if (v.h.n_pix == 0) return 0;
    AALaunch L{};
    if (int e = batch_aa_common(L, v, color, C, rast, pos, tri, opp, workspace, workspace_bytes, plan)) return e;
    if (!grad_out || !grad_color) return fail(-1, "null gradient buffer");
    if (C % 4 == 0 && (!aligned16(grad_out) || !aligned16(grad_color)))
        return fail(-1, "grad_out/grad_color must be 16-byte aligned when C %% 4 == 0");
    L.grad_out = grad_out;
    L.grad_color = grad_color;
    L.grad_pos = grad_pos;
    L.boost = pos_gradient_boost;
    L.reuse = reuse_forward && L.decisions != nullptr;
    int rc = launch_antialias_bwd(L, (hipStream_t)stream);
    if (rc) return fail(rc, "batch antialias backward launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

size_t drp_rasterize_bwd_workspace_size(int n_img, int H, int W, int F) {
    if (n_img <= 0 || F <= 0 || H <= 0 || W <= 0) return 0;
    return align256((size_t)n_img * (size_t)F * 3 * 16) +
           tile_records_bytes(n_img, H, W, 12, (long long)n_img * F) +
           align256((size_t)F * 3 * 4);
}

int drp_rasterize_bwd(void *stream, const float *pos, int n_img, int V, int pos_batched, const int32_t *tri, int F,
                      const int32_t *ranges_dev, int H, int W, const float *rast, const float *grad_rast,
                      const float *grad_db, float *grad_pos, const int32_t *boxes, const int32_t *inc_off,
                      const int32_t *inc, void *workspace, size_t workspace_bytes) {
    if (int e = check_dims(n_img, H, W)) return e;
    if (n_img == 0) return 0;
    if ((V > 0 && !pos) || (F > 0 && !tri) || !rast || !grad_rast || (V > 0 && !grad_pos))
        return fail(-1, "null pointer argument");
    if (F == 0 || V == 0) return 0;  // no triangle: nothing reaches grad_pos
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
        L.corner_slot = (int *)((char *)L.records + tile_records_bytes(n_img, H, W, 12, (long long)n_img * F));
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
    if ((V > 0 && !attr) || !rast || (F > 0 && !tri) || !out) return fail(-1, "null pointer argument");
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
    return interp_partial_bytes(n_img, F, C, attr_batched, views_per_group) +
           tile_records_bytes(n_img, H, W, 3 * C, (long long)n_img * F);
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
    if ((V > 0 && (!attr || !grad_attr)) || !rast || (F > 0 && !tri) || !grad_out || !grad_rast)
        return fail(-1, "null pointer argument");
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
    bool tri_path = F > 0 && boxes && inc_off && inc && n_diff == 0 && L.cv > 0;
    if (tri_path) {
        size_t need = drp_interpolate_bwd_workspace_size(n_img, H, W, F, C, attr_batched, L.views_per_group);
        if (!workspace || workspace_bytes < need)
            return fail(-1, "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
        if (C % 4 == 0 && (!aligned16(grad_out) || (V > 0 && !aligned16(attr))))
            return fail(-1, "grad_out and attr must be 16-byte aligned when C %% 4 == 0");
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

// interpolate backward and the backward of the rasterize call that produced
// rast, from one pass over the pixels (tile records carry both sums)
size_t drp_interpolate_rasterize_bwd_workspace_size(int n_img, int H, int W, int F, int C, int views_per_group) {
    if (n_img <= 0 || F <= 0 || C <= 0 || H <= 0 || W <= 0) return 0;
    const int vpg = views_per_block(views_per_group);
    const size_t groups = (size_t)((n_img + vpg - 1) / vpg);
    return align256(groups * (size_t)F * 3 * (size_t)C * 4) + align256((size_t)n_img * (size_t)F * 3 * 16) +
           tile_records_bytes(n_img, H, W, 3 * C + 12, (long long)n_img * F) + align256((size_t)F * 3 * 4);
}

int drp_interpolate_rasterize_bwd(void *stream, const float *attr, int n_img, int V, int C, const float *rast, int H,
                                  int W, const int32_t *tri, int F, const float *pos, int pos_batched,
                                  const float *grad_out, float *grad_attr, float *grad_rast, float *grad_pos,
                                  const int32_t *boxes, int views_per_group, const int32_t *inc_off,
                                  const int32_t *inc, void *workspace, size_t workspace_bytes) {
    if (int e = check_dims(n_img, H, W)) return e;
    if (n_img == 0) return 0;
    if (!(C == 4 || C == 8)) return fail(-1, "the fused backward takes C in {4, 8} (got %d)", C);
    if (F <= 0 || V <= 0) return fail(-1, "the fused backward needs triangles and vertices");
    if (!attr || !rast || !tri || !pos || !grad_out || !grad_attr || !grad_pos || !boxes || !inc_off || !inc)
        return fail(-1, "null pointer argument");
    if (!aligned16(rast) || (grad_rast && !aligned16(grad_rast)) || !aligned16(pos) || !aligned16(grad_pos) ||
        !aligned16(grad_out) || !aligned16(attr))
        return fail(-1, "rast, grad_rast, pos, grad_pos, grad_out and attr must be 16-byte aligned");
    const size_t need = drp_interpolate_rasterize_bwd_workspace_size(n_img, H, W, F, C, views_per_group);
    if (!workspace || workspace_bytes < need)
        return fail(-1, "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    const int vpg = views_per_block(views_per_group);
    const int groups = (n_img + vpg - 1) / vpg;
    char *b = (char *)workspace;
    float *partial_attr = (float *)b;
    b += align256((size_t)groups * (size_t)F * 3 * (size_t)C * 4);
    float *partial_pos = (float *)b;
    b += align256((size_t)n_img * (size_t)F * 3 * 16);
    void *records = b;
    b += tile_records_bytes(n_img, H, W, 3 * C + 12, (long long)n_img * F);
    int *slot = (int *)b;
    hipStream_t st = (hipStream_t)stream;
    if (int rc = launch_corner_slots(inc, inc_off, V, F, slot, st))
        return fail(rc, "corner slots launch failed: %s", hipGetErrorString(hipGetLastError()));
    TriBwdLaunch T{};
    T.b.n_img = n_img;
    T.b.V = V;
    T.b.F = F;
    T.b.H = H;
    T.b.W = W;
    T.b.pos_batched = pos_batched;
    T.pos = pos;
    T.tri = tri;
    T.rast = rast;
    T.grad_out = grad_out;
    T.attr = attr;
    T.V = V;
    T.attr_batched = 0;
    T.grad_rast_out = grad_rast;
    T.boxes = (const ItemBox *)boxes;
    T.partial = partial_attr;
    T.partial_pos = partial_pos;
    T.corner_slot = slot;
    T.n_img = n_img;
    T.F = F;
    T.H = H;
    T.W = W;
    T.views_per_block = vpg;
    tile_records_layout(T, records, n_img, H, W, (long long)n_img * F);
    int rc = launch_tri_bwd_fused(T, C, st);
    Fork f{};  // the two vertex gathers are independent
    const bool par = !rc && fork_begin(st, f);
    if (!rc) rc = launch_gather_attr(partial_attr, groups, F, V, C, inc_off, inc, grad_attr, par ? f.side : st);
    if (!rc) rc = launch_gather_pos(partial_pos, n_img, F, V, pos_batched, inc_off, grad_pos, st);
    if (par && fork_end(st, f)) rc = -2;
    if (rc) return fail(rc, "fused backward launch failed: %s", hipGetErrorString(hipGetLastError()));
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

int drp_count_covered(void *stream, const float *rast, int n_img, int H, int W, int32_t *counts) {
    if (int e = check_dims(n_img, H, W)) return e;
    if (n_img > 65535) return fail(-1, "at most 65535 images per call (got %d)", n_img);
    if (n_img == 0) return 0;
    if (!rast || !counts) return fail(-1, "null pointer argument");
    if (!aligned16(rast)) return fail(-1, "rast must be 16-byte aligned");
    int rc = launch_count_covered(rast, n_img, (long long)H * W, counts, (hipStream_t)stream);
    if (rc) return fail(rc, "count launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

static int tex_check(int D, int H, int W, int C, long long n, int wrap, int mode, bool cyclic_ok) {
    if (D < 1 || H < 1 || W < 1 || C < 1) return fail(-1, "texture dimensions must be >= 1");
    if ((long long)D * H * W * C >= (1ll << 40)) return fail(-1, "texture too large");
    if (n < 0) return fail(-1, "negative sample count");
    if (wrap < 0 || wrap > 3 || (wrap == 2 && !cyclic_ok))
        return fail(-1, "wrap must be border (0), reflection (1), %slatlong (3) (got %d)", cyclic_ok ? "cyclic (2), " : "",
                    wrap);
    if (mode < 0 || mode > (cyclic_ok ? 2 : 1))  // the 2-D entry points also take bicubic
        return fail(-1, "mode must be bilinear (0), nearest (1)%s (got %d)", cyclic_ok ? " or bicubic (2)" : "", mode);
    return 0;
}

int drp_sample2d_fwd(void *stream, const float *tex, int H, int W, int C, const float *uv, long long n, int wrap,
                     int mode, float *out) {
    if (int e = tex_check(1, H, W, C, n, wrap, mode, true)) return e;
    if (n == 0) return 0;
    if (!tex || !uv || !out) return fail(-1, "null pointer argument");
    int rc = launch_sample2d(tex, H, W, C, uv, n, wrap, mode, out, (hipStream_t)stream);
    if (rc) return fail(rc, "sample2d launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

int drp_sample2d_bwd(void *stream, const float *tex, int H, int W, int C, const float *uv, long long n, int wrap,
                     int mode, const float *grad_out, float *grad_tex, float *grad_uv) {
    if (int e = tex_check(1, H, W, C, n, wrap, mode, true)) return e;
    if (n == 0) return 0;
    if (!tex || !uv || !grad_out) return fail(-1, "null pointer argument");
    int rc = launch_sample2d_bwd(tex, H, W, C, uv, n, wrap, mode, grad_out, grad_tex, grad_uv, (hipStream_t)stream);
    if (rc) return fail(rc, "sample2d backward launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

int drp_sample3d_fwd(void *stream, const float *tex, int D, int H, int W, int C, const float *uvw, long long n,
                     int wrap, int mode, float *out) {
    if (int e = tex_check(D, H, W, C, n, wrap, mode, false)) return e;
    if (n == 0) return 0;
    if (!tex || !uvw || !out) return fail(-1, "null pointer argument");
    int rc = launch_sample3d(tex, D, H, W, C, uvw, n, wrap, mode, out, (hipStream_t)stream);
    if (rc) return fail(rc, "sample3d launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

int drp_sample3d_bwd(void *stream, const float *tex, int D, int H, int W, int C, const float *uvw, long long n,
                     int wrap, int mode, const float *grad_out, float *grad_tex, float *grad_uvw) {
    if (int e = tex_check(D, H, W, C, n, wrap, mode, false)) return e;
    if (n == 0) return 0;
    if (!tex || !uvw || !grad_out) return fail(-1, "null pointer argument");
    int rc = launch_sample3d_bwd(tex, D, H, W, C, uvw, n, wrap, mode, grad_out, grad_tex, grad_uvw,
                                 (hipStream_t)stream);
    if (rc) return fail(rc, "sample3d backward launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

size_t drp_stencil_partition_workspace_size(int H, int W, int n_owner) {
    if (H <= 0 || W <= 0 || n_owner < 0) return 0;
    return stencil_partition_ws_bytes((long long)H * W, n_owner);
}

int drp_stencil_partition(void *stream, const float *rast, int H, int W, const int32_t *stencils, int F, int n_owner,
                          int32_t *order, int32_t *starts, int32_t *ends, void *workspace, size_t workspace_bytes) {
    if (int e = check_dims(1, H, W)) return e;
    if (F < 0 || n_owner < 0 || n_owner >= (1 << 30)) return fail(-1, "bad F or owner count");
    if (!rast || !stencils || !order || !starts || !ends) return fail(-1, "null pointer argument");
    if (!aligned16(rast)) return fail(-1, "rast must be 16-byte aligned");
    const size_t need = stencil_partition_ws_bytes((long long)H * W, n_owner);
    if (!workspace || workspace_bytes < need)
        return fail(-1, "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    int rc = launch_stencil_partition(rast, (long long)H * W, stencils, F, n_owner, order, starts, ends, workspace,
                                      (hipStream_t)stream);
    if (rc) return fail(rc, "stencil partition launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

size_t drp_incidence_workspace_size(int V, int F) { return V > 0 && F >= 0 ? incidence_ws_bytes(V, F) : 0; }

int drp_incidence_build(void *stream, const int32_t *tri, int F, int V, int32_t *inc_off, int32_t *inc,
                        void *workspace, size_t workspace_bytes) {
    if (F < 0 || V < 0) return fail(-1, "negative F or V");
    if (F > 0x2aaaaaaa) return fail(-1, "too many triangles");
    if (V == 0) return 0;
    if ((F > 0 && !tri) || !inc_off || (F > 0 && !inc)) return fail(-1, "null pointer argument");
    const size_t need = incidence_ws_bytes(V, F);
    if (!workspace || workspace_bytes < need)
        return fail(-1, "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
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

// the antialias workspace: per-vertex pixel positions, per-triangle
// orientation and silhouette flags, per-pixel pair weights and decisions
static void aa_layout(AALaunch &L, void *workspace, int n_img, int H, int W, int V, int F, int pos_batched) {
    const size_t n_proj = pos_batched ? (size_t)n_img : 1;
    const size_t npx = (size_t)n_img * H * W;
    const size_t o_orient = align256(n_proj * (size_t)V * 16);
    const size_t o_silh = o_orient + align256(n_proj * (size_t)F), o_w = o_silh + align256(n_proj * (size_t)F);
    const size_t o_dec = o_w + align256(npx * 8);
    char *b = (char *)workspace;
    L.proj = (double2 *)b;
    L.orient = (signed char *)(b + o_orient);
    L.silh = (unsigned char *)(b + o_silh);
    L.weights = (float2 *)(b + o_w);
    // pair decisions pack the triangle into 27 bits; larger meshes recompute them in the backward
    L.decisions = F < kMaxDecisionTris ? (int2 *)(b + o_dec) : nullptr;
}

static int aa_common(AALaunch &L, void *stream, const float *color, int n_img, int H, int W, int C,
                     const float *rast, const float *pos, int V, int pos_batched, const int32_t *tri, int F,
                     const int32_t *opp, void *workspace, size_t workspace_bytes) {
    if (int e = check_dims(n_img, H, W)) return e;
    if (n_img > 65535) return fail(-1, "antialias takes at most 65535 images per call (got %d)", n_img);
    if (C < 1) return fail(-1, "color channel count must be >= 1");
    if (!color || !rast || (V > 0 && !pos) || (F > 0 && (!tri || !opp))) return fail(-1, "null pointer argument");
    if (!aligned16(pos) || !aligned16(rast)) return fail(-1, "pos/rast must be 16-byte aligned");
    if (C % 4 == 0 && !aligned16(color)) return fail(-1, "color must be 16-byte aligned when C %% 4 == 0");
    size_t need = drp_antialias_workspace_size(n_img, H, W, V, F, pos_batched);
    if (!workspace || workspace_bytes < need)
        return fail(-1, "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
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
    aa_layout(L, workspace, n_img, H, W, V, F, pos_batched);
    (void)stream;
    return 0;
}

static int aa_fwd(void *stream, const float *color, int n_img, int H, int W, int C, const float *rast,
                  const float *pos, int V, int pos_batched, const int32_t *tri, int F, const int32_t *opp, float *out,
                  void *workspace, size_t workspace_bytes, int prepared) {
    if (n_img == 0) return 0;
    AALaunch L{};
    if (int e = aa_common(L, stream, color, n_img, H, W, C, rast, pos, V, pos_batched, tri, F, opp, workspace,
                          workspace_bytes))
        return e;
    if (!out) return fail(-1, "null output");
    if (C % 4 == 0 && !aligned16(out)) return fail(-1, "out must be 16-byte aligned when C %% 4 == 0");
    L.out = out;
    L.prepared = prepared;
    int rc = launch_antialias(L, (hipStream_t)stream);
    if (rc) return fail(rc, "antialias launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

int drp_antialias_fwd(void *stream, const float *color, int n_img, int H, int W, int C, const float *rast,
                      const float *pos, int V, int pos_batched, const int32_t *tri, int F, const int32_t *opp,
                      float *out, void *workspace, size_t workspace_bytes) {
    return aa_fwd(stream, color, n_img, H, W, C, rast, pos, V, pos_batched, tri, F, opp, out, workspace,
                  workspace_bytes, 0);
}

int drp_antialias_prepare(void *stream, int n_img, int H, int W, const float *pos, int V, int pos_batched,
                          const int32_t *tri, int F, const int32_t *opp, void *workspace, size_t workspace_bytes) {
    if (int e = check_dims(n_img, H, W)) return e;
    if (n_img == 0) return 0;
    if (n_img > 65535) return fail(-1, "antialias takes at most 65535 images per call (got %d)", n_img);
    if ((V > 0 && !pos) || (F > 0 && (!tri || !opp))) return fail(-1, "null pointer argument");
    if (!aligned16(pos)) return fail(-1, "pos must be 16-byte aligned");
    if (pos_batched && (long long)n_img * V >= (1ll << 31))
        return fail(-1, "antialias with per-image positions takes fewer than 2^31 vertices in all");
    const size_t need = drp_antialias_workspace_size(n_img, H, W, V, F, pos_batched);
    if (!workspace || workspace_bytes < need)
        return fail(-1, "workspace too small: need %zu bytes, got %zu", need, workspace_bytes);
    AALaunch L{};
    L.pos = pos;
    L.tri = tri;
    L.opp = opp;
    L.n_img = n_img;
    L.H = H;
    L.W = W;
    L.V = V;
    L.F = F;
    L.pos_batched = pos_batched;
    aa_layout(L, workspace, n_img, H, W, V, F, pos_batched);
    int rc = launch_antialias_prepare(L, (hipStream_t)stream);
    if (rc) return fail(rc, "antialias prepare launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

int drp_antialias_fwd_prepared(void *stream, const float *color, int n_img, int H, int W, int C, const float *rast,
                               const float *pos, int V, int pos_batched, const int32_t *tri, int F,
                               const int32_t *opp, float *out, void *workspace, size_t workspace_bytes) {
    return aa_fwd(stream, color, n_img, H, W, C, rast, pos, V, pos_batched, tri, F, opp, out, workspace,
                  workspace_bytes, 1);
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

// ---- clip-space transform and the L2 image loss ----------------------------
int drp_clip_transform_fwd(void *stream, const float *xyz, int V, const float *VP, int n, float *pos) {
    if (V < 0 || n < 0) return fail(-1, "negative V (%d) or view count (%d)", V, n);
    if (V == 0 || n == 0) return 0;
    if (!xyz || !VP || !pos) return fail(-1, "null pointer argument");
    if (!aligned16(pos)) return fail(-1, "pos must be 16-byte aligned");
    int rc = launch_clip_fwd(xyz, V, VP, n, pos, (hipStream_t)stream);
    if (rc) return fail(rc, "clip transform launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

int drp_clip_transform_bwd(void *stream, const float *grad_pos, int V, const float *VP, int n, float *grad_xyz) {
    if (V < 0 || n < 0) return fail(-1, "negative V (%d) or view count (%d)", V, n);
    if (V == 0) return 0;
    if (!grad_xyz || (n > 0 && (!grad_pos || !VP))) return fail(-1, "null pointer argument");
    if (n > 0 && !aligned16(grad_pos)) return fail(-1, "grad_pos must be 16-byte aligned");
    int rc = launch_clip_bwd(grad_pos, V, VP, n, grad_xyz, (hipStream_t)stream);
    if (rc) return fail(rc, "clip transform backward launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

size_t drp_l2_loss_workspace_size(void) { return l2_loss_ws_bytes(); }

int drp_l2_loss_fwd(void *stream, const float *a, const float *b, long long n, float *loss, void *workspace,
                    size_t workspace_bytes) {
    if (n < 0) return fail(-1, "negative element count");
    if (!loss || (n > 0 && (!a || !b))) return fail(-1, "null pointer argument");
    if (!workspace || workspace_bytes < l2_loss_ws_bytes())
        return fail(-1, "workspace too small: need %zu bytes, got %zu", l2_loss_ws_bytes(), workspace_bytes);
    int rc = launch_l2_loss(a, b, n, loss, nullptr, workspace, (hipStream_t)stream);
    if (rc) return fail(rc, "l2 loss launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

int drp_l2_loss_fwd_grad(void *stream, const float *a, const float *b, long long n, float *loss, float *grad_a,
                         void *workspace, size_t workspace_bytes) {
    if (n < 0) return fail(-1, "negative element count");
    if (!loss || (n > 0 && (!a || !b || !grad_a))) return fail(-1, "null pointer argument");
    if (!workspace || workspace_bytes < l2_loss_ws_bytes())
        return fail(-1, "workspace too small: need %zu bytes, got %zu", l2_loss_ws_bytes(), workspace_bytes);
    int rc = launch_l2_loss(a, b, n, loss, grad_a, workspace, (hipStream_t)stream);
    if (rc) return fail(rc, "l2 loss launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

int drp_l2_loss_scale(void *stream, float *grad, long long n, const float *grad_loss) {
    if (n < 0) return fail(-1, "negative element count");
    if (n == 0) return 0;
    if (!grad || !grad_loss) return fail(-1, "null pointer argument");
    int rc = launch_l2_loss_scale(grad, n, grad_loss, (hipStream_t)stream);
    if (rc) return fail(rc, "l2 loss scale launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

int drp_l2_loss_bwd(void *stream, const float *a, const float *b, long long n, const float *grad_loss, float *grad_a,
                    float *grad_b) {
    if (n < 0) return fail(-1, "negative element count");
    if (n == 0) return 0;
    if (!a || !b || !grad_loss || (!grad_a && !grad_b)) return fail(-1, "null pointer argument");
    int rc = launch_l2_loss_bwd(a, b, n, grad_loss, grad_a, grad_b, (hipStream_t)stream);
    if (rc) return fail(rc, "l2 loss backward launch failed: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

}  // extern "C"
