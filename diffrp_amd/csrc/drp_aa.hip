// Analytic silhouette antialiasing (after Laine et al. 2020, "Modular Primitives
// for High-Performance Differentiable Rendering", sec. 3.3) and the edge ->
// opposite-triangle table it needs.
//
// Per horizontal / vertical pair of pixel centres (a, b) with different
// triangle ids: candidates are the triangles at a and b, nearer z/w first.
// For candidate T seen at `from`, the segment from -> to leaves T through the
// edge with the smallest crossing t_e = f(from) / (f(from) - f(to)) among the
// edges whose half-plane excludes `to`.  If that edge is a silhouette (no
// neighbour, or a neighbour of opposite screen orientation) the pair blends:
// t = clamp(t_e, 0, 1); t < 1/2 moves `from` towards `to` by (1/2 - t),
// otherwise `to` towards `from` by (t - 1/2).  Else the next candidate is tried.
//
// Kernels: k_aa_orient (per-triangle screen orientation), k_aa_silhouette (triangles with a possible
// silhouette edge), k_aa_pairs (one thread per pixel evaluates the pairs to its
// right and above; signed weight per pair), k_aa_gather (output colours in a
// fixed order), and the backward counterparts.
#include "drp_common.h"
#include "drp_kernels.h"

namespace drp {

// ----------------------------------------------------------------------------
// topology: open-addressing hash of undirected edges -> (up to) two faces
// ----------------------------------------------------------------------------
constexpr unsigned long long kEmpty = ~0ull;

__device__ __forceinline__ unsigned long long edge_key(int a, int b) {
    unsigned lo = (unsigned)min(a, b), hi = (unsigned)max(a, b);
    return ((unsigned long long)lo << 32) | hi;
}
__device__ __forceinline__ unsigned hash64(unsigned long long k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return (unsigned)k;
}

__global__ void k_topo_clear(unsigned long long *keys, int *cnt, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { keys[i] = kEmpty; cnt[i] = 0; }
}

__global__ void k_topo_insert(const int *__restrict__ tri, int F, unsigned long long *keys, int *cnt, int2 *slots,
                              unsigned mask) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= 3ll * F) return;
    int t = (int)(g / 3), e = (int)(g % 3);
    int a = tri[3 * (long long)t + e], b = tri[3 * (long long)t + (e == 2 ? 0 : e + 1)];
    unsigned long long k = edge_key(a, b);
    unsigned h = hash64(k) & mask;
    for (unsigned probe = 0; probe <= mask; ++probe) {
        unsigned long long old = atomicCAS(&keys[h], kEmpty, k);
        if (old == kEmpty || old == k) {
            int s = atomicAdd(&cnt[h], 1);
            if (s == 0) slots[h].x = (int)g;
            else if (s == 1) slots[h].y = (int)g;
            return;
        }
        h = (h + 1) & mask;
    }
}

__global__ void k_topo_lookup(const int *__restrict__ tri, int F, const unsigned long long *__restrict__ keys,
                              const int *__restrict__ cnt, const int2 *__restrict__ slots, unsigned mask,
                              int *__restrict__ opp) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= 3ll * F) return;
    int t = (int)(g / 3), e = (int)(g % 3);
    int a = tri[3 * (long long)t + e], b = tri[3 * (long long)t + (e == 2 ? 0 : e + 1)];
    unsigned long long k = edge_key(a, b);
    unsigned h = hash64(k) & mask;
    int res = -1;
    for (unsigned probe = 0; probe <= mask; ++probe) {
        unsigned long long kk = keys[h];
        if (kk == k) {
            if (cnt[h] == 2) {
                int2 s = slots[h];
                res = (s.x == (int)g ? s.y : s.x) / 3;
            }
            break;
        }
        if (kk == kEmpty) break;
        h = (h + 1) & mask;
    }
    opp[g] = res;
}

int launch_topology(const TopoLaunch &L, hipStream_t st) {
    if (L.F == 0) return 0;
    long long n = (long long)L.mask + 1;
    hipLaunchKernelGGL(k_topo_clear, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, L.keys, L.cnt, n);
    long long ne = 3ll * L.F;
    dim3 g((unsigned)((ne + 255) / 256));
    hipLaunchKernelGGL(k_topo_insert, g, dim3(256), 0, st, L.tri, L.F, L.keys, L.cnt, L.slots, L.mask);
    hipLaunchKernelGGL(k_topo_lookup, g, dim3(256), 0, st, L.tri, L.F, L.keys, L.cnt, L.slots, L.mask, L.opp);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ----------------------------------------------------------------------------
// orientation / silhouette precompute
// ----------------------------------------------------------------------------
__device__ __forceinline__ double dfma(double a, double b, double c) { return __builtin_fma(a, b, c); }

__device__ __forceinline__ int orient3(float4 a, float4 b, float4 c) {
    // sign of det[[x0,x1,x2],[y0,y1,y2],[w0,w1,w2]], in double (edge-on triangles decide silhouettes)
    double m0 = dfma(b.y, c.w, -((double)c.y * b.w));
    double m1 = dfma(a.y, c.w, -((double)c.y * a.w));
    double m2 = dfma(a.y, b.w, -((double)b.y * a.w));
    double d = dfma(a.x, m0, dfma(-(double)b.x, m1, (double)c.x * m2));
    return d > 0.0 ? 1 : (d < 0.0 ? -1 : 0);
}

// pixel coordinates of every vertex, in double (NaN when w <= 0): the pair
// search reads them for every candidate triangle
__global__ void k_aa_project(const float *__restrict__ pos, int V, int n_proj, int H, int W, UnevenMap um,
                             double2 *__restrict__ proj) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)n_proj * V) return;
    if (um.n) {  // uneven batch: the vertex's item sets the image size
        const int item = um_find(um.vert, um.n, (int)g);
        H = um.hw[2 * item];
        W = um.hw[2 * item + 1];
    }
    float4 p = reinterpret_cast<const float4 *>(pos)[g];
    double2 o;
    if (p.w > 0.f) {
        o.x = ((double)p.x / (double)p.w + 1.0) * ((double)W * 0.5);
        o.y = ((double)p.y / (double)p.w + 1.0) * ((double)H * 0.5);
    } else {
        o.x = __longlong_as_double(0x7ff8000000000000ll);
        o.y = o.x;
    }
    proj[g] = o;
}

__global__ void k_aa_orient(const float *__restrict__ pos, int V, const int *__restrict__ tri, int F, int n_proj,
                            signed char *__restrict__ orient) {
    // XCD k takes the k-th eighth of the (image, triangle) items: an image's vertices stay in one L2
    long long g = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (g >= (long long)n_proj * F) return;
    int img = (int)(g / F), t = (int)(g % F);
    const int *t3 = tri + 3 * (long long)t;
    int i0 = t3[0], i1 = t3[1], i2 = t3[2];
    signed char o = 0;
    if ((unsigned)i0 < (unsigned)V && (unsigned)i1 < (unsigned)V && (unsigned)i2 < (unsigned)V) {
        const float4 *P = reinterpret_cast<const float4 *>(pos) + (long long)img * V;
        o = (signed char)orient3(P[i0], P[i1], P[i2]);
    }
    orient[g] = o;
}

// silh[img][t]: bit e (e = 0, 1, 2) when edge e -> e+1 of t can be a silhouette
// in that image (no neighbour, or a neighbour of another screen orientation),
// bit 3 when t's screen orientation is negative; 0 for an edge-on triangle.
// A pair whose candidate triangle has no such edge cannot blend through it, and
// the pair search reads the edge test and the orientation from this one byte.
__global__ void k_aa_silhouette(const int *__restrict__ opp, const signed char *__restrict__ orient, int F,
                                int n_proj, unsigned char *__restrict__ silh) {
    long long g = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;  // as k_aa_orient
    if (g >= (long long)n_proj * F) return;
    int img = (int)(g / F), t = (int)(g % F);
    const signed char *O = orient + (long long)img * F;
    int o = O[t];
    unsigned char f = 0;
    if (o != 0) {
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            int n = opp[3 * (long long)t + e];
            if (!(n >= 0 && n < F && O[n] == o)) f |= (unsigned char)(1 << e);
        }
        if (f && o < 0) f |= 8;
    }
    silh[g] = f;
}

// ----------------------------------------------------------------------------
// pair evaluation (same arithmetic as the scalar reference of these rules)
// ----------------------------------------------------------------------------
struct PairRes {
    int recv;   // -1 none, 0: a receives, 1: b receives
    float wgt;
    int tri, edge, from_a, clamped;
};

struct AAImg {
    const float4 *rast;         // image rast
    const double2 *proj;        // image's vertices in pixel coordinates (NaN: w <= 0)
    const signed char *orient;  // image's triangle orientations
    const unsigned char *silh;  // image's silhouette-capable triangles
    int V, F, H, W;
};

__device__ __forceinline__ int rast_id_w(float w, int F) {  // triangle + 1 from rast's w, 0 for none
    int id = w > 0.f ? float_to_triidx(w) : 0;
    return id > F ? 0 : id;
}
__device__ __forceinline__ int rast_id(float4 p, int F) { return rast_id_w(p.w, F); }

// Pair (a, b) from the two pixels' rast values and their triangles' silhouette
// bytes (the caller loads them, so the loads of neighbouring pairs overlap).
// Both candidates' corner indices, then both candidates' vertex projections,
// are loaded before either is tested, so the pair waits for two rounds of
// loads, not up to four.
__device__ __forceinline__ PairRes eval_pair(const AAImg &I, const int *__restrict__ tri, float4 pa, float4 pb,
                                             int sa, int sb, int ra, int ca, int rb, int cb) {
    PairRes res;
    res.recv = -1;
    const int ida = rast_id(pa, I.F), idb = rast_id(pb, I.F);
    if (ida == idb) return res;
    int a_first = 1;
    if (ida == 0) a_first = 0;
    else if (idb != 0 && pb.z < pa.z) a_first = 0;
    int tc[2], mc[2], vi[2][3];
    bool ok[2];
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {  // candidate 0 first, then 1
        const int use_a = pass == 0 ? a_first : !a_first;
        const int id = use_a ? ida : idb;
        mc[pass] = use_a ? sa : sb;
        ok[pass] = id != 0 && (mc[pass] & 7);  // a candidate needs a silhouette-capable edge
        tc[pass] = ok[pass] ? id - 1 : 0;
        const int *t3 = tri + 3 * (long long)tc[pass];
        vi[pass][0] = t3[0];
        vi[pass][1] = t3[1];
        vi[pass][2] = t3[2];
    }
    // pair geometry in double: a float clip coordinate fixes a vertex to
    // ~W * 2^-25 pixel, and the crossing needs differences of nearby pixel
    // positions
    double2 q[2][3];
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
        ok[pass] = ok[pass] && (unsigned)vi[pass][0] < (unsigned)I.V && (unsigned)vi[pass][1] < (unsigned)I.V &&
                   (unsigned)vi[pass][2] < (unsigned)I.V;
#pragma unroll
        for (int k = 0; k < 3; ++k) q[pass][k] = ok[pass] ? I.proj[vi[pass][k]] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
        if (!ok[pass]) continue;
        const int use_a = pass == 0 ? a_first : !a_first;
        const int t = tc[pass], m = mc[pass];
        double X[3], Y[3];
        bool front = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            front = front && !isnan(q[pass][k].x);
            X[k] = q[pass][k].x;
            Y[k] = q[pass][k].y;
        }
        if (!front) continue;
        const int o = (m & 8) ? -1 : 1;
        double fx = (double)(use_a ? ca : cb) + 0.5, fy = (double)(use_a ? ra : rb) + 0.5;
        double tx = (double)(use_a ? cb : ca) + 0.5, ty = (double)(use_a ? rb : ra) + 0.5;
        int best = -1;
        double bt = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            int j = k == 2 ? 0 : k + 1;
            double ex = X[j] - X[k], ey = Y[j] - Y[k];
            double ff = dfma(ex, fy - Y[k], -(ey * (fx - X[k])));
            double ft = dfma(ex, ty - Y[k], -(ey * (tx - X[k])));
            if (o < 0) { ff = -ff; ft = -ft; }
            if (!(ft < 0.0)) continue;
            double te = ff / (ff - ft);
            if (best < 0 || te < bt) { best = k; bt = te; }
        }
        if (best < 0 || !((m >> best) & 1)) continue;  // no crossing, or the edge has a same-facing neighbour
        int clamped = 0;
        if (bt < 0.0) { bt = 0.0; clamped = 1; }
        if (bt > 1.0) { bt = 1.0; clamped = 1; }
        res.tri = t;
        res.edge = best;
        res.from_a = use_a;
        res.clamped = clamped;
        if (bt < 0.5) { res.recv = use_a ? 0 : 1; res.wgt = (float)(0.5 - bt); }
        else { res.recv = use_a ? 1 : 0; res.wgt = (float)(bt - 0.5); }
        return res;
    }
    return res;
}

// signed pair weight: > 0 b receives from a, < 0 a receives from b
__device__ __forceinline__ float signed_w(const PairRes &p) {
    return p.recv < 0 ? 0.f : (p.recv == 1 ? p.wgt : -p.wgt);
}

__device__ __forceinline__ AAImg aa_img(const AALaunch &L, int img) {
    AAImg I;
    if (L.um.n) {  // uneven batch: the item's pixels; vertices and triangles are global
        I.rast = reinterpret_cast<const float4 *>(L.rast) + L.um.pix[img];
        I.proj = L.proj;
        I.orient = L.orient;
        I.silh = L.silh;
        I.V = L.V;
        I.F = L.F;
        I.H = L.um.hw[2 * img];
        I.W = L.um.hw[2 * img + 1];
        return I;
    }
    long long npx = (long long)L.H * L.W;
    I.rast = reinterpret_cast<const float4 *>(L.rast) + img * npx;
    int pimg = L.pos_batched ? img : 0;
    I.proj = L.proj + (long long)pimg * L.V;
    I.orient = L.orient + (long long)pimg * L.F;
    I.silh = L.silh + (long long)pimg * L.F;
    I.V = L.V;
    I.F = L.F;
    I.H = L.H;
    I.W = L.W;
    return I;
}

// Decision of one pair as the backward needs it: 0 = no blend, else
// (tri + 1) << 5 | recv << 4 | clamped << 3 | from_a << 2 | edge.
__device__ __forceinline__ int pack_decision(const PairRes &p) {
    if (p.recv < 0) return 0;
    return ((p.tri + 1) << 5) | (p.recv << 4) | (p.clamped << 3) | (p.from_a << 2) | p.edge;
}
__device__ __forceinline__ PairRes unpack_decision(int d) {
    PairRes p;
    p.recv = d == 0 ? -1 : ((d >> 4) & 1);
    p.tri = (int)((unsigned)d >> 5) - 1;
    p.clamped = (d >> 3) & 1;
    p.from_a = (d >> 2) & 1;
    p.edge = d & 3;
    p.wgt = 0.f;
    return p;
}

__device__ __forceinline__ int silh_of(const AAImg &I, int id) { return id ? I.silh[id - 1] : 0; }
// the same without a branch around the load (callers guarantee F > 0), so that
// several of them issue together
__device__ __forceinline__ int silh_at(const AAImg &I, int id) {
    const int s = I.silh[id > 0 ? id - 1 : 0];
    return id > 0 ? s : 0;
}

// weights[pix] = (signed weight of pair (pix, right), of pair (pix, above));
// decisions[pix] (when kept) the two pairs' decisions for the backward.
// Blocks of 64 x 8 pixels, two rows four apart per thread (grid: columns,
// rows, images), so that a thread's loads for both pixels are in flight
// together.  Only pairs whose two triangles differ and that have a
// silhouette-capable edge on either side can blend; the block queues those
// (in LDS) and evaluates the queue densely, so the double-precision crossing
// runs on full waves, then every pixel writes its two results.
constexpr int kAAPairRows = 8;                 // rows of a pair block (kAAPairRows / 4 pixels per thread)
constexpr int kAAPairPix = 64 * kAAPairRows;   // pixels of a pair block
constexpr int kAAPairPer = kAAPairPix / 256;   // pixels per thread
constexpr int kAAPairDir = 9;                  // queue entry: local pixel | direction << kAAPairDir
static_assert(kAAPairPix <= (1 << kAAPairDir), "the queue's direction bit must lie above every local pixel index");

// eight waves per SIMD (64 VGPRs, no spill; 66 left it at seven): the pair
// search waits on chains of dependent loads, so more waves hide more of them
__global__ void __launch_bounds__(256, 8) k_aa_pairs(AALaunch L, float2 *__restrict__ weights,
                                                  int2 *__restrict__ decisions) {
    __shared__ int q[2 * kAAPairPix];  // local pixel | direction << kAAPairDir (0: right, 1: above)
    __shared__ float2 sw[kAAPairPix];
    __shared__ int2 sd[kAAPairPix];
    __shared__ int q_n;
    const int tid = threadIdx.x, lane = tid & 63;
    // one row of blocks, padded to a multiple of 8; XCD k takes the k-th eighth
    // of them, whole images in a uniform batch (their vertices and flags stay
    // in its L2, and a block's upper neighbour is its successor there)
    const int lin = xcd_remap(blockIdx.x, gridDim.x);
    int bx, by, img;
    if (L.um.n) {  // uneven batch: every item's blocks in turn
        if (lin >= L.total_blk) return;  // uniform per block
        img = um_find(L.um.blk, L.um.n, lin);
        const int lb = lin - L.um.blk[img], nbx = (L.um.hw[2 * img + 1] + 63) / 64;
        bx = lb % nbx;
        by = lb / nbx;
    } else {
        const int gx = (L.W + 63) / 64, gy = (L.H + kAAPairRows - 1) / kAAPairRows;
        if (lin >= gx * gy * L.n_img) return;  // uniform per block
        bx = lin % gx;
        by = (lin / gx) % gy;
        img = lin / (gx * gy);
    }
    const AAImg I = aa_img(L, img);
    const int H = I.H, W = I.W;
    const int c = bx * 64 + lane;
    bool in[kAAPairPer], has_r[kAAPairPer], has_u[kAAPairPer];
    long long q0[kAAPairPer];
    float w0[kAAPairPer], wr[kAAPairPer], wu[kAAPairPer];
#pragma unroll
    for (int j = 0; j < kAAPairPer; ++j) {  // pixel j of this thread: local pixel tid + 256 j
        const int r = by * kAAPairRows + (tid >> 6) + 4 * j;
        in[j] = c < W && r < H && I.F > 0;
        q0[j] = (long long)r * W + c;
        has_r[j] = in[j] && c + 1 < W;
        has_u[j] = in[j] && r + 1 < H;
        w0[j] = wr[j] = wu[j] = 0.f;
        if (in[j]) {  // the neighbours' ids come from rast's w alone; all loads issue before the first wait
            w0[j] = I.rast[q0[j]].w;
            wr[j] = I.rast[has_r[j] ? q0[j] + 1 : q0[j]].w;
            wu[j] = I.rast[has_u[j] ? q0[j] + W : q0[j]].w;
        }
    }
    int cand[kAAPairPer];  // bit 0: pair (pix, right) may blend; bit 1: pair (pix, above)
    int nq = 0;
#pragma unroll
    for (int j = 0; j < kAAPairPer; ++j) {
        cand[j] = 0;
        if (in[j]) {
            const int i0 = rast_id_w(w0[j], I.F), ir = rast_id_w(wr[j], I.F), iu = rast_id_w(wu[j], I.F);
            const int s0 = silh_at(I, i0), sr = silh_at(I, ir), su = silh_at(I, iu);
            if (has_r[j] && ir != i0 && ((s0 | sr) & 7)) cand[j] |= 1;
            if (has_u[j] && iu != i0 && ((s0 | su) & 7)) cand[j] |= 2;
        }
        nq += (cand[j] & 1) + (cand[j] >> 1);
    }
#pragma unroll
    for (int j = 0; j < kAAPairPer; ++j) {
        sw[tid + 256 * j] = make_float2(0.f, 0.f);
        sd[tid + 256 * j] = make_int2(0, 0);
    }
    if (tid == 0) q_n = 0;
    __syncthreads();
    // wave-aggregated append
    int x = nq;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    const int wtot = __shfl(x, 63, 64);
    int base = 0;
    if (lane == 63 && wtot > 0) base = atomicAdd(&q_n, wtot);
    base = __shfl(base, 63, 64) + x - nq;
#pragma unroll
    for (int j = 0; j < kAAPairPer; ++j) {
        if (cand[j] & 1) q[base++] = tid + 256 * j;
        if (cand[j] & 2) q[base++] = (tid + 256 * j) | (1 << kAAPairDir);
    }
    __syncthreads();
    const int n = q_n;
    for (int e = tid; e < n; e += 256) {
        const int m = q[e], lp = m & ((1 << kAAPairDir) - 1), dir = m >> kAAPairDir;
        const int ra = by * kAAPairRows + (lp >> 6), ca = bx * 64 + (lp & 63);
        const int rb = dir ? ra + 1 : ra, cb = dir ? ca : ca + 1;
        const float4 pa = I.rast[(long long)ra * W + ca], pb = I.rast[(long long)rb * W + cb];
        const int ia = rast_id(pa, I.F), ib = rast_id(pb, I.F);
        const int sa = silh_at(I, ia), sb = silh_at(I, ib);  // queued pairs exist only when F > 0
        const PairRes p = eval_pair(I, L.tri, pa, pb, sa, sb, ra, ca, rb, cb);
        if (dir) {
            sw[lp].y = signed_w(p);
            sd[lp].y = pack_decision(p);
        } else {
            sw[lp].x = signed_w(p);
            sd[lp].x = pack_decision(p);
        }
    }
    __syncthreads();
    const long long pix0 = L.um.n ? L.um.pix[img] : (long long)img * L.H * L.W;
#pragma unroll
    for (int j = 0; j < kAAPairPer; ++j) {
        const int r = by * kAAPairRows + (tid >> 6) + 4 * j;
        if (c >= W || r >= H) continue;
        weights[pix0 + q0[j]] = sw[tid + 256 * j];
        if (decisions) decisions[pix0 + q0[j]] = sd[tid + 256 * j];
    }
}

// channel chunk handled by one thread: 4 (one float4) when C is a multiple of 4
template <int CV>
struct AAChunk {
    static constexpr int CH = (CV > 0 && CV % 4 == 0) ? 4 : 1;
};
// the backward gather takes a whole pixel for 8 channels (two float4 per
// thread: 0.438 -> 0.424 ms for the antialias backward at 64 x 512^2; the
// forward gather is faster with four channels per thread)
template <int CV>
struct AAChunkBwd {
    static constexpr int CH = CV == 8 ? 8 : AAChunk<CV>::CH;
};

template <int CH>
__device__ __forceinline__ void ld(const float *p, float (&x)[CH]) {
    if (CH % 4 == 0) {
#pragma unroll
        for (int k = 0; k < CH; k += 4) {
            const float4 v = *reinterpret_cast<const float4 *>(p + k);
            x[k] = v.x; x[k + 1] = v.y; x[k + 2] = v.z; x[k + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < CH; ++i) x[i] = p[i];
    }
}
// streaming (non-temporal) stores: the output is not read again by this pass
template <int CH>
__device__ __forceinline__ void st(float *p, const float (&x)[CH]) {
    if (CH % 4 == 0) {
        typedef float v4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int k = 0; k < CH; k += 4) {
            const v4 v = {x[k], x[k + 1], x[k + 2], x[k + 3]};
            __builtin_nontemporal_store(v, reinterpret_cast<v4 *>(p + k));
        }
    } else
#pragma unroll
        for (int i = 0; i < CH; ++i) __builtin_nontemporal_store(x[i], p + i);
}

// one thread per (pixel, channel chunk): consecutive threads touch consecutive bytes
// row, column and image width of packed pixel g (uniform batch: images of
// H x W; uneven batch: the item map)
__device__ __forceinline__ void pixel_rc(const AALaunch &L, long long g, int &r, int &c, int &W) {
    if (L.um.n) {
        const int img = um_find(L.um.pix, L.um.n, g);
        W = L.um.hw[2 * img + 1];
        const long long q = g - L.um.pix[img];
        r = (int)(q / W);
        c = (int)(q - (long long)r * W);
        return;
    }
    const long long npx = (long long)L.H * L.W;
    const long long q = g % npx;
    W = L.W;
    r = (int)(q / L.W);
    c = (int)(q % L.W);
}

template <int CV>
__global__ void __launch_bounds__(256) k_aa_gather(AALaunch L, const float2 *__restrict__ weights) {
    constexpr int CH = AAChunk<CV>::CH;
    const int C = CV > 0 ? CV : L.C;
    const int nch = C / CH;
    // XCD k takes the k-th eighth of the blocks, so a pixel's lower neighbour
    // (read a row earlier in the same eighth) is usually still in that XCD's L2
    long long gt = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    long long g = gt / nch;
    if (g >= L.total_pix) return;
    int ch = (int)(gt - g * nch) * CH;
    int r, c, W;
    pixel_rc(L, g, r, c, W);
    const float *col = L.color + ch;
    float2 wp = weights[g];
    float wl = c > 0 ? weights[g - 1].x : 0.f;        // pair (left, p): > 0 -> p receives
    float wd = r > 0 ? weights[g - W].y : 0.f;        // pair (below, p): > 0 -> p receives
    float cp[CH], o[CH], nb[CH];
    ld<CH>(col + g * C, cp);
#pragma unroll
    for (int i = 0; i < CH; ++i) o[i] = cp[i];
    if (wl > 0.f) {
        ld<CH>(col + (g - 1) * C, nb);
#pragma unroll
        for (int i = 0; i < CH; ++i) o[i] = fma_(wl, nb[i] - cp[i], o[i]);
    }
    if (wp.x < 0.f) {
        ld<CH>(col + (g + 1) * C, nb);
#pragma unroll
        for (int i = 0; i < CH; ++i) o[i] = fma_(-wp.x, nb[i] - cp[i], o[i]);
    }
    if (wd > 0.f) {
        ld<CH>(col + (g - W) * C, nb);
#pragma unroll
        for (int i = 0; i < CH; ++i) o[i] = fma_(wd, nb[i] - cp[i], o[i]);
    }
    if (wp.y < 0.f) {
        ld<CH>(col + (g + W) * C, nb);
#pragma unroll
        for (int i = 0; i < CH; ++i) o[i] = fma_(-wp.y, nb[i] - cp[i], o[i]);
    }
    st<CH>(L.out + ch + g * C, o);
}

// Per-block accumulation of vertex gradients before the global atomics.
// Device-scope float atomics on a multi-XCD GPU are performed beyond the
// per-XCD L2 caches, so every one is a memory transaction; neighbouring
// silhouette pairs share vertices, and summing them per 32 x 32 tile first
// leaves one atomic per distinct vertex and component.
constexpr int kAAHashBits = 9;
constexpr int kAAHash = 1 << kAAHashBits;  // table slots per tile (a 32 x 32 tile touches ~50-100 vertices)

struct VertexAcc {
    int key[kAAHash];
    float gx[kAAHash], gy[kAAHash], gw[kAAHash];
};

__device__ __forceinline__ void vacc_add(VertexAcc &A, float *GP, int v, float gx, float gy, float gw) {
    unsigned h = ((unsigned)v * 2654435761u) >> (32 - kAAHashBits);  // top bits
    for (int probe = 0; probe < 32; ++probe) {
        int k = atomicCAS(&A.key[h], -1, v);
        if (k == -1 || k == v) {
            atomicAdd(&A.gx[h], gx);
            atomicAdd(&A.gy[h], gy);
            atomicAdd(&A.gw[h], gw);
            return;
        }
        h = (h + 1) & (kAAHash - 1);
    }
    float *gp = GP + 4 * (long long)v;  // table crowded: straight to memory
    atomicAdd(gp + 0, gx);
    atomicAdd(gp + 1, gy);
    atomicAdd(gp + 3, gw);
}

// Position gradient of one blending pair: dL/dw (from the receiver's output
// gradient and the colour difference) times dw/dt times dt/d(edge vertices).
// The pair's decisions are the forward's; the derivative is taken in double:
// t = fa / (fa - fb) is ill-conditioned for edges nearly parallel to the pair.
// The edge's vertex loads do not wait for dL/dw, and the result is returned
// rather than added, so the loads of a pixel's two pairs overlap.
struct PairGrad {
    int v[2];
    float gx[2], gy[2], gw[2];
    bool live;
};

__device__ __forceinline__ PairGrad pair_pos_grad(const AALaunch &L, const AAImg &I, const float4 *P, long long g,
                                                  int r, int c, int dir, const PairRes &pr) {
    PairGrad out;
    const int rb = dir ? r + 1 : r, cb = dir ? c : c + 1;
    const long long ga = g, gb = dir ? g + I.W : g + 1;
    const long long gr = pr.recv == 0 ? ga : gb, gs = pr.recv == 0 ? gb : ga;
    const int *t3 = L.tri + 3 * (long long)pr.tri;
    const int i = pr.edge, j = pr.edge == 2 ? 0 : pr.edge + 1;
    out.v[0] = t3[i];
    out.v[1] = t3[j];
    const double2 q0 = I.proj[out.v[0]], q1 = I.proj[out.v[1]];
    const float4 pv[2] = {P[out.v[0]], P[out.v[1]]};
    const double so = I.orient[pr.tri] < 0 ? -1.0 : 1.0;
    double dLdw = 0.0;
    if ((L.C & 3) == 0) {
        const float4 *g4 = reinterpret_cast<const float4 *>(L.grad_out + gr * L.C);
        const float4 *cs = reinterpret_cast<const float4 *>(L.color + gs * L.C);
        const float4 *cr = reinterpret_cast<const float4 *>(L.color + gr * L.C);
        for (int k4 = 0; k4 < (L.C >> 2); ++k4) {
            float4 a = g4[k4], b = cs[k4], e = cr[k4];
            dLdw += (double)a.x * (b.x - e.x) + (double)a.y * (b.y - e.y) + (double)a.z * (b.z - e.z) +
                    (double)a.w * (b.w - e.w);
        }
    } else {
        for (int ch = 0; ch < L.C; ++ch)
            dLdw += (double)L.grad_out[gr * L.C + ch] * (L.color[gs * L.C + ch] - L.color[gr * L.C + ch]);
    }
    out.live = dLdw != 0.0;
    const bool recv_is_from = (pr.recv == 0) == (pr.from_a != 0);
    const double dLdt = recv_is_from ? -dLdw : dLdw;
    const double fxp = (double)(pr.from_a ? c : cb) + 0.5, fyp = (double)(pr.from_a ? r : rb) + 0.5;
    const double txp = (double)(pr.from_a ? cb : c) + 0.5, typ = (double)(pr.from_a ? rb : r) + 0.5;
    const double ex = q1.x - q0.x, ey = q1.y - q0.y;
    const double fa = so * (ex * (fyp - q0.y) - ey * (fxp - q0.x));
    const double fb = so * (ex * (typ - q0.y) - ey * (txp - q0.x));
    const double den = fa - fb;
    const double dt_dfa = -fb / (den * den), dt_dfb = fa / (den * den);
    const double k = dLdt * so * (double)L.boost;
    double gx2[2], gy2[2];
    gx2[0] = k * (dt_dfa * (q1.y - fyp) + dt_dfb * (q1.y - typ));
    gy2[0] = k * (dt_dfa * (fxp - q1.x) + dt_dfb * (txp - q1.x));
    gx2[1] = k * (dt_dfa * (fyp - q0.y) + dt_dfb * (typ - q0.y));
    gy2[1] = k * (dt_dfa * (q0.x - fxp) + dt_dfb * (q0.x - txp));
    const double hw = (double)I.W * 0.5, hh = (double)I.H * 0.5;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        double iw = 1.0 / (double)pv[m].w;
        double gx = gx2[m] * hw * iw, gy = gy2[m] * hh * iw;
        out.gx[m] = (float)gx;
        out.gy[m] = (float)gy;
        out.gw[m] = (float)(-(gx * pv[m].x + gy * pv[m].y) * iw);
    }
    return out;
}

__device__ __forceinline__ void pair_grad_add(VertexAcc &A, float *GP, const PairGrad &pg) {
    if (!pg.live) return;
#pragma unroll
    for (int m = 0; m < 2; ++m) vacc_add(A, GP, pg.v[m], pg.gx[m], pg.gy[m], pg.gw[m]);
}

// Backward pairs, one block (256 threads) per 32 x 32 tile of one image.  With
// the forward's decisions, read them; otherwise evaluate each pair and store
// its weight.  The tile's blending pairs (a fifth of all pairs on a folded
// mesh) are queued in LDS and their position gradients computed by
// consecutive lanes, so the double-precision chain runs on full waves; the
// gradients of the edge's two vertices are summed in the tile's table,
// flushed with one atomic per vertex component.
constexpr int kAAT = 32;                 // backward tile edge
constexpr int kAAPix = kAAT * kAAT;      // pixels per block (4 per thread)
constexpr int kAAQueue = 2 * kAAPix;     // at most two pairs per pixel

__global__ void __launch_bounds__(256) k_aa_pairs_bwd(AALaunch L, float2 *__restrict__ weights,
                                                      const int2 *__restrict__ decisions, int tiles_x,
                                                      int tiles_per_img, int n_blocks_pad) {
    __shared__ VertexAcc A;
    __shared__ int q_tri[kAAQueue], q_meta[kAAQueue];  // triangle; pixel | dir << 10 | decision bits << 11
    __shared__ int q_n;
    const int tid = threadIdx.x;
    const int bid = xcd_remap(blockIdx.x, n_blocks_pad);
    int img, lt;
    bool live_block;
    if (L.um.n) {  // uneven batch: the item map names the tile's item
        live_block = bid < L.total_t32;
        img = live_block ? um_find(L.um.t32, L.um.n, bid) : 0;
        lt = bid - L.um.t32[img];
        tiles_x = (L.um.hw[2 * img + 1] + kAAT - 1) / kAAT;
    } else {
        img = bid / tiles_per_img;
        lt = bid - img * tiles_per_img;
        live_block = img < L.n_img;
    }
    const bool want_pos = L.grad_pos != nullptr;
    for (int e = tid; e < kAAHash; e += 256) {
        A.key[e] = -1;
        A.gx[e] = 0.f;
        A.gy[e] = 0.f;
        A.gw[e] = 0.f;
    }
    if (tid == 0) q_n = 0;
    __syncthreads();
    const int r0 = (lt / tiles_x) * kAAT, c0 = (lt % tiles_x) * kAAT;
    const AAImg I = aa_img(L, live_block ? img : 0);
    const int H = I.H, W = I.W;
    const long long pix0 = L.um.n ? L.um.pix[live_block ? img : 0] : (long long)img * H * W;
    for (int k = 0; k < kAAPix / 256; ++k) {
        const int px = tid + 256 * k, r = r0 + (px >> 5), c = c0 + (px & 31);
        // the pixel's blending, unclamped pairs: slot a, slot b (direction: 0 right, 1 above)
        PairRes qa, qb;
        int da = 0, db = 0, nq = 0;
        qa.recv = qb.recv = -1;
        if (live_block && r < H && c < W) {
            long long g = pix0 + (long long)r * W + c;
            if (decisions) {
                int2 d = want_pos ? decisions[g] : make_int2(0, 0);
                const PairRes p0 = unpack_decision(d.x), p1 = unpack_decision(d.y);
                const bool h0 = d.x != 0 && !p0.clamped, h1 = d.y != 0 && !p1.clamped;
                qa = h0 ? p0 : p1;
                da = h0 ? 0 : 1;
                qb = p1;
                db = 1;
                nq = (int)h0 + (int)h1;
            } else {
                float2 w = make_float2(0.f, 0.f);
                for (int dir = 0; dir < 2; ++dir) {
                    int rb = dir ? r + 1 : r, cb = dir ? c : c + 1;
                    if (rb >= H || cb >= W) continue;
                    const float4 pa = I.rast[(long long)r * W + c], pb = I.rast[(long long)rb * W + cb];
                    const int ia = rast_id(pa, I.F), ib = rast_id(pb, I.F);
                    PairRes pr = eval_pair(I, L.tri, pa, pb, silh_of(I, ia), ia != ib ? silh_of(I, ib) : 0, r, c,
                                           rb, cb);
                    float sw = signed_w(pr);
                    if (dir) w.y = sw; else w.x = sw;
                    if (pr.recv < 0 || pr.clamped || !want_pos) continue;
                    if (nq == 0) { qa = pr; da = dir; }
                    else { qb = pr; db = dir; }
                    ++nq;
                }
                weights[g] = w;
            }
        }
        if (!want_pos || !live_block) continue;  // uniform across the block
        // wave-aggregated append of the pixel's queued pairs
        const int lane = tid & 63;
        int x = nq;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        int wtot = __shfl(x, 63, 64), base = 0;
        if (lane == 63 && wtot > 0) base = atomicAdd(&q_n, wtot);
        base = __shfl(base, 63, 64) + x - nq;
        if (nq > 0) {
            q_tri[base] = qa.tri;
            q_meta[base] = px | (da << 10) | ((qa.recv << 4 | qa.clamped << 3 | qa.from_a << 2 | qa.edge) << 11);
        }
        if (nq > 1) {
            q_tri[base + 1] = qb.tri;
            q_meta[base + 1] = px | (db << 10) | ((qb.recv << 4 | qb.clamped << 3 | qb.from_a << 2 | qb.edge) << 11);
        }
    }
    if (!want_pos || !live_block) return;  // uniform across the block
    __syncthreads();
    {
        const int n = q_n;
        const float4 *P = reinterpret_cast<const float4 *>(L.pos) + (L.pos_batched ? (long long)img * L.V : 0);
        float *GP = L.grad_pos + (L.pos_batched ? (long long)img * L.V * 4 : 0);
        for (int qi = tid; qi < n; qi += 256) {
            const int m = q_meta[qi], px = m & (kAAPix - 1), dir = (m >> 10) & 1, bits = m >> 11;
            PairRes pr;
            pr.tri = q_tri[qi];
            pr.recv = (bits >> 4) & 1;
            pr.clamped = (bits >> 3) & 1;
            pr.from_a = (bits >> 2) & 1;
            pr.edge = bits & 3;
            const int pr_r = r0 + (px >> 5), pr_c = c0 + (px & 31);
            const long long g = pix0 + (long long)pr_r * W + pr_c;
            pair_grad_add(A, GP, pair_pos_grad(L, I, P, g, pr_r, pr_c, dir, pr));
        }
    }
    __syncthreads();
    float *G = L.grad_pos + (L.pos_batched ? (long long)img * L.V * 4 : 0);
    for (int e = tid; e < kAAHash; e += 256) {
        int v = A.key[e];
        if (v < 0) continue;
        float *gp = G + 4 * (long long)v;
        atomicAdd(gp + 0, A.gx[e]);
        atomicAdd(gp + 1, A.gy[e]);
        atomicAdd(gp + 3, A.gw[e]);
    }
}

template <int CV>
__global__ void __launch_bounds__(256) k_aa_gather_bwd(AALaunch L, const float2 *__restrict__ weights) {
    constexpr int CH = AAChunkBwd<CV>::CH;
    const int C = CV > 0 ? CV : L.C;
    const int nch = C / CH;
    // XCD k takes the k-th eighth of the blocks, so a pixel's lower neighbour
    // (read a row earlier in the same eighth) is usually still in that XCD's L2
    long long gt = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    long long g = gt / nch;
    if (g >= L.total_pix) return;
    int ch = (int)(gt - g * nch) * CH;
    int r, c, W;
    pixel_rc(L, g, r, c, W);
    const float *go = L.grad_out + ch;
    float2 wp = weights[g];
    float wl = c > 0 ? weights[g - 1].x : 0.f;
    float wd = r > 0 ? weights[g - W].y : 0.f;
    // p receives: - w g[p];  p is the source of a neighbour's blend: + w g[neighbour]
    float self = 0.f;
    if (wl > 0.f) self += wl;
    if (wp.x < 0.f) self -= wp.x;
    if (wd > 0.f) self += wd;
    if (wp.y < 0.f) self -= wp.y;
    float gp[CH], acc[CH], nb[CH];
    ld<CH>(go + g * C, gp);
#pragma unroll
    for (int i = 0; i < CH; ++i) acc[i] = gp[i] - self * gp[i];
    if (wl < 0.f) {  // left neighbour received from p
        ld<CH>(go + (g - 1) * C, nb);
#pragma unroll
        for (int i = 0; i < CH; ++i) acc[i] = fma_(-wl, nb[i], acc[i]);
    }
    if (wp.x > 0.f) {  // right neighbour received from p
        ld<CH>(go + (g + 1) * C, nb);
#pragma unroll
        for (int i = 0; i < CH; ++i) acc[i] = fma_(wp.x, nb[i], acc[i]);
    }
    if (wd < 0.f) {
        ld<CH>(go + (g - W) * C, nb);
#pragma unroll
        for (int i = 0; i < CH; ++i) acc[i] = fma_(-wd, nb[i], acc[i]);
    }
    if (wp.y > 0.f) {
        ld<CH>(go + (g + W) * C, nb);
#pragma unroll
        for (int i = 0; i < CH; ++i) acc[i] = fma_(wp.y, nb[i], acc[i]);
    }
    st<CH>(L.grad_color + ch + g * C, acc);
}

#define DRP_AA_DISPATCH(KERNEL, ...)                                                         \
    switch (L.cv) {                                                                          \
        case 1: hipLaunchKernelGGL(KERNEL<1>, grid, block, 0, st, __VA_ARGS__); break;       \
        case 3: hipLaunchKernelGGL(KERNEL<3>, grid, block, 0, st, __VA_ARGS__); break;       \
        case 4: hipLaunchKernelGGL(KERNEL<4>, grid, block, 0, st, __VA_ARGS__); break;       \
        case 8: hipLaunchKernelGGL(KERNEL<8>, grid, block, 0, st, __VA_ARGS__); break;       \
        case 16: hipLaunchKernelGGL(KERNEL<16>, grid, block, 0, st, __VA_ARGS__); break;     \
        default: hipLaunchKernelGGL(KERNEL<0>, grid, block, 0, st, __VA_ARGS__); break;      \
    }

static void launch_orient(const AALaunch &L, hipStream_t st) {
    int n_proj = L.pos_batched ? L.n_img : 1;
    long long nv = (long long)n_proj * L.V, nf = (long long)n_proj * L.F;
    if (nv) hipLaunchKernelGGL(k_aa_project, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st, L.pos, L.V, n_proj,
                               L.H, L.W, L.um, L.proj);
    if (nf) {
        const dim3 grid((unsigned)((((nf + 255) / 256) + 7) & ~7ll));  // a multiple of 8 blocks (xcd_remap)
        hipLaunchKernelGGL(k_aa_orient, grid, dim3(256), 0, st, L.pos, L.V, L.tri, L.F, n_proj, L.orient);
        hipLaunchKernelGGL(k_aa_silhouette, grid, dim3(256), 0, st, L.opp, L.orient, L.F, n_proj, L.silh);
    }
}

// the gathers' grid, padded to a multiple of 8 blocks (xcd_remap)
static dim3 gather_grid(const AALaunch &L, long long total, bool bwd = false) {
    int nch = (L.cv > 0 && L.cv % 4 == 0) ? L.cv / 4 : (L.cv > 0 ? L.cv : L.C);
    if (bwd && L.cv == 8) nch = 1;  // AAChunkBwd: a pixel per thread
    const long long blocks = (total * nch + 255) / 256;
    return dim3((unsigned)((blocks + 7) & ~7ll));
}

// a uniform launch's totals in the fields the uneven path fills from its plan
static AALaunch with_totals(const AALaunch &L) {
    AALaunch K = L;
    if (!K.um.n) K.total_pix = (long long)L.n_img * L.H * L.W;
    return K;
}

int launch_antialias(const AALaunch &L0, hipStream_t st) {
    const AALaunch L = with_totals(L0);
    const long long total = L.total_pix;
    if (total == 0) return 0;
    if (!L.prepared) launch_orient(L, st);
    const long long nblk = L.um.n ? (long long)L.total_blk
                                  : (long long)((L.W + 63) / 64) * ((L.H + kAAPairRows - 1) / kAAPairRows) * L.n_img;
    if (nblk >= (1ll << 31)) return -1;
    dim3 grid((unsigned)((nblk + 7) & ~7ll)), block(256);
    hipLaunchKernelGGL(k_aa_pairs, grid, block, 0, st, L, L.weights, L.decisions);
    grid = gather_grid(L, total);
    DRP_AA_DISPATCH(k_aa_gather, L, (const float2 *)L.weights)
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

// the per-vertex pixel positions and per-triangle orientation and silhouette
// flags alone, for a forward that follows with L.prepared set
int launch_antialias_prepare(const AALaunch &L0, hipStream_t st) {
    const AALaunch L = with_totals(L0);
    if (L.total_pix == 0) return 0;
    launch_orient(L, st);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_antialias_bwd(const AALaunch &L0, hipStream_t st) {
    const AALaunch L = with_totals(L0);
    const long long total = L.total_pix;
    if (total == 0) return 0;
    // with the forward's state, projections, orientations, weights and decisions are in place
    if (!L.reuse) launch_orient(L, st);
    int tiles_x = (L.W + kAAT - 1) / kAAT, tiles_per_img = tiles_x * ((L.H + kAAT - 1) / kAAT);
    long long nb = L.um.n ? (long long)L.total_t32 : (long long)tiles_per_img * L.n_img, nb_pad = (nb + 7) & ~7ll;
    dim3 block(256);
    // With the forward's weights the position pass (latency-bound) and the
    // colour gather (bandwidth-bound) are independent: they run side by side.
    // Otherwise the position pass evaluates the pairs and stores the weights
    // the gather reads.
    Fork f{};
    const bool par = L.reuse && fork_begin(st, f);
    hipLaunchKernelGGL(k_aa_pairs_bwd, dim3((unsigned)nb_pad), block, 0, par ? f.side : st, L, L.weights,
                       L.reuse ? (const int2 *)L.decisions : (const int2 *)nullptr, tiles_x, tiles_per_img,
                       (int)nb_pad);
    dim3 grid = gather_grid(L, total, true);
    DRP_AA_DISPATCH(k_aa_gather_bwd, L, (const float2 *)L.weights)
    if (par && fork_end(st, f)) return -2;
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace drp
