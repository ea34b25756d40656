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
// Kernels: k_aa_project (per-vertex pixel coordinates), k_aa_orient
// (per-triangle orientation), k_aa_silhouette (triangles with a possible
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
// projection / orientation precompute
// ----------------------------------------------------------------------------
__device__ __forceinline__ int orient3(float4 a, float4 b, float4 c) {
    // sign of det[[x0,x1,x2],[y0,y1,y2],[w0,w1,w2]]
    float m0 = fma_(b.y, c.w, -(c.y * b.w));
    float m1 = fma_(a.y, c.w, -(c.y * a.w));
    float m2 = fma_(a.y, b.w, -(b.y * a.w));
    float d = fma_(a.x, m0, fma_(-b.x, m1, c.x * m2));
    return d > 0.f ? 1 : (d < 0.f ? -1 : 0);
}

__global__ void k_aa_project(const float *__restrict__ pos, int V, int n_proj, int H, int W,
                             float2 *__restrict__ proj) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)n_proj * V) return;
    float4 p = reinterpret_cast<const float4 *>(pos)[g];
    float2 o;
    if (p.w > 0.f) {
        o.x = (p.x / p.w + 1.0f) * ((float)W * 0.5f);
        o.y = (p.y / p.w + 1.0f) * ((float)H * 0.5f);
    } else {
        o.x = __int_as_float(0x7fc00000);
        o.y = o.x;
    }
    proj[g] = o;
}

__global__ void k_aa_orient(const float *__restrict__ pos, int V, const int *__restrict__ tri, int F, int n_proj,
                            signed char *__restrict__ orient) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
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

// silh[img][t] = 1 when t has an edge that can be a silhouette in that image:
// no neighbour, or a neighbour of another screen orientation.  A pair whose
// candidate triangle has none cannot blend through it, so the pair evaluation
// skips the edge search for it (same outcome, far fewer gathers).
__global__ void k_aa_silhouette(const int *__restrict__ opp, const signed char *__restrict__ orient, int F,
                                int n_proj, unsigned char *__restrict__ silh) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)n_proj * F) return;
    int img = (int)(g / F), t = (int)(g % F);
    const signed char *O = orient + (long long)img * F;
    int o = O[t];
    unsigned char f = 0;
    if (o != 0) {
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            int n = opp[3 * (long long)t + e];
            if (!(n >= 0 && n < F && O[n] == o)) f = 1;
        }
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
    float fa, fb;
};

struct AAImg {
    const float4 *rast;         // image rast
    const float2 *proj;         // image's projected vertices
    const signed char *orient;  // image's triangle orientations
    const unsigned char *silh;  // image's silhouette-capable triangles
    int V, F, H, W;
};

__device__ __forceinline__ PairRes eval_pair(const AAImg &I, const int *__restrict__ tri,
                                             const int *__restrict__ opp, int ra, int ca, int rb, int cb) {
    PairRes res;
    res.recv = -1;
    float4 pa = I.rast[(long long)ra * I.W + ca], pb = I.rast[(long long)rb * I.W + cb];
    int ida = pa.w > 0.f ? float_to_triidx(pa.w) : 0;
    int idb = pb.w > 0.f ? float_to_triidx(pb.w) : 0;
    if (ida > I.F) ida = 0;
    if (idb > I.F) idb = 0;
    if (ida == idb) return res;
    int a_first = 1;
    if (ida == 0) a_first = 0;
    else if (idb != 0 && pb.z < pa.z) a_first = 0;
    for (int pass = 0; pass < 2; ++pass) {
        int use_a = pass == 0 ? a_first : !a_first;
        int id = use_a ? ida : idb;
        if (id == 0) continue;
        int t = id - 1;
        if (!I.silh[t]) continue;
        const int *t3 = tri + 3 * (long long)t;
        int vi[3] = {t3[0], t3[1], t3[2]};
        if ((unsigned)vi[0] >= (unsigned)I.V || (unsigned)vi[1] >= (unsigned)I.V || (unsigned)vi[2] >= (unsigned)I.V)
            continue;
        float2 q0 = I.proj[vi[0]], q1 = I.proj[vi[1]], q2 = I.proj[vi[2]];
        if (isnan(q0.x) || isnan(q1.x) || isnan(q2.x)) continue;
        int o = I.orient[t];
        if (o == 0) continue;
        float X[3] = {q0.x, q1.x, q2.x}, Y[3] = {q0.y, q1.y, q2.y};
        float fx = (float)(use_a ? ca : cb) + 0.5f, fy = (float)(use_a ? ra : rb) + 0.5f;
        float tx = (float)(use_a ? cb : ca) + 0.5f, ty = (float)(use_a ? rb : ra) + 0.5f;
        int best = -1;
        float bt = 0.f, bfa = 0.f, bfb = 0.f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            int j = k == 2 ? 0 : k + 1;
            float ex = X[j] - X[k], ey = Y[j] - Y[k];
            float ff = fma_(ex, fy - Y[k], -(ey * (fx - X[k])));
            float ft = fma_(ex, ty - Y[k], -(ey * (tx - X[k])));
            if (o < 0) { ff = -ff; ft = -ft; }
            if (!(ft < 0.f)) continue;
            float te = ff / (ff - ft);
            if (best < 0 || te < bt) { best = k; bt = te; bfa = ff; bfb = ft; }
        }
        if (best < 0) continue;
        int n = opp[3 * (long long)t + best];
        if (n >= 0 && n < I.F && I.orient[n] == o) continue;
        int clamped = 0;
        if (bt < 0.f) { bt = 0.f; clamped = 1; }
        if (bt > 1.f) { bt = 1.f; clamped = 1; }
        res.tri = t;
        res.edge = best;
        res.from_a = use_a;
        res.fa = bfa;
        res.fb = bfb;
        res.clamped = clamped;
        if (bt < 0.5f) { res.recv = use_a ? 0 : 1; res.wgt = 0.5f - bt; }
        else { res.recv = use_a ? 1 : 0; res.wgt = bt - 0.5f; }
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

// weights[pix] = (signed weight of pair (pix, right), of pair (pix, above))
__global__ void __launch_bounds__(256) k_aa_pairs(AALaunch L, float2 *__restrict__ weights) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long npx = (long long)L.H * L.W;
    if (g >= (long long)L.n_img * npx) return;
    int img = (int)(g / npx);
    long long q = g - img * npx;
    int r = (int)(q / L.W), c = (int)(q % L.W);
    AAImg I = aa_img(L, img);
    float2 w = make_float2(0.f, 0.f);
    if (c + 1 < L.W) w.x = signed_w(eval_pair(I, L.tri, L.opp, r, c, r, c + 1));
    if (r + 1 < L.H) w.y = signed_w(eval_pair(I, L.tri, L.opp, r, c, r + 1, c));
    weights[g] = w;
}

// channel chunk handled by one thread: 4 (one float4) when C is a multiple of 4
template <int CV>
struct AAChunk {
    static constexpr int CH = (CV > 0 && CV % 4 == 0) ? 4 : 1;
};

template <int CH>
__device__ __forceinline__ void ld(const float *p, float (&x)[CH]) {
    if (CH == 4) {
        float4 v = *reinterpret_cast<const float4 *>(p);
        x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
    } else {
#pragma unroll
        for (int i = 0; i < CH; ++i) x[i] = p[i];
    }
}
template <int CH>
__device__ __forceinline__ void st(float *p, const float (&x)[CH]) {
    if (CH == 4) *reinterpret_cast<float4 *>(p) = make_float4(x[0], x[1], x[2], x[3]);
    else
#pragma unroll
        for (int i = 0; i < CH; ++i) p[i] = x[i];
}

// one thread per (pixel, channel chunk): consecutive threads touch consecutive bytes
template <int CV>
__global__ void __launch_bounds__(256) k_aa_gather(AALaunch L, const float2 *__restrict__ weights) {
    constexpr int CH = AAChunk<CV>::CH;
    const int C = CV > 0 ? CV : L.C;
    const int nch = C / CH;
    long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long npx = (long long)L.H * L.W;
    long long g = gt / nch;
    if (g >= (long long)L.n_img * npx) return;
    int ch = (int)(gt - g * nch) * CH;
    int img = (int)(g / npx);
    long long q = g - img * npx;
    int r = (int)(q / L.W), c = (int)(q % L.W);
    const float *col = L.color + ch;
    float2 wp = weights[g];
    float wl = c > 0 ? weights[g - 1].x : 0.f;        // pair (left, p): > 0 -> p receives
    float wd = r > 0 ? weights[g - L.W].y : 0.f;      // pair (below, p): > 0 -> p receives
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
        ld<CH>(col + (g - L.W) * C, nb);
#pragma unroll
        for (int i = 0; i < CH; ++i) o[i] = fma_(wd, nb[i] - cp[i], o[i]);
    }
    if (wp.y < 0.f) {
        ld<CH>(col + (g + L.W) * C, nb);
#pragma unroll
        for (int i = 0; i < CH; ++i) o[i] = fma_(-wp.y, nb[i] - cp[i], o[i]);
    }
    st<CH>(L.out + ch + g * C, o);
}

// backward pairs: recompute each pair once, store its weight, scatter the
// position gradient of the crossing edge's two vertices
__global__ void __launch_bounds__(256) k_aa_pairs_bwd(AALaunch L, float2 *__restrict__ weights) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long npx = (long long)L.H * L.W;
    if (g >= (long long)L.n_img * npx) return;
    int img = (int)(g / npx);
    long long q = g - img * npx;
    int r = (int)(q / L.W), c = (int)(q % L.W);
    AAImg I = aa_img(L, img);
    float2 w = make_float2(0.f, 0.f);
    const float4 *P = reinterpret_cast<const float4 *>(L.pos) + (L.pos_batched ? (long long)img * L.V : 0);
    float *GP = L.grad_pos + (L.pos_batched ? (long long)img * L.V * 4 : 0);
    for (int dir = 0; dir < 2; ++dir) {
        int rb = dir ? r + 1 : r, cb = dir ? c : c + 1;
        if (rb >= L.H || cb >= L.W) continue;
        PairRes pr = eval_pair(I, L.tri, L.opp, r, c, rb, cb);
        float sw = signed_w(pr);
        if (dir) w.y = sw; else w.x = sw;
        if (pr.recv < 0 || pr.clamped || L.grad_pos == nullptr) continue;
        long long ga = g, gb = dir ? g + L.W : g + 1;
        long long gr = pr.recv == 0 ? ga : gb, gs = pr.recv == 0 ? gb : ga;
        float dLdw = 0.f;
        if ((L.C & 3) == 0) {
            const float4 *g4 = reinterpret_cast<const float4 *>(L.grad_out + gr * L.C);
            const float4 *cs = reinterpret_cast<const float4 *>(L.color + gs * L.C);
            const float4 *cr = reinterpret_cast<const float4 *>(L.color + gr * L.C);
            for (int k4 = 0; k4 < (L.C >> 2); ++k4) {
                float4 a = g4[k4], b = cs[k4], e = cr[k4];
                dLdw += a.x * (b.x - e.x) + a.y * (b.y - e.y) + a.z * (b.z - e.z) + a.w * (b.w - e.w);
            }
        } else {
            for (int ch = 0; ch < L.C; ++ch)
                dLdw += L.grad_out[gr * L.C + ch] * (L.color[gs * L.C + ch] - L.color[gr * L.C + ch]);
        }
        if (dLdw == 0.f) continue;
        bool recv_is_from = (pr.recv == 0) == (pr.from_a != 0);
        float dLdt = recv_is_from ? -dLdw : dLdw;
        float den = pr.fa - pr.fb;
        float dt_dfa = -pr.fb / (den * den), dt_dfb = pr.fa / (den * den);
        const int *t3 = L.tri + 3 * (long long)pr.tri;
        int i = pr.edge, j = pr.edge == 2 ? 0 : pr.edge + 1;
        float so = I.orient[pr.tri] < 0 ? -1.f : 1.f;
        float2 qi = I.proj[t3[i]], qj = I.proj[t3[j]];
        int ra = r, ca = c;
        float fxp = (float)(pr.from_a ? ca : cb) + 0.5f, fyp = (float)(pr.from_a ? ra : rb) + 0.5f;
        float txp = (float)(pr.from_a ? cb : ca) + 0.5f, typ = (float)(pr.from_a ? rb : ra) + 0.5f;
        float k = dLdt * so * L.boost;
        float gXi = k * (dt_dfa * (qj.y - fyp) + dt_dfb * (qj.y - typ));
        float gYi = k * (dt_dfa * (fxp - qj.x) + dt_dfb * (txp - qj.x));
        float gXj = k * (dt_dfa * (fyp - qi.y) + dt_dfb * (typ - qi.y));
        float gYj = k * (dt_dfa * (qi.x - fxp) + dt_dfb * (qi.x - txp));
        float hw = (float)L.W * 0.5f, hh = (float)L.H * 0.5f;
        int vv[2] = {t3[i], t3[j]};
        float gx2[2] = {gXi, gXj}, gy2[2] = {gYi, gYj};
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            float4 pv = P[vv[m]];
            float *gp = GP + 4 * (long long)vv[m];
            float iw = 1.f / pv.w;
            atomicAdd(gp + 0, gx2[m] * hw * iw);
            atomicAdd(gp + 1, gy2[m] * hh * iw);
            atomicAdd(gp + 3, -(gx2[m] * hw * pv.x + gy2[m] * hh * pv.y) * iw * iw);
        }
    }
    weights[g] = w;
}

template <int CV>
__global__ void __launch_bounds__(256) k_aa_gather_bwd(AALaunch L, const float2 *__restrict__ weights) {
    constexpr int CH = AAChunk<CV>::CH;
    const int C = CV > 0 ? CV : L.C;
    const int nch = C / CH;
    long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long npx = (long long)L.H * L.W;
    long long g = gt / nch;
    if (g >= (long long)L.n_img * npx) return;
    int ch = (int)(gt - g * nch) * CH;
    int img = (int)(g / npx);
    long long q = g - img * npx;
    int r = (int)(q / L.W), c = (int)(q % L.W);
    const float *go = L.grad_out + ch;
    float2 wp = weights[g];
    float wl = c > 0 ? weights[g - 1].x : 0.f;
    float wd = r > 0 ? weights[g - L.W].y : 0.f;
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
        ld<CH>(go + (g - L.W) * C, nb);
#pragma unroll
        for (int i = 0; i < CH; ++i) acc[i] = fma_(-wd, nb[i], acc[i]);
    }
    if (wp.y > 0.f) {
        ld<CH>(go + (g + L.W) * C, nb);
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

static void launch_project(const AALaunch &L, hipStream_t st) {
    int n_proj = L.pos_batched ? L.n_img : 1;
    long long nv = (long long)n_proj * L.V, nf = (long long)n_proj * L.F;
    if (nv) hipLaunchKernelGGL(k_aa_project, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st, L.pos, L.V, n_proj,
                               L.H, L.W, L.proj);
    if (nf) {
        hipLaunchKernelGGL(k_aa_orient, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, st, L.pos, L.V, L.tri, L.F,
                           n_proj, L.orient);
        hipLaunchKernelGGL(k_aa_silhouette, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, st, L.opp, L.orient,
                           L.F, n_proj, L.silh);
    }
}

static dim3 gather_grid(const AALaunch &L, long long total) {
    int nch = (L.cv > 0 && L.cv % 4 == 0) ? L.cv / 4 : (L.cv > 0 ? L.cv : L.C);
    return dim3((unsigned)((total * nch + 255) / 256));
}

int launch_antialias(const AALaunch &L, hipStream_t st) {
    long long total = (long long)L.n_img * L.H * L.W;
    if (total == 0) return 0;
    launch_project(L, st);
    dim3 grid((unsigned)((total + 255) / 256)), block(256);
    hipLaunchKernelGGL(k_aa_pairs, grid, block, 0, st, L, L.weights);
    grid = gather_grid(L, total);
    DRP_AA_DISPATCH(k_aa_gather, L, (const float2 *)L.weights)
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_antialias_bwd(const AALaunch &L, hipStream_t st) {
    long long total = (long long)L.n_img * L.H * L.W;
    if (total == 0) return 0;
    launch_project(L, st);
    dim3 grid((unsigned)((total + 255) / 256)), block(256);
    hipLaunchKernelGGL(k_aa_pairs_bwd, grid, block, 0, st, L, L.weights);
    grid = gather_grid(L, total);
    DRP_AA_DISPATCH(k_aa_gather_bwd, L, (const float2 *)L.weights)
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace drp
