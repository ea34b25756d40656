/*
 * CPU oracle for the differentiable rasterisation path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests, the smoke check and the benchmark's CPU-baseline leg load this
 * library, and only as the checker: nothing in the product links it.
 *
 * It restates, as plain scalar loops, the operators diffrp drives through
 * nvdiffrast.torch:
 *   rasterize   (diffrp/rendering/surface_deferred.py:240-266, DepthPeeler :260-265)
 *   rasterize backward (autograd of rast, reached through interpolator.py:42,92)
 *   interpolate (interpolator.py:76-84; the pixel formula of interpolator.py:39-45)
 *   antialias   (surface_deferred.py:758-775) and its topology (edge -> triangle) table
 *   tri-id float codec (interpolator.py:24-29)
 * nvdiffrast itself is absent, so its numerics are not restated; the conventions
 * below are this project's documented ones, and the HIP kernels implement the
 * same floating-point operation sequence so that triangle ids and depth agree
 * bit for bit.
 *
 * Conventions (shared with the HIP kernels):
 *  - buffers are bottom-row first; pixel (r, c) has NDC centre
 *    ((2c+1-W)/W, (2r+1-H)/H);
 *  - "direct" triangles (all w > 0 and |x|,|y| <= GUARD*w, GUARD = 4096) are
 *    snapped to 4-bit sub-pixel fixed point, X = rint(((x/w)+1) * (W/2) * 16)
 *    (the precision that reproduces the reference's 512^2 README render
 *    mask exactly, docs/source/assets/procedural-normals.png), and covered with
 *    exact int64 edge functions under a top-left rule (y up, CCW after
 *    orientation normalisation);
 *  - other triangles are covered with the 2D-homogeneous test: every e_i has the
 *    sign of sum_i e_i w_i;
 *  - perspective-correct barycentrics from unsnapped clip coordinates:
 *    X_k = x_k - w_k px, Y_k = y_k - w_k py (one fma each),
 *    e0 = X1 Y2 - X2 Y1, e1 = X2 Y0 - X0 Y2, e2 = X0 Y1 - X1 Y0,
 *    s = (e0 + e1) + e2, u = e0/s, v = e1/s, z/w = sum e_k z_k / sum e_k w_k;
 *  - fragments with s == 0 or z/w outside [-1, 1] are discarded;
 *  - depth test: smallest 64-bit key (orderable(z/w) << 32 | triangle) wins;
 *    depth peeling takes the smallest key strictly above the previous layer's.
 *
 * The file compiles twice: REAL=float (bit-exact checker) and, with
 * ORA_DOUBLE, REAL=double (finite-difference oracle; symbols end in _f64).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef ORA_DOUBLE
typedef double real;
#define SFX(n) n##_f64
#define FMA fma
#define RINT rint
#define FABS fabs
#else
typedef float real;
#define SFX(n) n##_f32
#define FMA fmaf
#define RINT rintf
#define FABS fabsf
#endif

#define ORA_GUARD ((real)4096)
#define ORA_SUBPIX 16   /* 4 sub-pixel bits; pixel centre at +8 */

/* ------------------------------------------------------------------------- */
/* shared (precision independent) helpers: defined once                        */
/* ------------------------------------------------------------------------- */
#ifndef ORA_DOUBLE
static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* float -> unsigned key with the same total order as the floats (-0 < +0). */
static inline uint32_t ord_f32(float f) {
    uint32_t b = f2u(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

/* Triangle-id codec of interpolator.py:24-29: n <= 2^24 -> (float)n,
 * otherwise the float whose bits are 0x4a800000 + n (continuous at 2^24). */
float ora_triidx_to_float(int32_t n) {
    return (n <= 0x01000000) ? (float)n : u2f(0x4a800000u + (uint32_t)n);
}
int32_t ora_float_to_triidx(float x) {
    return (x <= 16777216.0f) ? (int32_t)x : (int32_t)(f2u(x) - 0x4a800000u);
}
/* Threads of the oracle's loops.  Results do not depend on it: forward pixels
 * are computed on their own, the depth test is an order-free minimum, and the
 * backward passes sum into per-band buffers whose number is fixed by the
 * image height and the buffer size, added in band order. */
void ora_set_threads(int n) { omp_set_num_threads(n > 0 ? n : omp_get_num_procs()); }

void ora_encode_ids(const int32_t *n, float *out, int64_t count) {
    for (int64_t i = 0; i < count; ++i) out[i] = ora_triidx_to_float(n[i]);
}
void ora_decode_ids(const float *x, int32_t *out, int64_t count) {
    for (int64_t i = 0; i < count; ++i) out[i] = ora_float_to_triidx(x[i]);
}

/* Edge -> opposite-triangle table ("topology hash" of antialias).
 * opp[3t+k] = triangle sharing edge (v_k, v_{k+1 mod 3}) of t, or -1 when that
 * undirected edge has one triangle or more than two (boundary / non-manifold). */
typedef struct { int64_t key; int32_t slot; } ora_edge;
static int cmp_edge(const void *a, const void *b) {
    const ora_edge *x = (const ora_edge *)a, *y = (const ora_edge *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->slot < y->slot ? -1 : (x->slot > y->slot);
}
int ora_topology(const int32_t *tri, int F, int32_t *opp) {
    ora_edge *e = (ora_edge *)malloc(sizeof(ora_edge) * (size_t)(3 * (int64_t)F + 1));
    if (!e) return -1;
    for (int t = 0; t < F; ++t)
        for (int k = 0; k < 3; ++k) {
            int64_t a = tri[3 * t + k], b = tri[3 * t + (k + 1) % 3];
            int64_t lo = a < b ? a : b, hi = a < b ? b : a;
            e[3 * t + k].key = (lo << 32) | hi;
            e[3 * t + k].slot = 3 * t + k;
            opp[3 * t + k] = -1;
        }
    qsort(e, (size_t)3 * F, sizeof(ora_edge), cmp_edge);
    for (int64_t i = 0; i < 3 * (int64_t)F;) {
        int64_t j = i;
        while (j < 3 * (int64_t)F && e[j].key == e[i].key) ++j;
        if (j - i == 2) {
            opp[e[i].slot] = e[i + 1].slot / 3;
            opp[e[i + 1].slot] = e[i].slot / 3;
        }
        i = j;
    }
    free(e);
    return 0;
}
#endif

/* ------------------------------------------------------------------------- */
/* rasterize backward of one pixel, in double precision (both variants)        */
/* ------------------------------------------------------------------------- */
/* d(u, v, z/w) at pixel (r, c) of the triangle with clip corners (x, y, z, w)_k,
 * plus d rast_db when gd (already scaled by 2/W, 2/H) is given, chained to
 * d(x, y, z, w)_k.  u = e0/s, v = e1/s, z/w = sum e_k z_k / sum e_k w_k with
 * e_k the 2-D homogeneous edge functions of X_k = x_k - w_k px, Y_k = y_k - w_k py. */
static void pix_grad_d(const double x[3], const double y[3], const double z[3], const double w[3], int r, int c,
                       int H, int W, double gu, double gv, double gz, const double *gd, double out[3][4]) {
    double px = (double)(2 * c + 1 - W) / (double)W, py = (double)(2 * r + 1 - H) / (double)H;
    double X[3], Y[3], e[3];
    for (int k = 0; k < 3; ++k) {
        X[k] = fma(-w[k], px, x[k]);
        Y[k] = fma(-w[k], py, y[k]);
    }
    e[0] = fma(X[1], Y[2], -(X[2] * Y[1]));
    e[1] = fma(X[2], Y[0], -(X[0] * Y[2]));
    e[2] = fma(X[0], Y[1], -(X[1] * Y[0]));
    double s = (e[0] + e[1]) + e[2];
    double zn = fma(e[0], z[0], fma(e[1], z[1], e[2] * z[2]));
    double wn = fma(e[0], w[0], fma(e[1], w[1], e[2] * w[2]));
    double u = e[0] / s, v = e[1] / s, zw = zn / wn;
    double gX[3] = {0, 0, 0}, gY[3] = {0, 0, 0}, gw[3] = {0, 0, 0}, gs = 0;
    if (gd) {
        /* rast_db = (e_kx - u sx)/s etc. (scaled); chain into e_kx, e_ky, u, v, s */
        double ex[3], ey[3];
        ex[0] = fma(w[2], Y[1], -(w[1] * Y[2]));
        ex[1] = fma(w[0], Y[2], -(w[2] * Y[0]));
        ex[2] = fma(w[1], Y[0], -(w[0] * Y[1]));
        ey[0] = fma(w[1], X[2], -(w[2] * X[1]));
        ey[1] = fma(w[2], X[0], -(w[0] * X[2]));
        ey[2] = fma(w[0], X[1], -(w[1] * X[0]));
        double sx = ex[0] + ex[1] + ex[2], sy = ey[0] + ey[1] + ey[2];
        double Aux = ex[0] - u * sx, Auy = ey[0] - u * sy, Avx = ex[1] - v * sx, Avy = ey[1] - v * sy;
        double a = gd[0] / s, b = gd[1] / s, cc = gd[2] / s, d = gd[3] / s;
        double gex[3] = {a, cc, 0}, gey[3] = {b, d, 0};
        double gsx = -(a * u + cc * v), gsy = -(b * u + d * v);
        for (int k = 0; k < 3; ++k) { gex[k] += gsx; gey[k] += gsy; }
        gu += -(a * sx + b * sy);
        gv += -(cc * sx + d * sy);
        gs += -(a * Aux + b * Auy + cc * Avx + d * Avy) / s;
        /* e_kx, e_ky in terms of w, X, Y */
        gw[2] += gex[0] * Y[1]; gY[1] += gex[0] * w[2]; gw[1] -= gex[0] * Y[2]; gY[2] -= gex[0] * w[1];
        gw[0] += gex[1] * Y[2]; gY[2] += gex[1] * w[0]; gw[2] -= gex[1] * Y[0]; gY[0] -= gex[1] * w[2];
        gw[1] += gex[2] * Y[0]; gY[0] += gex[2] * w[1]; gw[0] -= gex[2] * Y[1]; gY[1] -= gex[2] * w[0];
        gw[1] += gey[0] * X[2]; gX[2] += gey[0] * w[1]; gw[2] -= gey[0] * X[1]; gX[1] -= gey[0] * w[2];
        gw[2] += gey[1] * X[0]; gX[0] += gey[1] * w[2]; gw[0] -= gey[1] * X[2]; gX[2] -= gey[1] * w[0];
        gw[0] += gey[2] * X[1]; gX[1] += gey[2] * w[0]; gw[1] -= gey[2] * X[0]; gX[0] -= gey[2] * w[1];
    }
    double G0 = (gu * (1 - u) + gv * (-v)) / s + gs + gz * (z[0] - zw * w[0]) / wn;
    double G1 = (gu * (-u) + gv * (1 - v)) / s + gs + gz * (z[1] - zw * w[1]) / wn;
    double G2 = (gu * (-u) + gv * (-v)) / s + gs + gz * (z[2] - zw * w[2]) / wn;
    gX[0] += G2 * Y[1] - G1 * Y[2];
    gX[1] += G0 * Y[2] - G2 * Y[0];
    gX[2] += G1 * Y[0] - G0 * Y[1];
    gY[0] += G1 * X[2] - G2 * X[1];
    gY[1] += G2 * X[0] - G0 * X[2];
    gY[2] += G0 * X[1] - G1 * X[0];
    for (int k = 0; k < 3; ++k) {
        out[k][0] = gX[k];
        out[k][1] = gY[k];
        out[k][2] = gz * e[k] / wn;
        out[k][3] = gw[k] - px * gX[k] - py * gY[k] - gz * zw * e[k] / wn;
    }
}

/* ------------------------------------------------------------------------- */
/* triangle setup                                                              */
/* ------------------------------------------------------------------------- */
typedef struct {
    real x[3], y[3], z[3], w[3];
    int cls;                 /* 0 culled, 1 direct (fixed point), 2 homogeneous */
    int64_t X[3], Y[3];      /* snapped vertices, CCW order (direct only)        */
    int tl[3];               /* top-left flag of edge k -> k+1 (direct only)     */
    int c0, c1, r0, r1;      /* candidate pixel box, inclusive                   */
} SFX(ora_setup);

static inline int64_t floor_div(int64_t a, int64_t b) {
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
    return q;
}

static void SFX(setup_tri)(const real *pos, int V, const int32_t *t3, int H, int W, SFX(ora_setup) *s) {
    s->cls = 0;
    for (int k = 0; k < 3; ++k) {
        int32_t vi = t3[k];
        if (vi < 0 || vi >= V) return;
        s->x[k] = pos[4 * (int64_t)vi + 0];
        s->y[k] = pos[4 * (int64_t)vi + 1];
        s->z[k] = pos[4 * (int64_t)vi + 2];
        s->w[k] = pos[4 * (int64_t)vi + 3];
        if (!isfinite(s->x[k]) || !isfinite(s->y[k]) || !isfinite(s->z[k]) || !isfinite(s->w[k])) return;
    }
    int direct = 1;
    for (int k = 0; k < 3; ++k)
        if (!(s->w[k] > 0 && FABS(s->x[k]) <= ORA_GUARD * s->w[k] && FABS(s->y[k]) <= ORA_GUARD * s->w[k]))
            direct = 0;
    if (direct) {
        real hw = (real)W * (real)0.5, hh = (real)H * (real)0.5;
        int64_t X[3], Y[3];
        for (int k = 0; k < 3; ++k) {
            real nx = s->x[k] / s->w[k], ny = s->y[k] / s->w[k];
            real fx = (nx + (real)1) * hw, fy = (ny + (real)1) * hh;
            X[k] = (int64_t)RINT(fx * (real)ORA_SUBPIX);
            Y[k] = (int64_t)RINT(fy * (real)ORA_SUBPIX);
        }
        int64_t area = (X[1] - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (Y[1] - Y[0]);
        if (area == 0) return;
        int ord[3] = {0, 1, 2};
        if (area < 0) { ord[1] = 2; ord[2] = 1; }
        for (int k = 0; k < 3; ++k) { s->X[k] = X[ord[k]]; s->Y[k] = Y[ord[k]]; }
        for (int k = 0; k < 3; ++k) {
            int64_t dx = s->X[(k + 1) % 3] - s->X[k], dy = s->Y[(k + 1) % 3] - s->Y[k];
            s->tl[k] = (dy < 0) || (dy == 0 && dx < 0);
        }
        int64_t mnx = X[0], mxx = X[0], mny = Y[0], mxy = Y[0];
        for (int k = 1; k < 3; ++k) {
            if (X[k] < mnx) mnx = X[k];
            if (X[k] > mxx) mxx = X[k];
            if (Y[k] < mny) mny = Y[k];
            if (Y[k] > mxy) mxy = Y[k];
        }
        int64_t c0 = -floor_div(-(mnx - 8), 16), c1 = floor_div(mxx - 8, 16);
        int64_t r0 = -floor_div(-(mny - 8), 16), r1 = floor_div(mxy - 8, 16);
        if (c0 < 0) c0 = 0;
        if (r0 < 0) r0 = 0;
        if (c1 > W - 1) c1 = W - 1;
        if (r1 > H - 1) r1 = H - 1;
        if (c0 > c1 || r0 > r1) return;
        s->c0 = (int)c0; s->c1 = (int)c1; s->r0 = (int)r0; s->r1 = (int)r1;
        s->cls = 1;
        return;
    }
    /* homogeneous: a conservative pixel box of the w > 0 part (coverage itself is
     * decided per pixel).  Reject triangles wholly beyond one frustum side (pixel
     * rays have |x|, |y| < w) or wholly behind the eye; bound the rest by the
     * projections of its w > 0 vertices, extended to infinity along the direction
     * of every edge that crosses w = 0. */
    const real *x = s->x, *y = s->y, *w = s->w;
    if ((x[0] > w[0] && x[1] > w[1] && x[2] > w[2]) || (-x[0] > w[0] && -x[1] > w[1] && -x[2] > w[2]) ||
        (y[0] > w[0] && y[1] > w[1] && y[2] > w[2]) || (-y[0] > w[0] && -y[1] > w[1] && -y[2] > w[2]))
        return;
    if (!(w[0] > 0) && !(w[1] > 0) && !(w[2] > 0)) return;
    real mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY;
    for (int k = 0; k < 3; ++k) {
        int j = (k + 1) % 3;
        if (w[k] > 0) {
            real nx = x[k] / w[k], ny = y[k] / w[k];
            if (nx < mnx) mnx = nx;
            if (nx > mxx) mxx = nx;
            if (ny < mny) mny = ny;
            if (ny > mxy) mxy = ny;
        }
        if ((w[k] > 0) != (w[j] > 0)) {
            real t = w[k] / (w[k] - w[j]);
            real dx = x[k] + t * (x[j] - x[k]), dy = y[k] + t * (y[j] - y[k]);
            if (dx > 0) mxx = INFINITY;
            if (dx < 0) mnx = -INFINITY;
            if (dy > 0) mxy = INFINITY;
            if (dy < 0) mny = -INFINITY;
            if (dx == 0 && dy == 0) { mnx = mny = -INFINITY; mxx = mxy = INFINITY; }
        }
    }
    if (mnx < -2) mnx = -2;
    if (mxx > 2) mxx = 2;
    if (mny < -2) mny = -2;
    if (mxy > 2) mxy = 2;
    int c0 = (int)floor((mnx + 1) * (real)0.5 * (real)W) - 1;
    int c1 = (int)floor((mxx + 1) * (real)0.5 * (real)W) + 1;
    int r0 = (int)floor((mny + 1) * (real)0.5 * (real)H) - 1;
    int r1 = (int)floor((mxy + 1) * (real)0.5 * (real)H) + 1;
    if (c0 < 0) c0 = 0;
    if (r0 < 0) r0 = 0;
    if (c1 > W - 1) c1 = W - 1;
    if (r1 > H - 1) r1 = H - 1;
    if (c0 > c1 || r0 > r1) return;
    s->c0 = c0; s->c1 = c1; s->r0 = r0; s->r1 = r1;
    s->cls = 2;
}

/* ------------------------------------------------------------------------- */
/* per-pixel barycentric evaluation                                            */
/* ------------------------------------------------------------------------- */
typedef struct { real X[3], Y[3], e[3], s, zn, wn, px, py; } SFX(ora_pix);

static inline real SFX(pix_ndc)(int c, int W) { return (real)(2 * c + 1 - W) / (real)W; }

static void SFX(eval_pixel)(const SFX(ora_setup) *t, int r, int c, int H, int W, SFX(ora_pix) *p) {
    p->px = SFX(pix_ndc)(c, W);
    p->py = SFX(pix_ndc)(r, H);
    for (int k = 0; k < 3; ++k) {
        p->X[k] = FMA(-t->w[k], p->px, t->x[k]);
        p->Y[k] = FMA(-t->w[k], p->py, t->y[k]);
    }
    p->e[0] = FMA(p->X[1], p->Y[2], -(p->X[2] * p->Y[1]));
    p->e[1] = FMA(p->X[2], p->Y[0], -(p->X[0] * p->Y[2]));
    p->e[2] = FMA(p->X[0], p->Y[1], -(p->X[1] * p->Y[0]));
    p->s = (p->e[0] + p->e[1]) + p->e[2];
    p->zn = FMA(p->e[0], t->z[0], FMA(p->e[1], t->z[1], p->e[2] * t->z[2]));
    p->wn = FMA(p->e[0], t->w[0], FMA(p->e[1], t->w[1], p->e[2] * t->w[2]));
}

static int SFX(covers)(const SFX(ora_setup) *t, int r, int c, const SFX(ora_pix) *p) {
    if (t->cls == 1) {
        int64_t Px = 16 * (int64_t)c + 8, Py = 16 * (int64_t)r + 8;
        for (int k = 0; k < 3; ++k) {
            int64_t Xi = t->X[k], Yi = t->Y[k], Xj = t->X[(k + 1) % 3], Yj = t->Y[(k + 1) % 3];
            int64_t E = (Xj - Xi) * (Py - Yi) - (Yj - Yi) * (Px - Xi);
            if (E - (t->tl[k] ? 0 : 1) < 0) return 0;
        }
        return 1;
    }
    if (t->cls == 2) {
        if (p->wn > 0) return p->e[0] >= 0 && p->e[1] >= 0 && p->e[2] >= 0;
        if (p->wn < 0) return p->e[0] <= 0 && p->e[1] <= 0 && p->e[2] <= 0;
    }
    return 0;
}

static inline const real *SFX(img_pos)(const real *pos, int pos_batched, int img, int V) {
    return pos_batched ? pos + (int64_t)img * V * 4 : pos;
}

/* Deterministic parallel accumulation of the backward passes: the rows of an
 * image are split into K bands (K fixed by H and the accumulator size, not by
 * the thread count), each band sums its pixels' contributions, in row-major
 * order, into a zeroed buffer of its own, and the bands are then added to the
 * output in band order. */
static int SFX(n_bands)(int H, int64_t n_acc) {
    int64_t cap = ((int64_t)256 << 20) / (n_acc * (int64_t)sizeof(real) + 1);
    int64_t K = H < 16 ? H : 16;
    if (K > cap) K = cap;
    return K < 1 ? 1 : (int)K;
}
static void SFX(add_bands)(real *out, const real *bands, int K, int64_t n_acc) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_acc; ++i) {
        real s = out[i];
        for (int b = 0; b < K; ++b) s += bands[(int64_t)b * n_acc + i];
        out[i] = s;
    }
}

/* ------------------------------------------------------------------------- */
/* rasterize forward                                                           */
/* ------------------------------------------------------------------------- */
/* rast: [n, H, W, 4] (u, v, z/w, id); in f32 the id channel holds the codec
 * of (triangle + 1), in f64 it holds (triangle + 1) as a plain number.
 * prev (nullable): previous depth layer in the same format (peeling).
 * db (nullable): [n, H, W, 4] = (du/dX, du/dY, dv/dX, dv/dY), pixel units. */
/* Winner's (u, v, z/w, id) and rast_db at pixel (r, c). */
static void SFX(write_winner)(const real *P, int V, const int32_t *tri, int t, real zw, int r, int c, int H, int W,
                              real *o, real *D) {
    SFX(ora_setup) s;
    SFX(setup_tri)(P, V, tri + 3 * (int64_t)t, H, W, &s);
    SFX(ora_pix) p;
    SFX(eval_pixel)(&s, r, c, H, W, &p);
    real u = p.e[0] / p.s, v = p.e[1] / p.s;
    o[0] = u;
    o[1] = v;
    o[2] = zw;
#ifdef ORA_DOUBLE
    o[3] = (real)(t + 1);
#else
    o[3] = ora_triidx_to_float(t + 1);
#endif
    if (D) {
        real ex[3], ey[3];
        ex[0] = FMA(s.w[2], p.Y[1], -(s.w[1] * p.Y[2]));
        ex[1] = FMA(s.w[0], p.Y[2], -(s.w[2] * p.Y[0]));
        ex[2] = FMA(s.w[1], p.Y[0], -(s.w[0] * p.Y[1]));
        ey[0] = FMA(s.w[1], p.X[2], -(s.w[2] * p.X[1]));
        ey[1] = FMA(s.w[2], p.X[0], -(s.w[0] * p.X[2]));
        ey[2] = FMA(s.w[0], p.X[1], -(s.w[1] * p.X[0]));
        real sx = (ex[0] + ex[1]) + ex[2], sy = (ey[0] + ey[1]) + ey[2];
        real kx = (real)2 / (real)W, ky = (real)2 / (real)H;
        D[0] = (FMA(-u, sx, ex[0]) / p.s) * kx;
        D[1] = (FMA(-u, sy, ey[0]) / p.s) * ky;
        D[2] = (FMA(-v, sx, ex[1]) / p.s) * kx;
        D[3] = (FMA(-v, sy, ey[1]) / p.s) * ky;
    }
}

#ifndef ORA_DOUBLE
static inline float inv_ord_f32(uint32_t k) { return u2f((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k); }

static inline void atomic_min_u64(uint64_t *p, uint64_t v) {
    uint64_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
}

/* float: the depth test is the minimum of 64-bit keys (orderable(z/w) << 32 |
 * triangle), so triangles run in parallel with an atomic minimum and the
 * result does not depend on the order; then the winners, per pixel. */
int SFX(ora_rasterize)(const real *pos, int n_img, int V, int pos_batched, const int32_t *tri, int F,
                       const int32_t *ranges, int H, int W, const real *prev, real *rast, real *db) {
    int64_t npx = (int64_t)H * W;
    uint64_t *key = (uint64_t *)malloc(sizeof(uint64_t) * npx);
    if (!key) return -1;
    for (int img = 0; img < n_img; ++img) {
        const real *P = SFX(img_pos)(pos, pos_batched, img, V);
        int f0 = ranges ? ranges[2 * img] : 0;
        int f1 = ranges ? ranges[2 * img] + ranges[2 * img + 1] : F;
        if (f0 < 0) f0 = 0;
        if (f1 > F) f1 = F;
        const real *pv = prev ? prev + (int64_t)img * npx * 4 : NULL;
        for (int64_t i = 0; i < npx; ++i) key[i] = ~0ull;
#pragma omp parallel for schedule(dynamic, 4096)
        for (int t = f0; t < f1; ++t) {
            SFX(ora_setup) s;
            SFX(setup_tri)(P, V, tri + 3 * (int64_t)t, H, W, &s);
            if (!s.cls) continue;
            for (int r = s.r0; r <= s.r1; ++r)
                for (int c = s.c0; c <= s.c1; ++c) {
                    SFX(ora_pix) p;
                    SFX(eval_pixel)(&s, r, c, H, W, &p);
                    if (!SFX(covers)(&s, r, c, &p)) continue;
                    if (p.s == 0) continue;
                    real zw = p.zn / p.wn;
                    if (!(zw >= (real)-1 && zw <= (real)1)) continue;
                    int64_t pi = (int64_t)r * W + c;
                    uint64_t k = ((uint64_t)ord_f32(zw) << 32) | (uint32_t)t;
                    if (pv) {
                        float pid = pv[4 * pi + 3];
                        if (pid <= 0) continue;
                        uint64_t pkey = ((uint64_t)ord_f32(pv[4 * pi + 2]) << 32) | (uint32_t)(ora_float_to_triidx(pid) - 1);
                        if (k <= pkey) continue;
                    }
                    atomic_min_u64(&key[pi], k);
                }
        }
        real *R = rast + (int64_t)img * npx * 4;
        real *D = db ? db + (int64_t)img * npx * 4 : NULL;
#pragma omp parallel for schedule(static)
        for (int r = 0; r < H; ++r)
            for (int c = 0; c < W; ++c) {
                int64_t pi = (int64_t)r * W + c;
                real *o = R + 4 * pi;
                if (key[pi] == ~0ull) {
                    o[0] = o[1] = o[2] = o[3] = 0;
                    if (D) D[4 * pi] = D[4 * pi + 1] = D[4 * pi + 2] = D[4 * pi + 3] = 0;
                    continue;
                }
                int t = (int)(uint32_t)(key[pi] & 0xffffffffu);
                SFX(write_winner)(P, V, tri, t, inv_ord_f32((uint32_t)(key[pi] >> 32)), r, c, H, W, o,
                                  D ? D + 4 * pi : NULL);
            }
    }
    free(key);
    return 0;
}
#else
/* double: a sequential minimum over (z/w, triangle) */
int SFX(ora_rasterize)(const real *pos, int n_img, int V, int pos_batched, const int32_t *tri, int F,
                       const int32_t *ranges, int H, int W, const real *prev, real *rast, real *db) {
    int64_t npx = (int64_t)H * W;
    real *bz = (real *)malloc(sizeof(real) * npx);
    int32_t *bt = (int32_t *)malloc(sizeof(int32_t) * npx);
    if (!bz || !bt) { free(bz); free(bt); return -1; }
    for (int img = 0; img < n_img; ++img) {
        const real *P = SFX(img_pos)(pos, pos_batched, img, V);
        int f0 = ranges ? ranges[2 * img] : 0;
        int f1 = ranges ? ranges[2 * img] + ranges[2 * img + 1] : F;
        if (f0 < 0) f0 = 0;
        if (f1 > F) f1 = F;
        const real *pv = prev ? prev + (int64_t)img * npx * 4 : NULL;
        for (int64_t i = 0; i < npx; ++i) bt[i] = -1;
        for (int t = f0; t < f1; ++t) {
            SFX(ora_setup) s;
            SFX(setup_tri)(P, V, tri + 3 * (int64_t)t, H, W, &s);
            if (!s.cls) continue;
            for (int r = s.r0; r <= s.r1; ++r)
                for (int c = s.c0; c <= s.c1; ++c) {
                    SFX(ora_pix) p;
                    SFX(eval_pixel)(&s, r, c, H, W, &p);
                    if (!SFX(covers)(&s, r, c, &p)) continue;
                    if (p.s == 0) continue;
                    real zw = p.zn / p.wn;
                    if (!(zw >= (real)-1 && zw <= (real)1)) continue;
                    int64_t pi = (int64_t)r * W + c;
                    if (pv) {
                        real pid = pv[4 * pi + 3], pz = pv[4 * pi + 2];
                        if (pid <= 0) continue;
                        int32_t pt = (int32_t)pid - 1;
                        if (!(zw > pz || (zw == pz && t > pt))) continue;
                    }
                    if (bt[pi] < 0 || zw < bz[pi] || (zw == bz[pi] && t < bt[pi])) { bz[pi] = zw; bt[pi] = t; }
                }
        }
        real *R = rast + (int64_t)img * npx * 4;
        real *D = db ? db + (int64_t)img * npx * 4 : NULL;
        for (int r = 0; r < H; ++r)
            for (int c = 0; c < W; ++c) {
                int64_t pi = (int64_t)r * W + c;
                real *o = R + 4 * pi;
                if (bt[pi] < 0) {
                    o[0] = o[1] = o[2] = o[3] = 0;
                    if (D) D[4 * pi] = D[4 * pi + 1] = D[4 * pi + 2] = D[4 * pi + 3] = 0;
                    continue;
                }
                SFX(write_winner)(P, V, tri, bt[pi], bz[pi], r, c, H, W, o, D ? D + 4 * pi : NULL);
            }
    }
    free(bz);
    free(bt);
    return 0;
}
#endif

static inline int32_t SFX(decode_id)(real x) {
#ifdef ORA_DOUBLE
    return (int32_t)x;
#else
    return ora_float_to_triidx(x);
#endif
}

/* Re-evaluate (u, v, z/w) and db at the pixels of a fixed id buffer: the
 * rasterisation as a smooth function of pos with visibility held constant
 * (used for finite differences). rast_in ids are read, rast_out/db written. */
int SFX(ora_rasterize_fixed)(const real *pos, int n_img, int V, int pos_batched, const int32_t *tri, int F,
                             int H, int W, const real *rast_in, real *rast_out, real *db) {
    int64_t npx = (int64_t)H * W;
    for (int img = 0; img < n_img; ++img) {
        const real *P = SFX(img_pos)(pos, pos_batched, img, V);
#pragma omp parallel for schedule(static)
        for (int r = 0; r < H; ++r)
            for (int c = 0; c < W; ++c) {
                int64_t pi = (int64_t)img * npx + (int64_t)r * W + c;
                const real *ri = rast_in + 4 * pi;
                real *o = rast_out + 4 * pi;
                int32_t id = ri[3] > 0 ? SFX(decode_id)(ri[3]) : 0;
                if (id <= 0 || id > F) {
                    o[0] = o[1] = o[2] = o[3] = 0;
                    if (db) db[4 * pi] = db[4 * pi + 1] = db[4 * pi + 2] = db[4 * pi + 3] = 0;
                    continue;
                }
                SFX(ora_setup) s;
                SFX(setup_tri)(P, V, tri + 3 * (int64_t)(id - 1), H, W, &s);
                SFX(ora_pix) p;
                SFX(eval_pixel)(&s, r, c, H, W, &p);
                real u = p.e[0] / p.s, v = p.e[1] / p.s;
                o[0] = u; o[1] = v; o[2] = p.zn / p.wn; o[3] = ri[3];
                if (db) {
                    real ex[3], ey[3];
                    ex[0] = FMA(s.w[2], p.Y[1], -(s.w[1] * p.Y[2]));
                    ex[1] = FMA(s.w[0], p.Y[2], -(s.w[2] * p.Y[0]));
                    ex[2] = FMA(s.w[1], p.Y[0], -(s.w[0] * p.Y[1]));
                    ey[0] = FMA(s.w[1], p.X[2], -(s.w[2] * p.X[1]));
                    ey[1] = FMA(s.w[2], p.X[0], -(s.w[0] * p.X[2]));
                    ey[2] = FMA(s.w[0], p.X[1], -(s.w[1] * p.X[0]));
                    real sx = (ex[0] + ex[1]) + ex[2], sy = (ey[0] + ey[1]) + ey[2];
                    real kx = (real)2 / (real)W, ky = (real)2 / (real)H;
                    db[4 * pi + 0] = (FMA(-u, sx, ex[0]) / p.s) * kx;
                    db[4 * pi + 1] = (FMA(-u, sy, ey[0]) / p.s) * ky;
                    db[4 * pi + 2] = (FMA(-v, sx, ex[1]) / p.s) * kx;
                    db[4 * pi + 3] = (FMA(-v, sy, ey[1]) / p.s) * ky;
                }
            }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* rasterize backward                                                          */
/* ------------------------------------------------------------------------- */
/* grad_pos accumulates (+=): [n, V, 4] when pos_batched else [V, 4].
 * grad_rast channels (u, v, z/w) are used; grad_db (nullable) adds the
 * derivative of rast_db.  The per-pixel chain runs in double precision in
 * both variants (pix_grad_d): thin triangles seen edge-on lose ~|e|/s of the
 * relative precision of their derivatives in float. */
int SFX(ora_rasterize_bwd)(const real *pos, int n_img, int V, int pos_batched, const int32_t *tri, int F,
                           int H, int W, const real *rast, const real *grad_rast, const real *grad_db,
                           real *grad_pos) {
    int64_t npx = (int64_t)H * W;
    const real kx = (real)2 / (real)W, ky = (real)2 / (real)H;
    const int64_t n_acc = (int64_t)V * 4;
    const int K = SFX(n_bands)(H, n_acc);
    real *bands = (real *)malloc(sizeof(real) * (size_t)(K * n_acc + 1));
    if (!bands) return -1;
    for (int img = 0; img < n_img; ++img) {
        const real *P = SFX(img_pos)(pos, pos_batched, img, V);
        real *G = pos_batched ? grad_pos + (int64_t)img * V * 4 : grad_pos;
        memset(bands, 0, sizeof(real) * (size_t)(K * n_acc));
#pragma omp parallel for schedule(dynamic, 1)
        for (int band = 0; band < K; ++band) {
            real *B = bands + (int64_t)band * n_acc;
            for (int r = (int)((int64_t)band * H / K); r < (int)((int64_t)(band + 1) * H / K); ++r)
                for (int c = 0; c < W; ++c) {
                    int64_t pi = (int64_t)img * npx + (int64_t)r * W + c;
                    int32_t id = rast[4 * pi + 3] > 0 ? SFX(decode_id)(rast[4 * pi + 3]) : 0;
                    if (id <= 0 || id > F) continue;
                    const int32_t *t3 = tri + 3 * (int64_t)(id - 1);
                    if (t3[0] < 0 || t3[0] >= V || t3[1] < 0 || t3[1] >= V || t3[2] < 0 || t3[2] >= V) continue;
                    double cx[3], cy[3], cz[3], cw[3];
                    for (int k = 0; k < 3; ++k) {
                        const real *pv = P + 4 * (int64_t)t3[k];
                        cx[k] = pv[0]; cy[k] = pv[1]; cz[k] = pv[2]; cw[k] = pv[3];
                    }
                    double gd[4], go[3][4];
                    if (grad_db) {
                        const real *q = grad_db + 4 * pi;
                        gd[0] = (double)(q[0] * kx); gd[1] = (double)(q[1] * ky);
                        gd[2] = (double)(q[2] * kx); gd[3] = (double)(q[3] * ky);
                    }
                    pix_grad_d(cx, cy, cz, cw, r, c, H, W, grad_rast[4 * pi + 0], grad_rast[4 * pi + 1],
                               grad_rast[4 * pi + 2], grad_db ? gd : NULL, go);
                    for (int k = 0; k < 3; ++k) {
                        real *g = B + 4 * (int64_t)t3[k];
                        for (int j = 0; j < 4; ++j) g[j] += (real)go[k][j];
                    }
                }
        }
        SFX(add_bands)(G, bands, K, n_acc);
    }
    free(bands);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* interpolate                                                                 */
/* ------------------------------------------------------------------------- */
/* attr: [V, C] (attr_batched=0) or [n, V, C]; out [n, H, W, C];
 * diff_idx (nullable, n_diff entries) with db: out_da [n, H, W, 2 n_diff]
 * = (dA/dX, dA/dY) for each selected channel.  Pixel formula restates
 * interpolator.py:43-45: out = (a0-a2) u + ((a1-a2) v + a2). */
int SFX(ora_interpolate)(const real *attr, int n_img, int V, int attr_batched, int C, const real *rast,
                         int H, int W, const int32_t *tri, int F, const real *db, const int32_t *diff_idx,
                         int n_diff, real *out, real *out_da) {
    int64_t npx = (int64_t)H * W;
    for (int img = 0; img < n_img; ++img) {
        const real *A = attr_batched ? attr + (int64_t)img * V * C : attr;
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < npx; ++q) {
            int64_t pi = (int64_t)img * npx + q;
            real *o = out + pi * C;
            real *od = out_da ? out_da + pi * 2 * n_diff : NULL;
            int32_t id = rast[4 * pi + 3] > 0 ? SFX(decode_id)(rast[4 * pi + 3]) : 0;
            const int32_t *t3 = (id > 0 && id <= F) ? tri + 3 * (int64_t)(id - 1) : NULL;
            if (t3 && (t3[0] < 0 || t3[0] >= V || t3[1] < 0 || t3[1] >= V || t3[2] < 0 || t3[2] >= V)) t3 = NULL;
            if (!t3) {
                for (int ch = 0; ch < C; ++ch) o[ch] = 0;
                if (od) for (int j = 0; j < 2 * n_diff; ++j) od[j] = 0;
                continue;
            }
            real u = rast[4 * pi], v = rast[4 * pi + 1];
            const real *a0 = A + (int64_t)t3[0] * C, *a1 = A + (int64_t)t3[1] * C, *a2 = A + (int64_t)t3[2] * C;
            for (int ch = 0; ch < C; ++ch) {
                real d1 = a1[ch] - a2[ch], d0 = a0[ch] - a2[ch];
                o[ch] = FMA(d0, u, FMA(d1, v, a2[ch]));
            }
            if (od) {
                const real *d = db + 4 * pi;
                for (int j = 0; j < n_diff; ++j) {
                    int ch = diff_idx[j];
                    real d0 = a0[ch] - a2[ch], d1 = a1[ch] - a2[ch];
                    od[2 * j + 0] = FMA(d0, d[0], d1 * d[2]);
                    od[2 * j + 1] = FMA(d0, d[1], d1 * d[3]);
                }
            }
        }
    }
    return 0;
}

/* grad_attr accumulates (+=) with attr's layout; grad_rast / grad_db are
 * written (z/w and id channels zero). */
int SFX(ora_interpolate_bwd)(const real *attr, int n_img, int V, int attr_batched, int C, const real *rast,
                             int H, int W, const int32_t *tri, int F, const real *db, const int32_t *diff_idx,
                             int n_diff, const real *grad_out, const real *grad_da, real *grad_attr,
                             real *grad_rast, real *grad_db) {
    int64_t npx = (int64_t)H * W;
    const int64_t n_acc = (int64_t)V * C;
    const int K = SFX(n_bands)(H, n_acc);
    real *bands = (real *)malloc(sizeof(real) * (size_t)(K * n_acc + 1));
    if (!bands) return -1;
    for (int img = 0; img < n_img; ++img) {
        const real *A = attr_batched ? attr + (int64_t)img * V * C : attr;
        real *GA = attr_batched ? grad_attr + (int64_t)img * V * C : grad_attr;
        memset(bands, 0, sizeof(real) * (size_t)(K * n_acc));
#pragma omp parallel for schedule(dynamic, 1)
        for (int band = 0; band < K; ++band) {
            real *B = bands + (int64_t)band * n_acc;
            for (int64_t q = (int64_t)band * H / K * W; q < (int64_t)(band + 1) * H / K * W; ++q) {
                int64_t pi = (int64_t)img * npx + q;
                real *gr = grad_rast + 4 * pi;
                gr[0] = gr[1] = gr[2] = gr[3] = 0;
                if (grad_db) grad_db[4 * pi] = grad_db[4 * pi + 1] = grad_db[4 * pi + 2] = grad_db[4 * pi + 3] = 0;
                int32_t id = rast[4 * pi + 3] > 0 ? SFX(decode_id)(rast[4 * pi + 3]) : 0;
                const int32_t *t3 = (id > 0 && id <= F) ? tri + 3 * (int64_t)(id - 1) : NULL;
                if (t3 && (t3[0] < 0 || t3[0] >= V || t3[1] < 0 || t3[1] >= V || t3[2] < 0 || t3[2] >= V)) t3 = NULL;
                if (!t3) continue;
                real u = rast[4 * pi], v = rast[4 * pi + 1];
                real w2 = ((real)1 - u) - v;
                const real *a0 = A + (int64_t)t3[0] * C, *a1 = A + (int64_t)t3[1] * C, *a2 = A + (int64_t)t3[2] * C;
                real *g0 = B + (int64_t)t3[0] * C, *g1 = B + (int64_t)t3[1] * C, *g2 = B + (int64_t)t3[2] * C;
                const real *g = grad_out + pi * C;
                real su = 0, sv = 0;
                for (int ch = 0; ch < C; ++ch) {
                    g0[ch] += u * g[ch];
                    g1[ch] += v * g[ch];
                    g2[ch] += w2 * g[ch];
                    su += g[ch] * (a0[ch] - a2[ch]);
                    sv += g[ch] * (a1[ch] - a2[ch]);
                }
                gr[0] = su;
                gr[1] = sv;
                if (grad_da && diff_idx && n_diff > 0) {
                    const real *d = db + 4 * pi;
                    const real *gd = grad_da + pi * 2 * n_diff;
                    real s0 = 0, s1 = 0, s2 = 0, s3 = 0;
                    for (int j = 0; j < n_diff; ++j) {
                        int ch = diff_idx[j];
                        real gx = gd[2 * j], gy = gd[2 * j + 1];
                        real d0 = a0[ch] - a2[ch], d1 = a1[ch] - a2[ch];
                        real k0 = gx * d[0] + gy * d[1], k1 = gx * d[2] + gy * d[3];
                        g0[ch] += k0;
                        g1[ch] += k1;
                        g2[ch] -= k0 + k1;
                        s0 += gx * d0; s1 += gy * d0; s2 += gx * d1; s3 += gy * d1;
                    }
                    if (grad_db) {
                        grad_db[4 * pi + 0] = s0; grad_db[4 * pi + 1] = s1;
                        grad_db[4 * pi + 2] = s2; grad_db[4 * pi + 3] = s3;
                    }
                }
            }
        }
        SFX(add_bands)(GA, bands, K, n_acc);
    }
    free(bands);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* antialias                                                                   */
/* ------------------------------------------------------------------------- */
/* Analytic silhouette antialiasing after Laine et al. 2020 ("Modular
 * Primitives for High-Performance Differentiable Rendering", sec. 3.3).
 * For a horizontal or vertical pair of pixel centres (a, b) with different
 * triangle ids, the candidates are the triangles at a and at b, nearer z/w
 * first (background last).  For candidate T seen at pixel `from`, the segment
 * from -> to leaves T through the edge with the smallest crossing parameter
 * t_e = f(from) / (f(from) - f(to)) among edges whose half-plane excludes `to`.
 * If that edge is a silhouette (no neighbour across it, or the neighbour has
 * the opposite screen orientation) the pair blends: with t clamped to [0,1],
 * t < 1/2 moves `from` towards `to` by (1/2 - t), otherwise `to` towards
 * `from` by (t - 1/2).  Otherwise the next candidate is tried. */
typedef struct {
    int recv;       /* -1 none, 0 = a receives, 1 = b receives */
    real wgt;
    int tri, edge;  /* triangle and edge (k -> k+1) that produced the blend  */
    int from_a;     /* the candidate triangle was seen at a                  */
    double fa, fb;  /* edge function at from / to                             */
    int clamped;    /* t was clamped: no position gradient                   */
} SFX(ora_pair);

/* Pair geometry (projection, orientation, edge functions, crossing) is
 * evaluated in double from the clip coordinates in both variants: a float clip
 * coordinate fixes a vertex to ~W * 2^-25 pixel, and the differences of
 * nearby pixel positions that the crossing needs keep few of those digits. */
static int SFX(proj_tri)(const real *P, int V, const int32_t *t3, int H, int W, double X[3], double Y[3]) {
    for (int k = 0; k < 3; ++k) {
        int32_t vi = t3[k];
        if (vi < 0 || vi >= V) return 0;
        double x = P[4 * (int64_t)vi], y = P[4 * (int64_t)vi + 1], w = P[4 * (int64_t)vi + 3];
        if (!(w > 0)) return 0;
        X[k] = (x / w + 1.0) * ((double)W * 0.5);
        Y[k] = (y / w + 1.0) * ((double)H * 0.5);
    }
    return 1;
}

/* sign of det[[x0,x1,x2],[y0,y1,y2],[w0,w1,w2]] (screen orientation for w > 0) */
static int SFX(orient)(const real *P, int V, const int32_t *t3) {
    double a[3][3];
    for (int k = 0; k < 3; ++k) {
        int32_t vi = t3[k];
        if (vi < 0 || vi >= V) return 0;
        a[k][0] = P[4 * (int64_t)vi]; a[k][1] = P[4 * (int64_t)vi + 1]; a[k][2] = P[4 * (int64_t)vi + 3];
    }
    double m0 = fma(a[1][1], a[2][2], -(a[2][1] * a[1][2]));
    double m1 = fma(a[0][1], a[2][2], -(a[2][1] * a[0][2]));
    double m2 = fma(a[0][1], a[1][2], -(a[1][1] * a[0][2]));
    double d = fma(a[0][0], m0, fma(-a[1][0], m1, a[2][0] * m2));
    return d > 0 ? 1 : (d < 0 ? -1 : 0);
}

static void SFX(eval_pair)(const real *P, int V, const int32_t *tri, int F, const int32_t *opp,
                           const real *rast_img, int H, int W, int ra, int ca, int rb, int cb, SFX(ora_pair) *res) {
    res->recv = -1;
    int64_t ia = (int64_t)ra * W + ca, ib = (int64_t)rb * W + cb;
    int32_t ida = rast_img[4 * ia + 3] > 0 ? SFX(decode_id)(rast_img[4 * ia + 3]) : 0;
    int32_t idb = rast_img[4 * ib + 3] > 0 ? SFX(decode_id)(rast_img[4 * ib + 3]) : 0;
    if (ida > F) ida = 0;
    if (idb > F) idb = 0;
    if (ida == idb) return;
    int order_a_first = 1;
    if (ida == 0) order_a_first = 0;
    else if (idb != 0 && rast_img[4 * ib + 2] < rast_img[4 * ia + 2]) order_a_first = 0;
    for (int pass = 0; pass < 2; ++pass) {
        int use_a = (pass == 0) ? order_a_first : !order_a_first;
        int32_t id = use_a ? ida : idb;
        if (id == 0) continue;
        int t = id - 1;
        const int32_t *t3 = tri + 3 * (int64_t)t;
        double X[3], Y[3];
        if (!SFX(proj_tri)(P, V, t3, H, W, X, Y)) continue;
        int o = SFX(orient)(P, V, t3);
        if (o == 0) continue;
        double fx = (double)(use_a ? ca : cb) + 0.5, fy = (double)(use_a ? ra : rb) + 0.5;
        double tx = (double)(use_a ? cb : ca) + 0.5, ty = (double)(use_a ? rb : ra) + 0.5;
        int best = -1;
        double bt = 0, bfa = 0, bfb = 0;
        for (int k = 0; k < 3; ++k) {
            int j = (k + 1) % 3;
            double ex = X[j] - X[k], ey = Y[j] - Y[k];
            double ff = fma(ex, fy - Y[k], -(ey * (fx - X[k])));
            double ft = fma(ex, ty - Y[k], -(ey * (tx - X[k])));
            if (o < 0) { ff = -ff; ft = -ft; }
            if (!(ft < 0)) continue;
            double te = ff / (ff - ft);
            if (best < 0 || te < bt) { best = k; bt = te; bfa = ff; bfb = ft; }
        }
        if (best < 0) continue;
        int32_t n = opp[3 * t + best];
        if (n >= 0 && n < F && SFX(orient)(P, V, tri + 3 * (int64_t)n) == o) continue;
        int clamped = 0;
        if (bt < 0) { bt = 0; clamped = 1; }
        if (bt > 1) { bt = 1; clamped = 1; }
        res->tri = t;
        res->edge = best;
        res->from_a = use_a;
        res->fa = bfa;
        res->fb = bfb;
        res->clamped = clamped;
        if (bt < 0.5) { res->recv = use_a ? 0 : 1; res->wgt = (real)(0.5 - bt); }
        else { res->recv = use_a ? 1 : 0; res->wgt = (real)(bt - 0.5); }
        return;
    }
}

/* color [n,H,W,C] -> out [n,H,W,C].  Pixel p's output sums, in this order,
 * the pairs (left, p), (p, right), (below, p), (p, above). */
int SFX(ora_antialias)(const real *color, int n_img, int H, int W, int C, const real *rast, const real *pos,
                       int V, int pos_batched, const int32_t *tri, int F, const int32_t *opp, real *out) {
    int64_t npx = (int64_t)H * W;
    for (int img = 0; img < n_img; ++img) {
        const real *P = SFX(img_pos)(pos, pos_batched, img, V);
        const real *R = rast + (int64_t)img * npx * 4;
        const real *Cc = color + (int64_t)img * npx * C;
        real *O = out + (int64_t)img * npx * C;
#pragma omp parallel for schedule(static)
        for (int r = 0; r < H; ++r)
            for (int c = 0; c < W; ++c) {
                int64_t q = (int64_t)r * W + c;
                real *o = O + q * C;
                for (int ch = 0; ch < C; ++ch) o[ch] = Cc[q * C + ch];
                int nb[4][4] = {{r, c - 1, r, c}, {r, c, r, c + 1}, {r - 1, c, r, c}, {r, c, r + 1, c}};
                for (int k = 0; k < 4; ++k) {
                    int ra = nb[k][0], ca = nb[k][1], rb = nb[k][2], cb = nb[k][3];
                    if (ra < 0 || ca < 0 || rb >= H || cb >= W) continue;
                    SFX(ora_pair) pr;
                    SFX(eval_pair)(P, V, tri, F, opp, R, H, W, ra, ca, rb, cb, &pr);
                    if (pr.recv < 0) continue;
                    int self_is_a = (ra == r && ca == c);
                    if ((pr.recv == 0) != self_is_a) continue;
                    int64_t qs = self_is_a ? (int64_t)rb * W + cb : (int64_t)ra * W + ca;
                    for (int ch = 0; ch < C; ++ch) {
                        real diff = Cc[qs * C + ch] - Cc[q * C + ch];
                        o[ch] = FMA(pr.wgt, diff, o[ch]);
                    }
                }
            }
    }
    return 0;
}

/* grad_color written; grad_pos accumulated (+=) with pos's layout.
 * Pass 1 evaluates the pairs to the right of and above every pixel (bands of
 * rows in parallel, position gradients into per-band buffers) and keeps each
 * pair's decision; pass 2 forms every pixel's colour gradient from the four
 * pairs it belongs to, in the order (left, p), (p, right), (below, p),
 * (p, above). */
int SFX(ora_antialias_bwd)(const real *color, int n_img, int H, int W, int C, const real *rast, const real *pos,
                           int V, int pos_batched, const int32_t *tri, int F, const int32_t *opp,
                           const real *grad_out, real *grad_color, real *grad_pos, real boost) {
    int64_t npx = (int64_t)H * W;
    const int64_t n_acc = (int64_t)V * 4;
    const int K = SFX(n_bands)(H, n_acc);
    real *bands = (real *)malloc(sizeof(real) * (size_t)(K * n_acc + 1));
    signed char *prc = (signed char *)malloc((size_t)(2 * npx + 1));
    real *pw = (real *)malloc(sizeof(real) * (size_t)(2 * npx + 1));
    if (!bands || !prc || !pw) { free(bands); free(prc); free(pw); return -1; }
    for (int img = 0; img < n_img; ++img) {
        const real *P = SFX(img_pos)(pos, pos_batched, img, V);
        real *GP = pos_batched ? grad_pos + (int64_t)img * V * 4 : grad_pos;
        const real *R = rast + (int64_t)img * npx * 4;
        const real *Cc = color + (int64_t)img * npx * C;
        const real *Go = grad_out + (int64_t)img * npx * C;
        real *Gc = grad_color + (int64_t)img * npx * C;
        memset(bands, 0, sizeof(real) * (size_t)(K * n_acc));
#pragma omp parallel for schedule(dynamic, 1)
        for (int band = 0; band < K; ++band) {
            real *GB = bands + (int64_t)band * n_acc;
            const int rlo = (int)((int64_t)band * H / K), rhi = (int)((int64_t)(band + 1) * H / K);
            for (int dir = 0; dir < 2; ++dir)
                for (int ra = rlo; ra < rhi; ++ra)
                    for (int ca = 0; ca < W; ++ca) {
                        int rb = dir ? ra + 1 : ra, cb = dir ? ca : ca + 1;
                        int64_t qa = (int64_t)ra * W + ca, qb = (int64_t)rb * W + cb;
                        prc[dir * npx + qa] = -1;
                        if (rb >= H || cb >= W) continue;
                        SFX(ora_pair) pr;
                        SFX(eval_pair)(P, V, tri, F, opp, R, H, W, ra, ca, rb, cb, &pr);
                        if (pr.recv < 0) continue;
                        prc[dir * npx + qa] = (signed char)pr.recv;
                        pw[dir * npx + qa] = pr.wgt;
                        int64_t qr = pr.recv == 0 ? qa : qb, qs = pr.recv == 0 ? qb : qa;
                        double dLdw = 0;
                        for (int ch = 0; ch < C; ++ch)
                            dLdw += (double)Go[qr * C + ch] * (double)(Cc[qs * C + ch] - Cc[qr * C + ch]);
                        if (pr.clamped) continue;
                        /* The decisions above are the forward's (in `real`); the
                         * derivative of the crossing t = fa / (fa - fb) is taken in
                         * double from the clip coordinates in both variants: it is
                         * ill-conditioned for edges nearly parallel to the pair. */
                        /* w = 1/2 - t when the receiver is `from`, t - 1/2 otherwise */
                        int recv_is_from = (pr.recv == 0) == (pr.from_a != 0);
                        double dLdt = recv_is_from ? -dLdw : dLdw;
                        const int32_t *t3 = tri + 3 * (int64_t)pr.tri;
                        int o = SFX(orient)(P, V, t3);
                        double so = o < 0 ? -1.0 : 1.0;
                        int i = pr.edge, j = (pr.edge + 1) % 3;
                        int vv[2] = {t3[i], t3[j]};
                        double hw = (double)W * 0.5, hh = (double)H * 0.5;
                        double cx[2], cy[2], cw[2], qx[2], qy[2];
                        for (int m = 0; m < 2; ++m) {
                            const real *pv = P + 4 * (int64_t)vv[m];
                            cx[m] = pv[0]; cy[m] = pv[1]; cw[m] = pv[3];
                            qx[m] = (cx[m] / cw[m] + 1.0) * hw;
                            qy[m] = (cy[m] / cw[m] + 1.0) * hh;
                        }
                        double fxp = (double)(pr.from_a ? ca : cb) + 0.5, fyp = (double)(pr.from_a ? ra : rb) + 0.5;
                        double txp = (double)(pr.from_a ? cb : ca) + 0.5, typ = (double)(pr.from_a ? rb : ra) + 0.5;
                        double ex = qx[1] - qx[0], ey = qy[1] - qy[0];
                        double fa = so * (ex * (fyp - qy[0]) - ey * (fxp - qx[0]));
                        double fb = so * (ex * (typ - qy[0]) - ey * (txp - qx[0]));
                        double den = fa - fb;
                        double dt_dfa = -fb / (den * den), dt_dfb = fa / (den * den);
                        /* d f / d(Xi, Yi, Xj, Yj) at P: so*(Yj-PY), so*(PX-Xj), so*(PY-Yi), so*(Xi-PX) */
                        double k = dLdt * so * (double)boost;
                        double gx2[2], gy2[2];
                        gx2[0] = k * (dt_dfa * (qy[1] - fyp) + dt_dfb * (qy[1] - typ));
                        gy2[0] = k * (dt_dfa * (fxp - qx[1]) + dt_dfb * (txp - qx[1]));
                        gx2[1] = k * (dt_dfa * (fyp - qy[0]) + dt_dfb * (typ - qy[0]));
                        gy2[1] = k * (dt_dfa * (qx[0] - fxp) + dt_dfb * (qx[0] - txp));
                        for (int m = 0; m < 2; ++m) {
                            real *g = GB + 4 * (int64_t)vv[m];
                            double gx = gx2[m] * hw / cw[m], gy = gy2[m] * hh / cw[m];
                            g[0] += (real)gx;
                            g[1] += (real)gy;
                            g[3] += (real)(-(gx * cx[m] + gy * cy[m]) / cw[m]);
                        }
                    }
        }
        SFX(add_bands)(GP, bands, K, n_acc);
#pragma omp parallel for schedule(static)
        for (int r = 0; r < H; ++r)
            for (int c = 0; c < W; ++c) {
                int64_t q = (int64_t)r * W + c;
                for (int ch = 0; ch < C; ++ch) Gc[q * C + ch] = Go[q * C + ch];
                /* (left, p), (p, right), (below, p), (p, above): pair index (dir, first pixel) */
                int64_t first[4] = {q - 1, q, q - W, q};
                int dirs[4] = {0, 0, 1, 1};
                int ok[4] = {c > 0, c + 1 < W, r > 0, r + 1 < H};
                for (int k = 0; k < 4; ++k) {
                    if (!ok[k]) continue;
                    int64_t qa = first[k], qb = dirs[k] ? qa + W : qa + 1;
                    int rc = prc[dirs[k] * npx + qa];
                    if (rc < 0) continue;
                    real w = pw[dirs[k] * npx + qa];
                    int64_t qr = rc == 0 ? qa : qb;
                    if (q == qr)
                        for (int ch = 0; ch < C; ++ch) Gc[q * C + ch] -= w * Go[qr * C + ch];
                    else
                        for (int ch = 0; ch < C; ++ch) Gc[q * C + ch] += w * Go[qr * C + ch];
                }
            }
    }
    free(bands);
    free(prc);
    free(pw);
    return 0;
}
