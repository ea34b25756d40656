/*
 * drp.h -- C-ABI of the MI355X (gfx950) differentiable rasterisation path.
 *
 * Plain pointers and sizes only.  Every pointer except `ranges` host copies is
 * device memory owned by the caller; work is enqueued on `stream` (a
 * hipStream_t) with no implicit synchronisation and no device-memory
 * allocation, so calls can be captured into a HIP graph.  Scratch memory is a
 * caller-provided workspace sized by the matching *_workspace_size query.
 * Independent kernels of one call (the antialias backward's two passes, the
 * fused backward's two vertex gathers) run beside each other on one auxiliary
 * stream per device, forked from and joined back to `stream` with events
 * (a fork / join node pair under capture); the stream is created on the
 * first call on a device that is not under capture and kept until the
 * process ends, and without it the kernels run on `stream` in order.
 *
 * Return value: 0 on success, < 0 on error (message: drp_last_error(), which
 * is thread-local).
 *
 * Layouts (float32 unless noted, all contiguous):
 *   pos   [V,4] clip space, or [n,V,4] when pos_batched (one buffer per image)
 *   tri   [F,3] int32 vertex indices
 *   rast  [n,H,W,4] = (u, v, z/w, id), bottom row first; id is the float
 *         encoding of (triangle index + 1), 0 = empty
 *   rast_db [n,H,W,4] = (du/dX, du/dY, dv/dX, dv/dY) in pixel units
 *   attr  [V,C] or [n,V,C] (attr_batched); out [n,H,W,C]
 *
 * Each entry point replaces an operator of the nvdiffrast.torch module that
 * diffrp calls; citations are the diffrp call sites (paths relative to the
 * diffrp source tree).
 */
#ifndef DRP_H
#define DRP_H

#include <stddef.h>
#include <stdint.h>

#define DRP_ABI_VERSION 1

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed call on this thread. */
const char *drp_last_error(void);
int drp_abi_version(void);

/* ---- rasterize -----------------------------------------------------------
 * Replaces dr.DepthPeeler(ctx, pos, tri, (h, w), ranges).rasterize_next_layer()
 * (diffrp/rendering/surface_deferred.py:259-265) and dr.rasterize; the context
 * object dr.RasterizeCudaContext (surface_deferred.py:219-226) has no state
 * here beyond the caller's workspace.
 *   ranges_dev: nullptr (all triangles for every image; pos 3-D = instanced),
 *     or device int32 [n,3] = (first triangle, count, spill-list offset) for
 *     range mode (pos 2-D); spill-list offsets are prefix sums of the counts and
 *     list_total their sum; max_range the largest count.
 *   prev: previous depth layer (depth peeling) or nullptr.
 *   rast_db: nullptr unless derivatives are wanted.
 *   tile_w x tile_h in {8x8, 16x8, 16x16, 32x8}; bin_cap >= 1 (performance
 *   only: full bins spill to a per-image list, a full list to a full scan;
 *   above 4 x tile_w x tile_h the fine raster also culls triangles that lie
 *   behind a tile's current depths, with the same results).
 *   boxes (nullable, int32 [n,F,3], 4-byte aligned): per (image, triangle) the
 *   packed pixel box (c0 | c1 << 16, r0 | r1 << 16), c0 > c1 when culled, and a
 *   hit flag, nonzero when the triangle won at least one pixel of the image;
 *   the triangle-centric backward passes read both.  In range mode the layout
 *   is [list_total,3], image i's range at its spill-list offset. */
size_t drp_rasterize_workspace_size(int n_img, int H, int W, long long list_total, int tile_w, int tile_h,
                                    int bin_cap);
int drp_rasterize_fwd(void *stream, const float *pos, int n_img, int V, int pos_batched, const int32_t *tri, int F,
                      const int32_t *ranges_dev, int H, int W, const float *prev, float *rast, float *rast_db,
                      void *workspace, size_t workspace_bytes, int tile_w, int tile_h, int bin_cap,
                      long long list_total, int max_range, int32_t *boxes);

/* ---- rasterize without allocation --------------------------------------
 * No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  drp_rasterize_fwd with the same arguments except the
 * workspace: the scratch is a device array of DRP_NOALLOC_SCRATCH_BYTES
 * inside the library, one per GPU, held until the process ends; calls whose
 * drp_rasterize_workspace_size exceeds it fail (-3) with a message.  Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU (two
 * streams, two threads) race.  The first call on a device looks the array's
 * address up (make that first call outside graph capture); the device is the
 * current one.  Results equal drp_rasterize_fwd's bit for bit. */
#define DRP_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int drp_rasterize_fwd_noalloc(void *stream, const float *pos, int n_img, int V, int pos_batched,
                              const int32_t *tri, int F, const int32_t *ranges_dev, int H, int W, const float *prev,
                              float *rast, float *rast_db, int tile_w, int tile_h, int bin_cap, long long list_total,
                              int max_range, int32_t *boxes);

/* ---- uneven batches ------------------------------------------------------
 * Many meshes of different triangle counts, each at its own resolution, in
 * one rasterizer launch (the batched form of the DepthPeeler calls at
 * diffrp/rendering/surface_deferred.py:259-265, one per mesh).  No nvdiffrast
 * counterpart: nvdiffrast's range mode shares one resolution.
 *   items (host int32 [n,5]) = (V_i, F_i, H_i, W_i, bin capacity of the
 *   item's tiles, 0 = from its density); pos = the items' [V_i,4]
 *   clip-space vertices concatenated, tri = their [F_i,3] triangles
 *   concatenated with item-local vertex indices.
 *   drp_batch_plan fills a host plan of drp_batch_plan_size bytes for a tile
 *   shape; the caller copies it unchanged to 16-byte aligned device memory
 *   (plan_dev) and passes both copies.
 *   rast [n_pix,4]: item i's H_i x W_i image at the sum of the previous items'
 *   pixel counts, bottom row first; with global_ids = 0 the ids are item-local
 *   (triangle - first triangle of the item + 1), so each image is exactly what
 *   rasterizing the item alone gives, and the per-item interpolate /
 *   antialias / backward calls apply to its slices; with global_ids = 1 they
 *   are the batch triangle + 1, the layout the drp_batch_* passes below take.
 *   prev (nullable) is the previous layer in the same layout (depth peeling).
 *   boxes (nullable) int32 [n_tri,3] as for drp_rasterize_fwd, indexed by the
 *   batch triangle. */
size_t drp_batch_plan_size(const int32_t *items, int n_img, int tile_w, int tile_h);
int drp_batch_plan(const int32_t *items, int n_img, int tile_w, int tile_h, void *plan, size_t plan_bytes);
int drp_batch_plan_info(const void *plan, long long *n_pix, long long *n_tiles, long long *n_tri, long long *n_vert);
size_t drp_rasterize_batch_workspace_size(const void *plan);
int drp_rasterize_batch_fwd(void *stream, const float *pos, const int32_t *tri, const void *plan,
                            const void *plan_dev, const float *prev, float *rast, float *rast_db, void *workspace,
                            size_t workspace_bytes, int32_t *boxes, int global_ids);

/* The rest of an uneven batch's forward and backward in one launch sequence
 * per pass (the batched form of diffrp's per-object interpolate at
 * diffrp/rendering/interpolator.py:83 and antialias at surface_deferred.py:775
 * and of their autograd, for meshes of different sizes at different
 * resolutions).  Buffers are the items' packed back to back as in the plan:
 * rast [n_pix,4] from drp_rasterize_batch_fwd with global_ids = 1, colours and
 * their gradients [n_pix,C], pos / grad_pos [n_vert,4] and attr / grad_attr
 * [n_vert,C] concatenated, tri [n_tri,3] with batch (global) vertex indices
 * (each item's triangles offset by the vertices before it), opp its topology
 * table (drp_topology_build), boxes from the rasterize call, inc_off / inc the
 * incidence table of tri (drp_incidence_build).  Gradients are summed in fixed
 * orders without atomics, except the antialias position gradient (one float
 * atomic per vertex component and 32 x 32 tile).  grad_attr, grad_pos are
 * accumulated (caller zero-fills); grad_rast, grad_color written.  The
 * interpolate backward takes C in {1, 2, 3, 4, 8, 16}.  The rasterize
 * backward also takes the item-local ids of global_ids = 0 (local_ids = 1),
 * and the gradient of rast_db (grad_db, nullable). */
int drp_batch_interpolate_fwd(void *stream, const void *plan, const float *attr, int C, const float *rast,
                              const int32_t *tri, float *out);
size_t drp_batch_interpolate_bwd_workspace_size(const void *plan, int C);
int drp_batch_interpolate_bwd(void *stream, const void *plan, const void *plan_dev, const float *attr, int C,
                              const float *rast, const int32_t *tri, const float *grad_out, float *grad_attr,
                              float *grad_rast, const int32_t *boxes, const int32_t *inc_off, const int32_t *inc,
                              void *workspace, size_t workspace_bytes);
size_t drp_batch_rasterize_bwd_workspace_size(const void *plan);
int drp_batch_rasterize_bwd(void *stream, const void *plan, const void *plan_dev, const float *pos, const int32_t *tri,
                            const float *rast, int local_ids, const float *grad_rast, const float *grad_db,
                            float *grad_pos, const int32_t *boxes, const int32_t *inc_off, const int32_t *inc,
                            void *workspace, size_t workspace_bytes);
/* The attribute and position backward of an uneven batch from one pass over
 * its pixels: drp_batch_interpolate_bwd and drp_batch_rasterize_bwd (global
 * ids, no grad_db) of the grad_rast the first would write, without writing
 * it (the batched form of drp_interpolate_rasterize_bwd); grad_attr and
 * grad_pos accumulated (caller zero-fills).  C in {4, 8}; pos [n_vert, 4] as
 * rasterized, tri the batch triangles (global vertex indices). */
size_t drp_batch_interpolate_rasterize_bwd_workspace_size(const void *plan, int C);
int drp_batch_interpolate_rasterize_bwd(void *stream, const void *plan, const void *plan_dev, const float *attr,
                                        int C, const float *pos, const float *rast, const int32_t *tri,
                                        const float *grad_out, float *grad_attr, float *grad_pos,
                                        const int32_t *boxes, const int32_t *inc_off, const int32_t *inc,
                                        void *workspace, size_t workspace_bytes);
size_t drp_batch_antialias_workspace_size(const void *plan);
int drp_batch_antialias_fwd(void *stream, const void *plan, const void *plan_dev, const float *color, int C,
                            const float *rast, const float *pos, const int32_t *tri, const int32_t *opp, float *out,
                            void *workspace, size_t workspace_bytes);
int drp_batch_antialias_bwd(void *stream, const void *plan, const void *plan_dev, const float *color, int C,
                            const float *rast, const float *pos, const int32_t *tri, const int32_t *opp,
                            const float *grad_out, float *grad_color, float *grad_pos, float pos_gradient_boost,
                            int reuse_forward, void *workspace, size_t workspace_bytes);

/* Backward of rasterize: autograd of `rast` as diffrp consumes it
 * (u, v through diffrp/rendering/interpolator.py:42 and :92).  Accumulates
 * into grad_pos (same layout as pos; caller zero-fills).  grad_db may be null.
 * With boxes (from the forward), the incidence table of tri and a workspace
 * of drp_rasterize_bwd_workspace_size bytes, and no ranges, the gradient is
 * summed per 16x16 tile and triangle, per triangle and per vertex in fixed
 * orders without atomics (deterministic); otherwise per-pixel float atomics
 * are used.  The per-pixel chain is evaluated in double precision. */
size_t drp_rasterize_bwd_workspace_size(int n_img, int H, int W, int F);
int drp_rasterize_bwd(void *stream, const float *pos, int n_img, int V, int pos_batched, const int32_t *tri, int F,
                      const int32_t *ranges_dev, int H, int W, const float *rast, const float *grad_rast,
                      const float *grad_db, float *grad_pos, const int32_t *boxes, const int32_t *inc_off,
                      const int32_t *inc, void *workspace, size_t workspace_bytes);

/* ---- texture sampling ------------------------------------------------------
 * Replace diffrp's sample2d / sample3d (diffrp/utils/shader_ops.py:198-310,
 * torch grid_sample underneath; called by GLTFSampler, gltf_material.py:15-22,
 * and image-based lighting).  Texture channels-last: [H,W,C] / [D,H,W,C];
 * coordinates uv [n,2] / uvw [n,3] with (0,0) the bottom-left corner (last
 * row).  wrap: 0 border, 1 reflection, 2 cyclic (2-D only), 3 latlong
 * (reflection, align_corners); mode: 0 bilinear, 1 nearest (no coordinate
 * gradient), 2 bicubic (2-D only; Keys A = -0.75, padding per tap).  out [n,C].  Backward: grad_tex accumulated (float atomics, may
 * be NULL), grad_uv / grad_uvw written (may be NULL). */
int drp_sample2d_fwd(void *stream, const float *tex, int H, int W, int C, const float *uv, long long n, int wrap,
                     int mode, float *out);
int drp_sample2d_bwd(void *stream, const float *tex, int H, int W, int C, const float *uv, long long n, int wrap,
                     int mode, const float *grad_out, float *grad_tex, float *grad_uv);
int drp_sample3d_fwd(void *stream, const float *tex, int D, int H, int W, int C, const float *uvw, long long n,
                     int wrap, int mode, float *out);
int drp_sample3d_bwd(void *stream, const float *tex, int D, int H, int W, int C, const float *uvw, long long n,
                     int wrap, int mode, const float *grad_out, float *grad_tex, float *grad_uvw);

/* Stencil partition of one layer [H,W,4] for the stencil-masked G-buffer path
 * (replaces the per-object owner masks and mask.nonzero() calls of
 * diffrp/rendering/surface_deferred.py:283-290 and interpolator.py:87-97):
 * owner(p) = stencils[triangle id + 1 of p] (0 = background, stencils has F+1
 * entries, values in [0, n_owner]); order [H*W] lists the pixels by owner and,
 * within an owner, in row-major order (nonzero()'s order); owner k's pixels
 * are order[starts[k] .. ends[k]) (starts/ends [n_owner+1], 0/0 when absent).
 * One read of starts/ends per layer replaces one host synchronisation per
 * object. */
size_t drp_stencil_partition_workspace_size(int H, int W, int n_owner);
int drp_stencil_partition(void *stream, const float *rast, int H, int W, const int32_t *stencils, int F, int n_owner,
                          int32_t *order, int32_t *starts, int32_t *ends, void *workspace, size_t workspace_bytes);

/* Covered pixels (id channel > 0) of each of n_img rast images into counts[n]
 * (int32, device; no host synchronisation).  Replaces the host-side emptiness
 * test of the depth-peeling loop, (rast.a <= 0).all() at diffrp/rendering/
 * surface_deferred.py:263: the caller reads the counts one layer late, while
 * the next layer is being rasterised. */
int drp_count_covered(void *stream, const float *rast, int n_img, int H, int W, int32_t *counts);

/* Vertex -> incident corner table of tri (CSR): inc_off [V+1], inc [3F] with
 * corner ids 3t+k grouped by vertex in ascending order.  Built once per mesh;
 * used by the deterministic backward gathers.  No reference counterpart: it is
 * the gather index of the backward passes that replace scatter atomics.  The
 * workspace holds the corner sort keys (16 bytes per corner) and the sort's
 * temporary storage. */
size_t drp_incidence_workspace_size(int V, int F);
int drp_incidence_build(void *stream, const int32_t *tri, int F, int V, int32_t *inc_off, int32_t *inc,
                        void *workspace, size_t workspace_bytes);

/* ---- interpolate ---------------------------------------------------------
 * Replaces dr.interpolate(attr, rast, tri[, rast_db, diff_attrs])
 * (diffrp/rendering/interpolator.py:83).  out_da [n,H,W,2*n_diff] holds
 * (dA/dX, dA/dY) of the channels listed in diff_idx (n_diff = 0: none). */
int drp_interpolate_fwd(void *stream, const float *attr, int n_img, int V, int attr_batched, int C,
                        const float *rast, int H, int W, const int32_t *tri, int F, const float *rast_db,
                        const int32_t *diff_idx, int n_diff, float *out, float *out_da);

/* Backward: grad_attr accumulated (caller zero-fills), grad_rast (and grad_db
 * when not null) written; channels 2, 3 of grad_rast are zero.  With boxes
 * (from the rasterize call that produced rast), the incidence table, no
 * diff_attrs, C in {1,2,3,4,8,16} and a workspace of
 * drp_interpolate_bwd_workspace_size bytes, grad_attr is gathered per triangle
 * (views_per_group views per partial) and per vertex without atomics. */
size_t drp_interpolate_bwd_workspace_size(int n_img, int H, int W, int F, int C, int attr_batched,
                                          int views_per_group);
int drp_interpolate_bwd(void *stream, const float *attr, int n_img, int V, int attr_batched, int C,
                        const float *rast, int H, int W, const int32_t *tri, int F, const float *rast_db,
                        const int32_t *diff_idx, int n_diff, const float *grad_out, const float *grad_da,
                        float *grad_attr, float *grad_rast, float *grad_db, const int32_t *boxes,
                        int views_per_group, const int32_t *inc_off, const int32_t *inc, void *workspace,
                        size_t workspace_bytes);

/* Interpolate backward together with the backward of the rasterize call that
 * produced rast, for interpolate's share of rast's gradient (the chain attr ->
 * interpolate -> rast -> rasterize -> pos of
 * diffrp/rendering/interpolator.py:42-45,83): one pass over the pixels
 * computes the attribute contributions, grad_rast (written, as
 * drp_interpolate_bwd writes it, unless grad_rast is null) and that
 * grad_rast's position contributions; one record search per (image group,
 * triangle) sums both.  grad_attr and grad_pos are accumulated (caller
 * zero-fills) and equal, bit for bit, drp_interpolate_bwd followed by
 * drp_rasterize_bwd on the grad_rast it wrote.  The rasterize backward is
 * linear in grad_rast, so gradients that rast receives from other consumers
 * can go through drp_rasterize_bwd on their own.  attr [V,C] (not batched),
 * C in {4, 8}; pos [V,4] or [n,V,4] (pos_batched); tri and boxes those of the
 * rasterize call. */
size_t drp_interpolate_rasterize_bwd_workspace_size(int n_img, int H, int W, int F, int C, int views_per_group);
int drp_interpolate_rasterize_bwd(void *stream, const float *attr, int n_img, int V, int C, const float *rast, int H,
                                  int W, const int32_t *tri, int F, const float *pos, int pos_batched,
                                  const float *grad_out, float *grad_attr, float *grad_rast, float *grad_pos,
                                  const int32_t *boxes, int views_per_group, const int32_t *inc_off,
                                  const int32_t *inc, void *workspace, size_t workspace_bytes);

/* ---- antialias -----------------------------------------------------------
 * Topology table for antialias: replaces dr.antialias_construct_topology_hash
 * (built implicitly by dr.antialias at diffrp/rendering/surface_deferred.py:775,
 * which passes no hash).  opp [F,3] int32: triangle across edge (v_k, v_k+1)
 * of each triangle, -1 for boundary / non-manifold edges. */
size_t drp_topology_workspace_size(int F);
int drp_topology_build(void *stream, const int32_t *tri, int F, int32_t *opp, void *workspace,
                       size_t workspace_bytes);

/* Replaces dr.antialias(color, rast, pos, tri) (diffrp/rendering/
 * surface_deferred.py:775, via nvdr_antialias :758-775).  color/out [n,H,W,C],
 * n <= 65535 per call. */
size_t drp_antialias_workspace_size(int n_img, int H, int W, int V, int F, int pos_batched);
int drp_antialias_fwd(void *stream, const float *color, int n_img, int H, int W, int C, const float *rast,
                      const float *pos, int V, int pos_batched, const int32_t *tri, int F, const int32_t *opp,
                      float *out, void *workspace, size_t workspace_bytes);

/* The per-vertex and per-triangle part of drp_antialias_fwd (pixel positions
 * of the vertices, screen orientation and silhouette flags of the triangles;
 * no counterpart in nvdiffrast, whose antialias computes it in the call), into
 * an antialias workspace, for a later drp_antialias_fwd_prepared on the same
 * pos, tri, opp and image size: it depends on those alone, so a caller can run
 * it beside the rasterizer (nvdr.rasterize does, on an auxiliary stream, for
 * contexts made with prepare_antialias=True). */
int drp_antialias_prepare(void *stream, int n_img, int H, int W, const float *pos, int V, int pos_batched,
                          const int32_t *tri, int F, const int32_t *opp, void *workspace, size_t workspace_bytes);
/* drp_rasterize_fwd of one layer with every triangle (no ranges, no previous
 * layer) that also runs drp_antialias_prepare for the same pos, tri and opp
 * into aa_workspace (no counterpart in nvdiffrast): its per-vertex and
 * per-triangle passes, which wait on memory, run on the auxiliary stream
 * beside the fine raster, which waits on instruction issue, and are joined
 * before the call's work ends.  nvdr.rasterize uses it for contexts made with
 * prepare_antialias=True. */
int drp_rasterize_fwd_aa_prepare(void *stream, const float *pos, int n_img, int V, int pos_batched,
                                 const int32_t *tri, int F, int H, int W, float *rast, float *rast_db,
                                 void *workspace, size_t workspace_bytes, int tile_w, int tile_h, int bin_cap,
                                 int32_t *boxes, const int32_t *opp, void *aa_workspace, size_t aa_workspace_bytes);
/* drp_antialias_fwd with the workspace already prepared by drp_antialias_prepare. */
int drp_antialias_fwd_prepared(void *stream, const float *color, int n_img, int H, int W, int C, const float *rast,
                               const float *pos, int V, int pos_batched, const int32_t *tri, int F,
                               const int32_t *opp, float *out, void *workspace, size_t workspace_bytes);

/* Backward of antialias: grad_color written; grad_pos accumulated (caller
 * zero-fills; may be null to skip), scaled by pos_gradient_boost.
 * reuse_forward != 0: the workspace still holds what drp_antialias_fwd left
 * there for the same rast, pos and tri (colour may differ), so the pair
 * search is not repeated (ignored for F >= 2^27 - 1, where the forward keeps
 * no pair decisions). */
int drp_antialias_bwd(void *stream, const float *color, int n_img, int H, int W, int C, const float *rast,
                      const float *pos, int V, int pos_batched, const int32_t *tri, int F, const int32_t *opp,
                      const float *grad_out, float *grad_color, float *grad_pos, float pos_gradient_boost,
                      int reuse_forward, void *workspace, size_t workspace_bytes);

/* ---- edge gradient -----------------------------------------------------------
 * Replaces the backward of _EdgeGradient (diffrp/rendering/surface_deferred.py:
 * 120-169, used by edge_gradient :777-804).  All images bottom row first,
 * one view: grad_out / rgb [H,W,C] (C in {1,3,4}), rast [H,W,4], frag_xy [H,W,2]
 * (per-pixel interpolated clip x/w, y/w), tris_xy [F,3,2] (projected corners);
 * writes grad_frag_xy [H,W,2].  The colour gradient is grad_out unchanged. */
int drp_edge_grad_bwd(void *stream, const float *grad_out, const float *rgb, int H, int W, int C, const float *rast,
                      const float *frag_xy, const float *tris_xy, int F, float *grad_frag_xy);

/* ---- clip-space transform ------------------------------------------------
 * Replaces SurfaceDeferredRenderSession.clip_space (diffrp/rendering/
 * surface_deferred.py:229-238, transform_point of diffrp/utils/shader_ops.py:
 * 123-136) for n views at once: xyz [V,3] world positions, VP [n,4,4]
 * row-major view-projection matrices, pos [n,V,4] = VP_b [x, y, z, 1]^T
 * (each row one fma chain m0 x + (m1 y + (m2 z + m3))).  Backward: grad_xyz
 * [V,3] written, the views summed in view order (deterministic). */
int drp_clip_transform_fwd(void *stream, const float *xyz, int V, const float *VP, int n, float *pos);
int drp_clip_transform_bwd(void *stream, const float *grad_pos, int V, const float *VP, int n, float *grad_xyz);

/* ---- image L2 loss ---------------------------------------------------------
 * No counterpart on the reference path: the loss of a fitting step,
 * loss = sum (a - b)^2 over n floats (one read of each image; the sum is
 * deterministic, accumulated in double and rounded to float), and its
 * backward grad_a = 2 g (a - b), grad_b = -grad_a (either may be null) with g
 * the device scalar *grad_loss. */
size_t drp_l2_loss_workspace_size(void);
int drp_l2_loss_fwd(void *stream, const float *a, const float *b, long long n, float *loss, void *workspace,
                    size_t workspace_bytes);
int drp_l2_loss_bwd(void *stream, const float *a, const float *b, long long n, const float *grad_loss, float *grad_a,
                    float *grad_b);
/* The forward and, in the same pass over a and b, grad_a = 2 (a - b): the
 * backward for g = 1, as drp_l2_loss_bwd computes it.  For another g,
 * drp_l2_loss_scale multiplies it in place (grad *= *grad_loss; a pass that
 * reads only *grad_loss when it is 1), which equals drp_l2_loss_bwd bit for
 * bit (scaling by 2 is exact).  diffrp_amd.loss.l2_loss uses the pair when an
 * input requires grad, keeping grad_a from forward to backward. */
int drp_l2_loss_fwd_grad(void *stream, const float *a, const float *b, long long n, float *loss, float *grad_a,
                         void *workspace, size_t workspace_bytes);
int drp_l2_loss_scale(void *stream, float *grad, long long n, const float *grad_loss);

#ifdef __cplusplus
}
#endif

#endif /* DRP_H */
