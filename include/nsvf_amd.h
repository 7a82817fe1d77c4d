/*
 * nsvf_amd.h - C ABI of the MI355X-native sparse-voxel rendering library
 * (libnsvf_amd.so, gfx950).
 *
 * Conventions for every entry point:
 *   - all pointers are device pointers unless stated otherwise; inputs are
 *     read-only and contiguous (row-major, innermost dimension last);
 *   - outputs are fully written by the call (pads included), so the caller
 *     need not pre-fill them;
 *   - `stream` is a hipStream_t (NULL = the default stream); calls are
 *     asynchronous with respect to the host unless stated otherwise;
 *   - the return value is 0 on success, or a negative NSVF_E* code; the text
 *     of the last error on the calling thread is nsvf_last_error().
 *
 * The first group replaces the reference's native module `fairnr.clib._ext`
 * (fairnr/clib/src/binding.cpp:11-21) op for op: same argument meaning, same
 * padded output layout.  The reference's `ball_intersect` and
 * `triangle_intersect` (binding.cpp:12,15) are not part of this path.
 */
#ifndef NSVF_AMD_H
#define NSVF_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NSVF_OK 0
#define NSVF_EINVAL (-1)   /* bad argument (shape, size, null pointer) */
#define NSVF_ELAUNCH (-2)  /* a HIP launch or runtime call failed */
#define NSVF_ESTACK (-3)   /* octree deeper than the traversal supports */
#define NSVF_ENOMEM (-4)   /* host allocation failed */

/* Text of the last error on this thread ("" when none). */
const char* nsvf_last_error(void);

/* Library version: major*10000 + minor*100 + patch. */
int nsvf_version(void);

/* ------------------------------------------------------------------------
 * Ray-voxel intersection
 * ---------------------------------------------------------------------- */

/* Brute-force ray/voxel-box intersection; keeps the first n_max hit voxels
 * of each ray in voxel-index order.
 * Replaces `_ext.aabb_intersect(ray_start, ray_dir, points, voxelsize, n_max)`
 * (fairnr/clib/src/intersect.cpp:49-75, kernel intersect_gpu.cu:125-167).
 *   ray_start, ray_dir : [B,M,3] f32
 *   points             : [Bp,N,3] f32 voxel centres, Bp = 1 (shared by all B
 *                        batches; the reference passes B expanded copies) or B
 *   idx                : [B,M,n_max] i32, -1 padded
 *   min_depth,max_depth: [B,M,n_max] f32, 0 padded                        */
int nsvf_aabb_intersect(const float* ray_start, const float* ray_dir,
                        const float* points, int B, int M, int N,
                        int points_shared, float voxelsize, int n_max,
                        int32_t* idx, float* min_depth, float* max_depth,
                        void* stream);

/* Ray / sparse-voxel-octree intersection by depth-first traversal from the
 * root (node T-1); leaves are reported in traversal order (child slot 7
 * first), the first n_max per ray.
 * Replaces `_ext.svo_intersect(ray_start, ray_dir, points, children,
 * voxelsize, n_max)` (fairnr/clib/src/intersect.cpp:84-112, kernel
 * intersect_gpu.cu:170-237).
 *   points   : [Bp,T,3] f32 node centres; children: [Bp,T,9] i32 (slots 0-7
 *              child ids or -1, slot 8 = box scale 1<<(depth+1), 1 = leaf)
 *   outputs as nsvf_aabb_intersect.
 * Subtrees deeper than 30 internal levels are not descended (a valid tree
 * over int32 coordinates has at most 31).                                  */
int nsvf_svo_intersect(const float* ray_start, const float* ray_dir,
                       const float* points, const int32_t* children, int B,
                       int M, int T, int tree_shared, float voxelsize,
                       int n_max, int32_t* idx, float* min_depth,
                       float* max_depth, void* stream);

/* Host-side octree build over integer voxel coordinates (half-voxel units).
 * Replaces `_ext.build_octree(center, points, depth)`
 * (fairnr/clib/src/octree.cpp:125-136).  Host pointers.  Two calls: the
 * first returns a handle and the node count T, the second copies
 * centers [T,3] i32 and children [T,9] i32 out and frees the handle.      */
int nsvf_build_octree(const int64_t* center, const int64_t* points, int N,
                      int depth, void** handle, int* num_nodes);
int nsvf_octree_take(void* handle, int32_t* centers, int32_t* children);

/* ------------------------------------------------------------------------
 * Ray sampling
 * ---------------------------------------------------------------------- */

/* Importance sampling of each ray's hit intervals by inverse CDF.
 * Replaces `_ext.inverse_cdf_sampling(pts_idx, min_depth, max_depth,
 * uniform_noise, probs, steps, fixed_step_size)`
 * (fairnr/clib/src/sample.cpp:58-95, kernel sample_gpu.cu:108-202).
 *   pts_idx [B,R,P] i32; min_depth, max_depth, probs [B,R,P] f32;
 *   steps [B,R] f32; noise [B,R,S] f32
 *   out_idx [B,R,S] i32 (-1 pad); out_depth, out_dists [B,R,S] f32 (0 pad)
 *   compat = 1 reproduces the reference kernel's trailing-loop behaviour
 *   (sample_gpu.cu:188,194); 0 applies its evident intent.
 *   flags: optional device i64[4] counters (see nsvf_uniform_ray_sampling),
 *   accumulated, never cleared by the call.                                */
int nsvf_inverse_cdf_sampling(const int32_t* pts_idx, const float* min_depth,
                              const float* max_depth, const float* noise,
                              const float* probs, const float* steps, int B,
                              int R, int P, int S, float fixed_step_size,
                              int compat, int32_t* out_idx, float* out_depth,
                              float* out_dists, int64_t* flags, void* stream);

/* Uniform stepping through each ray's hit intervals, merged with the
 * interval boundaries, keeping the in-voxel segments.
 * Replaces `_ext.uniform_ray_sampling(pts_idx, min_depth, max_depth,
 * uniform_noise, step_size, max_steps)` (fairnr/clib/src/sample.cpp:23-55,
 * kernel sample_gpu.cu:15-106).
 *   noise [B,R,S], S = max_steps; outputs [B,R,S]: idx -1 past the kept
 *   samples; depth/dists past them hold the merge-phase values, as in the
 *   reference.
 *   flags (optional, device i64[4]): [0] pts_idx reads outside the buffer
 *   (read as -1), [1] writes past max_steps (dropped), [2] merge loops that
 *   stopped on NaN input, [3] reserved.                                     */
int nsvf_uniform_ray_sampling(const int32_t* pts_idx, const float* min_depth,
                              const float* max_depth, const float* noise,
                              int B, int R, int P, int S, float step_size,
                              int32_t* out_idx, float* out_depth,
                              float* out_dists, int64_t* flags, void* stream);

/* Inverse-CDF sampling of N rows [N,P] in one launch, with the results of the
 * reference's Python wrapper `inverse_cdf_sampling` (fairnr/clib/__init__.py:
 * 231-300), which pads the rows to a multiple of `group` (200) with copies of
 * row 0, regroups them to [group, H/group, P] and calls the kernel on chunks
 * of `chunk` (800) columns: each ray's "first ray of the batch row" and "next
 * ray" (read by the compat-mode trailing loop) are derived from that layout.
 *   noise [N,S] or NULL (then every noise value is 0.5, the wrapper's
 *   deterministic mode); outputs [N,S] (idx -1, depth 0, dists 0 padded).    */
int nsvf_inverse_cdf_sampling_rows(const int32_t* pts_idx, const float* min_depth,
                                   const float* max_depth, const float* noise,
                                   const float* probs, const float* steps, int N,
                                   int P, int S, int group, int chunk,
                                   float fixed_step_size, int compat,
                                   int32_t* out_idx, float* out_depth,
                                   float* out_dists, int64_t* flags, void* stream);

/* ------------------------------------------------------------------------
 * Pruning scores (fairnr/modules/encoder.py:605-654, nsvf.py:143-146): for
 * voxels [v0, v0 + n) of the N voxels, the max of relu(sigma) over a bits^3 lattice inside
 * each voxel (offsets (2i+1-bits)/(bits-1) * vs/2 per axis, corners
 * included) evaluated with the field MLP, into vmax [n].  The reference keeps
 * a voxel when 1 - min exp(-relu sigma) > th, i.e. 1 - exp(-vmax) > th.
 * precision 0 = BF16 MFMA, 1 = FP32 MFMA (the reference computes in FP32);
 * wfrag/bias as for nsvf_field_eval; centers are the x-shifted centres.
 * bits a multiple of 4 (the reference uses 16).
 * ---------------------------------------------------------------------- */
int nsvf_field_prune(int precision, int D, int H, int L, const float* centers,
                     const int32_t* feats, const float* values, int64_t N,
                     float voxel_size, int64_t v0, int64_t n, int bits,
                     const void* wfrag, const float* bias, float* vmax,
                     void* stream);

/* ------------------------------------------------------------------------
 * Splitting (fairnr/data/geometry.py:250-274, encoder.py:656-676): each of N
 * voxels of size vs becomes 8 children of size vs/2 (new_points [8N,3] =
 * p +- vs/4, child order 4x+2y+z), the corner keys are rebuilt (unique,
 * sorted) from the children's corners (new_feats [8N,8] = key ids, as the
 * reference's unique inverse), and the new embeddings are the trilinear
 * interpolation of a parent's corner embeddings at each new corner.  Each
 * new key takes the lowest-index parent among the voxels sharing it (the
 * reference's scatter_ leaves that choice to the device).
 *   nsvf_split_voxels writes new_points, new_feats and the new key count K'
 *   (device int32 *num_keys), and keeps the sorted keys in the workspace;
 *   nsvf_split_values then writes new_keys [K',3] (int64, quarter-voxel
 *   integer coordinates relative to the minimum point) and new_values
 *   [K',D].  Both take the same workspace of nsvf_split_workspace_bytes(N)
 *   bytes (-1 on error).  nsvf_split_voxels reads the points' extent back to
 *   the host once (the packed keys hold 21 bits per axis).
 * ---------------------------------------------------------------------- */
int64_t nsvf_split_workspace_bytes(int64_t N, void* stream);
int nsvf_split_voxels(const float* points, int64_t N, float voxel_size,
                      void* workspace, int64_t workspace_bytes, float* new_points,
                      int32_t* new_feats, int32_t* num_keys, void* stream);
int nsvf_split_values(const float* points, const int32_t* feats,
                      const float* values, int64_t N, int D, float voxel_size,
                      const void* workspace, int64_t workspace_bytes,
                      int64_t K_new, int64_t* new_keys, float* new_values,
                      void* stream);

/* ------------------------------------------------------------------------
 * Trilinear embedding gather (fairnr/modules/encoder.py:582-590 with
 * fairnr/data/geometry.py:195-200; the reference runs it as torch ops).
 *   s_idx [M] voxel ids; xyz [M,3] sample points; centers [N,3] (x-shifted
 *   voxel centres); feats [N,8] corner key ids (corner order 4x+2y+z);
 *   values [K,D]; out [M,D] = sum_c w_c values[feats[v][c]]; a sample
 *   with s_idx -1 (an unused slot) gives a zero row.
 * ---------------------------------------------------------------------- */
int nsvf_trilinear_gather(const int32_t* s_idx, const float* xyz, int64_t M,
                          const float* centers, const int32_t* feats,
                          const float* values, int D, float voxel_size,
                          float* out, void* stream);

/* Backward of nsvf_trilinear_gather: grad_values [K,D] (fully written) =
 * sum over samples and corners of w_c * grad_out[m] into key feats[v][c] --
 * what torch's F.embedding backward forms for the reference's gather
 * (encoder.py:588-590), here deterministic and atomic-free (samples sorted by
 * voxel, per-(voxel, corner) partials, per-key sums in a fixed order).
 * Samples with s_idx outside [0, N) contribute nothing.  D a multiple of 4,
 * at most 128; grad_out and grad_values 16-byte aligned.  workspace: device
 * memory of nsvf_trilinear_gather_backward_workspace(M, N, K, D, stream)
 * bytes (that query returns -1 on error).                                    */
int64_t nsvf_trilinear_gather_backward_workspace(int64_t M, int64_t N, int64_t K,
                                                 int D, void* stream);
int nsvf_trilinear_gather_backward(const int32_t* s_idx, const float* xyz,
                                   int64_t M, const float* centers,
                                   const int32_t* feats, int64_t N, int64_t K,
                                   int D, float voxel_size,
                                   const float* grad_out, float* grad_values,
                                   void* workspace, int64_t workspace_bytes,
                                   void* stream);

/* The per-key order of the (voxel, corner) incidences depends only on the
 * scene's corner keys: nsvf_trilinear_gather_backward_plan computes it once
 * into `plan` (plan_bytes(N, K) bytes; the workspace as for the backward)
 * and nsvf_trilinear_gather_backward_planned takes it instead of sorting
 * the incidences on every call -- the same gradients, bit for bit.  A plan
 * stays valid while feats, N and K do.                                     */
int64_t nsvf_trilinear_gather_backward_plan_bytes(int64_t N, int64_t K);
int nsvf_trilinear_gather_backward_plan(const int32_t* feats, int64_t N,
                                        int64_t K, void* plan,
                                        int64_t plan_bytes, void* workspace,
                                        int64_t workspace_bytes, void* stream);
int nsvf_trilinear_gather_backward_planned(const int32_t* s_idx,
                                           const float* xyz, int64_t M,
                                           const float* centers,
                                           const int32_t* feats, int64_t N,
                                           int64_t K, int D, float voxel_size,
                                           const float* grad_out,
                                           float* grad_values, const void* plan,
                                           void* workspace,
                                           int64_t workspace_bytes,
                                           void* stream);

/* ------------------------------------------------------------------------
 * Fused render pipeline.  No counterpart in the reference's native module:
 * these replace the Python stages of one view (fairnr/models/
 * fairnr_model.py:142-186 with encoder.py:498-592, nsvf.py:43-110,
 * nerf.py:64-79 and renderer.py:77-252).  Sample streams are compacted per
 * hit ray (offset, count) instead of padded [R,S] tensors.
 * ---------------------------------------------------------------------- */

/* Device-wide exclusive prefix sum of n int32 values (out may alias in);
 * the grand total goes to *total (device).  workspace: device int32[
 * nsvf_scan_workspace_ints(n)].                                            */
int64_t nsvf_scan_workspace_ints(int64_t n);
int nsvf_scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n,
                            int32_t* workspace, int32_t* total, void* stream);

/* Octree traversal with the encoder's post-processing: per ray the hits
 * sorted by entry depth (stable with respect to the reference's traversal
 * order), their count, the sum of segment lengths, a hit flag, and an upper
 * bound on the ray's coarse sample count (ceil(sum / step_size) + count + 2,
 * 0 for rays that hit nothing).  Replaces the octree branch of
 * fairnr/modules/encoder.py:498-523 (svo_ray_intersect, intersect.cpp:84-112
 * and intersect_gpu.cu:170-237, then the depth sort) and the segment sum
 * of fairnr/models/nsvf.py:66-74.
 *   width: 0, or the pixel width of the image whose rows the R rays are
 *   (R a multiple of width; several views stack as more rows): lanes then
 *   take the rays in 8x8 pixel tiles, for coherent traversal.  The outputs
 *   do not depend on it.
 *   records [T + 1][8][2] float4: the octree of the reference layout
 *   (centres [T,3], children [T,9], root T-1) packed as node records -- for
 *   node k and child slot s, {min x, y, z, scale bits} and {max x, y, z,
 *   child id bits} of child children[k][s] (id -1: empty slot); record T
 *   holds the root in slot 0.  A leaf's box is fp32(c -+ h), an internal
 *   node's the exact bounding box of its leaves' boxes; scale =
 *   children[k][8].  16-byte aligned.  Rays whose result depends on
 *   traversal order (n_max reached, equal entry depths) are traversed again
 *   in the reference's child order.
 *   depth: the octree's depth (the root's children[8] = 1 << (depth + 1));
 *   the traversal holds trees of depth <= 22 and returns NSVF_ESTACK for
 *   deeper ones (build_octree's trees are at most 19 deep).
 *   hits: [R, n_max] records {float lo, float hi, int32 idx, int32 pad};
 *   only the first count[r] records of a row are written.                  */
int nsvf_render_hits(const float* ray_start, const float* ray_dir, int64_t R,
                     int width, const void* records, int T, int depth, int n_max,
                     float step_size, void* hits, int32_t* count, float* tot,
                     int32_t* hit_flag, int32_t* bound, void* stream);
/* nsvf_render_hits with the traversal's grouping of rays given: lane t of
 * the kernel traces ray order[t] (order [R] int32 on the device, a
 * permutation of 0..R-1; entries outside [0, R) are skipped, so their rays
 * get no output).  For rays that are not image rows -- training batches of
 * random pixels from several views (fairnr/criterions/rendering_loss.py
 * renders such batches) -- a grouping by direction and origin keeps each
 * wave's rays coherent; every output is the same as nsvf_render_hits'.     */
int nsvf_render_hits_ordered(const float* ray_start, const float* ray_dir,
                             int64_t R, const int32_t* order, const void* records,
                             int T, int depth, int n_max, float step_size,
                             void* hits, int32_t* count, float* tot,
                             int32_t* hit_flag, int32_t* bound, void* stream);
/* Sort keys for such a grouping, one pass over the rays: keys[r] (int64) =
 * direction octant << 42 | origin cell << 30 | Morton code of the direction
 * (10 bits per axis over [-1, 1]), the origin cell on a 16^3 grid over the
 * box bounds = {min x, y, z, max x, y, z} (6 floats on the device, e.g. the
 * origins' bounds).  The ray indices sorted by key are an order for
 * nsvf_render_hits_ordered that keeps each wave's rays close in origin and
 * direction.  (No counterpart in the reference, whose traversal is a
 * thread per ray.)                                                         */
int nsvf_ray_order_keys(const float* ray_start, const float* ray_dir, int64_t R,
                        const float* bounds, int64_t* keys, void* stream);

/* One view's pinhole camera, as the reference builds its rays
 * (fairnr/renderer.py:91-112, data/geometry.py:68-122, data_utils.py:57-65):
 * intrinsics fx, fy, cx, cy of the full H x W image; inv_rt the
 * camera-to-world pose [R | t], row-major 3 x 4; the rays are the h x w
 * grid, pixel (row i, column j) at uv = (j * float(W / w), i * float(H / h))
 * (float32 products), in row-major order r = i * w + j.                    */
typedef struct nsvf_camera {
  float fx, fy, cx, cy;
  float inv_rt[12];
  int H, W, h, w;
} nsvf_camera;

/* Rays of a camera (ImageReader.forward / get_ray_direction,
 * fairnr/modules/reader.py:92-119, geometry.py:116-122): ray_start[r] = t,
 * ray_dir[r] = d / (|d| + 1e-8) with d = (R [x, y, 1] + t) - t,
 * x = (u - cx) / fx, y = (v - cy) / fy; FP32, products and sums rounded
 * separately, sums left to right.  cam is a host pointer; outputs
 * [h * w, 3] on the device.                                               */
int nsvf_generate_rays(const nsvf_camera* cam, float* ray_start,
                       float* ray_dir, void* stream);

/* nsvf_render_hits for the rays of a camera, generated inside the traversal
 * kernel (same arithmetic as nsvf_generate_rays) and written to ray_start /
 * ray_dir [h * w, 3] for the later stages; lanes take 8 x 8 pixel tiles.
 * Outputs as nsvf_render_hits with R = h * w.                             */
int nsvf_render_hits_camera(const nsvf_camera* cam, const void* records, int T,
                            int depth, int n_max, float step_size,
                            float* ray_start, float* ray_dir, void* hits,
                            int32_t* count, float* tot, int32_t* hit_flag,
                            int32_t* bound, void* stream);

/* list[rank[r]] = r for every ray with hit_flag[r] != 0.                  */
int nsvf_render_compact(const int32_t* hit_flag, const int32_t* rank,
                        int64_t R, int32_t* list, void* stream);

/* Sample streams: records {int32 voxel id (-1 = unused slot), float depth,
 * float length, int32 ray id}; ray r owns slots [off[r], off[r] + bound[r])
 * (off = exclusive scan of the bounds), the first cnt[r] hold its samples.
 *
 * Noise (both samplers; fairnr/clib/__init__.py:250-256): step i of a ray
 * sits at cdf (i + u_i) / steps.  noise_seed < 0: u_i = 0.5 (evaluation, or
 * --deterministic-step).  noise_seed in [0, 2^32): training-mode noise,
 * u_i = clamp((pcg(i + pcg(r + pcg(seed))) >> 8) / 2^24, 0.001, 0.999) for
 * frame ray r, with pcg the PCG hash of Jarzynski and Olano (JCGT 2020):
 * counter-based, so any implementation can reproduce the draws, where the
 * reference draws torch's uniform_() into a [rays, max_steps] tensor.

 * Coarse inverse-CDF sampling of the hit rays (fairnr/modules/encoder.py:
 * 538-547; steps = sum of segment lengths / step_size, probs = length /
 * sum).  Writes c_cnt[r], the fine-pass bound f_bound[r] =
 * ceil(fine_steps) + c_cnt[r] + 2
 * and *max_count = the longest row.  *n_hit_dev (device) is the number of
 * hit rays, n_hit_max >= it.  Segments (n_seg in 1..16, else 0 and NULLs):
 * several views rendered as one batch, segment s holding the frame rays
 * [seg_bounds[s], seg_bounds[s + 1]); seg_max[s] (zeroed by the caller)
 * receives the longest row of segment s.                                  */
int nsvf_render_sample_coarse(const int32_t* hit_list, const int32_t* n_hit_dev,
                              int64_t n_hit_max, const void* hits,
                              const int32_t* count, const float* tot,
                              const int32_t* bound, const int32_t* c_off,
                              int n_max, float step_size, int compat,
                              float fine_steps, void* samples, int32_t* c_cnt,
                              int32_t* f_bound, int32_t* max_count,
                              const int32_t* seg_bounds, int n_seg,
                              int32_t* seg_max, int64_t noise_seed,
                              void* stream);

/* Fine (hierarchical) inverse-CDF sampling: bins are each ray's coarse
 * samples [depth -+ length/2] with probs (p + 1e-5) / column sum over the
 * *c_max (device) columns; `fine_steps` CDF steps (fairnr/models/nerf.py:
 * 64-79); noise as above.  Writes the fine stream and f_cnt[r].
 * Field reuse (optional src, int16, fine-slot aligned): the fine pass
 * evaluates the coarse pass's field (nerf.py:58), whose value depends only
 * on ray, voxel and depth, and a coarse bin that no fine step splits comes
 * back as a fine sample equal to its coarse sample whenever (lo + hi) / 2
 * rounds back to the coarse depth.  src[slot] = k: the sample is coarse
 * bin k of its ray again (its field value is c_field[c_off[r] + k]);
 * -1: to evaluate; -2: unused slot.  The field list of the -1 slots comes
 * from nsvf_render_eval_list; nsvf_render_composite_reuse reads the reused
 * values from the coarse field.
 * Segments as for nsvf_render_sample_coarse: a ray of segment s normalises
 * its bins over seg_max[s] columns (its own view's longest coarse row)
 * instead of *c_max, so each view of a batch gets its own result.
 * c_prob: the coarse samples' compositing weights, slot-aligned with the
 * coarse stream.  Without c_field it holds them (nsvf_render_composite's
 * prob_out).  With c_field (float4 per coarse slot, the coarse field) the
 * sampler composites each ray's coarse row itself, with the same arithmetic
 * and the same bits as nsvf_render_composite, and writes the weights to
 * c_prob: in a hierarchical view the coarse composite's colour is replaced
 * by the fine one, so the coarse composite is then not needed.              */
int nsvf_render_sample_fine(const int32_t* hit_list, const int32_t* n_hit_dev,
                            int64_t n_hit_max, const int32_t* c_off,
                            const int32_t* c_cnt, const int32_t* c_max,
                            const void* coarse, const void* c_field,
                            float* c_prob, const int32_t* f_bound,
                            const int32_t* f_off, float fine_steps,
                            int compat, void* samples, int32_t* f_cnt,
                            int16_t* src, const int32_t* seg_bounds, int n_seg,
                            const int32_t* seg_max, int64_t noise_seed,
                            void* stream);

/* The field's evaluation list of a marked fine stream: the slots s < *n_dev
 * with src[s] == -1, in slot order.  Three steps:
 * nsvf_render_eval_list_count writes per-block counts (block_cnt[
 * nsvf_render_eval_list_blocks(n_max)]), nsvf_scan_exclusive_i32 turns them
 * into block_off (its total = the list length), nsvf_render_eval_list
 * writes the list.  No counterpart in the reference, which evaluates every
 * fine sample.                                                              */
int64_t nsvf_render_eval_list_blocks(int64_t n_max);
int nsvf_render_eval_list_count(const int16_t* src, const int32_t* n_dev,
                                int64_t n_max, int32_t* block_cnt,
                                void* stream);
int nsvf_render_eval_list(const int16_t* src, const int32_t* n_dev,
                          int64_t n_max, const int32_t* block_off,
                          int32_t* list, void* stream);

/* f_field[slot] = c_field[c_off[r] + src[slot]] for the reused slots of
 * each hit ray (for callers that keep the fine field; the reuse composite
 * does not need it).                                                        */
int nsvf_render_reuse_fill(const int32_t* hit_list, const int32_t* n_hit_dev,
                           int64_t n_hit_max, const int32_t* c_off,
                           const int32_t* f_off, const int32_t* f_cnt,
                           const int16_t* src, const float* c_field,
                           float* f_field, void* stream);

/* Field evaluation of a sample stream: trilinear gather fused with the MLP
 * (L hidden ReLU layers of width H on a D-wide embedding, 4-wide head).
 *   precision 0: BF16 MFMA with FP32 accumulation, wfrag = packed bf16;
 *   precision 1: FP32 MFMA, wfrag = packed f32;
 *   precision 2: split BF16 ("bf16x3"): every product W X as
 *     Wl Xh + Wh Xl + Wh Xh with W = Wh + Wl, X = Xh + Xl (bf16 parts),
 *     FP32 accumulation; wfrag = packed bf16 hi / lo fragment pairs; needs
 *     a workspace and nsvf_field_bf16x3_supported(D, H, L).
 *   samples: *m_dev records (m_max >= *m_dev); unused slots are skipped.
 *   list (optional): *m_dev slot indices -- then sample s is the record at
 *   slot list[s] and its result goes to out[list[s]]; other slots of out
 *   are left as they are (nsvf_render_eval_list builds such a list).
 *   out [M,4] = (sigma, sigmoid(r), sigmoid(g), sigmoid(b)).
 *   workspace (device, optional): with at least
 *   nsvf_field_workspace_bytes(...) bytes, BF16 runs as a gather kernel
 *   writing operand tiles plus an MLP kernel; with NULL, as one fused kernel
 *   (the run-staged gather and the MLP in one kernel where
 *   nsvf_field_fused_supported(D, H, L), else a per-lane gather).  Both
 *   give the same bits.
 *   Supported shapes: nsvf_field_supported(D, H, L).                        */
int nsvf_field_supported(int D, int H, int L);
int nsvf_field_fused_supported(int D, int H, int L);
int nsvf_field_bf16x3_supported(int D, int H, int L);
int64_t nsvf_field_workspace_bytes(int precision, int D, int H, int L,
                                   int64_t m_max);
int nsvf_field_eval(int precision, int D, int H, int L, const float* centers,
                    const int32_t* feats, const float* values, float voxel_size,
                    const float* ray_start, const float* ray_dir,
                    const void* samples, const int32_t* list,
                    const int32_t* m_dev, int64_t m_max, const void* wfrag,
                    const float* bias, float* out, void* workspace,
                    int64_t workspace_bytes, void* stream);

/* nsvf_field_eval without a caller workspace and without allocating: the
 * BF16 (0) and split-BF16 (2) paths stage their operand tiles in a static
 * device array of nsvf_field_static_scratch_bytes() bytes (64 MiB) that
 * every GPU the library runs on holds until the process ends; FP32 (1)
 * needs none.  Larger m_max is refused with NSVF_EINVAL (about 524,288
 * split-BF16 or 1,048,576 BF16 samples at D = 32).  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid it).
 * Calls on one GPU share the array: calls that overlap on that GPU (two
 * streams, or two host threads) race and must be ordered by the caller.
 * Same arguments and results as nsvf_field_eval otherwise.                 */
int64_t nsvf_field_static_scratch_bytes(void);
int nsvf_field_eval_static(int precision, int D, int H, int L,
                           const float* centers, const int32_t* feats,
                           const float* values, float voxel_size,
                           const float* ray_start, const float* ray_dir,
                           const void* samples, const int32_t* list,
                           const int32_t* m_dev, int64_t m_max,
                           const void* wfrag, const float* bias, float* out,
                           void* stream);

/* The two kernels of the workspace BF16 path, callable one by one (so a
 * caller can time or overlap them): the gather writes the MLP's operand
 * tiles into the workspace, the MLP reads them.  Same arguments as
 * nsvf_field_eval.                                                          */
int nsvf_field_embed_bf16(int D, int H, int L, const float* centers,
                          const int32_t* feats, const float* values,
                          float voxel_size, const float* ray_start,
                          const float* ray_dir, const void* samples,
                          const int32_t* list, const int32_t* m_dev,
                          int64_t m_max, void* workspace,
                          int64_t workspace_bytes, void* stream);
int nsvf_field_mlp_bf16(int D, int H, int L, const void* workspace,
                        int64_t workspace_bytes, const int32_t* m_dev,
                        int64_t m_max, const void* wfrag, const float* bias,
                        float* out, const int32_t* list, void* stream);
/* The two kernels of the split-BF16 path (precision 2), one by one: the
 * gather writes hi and lo operand tiles, the MLP reads them.  Same
 * arguments as the BF16 pair; workspace of nsvf_field_workspace_bytes(2,
 * ...) bytes.                                                               */
int nsvf_field_embed_bf16x3(int D, int H, int L, const float* centers,
                            const int32_t* feats, const float* values,
                            float voxel_size, const float* ray_start,
                            const float* ray_dir, const void* samples,
                            const int32_t* list, const int32_t* m_dev,
                            int64_t m_max, void* workspace,
                            int64_t workspace_bytes, void* stream);
int nsvf_field_mlp_bf16x3(int D, int H, int L, const void* workspace,
                          int64_t workspace_bytes, const int32_t* m_dev,
                          int64_t m_max, const void* wfrag, const float* bias,
                          float* out, const int32_t* list, void* stream);
/* The split-BF16 MLP's operand tiles from given features (feats [M, D]
 * row-major float32, e.g. training's gather output; field.py:217-279 takes
 * them as its input), written into the workspace as
 * nsvf_field_embed_bf16x3 writes gathered ones, for nsvf_field_mlp_bf16x3
 * with m_max = M.                                                           */
int nsvf_field_tiles_bf16x3(int D, int H, int L, const float* feats, int64_t M,
                            void* workspace, int64_t workspace_bytes,
                            void* stream);

/* Front-to-back compositing of each hit ray's samples (one wavefront per
 * ray): free energy relu(sigma) * length * 7, probs, and the ray's rgb,
 * depth and missed = 1 - sum(probs), indexed by ray id.  prob_out [M]
 * optional (slot-aligned with the stream).                                  */
int nsvf_render_composite(const int32_t* hit_list, const int32_t* n_hit_dev,
                          int64_t n_hit_max, const int32_t* off,
                          const int32_t* cnt, const void* samples,
                          const float* field, float* prob_out, float* ray_rgb,
                          float* ray_depth, float* ray_missed, void* stream);

/* nsvf_render_composite over a fine stream marked for field reuse: a slot
 * with src[slot] = k >= 0 takes its field value from c_field[c_off[r] + k]
 * (coarse bin k of its ray), the others from field.                        */
int nsvf_render_composite_reuse(const int32_t* hit_list,
                                const int32_t* n_hit_dev, int64_t n_hit_max,
                                const int32_t* off, const int32_t* cnt,
                                const void* samples, const float* field,
                                const int16_t* src, const int32_t* c_off,
                                const float* c_field, float* prob_out,
                                float* ray_rgb, float* ray_depth,
                                float* ray_missed, void* stream);

/* Train-set voxel statistics (encoder.py:594-603 track_voxel_probs, called
 * at renderer.py:250-251): max_voxel_probs[v] = max(max_voxel_probs[v], max
 * over the hit rays of the sum of the ray's sample probabilities in voxel
 * v), probs slot-aligned with the stream (nsvf_render_composite's
 * prob_out); -1 samples ignored.  Sums run left to right in sample order
 * (the reference's scatter_add order is unspecified).  max_voxel_probs
 * [N] on the device, values >= 0; pruning keeps voxels above th
 * (encoder.py:614-615).                                                    */
int nsvf_track_voxel_probs(const int32_t* hit_list, const int32_t* n_hit_dev,
                           int64_t n_hit_max, const int32_t* off,
                           const int32_t* cnt, const void* samples,
                           const float* probs, float* max_voxel_probs,
                           int64_t N, void* stream);

/* Backward of nsvf_render_composite (the autograd of renderer.py:119-121,
 * 195-218): from the gradients of the ray outputs (grad_rgb [R,3],
 * grad_depth [R], grad_missed [R], indexed by ray id) to grad_field [M,4]
 * = (d/dsigma, d/dr, d/dg, d/db) of each used slot.  Transmittance is
 * recomputed, not stored; slots past a ray's count are not written.         */
int nsvf_render_composite_backward(const int32_t* hit_list,
                                   const int32_t* n_hit_dev, int64_t n_hit_max,
                                   const int32_t* off, const int32_t* cnt,
                                   const void* samples, const float* field,
                                   const float* grad_rgb, const float* grad_depth,
                                   const float* grad_missed, float* grad_field,
                                   void* stream);

/* Early ray termination as the reference's chunked ray marching
 * (fairnr/modules/renderer.py:136-174 with --raymarching-tolerance > 0;
 * forward():234-247 splits the rows): the n_hit rows of a pass (row n = hit
 * ray hit_list[n], its samples at stream slots off[r] + j, j < cnt[r]) go
 * in groups of `chunk` rows; in a group the sample columns go in chunks of
 * at most `chunk` hits of rows still marching (a chunk holds at least one
 * column); after each chunk a row's float32 total of free energy
 * relu(sigma) * dist * 7 grows by its chunk's sum (left to right), and a row
 * whose total exceeds `threshold` (= float32(-log(tolerance))) stops: the
 * reference evaluates none of its later samples.
 * The entry points take the pass's field with every sample evaluated
 * (float4 per slot, sigma first; a sample's value does not depend on the
 * others) and replay the marching:
 *   nsvf_march_begin  active[n] = min(cnt, S) + 1 (nonzero while row n
 *                     marches, 0 once it stops), total[n] = 0,
 *                     hist[g][c] = rows of group g with c samples
 *                     (hist zeroed by the caller,
 *                     nsvf_march_hist_ints ints; groups = ceil(n_hit /
 *                     chunk), S = the pass's longest row);
 *   nsvf_march_step   one chunk of every group (span[2 groups] and next
 *                     [groups], next zeroed before the first step): the
 *                     chunk's end column from the histogram, then each
 *                     marching row's chunk sum, its samples marked
 *                     (mark[slot] = 1, mark zeroed by the caller) and the
 *                     stop test; at most S steps, a group is done once
 *                     next[g] == S;
 *   nsvf_march_end    zeroes the field of the unmarked samples, which is
 *                     what the reference's compositing then sees.       */
int64_t nsvf_march_hist_ints(int64_t n_hit, int chunk, int S);
int nsvf_march_begin(const int32_t* hit_list, int64_t n_hit, const int32_t* cnt,
                     int chunk, int S, int32_t* active, float* total,
                     int32_t* hist, void* stream);
int nsvf_march_step(const int32_t* hit_list, int64_t n_hit, const int32_t* off,
                    const int32_t* cnt, const void* samples, const float* field,
                    int chunk, int S, float threshold, int32_t* active,
                    float* total, int32_t* hist, int32_t* span, int32_t* next,
                    int8_t* mark, void* stream);
int nsvf_march_end(const int32_t* hit_list, int64_t n_hit, const int32_t* off,
                   const int32_t* cnt, const int8_t* mark, float* field,
                   void* stream);

/* The field MLP for training (fairnr/modules/field.py:217-279: L Linear +
 * ReLU layers of width H over D inputs, a 4-wide head; output (sigma,
 * sigmoid(r, g, b)), which the reference differentiates with torch
 * autograd), forward and backward on FP32 MFMA.  Built for D=32, H=128,
 * L=3 (nsvf_mlp_train_supported); other shapes return NSVF_EINVAL and the
 * size queries -1.
 * Parameters in torch layout: weights[l] [out][in] and biases[l] [out] for
 * l = 0..L (device pointers, host arrays of L + 1).  nsvf_mlp_train_pack
 * writes the kernels' operand fragments into `packed` (packed_floats).
 * Forward: out [M][4] from emb [M][D].  Backward: from dout [M][4] (the
 * gradient of the output) to demb [M][D] and dparams (param_floats, the
 * gradients of W_0, b_0, ..., W_L, b_L in that order and layout); the
 * activations are recomputed from emb, and the parameter gradients summed
 * over samples in a fixed order through `partial` (partial_floats(M)), so
 * they are deterministic.                                                  */
int nsvf_mlp_train_supported(int D, int H, int L);
int64_t nsvf_mlp_train_packed_floats(int D, int H, int L);
int64_t nsvf_mlp_train_param_floats(int D, int H, int L);
int64_t nsvf_mlp_train_partial_floats(int D, int H, int L, int64_t M);
int nsvf_mlp_train_pack(int D, int H, int L, const float* const* weights,
                        const float* const* biases, float* packed, void* stream);
int nsvf_mlp_train_forward(int D, int H, int L, const float* packed,
                           const float* emb, int64_t M, float* out, void* stream);
int nsvf_mlp_train_backward(int D, int H, int L, const float* packed,
                            const float* emb, int64_t M, const float* dout,
                            float* demb, float* partial, int64_t partial_floats,
                            float* dparams, void* stream);

/* The same backward with the activations' recomputation, the input
 * gradients W^T dz and the parameter gradients in split BF16 (three BF16
 * MFMAs per product on hi / lo operand parts, hi = bf16(f), lo = bf16(f -
 * hi), FP32 accumulation: about FP32's accuracy); the head's recomputed
 * logits and input gradient stay FP32; still deterministic.  No
 * counterpart in the reference (its backward is torch autograd over FP32
 * Linear layers).  x3w: x3_frags(D, H, L) fragments of 1024 bf16 bit
 * patterns (512 hi, then 512 lo), 16-byte aligned, in the order W_0 [H/16
 * m-tiles][D/32 k-tiles], W_1..W_{L-1} [H/16][H/32], then the transposes
 * W_0^T [D/16][H/32], W_1^T..W_{L-1}^T [H/16][H/32]; element (lane, i) of
 * fragment (mt, kt) of a matrix A is A[16 mt + lane % 16][32 kt + k(i)]
 * with k(i) = 4 (lane / 16) + i for i < 4, 16 + 4 (lane / 16) + i - 4
 * otherwise.  `packed` is still nsvf_mlp_train_pack's (biases, head).      */
int64_t nsvf_mlp_train_x3_frags(int D, int H, int L);
int nsvf_mlp_train_backward_x3(int D, int H, int L, const float* packed,
                               const int16_t* x3w, const float* emb, int64_t M,
                               const float* dout, float* demb, float* partial,
                               int64_t partial_floats, float* dparams,
                               void* stream);

/* Frame outputs for all R rays: hit rays' composite plus missed * background
 * (fairnr/models/nsvf.py:89-101); rays that hit nothing get the background. */
int nsvf_render_finish(int64_t R, const int32_t* hit_flag, const float* ray_rgb,
                       const float* ray_depth, const float* ray_missed,
                       float bg_r, float bg_g, float bg_b, float bg_depth,
                       float* rgb, float* depth, float* missed, void* stream);

/* ------------------------------------------------------------------------
 * Comparator kernels: the same ops in the reference's naive form (one thread
 * per ray, the reference's launch geometry of one workgroup per batch row,
 * no LDS, a 256-entry per-thread stack in scratch memory), used only to
 * measure the plain FP32 baseline.  Inputs are per-batch copies as the
 * reference's Python wrappers pass them; outputs as the fast ops
 * (inverse-CDF: the caller pre-fills idx -1 and depths 0, as sample.cpp:81-90
 * does).
 * ---------------------------------------------------------------------- */
int nsvf_plain_aabb_intersect(const float* ray_start, const float* ray_dir,
                              const float* points, int B, int M, int N,
                              float voxelsize, int n_max, int32_t* idx,
                              float* min_depth, float* max_depth, void* stream);
int nsvf_plain_svo_intersect(const float* ray_start, const float* ray_dir,
                             const float* points, const int32_t* children,
                             int B, int M, int T, float voxelsize, int n_max,
                             int32_t* idx, float* min_depth, float* max_depth,
                             void* stream);
int nsvf_plain_inverse_cdf_sampling(const int32_t* pts_idx,
                                    const float* min_depth,
                                    const float* max_depth, const float* noise,
                                    const float* probs, const float* steps,
                                    int B, int R, int P, int S,
                                    float fixed_step_size, int compat,
                                    int32_t* out_idx, float* out_depth,
                                    float* out_dists, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NSVF_AMD_H */
