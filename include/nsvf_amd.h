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
 * Returns NSVF_ESTACK if the tree is deeper than 30 levels.                 */
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

#ifdef __cplusplus
}
#endif

#endif /* NSVF_AMD_H */
