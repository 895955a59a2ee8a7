/*
 * CPU oracle for the voxel-hashing fusion pipeline: test infrastructure only.
 *
 * This is a serial C++ restatement of the reference's fusion logic
 * (include/map_struct.h, src/map_struct.cu, src/depth_fusion.cu,
 * src/ray_tracing.cu, src/marching_cubes.cu, src/image_proc.cu). It exists to
 * CHECK the HIP library: only tests, the smoke entry point and the CPU-baseline
 * leg of the benchmark may load it. Nothing in the product links or calls it.
 *
 * Float semantics: every expression keeps the reference's operand order and
 * the order Eigen >= 3.3 evaluates fixed-size 3-vector reductions in
 * (x + (y + z)); the file is compiled with -ffp-contract=off so that no
 * multiply-add is fused. Allocation is serial (pixel row-major order), which is
 * the one deterministic reading of the reference's try-lock insert.
 */
#ifndef VM_ORACLE_H
#define VM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_map or_map;

/* voxel storage formats: 0 = reference Voxel {short sdf; uchar wt} (4 B),
 * 1 = reference VoxelRGB {short sdf; uchar wt; uchar rgb[3]} (6 B),
 * 2 = {float sdf; float wt; uchar rgb[3]; pad} (12 B) */
or_map* or_map_create(int n_entry, int n_bucket, int n_block, float voxel_size,
                      float trunc_dist, float depth_min, float depth_max,
                      int voxel_format, float max_weight);
void or_map_destroy(or_map* m);
void or_map_reset(or_map* m);
void or_set_threads(int n);

/* Fuse one frame: depth is float metres, row-major W*H (NaN or out of
 * [depth_min, depth_max] = invalid); rgb is W*H*3 bytes or NULL.
 * K = {fx, fy, cx, cy}; T = camera-to-world, column-major 4x4.
 * Returns the visible-block count. */
int or_fuse(or_map* m, const float* depth, const uint8_t* rgb, int W, int H,
            const float* K, const float* T);
/* Allocation stage alone (line tracing) */
void or_allocate(or_map* m, const float* depth, int W, int H, const float* K,
                 const float* T);
/* Visibility stage alone: compacts visible entries, returns the count */
int or_check_visible(or_map* m, int W, int H, const float* K, const float* T);
/* Raycast stage alone into vmap (H*W*4 floats, camera frame) */
void or_raycast(or_map* m, float* vmap, int W, int H, const float* K,
                const float* T);
/* Facade RenderScene: visibility, gate on count, raycast. Returns count. */
int or_render_scene(or_map* m, float* vmap, int W, int H, const float* K,
                    const float* T);
void or_compute_normals(const float* vmap, float* nmap, int W, int H);
/* Phong shading of (vmap, nmap) to RGBA8 */
void or_render_shading(const float* vmap, const float* nmap, uint8_t* rgba,
                       int W, int H);
/* Marching cubes: returns the triangle count (clamped to cap). verts/norms
 * hold cap*9 floats. *block_count receives the number of allocated entries. */
int64_t or_polygonize(or_map* m, float* verts, float* norms, int64_t cap,
                      int* block_count);
void or_surface_points(const float* verts, float* points, int64_t n);
/* Rendering-block z-range maps at 1/8 resolution (cols8 x rows8 each);
 * returns the rendering-block count. Uses the current visible table. */
int or_project_rendering_blocks(or_map* m, float* zrange_x, float* zrange_y,
                                int cols8, int rows8, const float* K,
                                const float* T_world_to_cam);

/* The raycast march for a list of pixels: FP32 (the float raycast's
 * decisions) or FP64 with TSDF perturbations (voxel slot pidx[k] += pval[k]).
 * out_v: n * 4 camera-frame vertices (x = NaN on a miss); sig: signature of
 * the march's discrete decisions; hit: 1 on a hit. T (and K) are the same
 * camera-to-world / intrinsics as or_raycast, in float or double. */
void or_raycast_pixels_f32(or_map* m, int n, const int* px, const int* py, const float* K, const float* T,
                           float* out_v, uint64_t* sig, int* hit);
void or_raycast_pixels_f64(or_map* m, int n, const int* px, const int* py, const double* K, const double* T,
                           const int64_t* pidx, const double* pval, int npert, double* out_v, uint64_t* sig,
                           int* hit);
/* the pose refinement's depth (march, corrections, three Newton steps onto
 * the trilinear zero level) of pixels in FP64; T column-major double */
void or_surface_depth_f64(or_map* m, int n, const int* px, const int* py, const double* K, const double* T,
                          double* out_z, int* ok);

/* createBlock (map_struct.h:304-354) for one key; returns its voxel ptr or -1 */
int or_insert_block(or_map* m, int x, int y, int z);

/* state access for tests */
int or_map_counters(or_map* m, int* out); /* heapPtr, excessPtr, visibleCount,
                                              aliasEvents, allocFailures,
                                              chainBoundHits */
int32_t* or_map_hash_table(or_map* m); /* n_entry * 5 ints {ptr,offset,x,y,z} */
int32_t* or_map_heap(or_map* m);
uint8_t* or_map_voxels(or_map* m);
int32_t* or_map_visible_table(or_map* m);
int or_voxel_bytes(int voxel_format);

/* primitives exposed for closed-form tests */
int or_hash(int x, int y, int z, int n_bucket);
void or_voxel_to_block(int vx, int vy, int vz, int* out3);
void or_voxel_to_local(int vx, int vy, int vz, int* out3);
int or_float2int_rd(float x);
void or_pose_inverse(const float* T, float* out);
int or_mc_edge_table(int cube);
int or_mc_tri_table(int cube, int i);

#ifdef __cplusplus
}
#endif
#endif
