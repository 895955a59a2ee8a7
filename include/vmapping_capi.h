/*
 * vmapping_capi.h -- C ABI of the MI355X (gfx950) voxel-hashing fusion
 * pipeline. Plain pointers and sizes only; every image pointer is DEVICE
 * memory (HBM), every pose/intrinsics array is HOST memory. All calls are
 * asynchronous on `stream` (a hipStream_t, NULL = the null stream) unless an
 * output count pointer is passed, which synchronises that stream.
 *
 * Return value: 0 on success, a negative VM_ERR_* code on failure; the
 * message of the last failure on the calling thread is vm_last_error().
 * Calls on one map must be ordered by the caller (one stream, or events
 * between streams): the map's internal work buffers (raycast tile bounds,
 * mesh work, adjoint scratch) are shared by its calls, as the reference's
 * MapStruct arrays are.
 * Unlike the reference (which prints and calls exit(0) on any CUDA error,
 * src/cuda_utils.h:16-26), errors are returned, never fatal.
 *
 * Conventions kept from the reference:
 *  - camToWorld / worldToCam are column-major 4x4 float (Eigen::Matrix4f);
 *  - K = {fx, fy, cx, cy} = {K(0,0), K(1,1), K(0,2), K(1,2)};
 *  - depth is metres (float) or raw units (uint16) times depth_scale; 0, NaN
 *    or values outside [depth_min, depth_max] are invalid;
 *  - vertex maps are float4 per pixel in the camera frame, x = NaN on a miss;
 *  - meshes are triangle soups: 9 floats per triangle, world metres.
 */
#ifndef VMAPPING_CAPI_H
#define VMAPPING_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VM_OK 0
#define VM_ERR_INVALID_ARG (-1)
#define VM_ERR_HIP (-2)
#define VM_ERR_OUT_OF_MEMORY (-3)
#define VM_ERR_STATE (-4)

/* voxel storage formats */
#define VM_VOXEL_I16_U8 0     /* vmap::Voxel    {short sdf; uchar wt}            4 B */
#define VM_VOXEL_I16_U8_RGB 1 /* vmap::VoxelRGB {short sdf; uchar wt; uchar[3]}  6 B */
#define VM_VOXEL_F32_F32_RGB8 2 /* {float sdf; float wt; uchar rgb[3]; pad}     12 B;
                                   in device memory each block holds its 512
                                   {sdf, wt} pairs (4 KiB), then its 512
                                   {rgb, pad} words (2 KiB); transfers to and
                                   from the host use the 12-byte records */

/* depth image formats */
#define VM_DEPTH_F32 0 /* float metres (the reference's GpuMat CV_32FC1) */
#define VM_DEPTH_U16 1 /* uint16 raw units, metres = raw * depth_scale */

/* status bits (vm_map_stats.status); sticky until vm_map_reset */
#define VM_STATUS_HEAP_FULL 1
#define VM_STATUS_EXCESS_FULL 2
#define VM_STATUS_CHAIN_BOUND 4
#define VM_STATUS_MESH_CLAMPED 8

typedef struct vm_map vm_map;

/* Map geometry. Replaces MapStruct::create(nEntry, nBucket, nVBlock,
 * voxelSize, truncationDist), include/map_struct.h:50 / src/map_struct.cu:44-66,
 * plus the compile-time MIN_DEPTH / MAX_DEPTH of include/map_struct.h:9-10. */
typedef struct vm_map_desc {
    int32_t n_entry;   /* hash slots: n_bucket heads + excess list */
    int32_t n_bucket;  /* primary buckets, 0 < n_bucket <= n_entry */
    int32_t n_block;   /* voxel-block heap capacity, <= 2^22 */
    float voxel_size;  /* metres */
    float trunc_dist;  /* metres */
    float depth_min;   /* 0.1 in the reference */
    float depth_max;   /* 5.0 in the reference */
    int32_t voxel_format;
    float max_weight;  /* weight cap of VM_VOXEL_F32_F32_RGB8 (255 by default);
                          the uchar formats saturate at 255 */
} vm_map_desc;

/* Device arrays of a map, laid out as the reference's MapStruct<TVoxel>
 * members (include/map_struct.h:54-67). Owned by the map. */
typedef struct vm_map_arrays {
    void* hash_table;     /* HashEntry[n_entry]  {int ptr, offset, x, y, z} */
    void* visible_table;  /* HashEntry[n_entry]  compacted visible entries */
    int32_t* heap;        /* int[n_block] free-block stack */
    int32_t* heap_ptr;    /* int */
    int32_t* excess_ptr;  /* int */
    int32_t* bucket_mutex;/* int[n_bucket] */
    uint32_t* visible_block_num; /* uint */
    void* voxel_block;    /* voxels[n_block * 512] (the float format's blocks split, see above) */
    int32_t* block_pos;   /* int4[n_block] {x, y, z, slot} of allocated blocks */
} vm_map_arrays;

typedef struct vm_map_stats {
    int32_t allocated_blocks; /* n_block - 1 - heap_ptr */
    int32_t heap_ptr;
    int32_t excess_used;      /* excess_ptr */
    int32_t visible_blocks;   /* last visibility pass */
    int32_t status;           /* VM_STATUS_* bits */
    int32_t frames_fused;
    int64_t sum_visible;      /* sum of visible counts over fused frames */
} vm_map_stats;

const char* vm_last_error(void);
/* library version string */
const char* vm_version(void);

/* MapStruct::create + reset (map_struct.cu:29-66). */
int vm_map_create(const vm_map_desc* desc, vm_map** out_map);
/* MapStruct::release (map_struct.cu:68-79). NULL is accepted. */
int vm_map_destroy(vm_map* map);
/* MapStruct::reset (map_struct.cu:29-42): empty table, full heap, voxels 0. */
int vm_map_reset(vm_map* map, void* stream);
int vm_map_get_desc(const vm_map* map, vm_map_desc* out_desc);
int vm_map_get_arrays(const vm_map* map, vm_map_arrays* out_arrays);
/* Counters; synchronises `stream`. getVisibleBlockCount (map_struct.cu:81-85)
 * is stats.visible_blocks. */
int vm_map_get_stats(vm_map* map, void* stream, vm_map_stats* out_stats);

/* FuseImage<TVoxel> (depth_fusion.cu:199-301) and, with rgb != NULL,
 * FuseDepthAndImage<TVoxel> (depth_fusion.cu:383-488): allocate the blocks
 * traced by every valid pixel, compact the visible blocks, integrate TSDF,
 * weight (and colour). rgb: 3 bytes per pixel, device, pitch rgb_pitch; it is
 * ignored by the VM_VOXEL_I16_U8 format. If out_visible_count is not NULL the
 * stream is synchronised and the visible-block count returned there (the
 * reference's return value). */
int vm_fuse_depth(vm_map* map, const void* depth, int64_t depth_pitch, int32_t width, int32_t height,
                  int32_t depth_format, float depth_scale, const void* rgb, int64_t rgb_pitch,
                  const float K[4], const float cam_to_world[16], void* stream, uint32_t* out_visible_count);

/* Stage API, as the reference's internal device_functions.h split: */
/* CreateBlockLineTracingFunctor launch alone (depth_fusion.cu:214-236). */
int vm_allocate_blocks(vm_map* map, const void* depth, int64_t depth_pitch, int32_t width, int32_t height,
                       int32_t depth_format, float depth_scale, const float K[4], const float cam_to_world[16],
                       void* stream);
/* The rest of FuseImage after allocation (CheckEntryVisibilityFunctor +
 * DepthFusionFunctor / ImageFusionFunctor, depth_fusion.cu:241-298, :427-484)
 * as one fused pass. vm_allocate_blocks followed by vm_integrate equals
 * vm_fuse_depth. */
int vm_integrate(vm_map* map, const void* depth, int64_t depth_pitch, int32_t width, int32_t height,
                 int32_t depth_format, float depth_scale, const void* rgb, int64_t rgb_pitch, const float K[4],
                 const float cam_to_world[16], void* stream, uint32_t* out_visible_count);
/* CheckVisibleBlocks<TVoxel> (depth_fusion.cu:494-539). */
int vm_check_visible_blocks(vm_map* map, int32_t width, int32_t height, const float K[4],
                            const float cam_to_world[16], void* stream, uint32_t* out_visible_count);

/* RenderScene<TVoxel> (ray_tracing.cu:403-437): raycast every pixel into
 * vmap (float4, camera frame, x = NaN on a miss). A first pass over the
 * allocated blocks bounds, per 8x8 pixel tile, the ray distances at which a
 * sample can read an allocated voxel; rays skip the stretch before and stop
 * (as misses) after it, with results identical to the full march. The
 * bounds live in a per-map buffer sized at vm_map_create for images of up to
 * 2^16 tiles (4.2 Mpixel); a larger image makes the call allocate. */
int vm_raycast(vm_map* map, float* vmap, int64_t vmap_pitch, int32_t width, int32_t height, const float K[4],
               const float cam_to_world[16], void* stream);
/* VoxelMapping::RenderScene (vmapping.cpp:83-104): visible-block check, then
 * the raycast unless nothing is visible (vmap is then left untouched). The
 * visibility pass is skipped when the map has not changed since a pass at the
 * same pose, size and intrinsics (the result would be identical). */
int vm_render_scene(vm_map* map, float* vmap, int64_t vmap_pitch, int32_t width, int32_t height,
                    const float K[4], const float cam_to_world[16], void* stream);
/* ProjectRenderingBlocks<TVoxel> (ray_tracing.cu:171-234) over the current
 * visible table: z-range maps at 1/8 resolution (cols8 x rows8 floats each,
 * pitch in bytes). world_to_cam as the reference passes it. */
int vm_project_rendering_blocks(vm_map* map, float* zrange_x, float* zrange_y, int64_t pitch, int32_t cols8,
                                int32_t rows8, const float K[4], const float world_to_cam[16], void* stream,
                                uint32_t* out_rendering_block_count);

/* vmap::computeNormal (image_proc.cu:95-104). The last row and column are
 * not written, as in the reference. */
int vm_compute_normals(const float* vmap, int64_t vmap_pitch, float* nmap, int64_t nmap_pitch, int32_t width,
                       int32_t height, void* stream);
/* vmap::renderScene (image_proc.cu:63-72): Phong shading to RGBA8. */
int vm_render_shading(const float* vmap, int64_t vmap_pitch, const float* nmap, int64_t nmap_pitch,
                      uint8_t* rgba, int64_t rgba_pitch, int32_t width, int32_t height, void* stream);

/* Polygonize<TVoxel> (marching_cubes.cu:251-305): triangle soup of every
 * allocated block into verts / norms (device, cap_triangles * 9 floats each).
 * Synchronises; returns the triangle count clamped to cap_triangles and the
 * number of allocated blocks. The triangle order is a function of the map
 * (the reference's depends on atomic timing): repeated calls give identical
 * buffers. The first call allocates a work buffer of about 36 B per pool
 * block, kept until vm_map_destroy. */
int vm_polygonize(vm_map* map, float* verts, float* norms, int64_t cap_triangles, void* stream,
                  uint32_t* out_block_count, uint32_t* out_triangle_count);
/* vm_polygonize without any device allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (serving loops
 * and harnesses that forbid hipMalloc, set-up of captured work). Same results
 * as vm_polygonize, into the caller's verts / norms; the mesh work buffer is
 * a fixed array in the library (18 MiB per device, held until the process
 * ends) instead of one allocated by the first call. Maps with more than
 * VM_NOALLOC_MAX_BLOCKS allocated blocks are rejected (VM_ERR_INVALID_ARG).
 * The map's device must be current. All calls on one device share that
 * array: calls that overlap on one device (from other streams or threads)
 * race. Synchronises `stream` before the launches (to check the block
 * count) and after them. */
#define VM_NOALLOC_MAX_BLOCKS (1 << 19)
int vm_polygonize_noalloc(vm_map* map, float* verts, float* norms, int64_t cap_triangles, void* stream,
                          uint32_t* out_block_count, uint32_t* out_triangle_count);
/* GetSurfacePoints (marching_cubes.cu:330-344): triangle centroids. */
int vm_surface_points(const float* verts, float* points, int64_t n_triangles, void* stream);

/* Adjoint of the raycast and the vertex-map normals (no counterpart in the
 * reference): given the forward vertex map `vmap` of vm_raycast at this pose
 * and upstream gradients d loss / d vertex (grad_vertex, float4 per pixel,
 * xyz used) and d loss / d normal (grad_normal, float4 per pixel; either may
 * be NULL), ACCUMULATE
 *   grad_tsdf[i] += d loss / d sdf of voxel slot i (device, n_block * 512
 *                   floats, indexed like voxel_block), and
 *   grad_pose[k] += d loss / d xi_k (device, 6 doubles) for
 *                   camToWorld = T exp(xi), xi = (rho_x, rho_y, rho_z, phi_x, phi_y, phi_z).
 * The march is recomputed with its discrete decisions (cells, interpolation
 * test, break step) held fixed; memory is O(pixels): a work buffer of about
 * 1.5 KiB per pixel (48 records of 32 B), allocated by the first call that
 * needs it and kept until vm_map_destroy. Float voxel format only.
 * Pixels whose march has more than 48 value-dependent steps contribute no
 * gradient and are counted in overflow_pixels; the counts are kept only when
 * out_stats is given (the call then waits for them). */
typedef struct vm_backward_stats {
    int32_t hit_pixels;
    int32_t overflow_pixels;
} vm_backward_stats;
int vm_raycast_backward(vm_map* map, const float* vmap, int64_t vmap_pitch, const float* grad_vertex,
                        int64_t grad_vertex_pitch, const float* grad_normal, int64_t grad_normal_pitch, int32_t width,
                        int32_t height, const float K[4], const float cam_to_world[16], float* grad_tsdf,
                        double* grad_pose, void* stream, vm_backward_stats* out_stats);

/* Batched pose refinement (no counterpart in the reference): n_hyp rigid
 * camera-to-world hypotheses (HOST, column-major float[16] each; every entry
 * of R^T R - I within 1e-5, else VM_ERR_INVALID_ARG) are refined
 * against one depth image (device, formats as vm_fuse_depth) by `iters`
 * Gauss-Newton iterations on the depth residual r = z_render(T) - depth over
 * pixels (x, y) with x and y multiples of `stride`. z_render is the camera z
 * of the map's zero level along the pixel's ray: vm_raycast's hit refined by
 * three Newton steps on the trilinear TSDF. Its Jacobian by the twist xi of
 * T exp(xi) is the implicit-surface derivative dt = -grad s . (do + t dd) /
 * (grad s . d) through the ray's origin and direction. Residuals with
 * |r| >= max_residual metres, misses and depth outside [depth_min,
 * min(depth_max, max_depth)] are left out (max_depth <= 0: depth_max) --
 * with depth noise that grows with range, a max_depth where the noise stays
 * below half the truncation keeps the fit on the part of the map that
 * fusion did not bias. Each iteration is one Gauss-Newton launch over all
 * hypotheses (after a small launch that bounds each hypothesis's ray
 * distances, as vm_raycast's tile bounds do; results are the same bit for
 * bit), then a Levenberg-Marquardt decision per hypothesis: a
 * new pose is kept only if it does not raise the truncated cost (sum of r^2
 * below max_residual, max_residual^2 for every other pixel with a valid
 * depth, over the valid pixels), else the previous pose is restored and the
 * damping (initially 1e-3 of diag(J^T J)) grows tenfold; accepted steps
 * divide it by ten. The damped 6x6 normal equations are solved in double and
 * T <- T exp(delta). out_poses (HOST, n_hyp * 16, column-major) receives the
 * kept poses; out_cost (HOST, n_hyp * (iters + 1) * 2 floats) the {mean r^2,
 * residual count} of the pose kept after each evaluation (mean -1 when no
 * residual is left); out_system (HOST, may be NULL, n_hyp * (iters + 1) * 27
 * doubles) that pose's normal equations: the 21 upper-triangle entries of
 * J^T J row by row, then J^T r. Synchronous. */
int vm_refine_poses(vm_map* map, const void* depth, int64_t depth_pitch, int32_t width, int32_t height,
                    int32_t depth_format, float depth_scale, const float K[4], const float* init_poses, int32_t n_hyp,
                    int32_t iters, int32_t stride, float max_residual, float max_depth, float* out_poses,
                    float* out_cost, double* out_system, void* stream);

/* State transfer for tests and tools: copy a whole map array to/from HOST
 * memory. which: 0 hash table, 1 heap, 2 voxels, 3 counters (16 ints),
 * 4 visible table, 5 block positions. Voxels travel as the format's records
 * (12-byte {sdf, wt, rgb, pad} for VM_VOXEL_F32_F32_RGB8). vm_map_upload
 * rebuilds the block index from the hash table after uploading the hash
 * table. Synchronous. */
int vm_map_download(vm_map* map, int32_t which, void* host_dst, int64_t bytes);
int vm_map_upload(vm_map* map, int32_t which, const void* host_src, int64_t bytes);
/* Voxels of blocks [first_block, first_block + n_blocks) to/from HOST memory
 * (n_blocks * 512 * voxel bytes). No counterpart in the reference: with the
 * whole-array calls above they save and load a map holding only its
 * allocated blocks (ids [0, allocated_blocks)). Synchronous. */
int vm_map_download_blocks(vm_map* map, int32_t first_block, int32_t n_blocks, void* host_dst);
int vm_map_upload_blocks(vm_map* map, int32_t first_block, int32_t n_blocks, const void* host_src);

/* ---------------------------------------------------------------------------
 * Multi-scene batches (no counterpart in the reference: its API fuses one
 * scene per call). Many independent maps of uneven image size, voxel size and
 * table size are advanced in lockstep: one step runs allocation, visibility,
 * integration, raycast and normals for every scene at that scene's current
 * frame, with each stage's work from all scenes in one launch. Per-frame
 * inputs live in device memory (frames stored at a fixed stride, poses
 * uploaded at creation) and the frame index is a device counter, so a step
 * can be captured once into a hipGraph and replayed. Per scene the results
 * equal vm_fuse_depth + vm_render_scene + vm_compute_normals at that frame.
 * All maps of a batch share one voxel format, one depth format and colour
 * (all scenes or none); at most 64 scenes; a map belongs to one batch. */
typedef struct vm_batch vm_batch;

typedef struct vm_batch_scene {
    vm_map* map;
    int32_t width, height;
    float K[4];
    int32_t depth_format;
    float depth_scale;
    const void* depth_frames;   /* device: frame f at depth_frames + f * depth_frame_stride */
    int64_t depth_pitch;
    int64_t depth_frame_stride;
    const void* rgb_frames;     /* device or NULL: 3 bytes per pixel */
    int64_t rgb_pitch;
    int64_t rgb_frame_stride;
    const float* poses;         /* host: n_frames column-major camera-to-world 4x4 */
    int32_t n_frames;           /* steps past the last frame leave the scene unchanged */
    float* vmap;                /* device outputs, float4 per pixel */
    int64_t vmap_pitch;
    float* nmap;
    int64_t nmap_pitch;
} vm_batch_scene;

/* launch shape of a batch step (chosen by vm_batch_tune) */
typedef struct vm_batch_tuning {
    int32_t ray_tile_width;       /* 8 (8x8 rays per wave), 16 (16x4) or 32 (32x2) */
    int32_t xcd_bands;            /* 1: give each XCD a contiguous range of ray tiles */
    int32_t visible_wg_per_cu;    /* visibility workgroups per compute unit */
    int32_t integrate_wg_per_cu;  /* integration workgroups (4 waves) per compute unit */
    int32_t scene_groups;         /* 1..4: scenes dealt to this many groups (largest image first, to the
                                     group with the fewest pixels), each group's kernels on a stream of
                                     its own, forked from and joined back into the caller's stream */
    int32_t persistent;           /* 0: a step is five kernels; 1: one kernel whose workgroups are all
                                     resident, the stages separated by grid barriers (needs
                                     scene_groups 1); 2: the same, launched cooperatively (not capturable) */
    int32_t persist_wg_per_cu;    /* 1..32: resident workgroups per compute unit of the one-kernel step,
                                     capped by its occupancy */
} vm_batch_tuning;

int vm_batch_create(const vm_batch_scene* scenes, int32_t n_scenes, vm_batch** out_batch);
int vm_batch_destroy(vm_batch* batch);
/* device-side frame index used by the next step (stream-ordered) */
int vm_batch_set_frame(vm_batch* batch, int32_t frame, void* stream);
/* reads the frame index back (synchronizes the stream); VM_ERR_STATE when a
   one-kernel step's grid-barrier wait has run out (the frame is still set) */
int vm_batch_get_frame(vm_batch* batch, void* stream, int32_t* out_frame);
/* one step for all scenes, then frame += 1; replays the captured graph if any */
int vm_batch_step(vm_batch* batch, void* stream);
/* capture one step into a hipGraph (stream must not be NULL) */
int vm_batch_capture(vm_batch* batch, void* stream);
int vm_batch_get_tuning(vm_batch* batch, vm_batch_tuning* out_tuning);
/* changes the launch shape; drops a captured graph */
int vm_batch_set_tuning(vm_batch* batch, const vm_batch_tuning* tuning);
/* Autotuner: times the 18 launch shapes (ray tile 8x8 / 16x4 / 32x2, XCD
 * bands on / off, 4 / 8 / 16 integration workgroups per CU) with one scene
 * group, then 2, 3 and 4 scene groups with the fastest shape, then the
 * one-kernel step (persistent 1, 4 workgroups per CU) with that shape,
 * `steps` steps each (the steps fuse frames like any step: up to 22 steps
 * times `steps`), and applies the fastest; drops a captured graph. */
int vm_batch_tune(vm_batch* batch, void* stream, int32_t steps, vm_batch_tuning* out_best);

#ifdef __cplusplus
}
#endif
#endif
