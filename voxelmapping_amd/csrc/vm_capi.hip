// C ABI of the gfx950 fusion pipeline (declared in include/vmapping_capi.h).
// Owns map allocation, argument checking, host-side constants (intrinsic
// reciprocals, pose inverse) and stage sequencing. No host synchronisation
// happens unless the caller asks for a count.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "vm_batch.h"
#include "vm_common.h"
#include "vm_dev_ray.h"
#include "vm_launch.h"
#include "vmapping_capi.h"

using namespace vmamd;

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    g_error = std::string(what) + ": " + hipGetErrorString(e);
    return VM_ERR_HIP;
}

#define VM_HIP(call, what)                      \
    do {                                        \
        hipError_t e_ = (call);                 \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

Affine affine_of(const float* T) {
    Affine a;
    a.r00 = T[0]; a.r01 = T[4]; a.r02 = T[8];  a.t0 = T[12];
    a.r10 = T[1]; a.r11 = T[5]; a.r12 = T[9];  a.t1 = T[13];
    a.r20 = T[2]; a.r21 = T[6]; a.r22 = T[10]; a.t2 = T[14];
    return a;
}

// General 4x4 inverse by cofactors in double, rounded to float. The
// reference uses Eigen's Matrix4f::inverse() (depth_fusion.cu:251,
// ray_tracing.cu:414); this is the correctly rounded form of the same matrix.
bool inverse4(const float* Tf, float* out) {
    double m[16];
    for (int i = 0; i < 16; ++i) m[i] = Tf[i];
    auto A = [&](int r, int c) { return m[c * 4 + r]; };
    double cof[4][4];
    for (int r = 0; r < 4; ++r) {
        for (int c = 0; c < 4; ++c) {
            int rr[3], cc[3], k = 0, l = 0;
            for (int i = 0; i < 4; ++i) if (i != r) rr[k++] = i;
            for (int j = 0; j < 4; ++j) if (j != c) cc[l++] = j;
            double d = A(rr[0], cc[0]) * (A(rr[1], cc[1]) * A(rr[2], cc[2]) - A(rr[1], cc[2]) * A(rr[2], cc[1])) -
                       A(rr[0], cc[1]) * (A(rr[1], cc[0]) * A(rr[2], cc[2]) - A(rr[1], cc[2]) * A(rr[2], cc[0])) +
                       A(rr[0], cc[2]) * (A(rr[1], cc[0]) * A(rr[2], cc[1]) - A(rr[1], cc[1]) * A(rr[2], cc[0]));
            cof[r][c] = ((r + c) & 1) ? -d : d;
        }
    }
    double det = 0;
    for (int c = 0; c < 4; ++c) det += A(0, c) * cof[0][c];
    if (det == 0.0 || !std::isfinite(det)) return false;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[c * 4 + r] = (float)(cof[c][r] / det);
    return true;
}

Intrinsics intrinsics_of(const float* K) {
    Intrinsics k;
    k.fx = K[0]; k.fy = K[1]; k.cx = K[2]; k.cy = K[3];
    k.ifx = (float)(1.0 / (double)K[0]);
    k.ify = (float)(1.0 / (double)K[1]);
    return k;
}

bool finite_all(const float* v, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

}  // namespace

struct vm_map {
    vm_map_desc desc;
    MapView view;
    int device;
    // visibility reuse: the visible table is a pure function of the allocated
    // set and (size, K, pose); `epoch` counts allocated-set changes
    uint64_t epoch = 0;
    bool visValid = false;
    uint64_t visEpoch = 0;
    int visW = 0, visH = 0;
    float visK[4];
    float visT[16];
    // set by an allocation launch, which zeroes the visible count on the device
    bool visCountZeroed = false;
    // scratch of the raycast adjoint, grown on demand
    void* scratch = nullptr;
    size_t scratchBytes = 0;
    // per-tile ray-distance bounds of the raycast (keys of 2 (tiles + 1)
    // words), grown on demand; one epoch per raycast
    unsigned long long* rayKeys = nullptr;
    int rayKeysTiles = 0;
    unsigned rayEpoch = 0;
    // the bounds of the last raycast: the adjoint reuses them while the
    // allocated set (epoch), image size, intrinsics and pose are the same and
    // no bounds were computed since (rayEpoch)
    bool rbValid = false;
    uint64_t rbMapEpoch = 0;
    unsigned rbEpoch = 0;
    int rbW = 0, rbH = 0;
    float rbK[4];
    float rbT[16];
    // work buffer of the mesh extraction (neighbour offsets and triangle
    // counts, about 9 ints per pool block), allocated by the first polygonize
    int* meshWork = nullptr;
};

namespace {

int check_map(const vm_map* m) {
    if (!m) return fail(VM_ERR_INVALID_ARG, "map is NULL");
    return VM_OK;
}

// The raycast's per-tile bounds for a W x H image with a fresh epoch. The
// keys are (re)allocated only when the image has more tiles than any before
// (room for 2^16 tiles at first) or when the epoch counter wraps; a fresh
// buffer is filled on the stream (near keys all ones, far keys zero), which
// reads as "no block" for every epoch.
int ray_bounds_for(vm_map* m, int W, int H, hipStream_t s, RayBounds* out) {
    const int n = ray_bounds_tiles(W, H);
    m->rbValid = false;  // the caller records the bounds it launches
    if (n > m->rayKeysTiles || m->rayEpoch == 0xfffffffeu) {
        const int cap = std::max(n, 1 << 16);
        if (cap > m->rayKeysTiles) {
            // raycasts already queued on any stream may still read the old keys
            if (m->rayKeys) {
                VM_HIP(hipDeviceSynchronize(), "device synchronize");
                VM_HIP(hipFree(m->rayKeys), "raycast bounds free");
            }
            m->rayKeys = nullptr;
            m->rayKeysTiles = 0;
            hipError_t e = hipMalloc((void**)&m->rayKeys, sizeof(unsigned long long) * 2 * ((size_t)cap + 1));
            if (e != hipSuccess) return hip_fail(e, "raycast bounds allocation");
            m->rayKeysTiles = cap;
        }
        const int64_t words = 2 * ((int64_t)m->rayKeysTiles + 1);  // 32-bit words per half
        VM_HIP(launch_fill_u32(m->rayKeys, 0xffffffffu, words, s), "raycast bounds fill");
        VM_HIP(launch_fill_u32(m->rayKeys + m->rayKeysTiles + 1, 0u, words, s), "raycast bounds fill");
        m->rayEpoch = 0;
    }
    m->rayEpoch++;
    // the layout follows the capacity, not this image's tile count, so a key
    // left by a launch at another image size sits in a slot of an older
    // epoch, which every newer key replaces
    *out = RayBounds{m->rayKeys, m->rayKeysTiles, m->rayEpoch};
    return VM_OK;
}

void remember_bounds(vm_map* m, int W, int H, const float K[4], const float T[16]) {
    m->rbValid = true;
    m->rbMapEpoch = m->epoch;
    m->rbEpoch = m->rayEpoch;
    m->rbW = W;
    m->rbH = H;
    std::memcpy(m->rbK, K, sizeof(m->rbK));
    std::memcpy(m->rbT, T, sizeof(m->rbT));
}

bool bounds_current(const vm_map* m, int W, int H, const float K[4], const float T[16]) {
    return m->rbValid && m->rbMapEpoch == m->epoch && m->rbEpoch == m->rayEpoch && m->rbW == W && m->rbH == H &&
           !std::memcmp(m->rbK, K, sizeof(m->rbK)) && !std::memcmp(m->rbT, T, sizeof(m->rbT));
}

bool rigid_rotation(const float* T);

// The visibility launch of a fusion also computes the ray bounds of a
// raycast at the fusion's pose, image size and intrinsics (fresh epoch), so
// that the raycast which usually follows at that view skips its own pass
// over the allocated blocks (bounds_current then holds). Not for a pose
// whose rotation is not orthonormal (the bounds assume it): *use false.
int fusion_bounds(vm_map* m, int W, int H, const float K[4], const float T[16], hipStream_t s, BoundsAtPose* out,
                  bool* use) {
    *use = rigid_rotation(T);
    if (!*use) return VM_OK;
    if (int rc = ray_bounds_for(m, W, H, s, &out->rb)) return rc;
    const float vsInv = (float)(1.0 / (double)m->view.voxelSize);  // as the raycast computes it
    const Affine P = affine_of(T);
    out->centre = make_float3(P.t0 * vsInv, P.t1 * vsInv, P.t2 * vsInv);
    out->grow = ray_drift_margin(W, H, intrinsics_of(K), P.t0, P.t1, P.t2, vsInv);
    return VM_OK;
}

// The rotation block of a column-major camToWorld is orthonormal: every
// entry of R^T R - I within 1e-5 (float roundings of a rotation sit near
// 1e-7). The raycast's per-tile ray-distance bounds assume it.
bool rigid_rotation(const float* T) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double d = 0.0;
            for (int k = 0; k < 3; ++k) d += (double)T[i * 4 + k] * (double)T[j * 4 + k];
            if (std::fabs(d - (i == j ? 1.0 : 0.0)) > 1e-5) return false;
        }
    return true;
}

int check_pose_k(const float* K, const float* T) {
    if (!K || !T) return fail(VM_ERR_INVALID_ARG, "K and pose must not be NULL");
    if (!finite_all(K, 4) || !finite_all(T, 16)) return fail(VM_ERR_INVALID_ARG, "K and pose must be finite");
    if (K[0] == 0.f || K[1] == 0.f) return fail(VM_ERR_INVALID_ARG, "focal lengths must be non-zero");
    return VM_OK;
}

int check_image(const void* p, int64_t pitch, int w, int h, int bpp, const char* name) {
    if (w <= 0 || h <= 0) return fail(VM_ERR_INVALID_ARG, std::string(name) + ": width and height must be > 0");
    if (w > 32767 || h > 32767) return fail(VM_ERR_INVALID_ARG, std::string(name) + ": image larger than 32767");
    if (!p) return fail(VM_ERR_INVALID_ARG, std::string(name) + " is NULL");
    if (pitch < (int64_t)w * bpp) return fail(VM_ERR_INVALID_ARG, std::string(name) + ": pitch smaller than a row");
    return VM_OK;
}

void remember_visibility(vm_map* m, int W, int H, const float* K, const float* T) {
    m->visCountZeroed = false;  // a visibility pass has counted into it
    m->visValid = true;
    m->visEpoch = m->epoch;
    m->visW = W;
    m->visH = H;
    std::memcpy(m->visK, K, sizeof(m->visK));
    std::memcpy(m->visT, T, sizeof(m->visT));
}

bool visibility_current(const vm_map* m, int W, int H, const float* K, const float* T) {
    return m->visValid && m->visEpoch == m->epoch && m->visW == W && m->visH == H &&
           std::memcmp(m->visK, K, sizeof(m->visK)) == 0 && std::memcmp(m->visT, T, sizeof(m->visT)) == 0;
}

int check_polygonize_args(const vm_map* m, const float* verts, const float* norms, int64_t cap) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    if (cap < 0) return fail(VM_ERR_INVALID_ARG, "cap_triangles must be >= 0");
    if (cap > 0 && (!verts || !norms)) return fail(VM_ERR_INVALID_ARG, "verts and norms must not be NULL");
    return VM_OK;
}

// After the mesh launches on s: wait, read the counters, flag a clamped
// mesh, return the block count and the triangles kept.
int polygonize_finish(vm_map* m, int64_t cap, hipStream_t s, uint32_t* out_blocks, uint32_t* out_tris) {
    int c[kCounterInts];
    VM_HIP(hipMemcpyAsync(c, m->view.ctr, sizeof(c), hipMemcpyDeviceToHost, s), "counter copy");
    VM_HIP(hipStreamSynchronize(s), "stream synchronize");
    const int64_t tris = (int64_t)((uint64_t)(uint32_t)c[kTriangleTotalLo] | ((uint64_t)(uint32_t)c[kTriangleTotalHi] << 32));
    const int64_t kept = std::min<int64_t>(std::min<int64_t>(tris, cap), 0xffffffffll);
    if (tris > kept) {
        int st = c[kStatus] | kStatusMeshClamped;
        VM_HIP(hipMemcpyAsync(m->view.ctr + kStatus, &st, sizeof(int), hipMemcpyHostToDevice, s), "status update");
        VM_HIP(hipStreamSynchronize(s), "stream synchronize");
    }
    if (out_blocks) *out_blocks = (uint32_t)(m->view.nBlock - 1 - c[kHeapPtr]);
    if (out_tris) *out_tris = (uint32_t)kept;
    return VM_OK;
}

int read_counter(vm_map* m, int idx, hipStream_t s, int* out) {
    VM_HIP(hipMemcpyAsync(out, m->view.ctr + idx, sizeof(int), hipMemcpyDeviceToHost, s), "counter copy");
    VM_HIP(hipStreamSynchronize(s), "stream synchronize");
    return VM_OK;
}

}  // namespace

extern "C" {

const char* vm_last_error(void) { return g_error.c_str(); }
const char* vm_version(void) { return "vmapping-amd 0.1 gfx950"; }

int vm_map_create(const vm_map_desc* d, vm_map** out) {
    if (!d || !out) return fail(VM_ERR_INVALID_ARG, "desc and out_map must not be NULL");
    *out = nullptr;
    if (d->n_bucket <= 0 || d->n_entry < d->n_bucket)
        return fail(VM_ERR_INVALID_ARG, "need 0 < n_bucket <= n_entry");
    if (d->n_block <= 0 || d->n_block > (1 << 22))
        return fail(VM_ERR_INVALID_ARG, "n_block must be in [1, 2^22] (ptr = block * 512 is an int)");
    if (!(d->voxel_size > 0.f) || !(d->trunc_dist > 0.f) || !std::isfinite(d->voxel_size) ||
        !std::isfinite(d->trunc_dist))
        return fail(VM_ERR_INVALID_ARG, "voxel_size and trunc_dist must be finite and > 0");
    if (!(d->depth_min >= 0.f) || !(d->depth_max > d->depth_min))
        return fail(VM_ERR_INVALID_ARG, "need 0 <= depth_min < depth_max");
    if (d->voxel_format < 0 || d->voxel_format > 2) return fail(VM_ERR_INVALID_ARG, "unknown voxel_format");
    if (d->voxel_format == VM_VOXEL_F32_F32_RGB8 && !(d->max_weight >= 1.f))
        return fail(VM_ERR_INVALID_ARG, "max_weight must be >= 1");

    vm_map* m = new vm_map;
    m->desc = *d;
    if (m->desc.voxel_format != VM_VOXEL_F32_F32_RGB8) m->desc.max_weight = 255.f;
    MapView& v = m->view;
    std::memset(&v, 0, sizeof(v));
    v.nEntry = d->n_entry;
    v.nBucket = d->n_bucket;
    v.nBlock = d->n_block;
    v.bucketMask = (d->n_bucket & (d->n_bucket - 1)) == 0 ? d->n_bucket - 1 : -1;
    v.voxelSize = d->voxel_size;
    v.trunc = d->trunc_dist;
    v.depthMin = d->depth_min;
    v.depthMax = d->depth_max;
    v.maxWeight = m->desc.max_weight;
    v.fmt = d->voxel_format;
    v.occMask = brick_words(v.nBlock) - 1;
    if (hipError_t e = hipGetDevice(&m->device); e != hipSuccess) {
        vm_map_destroy(m);
        return hip_fail(e, "current device");
    }

    const size_t vb = (size_t)voxel_bytes(v.fmt);
    struct { void** p; size_t bytes; const char* name; } allocs[] = {
        {(void**)&v.table, sizeof(HashEntry) * (size_t)v.nEntry, "hash table"},
        {(void**)&v.visible, sizeof(HashEntry) * (size_t)v.nEntry, "visible table"},
        {(void**)&v.heap, sizeof(int) * (size_t)v.nBlock, "heap"},
        {(void**)&v.mutex, sizeof(int) * (size_t)v.nBucket, "bucket mutex"},
        {(void**)&v.voxels, vb * kBlockVox * (size_t)v.nBlock, "voxel blocks"},
        {(void**)&v.blockPos, sizeof(int4) * (size_t)v.nBlock, "block positions"},
        {(void**)&v.ctr, sizeof(int) * kCounterInts, "counters"},
        {(void**)&v.occ, sizeof(unsigned long long) * ((size_t)v.occMask + 1), "occupancy bricks"},
    };
    for (auto& a : allocs) {
        hipError_t e = hipMalloc(a.p, a.bytes);
        if (e != hipSuccess) {
            std::string msg = std::string("hipMalloc of ") + a.name + " failed: " + hipGetErrorString(e);
            vm_map_destroy(m);
            return fail(VM_ERR_OUT_OF_MEMORY, msg);
        }
    }
    // the raycast's tile bounds for images up to 2^16 bounds tiles, so that a
    // raycast of such an image allocates nothing
    RayBounds rb0;
    int rc = ray_bounds_for(m, 1, 1, nullptr, &rb0);
    if (rc == VM_OK) rc = vm_map_reset(m, nullptr);
    if (rc != VM_OK) {
        std::string msg = g_error;
        vm_map_destroy(m);
        return fail(rc, msg);
    }
    hipError_t e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        vm_map_destroy(m);
        return hip_fail(e, "map reset");
    }
    *out = m;
    return VM_OK;
}

int vm_map_destroy(vm_map* m) {
    if (!m) return VM_OK;
    MapView& v = m->view;
    void* ptrs[] = {v.table, v.visible, v.heap, v.mutex, v.voxels, v.blockPos, v.ctr, v.occ, m->scratch,
                    m->rayKeys, m->meshWork};
    // teardown: a failed free leaves nothing for the caller to recover
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete m;
    return VM_OK;
}

int vm_map_reset(vm_map* m, void* stream) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    VM_HIP(launch_reset(m->view, (hipStream_t)stream), "map reset");
    m->epoch++;
    m->visValid = false;
    return VM_OK;
}

int vm_map_get_desc(const vm_map* m, vm_map_desc* out) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    if (!out) return fail(VM_ERR_INVALID_ARG, "out_desc is NULL");
    *out = m->desc;
    return VM_OK;
}

int vm_map_get_arrays(const vm_map* m, vm_map_arrays* o) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    if (!o) return fail(VM_ERR_INVALID_ARG, "out_arrays is NULL");
    const MapView& v = m->view;
    o->hash_table = v.table;
    o->visible_table = v.visible;
    o->heap = v.heap;
    o->heap_ptr = v.ctr + kHeapPtr;
    o->excess_ptr = v.ctr + kExcessPtr;
    o->bucket_mutex = v.mutex;
    o->visible_block_num = (uint32_t*)(v.ctr + kVisibleCount);
    o->voxel_block = v.voxels;
    o->block_pos = (int32_t*)v.blockPos;
    return VM_OK;
}

int vm_map_get_stats(vm_map* m, void* stream, vm_map_stats* o) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    if (!o) return fail(VM_ERR_INVALID_ARG, "out_stats is NULL");
    int c[kCounterInts];
    hipStream_t s = (hipStream_t)stream;
    VM_HIP(hipMemcpyAsync(c, m->view.ctr, sizeof(c), hipMemcpyDeviceToHost, s), "counter copy");
    VM_HIP(hipStreamSynchronize(s), "stream synchronize");
    o->heap_ptr = c[kHeapPtr];
    o->allocated_blocks = m->view.nBlock - 1 - c[kHeapPtr];
    o->excess_used = c[kExcessPtr];
    o->visible_blocks = c[kVisibleCount];
    o->status = c[kStatus];
    o->frames_fused = c[kFramesFused];
    uint64_t lo = (uint32_t)c[kSumVisibleLo], hi = (uint32_t)c[kSumVisibleHi];
    o->sum_visible = (int64_t)(lo | (hi << 32));
    return VM_OK;
}

int vm_allocate_blocks(vm_map* m, const void* depth, int64_t pitch, int32_t W, int32_t H, int32_t fmt, float scale,
                       const float K[4], const float T[16], void* stream) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    if (fmt != VM_DEPTH_F32 && fmt != VM_DEPTH_U16) return fail(VM_ERR_INVALID_ARG, "unknown depth_format");
    if (int rc = check_image(depth, pitch, W, H, fmt == VM_DEPTH_F32 ? 4 : 2, "depth")) return rc;
    if (int rc = check_pose_k(K, T)) return rc;
    DepthImage d{(const uint8_t*)depth, pitch, W, H, fmt, scale};
    VM_HIP(launch_allocate(m->view, d, intrinsics_of(K), affine_of(T), (hipStream_t)stream), "allocation");
    m->epoch++;
    m->visCountZeroed = true;
    return VM_OK;
}

int vm_integrate(vm_map* m, const void* depth, int64_t pitch, int32_t W, int32_t H, int32_t fmt, float scale,
                 const void* rgb, int64_t rgb_pitch, const float K[4], const float T[16], void* stream,
                 uint32_t* out_visible) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    if (fmt != VM_DEPTH_F32 && fmt != VM_DEPTH_U16) return fail(VM_ERR_INVALID_ARG, "unknown depth_format");
    if (int rc = check_image(depth, pitch, W, H, fmt == VM_DEPTH_F32 ? 4 : 2, "depth")) return rc;
    if (rgb)
        if (int rc = check_image(rgb, rgb_pitch, W, H, 3, "rgb")) return rc;
    if (int rc = check_pose_k(K, T)) return rc;
    float Ti[16];
    if (!inverse4(T, Ti)) return fail(VM_ERR_INVALID_ARG, "camToWorld is singular");
    hipStream_t s = (hipStream_t)stream;
    // the allocation kernel zeroes the visible count; without it, zero it here
    if (!m->visCountZeroed) VM_HIP(launch_fill_u32(m->view.ctr + kVisibleCount, 0u, 1, s), "count reset");
    m->visCountZeroed = false;
    DepthImage d{(const uint8_t*)depth, pitch, W, H, fmt, scale};
    ColorImage c{(const uint8_t*)rgb, rgb_pitch};
    BoundsAtPose bp;
    bool withBounds = false;
    if (int rc = fusion_bounds(m, W, H, K, T, s, &bp, &withBounds)) return rc;
    VM_HIP(launch_visible_fuse(m->view, d, c, intrinsics_of(K), affine_of(Ti), s, withBounds ? &bp : nullptr),
           "visibility + integration");
    remember_visibility(m, W, H, K, T);
    if (withBounds) remember_bounds(m, W, H, K, T);
    if (out_visible) {
        int n = 0;
        if (int rc = read_counter(m, kVisibleCount, s, &n)) return rc;
        *out_visible = (uint32_t)n;
    }
    return VM_OK;
}

int vm_fuse_depth(vm_map* m, const void* depth, int64_t pitch, int32_t W, int32_t H, int32_t fmt, float scale,
                  const void* rgb, int64_t rgb_pitch, const float K[4], const float T[16], void* stream,
                  uint32_t* out_visible) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    if (fmt != VM_DEPTH_F32 && fmt != VM_DEPTH_U16) return fail(VM_ERR_INVALID_ARG, "unknown depth_format");
    if (int rc = check_image(depth, pitch, W, H, fmt == VM_DEPTH_F32 ? 4 : 2, "depth")) return rc;
    if (rgb)
        if (int rc = check_image(rgb, rgb_pitch, W, H, 3, "rgb")) return rc;
    if (int rc = check_pose_k(K, T)) return rc;
    float Ti[16];
    if (!inverse4(T, Ti)) return fail(VM_ERR_INVALID_ARG, "camToWorld is singular");
    hipStream_t s = (hipStream_t)stream;
    const Intrinsics k = intrinsics_of(K);
    DepthImage d{(const uint8_t*)depth, pitch, W, H, fmt, scale};
    ColorImage c{(const uint8_t*)rgb, rgb_pitch};
    VM_HIP(launch_allocate(m->view, d, k, affine_of(T), s), "allocation");
    m->epoch++;
    BoundsAtPose bp;
    bool withBounds = false;
    if (int rc = fusion_bounds(m, W, H, K, T, s, &bp, &withBounds)) return rc;
    VM_HIP(launch_visible_fuse(m->view, d, c, k, affine_of(Ti), s, withBounds ? &bp : nullptr),
           "visibility + integration");
    remember_visibility(m, W, H, K, T);
    if (withBounds) remember_bounds(m, W, H, K, T);
    if (out_visible) {
        int n = 0;
        if (int rc = read_counter(m, kVisibleCount, s, &n)) return rc;
        *out_visible = (uint32_t)n;
    }
    return VM_OK;
}

int vm_check_visible_blocks(vm_map* m, int32_t W, int32_t H, const float K[4], const float T[16], void* stream,
                            uint32_t* out_visible) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    if (W <= 0 || H <= 0) return fail(VM_ERR_INVALID_ARG, "width and height must be > 0");
    if (int rc = check_pose_k(K, T)) return rc;
    float Ti[16];
    if (!inverse4(T, Ti)) return fail(VM_ERR_INVALID_ARG, "camToWorld is singular");
    hipStream_t s = (hipStream_t)stream;
    VM_HIP(launch_visible(m->view, W, H, intrinsics_of(K), affine_of(Ti), s), "visibility");
    remember_visibility(m, W, H, K, T);
    if (out_visible) {
        int n = 0;
        if (int rc = read_counter(m, kVisibleCount, s, &n)) return rc;
        *out_visible = (uint32_t)n;
    }
    return VM_OK;
}

int vm_raycast(vm_map* m, float* vmap, int64_t pitch, int32_t W, int32_t H, const float K[4], const float T[16],
               void* stream) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    if (int rc = check_image(vmap, pitch, W, H, 16, "vmap")) return rc;
    if (pitch % 16) return fail(VM_ERR_INVALID_ARG, "vmap pitch must be a multiple of 16");
    if (int rc = check_pose_k(K, T)) return rc;
    float Ti[16];
    if (!inverse4(T, Ti)) return fail(VM_ERR_INVALID_ARG, "camToWorld is singular");
    RayBounds rb;
    const bool rigid = rigid_rotation(T);
    const bool ready = rigid && bounds_current(m, W, H, K, T);
    if (ready) rb = RayBounds{m->rayKeys, m->rayKeysTiles, m->rbEpoch};
    else if (int rc = ray_bounds_for(m, W, H, (hipStream_t)stream, &rb)) return rc;
    VM_HIP(launch_raycast(m->view, vmap, pitch, W, H, intrinsics_of(K), affine_of(T), affine_of(Ti), false,
                          (hipStream_t)stream, rigid ? &rb : nullptr, ready),
           "raycast");
    if (rigid && !ready) remember_bounds(m, W, H, K, T);
    return VM_OK;
}

int vm_render_scene(vm_map* m, float* vmap, int64_t pitch, int32_t W, int32_t H, const float K[4],
                    const float T[16], void* stream) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    if (int rc = check_image(vmap, pitch, W, H, 16, "vmap")) return rc;
    if (pitch % 16) return fail(VM_ERR_INVALID_ARG, "vmap pitch must be a multiple of 16");
    if (int rc = check_pose_k(K, T)) return rc;
    float Ti[16];
    if (!inverse4(T, Ti)) return fail(VM_ERR_INVALID_ARG, "camToWorld is singular");
    hipStream_t s = (hipStream_t)stream;
    const Intrinsics k = intrinsics_of(K);
    if (!visibility_current(m, W, H, K, T)) {
        VM_HIP(launch_visible(m->view, W, H, k, affine_of(Ti), s), "visibility");
        remember_visibility(m, W, H, K, T);
    }
    RayBounds rb;
    const bool rigid = rigid_rotation(T);
    const bool ready = rigid && bounds_current(m, W, H, K, T);
    if (ready) rb = RayBounds{m->rayKeys, m->rayKeysTiles, m->rbEpoch};
    else if (int rc = ray_bounds_for(m, W, H, s, &rb)) return rc;
    VM_HIP(launch_raycast(m->view, vmap, pitch, W, H, k, affine_of(T), affine_of(Ti), true, s,
                          rigid ? &rb : nullptr, ready),
           "raycast");
    if (rigid && !ready) remember_bounds(m, W, H, K, T);
    return VM_OK;
}

int vm_project_rendering_blocks(vm_map* m, float* zx, float* zy, int64_t pitch, int32_t cols8, int32_t rows8,
                                const float K[4], const float Twc[16], void* stream, uint32_t* out_count) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    if (int rc = check_image(zx, pitch, cols8, rows8, 4, "zrange_x")) return rc;
    if (int rc = check_image(zy, pitch, cols8, rows8, 4, "zrange_y")) return rc;
    if (int rc = check_pose_k(K, Twc)) return rc;
    hipStream_t s = (hipStream_t)stream;
    VM_HIP(launch_rendering_blocks(m->view, zx, zy, pitch, cols8, rows8, intrinsics_of(K), affine_of(Twc), s),
           "rendering blocks");
    if (out_count) {
        int n = 0;
        if (int rc = read_counter(m, kRenderBlockCount, s, &n)) return rc;
        *out_count = (uint32_t)n;
    }
    return VM_OK;
}

int vm_compute_normals(const float* vmap, int64_t vp, float* nmap, int64_t np, int32_t W, int32_t H,
                       void* stream) {
    if (int rc = check_image(vmap, vp, W, H, 16, "vmap")) return rc;
    if (int rc = check_image(nmap, np, W, H, 16, "nmap")) return rc;
    if (vp % 16 || np % 16) return fail(VM_ERR_INVALID_ARG, "vmap/nmap pitch must be a multiple of 16");
    VM_HIP(launch_normals(vmap, vp, nmap, np, W, H, (hipStream_t)stream), "normals");
    return VM_OK;
}

int vm_render_shading(const float* vmap, int64_t vp, const float* nmap, int64_t np, uint8_t* rgba, int64_t rp,
                      int32_t W, int32_t H, void* stream) {
    if (int rc = check_image(vmap, vp, W, H, 16, "vmap")) return rc;
    if (int rc = check_image(nmap, np, W, H, 16, "nmap")) return rc;
    if (int rc = check_image(rgba, rp, W, H, 4, "image")) return rc;
    if (vp % 16 || np % 16 || rp % 4) return fail(VM_ERR_INVALID_ARG, "misaligned pitch");
    VM_HIP(launch_shading(vmap, vp, nmap, np, rgba, rp, W, H, (hipStream_t)stream), "shading");
    return VM_OK;
}

int vm_polygonize(vm_map* m, float* verts, float* norms, int64_t cap, void* stream, uint32_t* out_blocks,
                  uint32_t* out_tris) {
    if (int rc = check_polygonize_args(m, verts, norms, cap)) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int nb = m->view.nBlock;
    if (!m->meshWork)
        VM_HIP(hipMalloc((void**)&m->meshWork, polygonize_work_ints(nb, polygonize_groups(nb)) * sizeof(int)),
               "mesh work allocation");
    VM_HIP(launch_polygonize(m->view, m->meshWork, nb, verts, norms, cap, s), "polygonize");
    return polygonize_finish(m, cap, s, out_blocks, out_tris);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// marching cubes without device allocation

namespace {

constexpr int kNoAllocMaxBlocks = VM_NOALLOC_MAX_BLOCKS;
constexpr int kNoAllocMaxGroups = 4096;  // emitting workgroups (1536 on 256 compute units)
constexpr size_t kNoAllocInts = 9 * (size_t)kNoAllocMaxBlocks + 1 + 2 * (size_t)kNoAllocMaxGroups;
constexpr int kMaxDevices = 64;

}  // namespace

// The mesh work buffer of vm_polygonize_noalloc: 18 MiB per device that
// loads the library, held until the process ends. Host code never uses
// this name as an address: each device's copy is found with
// hipGetSymbolAddress (noalloc_scratch below).
static __device__ __align__(256) int scratch[kNoAllocInts];

namespace {

std::mutex g_scratchMutex;
int* g_scratchPtr[kMaxDevices];

// the address of this device's `scratch`, looked up once per device with
// the device current
int noalloc_scratch(int dev, int** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(VM_ERR_INVALID_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "scratch symbol address");
        g_scratchPtr[dev] = (int*)p;
    }
    *out = g_scratchPtr[dev];
    return VM_OK;
}

}  // namespace

extern "C" {

int vm_polygonize_noalloc(vm_map* m, float* verts, float* norms, int64_t cap, void* stream, uint32_t* out_blocks,
                          uint32_t* out_tris) {
    if (int rc = check_polygonize_args(m, verts, norms, cap)) return rc;
    int dev = -1;
    VM_HIP(hipGetDevice(&dev), "current device");
    if (dev != m->device) return fail(VM_ERR_INVALID_ARG, "the map's device must be the current device");
    int* work = nullptr;
    if (int rc = noalloc_scratch(dev, &work)) return rc;
    if (polygonize_groups(m->view.nBlock) > kNoAllocMaxGroups)
        return fail(VM_ERR_INVALID_ARG, "too many compute units for the fixed mesh work array");
    hipStream_t s = (hipStream_t)stream;
    // the allocated blocks must fit the fixed work array
    int heapPtr = 0;
    if (int rc = read_counter(m, kHeapPtr, s, &heapPtr)) return rc;
    const int64_t A = (int64_t)m->view.nBlock - 1 - heapPtr;
    if (A > kNoAllocMaxBlocks)
        return fail(VM_ERR_INVALID_ARG, "vm_polygonize_noalloc: " + std::to_string(A) +
                                            " allocated blocks exceed VM_NOALLOC_MAX_BLOCKS (" +
                                            std::to_string(kNoAllocMaxBlocks) + "); use vm_polygonize");
    VM_HIP(launch_polygonize(m->view, work, kNoAllocMaxBlocks, verts, norms, cap, s), "polygonize");
    return polygonize_finish(m, cap, s, out_blocks, out_tris);
}

int vm_raycast_backward(vm_map* m, const float* vmap, int64_t vpitch, const float* gv, int64_t gvpitch,
                        const float* gn, int64_t gnpitch, int32_t W, int32_t H, const float K[4], const float T[16],
                        float* grad_tsdf, double* grad_pose, void* stream, vm_backward_stats* out_stats) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    if (m->view.fmt != VM_VOXEL_F32_F32_RGB8)
        return fail(VM_ERR_INVALID_ARG, "the raycast adjoint needs the float voxel format");
    if (int rc = check_image(vmap, vpitch, W, H, 16, "vmap")) return rc;
    if (gv)
        if (int rc = check_image(gv, gvpitch, W, H, 16, "grad_vertex")) return rc;
    if (gn)
        if (int rc = check_image(gn, gnpitch, W, H, 16, "grad_normal")) return rc;
    if (vpitch % 16 || (gv && gvpitch % 16) || (gn && gnpitch % 16))
        return fail(VM_ERR_INVALID_ARG, "map pitches must be multiples of 16");
    if (!grad_tsdf || !grad_pose) return fail(VM_ERR_INVALID_ARG, "grad_tsdf and grad_pose must not be NULL");
    if (int rc = check_pose_k(K, T)) return rc;
    float Ti[16];
    if (!inverse4(T, Ti)) return fail(VM_ERR_INVALID_ARG, "camToWorld is singular");
    hipStream_t s = (hipStream_t)stream;
    const size_t need = raycast_backward_scratch_bytes(W, H) + 256;
    if (need > m->scratchBytes) {
        VM_HIP(hipStreamSynchronize(s), "stream synchronize");
        if (m->scratch) VM_HIP(hipFree(m->scratch), "scratch free");
        m->scratch = nullptr;
        m->scratchBytes = 0;
        VM_HIP(hipMalloc(&m->scratch, need), "scratch allocation");
        m->scratchBytes = need;
    }
    int* stats = out_stats ? (int*)((uint8_t*)m->scratch + need - 256) : nullptr;
    if (stats) VM_HIP(launch_fill_u32(stats, 0u, 2, s), "stats reset");
    // the bounds of the forward raycast at this pose when nothing has changed
    // since, else fresh ones (none for a non-rigid pose: the forward raycast
    // marches in full too)
    RayBounds rb{nullptr, 0, 0u};
    const bool rigid = rigid_rotation(T);
    const bool reuse = rigid && bounds_current(m, W, H, K, T);
    if (reuse) {
        rb = RayBounds{m->rayKeys, m->rayKeysTiles, m->rbEpoch};
    } else {
        if (int rc = ray_bounds_for(m, W, H, s, &rb)) return rc;
        if (!rigid) rb.keys = nullptr;
    }
    VM_HIP(launch_raycast_backward(m->view, vmap, vpitch, gv, gvpitch, gn, gnpitch, W, H, intrinsics_of(K),
                                   affine_of(T), affine_of(Ti), grad_tsdf, grad_pose, m->scratch, stats, rb,
                                   !reuse, s),
           "raycast backward");
    if (rigid && !reuse) remember_bounds(m, W, H, K, T);
    if (out_stats) {
        int st[2];
        VM_HIP(hipMemcpyAsync(st, stats, sizeof(st), hipMemcpyDeviceToHost, s), "stats copy");
        VM_HIP(hipStreamSynchronize(s), "stream synchronize");
        out_stats->hit_pixels = st[0];
        out_stats->overflow_pixels = st[1];
    }
    return VM_OK;
}

int vm_refine_poses(vm_map* m, const void* depth, int64_t pitch, int32_t W, int32_t H, int32_t fmt, float scale,
                    const float K[4], const float* init_poses, int32_t n_hyp, int32_t iters, int32_t stride,
                    float max_residual, float max_depth, float* out_poses, float* out_cost, double* out_system,
                    void* stream) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    if (fmt != VM_DEPTH_F32 && fmt != VM_DEPTH_U16) return fail(VM_ERR_INVALID_ARG, "unknown depth_format");
    if (int rc = check_image(depth, pitch, W, H, fmt == VM_DEPTH_F32 ? 4 : 2, "depth")) return rc;
    if (n_hyp < 1 || n_hyp > 4096) return fail(VM_ERR_INVALID_ARG, "n_hyp must be in [1, 4096]");
    if (iters < 0 || iters > 100) return fail(VM_ERR_INVALID_ARG, "iters must be in [0, 100]");
    if (stride < 1) return fail(VM_ERR_INVALID_ARG, "stride must be >= 1");
    if (!(max_residual > 0.f)) return fail(VM_ERR_INVALID_ARG, "max_residual must be > 0");
    if (!(max_depth == max_depth)) return fail(VM_ERR_INVALID_ARG, "max_depth must not be NaN");
    if (!init_poses || !out_poses || !out_cost)
        return fail(VM_ERR_INVALID_ARG, "init_poses, out_poses and out_cost must not be NULL");
    // column-major float -> row-major double; every hypothesis must be finite
    std::vector<double> Tin((size_t)n_hyp * 16);
    for (int h = 0; h < n_hyp; ++h) {
        if (int rc = check_pose_k(K, init_poses + (size_t)h * 16)) return rc;
        // the refinement inverts hypotheses as [R^T, -R^T t]
        if (!rigid_rotation(init_poses + (size_t)h * 16))
            return fail(VM_ERR_INVALID_ARG, "pose hypotheses must be rigid (orthonormal rotation)");
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) Tin[(size_t)h * 16 + i * 4 + j] = init_poses[(size_t)h * 16 + j * 4 + i];
    }
    hipStream_t s = (hipStream_t)stream;
    const size_t need = refine_poses_scratch_bytes(W, H, n_hyp, iters, stride);
    if (need > m->scratchBytes) {
        VM_HIP(hipStreamSynchronize(s), "stream synchronize");
        if (m->scratch) VM_HIP(hipFree(m->scratch), "scratch free");
        m->scratch = nullptr;
        m->scratchBytes = 0;
        VM_HIP(hipMalloc(&m->scratch, need), "scratch allocation");
        m->scratchBytes = need;
    }
    DepthImage d{(const uint8_t*)depth, pitch, W, H, fmt, scale};
    double* posesD = nullptr;
    float* cost = nullptr;
    double* system = nullptr;
    // pixels deeper than max_depth are left out like those beyond depth_max
    MapView view = m->view;
    if (max_depth > 0.f && max_depth < view.depthMax) view.depthMax = max_depth;
    VM_HIP(launch_refine_poses(view, d, intrinsics_of(K), Tin.data(), n_hyp, iters, stride, max_residual, 1e-3,
                               m->scratch, &posesD, &cost, &system, s),
           "pose refinement");
    if (out_system)
        VM_HIP(hipMemcpyAsync(out_system, system, (size_t)n_hyp * (iters + 1) * 27 * sizeof(double),
                              hipMemcpyDeviceToHost, s),
               "system copy");
    std::vector<double> Tout((size_t)n_hyp * 16);
    VM_HIP(hipMemcpyAsync(Tout.data(), posesD, Tout.size() * sizeof(double), hipMemcpyDeviceToHost, s), "pose copy");
    VM_HIP(hipMemcpyAsync(out_cost, cost, (size_t)n_hyp * (iters + 1) * 2 * sizeof(float), hipMemcpyDeviceToHost, s),
           "cost copy");
    VM_HIP(hipStreamSynchronize(s), "stream synchronize");
    for (int h = 0; h < n_hyp; ++h)
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                out_poses[(size_t)h * 16 + j * 4 + i] = (float)Tout[(size_t)h * 16 + i * 4 + j];
    return VM_OK;
}

int vm_surface_points(const float* verts, float* points, int64_t n, void* stream) {
    if (n < 0) return fail(VM_ERR_INVALID_ARG, "n_triangles must be >= 0");
    if (n > 0 && (!verts || !points)) return fail(VM_ERR_INVALID_ARG, "verts and points must not be NULL");
    VM_HIP(launch_surface_points(verts, points, n, (hipStream_t)stream), "surface points");
    return VM_OK;
}

// Voxels of blocks [first, first + n) between the device pool and HOST
// memory. The float format travels as 12-byte records {float sdf; float wt;
// uchar rgb[3]; uchar pad}; on the device its blocks hold the {sdf, wt}
// pairs and the colour words as two arrays (Vox<kFmtF32>), so it is
// converted in chunks through a host buffer. The other formats are copied
// as they are.
static int voxels_transfer(vm_map* m, int64_t first, int64_t n, void* host, bool toHost) {
    const MapView& v = m->view;
    const size_t per = (size_t)voxel_bytes(v.fmt) * kBlockVox;
    uint8_t* dev = v.voxels + (size_t)first * per;
    VM_HIP(hipDeviceSynchronize(), "synchronize");
    if (n <= 0) return VM_OK;
    if (v.fmt != kFmtF32) {
        VM_HIP(hipMemcpy(toHost ? host : (void*)dev, toHost ? (const void*)dev : host, (size_t)n * per,
                         toHost ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice),
               toHost ? "download" : "upload");
        return VM_OK;
    }
    const int64_t chunk = 8192;  // blocks per pass: 48 MiB
    std::vector<uint8_t> tmp((size_t)std::min<int64_t>(n, chunk) * per);
    const size_t colour = (size_t)Vox<kFmtF32>::kColorOffset;
    for (int64_t b0 = 0; b0 < n; b0 += chunk) {
        const int64_t c = std::min<int64_t>(chunk, n - b0);
        uint8_t* h = (uint8_t*)host + (size_t)b0 * per;
        if (toHost) {
            VM_HIP(hipMemcpy(tmp.data(), dev + (size_t)b0 * per, (size_t)c * per, hipMemcpyDeviceToHost), "download");
            for (int64_t k = 0; k < c; ++k) {
                const uint8_t* blk = tmp.data() + (size_t)k * per;
                uint8_t* rec = h + (size_t)k * per;
                for (int l = 0; l < kBlockVox; ++l) {
                    std::memcpy(rec + (size_t)l * 12, blk + (size_t)l * 8, 8);
                    std::memcpy(rec + (size_t)l * 12 + 8, blk + colour + (size_t)l * 4, 4);
                }
            }
        } else {
            for (int64_t k = 0; k < c; ++k) {
                uint8_t* blk = tmp.data() + (size_t)k * per;
                const uint8_t* rec = h + (size_t)k * per;
                for (int l = 0; l < kBlockVox; ++l) {
                    std::memcpy(blk + (size_t)l * 8, rec + (size_t)l * 12, 8);
                    std::memcpy(blk + colour + (size_t)l * 4, rec + (size_t)l * 12 + 8, 4);
                }
            }
            VM_HIP(hipMemcpy(dev + (size_t)b0 * per, tmp.data(), (size_t)c * per, hipMemcpyHostToDevice), "upload");
        }
    }
    return VM_OK;
}

static int array_of(vm_map* m, int which, void** p, int64_t* bytes) {
    const MapView& v = m->view;
    switch (which) {
        case 0: *p = v.table; *bytes = (int64_t)sizeof(HashEntry) * v.nEntry; return VM_OK;
        case 1: *p = v.heap; *bytes = (int64_t)sizeof(int) * v.nBlock; return VM_OK;
        case 2: *p = v.voxels; *bytes = (int64_t)voxel_bytes(v.fmt) * kBlockVox * v.nBlock; return VM_OK;
        case 3: *p = v.ctr; *bytes = (int64_t)sizeof(int) * kCounterInts; return VM_OK;
        case 4: *p = v.visible; *bytes = (int64_t)sizeof(HashEntry) * v.nEntry; return VM_OK;
        case 5: *p = v.blockPos; *bytes = (int64_t)sizeof(int4) * v.nBlock; return VM_OK;
        default: return fail(VM_ERR_INVALID_ARG, "unknown array id");
    }
}

int vm_map_download(vm_map* m, int32_t which, void* dst, int64_t bytes) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    void* p;
    int64_t n;
    if (int rc = array_of(m, which, &p, &n)) return rc;
    if (!dst || bytes != n) return fail(VM_ERR_INVALID_ARG, "host buffer size does not match the array");
    if (which == 2) return voxels_transfer(m, 0, m->view.nBlock, dst, true);
    VM_HIP(hipDeviceSynchronize(), "synchronize");
    VM_HIP(hipMemcpy(dst, p, (size_t)n, hipMemcpyDeviceToHost), "download");
    return VM_OK;
}

int vm_map_upload(vm_map* m, int32_t which, const void* src, int64_t bytes) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    void* p;
    int64_t n;
    if (int rc = array_of(m, which, &p, &n)) return rc;
    if (!src || bytes != n) return fail(VM_ERR_INVALID_ARG, "host buffer size does not match the array");
    // reject state that would make a kernel index outside the map's arrays
    const MapView& v = m->view;
    const int maxOff = v.nEntry - v.nBucket;
    if (which == 0 || which == 4) {
        const HashEntry* e = (const HashEntry*)src;
        for (int i = 0; i < v.nEntry; ++i) {
            const bool ptrOk = e[i].ptr == -1 || (e[i].ptr >= 0 && e[i].ptr % kBlockVox == 0 &&
                                                  e[i].ptr / kBlockVox < v.nBlock);
            if (!ptrOk || e[i].offset < -1 || e[i].offset > maxOff)
                return fail(VM_ERR_INVALID_ARG, "hash entry " + std::to_string(i) + " out of range");
        }
    } else if (which == 1) {
        const int* h = (const int*)src;
        for (int i = 0; i < v.nBlock; ++i)
            if (h[i] < -1 || h[i] >= v.nBlock) return fail(VM_ERR_INVALID_ARG, "heap entry out of range");
    } else if (which == 3) {
        const int* c = (const int*)src;
        if (c[kHeapPtr] < -1 || c[kHeapPtr] >= v.nBlock || c[kExcessPtr] < 0 || c[kExcessPtr] > maxOff ||
            c[kVisibleCount] < 0 || c[kVisibleCount] > v.nEntry)
            return fail(VM_ERR_INVALID_ARG, "counters out of range");
    } else if (which == 5) {
        const int4* b = (const int4*)src;
        for (int i = 0; i < v.nBlock; ++i)
            if (b[i].w < -1 || b[i].w >= v.nEntry) return fail(VM_ERR_INVALID_ARG, "block position out of range");
    }
    if (which == 2) {
        if (int rc = voxels_transfer(m, 0, v.nBlock, const_cast<void*>(src), false)) return rc;
    } else {
        VM_HIP(hipDeviceSynchronize(), "synchronize");
        VM_HIP(hipMemcpy(p, src, (size_t)n, hipMemcpyHostToDevice), "upload");
    }
    if (which == 0) {
        VM_HIP(launch_rebuild_index(m->view, nullptr), "block index rebuild");
        VM_HIP(hipStreamSynchronize(nullptr), "synchronize");
    }
    m->epoch++;
    m->visValid = false;
    return VM_OK;
}

static int voxel_range(vm_map* m, int32_t first, int32_t count, const void* host, uint8_t** dev, size_t* bytes) {
    if (check_map(m)) return VM_ERR_INVALID_ARG;
    const MapView& v = m->view;
    if (!host || first < 0 || count < 0 || first > v.nBlock || count > v.nBlock - first)
        return fail(VM_ERR_INVALID_ARG, "block range outside the voxel pool");
    const size_t per = (size_t)voxel_bytes(v.fmt) * kBlockVox;
    *dev = v.voxels + (size_t)first * per;
    *bytes = (size_t)count * per;
    return VM_OK;
}

int vm_map_download_blocks(vm_map* m, int32_t first_block, int32_t n_blocks, void* dst) {
    uint8_t* p;
    size_t n;
    if (int rc = voxel_range(m, first_block, n_blocks, dst, &p, &n)) return rc;
    return voxels_transfer(m, first_block, n_blocks, dst, true);
}

int vm_map_upload_blocks(vm_map* m, int32_t first_block, int32_t n_blocks, const void* src) {
    uint8_t* p;
    size_t n;
    if (int rc = voxel_range(m, first_block, n_blocks, src, &p, &n)) return rc;
    if (int rc = voxels_transfer(m, first_block, n_blocks, const_cast<void*>(src), false)) return rc;
    m->epoch++;
    m->visValid = false;
    return VM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// multi-scene batches

constexpr int kBatchCtlInts = 512;

struct vm_batch {
    std::vector<vm_map*> maps;
    std::vector<vm_batch_scene> desc;
    std::vector<BatchScene> host;
    BatchScene* dScenes = nullptr;
    Affine* dPoses = nullptr;  // per scene: nFrames camToWorld, then nFrames worldToCam
    unsigned long long* dRayKeys = nullptr;  // per scene: 2 (tiles + 1) ray-bound keys
    int* dFrame = nullptr;  // [0] frame index, [1] steps run (ray-bound epoch - 1); one-kernel step:
                            // [2] status, [32] grid-barrier arrivals, [64] launches run,
                            // [128 + 32 k] ray-tile counters (k < 8)
    std::vector<size_t> keyOff;     // first key of each scene in dRayKeys
    std::vector<size_t> keyHalf;    // near (= far) keys of each scene: tiles + 1
    uint64_t hostSteps = 0;         // steps issued since the keys were last filled
    std::vector<BatchView> groups;  // one view per scene group (contiguous in dScenes)
    BatchTuning tuning{8, 1, 4, 8, 1, 0, 4};
    int fmt = 0, df = 0;
    bool color = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipStream_t> side;  // streams of groups 1.., created on first use
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
};

namespace {

void batch_drop_graph(vm_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// (re)build the per-scene device descriptors for the current tuning. Scenes
// are dealt to sceneGroups groups, largest image first, each to the group
// with the fewest pixels so far; the device array holds the groups one after
// the other, and each scene's tile offsets are relative to its group.
int batch_upload(vm_batch* b) {
    const int n = (int)b->maps.size();
    const int G = std::max(1, std::min(b->tuning.sceneGroups, n));
    std::vector<int> bySize(n);
    for (int i = 0; i < n; ++i) bySize[i] = i;
    std::stable_sort(bySize.begin(), bySize.end(), [&](int x, int y) {
        return (int64_t)b->desc[x].width * b->desc[x].height > (int64_t)b->desc[y].width * b->desc[y].height;
    });
    std::vector<int> groupOf(n);
    std::vector<int64_t> load(G, 0);
    for (int i : bySize) {
        const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        groupOf[i] = g;
        load[g] += (int64_t)b->desc[i].width * b->desc[i].height;
    }
    const int tw = b->tuning.rayTileW, th = 64 / tw;
    std::vector<BatchScene> dev;
    b->groups.clear();
    for (int g = 0; g < G; ++g) {
        BatchView v{};
        v.scenes = b->dScenes + dev.size();
        v.frame = b->dFrame;
        v.steps = (const unsigned*)(b->dFrame + 1);
        int tile16 = 0, ray = 0, chunks = 0, vis = 0, cnt = 0;
        for (int i = 0; i < n; ++i) {
            if (groupOf[i] != g) continue;
            BatchScene S = b->host[i];
            const vm_batch_scene& d = b->desc[i];
            S.tile16First = tile16;
            S.tilesX16 = (d.width + 15) / 16;
            tile16 += S.tilesX16 * ((d.height + 15) / 16);
            S.rayTileFirst = ray;
            ray += ((d.width + tw - 1) / tw) * ((d.height + th - 1) / th);
            chunks += (S.m.nBlock + 255) / 256;
            vis += S.m.nBlock < S.m.nEntry ? S.m.nBlock : S.m.nEntry;
            dev.push_back(S);
            ++cnt;
        }
        v.nScenes = cnt;
        v.totalTile16 = tile16;
        v.totalRayTiles = ray;
        v.rayTilesPerXcd = (ray + 7) / 8;
        v.maxChunks = chunks;
        v.maxVisible = vis;
        b->groups.push_back(v);
    }
    VM_HIP(hipMemcpy(b->dScenes, dev.data(), sizeof(BatchScene) * n, hipMemcpyHostToDevice), "batch upload");
    return VM_OK;
}

PersistCtl batch_ctl(vm_batch* b) {
    return PersistCtl{b->dFrame + 2, (unsigned*)(b->dFrame + 32), (unsigned*)(b->dFrame + 64),
                      (unsigned*)(b->dFrame + 128)};
}

// one step: group 0 on s, group g on side stream g - 1, forked from and
// joined back into s by events (parallel branches when captured), then the
// frame counter increment on s
hipError_t batch_launch_all(vm_batch* b, hipStream_t s) {
    const int G = (int)b->groups.size();
    if (b->tuning.persistent)
        return batch_launch_persistent(b->groups[0], batch_ctl(b), b->tuning, b->fmt, b->df, b->color, s);
    if (G == 1) return batch_launch_step(b->groups[0], b->tuning, b->fmt, b->df, b->color, s);
    hipError_t e = hipSuccess;
    if (!b->fork) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    while (e == hipSuccess && (int)b->side.size() < G - 1) {
        hipStream_t t = nullptr;
        hipEvent_t ev = nullptr;
        e = hipStreamCreateWithFlags(&t, hipStreamNonBlocking);
        if (e == hipSuccess) b->side.push_back(t);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) b->join.push_back(ev);
    }
    if (e != hipSuccess) return e;
    if ((e = hipEventRecord(b->fork, s)) != hipSuccess) return e;
    for (int g = 1; g < G; ++g)
        if ((e = hipStreamWaitEvent(b->side[g - 1], b->fork, 0)) != hipSuccess) return e;
    for (int g = 0; g < G; ++g)
        if ((e = batch_launch_step(b->groups[g], b->tuning, b->fmt, b->df, b->color, g ? b->side[g - 1] : s,
                                   false)) != hipSuccess)
            return e;
    for (int g = 1; g < G; ++g) {
        if ((e = hipEventRecord(b->join[g - 1], b->side[g - 1])) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(s, b->join[g - 1], 0)) != hipSuccess) return e;
    }
    return batch_advance(b->dFrame, s);
}

// Fill every scene's ray-bound keys with "no block" for every epoch (near
// halves all ones, far halves zero) and zero the device step count, on s.
hipError_t batch_reset_keys(vm_batch* b, hipStream_t s) {
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < b->keyOff.size() && e == hipSuccess; ++i) {
        unsigned long long* k = b->dRayKeys + b->keyOff[i];
        const int64_t words = 2 * (int64_t)b->keyHalf[i];  // 32-bit words per half
        e = launch_fill_u32(k, 0xffffffffu, words, s);
        if (e == hipSuccess) e = launch_fill_u32(k + b->keyHalf[i], 0u, words, s);
    }
    if (e == hipSuccess) e = batch_set_frame(b->dFrame + 1, 0, s);
    b->hostSteps = 0;
    return e;
}

void batch_touch_maps(vm_batch* b) {
    for (vm_map* m : b->maps) {
        m->epoch++;
        m->visValid = false;
        m->visCountZeroed = false;
    }
}

}  // namespace

extern "C" {

int vm_batch_create(const vm_batch_scene* scenes, int32_t n, vm_batch** out) {
    if (!scenes || !out) return fail(VM_ERR_INVALID_ARG, "scenes and out_batch must not be NULL");
    *out = nullptr;
    if (n <= 0 || n > batch_max_scenes())
        return fail(VM_ERR_INVALID_ARG, "scene count must be in [1, " + std::to_string(batch_max_scenes()) + "]");
    const int fmt = scenes[0].map ? scenes[0].map->view.fmt : 0;
    const int df = scenes[0].depth_format;
    const bool color = scenes[0].rgb_frames != nullptr;
    size_t nPoses = 0;
    for (int i = 0; i < n; ++i) {
        const vm_batch_scene& d = scenes[i];
        const std::string who = "scene " + std::to_string(i) + ": ";
        if (!d.map) return fail(VM_ERR_INVALID_ARG, who + "map is NULL");
        for (int j = 0; j < i; ++j)
            if (scenes[j].map == d.map) return fail(VM_ERR_INVALID_ARG, who + "map used twice");
        if (d.map->view.fmt != fmt) return fail(VM_ERR_INVALID_ARG, who + "all maps must share a voxel format");
        if (d.depth_format != df) return fail(VM_ERR_INVALID_ARG, who + "all scenes must share a depth format");
        if ((d.rgb_frames != nullptr) != color) return fail(VM_ERR_INVALID_ARG, who + "colour for all or none");
        if (color && fmt == VM_VOXEL_I16_U8) return fail(VM_ERR_INVALID_ARG, who + "format stores no colour");
        if (df != VM_DEPTH_F32 && df != VM_DEPTH_U16) return fail(VM_ERR_INVALID_ARG, who + "unknown depth format");
        if (int rc = check_image(d.depth_frames, d.depth_pitch, d.width, d.height, df == VM_DEPTH_F32 ? 4 : 2,
                                 (who + "depth").c_str()))
            return rc;
        if (color)
            if (int rc = check_image(d.rgb_frames, d.rgb_pitch, d.width, d.height, 3, (who + "rgb").c_str()))
                return rc;
        if (int rc = check_image(d.vmap, d.vmap_pitch, d.width, d.height, 16, (who + "vmap").c_str())) return rc;
        if (int rc = check_image(d.nmap, d.nmap_pitch, d.width, d.height, 16, (who + "nmap").c_str())) return rc;
        if (d.vmap_pitch % 16 || d.nmap_pitch % 16) return fail(VM_ERR_INVALID_ARG, who + "misaligned map pitch");
        if (d.n_frames <= 0 || !d.poses) return fail(VM_ERR_INVALID_ARG, who + "needs frames and poses");
        if (d.depth_frame_stride < d.depth_pitch * d.height || (color && d.rgb_frame_stride < d.rgb_pitch * d.height))
            return fail(VM_ERR_INVALID_ARG, who + "frame stride smaller than a frame");
        if (int rc = check_pose_k(d.K, d.poses)) return rc;
        nPoses += 2 * (size_t)d.n_frames;
    }
    vm_batch* b = new vm_batch;
    b->fmt = fmt;
    b->df = df;
    b->color = color;
    std::vector<Affine> poses(nPoses);
    size_t off = 0;
    for (int i = 0; i < n; ++i) {
        const vm_batch_scene& d = scenes[i];
        for (int f = 0; f < d.n_frames; ++f) {
            const float* T = d.poses + 16 * (size_t)f;
            float Ti[16];
            if (!finite_all(T, 16) || !inverse4(T, Ti)) {
                delete b;
                return fail(VM_ERR_INVALID_ARG, "scene " + std::to_string(i) + ": pose " + std::to_string(f) +
                                                    " is not finite or singular");
            }
            if (!rigid_rotation(T)) {
                delete b;
                return fail(VM_ERR_INVALID_ARG, "scene " + std::to_string(i) + ": pose " + std::to_string(f) +
                                                    " has a rotation that is not orthonormal (|R^T R - I| > 1e-5)");
            }
            poses[off + f] = affine_of(T);
            poses[off + d.n_frames + f] = affine_of(Ti);
        }
        off += 2 * (size_t)d.n_frames;
    }
    // ray-bound keys of every scene (2 (tiles + 1) each), filled with "no
    // block" for every epoch; an epoch is the step count + 1, from 1 on
    size_t nKeys = 0;
    for (int i = 0; i < n; ++i) {
        b->keyOff.push_back(nKeys);
        b->keyHalf.push_back((size_t)ray_bounds_tiles(scenes[i].width, scenes[i].height) + 1);
        nKeys += 2 * b->keyHalf.back();
    }
    hipError_t e = hipMalloc((void**)&b->dPoses, sizeof(Affine) * nPoses);
    if (e == hipSuccess) e = hipMalloc((void**)&b->dRayKeys, sizeof(unsigned long long) * nKeys);
    if (e == hipSuccess) e = hipMalloc((void**)&b->dScenes, sizeof(BatchScene) * n);
    if (e == hipSuccess) e = hipMalloc((void**)&b->dFrame, sizeof(int) * kBatchCtlInts);
    if (e == hipSuccess) e = hipMemcpy(b->dPoses, poses.data(), sizeof(Affine) * nPoses, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(b->dFrame, 0, sizeof(int) * kBatchCtlInts);
    if (e == hipSuccess) e = batch_reset_keys(b, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        vm_batch_destroy(b);
        return hip_fail(e, "batch allocation");
    }
    off = 0;
    for (int i = 0; i < n; ++i) {
        const vm_batch_scene& d = scenes[i];
        b->maps.push_back(d.map);
        b->desc.push_back(d);
        BatchScene S;
        std::memset(&S, 0, sizeof(S));
        S.m = d.map->view;
        S.K = intrinsics_of(d.K);
        S.d = DepthImage{(const uint8_t*)d.depth_frames, d.depth_pitch, d.width, d.height, df, d.depth_scale};
        S.c = ColorImage{(const uint8_t*)d.rgb_frames, d.rgb_pitch};
        S.depthFrameStride = d.depth_frame_stride;
        S.rgbFrameStride = d.rgb_frame_stride;
        S.poses = b->dPoses + off;
        S.posesInv = b->dPoses + off + d.n_frames;
        S.nFrames = d.n_frames;
        S.vmap = d.vmap;
        S.vpitch = d.vmap_pitch;
        S.nmap = d.nmap;
        S.npitch = d.nmap_pitch;
        S.vsInv = (float)(1.0 / (double)S.m.voxelSize);
        S.rtStep = S.m.trunc / S.m.voxelSize;
        S.rayKeys = b->dRayKeys + b->keyOff[i];
        S.rayTiles = ray_bounds_tiles(d.width, d.height);
        b->host.push_back(S);
        off += 2 * (size_t)d.n_frames;
    }
    if (int rc = batch_upload(b)) {
        vm_batch_destroy(b);
        return rc;
    }
    *out = b;
    return VM_OK;
}

int vm_batch_destroy(vm_batch* b) {
    if (!b) return VM_OK;
    batch_drop_graph(b);
    // teardown: failures here leave nothing for the caller to recover
    for (hipStream_t t : b->side) {
        (void)hipStreamSynchronize(t);
        (void)hipStreamDestroy(t);
    }
    for (hipEvent_t ev : b->join) (void)hipEventDestroy(ev);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->dScenes) (void)hipFree(b->dScenes);
    if (b->dPoses) (void)hipFree(b->dPoses);
    if (b->dRayKeys) (void)hipFree(b->dRayKeys);
    if (b->dFrame) (void)hipFree(b->dFrame);
    delete b;
    return VM_OK;
}

int vm_batch_set_frame(vm_batch* b, int32_t frame, void* stream) {
    if (!b) return fail(VM_ERR_INVALID_ARG, "batch is NULL");
    if (frame < 0) return fail(VM_ERR_INVALID_ARG, "frame must be >= 0");
    VM_HIP(batch_set_frame(b->dFrame, frame, (hipStream_t)stream), "set frame");
    return VM_OK;
}

int vm_batch_get_frame(vm_batch* b, void* stream, int32_t* frame) {
    if (!b || !frame) return fail(VM_ERR_INVALID_ARG, "batch and frame must not be NULL");
    hipStream_t s = (hipStream_t)stream;
    int w[3];
    VM_HIP(hipMemcpyAsync(w, b->dFrame, sizeof(w), hipMemcpyDeviceToHost, s), "frame copy");
    VM_HIP(hipStreamSynchronize(s), "stream synchronize");
    *frame = w[0];
    if (w[2])
        return fail(VM_ERR_STATE, "a one-kernel step's grid barrier wait ran out (its workgroups were not all "
                                  "resident): that step's results are invalid");
    return VM_OK;
}

int vm_batch_step(vm_batch* b, void* stream) {
    if (!b) return fail(VM_ERR_INVALID_ARG, "batch is NULL");
    hipStream_t s = (hipStream_t)stream;
    // the 32-bit epoch (step count + 1) must never wrap onto the fill
    // pattern: refill the keys long before it could
    if (++b->hostSteps >= 0xfffff000ull) VM_HIP(batch_reset_keys(b, s), "ray-bound key reset");
    if (b->exec) VM_HIP(hipGraphLaunch(b->exec, s), "graph launch");
    else VM_HIP(batch_launch_all(b, s), "batch step");
    batch_touch_maps(b);
    return VM_OK;
}

int vm_batch_capture(vm_batch* b, void* stream) {
    if (!b) return fail(VM_ERR_INVALID_ARG, "batch is NULL");
    if (!stream) return fail(VM_ERR_INVALID_ARG, "capture needs a non-null stream");
    if (b->tuning.persistent == 2)
        return fail(VM_ERR_STATE, "a cooperative launch is not captured: use persistent 1 for a graph");
    hipStream_t s = (hipStream_t)stream;
    batch_drop_graph(b);
    VM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
    hipError_t e = batch_launch_all(b, s);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(e, "captured step");
    }
    if (e2 != hipSuccess) return hip_fail(e2, "end capture");
    b->graph = g;
    VM_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "graph instantiate");
    return VM_OK;
}

int vm_batch_get_tuning(vm_batch* b, vm_batch_tuning* out) {
    if (!b || !out) return fail(VM_ERR_INVALID_ARG, "batch and out must not be NULL");
    out->ray_tile_width = b->tuning.rayTileW;
    out->xcd_bands = b->tuning.xcdBands;
    out->visible_wg_per_cu = b->tuning.visibleWgPerCu;
    out->integrate_wg_per_cu = b->tuning.integrateWgPerCu;
    out->scene_groups = b->tuning.sceneGroups;
    out->persistent = b->tuning.persistent;
    out->persist_wg_per_cu = b->tuning.persistWgPerCu;
    return VM_OK;
}

int vm_batch_set_tuning(vm_batch* b, const vm_batch_tuning* t) {
    if (!b || !t) return fail(VM_ERR_INVALID_ARG, "batch and tuning must not be NULL");
    if (t->ray_tile_width != 8 && t->ray_tile_width != 16 && t->ray_tile_width != 32)
        return fail(VM_ERR_INVALID_ARG, "ray_tile_width must be 8, 16 or 32");
    if (t->visible_wg_per_cu < 1 || t->visible_wg_per_cu > 64 || t->integrate_wg_per_cu < 1 ||
        t->integrate_wg_per_cu > 64)
        return fail(VM_ERR_INVALID_ARG, "workgroups per CU must be in [1, 64]");
    if (t->scene_groups < 1 || t->scene_groups > 4) return fail(VM_ERR_INVALID_ARG, "scene_groups must be in [1, 4]");
    if (t->persistent < 0 || t->persistent > 2) return fail(VM_ERR_INVALID_ARG, "persistent must be 0, 1 or 2");
    if (t->persistent && t->scene_groups != 1)
        return fail(VM_ERR_INVALID_ARG, "the one-kernel step runs one scene group");
    if (t->persist_wg_per_cu < 1 || t->persist_wg_per_cu > 32)
        return fail(VM_ERR_INVALID_ARG, "persist_wg_per_cu must be in [1, 32]");
    VM_HIP(hipDeviceSynchronize(), "synchronize");  // steps in flight read the scene array
    batch_drop_graph(b);
    b->tuning = BatchTuning{t->ray_tile_width, t->xcd_bands ? 1 : 0, t->visible_wg_per_cu, t->integrate_wg_per_cu,
                            t->scene_groups,   t->persistent,        t->persist_wg_per_cu};
    // the grid-barrier targets assume one grid size since the count was last cleared
    VM_HIP(hipMemset(b->dFrame + 32, 0, sizeof(int) * (kBatchCtlInts - 32)), "barrier clear");
    return batch_upload(b);
}

// Autotuner: time every candidate launch shape over `steps` consecutive
// steps (which fuse frames, like any step) and keep the fastest: 18 launch
// shapes with one scene group, 2..4 groups with the best shape, then the
// one-kernel step (22 trials with 4 or more scenes).
int vm_batch_tune(vm_batch* b, void* stream, int32_t steps, vm_batch_tuning* best_out) {
    if (!b) return fail(VM_ERR_INVALID_ARG, "batch is NULL");
    if (steps < 1) return fail(VM_ERR_INVALID_ARG, "steps must be >= 1");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    VM_HIP(hipEventCreate(&e0), "event");
    VM_HIP(hipEventCreate(&e1), "event");
    vm_batch_tuning best{8, 1, 4, 8, 1, 0, 4};
    float bestMs = 1e30f;
    const int tws[] = {8, 16, 32};
    const int xcds[] = {1, 0};
    const int ints[] = {4, 8, 16};
    int rc = VM_OK;
    // time `steps` steps of tuning t; keep it if fastest so far
    auto trial = [&](const vm_batch_tuning& t) -> int {
        int r;
        if ((r = vm_batch_set_tuning(b, &t))) return r;
        if (hipEventRecord(e0, s) != hipSuccess) return fail(VM_ERR_HIP, "event record");
        for (int k = 0; k < steps; ++k)
            if ((r = vm_batch_step(b, stream))) return r;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            return fail(VM_ERR_HIP, "event timing");
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return fail(VM_ERR_HIP, "event timing");
        if (ms < bestMs) { bestMs = ms; best = t; }
        return VM_OK;
    };
    // the 18 launch shapes with one group, then 2..4 groups with the best shape
    for (int tw : tws)
        for (int xb : xcds)
            for (int iw : ints)
                if ((rc = trial(vm_batch_tuning{tw, xb, 4, iw, 1, 0, 4}))) goto done;
    for (int g = 2; g <= 4 && g <= (int)b->maps.size(); ++g) {
        vm_batch_tuning t = best;
        t.scene_groups = g;
        if ((rc = trial(t))) goto done;
    }
    {  // the one-kernel step with the best ray-tile shape
        vm_batch_tuning t = best;
        t.scene_groups = 1;
        t.persistent = 1;
        t.persist_wg_per_cu = 4;
        if ((rc = trial(t))) goto done;
    }
    rc = vm_batch_set_tuning(b, &best);
    if (rc == VM_OK && best_out) *best_out = best;
done:
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

}  // extern "C"
