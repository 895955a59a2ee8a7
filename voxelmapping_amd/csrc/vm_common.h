// Shared device-side definitions for the gfx950 fusion pipeline.
//
// Arithmetic contract: every floating-point expression keeps the operand
// order of the reference (include/map_struct.h, src/*.cu) and of Eigen >= 3.3
// fixed-size reductions (a0 + (a1 + a2)); the library is compiled with
// -ffp-contract=off and IEEE-rounded division/sqrt, so each expression rounds
// exactly as written. Integer hashing and block/case indexing are exact.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <climits>

namespace vmamd {

constexpr int kBlockSide = 8;    // BLOCK_SIZE, map_struct.h:6
constexpr int kBlockVox = 512;   // BLOCK_SIZE_3, map_struct.h:7
constexpr int kEmptyKey = INT_MIN;  // position of a never-used hash slot
constexpr int kClaimed = -2;        // ptr of a slot claimed by an insert in flight

// HashEntry, map_struct.h:30-35: {int ptr; int offset; Vector3i pos} (20 B)
struct HashEntry {
    int ptr;
    int offset;
    int x, y, z;
};
static_assert(sizeof(HashEntry) == 20, "HashEntry layout");

// Device counter block, one allocation per map. The reference keeps heapPtr,
// excessPtr and visibleBlockNum as separate one-int allocations
// (map_struct.cu:44-52); they live here at fixed indices.
enum Counter : int {
    kHeapPtr = 0,      // top of the free-block stack (map_struct.h:61)
    kExcessPtr = 1,    // next excess-list offset (map_struct.h:60)
    kVisibleCount = 2, // compacted visible entries (map_struct.h:67)
    kStatus = 3,       // sticky status bits, see StatusBits
    kTriangleCount = 4,  // last mesh's triangle count, saturated at INT_MAX
    kRenderBlockCount = 5,
    kFramesFused = 6,
    kSumVisibleLo = 8, // 64-bit running sum of visible counts (stats)
    kSumVisibleHi = 9,
    kTriangleTotalLo = 10,  // 64-bit triangle count of the last mesh
    kTriangleTotalHi = 11,
    kCounterInts = 16,
};

enum StatusBits : int {
    kStatusHeapFull = 1,     // a block allocation found the heap empty
    kStatusExcessFull = 2,   // an excess-list allocation found the list full
    kStatusChainBound = 4,   // a chain walk exceeded the table length
    kStatusMeshClamped = 8,  // polygonize produced more triangles than fit
};

enum VoxelFormat : int {
    kFmtRef = 0,     // vmap::Voxel    {short sdf; uchar wt}           4 B
    kFmtRefRGB = 1,  // vmap::VoxelRGB {short sdf; uchar wt; uchar3}   6 B
    kFmtF32 = 2,     // float sdf, float wt, uchar rgb[3] + pad: 12 B, split per block (Vox<kFmtF32>)
};

// The raycast's ray-distance bounds are kept per 8x8 pixel tile, the rays of
// one wave of the raycast; the multi-scene step keeps them per 16x16 tile
// (fewer tile writes in its bounds pass, which weigh more there than the
// looser limits in its raycast).
constexpr int kBoundsTileShift = 3;
constexpr int kBatchBoundsTileShift = 4;

// Plain-old-data view of one map, passed to kernels by value.
struct MapView {
    HashEntry* table;      // nEntry slots: nBucket heads, then the excess list
    HashEntry* visible;    // compacted copies of visible entries
    int* heap;             // nBlock free-block stack
    int* mutex;            // nBucket bucket locks
    uint8_t* voxels;       // nBlock * 512 voxels of the map's format
    int4* blockPos;        // block id -> {x, y, z, hash slot}; x = kEmptyKey if dead
    int* ctr;              // Counter block
    unsigned long long* occ;  // occupancy bricks, occMask + 1 words (see brick_slot)
    uint32_t occMask;
    int nEntry, nBucket, nBlock;
    int bucketMask;        // nBucket - 1 when nBucket is a power of two, else -1
    float voxelSize, trunc, depthMin, depthMax, maxWeight;
    int fmt;
};

// Per-tile ray-distance bounds of one raycast (see vm_dev_ray.h): 2 (nTiles
// + 1) epoch-stamped 64-bit keys; nTiles is the capacity, at least the
// image's tile count.
struct RayBounds {
    unsigned long long* keys;
    int nTiles;
    unsigned epoch;
};
constexpr int kBoundsChunk = 32;  // allocated blocks per workgroup pass of the single-map bounds kernel

struct Intrinsics {
    float fx, fy, cx, cy;  // K(0,0), K(1,1), K(0,2), K(1,2)
    float ifx, ify;        // (float)(1.0 / K(0,0)): double divide, depth_fusion.cu:209
};

// Affine part of a column-major 4x4: rows of R and t.
struct Affine {
    float r00, r01, r02, t0;
    float r10, r11, r12, t1;
    float r20, r21, r22, t2;
};

enum DepthFormat : int { kDepthF32 = 0, kDepthU16 = 1 };

struct DepthImage {
    const uint8_t* data;
    int64_t pitch;  // bytes between rows
    int width, height;
    int format;     // DepthFormat
    float scale;    // metres per unit for kDepthU16
};

struct ColorImage {
    const uint8_t* data;  // 3 bytes per pixel, or null
    int64_t pitch;
};

// ---------------------------------------------------------------------------
// exact helpers (host + device)

__host__ __device__ inline float sum3(float a, float b, float c) { return a + (b + c); }

__host__ __device__ inline float3 xform(const Affine& T, float3 v) {
    float3 o;
    o.x = T.t0 + sum3(T.r00 * v.x, T.r01 * v.y, T.r02 * v.z);
    o.y = T.t1 + sum3(T.r10 * v.x, T.r11 * v.y, T.r12 * v.z);
    o.z = T.t2 + sum3(T.r20 * v.x, T.r21 * v.y, T.r22 * v.z);
    return o;
}

// float -> int conversion toward zero, NaN -> 0, out-of-range saturating.
// That is what v_cvt_i32_f32 computes, so the device uses the instruction
// itself (a C++ cast of an out-of-range float would be undefined); the host
// spells it out as selects over a clamped cast (2147483520 is the largest
// float below 2^31).
__host__ __device__ inline int f2i_sat(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    int r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
#else
    const int r = (int)fminf(fmaxf(x, -2147483648.0f), 2147483520.0f);
    const int s = x >= 2147483648.0f ? INT_MAX : r;
    return x != x ? 0 : s;
#endif
}
__host__ __device__ inline int floor_i(float x) { return f2i_sat(floorf(x)); }

// hash, map_struct.h:104-109: wrap-around int32 products, C remainder made
// non-negative. For a power-of-two bucket count the remainder is a mask.
__host__ __device__ inline int hash_block(int x, int y, int z, int nBucket, int mask) {
    uint32_t r = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349669u) ^ ((uint32_t)z * 83492791u);
    if (mask >= 0) return (int)(r & (uint32_t)mask);
    int res = (int)r % nBucket;
    return res < 0 ? res + nBucket : res;
}

// Occupancy bricks: one 64-bit word per 4x4x4 voxel blocks, in a table of
// occMask + 1 words indexed by a hash of the brick position. Words are not
// keyed, so bricks may share one; a block's bit is set when the block is
// published and never cleared until reset. A clear bit therefore proves that
// no hash entry holds the block; a set bit is answered by the hash table.
__host__ __device__ inline uint32_t brick_slot(int bx, int by, int bz, uint32_t mask) {
    const uint32_t h = ((uint32_t)(bx >> 2) * 73856093u) ^ ((uint32_t)(by >> 2) * 19349669u) ^
                       ((uint32_t)(bz >> 2) * 83492791u);
    return h & mask;
}
__host__ __device__ inline unsigned long long brick_bit(int bx, int by, int bz) {
    return 1ull << ((bx & 3) | ((by & 3) << 2) | ((bz & 3) << 4));
}
// words for a pool of nBlock blocks: a power of two in [2^12, 2^24]
__host__ __device__ inline uint32_t brick_words(int nBlock) {
    uint32_t w = 1u << 12;
    while (w < (uint32_t)nBlock / 4u && w < (1u << 24)) w <<= 1;
    return w;
}
// set a published block's occupancy bit
__device__ inline void occupancy_mark(const MapView& m, int x, int y, int z) {
    atomicOr(m.occ + brick_slot(x, y, z, m.occMask), brick_bit(x, y, z));
}

// voxelPosToBlockPos (map_struct.h:131-141) is floor division by 8, i.e. an
// arithmetic shift; voxelPosToLocalPos (:148-162) is the Euclidean remainder.
__host__ __device__ inline int vox_to_blk(int v) { return v >> 3; }
__host__ __device__ inline int vox_to_loc(int v) { return v & 7; }
// localPosToLocalIdx, map_struct.h:164-167
__host__ __device__ inline int local_index(int lx, int ly, int lz) { return (lz << 6) | (ly << 3) | lx; }

__host__ __device__ inline int voxel_bytes(int fmt) { return fmt == kFmtRef ? 4 : fmt == kFmtRefRGB ? 6 : 12; }

// Allocated block ids are [0, A) with A = nBlock - 1 - heapPtr (the k-th pop
// of the stack yields block k); clamped so a corrupted counter cannot index
// past the block arrays.
__device__ inline int allocated_count(const int* ctr, int nBlock) {
    const int a = nBlock - 1 - ctr[0];
    return a < 0 ? 0 : (a > nBlock ? nBlock : a);
}

// ---------------------------------------------------------------------------
// voxel storage traits

template <int FMT> struct Vox;

// vmap::Voxel: bytes 0-1 sdf (int16, TSDF * 32767), byte 2 weight, byte 3 pad
template <> struct Vox<kFmtRef> {
    static constexpr int kBytes = 4;
    static constexpr bool kHasColor = false;
    __device__ static inline float sdf_of(int16_t s) { return (float)s / 32767.0f; }
    __device__ static inline void read(const uint8_t* base, int64_t i, float& sdf, bool& valid) {
        uint32_t w = *(const uint32_t*)(base + i * 4);
        sdf = sdf_of((int16_t)(w & 0xffffu));
        valid = ((w >> 16) & 0xffu) != 0;
    }
    // voxel `local` of the block at pool offset ptr (a multiple of 512)
    __device__ static inline void read_local(const uint8_t* base, int ptr, int local, float& sdf, bool& valid) {
        read(base, (int64_t)ptr + local, sdf, valid);
    }
};
// vmap::VoxelRGB: bytes 0-1 sdf, 2 weight, 3-5 rgb (6 B, 2-byte aligned)
template <> struct Vox<kFmtRefRGB> {
    static constexpr int kBytes = 6;
    static constexpr bool kHasColor = true;
    __device__ static inline void read(const uint8_t* base, int64_t i, float& sdf, bool& valid) {
        const uint16_t* p = (const uint16_t*)(base + i * 6);
        sdf = (float)(int16_t)p[0] / 32767.0f;
        valid = (p[1] & 0xffu) != 0;
    }
    __device__ static inline void read_local(const uint8_t* base, int ptr, int local, float& sdf, bool& valid) {
        read(base, (int64_t)ptr + local, sdf, valid);
    }
};
// float TSDF + float weight + rgb8, 12 B per voxel, stored per block as two
// arrays: 512 {sdf, weight} float pairs (4 KiB), then 512 colour words
// {r, g, b, pad} (2 KiB). Voxel i (a block's ptr + local index) is local
// i % 512 of block i / 512. Readers of the field (raycast, adjoint, pose,
// marching cubes) fetch one 8-byte pair and no colour bytes.
template <> struct Vox<kFmtF32> {
    static constexpr int kBytes = 12;
    static constexpr bool kHasColor = true;
    static constexpr int64_t kBlockBytes = (int64_t)kBlockVox * 12;
    static constexpr int kColorOffset = kBlockVox * 8;  // of the colour words within a block
    __host__ __device__ static inline int64_t field_offset(int64_t i) { return (i >> 9) * kBlockBytes + (i & 511) * 8; }
    __device__ static inline void read(const uint8_t* base, int64_t i, float& sdf, bool& valid) {
        const float2 v = *(const float2*)(base + field_offset(i));
        sdf = v.x;
        valid = v.y != 0.f;
    }
    // voxel `local` of the block at pool offset ptr (a multiple of 512): the
    // block's byte offset from its 32-bit id, the voxel's within the block
    __device__ static inline void read_local(const uint8_t* base, int ptr, int local, float& sdf, bool& valid) {
        const float2 v = *(const float2*)(base + (uint64_t)(uint32_t)(ptr >> 9) * (uint64_t)kBlockBytes +
                                          (uint32_t)(local << 3));
        sdf = v.x;
        valid = v.y != 0.f;
    }
};

}  // namespace vmamd
