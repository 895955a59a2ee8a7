// Drop-in replacement of the reference's include/map_struct.h: constants,
// voxel / hash-entry layouts and MapStruct<TVoxel>, backed by the gfx950
// library through include/vmapping_capi.h.
//
// Layouts are byte-identical to the reference (Voxel 4 B, VoxelRGB 6 B,
// HashEntry 20 B); VoxelF32 is an additional float format. MapStruct keeps
// the reference's members; like the reference it is a shallow handle: copies
// alias the same device map and whoever called create() calls release().
#ifndef VMAPPING_MAP_STRUCT_H
#define VMAPPING_MAP_STRUCT_H

#include <stdexcept>
#include <string>

#include "vm_types.h"
#include "vmapping_capi.h"

#define BLOCK_SIZE 8
#define BLOCK_SIZE_3 512
#define BLOCK_SIZE_M1 7
#define MAX_DEPTH 5.f
#define MIN_DEPTH 0.1f

namespace vmap {

// include/map_struct.h:17-21
struct Voxel {
    short sdf;
    unsigned char wt;
};
// include/map_struct.h:23-28
struct VoxelRGB {
    short sdf;
    unsigned char wt;
    Vector3b rgb;
};
// float TSDF / float weight / colour (no counterpart in the reference): the
// record vm_map_download / vm_map_upload transfer. In device memory a
// MapStruct<VoxelF32>'s voxelBlock holds each block as 512 {sdf, wt} float
// pairs followed by 512 {rgb, pad} words (VM_VOXEL_F32_F32_RGB8).
struct VoxelF32 {
    float sdf;
    float wt;
    Vector3b rgb;
    unsigned char pad;
};
// include/map_struct.h:30-35
struct HashEntry {
    int ptr;
    int offset;
    Vector3i pos;
};
// include/map_struct.h:37-42
struct RenderingBlock {
    short upper_left[2];
    short lower_right[2];
    float zrange[2];
};
static_assert(sizeof(Voxel) == 4 && sizeof(VoxelRGB) == 6 && sizeof(VoxelF32) == 12, "voxel layouts");
static_assert(sizeof(HashEntry) == 20, "hash entry layout");

template <class TVoxel> struct VoxelFormatOf;
template <> struct VoxelFormatOf<Voxel> { static constexpr int value = VM_VOXEL_I16_U8; };
template <> struct VoxelFormatOf<VoxelRGB> { static constexpr int value = VM_VOXEL_I16_U8_RGB; };
template <> struct VoxelFormatOf<VoxelF32> { static constexpr int value = VM_VOXEL_F32_F32_RGB8; };

inline void vm_check(int rc) {
    if (rc != VM_OK) throw std::runtime_error(std::string("vmapping: ") + vm_last_error());
}

// include/map_struct.h:44-68
template <class TVoxel>
struct MapStruct {
    void release() {
        if (map) vm_check(vm_map_destroy(map));
        map = nullptr;
        nBucket = 0;
    }
    bool empty() { return nBucket == 0; }
    void reset() { vm_check(vm_map_reset(map, nullptr)); }
    void create(int nEntry_, int nBucket_, int nVBlock_, float voxelSize_, float truncationDist_) {
        vm_map_desc d;
        d.n_entry = nEntry_;
        d.n_bucket = nBucket_;
        d.n_block = nVBlock_;
        d.voxel_size = voxelSize_;
        d.trunc_dist = truncationDist_;
        d.depth_min = MIN_DEPTH;
        d.depth_max = MAX_DEPTH;
        d.voxel_format = VoxelFormatOf<TVoxel>::value;
        d.max_weight = 255.f;
        vm_check(vm_map_create(&d, &map));
        vm_map_arrays a;
        vm_check(vm_map_get_arrays(map, &a));
        heap = a.heap;
        heapPtr = a.heap_ptr;
        excessPtr = a.excess_ptr;
        bucketMutex = a.bucket_mutex;
        voxelBlock = (TVoxel*)a.voxel_block;
        hashTable = (HashEntry*)a.hash_table;
        visibleTable = (HashEntry*)a.visible_table;
        visibleBlockNum = a.visible_block_num;
        nEntry = nEntry_;
        nBucket = nBucket_;
        nVBlock = nVBlock_;
        voxelSize = voxelSize_;
        truncationDist = truncationDist_;
    }
    void getVisibleBlockCount(unsigned int& hostData) {
        vm_map_stats s;
        vm_check(vm_map_get_stats(map, nullptr, &s));
        hostData = (unsigned int)s.visible_blocks;
    }
    void resetVisibleBlockCount() {
        vm_check(hipMemset(visibleBlockNum, 0, sizeof(unsigned int)) == hipSuccess ? VM_OK : VM_ERR_HIP);
    }

    int nBucket = 0;
    int nEntry = 0;
    int nVBlock = 0;
    float voxelSize = 0.f;
    float truncationDist = 0.f;

    int* heap = nullptr;
    int* excessPtr = nullptr;
    int* heapPtr = nullptr;
    int* bucketMutex = nullptr;
    TVoxel* voxelBlock = nullptr;
    HashEntry* hashTable = nullptr;
    HashEntry* visibleTable = nullptr;
    unsigned int* visibleBlockNum = nullptr;

    vm_map* map = nullptr;  // the C-ABI handle behind the arrays
};

}  // namespace vmap

#endif
