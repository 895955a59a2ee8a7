// Device-side descriptors of the multi-scene pipeline (vm_batch.hip).
#pragma once

#include "vm_common.h"

namespace vmamd {

// one scene of a batch, in device memory
struct BatchScene {
    MapView m;
    Intrinsics K;
    DepthImage d;             // data = frame 0; frame f at data + f * depthFrameStride
    ColorImage c;             // data = frame 0 or null
    int64_t depthFrameStride;
    int64_t rgbFrameStride;
    const Affine* poses;      // [nFrames] camera-to-world
    const Affine* posesInv;   // [nFrames] world-to-camera
    int nFrames;
    float* vmap;
    int64_t vpitch;
    float* nmap;
    int64_t npitch;
    float vsInv, rtStep;      // (float)(1.0 / voxelSize), trunc / voxelSize
    int tile16First;          // first 16x16 tile of this scene in the flat tile space
    int tilesX16;
    int rayTileFirst;         // first ray tile of this scene
    // per-tile ray-distance bounds (vm_dev_ray.h) of the scene's pixel
    // tiles; the epoch of a step is the batch's step count + 1
    unsigned long long* rayKeys;
    int rayTiles;
};

struct BatchView {
    const BatchScene* scenes;  // device array
    int nScenes;
    int* frame;                // device-side frame index
    const unsigned* steps;     // device-side count of steps run, never rewound: the ray-bound epoch of the
                               // next step is steps + 1, so keys of an earlier step never look current,
                               // whatever frame the batch is set back to
    int totalTile16;
    int totalRayTiles;
    int rayTilesPerXcd;
    int maxChunks;             // sum over scenes of ceil(nBlock / 256)
    int maxVisible;            // sum over scenes of min(nBlock, nEntry)
};

// launch shape picked by the autotuner
struct BatchTuning {
    int rayTileW;           // ray tile width: 8 (8x8), 16 (16x4) or 32 (32x2)
    int xcdBands;           // remap ray tiles into per-XCD contiguous bands
    int visibleWgPerCu;     // visibility workgroups per compute unit
    int integrateWgPerCu;   // integration workgroups (4 waves each) per CU
    int sceneGroups;        // scenes split into this many groups, each on its own stream
    int persistent;         // 0: five kernels; 1: one kernel with grid barriers; 2: the same, cooperative launch
    int persistWgPerCu;     // resident workgroups per CU of the one-kernel step (capped by occupancy)
};

// device words of the one-kernel step
struct PersistCtl {
    int* status;            // bit 0: a grid barrier wait ran out
    unsigned* arrivals;     // grid-barrier arrivals, never reset
    unsigned* seq;          // launches run (the barrier targets' base)
    unsigned* queues;       // ray-tile counters, one per XCD band, kQueueStride apart
};
constexpr int kQueueStride = 32;  // one 128-byte line per counter

// the five stage kernels of one step over the scenes of B; with advance, then
// the frame counter increment
hipError_t batch_launch_step(const BatchView& B, const BatchTuning& t, int fmt, int df, bool color, hipStream_t s,
                             bool advance = true);
// the same step as one kernel (stages separated by grid barriers), which also
// advances the frame counter; needs one scene group
hipError_t batch_launch_persistent(const BatchView& B, const PersistCtl& P, const BatchTuning& t, int fmt, int df,
                                   bool color, hipStream_t s);
hipError_t batch_advance(int* frame, hipStream_t s);
hipError_t batch_set_frame(int* frame, int v, hipStream_t s);
int batch_max_scenes();

}  // namespace vmamd
