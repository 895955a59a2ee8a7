// Device functions of allocation, visibility and integration, shared by the
// single-map kernels (vm_fusion.hip) and the multi-scene kernels
// (vm_batch.hip).
//
// Allocation (CreateBlockLineTracingFunctor, depth_fusion.cu:8-60 with
// createBlock, map_struct.h:304-354) keeps the reference's table layout:
// bucket heads in slots [0, nBucket), excess offset o in slot nBucket + o - 1,
// a stack heap whose k-th pop yields block k. Two things differ from the
// reference's try-lock, both so that no block is ever lost:
//   * a bucket whose lock is busy is retried instead of dropped;
//   * slots are claimed by CAS on `ptr`, so excess offset 0 (which lands on
//     the last bucket head) is skipped when that head is already occupied
//     instead of overwriting it.
// The allocated key set therefore equals the union of traced blocks (given
// capacity), which is what the serial reading of the reference produces.
#pragma once

#include "vm_common.h"

namespace vmamd {

__device__ inline int ld(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void st(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Lock-free lookup. Slots are keyed once and never re-keyed, so a matching
// position identifies the slot for good; ptr >= 0 says it is published.
// Returns the slot, -1 when absent, -2 when the chain bound was hit. The five
// fields of an entry are read together (one memory round trip per entry;
// relaxed loads of different words are unordered either way).
__device__ inline int probe(const MapView& m, int x, int y, int z, int h) {
    int slot = h;
    int guard = m.nEntry - m.nBucket + 1;
    while (true) {
        HashEntry* e = m.table + slot;
        const int ex = ld(&e->x), ey = ld(&e->y), ez = ld(&e->z), p = ld(&e->ptr), off = ld(&e->offset);
        if ((ex == x) & (ey == y) & (ez == z) & (p >= 0)) return slot;
        if (off < 0) return -1;
        if (--guard < 0) {
            atomicOr(&m.ctr[kStatus], kStatusChainBound);
            return -2;
        }
        slot = m.nBucket + off - 1;
    }
}

// AllocateHashEntry's heap pop, map_struct.h:281-295
__device__ inline int pop_block(const MapView& m) {
    int cur = atomicSub(&m.ctr[kHeapPtr], 1);
    if (cur >= 0) {
        int b = m.heap[cur];
        if (b != -1) return b;
    }
    atomicAdd(&m.ctr[kHeapPtr], 1);
    atomicOr(&m.ctr[kStatus], kStatusHeapFull);
    return -1;
}

__device__ inline void publish(const MapView& m, int slot, int x, int y, int z, int blk) {
    HashEntry* e = m.table + slot;
    st(&e->x, x);
    st(&e->y, y);
    st(&e->z, z);
    m.blockPos[blk] = make_int4(x, y, z, slot);
    occupancy_mark(m, x, y, z);
    __hip_atomic_store(&e->ptr, blk * kBlockVox, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Body of createBlock run while holding bucket h's lock. Returns false when a
// slot claim lost a race (only possible on the aliased slot nBucket - 1) and
// the insert must be retried.
__device__ inline bool locked_insert(const MapView& m, int x, int y, int z, int h) {
    int slot = h, empty = -1, tail;
    int guard = m.nEntry - m.nBucket + 1;
    while (true) {
        HashEntry* e = m.table + slot;
        const int p = ld(&e->ptr), ex = ld(&e->x), ey = ld(&e->y), ez = ld(&e->z), off = ld(&e->offset);
        if ((p >= 0) & (ex == x) & (ey == y) & (ez == z)) return true;
        if (p == -1 && empty < 0) empty = slot;
        if (off < 0) { tail = slot; break; }
        if (--guard < 0) {
            atomicOr(&m.ctr[kStatus], kStatusChainBound);
            return true;
        }
        slot = m.nBucket + off - 1;
    }
    if (empty >= 0) {
        // map_struct.h:329-336: reuse the first empty entry of the chain
        if (atomicCAS(&m.table[empty].ptr, -1, kClaimed) != -1) return false;
        int blk = pop_block(m);
        if (blk < 0) { st(&m.table[empty].ptr, -1); return true; }
        publish(m, empty, x, y, z, blk);
        return true;
    }
    // map_struct.h:337-353: new excess entry linked behind the chain tail
    while (true) {
        int off = atomicAdd(&m.ctr[kExcessPtr], 1);
        if (off + m.nBucket >= m.nEntry) {
            atomicSub(&m.ctr[kExcessPtr], 1);
            atomicOr(&m.ctr[kStatus], kStatusExcessFull);
            return true;
        }
        int s2 = m.nBucket + off - 1;
        if (atomicCAS(&m.table[s2].ptr, -1, kClaimed) != -1) continue;  // aliased, occupied head
        int blk = pop_block(m);
        if (blk < 0) { st(&m.table[s2].ptr, -1); return true; }
        publish(m, s2, x, y, z, blk);
        // link behind the tail; a chain that merged through the aliased slot
        // can gain a tail concurrently, so the link is a CAS that follows on
        int t = tail, g = m.nEntry - m.nBucket + 1;
        while (g-- > 0) {
            int old = atomicCAS(&m.table[t].offset, -1, off);
            if (old == -1) break;
            t = m.nBucket + old - 1;
        }
        return true;
    }
}

// createBlock with a retried lock: the lane that wins the bucket lock runs
// the critical section and releases it in the same loop iteration, so lanes
// of one wave contending for one bucket cannot deadlock.
__device__ inline void insert_block(const MapView& m, int x, int y, int z) {
    int h = hash_block(x, y, z, m.nBucket, m.bucketMask);
    if (probe(m, x, y, z, h) != -1) return;
    bool done = false;
    while (!done) {
        if (atomicCAS(&m.mutex[h], 0, 1) == 0) {
            done = locked_insert(m, x, y, z, h);
            __threadfence();
            atomicExch(&m.mutex[h], 0);
        } else if (probe(m, x, y, z, h) != -1) {
            done = true;  // the lock holder inserted it
        }
    }
}

template <int DF>
__device__ inline float load_depth(const DepthImage& d, int x, int y) {
    const uint8_t* row = d.data + (int64_t)y * d.pitch;
    if constexpr (DF == kDepthF32) return ((const float*)row)[x];
    else return (float)((const uint16_t*)row)[x] * d.scale;
}

// CreateBlockLineTracingFunctor::operator() for pixel (x, y),
// depth_fusion.cu:32-59, run by a whole wave whose lanes cover a 16 x 4
// pixel tile (lane = 16 * row + column): every lane must call this, a lane
// without a pixel with inImage = false. Each lane walks its segment's blocks
// in order, consecutive duplicates being no-ops as in the reference loop, and
// keeps its first kLaneKeys distinct blocks (later ones are inserted at
// once). Neighbouring pixels cross mostly the same blocks, so a lane skips a
// kept block that the lane to its left or above also keeps: the lowest lane
// keeping a block never skips it, so the inserted set is still the union of
// the lanes' blocks, with a hash probe per block and tile row run instead of
// per pixel.
constexpr int kLaneKeys = 4;

template <int DF>
__device__ __forceinline__ void allocate_pixel(const MapView& m, const DepthImage& d, const Intrinsics& K, const Affine& T,
                                      int x, int y, bool inImage) {
    const int lane = threadIdx.x & 63;
    int kx[kLaneKeys], ky[kLaneKeys], kz[kLaneKeys];
#pragma unroll
    for (int k = 0; k < kLaneKeys; ++k) kx[k] = ky[k] = kz[k] = 0;
    int n = 0;
    if (inImage) {
        const float dist = load_depth<DF>(d, x, y);
        const float truncHalf = m.trunc * 0.5f;  // (float)(trunc * 0.5), depth_fusion.cu:226
        const float dNear = fmaxf(m.depthMin, dist - truncHalf);
        const float dFar = fmaxf(m.depthMin, dist + truncHalf);
        if (!(isnan(dist) || dist < m.depthMin || dist > m.depthMax) && dNear < dFar) {
            // unproject (map_struct.h:189-194), camToWorld, worldPtToVoxelPos (:116-120)
            const float ux = K.ifx * ((float)x - K.cx), uy = K.ify * ((float)y - K.cy);
            const float3 wn = xform(T, make_float3(ux * dNear, uy * dNear, dNear));
            const float3 wf = xform(T, make_float3(ux * dFar, uy * dFar, dFar));
            const float vs = m.voxelSize;
            const int pnx = floor_i(wn.x / vs), pny = floor_i(wn.y / vs), pnz = floor_i(wn.z / vs);
            const int pfx = floor_i(wf.x / vs), pfy = floor_i(wf.y / vs), pfz = floor_i(wf.z / vs);
            float dx = (float)(pfx - pnx), dy = (float)(pfy - pny), dz = (float)(pfz - pnz);
            const float len = sqrtf(sum3(dx * dx, dy * dy, dz * dz));
            // (int)ceil(2.0 * len): 2*len is exact in float, so ceilf matches the double form
            const int steps = f2i_sat(ceilf(2.0f * len));
            const float den = (float)(steps - 1);
            dx = dx / den; dy = dy / den; dz = dz / den;
            float px = (float)pnx, py = (float)pny, pz = (float)pnz;
            int lbx = INT_MIN, lby = INT_MIN, lbz = INT_MIN;
            for (int s = 0; s < steps; ++s) {
                // pt.cast<int>() truncates toward zero (depth_fusion.cu:58)
                const int bx = vox_to_blk(f2i_sat(px)), by = vox_to_blk(f2i_sat(py)), bz = vox_to_blk(f2i_sat(pz));
                if (bx != lbx || by != lby || bz != lbz) {  // consecutive duplicates are no-ops
                    if (n < kLaneKeys) {
#pragma unroll
                        for (int k = 0; k < kLaneKeys; ++k)
                            if (k == n) { kx[k] = bx; ky[k] = by; kz[k] = bz; }
                        ++n;
                    } else {
                        insert_block(m, bx, by, bz);
                    }
                    lbx = bx; lby = by; lbz = bz;
                }
                px += dx; py += dy; pz += dz;
            }
        }
    }
    // the kept blocks of the lanes to the left (same row) and above
    const int nl = __shfl_up(n, 1, 64), nu = __shfl_up(n, 16, 64);
    const bool hasL = (lane & 15) != 0, hasU = lane >= 16;
    bool skip[kLaneKeys];
#pragma unroll
    for (int k = 0; k < kLaneKeys; ++k) skip[k] = false;
#pragma unroll
    for (int j = 0; j < kLaneKeys; ++j) {
        const int lx = __shfl_up(kx[j], 1, 64), ly = __shfl_up(ky[j], 1, 64), lz = __shfl_up(kz[j], 1, 64);
        const int ux = __shfl_up(kx[j], 16, 64), uy = __shfl_up(ky[j], 16, 64), uz = __shfl_up(kz[j], 16, 64);
        const bool vl = hasL && j < nl, vu = hasU && j < nu;
#pragma unroll
        for (int k = 0; k < kLaneKeys; ++k)
            skip[k] = skip[k] || (vl && kx[k] == lx && ky[k] == ly && kz[k] == lz) ||
                      (vu && kx[k] == ux && ky[k] == uy && kz[k] == uz);
    }
#pragma unroll
    for (int k = 0; k < kLaneKeys; ++k)
        if (k < n && !skip[k]) insert_block(m, kx[k], ky[k], kz[k]);
}

// checkBlockVisible, map_struct.h:204-244
__device__ inline bool block_visible(const MapView& m, int bx, int by, int bz, const Affine& Tinv,
                                     const Intrinsics& K, int cols, int rows) {
    const float scale = m.voxelSize * 8.0f;
    const float maxu = (float)(cols - 1), maxv = (float)(rows - 1);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float3 w = make_float3((float)(bx + (c & 1)) * scale, (float)(by + ((c >> 1) & 1)) * scale,
                               (float)(bz + ((c >> 2) & 1)) * scale);
        float3 p = xform(Tinv, w);
        float u = K.fx * p.x / p.z + K.cx;
        float v = K.fy * p.y / p.z + K.cy;
        bool out = u < 0.f || v < 0.f || u > maxu || v > maxv || p.z < m.depthMin || p.z > m.depthMax;
        if (!out) return true;
    }
    return false;
}

// Integration of one visible block by one wave: lane = (x, y), 8 z-slices
// (DepthFusionFunctor :148-196 / ImageFusionFunctor :328-380).
//
// Every read is issued before any of its results is needed, so a lane has
// two memory round trips instead of one or two per slice: the eight depth
// reads (a projection outside the image reads pixel (0, 0), unused), then
// the eight voxels of the lane's column and the eight colour pixels
// together. Voxels whose depth test fails are read but never written.
template <int FMT, int DF, bool COLOR>
__device__ __forceinline__ void integrate_block(const MapView& m, const DepthImage& d, const ColorImage& c,
                                       const Intrinsics& K, const Affine& Tinv, int blk, int4 key, int lane) {
    const int lx = lane & 7, ly = lane >> 3;
    const float vs = m.voxelSize, tr = m.trunc;
    const int vx = key.x * 8 + lx, vy = key.y * 8 + ly, vz0 = key.z * 8;
    float dist[8], pz[8];
    int pix[8];
    unsigned inside = 0;
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        const float3 p = xform(Tinv, make_float3((float)vx * vs, (float)vy * vs, (float)(vz0 + z) * vs));
        // __float2int_rd(fx * x / z + cx + 0.5): the double add rounds the exact
        // sum once to float, exactly like a float add
        const int u = floor_i(K.fx * p.x / p.z + K.cx + 0.5f);
        const int v = floor_i(K.fy * p.y / p.z + K.cy + 0.5f);
        const bool in = !(u < 0 || v < 0 || u > d.width - 1 || v > d.height - 1);
        const int uc = in ? u : 0, vc = in ? v : 0;
        inside |= (unsigned)in << z;
        pix[z] = vc * 65536 + uc;
        pz[z] = p.z;
        dist[z] = load_depth<DF>(d, uc, vc);
    }
    float sdfNew[8];
    unsigned ok = 0;
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        const float dd = dist[z];
        const float sdf = dd - pz[z];
        const bool good = ((inside >> z) & 1u) && !(isnan(dd) || dd > m.depthMax || dd < m.depthMin) && !(sdf < -tr);
        sdfNew[z] = fminf(1.0f, sdf / tr);
        ok |= (unsigned)good << z;
    }
    if (ok == 0) return;
    uint32_t rgb[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        rgb[z] = 0;
        if constexpr (COLOR) {
            const int u = pix[z] & 0xffff, v = pix[z] >> 16;
            const uint8_t* cp = c.data + (int64_t)v * c.pitch + u * 3;
            rgb[z] = (uint32_t)cp[0] | ((uint32_t)cp[1] << 8) | ((uint32_t)cp[2] << 16);
        }
    }
    uint8_t* base = m.voxels + ((int64_t)blk * kBlockVox + lane) * Vox<FMT>::kBytes;
    constexpr int64_t kSlice = 64 * Vox<FMT>::kBytes;
    if constexpr (FMT == kFmtF32) {
        // the block's {sdf, weight} pairs and colour words, z-slice z at + 64 z
        uint8_t* blockBase = m.voxels + (int64_t)blk * Vox<kFmtF32>::kBlockBytes;
        float2* fld = (float2*)blockBase + lane;
        uint32_t* col = (uint32_t*)(blockBase + Vox<kFmtF32>::kColorOffset) + lane;
        float2 ov[8];
        uint32_t oc[8];
#pragma unroll
        for (int z = 0; z < 8; ++z) {
            ov[z] = fld[z * 64];
            oc[z] = COLOR ? col[z * 64] : 0u;
        }
#pragma unroll
        for (int z = 0; z < 8; ++z) {
            if (!((ok >> z) & 1u)) continue;
            const float sdf = sdfNew[z];
            const float os = ov[z].x, ow = ov[z].y;
            float ns, nw;
            if (ow == 0.f) { ns = sdf; nw = 1.f; }
            else { ns = (os * ow + sdf * 1.0f) / (ow + 1.0f); nw = fminf(m.maxWeight, ow + 1.0f); }
            fld[z * 64] = make_float2(ns, nw);
            if constexpr (COLOR) {
                const uint32_t old = oc[z];
                uint32_t r = rgb[z] & 0xffu, g = (rgb[z] >> 8) & 0xffu, b = rgb[z] >> 16;
                if (ow != 0.f) {
                    r = (uint32_t)(0.7f * (float)(old & 0xffu) + 0.3f * (float)r);
                    g = (uint32_t)(0.7f * (float)((old >> 8) & 0xffu) + 0.3f * (float)g);
                    b = (uint32_t)(0.7f * (float)((old >> 16) & 0xffu) + 0.3f * (float)b);
                }
                col[z * 64] = (old & 0xff000000u) | r | (g << 8) | (b << 16);
            }
        }
    } else if constexpr (FMT == kFmtRef) {
        uint32_t ov[8];
#pragma unroll
        for (int z = 0; z < 8; ++z) ov[z] = *(const uint32_t*)(base + z * kSlice);
#pragma unroll
        for (int z = 0; z < 8; ++z) {
            if (!((ok >> z) & 1u)) continue;
            const uint32_t old = ov[z];
            const float os = (float)(int16_t)(old & 0xffffu) / 32767.0f;
            const int ow = (old >> 16) & 0xff;
            const float sdf = sdfNew[z];
            float ns;
            int nw;
            if (ow == 0) { ns = sdf; nw = 1; }
            else { ns = (os * (float)ow + sdf * 1.0f) / (float)(ow + 1); nw = min(255, ow + 1); }
            const uint32_t ps = (uint16_t)(int16_t)f2i_sat(ns * 32767.0f);
            *(uint32_t*)(base + z * kSlice) = (old & 0xff000000u) | ((uint32_t)nw << 16) | ps;
        }
    } else {  // kFmtRefRGB
        uint16_t h0[8], h1[8], h2[8];
#pragma unroll
        for (int z = 0; z < 8; ++z) {
            const uint16_t* h = (const uint16_t*)(base + z * kSlice);
            h0[z] = h[0];
            h1[z] = h[1];
            h2[z] = COLOR ? h[2] : (uint16_t)0;
        }
#pragma unroll
        for (int z = 0; z < 8; ++z) {
            if (!((ok >> z) & 1u)) continue;
            uint16_t* h = (uint16_t*)(base + z * kSlice);
            const float os = (float)(int16_t)h0[z] / 32767.0f;
            const int ow = h1[z] & 0xff;
            const float sdf = sdfNew[z];
            float ns;
            int nw;
            if (ow == 0) { ns = sdf; nw = 1; }
            else { ns = (os * (float)ow + sdf * 1.0f) / (float)(ow + 1); nw = min(255, ow + 1); }
            h[0] = (uint16_t)(int16_t)f2i_sat(ns * 32767.0f);
            if constexpr (COLOR) {
                uint32_t r = rgb[z] & 0xffu, g = (rgb[z] >> 8) & 0xffu, b = rgb[z] >> 16;
                if (ow != 0) {
                    r = (uint32_t)(0.7f * (float)(h1[z] >> 8) + 0.3f * (float)r);
                    g = (uint32_t)(0.7f * (float)(h2[z] & 0xffu) + 0.3f * (float)g);
                    b = (uint32_t)(0.7f * (float)(h2[z] >> 8) + 0.3f * (float)b);
                }
                h[1] = (uint16_t)(nw | (r << 8));
                h[2] = (uint16_t)(g | (b << 8));
            } else {
                h[1] = (uint16_t)((h1[z] & 0xff00u) | (uint32_t)nw);
            }
        }
    }
}

// LDS of one 256-thread workgroup running the visibility chunks
struct VisibleShared {
    int wave[4];
    int base;
};

// One 256-block chunk of the allocated list [chunk, chunk + 256): test each
// block (one per lane) and compact the visible ones with wave ballots into
// the map's visible table (one atomic per chunk). Returns the chunk's visible
// count. Must be called by all 256 threads of the workgroup.
__device__ __forceinline__ int visible_chunk(const MapView& m, const Intrinsics& K, const Affine& Tinv, int cols, int rows,
                                    int chunk, int A, VisibleShared& sh) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int b = chunk + threadIdx.x;
    bool vis = false;
    int4 key = make_int4(0, 0, 0, -1);
    if (b < A) {
        key = m.blockPos[b];
        if (key.x != kEmptyKey) vis = block_visible(m, key.x, key.y, key.z, Tinv, K, cols, rows);
    }
    const unsigned long long mask = __ballot(vis);
    if (lane == 0) sh.wave[wave] = __popcll(mask);
    __syncthreads();
    const int w0 = sh.wave[0], w1 = sh.wave[1], w2 = sh.wave[2], w3 = sh.wave[3];
    const int total = w0 + w1 + w2 + w3;
    const int woff = (wave > 0 ? w0 : 0) + (wave > 1 ? w1 : 0) + (wave > 2 ? w2 : 0);
    if (threadIdx.x == 0 && total > 0) sh.base = atomicAdd(&m.ctr[kVisibleCount], total);
    __syncthreads();
    if (vis) {
        const int r = woff + __popcll(mask & ((1ull << lane) - 1ull));
        HashEntry e;
        e.ptr = b * kBlockVox;
        e.offset = m.table[key.w].offset;
        e.x = key.x; e.y = key.y; e.z = key.z;
        m.visible[sh.base + r] = e;
    }
    __syncthreads();
    return total;
}

}  // namespace vmamd
