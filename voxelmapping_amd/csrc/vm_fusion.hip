// Map reset, block allocation and the fused visibility + TSDF integration
// pass for gfx950.
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
#include "vm_common.h"
#include "vm_launch.h"

namespace vmamd {

namespace {

__device__ inline int ld(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void st(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// ---------------------------------------------------------------------------
// reset (map_struct.cu:7-42)

__global__ void k_reset_table(HashEntry* t, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    t[i].ptr = -1;
    t[i].offset = -1;
    t[i].x = kEmptyKey;
    t[i].y = kEmptyKey;
    t[i].z = kEmptyKey;
}

// heap[i] = nBlock - 1 - i and heapPtr = nBlock - 1, so pops return 0, 1, 2...
__global__ void k_reset_heap(int* heap, int4* blockPos, int* ctr, int* mutex, int nBlock, int nBucket) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nBlock) {
        heap[i] = nBlock - 1 - i;
        blockPos[i] = make_int4(kEmptyKey, kEmptyKey, kEmptyKey, -1);
    }
    if (i < nBucket) mutex[i] = 0;
    if (i < kCounterInts) ctr[i] = (i == kHeapPtr) ? nBlock - 1 : 0;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t n4 = n >> 2;
    uint4 v4 = make_uint4(v, v, v, v);
    for (int64_t j = i; j < n4; j += stride) ((uint4*)p)[j] = v4;
    for (int64_t j = (n4 << 2) + i; j < n; j += stride) p[j] = v;
}

__global__ void k_rebuild_index(MapView m) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.nEntry) return;
    HashEntry e = m.table[i];
    if (e.ptr >= 0 && e.ptr / kBlockVox < m.nBlock) m.blockPos[e.ptr / kBlockVox] = make_int4(e.x, e.y, e.z, i);
}

// ---------------------------------------------------------------------------
// allocation

// Lock-free lookup. Slots are keyed once and never re-keyed, so a matching
// position identifies the slot for good; ptr >= 0 says it is published.
// Returns the slot, -1 when absent, -2 when the chain bound was hit.
__device__ int probe(const MapView& m, int x, int y, int z, int h) {
    int slot = h;
    int guard = m.nEntry - m.nBucket + 1;
    while (true) {
        HashEntry* e = m.table + slot;
        if (ld(&e->x) == x && ld(&e->y) == y && ld(&e->z) == z && ld(&e->ptr) >= 0) return slot;
        int off = ld(&e->offset);
        if (off < 0) return -1;
        if (--guard < 0) {
            atomicOr(&m.ctr[kStatus], kStatusChainBound);
            return -2;
        }
        slot = m.nBucket + off - 1;
    }
}

// AllocateHashEntry's heap pop, map_struct.h:281-295
__device__ int pop_block(const MapView& m) {
    int cur = atomicSub(&m.ctr[kHeapPtr], 1);
    if (cur >= 0) {
        int b = m.heap[cur];
        if (b != -1) return b;
    }
    atomicAdd(&m.ctr[kHeapPtr], 1);
    atomicOr(&m.ctr[kStatus], kStatusHeapFull);
    return -1;
}

__device__ void publish(const MapView& m, int slot, int x, int y, int z, int blk) {
    HashEntry* e = m.table + slot;
    st(&e->x, x);
    st(&e->y, y);
    st(&e->z, z);
    m.blockPos[blk] = make_int4(x, y, z, slot);
    __hip_atomic_store(&e->ptr, blk * kBlockVox, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Body of createBlock run while holding bucket h's lock. Returns false when a
// slot claim lost a race (only possible on the aliased slot nBucket - 1) and
// the insert must be retried.
__device__ bool locked_insert(const MapView& m, int x, int y, int z, int h) {
    int slot = h, empty = -1, tail;
    int guard = m.nEntry - m.nBucket + 1;
    while (true) {
        HashEntry* e = m.table + slot;
        int p = ld(&e->ptr);
        if (p >= 0 && ld(&e->x) == x && ld(&e->y) == y && ld(&e->z) == z) return true;
        if (p == -1 && empty < 0) empty = slot;
        int off = ld(&e->offset);
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

__device__ void insert_block(const MapView& m, int x, int y, int z) {
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

// one thread per pixel, 16x16 pixel tiles
template <int DF>
__global__ __launch_bounds__(256) void k_allocate(MapView m, DepthImage d, Intrinsics K, Affine T) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        m.ctr[kVisibleCount] = 0;
        m.ctr[kFramesFused] += 1;
    }
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= d.width || y >= d.height) return;
    const float dist = load_depth<DF>(d, x, y);
    if (isnan(dist) || dist < m.depthMin || dist > m.depthMax) return;
    const float truncHalf = m.trunc * 0.5f;  // (float)(trunc * 0.5), depth_fusion.cu:226
    const float dNear = fmaxf(m.depthMin, dist - truncHalf);
    const float dFar = fmaxf(m.depthMin, dist + truncHalf);
    if (dNear >= dFar) return;
    // unproject (map_struct.h:189-194), camToWorld, worldPtToVoxelPos (:116-120)
    const float ux = K.ifx * ((float)x - K.cx), uy = K.ify * ((float)y - K.cy);
    float3 wn = xform(T, make_float3(ux * dNear, uy * dNear, dNear));
    float3 wf = xform(T, make_float3(ux * dFar, uy * dFar, dFar));
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
            insert_block(m, bx, by, bz);
            lbx = bx; lby = by; lbz = bz;
        }
        px += dx; py += dy; pz += dz;
    }
}

// ---------------------------------------------------------------------------
// visibility (checkBlockVisible, map_struct.h:204-244)

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

// ---------------------------------------------------------------------------
// integration of one visible block by one wave: lane = (x, y), 8 z-slices
// (DepthFusionFunctor :148-196 / ImageFusionFunctor :328-380)

template <int FMT, int DF, bool COLOR>
__device__ inline void integrate_block(const MapView& m, const DepthImage& d, const ColorImage& c,
                                       const Intrinsics& K, const Affine& Tinv, int blk, int4 key,
                                       int lane) {
    const int lx = lane & 7, ly = lane >> 3;
    const float vs = m.voxelSize, tr = m.trunc;
    const int vx = key.x * 8 + lx, vy = key.y * 8 + ly, vz0 = key.z * 8;
    float sdfNew[8];
    int pix[8];
    unsigned ok = 0;
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        float3 p = xform(Tinv, make_float3((float)vx * vs, (float)vy * vs, (float)(vz0 + z) * vs));
        // __float2int_rd(fx * x / z + cx + 0.5): the double add rounds the exact
        // sum once to float, exactly like a float add
        const int u = floor_i(K.fx * p.x / p.z + K.cx + 0.5f);
        const int v = floor_i(K.fy * p.y / p.z + K.cy + 0.5f);
        pix[z] = 0;
        sdfNew[z] = 0.f;
        if (u < 0 || v < 0 || u > d.width - 1 || v > d.height - 1) continue;
        const float dist = load_depth<DF>(d, u, v);
        if (isnan(dist) || dist > m.depthMax || dist < m.depthMin) continue;
        float sdf = dist - p.z;
        if (sdf < -tr) continue;
        sdfNew[z] = fminf(1.0f, sdf / tr);
        pix[z] = v * 65536 + u;
        ok |= 1u << z;
    }
    uint8_t* base = m.voxels + ((int64_t)blk * kBlockVox + lane) * Vox<FMT>::kBytes;
    constexpr int64_t kSlice = 64 * Vox<FMT>::kBytes;
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        if (!(ok & (1u << z))) continue;
        uint8_t* q = base + z * kSlice;
        const float sdf = sdfNew[z];
        uint8_t cr = 0, cg = 0, cb = 0;
        if constexpr (COLOR) {
            const int u = pix[z] & 0xffff, v = pix[z] >> 16;
            const uint8_t* cp = c.data + (int64_t)v * c.pitch + u * 3;
            cr = cp[0]; cg = cp[1]; cb = cp[2];
        }
        if constexpr (FMT == kFmtF32) {
            float* f = (float*)q;
            const float os = f[0], ow = f[1];
            float ns, nw;
            if (ow == 0.f) { ns = sdf; nw = 1.f; }
            else { ns = (os * ow + sdf * 1.0f) / (ow + 1.0f); nw = fminf(m.maxWeight, ow + 1.0f); }
            f[0] = ns;
            f[1] = nw;
            if constexpr (COLOR) {
                uint32_t* cw = (uint32_t*)(q + 8);
                uint32_t old = *cw;
                uint32_t r = cr, g = cg, b = cb;
                if (ow != 0.f) {
                    r = (uint32_t)(0.7f * (float)(old & 0xffu) + 0.3f * (float)cr);
                    g = (uint32_t)(0.7f * (float)((old >> 8) & 0xffu) + 0.3f * (float)cg);
                    b = (uint32_t)(0.7f * (float)((old >> 16) & 0xffu) + 0.3f * (float)cb);
                }
                *cw = (old & 0xff000000u) | r | (g << 8) | (b << 16);
            }
        } else if constexpr (FMT == kFmtRef) {
            uint32_t* w32 = (uint32_t*)q;
            const uint32_t old = *w32;
            const float os = (float)(int16_t)(old & 0xffffu) / 32767.0f;
            const int ow = (old >> 16) & 0xff;
            float ns;
            int nw;
            if (ow == 0) { ns = sdf; nw = 1; }
            else { ns = (os * (float)ow + sdf * 1.0f) / (float)(ow + 1); nw = min(255, ow + 1); }
            const uint32_t ps = (uint16_t)(int16_t)f2i_sat(ns * 32767.0f);
            *w32 = (old & 0xff000000u) | ((uint32_t)nw << 16) | ps;
        } else {  // kFmtRefRGB
            uint16_t* h = (uint16_t*)q;
            const uint16_t h1 = h[1];
            const float os = (float)(int16_t)h[0] / 32767.0f;
            const int ow = h1 & 0xff;
            float ns;
            int nw;
            if (ow == 0) { ns = sdf; nw = 1; }
            else { ns = (os * (float)ow + sdf * 1.0f) / (float)(ow + 1); nw = min(255, ow + 1); }
            h[0] = (uint16_t)(int16_t)f2i_sat(ns * 32767.0f);
            if constexpr (COLOR) {
                const uint16_t h2 = h[2];
                uint32_t r = cr, g = cg, b = cb;
                if (ow != 0) {
                    r = (uint32_t)(0.7f * (float)(h1 >> 8) + 0.3f * (float)cr);
                    g = (uint32_t)(0.7f * (float)(h2 & 0xffu) + 0.3f * (float)cg);
                    b = (uint32_t)(0.7f * (float)(h2 >> 8) + 0.3f * (float)cb);
                }
                h[1] = (uint16_t)(nw | (r << 8));
                h[2] = (uint16_t)(g | (b << 8));
            } else {
                h[1] = (uint16_t)((h1 & 0xff00u) | (uint32_t)nw);
            }
        }
    }
}

// Visibility over the allocated blocks [0, A) instead of the whole hash
// table: the allocated ids are exactly the popped heap entries. Each 256-block
// chunk is tested (one block per lane), compacted with wave ballots, appended
// to the visible table with one atomic per chunk, and then (FUSE) integrated
// by the workgroup's four waves. Persistent grid: the block count stays on
// the device.
template <bool FUSE, int FMT, int DF, bool COLOR>
__global__ __launch_bounds__(256) void k_visible(MapView m, DepthImage d, ColorImage c, Intrinsics K,
                                                 Affine Tinv, int cols, int rows) {
    __shared__ int s_wave[4];
    __shared__ int s_base;
    __shared__ int s_blk[256];
    __shared__ int4 s_key[256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int A = allocated_count(m.ctr, m.nBlock);
    long long visSum = 0;
    for (int chunk = blockIdx.x * 256; chunk < A; chunk += gridDim.x * 256) {
        const int b = chunk + threadIdx.x;
        bool vis = false;
        int4 key = make_int4(0, 0, 0, -1);
        if (b < A) {
            key = m.blockPos[b];
            if (key.x != kEmptyKey) vis = block_visible(m, key.x, key.y, key.z, Tinv, K, cols, rows);
        }
        const unsigned long long mask = __ballot(vis);
        if (lane == 0) s_wave[wave] = __popcll(mask);
        __syncthreads();
        const int w0 = s_wave[0], w1 = s_wave[1], w2 = s_wave[2], w3 = s_wave[3];
        const int total = w0 + w1 + w2 + w3;
        const int woff = (wave > 0 ? w0 : 0) + (wave > 1 ? w1 : 0) + (wave > 2 ? w2 : 0);
        if (threadIdx.x == 0 && total > 0) s_base = atomicAdd(&m.ctr[kVisibleCount], total);
        __syncthreads();
        if (vis) {
            const int r = woff + __popcll(mask & ((1ull << lane) - 1ull));
            s_blk[r] = b;
            s_key[r] = key;
            HashEntry e;
            e.ptr = b * kBlockVox;
            e.offset = m.table[key.w].offset;
            e.x = key.x; e.y = key.y; e.z = key.z;
            m.visible[s_base + r] = e;
        }
        visSum += total;
        if constexpr (FUSE) {
            __syncthreads();
            for (int i = wave; i < total; i += 4)
                integrate_block<FMT, DF, COLOR>(m, d, c, K, Tinv, s_blk[i], s_key[i], lane);
        }
        __syncthreads();
    }
    if constexpr (FUSE) {
        if (threadIdx.x == 0 && visSum > 0)
            atomicAdd((unsigned long long*)&m.ctr[kSumVisibleLo], (unsigned long long)visSum);
    }
}

int grid_for_blocks(const MapView& m) {
    int cu = num_compute_units();
    int maxChunks = (m.nBlock + 255) / 256;
    int g = cu * 8;
    return g < maxChunks ? g : (maxChunks > 0 ? maxChunks : 1);
}

template <int FMT, int DF, bool COLOR>
hipError_t launch_vf(const MapView& m, const DepthImage& d, const ColorImage& c, const Intrinsics& K,
                     const Affine& Tinv, hipStream_t s) {
    hipLaunchKernelGGL((k_visible<true, FMT, DF, COLOR>), dim3(grid_for_blocks(m)), dim3(256), 0, s, m, d, c,
                       K, Tinv, d.width, d.height);
    return hipGetLastError();
}

template <int FMT, int DF>
hipError_t launch_vf_c(const MapView& m, const DepthImage& d, const ColorImage& c, const Intrinsics& K,
                       const Affine& Tinv, hipStream_t s) {
    if (c.data && Vox<FMT>::kHasColor) return launch_vf<FMT, DF, Vox<FMT>::kHasColor>(m, d, c, K, Tinv, s);
    return launch_vf<FMT, DF, false>(m, d, c, K, Tinv, s);
}

template <int FMT>
hipError_t launch_vf_d(const MapView& m, const DepthImage& d, const ColorImage& c, const Intrinsics& K,
                       const Affine& Tinv, hipStream_t s) {
    if (d.format == kDepthU16) return launch_vf_c<FMT, kDepthU16>(m, d, c, K, Tinv, s);
    return launch_vf_c<FMT, kDepthF32>(m, d, c, K, Tinv, s);
}

}  // namespace

int num_compute_units() {
    static int cu = 0;
    if (cu == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0)
            cu = 256;
    }
    return cu;
}

hipError_t launch_fill_u32(void* dst, uint32_t value, int64_t count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    int64_t blocks = (count / 4 + 255) / 256;
    int g = (int)(blocks < 4096 ? (blocks > 0 ? blocks : 1) : 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(256), 0, s, (uint32_t*)dst, value, count);
    return hipGetLastError();
}

hipError_t launch_reset(const MapView& m, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_table, dim3((m.nEntry + 255) / 256), dim3(256), 0, s, m.table, m.nEntry);
    int n = m.nBlock > m.nBucket ? m.nBlock : m.nBucket;
    if (n < kCounterInts) n = kCounterInts;
    hipLaunchKernelGGL(k_reset_heap, dim3((n + 255) / 256), dim3(256), 0, s, m.heap, m.blockPos, m.ctr, m.mutex,
                       m.nBlock, m.nBucket);
    int64_t bytes = (int64_t)m.nBlock * kBlockVox * voxel_bytes(m.fmt);
    hipError_t e = launch_fill_u32(m.voxels, 0u, bytes / 4, s);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t launch_rebuild_index(const MapView& m, hipStream_t s) {
    hipLaunchKernelGGL(k_rebuild_index, dim3((m.nEntry + 255) / 256), dim3(256), 0, s, m);
    return hipGetLastError();
}

hipError_t launch_allocate(const MapView& m, const DepthImage& d, const Intrinsics& K, const Affine& T,
                           hipStream_t s) {
    dim3 grid((d.width + 15) / 16, (d.height + 15) / 16);
    if (d.format == kDepthU16)
        hipLaunchKernelGGL(k_allocate<kDepthU16>, grid, dim3(256), 0, s, m, d, K, T);
    else
        hipLaunchKernelGGL(k_allocate<kDepthF32>, grid, dim3(256), 0, s, m, d, K, T);
    return hipGetLastError();
}

hipError_t launch_visible_fuse(const MapView& m, const DepthImage& d, const ColorImage& c, const Intrinsics& K,
                               const Affine& Tinv, hipStream_t s) {
    switch (m.fmt) {
        case kFmtRef: return launch_vf_d<kFmtRef>(m, d, c, K, Tinv, s);
        case kFmtRefRGB: return launch_vf_d<kFmtRefRGB>(m, d, c, K, Tinv, s);
        default: return launch_vf_d<kFmtF32>(m, d, c, K, Tinv, s);
    }
}

hipError_t launch_visible(const MapView& m, int cols, int rows, const Intrinsics& K, const Affine& Tinv,
                          hipStream_t s) {
    // a zero-count reset first: the allocation kernel is not run here
    hipError_t e = launch_fill_u32(m.ctr + kVisibleCount, 0u, 1, s);
    if (e != hipSuccess) return e;
    DepthImage d{nullptr, 0, cols, rows, kDepthF32, 1.f};
    ColorImage c{nullptr, 0};
    hipLaunchKernelGGL((k_visible<false, kFmtRef, kDepthF32, false>), dim3(grid_for_blocks(m)), dim3(256), 0, s,
                       m, d, c, K, Tinv, cols, rows);
    return hipGetLastError();
}

}  // namespace vmamd
