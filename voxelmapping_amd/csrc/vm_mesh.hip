// Marching cubes (BuildVertexArray / Polygonize, marching_cubes.cu:9-305),
// triangle centroids (:314-344), vertex-map normals and Phong shading
// (image_proc.cu:8-104) for gfx950.
//
// Marching cubes: four launches. k_mc_links resolves every allocated block's
// 7 +x/+y/+z neighbours (all hash probes in flight at once); k_mc_count and
// k_mc_emit read each block's 8x8x8 voxels plus its neighbours' face layers
// (a 9^3 halo) into 256-thread workgroups' LDS, next to the LDS-resident
// case tables. The counting pass classifies the cubes and records per-block
// triangle counts, k_mc_scan sums them per emitting workgroup (a fixed set
// of blocks each) and turns the sums into workgroup offsets, and the
// emitting pass builds the triangles of the blocks that have any, spread
// evenly over the workgroup's threads, and stores them through LDS as
// contiguous dwords. Nothing contends for a shared counter (one atomic per
// block on a single address cost more than the rest of the kernel), and the
// output order is a function of the map, where the reference's depends on
// atomic timing; triangle contents, winding and normals follow :209-242
// exactly.
#include "vm_common.h"
#include "vm_dev_ray.h"
#include "vm_launch.h"
#include "vm_mc_edge.h"
#include "vm_mc_table.h"

namespace vmamd {

namespace {

__constant__ McTables c_mc = kMcTables;

constexpr int kHalo = 9 * 9 * 9;
constexpr int kMcThreads = 256;           // one workgroup (4 waves) per block
constexpr int kMcMaxTri = kBlockVox * 5;  // at most 5 triangles per cube
constexpr int kMcGroupsPerCu = 6;         // emit pass: ~23.7 KB of LDS per workgroup

// Allocated blocks the mesh passes take: the allocated count, at most the
// capacity of the work buffer (the caller checks it fits).
__device__ inline int mesh_blocks(const int* ctr, int nBlock, int capBlocks) {
    const int a = allocated_count(ctr, nBlock);
    return a < capBlocks ? a : capBlocks;
}

// Voxel offsets of every allocated block and its 7 +x/+y/+z neighbours
// (links[8 b + n], n = dx | dy << 1 | dz << 2; -1: absent; links[8 b] = -2:
// dead block), one thread per (block, neighbour): the hash probes of the
// mesh, all in flight at once.
__global__ void k_mc_links(MapView m, int* links, int capBlocks) {
    const int A = mesh_blocks(m.ctr, m.nBlock, capBlocks);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)A * 8;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int b = (int)(i >> 3), n = (int)(i & 7);
        const int4 key = m.blockPos[b];
        int p;
        if (key.x == kEmptyKey) p = n == 0 ? -2 : -1;
        else if (n == 0) p = b * kBlockVox;
        else p = find_block(m, key.x + (n & 1), key.y + ((n >> 1) & 1), key.z + (n >> 2));
        links[i] = p;
    }
}

// Blocks of workgroup w of a g-workgroup counting or emitting pass (g a
// multiple of 8). The dispatcher places workgroup w on XCD w % 8; XCD x
// takes the contiguous block-id range [A x / 8, A (x + 1) / 8) -- blocks
// allocated together, mostly neighbours in space -- and its workgroups step
// through it side by side, so the face layers a block reads from its
// neighbours were mostly fetched into that XCD's own L2 moments before.
struct McRange {
    int first, end, step;
};
__device__ inline McRange mc_range(int w, int g, int A) {
    const int x = w & 7;
    const int lo = (int)((int64_t)A * x / 8), hi = (int)((int64_t)A * (x + 1) / 8);
    return McRange{lo + (w >> 3), hi, g >> 3};
}

// LDS of the counting and emitting passes
struct McShared {
    float sdf[kHalo];
    uint8_t ok[kHalo];
    uint8_t cases[kBlockVox];
    // the case tables, read with per-thread case indices, stay in LDS for
    // the workgroup's whole life (loaded once)
    uint64_t tri[256];
    // the current block's neighbour offsets and key (x, y, z, triangle
    // count), double-buffered by iteration parity: a wave that skips a
    // block cannot overwrite what a slower wave still reads
    int nb[2][8];
    int key[2][4];
    int wsum[kMcThreads / 64];
};

__device__ inline void mc_load_tables(McShared& sh, int tid) {
    for (int i = tid; i < 256; i += kMcThreads) sh.tri[i] = c_mc.tri[i];
}

// the 9^3 halo of the block whose neighbour offsets are sh.nb[ph], in three
// independent rounds of loads; ends with a barrier
template <int FMT>
__device__ inline void mc_load_halo(const MapView& m, McShared& sh, int ph, int tid) {
    float sv[3];
    bool okv[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int i = tid + r * kMcThreads;
        sv[r] = 1.f;
        okv[r] = false;
        if (i < kHalo) {
            const int hx = i % 9, hy = (i / 9) % 9, hz = i / 81;
            const int p = sh.nb[ph][(hx >> 3) | ((hy >> 3) << 1) | ((hz >> 3) << 2)];
            if (p >= 0) Vox<FMT>::read(m.voxels, (int64_t)p + local_index(hx & 7, hy & 7, hz & 7), sv[r], okv[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int i = tid + r * kMcThreads;
        if (i < kHalo) {
            sh.sdf[i] = sv[r];
            sh.ok[i] = okv[r];
        }
    }
    __syncthreads();
}

// corner offsets in Bourke order (read_sdf_list, marching_cubes.cu:68-104)
__device__ inline int mc_corner(int k) {
    return k == 0 ? 0 : k == 1 ? 1 : k == 2 ? 1 + 9 : k == 3 ? 9 : k == 4 ? 81 : k == 5 ? 1 + 81 : k == 6 ? 1 + 9 + 81 : 9 + 81;
}

// Cases of cubes tid and tid + 256 (cube c: x = c & 7, y = (c >> 3) & 7,
// z = c >> 6) into sh.cases, their triangle counts into nt, then a
// block-wide scan: returns the thread's first triangle in the block and sets
// the block's total. Ends with a barrier.
__device__ inline int mc_classify(McShared& sh, int tid, int nt[2], int& total) {
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int c = tid + q * kMcThreads;
        const int h0 = (c & 7) + 9 * ((c >> 3) & 7) + 81 * (c >> 6);
        int cube = 0;
        bool all = true;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            all = all && sh.ok[h0 + mc_corner(k)];
            if (sh.sdf[h0 + mc_corner(k)] < 0.f) cube |= 1 << k;
        }
        if (!all) cube = 0;
        sh.cases[c] = (uint8_t)cube;
        nt[q] = (int)(sh.tri[cube] & 15u);  // 0 for cases 0 and 255 (edgeTable 0)
    }
    const int mine = nt[0] + nt[1];
    int incl = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
    }
    if (lane == 63) sh.wsum[wave] = incl;
    __syncthreads();
    int off = incl - mine;
    total = 0;
#pragma unroll
    for (int w = 0; w < kMcThreads / 64; ++w) {
        const int v = sh.wsum[w];
        off += w < wave ? v : 0;
        total += v;
    }
    return off;
}

// Counting pass: the triangle count of every block (blockTri[b]), over a
// grid of its own.
template <int FMT>
__global__ __launch_bounds__(kMcThreads) void k_mc_count(MapView m, const int* links, int* blockTri, int capBlocks) {
    __shared__ McShared sh;
    const int tid = threadIdx.x;
    mc_load_tables(sh, tid);
    const int A = mesh_blocks(m.ctr, m.nBlock, capBlocks);
    const McRange r = mc_range(blockIdx.x, gridDim.x, A);
    // per-lane (vector) loads: their wait is at the use, one block later
    int nbNext = -1;
    if (r.first < r.end && tid < 8) nbNext = links[(int64_t)r.first * 8 + tid];
    int ph = 0;
    for (int b = r.first; b < r.end; b += r.step, ph ^= 1) {
        const int nb = nbNext;
        const int bn = b + r.step;
        if (bn < r.end && tid < 8) nbNext = links[(int64_t)bn * 8 + tid];
        if (tid < 8) sh.nb[ph][tid] = nb;
        __syncthreads();
        if (sh.nb[ph][0] == -2) {  // dead block
            if (tid == 0) blockTri[b] = 0;
            continue;
        }
        mc_load_halo<FMT>(m, sh, ph, tid);
        int nt[2], total;
        mc_classify(sh, tid, nt, total);
        if (tid == 0) blockTri[b] = total;
    }
}

// Offsets of the emitting pass's g workgroups (one 1024-thread workgroup):
// workgroup w takes the blocks of mc_range(w, g, A), so its count is the sum
// of theirs; an exclusive scan over w gives the offsets, the sum the
// triangle count.
// Offsets and the total are 64-bit (2^22 blocks of up to 2560 triangles
// exceed 2^32); a workgroup's own sum, at most 2560 per block of its share,
// fits 32 bits.
__global__ __launch_bounds__(1024) void k_mc_scan(const int* blockTri, int g, unsigned long long* wgBase, int* ctr,
                                                  int nBlock, int capBlocks) {
    __shared__ unsigned long long s_w[16];
    __shared__ unsigned long long s_carry;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int A = mesh_blocks(ctr, nBlock, capBlocks);
    if (tid == 0) s_carry = 0ull;
    __syncthreads();
    for (int c0 = 0; c0 < g; c0 += 1024) {
        const int i = c0 + tid;
        unsigned v = 0u;
        if (i < g) {
            const McRange r = mc_range(i, g, A);
            for (int b = r.first; b < r.end; b += r.step) v += (unsigned)blockTri[b];
        }
        unsigned long long incl = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned long long x = __shfl_up(incl, d, 64);
            if (lane >= d) incl += x;
        }
        if (lane == 63) s_w[wave] = incl;
        __syncthreads();
        unsigned long long off = s_carry, tot = 0ull;
#pragma unroll
        for (int w = 0; w < 16; ++w) {
            const unsigned long long x = s_w[w];
            off += w < wave ? x : 0ull;
            tot += x;
        }
        if (i < g) wgBase[i] = off + incl - v;
        __syncthreads();
        if (tid == 0) s_carry += tot;
        __syncthreads();
    }
    if (tid == 0) {
        const unsigned long long t = s_carry;
        ctr[kTriangleCount] = t > 0x7fffffffull ? 0x7fffffff : (int)t;
        ctr[kTriangleTotalLo] = (int)(unsigned)t;
        ctr[kTriangleTotalHi] = (int)(unsigned)(t >> 32);
    }
}

// Emitting pass, one 256-thread workgroup per block, workgroup w taking
// the blocks of mc_range(w, grid, A) (the order k_mc_scan assumes): blocks without
// triangles are skipped; for the others
// the halo is read and classified again, an LDS map lists (cube, k) per
// triangle, each thread builds one triangle per pass of 256, staged in LDS,
// and the pass is stored as contiguous dwords (coalesced) at the
// workgroup's own offset, up to `cap`. Triangle order is a function of the
// map alone: workgroup, then block, then cube, then the case's order.
template <int FMT>
__global__ __launch_bounds__(kMcThreads) void k_mc_emit(MapView m, const int* links, const int* blockTri,
                                                         const unsigned long long* wgBase, float* verts, float* norms,
                                                         int64_t cap, int capBlocks) {
    __shared__ McShared sh;
    __shared__ uint16_t s_map[kMcMaxTri];  // triangle -> (cube << 3) | k
    __shared__ float s_v[kMcThreads * 9];
    __shared__ float s_n[kMcThreads * 3];
    const int tid = threadIdx.x;
    mc_load_tables(sh, tid);
    const int A = mesh_blocks(m.ctr, m.nBlock, capBlocks);
    const float vs = m.voxelSize;
    int64_t base = (int64_t)wgBase[blockIdx.x];
    const McRange r = mc_range(blockIdx.x, gridDim.x, A);
    // per-lane (vector) loads: their wait is at the use, one block later
    int nbNext = -1, kcNext = 0;
    auto prefetch = [&](int bb) {
        if (tid < 8) nbNext = links[(int64_t)bb * 8 + tid];
        if (tid < 3) kcNext = ((const int*)m.blockPos)[(int64_t)bb * 4 + tid];
        if (tid == 3) kcNext = blockTri[bb];
    };
    if (r.first < r.end) prefetch(r.first);
    int ph = 0;
    for (int b = r.first; b < r.end; b += r.step, ph ^= 1) {
        const int nb = nbNext, kc = kcNext;
        if (b + r.step < r.end) prefetch(b + r.step);
        if (tid < 8) sh.nb[ph][tid] = nb;
        if (tid < 4) sh.key[ph][tid] = kc;
        __syncthreads();
        if (sh.key[ph][3] == 0) continue;  // no triangles (dead blocks included)
        const int kx = sh.key[ph][0], ky = sh.key[ph][1], kz = sh.key[ph][2];
        mc_load_halo<FMT>(m, sh, ph, tid);
        int nt[2], total;
        const int off = mc_classify(sh, tid, nt, total);
        for (int k = 0; k < nt[0]; ++k) s_map[off + k] = (uint16_t)((tid << 3) | k);
        for (int k = 0; k < nt[1]; ++k) s_map[off + nt[0] + k] = (uint16_t)(((tid + kMcThreads) << 3) | k);
        __syncthreads();
        for (int j0 = 0; j0 < total; j0 += kMcThreads) {
            const int j = j0 + tid;
            if (j < total) {
                const int e = s_map[j], c = e >> 3, k = e & 7;
                const uint64_t tri = sh.tri[sh.cases[c]];
                const int lx = c & 7, ly = (c >> 3) & 7, lz = c >> 6;
                const float* hs = sh.sdf + lx + 9 * ly + 81 * lz;
                const float3 pos = make_float3((float)(kx * 8 + lx), (float)(ky * 8 + ly), (float)(kz * 8 + lz));
                const int ea = (int)((tri >> (4 + 12 * k)) & 15u);
                const int eb = (int)((tri >> (8 + 12 * k)) & 15u);
                const int ec = (int)((tri >> (12 + 12 * k)) & 15u);
                const float3 va = mc_edge_vertex(ea, pos, hs), vb = mc_edge_vertex(eb, pos, hs),
                             vc = mc_edge_vertex(ec, pos, hs);
                // stored reversed: (c, b, a), scaled to metres
                const float3 p0 = make_float3(vc.x * vs, vc.y * vs, vc.z * vs);
                const float3 p1 = make_float3(vb.x * vs, vb.y * vs, vb.z * vs);
                const float3 p2 = make_float3(va.x * vs, va.y * vs, va.z * vs);
                const float3 n = normalized3(cross3(make_float3(p1.x - p0.x, p1.y - p0.y, p1.z - p0.z),
                                                    make_float3(p2.x - p0.x, p2.y - p0.y, p2.z - p0.z)));
                float* o = s_v + tid * 9;
                o[0] = p0.x; o[1] = p0.y; o[2] = p0.z;
                o[3] = p1.x; o[4] = p1.y; o[5] = p1.z;
                o[6] = p2.x; o[7] = p2.y; o[8] = p2.z;
                s_n[tid * 3 + 0] = n.x;
                s_n[tid * 3 + 1] = n.y;
                s_n[tid * 3 + 2] = n.z;
            }
            __syncthreads();
            // triangles [base + j0, base + j0 + n) of this pass, clipped at cap
            const int64_t first = base + j0;
            int64_t n = total - j0 < kMcThreads ? total - j0 : kMcThreads;
            if (cap - first < n) n = cap - first;
            const int words = n > 0 ? (int)n * 9 : 0;
            float* ov = verts + first * 9;
            float* on = norms + first * 9;
            for (int i = tid; i < words; i += kMcThreads) {
                const int t = i / 9, c = i - t * 9;
                ov[i] = s_v[i];
                on[i] = s_n[t * 3 + c % 3];
            }
            __syncthreads();
        }
        base += total;
    }
}

__global__ void k_surface_points(const float* v, float* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* a = v + i * 9;
#pragma unroll
    for (int j = 0; j < 3; ++j) p[i * 3 + j] = ((a[j] + a[3 + j]) + a[6 + j]) / 3.0f;
}

// computeNormalKernel, image_proc.cu:74-93
__global__ void k_normals(const float* vmap, int64_t vpitch, float* nmap, int64_t npitch, int W, int H) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W - 1 || y >= H - 1) return;
    normal_pixel(vmap, vpitch, nmap, npitch, W, H, x, y);
}

__device__ inline float saturate01(float v) { return v != v ? 0.f : fminf(1.f, fmaxf(0.f, v)); }
__device__ inline float dot3(float3 a, float3 b) { return sum3(a.x * b.x, a.y * b.y, a.z * b.z); }

// renderPoint / renderSceneKernel, image_proc.cu:8-61; light at (5, 5, 5)
__global__ void k_shading(const float* vmap, int64_t vpitch, const float* nmap, int64_t npitch, uint8_t* rgba,
                          int64_t rpitch, int W, int H) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const float4 P4 = *((const float4*)((const uint8_t*)vmap + (int64_t)y * vpitch) + x);
    const float4 N4 = *((const float4*)((const uint8_t*)nmap + (int64_t)y * npitch) + x);
    const float3 P = make_float3(P4.x, P4.y, P4.z), N = make_float3(N4.x, N4.y, N4.z);
    float3 col = make_float3(4.f / 255.f, 2.f / 255.f, 2.f / 255.f);
    if (!isnan(P.x)) {
        const float Ka = 0.3f, Kd = 0.5f, Ks = 0.2f, n = 20.f, Ax = 1.f, Dx = 1.f, Sx = 1.f, Lx = 1.f;
        const float3 L = normalized3(make_float3(5.f - P.x, 5.f - P.y, 5.f - P.z));
        const float3 V = normalized3(make_float3(0.f - P.x, 0.f - P.y, 0.f - P.z));
        const float nl = dot3(N, L);
        const float3 R = normalized3(make_float3(2.f * N.x * nl - L.x, 2.f * N.y * nl - L.y, 2.f * N.z * nl - L.z));
        const float I = Ax * Ka * Dx + Lx * Kd * Dx * fmaxf(0.f, nl) + Lx * Ks * Sx * powf(fmaxf(0.f, dot3(R, V)), n);
        col = make_float3(I, I, I);
    }
    uchar4 o;
    o.x = (uint8_t)(saturate01(col.x) * 255.f);
    o.y = (uint8_t)(saturate01(col.y) * 255.f);
    o.z = (uint8_t)(saturate01(col.z) * 255.f);
    o.w = 255;
    *((uchar4*)(rgba + (int64_t)y * rpitch) + x) = o;
}

}  // namespace

// workgroups of a counting or emitting pass: perCu per compute unit, fewer
// for a small pool, a multiple of 8 (mc_range)
static int mc_groups(int nBlock, int perCu) {
    int g = num_compute_units() * perCu;
    if (g > nBlock) g = nBlock;
    return ((g + 7) / 8) * 8;
}

int polygonize_groups(int nBlock) { return mc_groups(nBlock, kMcGroupsPerCu); }

// links (8 per block of capacity), blockTri (1 per block), then the emitting
// workgroups' 64-bit offsets at an even int index
size_t polygonize_work_ints(int capBlocks, int groups) { return (size_t)capBlocks * 9 + 1 + 2 * (size_t)groups; }

hipError_t launch_polygonize(const MapView& m, int* work, int capBlocks, float* verts, float* norms, int64_t cap,
                             hipStream_t s) {
    const int g = polygonize_groups(m.nBlock);
    const int gc = mc_groups(m.nBlock, 8);  // counting pass: 36 VGPRs, ~6.3 KB of LDS
    int* links = work;
    int* blockTri = work + (size_t)capBlocks * 8;
    unsigned long long* wgBase = (unsigned long long*)(work + (((size_t)capBlocks * 9 + 1) & ~(size_t)1));
    hipLaunchKernelGGL(k_mc_links, dim3(num_compute_units() * 8), dim3(256), 0, s, m, links, capBlocks);
    const dim3 block(kMcThreads);
    switch (m.fmt) {
        case kFmtRef:
            hipLaunchKernelGGL(k_mc_count<kFmtRef>, dim3(gc), block, 0, s, m, links, blockTri, capBlocks);
            break;
        case kFmtRefRGB:
            hipLaunchKernelGGL(k_mc_count<kFmtRefRGB>, dim3(gc), block, 0, s, m, links, blockTri, capBlocks);
            break;
        default: hipLaunchKernelGGL(k_mc_count<kFmtF32>, dim3(gc), block, 0, s, m, links, blockTri, capBlocks);
    }
    hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(1024), 0, s, blockTri, g, wgBase, m.ctr, m.nBlock, capBlocks);
    const dim3 grid(g);
    switch (m.fmt) {
        case kFmtRef:
            hipLaunchKernelGGL(k_mc_emit<kFmtRef>, grid, block, 0, s, m, links, blockTri, wgBase, verts, norms, cap,
                               capBlocks);
            break;
        case kFmtRefRGB:
            hipLaunchKernelGGL(k_mc_emit<kFmtRefRGB>, grid, block, 0, s, m, links, blockTri, wgBase, verts, norms,
                               cap, capBlocks);
            break;
        default:
            hipLaunchKernelGGL(k_mc_emit<kFmtF32>, grid, block, 0, s, m, links, blockTri, wgBase, verts, norms, cap,
                               capBlocks);
    }
    return hipGetLastError();
}

hipError_t launch_surface_points(const float* verts, float* points, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_surface_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, verts, points, n);
    return hipGetLastError();
}

hipError_t launch_normals(const float* vmap, int64_t vpitch, float* nmap, int64_t npitch, int W, int H,
                          hipStream_t s) {
    dim3 grid((W + 15) / 16, (H + 15) / 16);
    hipLaunchKernelGGL(k_normals, grid, dim3(256), 0, s, vmap, vpitch, nmap, npitch, W, H);
    return hipGetLastError();
}

hipError_t launch_shading(const float* vmap, int64_t vpitch, const float* nmap, int64_t npitch, uint8_t* rgba,
                          int64_t rpitch, int W, int H, hipStream_t s) {
    dim3 grid((W + 15) / 16, (H + 15) / 16);
    hipLaunchKernelGGL(k_shading, grid, dim3(256), 0, s, vmap, vpitch, nmap, npitch, rgba, rpitch, W, H);
    return hipGetLastError();
}

}  // namespace vmamd
