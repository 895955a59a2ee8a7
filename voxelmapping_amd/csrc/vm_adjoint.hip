// Adjoint of the raycaster: d(vertex map, normal map) / d(TSDF voxel values,
// camera pose twist), for the float voxel format. No counterpart in the
// reference; it differentiates its march (ray_tracing.cu:323-400) and its
// normals (image_proc.cu:74-93) with the discrete decisions (cells, the
// interpolate-or-not test, the break step) held fixed.
//
// Forward march, per pixel, in voxel units: r_0 = bs, r_{k+1} = r_k + s_k dir
// with s_k = max(rtStep * sdf_k, 1) when the sample is valid, else 1; the
// break step k* is the first with sdf <= 0; then r_a = r_k* + rtStep sdf_k* dir,
// r_b = r_a + rtStep interp(r_a) dir, and the vertex is v = T^-1 (r_b / vsInv).
// A step is "active" when its length depends on voxel values (valid and
// rtStep * sdf > 1). Only active steps (and the break) carry gradient; the
// others translate by exactly dir.
//
// Backward, per pixel: the march is recomputed with the same code path as
// the forward kernel, active steps are pushed to a per-pixel stack in HBM
// (position, type, the count of unit steps after it, and the sample's value
// and spatial gradient, so that the sweep reads no voxels), and the reverse
// sweep applies lambda_k = lambda_{k+1} + (lambda_{k+1}.dir) rtStep grad s(r_k),
// scattering (lambda.dir) rtStep w_j into dTSDF[j] with float atomics, and
// accumulating d/d(dir), d/d(start) and the output transform into the pose
// twist of camToWorld = T0 exp(xi), xi = (rho, phi). Memory is O(pixels).
#include "vm_common.h"
#include "vm_dev_ray.h"
#include "vm_launch.h"

namespace vmamd {

namespace {

constexpr int kStackDepth = 48;

__device__ inline float3 add3(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline float3 mul3(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ inline float dot3f(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float3 crs3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// R * v and R^T * v of an Affine's rotation
__device__ inline float3 rot(const Affine& T, float3 v) {
    return make_float3(T.r00 * v.x + T.r01 * v.y + T.r02 * v.z, T.r10 * v.x + T.r11 * v.y + T.r12 * v.z,
                       T.r20 * v.x + T.r21 * v.y + T.r22 * v.z);
}
__device__ inline float3 rotT(const Affine& T, float3 v) {
    return make_float3(T.r00 * v.x + T.r10 * v.y + T.r20 * v.z, T.r01 * v.x + T.r11 * v.y + T.r21 * v.z,
                       T.r02 * v.x + T.r12 * v.y + T.r22 * v.z);
}

// dTSDF contributions of one wave, summed per voxel in LDS (neighbouring rays
// touch the same voxels) and added to global memory once per voxel by
// flush(). An open-addressed table of kAccSlots voxel indices; a contribution
// that finds no slot within kAccProbes goes to global memory at once.
constexpr int kAccSlots = 512;
constexpr int kAccProbes = 8;
struct GradAcc {
    int* keys;     // LDS, -1 = free
    float* vals;   // LDS
    float* g;      // global dTSDF
    __device__ inline void add(int64_t idx, float v) {
        const int key = (int)idx;  // voxel slots of the float format are < 2^31
        const uint32_t h = ((uint32_t)key * 2654435761u) >> 23;  // 9 bits
        for (int k = 0; k < kAccProbes; ++k) {
            const int slot = (int)((h + k) & (kAccSlots - 1));
            const int prev = atomicCAS(keys + slot, -1, key);
            if (prev == -1 || prev == key) {
                atomicAdd(vals + slot, v);
                return;
            }
        }
        atomicAdd(g + idx, v);
    }
};

// one reverse step: position q carried a sample s(q) used as rtStep * s(q) * dir
// with multiplier lam; returns the added position adjoint, scatters dTSDF
__device__ inline float3 reverse_sample(RayReader<kFmtF32>& rd, float3 q, bool interp, float3 lam, float3 dir,
                                        float rtStep, GradAcc& acc) {
    const float a = dot3f(lam, dir) * rtStep;
    if (interp) {
        Trilinear t;
        trilinear(rd, q, t);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (t.ok[k] && t.w[k] != 0.f) acc.add(t.idx[k], a * t.w[k]);
        return mul3(t.grad, a);
    }
    const int64_t i = rd.voxel(floor_i(q.x), floor_i(q.y), floor_i(q.z));
    if (i >= 0) acc.add(i, a);
    return make_float3(0.f, 0.f, 0.f);
}

// Adjoint of computeNormalKernel gathered per vertex, for a T x T tile of
// vertices at (X0, Y0): g_v[q] = g_vertex[q] + the contributions of every
// normal pixel whose stencil reads q. Normals computed from a missed vertex
// (x = NaN) pass no gradient. The tile's vertices are reached by the normal
// pixels of the tile grown by one; each of those (T + 2)^2 normal pixels'
// two stencil-difference gradients (ga for v(x+1) - v(x-1), gb for
// v(y+1) - v(y-1)) is computed once into LDS by normals_adjoint_halo, and
// each vertex then gathers what its neighbours' stencils give it
// (vertex_grad).
template <int T>
struct NormalsAdjointShared {
    float3 ga[(T + 2) * (T + 2)], gb[(T + 2) * (T + 2)];
    unsigned char on[(T + 2) * (T + 2)];
};

__device__ inline float3 vmap_at(const float* vmap, int64_t vpitch, int yy, int xx) {
    const float4 v = *((const float4*)((const uint8_t*)vmap + (int64_t)yy * vpitch) + xx);
    return make_float3(v.x, v.y, v.z);
}

// the halo's stencil gradients, by `nthreads` threads (index tid); the caller
// synchronizes before vertex_grad
template <int T>
__device__ inline void normals_adjoint_halo(const float* vmap, int64_t vpitch, const float* gn, int64_t gnpitch, int W,
                                            int H, int X0, int Y0, int tid, int nthreads,
                                            NormalsAdjointShared<T>& sh) {
    constexpr int NH = T + 2;
    for (int t = tid; t < NH * NH; t += nthreads) {
        const int px = X0 - 1 + t % NH, py = Y0 - 1 + t / NH;
        bool on = false;
        if (gn && px >= 0 && py >= 0 && px < W - 1 && py < H - 1) {  // the pixels computeNormal writes
            const float4 g4 = *((const float4*)((const uint8_t*)gn + (int64_t)py * gnpitch) + px);
            const float3 g = make_float3(g4.x, g4.y, g4.z);
            if (g.x != 0.f || g.y != 0.f || g.z != 0.f) {
                const int x10 = max(px - 1, 0), x01 = min(px + 1, W), y10 = max(py - 1, 0), y01 = min(py + 1, H);
                const float3 v00 = vmap_at(vmap, vpitch, py, x10), v01 = vmap_at(vmap, vpitch, py, x01);
                const float3 v10 = vmap_at(vmap, vpitch, y10, px), v11 = vmap_at(vmap, vpitch, y01, px);
                if (!(isnan(v00.x) || isnan(v01.x) || isnan(v10.x) || isnan(v11.x))) {
                    const float3 a = sub3(v01, v00), b = sub3(v11, v10);
                    const float3 cr = crs3(a, b);
                    const float n2 = sum3(cr.x * cr.x, cr.y * cr.y, cr.z * cr.z);
                    float3 gc = g;
                    if (n2 > 0.f) {
                        const float nn = sqrtf(n2);
                        const float3 nrm = mul3(cr, 1.0f / nn);
                        gc = mul3(sub3(g, mul3(nrm, dot3f(nrm, g))), 1.0f / nn);
                    }
                    sh.ga[t] = crs3(b, gc);
                    sh.gb[t] = crs3(gc, a);
                    on = true;
                }
            }
        }
        sh.on[t] = on;
    }
}

// total vertex gradient of vertex (x, y) of the tile
template <int T>
__device__ inline float3 vertex_grad(const float* gvin, int64_t gvpitch, int W, int H, int X0, int Y0, int x, int y,
                                     const NormalsAdjointShared<T>& sh) {
    constexpr int NH = T + 2;
    float3 acc = make_float3(0.f, 0.f, 0.f);
    if (gvin) {
        const float4 g = *((const float4*)((const uint8_t*)gvin + (int64_t)y * gvpitch) + x);
        acc = make_float3(g.x, g.y, g.z);
    }
    // normal pixels (px, py) whose stencil can read (x, y)
    const int cand[5][2] = {{x, y}, {x - 1, y}, {x + 1, y}, {x, y - 1}, {x, y + 1}};
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const int px = cand[c][0], py = cand[c][1];
        const int t = (py - Y0 + 1) * NH + (px - X0 + 1);
        if (!sh.on[t]) continue;  // off the written pixels, no upstream gradient, or a missed vertex
        const int x10 = max(px - 1, 0), x01 = min(px + 1, W), y10 = max(py - 1, 0), y01 = min(py + 1, H);
        const float3 ga = sh.ga[t], gb = sh.gb[t];
        if (py == y && x01 == x) acc = add3(acc, ga);
        if (py == y && x10 == x) acc = sub3(acc, ga);
        if (px == x && y01 == y) acc = add3(acc, gb);
        if (px == x && y10 == y) acc = sub3(acc, gb);
    }
    return acc;
}

// An active step's sample: position, value and (trilinear samples) spatial
// gradient. An active sample is valid, so a trilinear one read 8 observed
// corners: its weights and voxel indices are all the sweep still needs.
struct StackRec {
    float x, y, z;
    int meta;  // bit 0: interp sample; bits 1..: unit steps after this step
    float value;
    float gx, gy, gz;
};

// One wave per 8x8-pixel tile, tiles in the raycast's XCD-banded order; each
// wave writes its 6 pose parts.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_raycast_backward(MapView m, const float* vmap, int64_t vpitch,
                                                         const float* gVertex, int64_t gvpitch, const float* gNormal,
                                                         int64_t gnpitch, int W, int H,
                                                         Intrinsics K, Affine pose, Affine Tinv, float vsInv,
                                                         float rtStep, float* gTsdf, float* poseParts,
                                                         StackRec* stack, int* stats, int tilesX, int tilesPerXcd,
                                                         int nTiles, RayBounds rb, float margin) {
    __shared__ int s_keys[kAccSlots];
    __shared__ float s_vals[kAccSlots];
    const int lane = threadIdx.x;
    for (int i = lane; i < kAccSlots; i += 64) {
        s_keys[i] = -1;
        s_vals[i] = 0.f;
    }
    __syncthreads();
    GradAcc acc{s_keys, s_vals, gTsdf};
    const int g = blockIdx.x;
    const int tile = (g & 7) * tilesPerXcd + (g >> 3);
    const int X0 = (tile % tilesX) * 8, Y0 = (tile / tilesX) * 8;
    const int x = X0 + (lane & 7);
    const int y = Y0 + (lane >> 3);
    // the vertex gradients of the tile: the upstream vertex gradient plus the
    // normals' adjoint (computeNormal's stencil over the tile grown by one)
    __shared__ NormalsAdjointShared<8> nsh;
    if (tile < nTiles) normals_adjoint_halo<8>(vmap, vpitch, gNormal, gnpitch, W, H, X0, Y0, lane, 64, nsh);
    __syncthreads();
    float part[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    bool overflow = false, hit = false;
    if (tile < nTiles && x < W && y < H) {
        const float3 gvec = vertex_grad<8>(gVertex, gvpitch, W, H, X0, Y0, x, y, nsh);
        if (gvec.x != 0.f || gvec.y != 0.f || gvec.z != 0.f) {
            const int64_t P = (int64_t)W * H, pix = (int64_t)y * W + x;
            // ---- forward recompute (same decisions as raycast_pixel) ----
            const float zNear = 0.1f, zFar = 5.0f;
            const float ux = K.ifx * ((float)x - K.cx), uy = K.ify * ((float)y - K.cy);
            const float3 pn = make_float3(ux * zNear, uy * zNear, zNear);
            float dist_s = norm3(pn) * vsInv;
            float3 bs = xform(pose, pn);
            bs = make_float3(bs.x * vsInv, bs.y * vsInv, bs.z * vsInv);
            const float3 pf = make_float3(ux * zFar, uy * zFar, zFar);
            // the ray-distance bounds of the tile end the march early as the
            // forward raycast does (raycast_pixel); a miss has no adjoint
            float nearLim, farLim;
            ray_bounds_limits(rb, bounds_tile(x, y, W), margin, nearLim, farLim);
            const float distEnd = fminf(norm3(pf) * vsInv, farLim);
            float3 be = xform(pose, pf);
            be = make_float3(be.x * vsInv, be.y * vsInv, be.z * vsInv);
            const float3 dvec = make_float3(be.x - bs.x, be.y - bs.y, be.z - bs.z);
            float3 dir = dvec;
            float dn = 0.f;
            {
                const float n2 = sum3(dir.x * dir.x, dir.y * dir.y, dir.z * dir.z);
                if (n2 > 0.f) {
                    dn = sqrtf(n2);
                    dir = make_float3(dir.x / dn, dir.y / dn, dir.z / dn);
                }
            }
            const float3 inv = make_float3(1.0f / fabsf(dir.x), 1.0f / fabsf(dir.y), 1.0f / fabsf(dir.z));
            float3 res = bs;
            RayReader<kFmtF32> rd(m);
            float sdf = 1.0f, last;
            bool valid = false, interpUsed = false;
            float step;
            int depth = 0, lead = 0, units = 0;  // units: unit steps since the last pushed record
            // the last pushed record stays in registers until the unit steps
            // after it are known (the next push, or the end of the march)
            StackRec pend{0.f, 0.f, 0.f, 0, 0.f, 0.f, 0.f, 0.f};
            {
                // unit steps below the near bound (ray_skip), counted as units
                const float q = fminf(nearLim, distEnd - 2.0f) - dist_s;
                int n = q >= 1.0f ? (int)q : 0;
                units += n;
                for (; n >= 8; n -= 8) {
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        res = make_float3(res.x + dir.x, res.y + dir.y, res.z + dir.z);
                        dist_s += 1.0f;
                    }
                }
                for (; n > 0; --n) {
                    res = make_float3(res.x + dir.x, res.y + dir.y, res.z + dir.z);
                    dist_s += 1.0f;
                }
            }
            // walks through empty blocks one box per pass, side by side with
            // the other lanes' samples (march, vm_dev_ray.h)
            Walk wk{0, 0, 0, 0, false};
            while (true) {
                if (wk.on) wk.on = walk_box(rd, res, dir, inv, dist_s, distEnd, wk.bx, wk.by, wk.bz, units);
                if (wk.on) continue;
                if (!(dist_s < distEnd)) break;
                last = sdf;
                const int vx = floor_i(res.x), vy = floor_i(res.y), vz = floor_i(res.z);
                const int bx = vox_to_blk(vx), by = vox_to_blk(vy), bz = vox_to_blk(vz);
                const int p = rd.block(bx, by, bz);
                if (p < 0 && exact_range(bx) && exact_range(by) && exact_range(bz)) {
                    wk = Walk{bx, by, bz, 0, true};
                    sdf = 1.0f;
                    valid = false;
                    continue;
                }
                // the forward's sample (RayReader::sample), with its
                // trilinear read taking the sample as its first corner
                sdf = rd.template sample<true>(res, vx, vy, vz, p, valid);
                interpUsed = rd.lastInterp;
                if (sdf <= 0.0f) break;
                if (sdf >= 0.f && last < 0.f) break;  // the reference returns here: no hit
                step = valid ? fmaxf(sdf * rtStep, 1.0f) : 1.0f;
                const bool active = valid && sdf * rtStep > 1.0f;
                if (active) {
                    if (depth > 0) {
                        pend.meta |= units << 1;
                        stack[(int64_t)(depth - 1) * P + pix] = pend;
                    } else {
                        lead = units;
                    }
                    if (depth < kStackDepth) {
                        const float3 g = rd.lastGrad;
                        pend = StackRec{res.x, res.y, res.z, interpUsed ? 1 : 0, sdf, g.x, g.y, g.z};
                        ++depth;
                    } else {
                        overflow = true;
                    }
                    units = 0;
                } else {
                    ++units;
                }
                res = make_float3(res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z);
                dist_s += step;
                // the unobserved samples after an unobserved one in the same
                // block, read together: unit steps, sdf stays 1
                if (!valid && p >= 0) units += run_samples(rd, res, dist_s, dir, distEnd, p, bx, by, bz, 1.0f, false);
            }
            // the last record stays in registers (pend): the reverse sweep
            // starts from it
            if (sdf <= 0.0f && !overflow) {
                const float3 rk = res;
                const bool kInterp = interpUsed;
                const float sk = sdf;
                step = sdf * rtStep;
                const float3 ra = make_float3(res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z);
                Trilinear ta;
                trilinear(rd, ra, ta);
                const float sb = ta.value;
                step = sb * rtStep;
                const float3 rb = make_float3(ra.x + step * dir.x, ra.y + step * dir.y, ra.z + step * dir.z);
                if (ta.valid) {
                    hit = true;
                    const float3 w = make_float3(rb.x / vsInv, rb.y / vsInv, rb.z / vsInv);
                    const float3 u = xform(Tinv, w);  // camera-frame vertex
                    // output transform v = R^T (w - t): d/dw = R g; twist part
                    float3 lam = mul3(rot(pose, gvec), 1.0f / vsInv);
                    float3 gRho = mul3(gvec, -1.f), gPhi = crs3(gvec, u);
                    float3 gDir = make_float3(0.f, 0.f, 0.f);
                    // r_b = r_a + rtStep s_b(r_a) dir
                    {
                        const float a = dot3f(lam, dir) * rtStep;
#pragma unroll
                        for (int k = 0; k < 8; ++k)
                            if (ta.ok[k] && ta.w[k] != 0.f) acc.add(ta.idx[k], a * ta.w[k]);
                        gDir = add3(gDir, mul3(lam, rtStep * sb));
                        lam = add3(lam, mul3(ta.grad, a));
                    }
                    // r_a = r_k* + rtStep s_k* dir
                    gDir = add3(gDir, mul3(lam, rtStep * sk));
                    lam = add3(lam, reverse_sample(rd, rk, kInterp, lam, dir, rtStep, acc));
                    // unit steps before the break, then the active steps in reverse
                    gDir = add3(gDir, mul3(lam, (float)units));
                    // each record's load is issued one record ahead, so that it
                    // is not a round trip of its own in the sweep
                    StackRec next = pend;
                    for (int d = depth - 1; d >= 0; --d) {
                        const StackRec rec = next;
                        if (d > 0) next = stack[(int64_t)(d - 1) * P + pix];
                        const float3 q = make_float3(rec.x, rec.y, rec.z);
                        // the record's value for d/d dir, its corner weights
                        // and gradient for dTSDF and lambda
                        const float a = dot3f(lam, dir) * rtStep;
                        gDir = add3(gDir, mul3(lam, rec.value * rtStep + (float)(rec.meta >> 1)));
                        if (rec.meta & 1) {
                            int64_t idx[8];
                            float w[8];
                            trilinear_corners(rd, q, idx);
                            trilinear_weights(q.x - floorf(q.x), q.y - floorf(q.y), q.z - floorf(q.z), w);
#pragma unroll
                            for (int k = 0; k < 8; ++k)
                                if (idx[k] >= 0 && w[k] != 0.f) acc.add(idx[k], a * w[k]);
                            lam = add3(lam, mul3(make_float3(rec.gx, rec.gy, rec.gz), a));
                        } else {
                            const int64_t i = rd.voxel(floor_i(q.x), floor_i(q.y), floor_i(q.z));
                            if (i >= 0) acc.add(i, a);
                        }
                    }
                    gDir = add3(gDir, mul3(lam, (float)(depth > 0 ? lead : 0)));
                    // dir = dvec / |dvec|; bs gets lam - d/d dvec, be gets d/d dvec
                    float3 gD = make_float3(0.f, 0.f, 0.f);
                    if (dn > 0.f) gD = mul3(sub3(gDir, mul3(dir, dot3f(dir, gDir))), 1.0f / dn);
                    const float3 gBs = sub3(lam, gD), gBe = gD;
                    // bs = vsInv (R p_n + t), be = vsInv (R p_f + t), T = T0 exp(xi)
                    const float3 hBs = mul3(rotT(pose, gBs), vsInv), hBe = mul3(rotT(pose, gBe), vsInv);
                    gRho = add3(gRho, add3(hBs, hBe));
                    gPhi = add3(gPhi, add3(crs3(pn, hBs), crs3(pf, hBe)));
                    part[0] = gRho.x; part[1] = gRho.y; part[2] = gRho.z;
                    part[3] = gPhi.x; part[4] = gPhi.y; part[5] = gPhi.z;
                }
            }
        }
    }
    // the wave's dTSDF sums, one global atomic per voxel
    __syncthreads();
    for (int i = lane; i < kAccSlots; i += 64)
        if (s_keys[i] >= 0) atomicAdd(gTsdf + s_keys[i], s_vals[i]);
    // wave reduction of the pose parts into this wave's slot
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        float v = part[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) poseParts[(int64_t)blockIdx.x * 6 + k] = v;
    }
    const unsigned long long hits = __ballot(hit), ovf = __ballot(overflow);
    if (lane == 0 && stats) {  // counted only when the caller asks
        if (hits) atomicAdd(&stats[0], __popcll(hits));
        if (ovf) atomicAdd(&stats[1], __popcll(ovf));
    }
}

// the 6 pose components summed over the waves' parts in a fixed order (each
// thread a strided slice in double, then a tree): deterministic
constexpr int kReduceThreads = 1024;
__global__ __launch_bounds__(kReduceThreads) void k_pose_reduce(const float* parts, int n, double* gPose) {
    __shared__ double s[kReduceThreads];
    const int k = blockIdx.x;  // component
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kReduceThreads) acc += (double)parts[(int64_t)i * 6 + k];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = kReduceThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) gPose[k] += s[0];
}

}  // namespace

// scratch layout: [stack | pose parts (256-byte aligned)]
// one launch slot per 8x8 tile, rounded up to whole XCD bands
static size_t backward_waves(int W, int H) {
    const size_t tiles = (size_t)((W + 7) / 8) * ((H + 7) / 8);
    return (tiles + 7) / 8 * 8;
}
static size_t parts_bytes(int W, int H) { return ((sizeof(float) * 6 * backward_waves(W, H) + 255) / 256) * 256; }

size_t raycast_backward_scratch_bytes(int W, int H) {
    return sizeof(StackRec) * kStackDepth * (size_t)W * H + parts_bytes(W, H);
}

hipError_t launch_raycast_backward(const MapView& m, const float* vmap, int64_t vpitch, const float* gVertex,
                                   int64_t gvpitch, const float* gNormal, int64_t gnpitch, int W, int H,
                                   const Intrinsics& K, const Affine& pose, const Affine& Tinv, float* gTsdf,
                                   double* gPose, void* scratch, int* stats, const RayBounds& rb,
                                   bool launchBounds, hipStream_t s) {
    const int tilesX = (W + 7) / 8, nTiles = tilesX * ((H + 7) / 8);
    const int waves = (int)backward_waves(W, H);
    uint8_t* p = (uint8_t*)scratch;
    StackRec* stack = (StackRec*)p;
    p += sizeof(StackRec) * kStackDepth * (size_t)W * H;
    float* parts = (float*)p;
    const float vsInv = (float)(1.0 / (double)m.voxelSize);
    const float rtStep = m.trunc / m.voxelSize;
    const float margin = ray_drift_margin(W, H, K, pose.t0, pose.t1, pose.t2, vsInv);
    if (launchBounds) {
        hipError_t e = launch_ray_bounds(m, rb, W, H, K, pose, Tinv, margin, s);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_raycast_backward, dim3((unsigned)waves), dim3(64), 0, s, m, vmap, vpitch, gVertex, gvpitch,
                       gNormal, gnpitch, W, H, K, pose, Tinv, vsInv, rtStep, gTsdf, parts, stack, stats, tilesX,
                       waves / 8, nTiles, rb, margin);
    hipLaunchKernelGGL(k_pose_reduce, dim3(6), dim3(kReduceThreads), 0, s, parts, waves, gPose);
    return hipGetLastError();
}

}  // namespace vmamd
