// Device functions of the raycaster (MapRenderingDelegate,
// ray_tracing.cu:255-401), shared by the single-map and multi-scene kernels.
//
// Every ray runs the reference's march step for step: the same sample
// positions, the same reads, the same float operations in the same order.
// Three things make it cheaper without changing any value:
//   * each lane keeps the last two voxel blocks it resolved, so a sample that
//     stays inside a block costs no hash probe;
//   * a block whose occupancy-brick bit is clear is absent without a probe,
//     and the lane keeps the brick word it read last;
//   * inside a block that is not allocated every sample reads "invalid, 1.0"
//     and steps exactly one voxel (res += 1.0f * dir == res + dir), so those
//     iterations run as a tight loop of additions with a float-bounds test for
//     leaving the block (or the brick octant or whole 4x4x4-block brick
//     when none of their blocks is present), and no reads at all.
#pragma once

#include "vm_common.h"

namespace vmamd {

// The excess chain behind a bucket head that did not hold the key: the tail of
// findEntry, map_struct.h:356-374 (read-only walk, bounded); ptr or -1.
// (The key tests use & rather than && so that the five fields of an entry are
// loaded together instead of the position only after ptr has arrived.)
__device__ inline int chain_walk(const MapView& m, int off, int bx, int by, int bz) {
    int guard = m.nEntry - m.nBucket + 1;
    while (off >= 0 && --guard >= 0) {
        const HashEntry* e = m.table + (m.nBucket + off - 1);
        const int p = e->ptr, x = e->x, y = e->y, z = e->z;
        if ((p >= 0) & (x == bx) & (y == by) & (z == bz)) return p;
        off = e->offset;
    }
    return -1;
}

// findEntry, map_struct.h:356-374; returns ptr or -1
__device__ inline int find_block(const MapView& m, int bx, int by, int bz) {
    const HashEntry* e = m.table + hash_block(bx, by, bz, m.nBucket, m.bucketMask);
    const int p = e->ptr, off = e->offset, x = e->x, y = e->y, z = e->z;
    if ((p >= 0) & (x == bx) & (y == by) & (z == bz)) return p;
    return chain_walk(m, off, bx, by, bz);
}

// Block lookups of one ray. Each lane keeps the last two blocks it resolved
// and the occupancy-brick word of the last brick it touched: a block whose
// bit is clear is absent without a hash probe, so a march through empty space
// reads one 8-byte word per 4x4x4 blocks (shared by the neighbouring rays of
// the wave) instead of probing a hash bucket per block.
template <int FMT>
struct RayReader {
    const MapView& m;
    int cx, cy, cz, cp;  // most recent block
    int px, py, pz, pp;  // the one before
    int qx, qy, qz;      // brick (block >> 2) of `word`; INT_MIN: none yet
    unsigned long long word;
    bool lastInterp;     // the last sample() took a trilinear read
    float3 lastGrad;     // its spatial gradient (sample<true> only)

    __device__ explicit RayReader(const MapView& mv)
        : m(mv), cx(INT_MIN), cy(INT_MIN), cz(INT_MIN), cp(-1), px(INT_MIN), py(INT_MIN), pz(INT_MIN), pp(-1),
          qx(INT_MIN), qy(INT_MIN), qz(INT_MIN), word(0), lastInterp(false),
          lastGrad(make_float3(0.f, 0.f, 0.f)) {}

    // false: no hash entry holds block (bx, by, bz)
    __device__ inline bool maybe_present(int bx, int by, int bz) {
        const int x = bx >> 2, y = by >> 2, z = bz >> 2;
        if (x != qx || y != qy || z != qz) {
            word = m.occ[brick_slot(bx, by, bz, m.occMask)];
            qx = x; qy = y; qz = z;
        }
        return (word & brick_bit(bx, by, bz)) != 0;
    }

    // true: block (bx, by, bz) is in the brick held and its bit is clear
    __device__ inline bool known_absent(int bx, int by, int bz) const {
        return (bx >> 2) == qx && (by >> 2) == qy && (bz >> 2) == qz && (word & brick_bit(bx, by, bz)) == 0;
    }

    // true: the brick word held is that of block (bx, by, bz)
    __device__ inline bool holds(int bx, int by, int bz) const {
        return (bx >> 2) == qx && (by >> 2) == qy && (bz >> 2) == qz;
    }

    __device__ inline int block(int bx, int by, int bz) {
        if (bx == cx && by == cy && bz == cz) return cp;
        if (bx == px && by == py && bz == pz) {
            int t;
            t = cx; cx = px; px = t;
            t = cy; cy = py; py = t;
            t = cz; cz = pz; pz = t;
            t = cp; cp = pp; pp = t;
            return cp;
        }
        const int p = maybe_present(bx, by, bz) ? find_block(m, bx, by, bz) : -1;
        px = cx; py = cy; pz = cz; pp = cp;
        cx = bx; cy = by; cz = bz; cp = p;
        return p;
    }

    // voxel index of integer voxel (vx, vy, vz), -1 when its block is absent
    __device__ inline int64_t voxel(int vx, int vy, int vz) {
        const int p = block(vox_to_blk(vx), vox_to_blk(vy), vox_to_blk(vz));
        if (p < 0) return -1;
        return (int64_t)p + local_index(vox_to_loc(vx), vox_to_loc(vy), vox_to_loc(vz));
    }


    // read_sdf_interped, ray_tracing.cu:288-321, at q with floor(q) =
    // (X0, Y0, Z0) in block (bx, by, bz) at pool offset p (-1: absent).
    // floor(q + e_k) is computed per axis exactly as the reference adds the
    // unit offset before flooring. Corner k's block differs from (bx, by, bz)
    // only along the axes where floor(q + 1) crosses into the next block, so
    // only those neighbours are resolved (one lookup for the usual single
    // crossing, none for the two thirds of samples whose corners all lie in
    // the block) instead of one cache lookup per corner. With HAVE0 the value
    // of corner 0 (the voxel floor(q) itself) is the caller's s0 / ok0, the
    // march's own sample of it.
    // pb[c], c = i | j << 1 | l << 2: the pool offset of the block of the
    // corners whose coordinates are X_i, Y_j, Z_l (i, j, l in {0, 1}), given
    // the block (bx, by, bz) of (X0, Y0, Z0) at offset p and the corners
    // X1, Y1, Z1. A combination is looked up only when every axis it sets
    // crosses into the next block; otherwise it is the block of c with the
    // bits of the non-crossing axes cleared.
    __device__ __forceinline__ void corner_blocks(int X1, int Y1, int Z1, int bx, int by, int bz, int p, int pb[8]) {
        const int bx1 = vox_to_blk(X1), by1 = vox_to_blk(Y1), bz1 = vox_to_blk(Z1);
        const bool sx = bx1 != bx, sy = by1 != by, sz = bz1 != bz;
        pb[0] = p;
        pb[1] = sx ? block(bx1, by, bz) : p;
        pb[2] = sy ? block(bx, by1, bz) : p;
        pb[3] = sx && sy ? block(bx1, by1, bz) : sx ? pb[1] : pb[2];
        pb[4] = sz ? block(bx, by, bz1) : p;
        pb[5] = sx && sz ? block(bx1, by, bz1) : sx ? pb[1] : pb[4];
        pb[6] = sy && sz ? block(bx, by1, bz1) : sy ? pb[2] : pb[4];
        pb[7] = sx && sy && sz ? block(bx1, by1, bz1)
                : sx && sy   ? pb[3]
                : sx && sz   ? pb[5]
                : sy && sz   ? pb[6]
                : sx         ? pb[1]
                : sy         ? pb[2]
                : pb[4];
    }

    template <bool HAVE0, bool GRAD = false>
    __device__ __forceinline__ float interp_in(float3 q, int X0, int Y0, int Z0, int bx, int by, int bz, int p,
                                               float s0, bool ok0, bool& valid) {
        const float fx = q.x - floorf(q.x), fy = q.y - floorf(q.y), fz = q.z - floorf(q.z);
        const int X1 = floor_i(q.x + 1.f), Y1 = floor_i(q.y + 1.f), Z1 = floor_i(q.z + 1.f);
        int pb[8];
        corner_blocks(X1, Y1, Z1, bx, by, bz, p, pb);
        const int lx0 = vox_to_loc(X0), lx1 = vox_to_loc(X1);
        const int ly0 = vox_to_loc(Y0), ly1 = vox_to_loc(Y1);
        const int lz0 = vox_to_loc(Z0), lz1 = vox_to_loc(Z1);
        float s[8];
        bool all = true;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            float v = 1.0f;
            bool ok = false;
            if (HAVE0 && k == 0) {
                v = s0;
                ok = ok0;
            } else {
                const int b = pb[k];
                if (b >= 0)
                    Vox<FMT>::read_local(m.voxels, b,
                                         local_index((k & 1) ? lx1 : lx0, (k & 2) ? ly1 : ly0, (k & 4) ? lz1 : lz0),
                                         v, ok);
            }
            s[k] = ok ? v : 1.0f;
            all = all && ok;
        }
        valid = all;
        if (GRAD) lastGrad = blend_grad(fx, fy, fz, s);
        return blend(fx, fy, fz, s);
    }

    // read_sdf_interped at an arbitrary q
    __device__ inline float interp(float3 q, bool& valid) {
        const int X0 = floor_i(q.x), Y0 = floor_i(q.y), Z0 = floor_i(q.z);
        const int bx = vox_to_blk(X0), by = vox_to_blk(Y0), bz = vox_to_blk(Z0);
        return interp_in<false>(q, X0, Y0, Z0, bx, by, bz, block(bx, by, bz), 1.0f, false, valid);
    }

    // the trilinear blend of read_sdf_interped, ray_tracing.cu:305-320
    __device__ static inline float blend(float fx, float fy, float fz, const float s[8]) {
        float r0 = (1.0f - fx) * s[0] + fx * s[1];
        float r1 = (1.0f - fx) * s[2] + fx * s[3];
        const float r2 = (1.0f - fy) * r0 + fy * r1;
        r0 = (1.0f - fx) * s[4] + fx * s[5];
        r1 = (1.0f - fx) * s[6] + fx * s[7];
        const float r3 = (1.0f - fy) * r0 + fy * r1;
        return (1.0f - fz) * r2 + fz * r3;
    }

    // d blend / d q: the gradient of the same trilinear form (as trilinear()
    // computes it)
    __device__ static inline float3 blend_grad(float fx, float fy, float fz, const float s[8]) {
        const float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
        const float r0 = gx * s[0] + fx * s[1], r1 = gx * s[2] + fx * s[3];
        const float q0 = gx * s[4] + fx * s[5], q1 = gx * s[6] + fx * s[7];
        return make_float3(gz * (gy * (s[1] - s[0]) + fy * (s[3] - s[2])) + fz * (gy * (s[5] - s[4]) + fy * (s[7] - s[6])),
                           gz * (r1 - r0) + fz * (q1 - q0), (gy * q0 + fy * q1) - (gy * r0 + fy * r1));
    }

    // The march's sample at q, floor(q) = (vx, vy, vz) in block p (ptr or
    // -1): read_sdf of floor(q) (ray_tracing.cu:272-286: the voxel's value
    // when its weight is nonzero, else (1.0, invalid)), and
    // read_sdf_interped of q when that value is within [-0.5, 0.5]
    // (ray_tracing.cu:358-362).
    // GRAD: a trilinear read also leaves its spatial gradient in lastGrad.
    template <bool GRAD = false>
    __device__ inline float sample(float3 q, int vx, int vy, int vz, int p, bool& valid) {
        const int lx = vox_to_loc(vx), ly = vox_to_loc(vy), lz = vox_to_loc(vz);
        float sdf = 1.0f;
        bool v = false;
        if (p >= 0) Vox<FMT>::read_local(m.voxels, p, local_index(lx, ly, lz), sdf, v);
        valid = v;
        if (!v) sdf = 1.0f;
        lastInterp = sdf <= 0.5f && sdf >= -0.5f;
        if (lastInterp)
            sdf = interp_in<true, GRAD>(q, vx, vy, vz, vox_to_blk(vx), vox_to_blk(vy), vox_to_blk(vz), p, sdf,
                                        valid, valid);
        return sdf;
    }
};

__device__ inline float norm3(float3 v) { return sqrtf(sum3(v.x * v.x, v.y * v.y, v.z * v.z)); }

// |v| below 2^24: integer block bounds are exact floats there
__device__ inline bool exact_range(int b) { return b > -(1 << 20) && b < (1 << 20); }

// Trilinear sample at q with its voxel indices, weights and spatial
// gradient (d sdf / d q): the same corners and the same operations as
// RayReader::interp, so `value` and `valid` equal what interp returns.
struct Trilinear {
    int64_t idx[8];
    float w[8];
    bool ok[8];
    float value;
    float3 grad;
    bool valid;
};

// the voxel indices of the 8 corners of a trilinear read at q (-1: block
// absent), as trilinear() resolves them
template <int FMT>
__device__ inline void trilinear_corners(RayReader<FMT>& rd, float3 q, int64_t idx[8]) {
    const int X0 = floor_i(q.x), X1 = floor_i(q.x + 1.f);
    const int Y0 = floor_i(q.y), Y1 = floor_i(q.y + 1.f);
    const int Z0 = floor_i(q.z), Z1 = floor_i(q.z + 1.f);
    {
        const int bx = vox_to_blk(X0), by = vox_to_blk(Y0), bz = vox_to_blk(Z0);
        int pb[8];
        rd.corner_blocks(X1, Y1, Z1, bx, by, bz, rd.block(bx, by, bz), pb);
        const int lx0 = vox_to_loc(X0), lx1 = vox_to_loc(X1);
        const int ly0 = vox_to_loc(Y0), ly1 = vox_to_loc(Y1);
        const int lz0 = vox_to_loc(Z0), lz1 = vox_to_loc(Z1);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            idx[k] = pb[k] < 0 ? -1
                               : (int64_t)pb[k] + local_index((k & 1) ? lx1 : lx0, (k & 2) ? ly1 : ly0,
                                                              (k & 4) ? lz1 : lz0);
    }
}

// the trilinear weights of the 8 corners at fractional offsets (fx, fy, fz)
__device__ inline void trilinear_weights(float fx, float fy, float fz, float w[8]) {
    const float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
    w[0] = gx * gy * gz; w[1] = fx * gy * gz; w[2] = gx * fy * gz; w[3] = fx * fy * gz;
    w[4] = gx * gy * fz; w[5] = fx * gy * fz; w[6] = gx * fy * fz; w[7] = fx * fy * fz;
}

template <int FMT>
__device__ inline void trilinear(RayReader<FMT>& rd, float3 q, Trilinear& t) {
    const float fx = q.x - floorf(q.x), fy = q.y - floorf(q.y), fz = q.z - floorf(q.z);
    trilinear_corners(rd, q, t.idx);
    float s[8];
    t.valid = true;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float v = 1.0f;
        bool ok = false;
        if (t.idx[k] >= 0) Vox<FMT>::read(rd.m.voxels, t.idx[k], v, ok);
        s[k] = ok ? v : 1.0f;
        t.ok[k] = ok;
        t.valid = t.valid && ok;
    }
    const float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
    float r0 = gx * s[0] + fx * s[1];
    float r1 = gx * s[2] + fx * s[3];
    const float r2 = gy * r0 + fy * r1;
    const float q0 = gx * s[4] + fx * s[5];
    const float q1 = gx * s[6] + fx * s[7];
    const float r3 = gy * q0 + fy * q1;
    t.value = gz * r2 + fz * r3;
    trilinear_weights(fx, fy, fz, t.w);
    t.grad.x = gz * (gy * (s[1] - s[0]) + fy * (s[3] - s[2])) + fz * (gy * (s[5] - s[4]) + fy * (s[7] - s[6]));
    t.grad.y = gz * (r1 - r0) + fz * (q1 - q0);
    t.grad.z = r3 - r2;
}

// The march's samples in blocks that no hash entry holds (read_sdf gives
// (1.0, invalid), the step is one voxel). The position res lies in block
// (cbx, cby, cbz), known to be absent, whose bounds are exact floats. One box
// of the walk: the same additions as the reference's loop through the
// largest box known to be empty around that block, ending with the first
// sample outside it. Returns true while the walk can go on (the sample is
// below distEnd, in a block known to be absent: (cbx, cby, cbz) is then that
// block), false once the march must take the next sample itself. Adds the
// unit steps taken to `steps`. inv = 1 / |dir| per axis.
template <int FMT>
__device__ __forceinline__ bool walk_box(const RayReader<FMT>& rd, float3& res, const float3 dir, const float3 inv,
                                         float& dist_s, const float distEnd, int& cbx, int& cby, int& cbz,
                                         int& steps) {
    // The walk goes on across a face into the next block when that block's
    // bit in the brick word the lane holds is clear. The box known to be
    // empty is the largest of: the whole brick (32^3 voxels) when its word is
    // zero, the brick octant (2x2x2 blocks, bit mask 0x330033 shifted to the
    // octant) holding the current block when its eight bits are clear, the
    // current block.
    const bool held = rd.holds(cbx, cby, cbz);
    const bool brick = held && rd.word == 0ull;
    const bool oct = held && (rd.word & (0x330033ull << (((cbx >> 1) & 1) * 2 + ((cby >> 1) & 1) * 8 +
                                                        ((cbz >> 1) & 1) * 32))) == 0ull;
    const float edge = brick ? 32.f : oct ? 16.f : 8.f;
    const int am = brick ? ~3 : oct ? ~1 : ~0;  // block coordinates of the box's low corner
    const float lx = (float)((cbx & am) * 8), ly = (float)((cby & am) * 8), lz = (float)((cbz & am) * 8);
    const float hx = lx + edge, hy = ly + edge, hz = lz + edge;
    // Unit steps that stay strictly inside the box and below distEnd
    // whatever the rounding of the additions: k <= 64 additions to
    // coordinates below B = max(|l|, |h|) + 1 move them at most
    // 64 * 2^-24 * B from the exact k * dir, which the margins cover. They
    // are the same additions as the checked steps, without the checks.
    {
        const float ex = (fmaxf(fabsf(lx), fabsf(hx)) + 1.f) * 0x1p-18f + 0x1p-16f;
        const float ey = (fmaxf(fabsf(ly), fabsf(hy)) + 1.f) * 0x1p-18f + 0x1p-16f;
        const float ez = (fmaxf(fabsf(lz), fabsf(hz)) + 1.f) * 0x1p-18f + 0x1p-16f;
        const float kx = ((dir.x > 0.f ? hx - res.x : res.x - lx) - ex) * inv.x;
        const float ky = ((dir.y > 0.f ? hy - res.y : res.y - ly) - ey) * inv.y;
        const float kz = ((dir.z > 0.f ? hz - res.z : res.z - lz) - ez) * inv.z;
        // minimum that keeps a NaN (no unchecked steps then)
        float q = fminf((distEnd - dist_s) - 2.0f, 64.0f);
        q = kx >= q ? q : kx;
        q = ky >= q ? q : ky;
        q = kz >= q ? q : kz;
        int n = q >= 1.0f ? (int)q : 0;
        steps += n + 1;
        for (; n >= 4; n -= 4) {
            res = make_float3(res.x + dir.x, res.y + dir.y, res.z + dir.z);
            dist_s += 1.0f;
            res = make_float3(res.x + dir.x, res.y + dir.y, res.z + dir.z);
            dist_s += 1.0f;
            res = make_float3(res.x + dir.x, res.y + dir.y, res.z + dir.z);
            dist_s += 1.0f;
            res = make_float3(res.x + dir.x, res.y + dir.y, res.z + dir.z);
            dist_s += 1.0f;
        }
        for (; n > 0; --n) {
            res = make_float3(res.x + dir.x, res.y + dir.y, res.z + dir.z);
            dist_s += 1.0f;
        }
    }
    res = make_float3(res.x + dir.x, res.y + dir.y, res.z + dir.z);
    dist_s += 1.0f;
    if (!(dist_s < distEnd)) return false;
    if (res.x >= lx && res.x < hx && res.y >= ly && res.y < hy && res.z >= lz && res.z < hz) return true;
    cbx = vox_to_blk(floor_i(res.x));
    cby = vox_to_blk(floor_i(res.y));
    cbz = vox_to_blk(floor_i(res.z));
    return exact_range(cbx) && exact_range(cby) && exact_range(cbz) && rd.known_absent(cbx, cby, cbz);
}

// State of one ray of MapRenderingDelegate::operator() (ray_tracing.cu:330-401)
// between samples, so that a lane can run a ray in pieces.
struct Ray {
    float3 res, dir, inv;  // inv = 1 / |dir| per axis (+inf for a zero component)
    float dist_s, distEnd, sdf;
    bool valid;
};

// Ray setup for pixel (x, y). vsInv = (float)(1.0 / voxelSize)
// (ray_tracing.cu:430).
__device__ inline void ray_begin(Ray& r, int x, int y, const Intrinsics& K, const Affine& pose, float vsInv) {
    const float zNear = 0.1f, zFar = 5.0f;  // MIN_DEPTH / MAX_DEPTH, ray_tracing.cu:340-341
    const float ux = K.ifx * ((float)x - K.cx), uy = K.ify * ((float)y - K.cy);
    float3 pt = make_float3(ux * zNear, uy * zNear, zNear);
    r.dist_s = norm3(pt) * vsInv;
    float3 bs = xform(pose, pt);
    bs = make_float3(bs.x * vsInv, bs.y * vsInv, bs.z * vsInv);
    pt = make_float3(ux * zFar, uy * zFar, zFar);
    r.distEnd = norm3(pt) * vsInv;
    float3 be = xform(pose, pt);
    be = make_float3(be.x * vsInv, be.y * vsInv, be.z * vsInv);
    float3 dir = make_float3(be.x - bs.x, be.y - bs.y, be.z - bs.z);
    {
        // Eigen normalized(): divide by the norm when the squared norm is > 0
        const float n2 = sum3(dir.x * dir.x, dir.y * dir.y, dir.z * dir.z);
        if (n2 > 0.f) {
            const float n = sqrtf(n2);
            dir = make_float3(dir.x / n, dir.y / n, dir.z / n);
        }
    }
    r.dir = dir;
    r.inv = make_float3(1.0f / fabsf(dir.x), 1.0f / fabsf(dir.y), 1.0f / fabsf(dir.z));
    r.res = bs;
    r.sdf = 1.0f;
    r.valid = false;
}

enum RayStatus : int { kRayGoing = 0, kRaySurface = 1, kRayMiss = 2 };

// Runs of predictable samples (the forward march, and the raycast adjoint's
// recompute of it: unit steps only). A sample in an
// allocated block that reads "invalid" (an unobserved voxel, or an
// interpolation with an unobserved corner) steps exactly one voxel, and on
// the rays that take longest it is mostly followed by more of them: rays
// crossing the unobserved side of fused surfaces. Likewise a valid voxel of
// exactly 1.0 (free space in front of a surface) steps rtStep. The
// positions of such followers are known before their reads return, so the
// next kRunSpec sample positions inside the same block are computed with the
// loop's own additions, their voxels read together, and the prefix that
// reads the same kind is taken (each such sample only steps); the march
// resumes at the first sample that reads anything else, the block's end or
// distEnd. One memory round trip then covers up to kRunSpec samples that the
// loop reads one after the other. Positions, distances and results are the
// loop's own, bit for bit. Which runs ray_step takes is a template choice:
// kRunsNone, kRunsInvalid (unit steps), kRunsAll (also 1.0 steps).
enum RunKinds : int { kRunsNone = 0, kRunsInvalid = 1, kRunsAll = 2 };
constexpr int kRunSpec = 4;

// A march position res (distance dist) is just past a sample of voxel block
// (bx, by, bz) (pool offset p) that read "invalid" (valid = false, step 1)
// or a valid 1.0 (step rtStep), and took its step `step`. Up to kRunSpec
// further samples of the loop lie in the same block; their voxels are read
// together, and the leading ones that read the same (and so take the same
// step, with nothing else changing) are taken: res and dist advance by the
// same additions as the loop's. Returns the number of samples taken.
template <int FMT>
__device__ inline int run_samples(const RayReader<FMT>& rd, float3& res, float& dist, float3 dir, float distEnd,
                                  int p, int bx, int by, int bz, float step, bool valid) {
    int loc[kRunSpec];
    int n = 0;
    float3 q = res;
    float d = dist;
#pragma unroll
    for (int k = 0; k < kRunSpec; ++k) {
        loc[k] = 0;
        if (n == k && d < distEnd) {  // the loop condition of sample k
            const int qx = floor_i(q.x), qy = floor_i(q.y), qz = floor_i(q.z);
            if (vox_to_blk(qx) == bx && vox_to_blk(qy) == by && vox_to_blk(qz) == bz) {
                loc[k] = local_index(vox_to_loc(qx), vox_to_loc(qy), vox_to_loc(qz));
                q = make_float3(q.x + step * dir.x, q.y + step * dir.y, q.z + step * dir.z);
                d += step;
                n = k + 1;
            }
        }
    }
    float sv[kRunSpec];
    bool ok[kRunSpec];
#pragma unroll
    for (int k = 0; k < kRunSpec; ++k) {
        sv[k] = 1.0f;
        ok[k] = false;
        if (k < n) Vox<FMT>::read_local(rd.m.voxels, p, loc[k], sv[k], ok[k]);
    }
    int acc = 0;
#pragma unroll
    for (int k = 0; k < kRunSpec; ++k)
        if (acc == k && k < n && (valid ? (ok[k] && sv[k] == 1.0f) : !ok[k])) acc = k + 1;
    for (int k = 0; k < acc; ++k) {
        res = make_float3(res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z);
        dist += step;
    }
    return acc;
}

template <int FMT>
__device__ inline void ray_run(Ray& r, const RayReader<FMT>& rd, int p, int bx, int by, int bz, float step,
                               bool valid) {
    run_samples(rd, r.res, r.dist_s, r.dir, r.distEnd, p, bx, by, bz, step, valid);
}

// A walk through empty blocks that a lane has started and takes one box at
// a time (walk_box): the block the position is in, and whether it is on.
struct Walk {
    int bx, by, bz, steps;
    bool on;
};

// One pass of the march loop body (ray_tracing.cu:355-379): a sample, or a
// run of samples in empty blocks. kRaySurface: the loop broke on sdf <= 0;
// kRayMiss: it returned on a sign change or ran out of distance.
// rtStep = trunc / voxelSize (ray_tracing.cu:431). RUNS (RunKinds): the
// runs of samples taken after a sample that starts one (ray_run) -- for the
// forward raycast only, as a run's samples are not visited one by one.
// A walk through empty blocks is only started here (wk.on, its first block,
// the sample state it leaves: (1.0, invalid)); the caller takes its boxes
// (walk_box), one per pass of its loop.
template <int FMT, int RUNS = kRunsNone>
__device__ __forceinline__ int ray_step(Ray& r, RayReader<FMT>& rd, float rtStep, Walk& wk) {
    if (!(r.dist_s < r.distEnd)) return kRayMiss;  // loop condition; sdf > 0 or NaN here
    const float last = r.sdf;
    const int vx = floor_i(r.res.x), vy = floor_i(r.res.y), vz = floor_i(r.res.z);
    const int bx = vox_to_blk(vx), by = vox_to_blk(vy), bz = vox_to_blk(vz);
    const int p = rd.block(bx, by, bz);
    if (p < 0 && exact_range(bx) && exact_range(by) && exact_range(bz)) {
        wk.on = true;
        wk.bx = bx;
        wk.by = by;
        wk.bz = bz;
        r.sdf = 1.0f;
        r.valid = false;
        return kRayGoing;
    }
    bool valid;
    const float sdf = rd.sample(r.res, vx, vy, vz, p, valid);
    r.sdf = sdf;
    r.valid = valid;
    if (sdf <= 0.0f) return kRaySurface;
    if (sdf >= 0.f && last < 0.f) return kRayMiss;
    const float step = valid ? fmaxf(sdf * rtStep, 1.0f) : 1.0f;
    r.res = make_float3(r.res.x + step * r.dir.x, r.res.y + step * r.dir.y, r.res.z + step * r.dir.z);
    r.dist_s += step;
    if (RUNS != kRunsNone && p >= 0 && (!valid || (RUNS == kRunsAll && sdf == 1.0f)))
        ray_run(r, rd, p, bx, by, bz, step, valid);
    return kRayGoing;
}

// The two zero-crossing corrections and the output (ray_tracing.cu:381-399)
// after kRaySurface: the camera-frame vertex, or false when the last
// interpolation is invalid.
template <int FMT>
__device__ inline bool ray_surface(Ray& r, RayReader<FMT>& rd, const Affine& Tinv, float vsInv, float rtStep,
                                   float4& v) {
    float step = r.sdf * rtStep;
    float3 res = make_float3(r.res.x + step * r.dir.x, r.res.y + step * r.dir.y, r.res.z + step * r.dir.z);
    bool valid;
    const float sdf = rd.interp(res, valid);
    step = sdf * rtStep;
    res = make_float3(res.x + step * r.dir.x, res.y + step * r.dir.y, res.z + step * r.dir.z);
    if (!valid) return false;
    const float3 w = make_float3(res.x / vsInv, res.y / vsInv, res.z / vsInv);
    const float3 c = xform(Tinv, w);
    v = make_float4(c.x, c.y, c.z, 1.0f);
    return true;
}

// The loop's first samples of a ray while dist_s stays below nearLim, a
// distance below which no sample can read a voxel of an allocated block:
// each reads (1.0, invalid) and steps one voxel, so they run as the same
// additions without reads. Samples k = 0..n-1 sit at dist_s + k < nearLim
// and below distEnd - 2, where the loop condition holds.
__device__ inline void ray_skip(Ray& r, float nearLim) {
    const float q = fminf(nearLim, r.distEnd - 2.0f) - r.dist_s;
    int n = q >= 1.0f ? (int)q : 0;
    for (; n >= 8; n -= 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            r.res = make_float3(r.res.x + r.dir.x, r.res.y + r.dir.y, r.res.z + r.dir.z);
            r.dist_s += 1.0f;
        }
    }
    for (; n > 0; --n) {
        r.res = make_float3(r.res.x + r.dir.x, r.res.y + r.dir.y, r.res.z + r.dir.z);
        r.dist_s += 1.0f;
    }
}

// MapRenderingDelegate::operator() for pixel (x, y); out points at the
// pixel's float4. A miss writes only x = NaN, as the reference does.
// [nearLim, farLim] (in the unit of dist_s) brackets every sample that can
// read a voxel of an allocated block (ray_bounds, vm_raycast.hip): samples
// before nearLim are skipped, and a ray that reaches farLim can only read
// (1.0, invalid) from there to distEnd, which ends in a miss, so the march
// stops there.
// (The per-pixel and per-block functions of these headers are force-inlined:
// in a file whose kernels call one of them several times the compiler
// otherwise calls it out of line, with a stack frame, which cost the batched
// raycast a quarter of its time.)
template <int FMT, int RUNS = kRunsNone>
__device__ __forceinline__ void raycast_pixel(const MapView& m, float* out, int x, int y, const Intrinsics& K,
                                     const Affine& pose, const Affine& Tinv, float vsInv, float rtStep,
                                     float nearLim = 0.0f, float farLim = INFINITY) {
    Ray r;
    ray_begin(r, x, y, K, pose, vsInv);
    r.distEnd = fminf(r.distEnd, farLim);
    ray_skip(r, nearLim);
    RayReader<FMT> rd(m);
    // The march loop (ray_tracing.cu:355-379) takes, per pass, one box of a
    // walk through empty blocks for the lanes that are walking and one march
    // step for the others, so that the lanes of a wave walk side by side: a
    // whole walk run at once by one lane is a loop of additions that its
    // other lanes wait through, one lane after the other. (Written out here
    // and in the pose kernel: as a shared inline function it compiled to a
    // raycast 12 % slower.)
    int st = kRayGoing;
    Walk wk{0, 0, 0, 0, false};
    do {
        if (wk.on) wk.on = walk_box(rd, r.res, r.dir, r.inv, r.dist_s, r.distEnd, wk.bx, wk.by, wk.bz, wk.steps);
        if (!wk.on) st = ray_step<FMT, RUNS>(r, rd, rtStep, wk);
    } while (st == kRayGoing);
    float4 v;
    if (st == kRaySurface && ray_surface(r, rd, Tinv, vsInv, rtStep, v)) *(float4*)out = v;
    else out[0] = __int_as_float(0x7fffffff);
}

// ---- per-tile ray-distance bounds ------------------------------------------
//
// For one raycast, per 8x8 pixel tile, the range of ray distance (voxels
// along the ray, the unit of the march's dist_s) in which a ray of the tile
// can read a voxel of an allocated block. Entries are 64-bit keys stamped
// with the raycast's epoch, so a launch needs no clearing pass: near keys
// (~epoch << 32 | near bits) take atomicMin, far keys (epoch << 32 | far
// bits) atomicMax, and a key of another epoch reads as "no block" (near +inf,
// far 0). Non-negative floats order as their bit patterns. An entry that
// carries the current epoch from an earlier launch only widens the range.
// Layout: near keys of tiles 0..n-1 and the all-tiles slot n, then the far
// keys in the same order (2 (n + 1) words), n = RayBounds::nTiles (the
// capacity, at least the image's tile count); struct RayBounds, vm_common.h.

// The bounds compare a box's world distance from the camera centre with the
// march's dist_s, measured along the camera-frame ray: they hold for a pose
// whose rotation is orthonormal. Callers pass no keys for other poses
// (|R^T R - I| above 1e-5, which moves a distance of at most 2500 voxels by
// less than 0.05 voxel, inside the one voxel of slack below).
//
// Bound, in voxels, on how far the march's rounded positions and its dist_s
// drift from the exact ray: at most dEnd additions (every step is >= 1
// voxel) round a coordinate of magnitude <= cmax by <= ulp(cmax) / 2 each and
// a step * dir product by <= 2^-24 * step; sqrt(3) for the three axes; 8 ulp
// for the ray's start and direction; one voxel of slack. The longest ray,
// |(ifx (x - cx), ify (y - cy), 1)| * 5 m / vs, is convex in (x, y), so its
// maximum is at a corner pixel; the 1.001 factors cover the float rounding
// of this estimate itself.
__host__ __device__ inline float ray_drift_margin(int W, int H, const Intrinsics& K, float tx, float ty, float tz,
                                                  float vsInv) {
    float dEnd = 0.0f;
    for (int k = 0; k < 4; ++k) {
        const float ux = K.ifx * ((float)((k & 1) ? W - 1 : 0) - K.cx);
        const float uy = K.ify * ((float)((k & 2) ? H - 1 : 0) - K.cy);
        dEnd = fmaxf(dEnd, sqrtf(ux * ux + uy * uy + 1.0f) * 5.0f * vsInv);
    }
    dEnd = dEnd * 1.001f + 1.0f;
    const float cmax = sqrtf(tx * tx + ty * ty + tz * tz) * vsInv * 1.001f + dEnd + 1.0f;
    int e = 0;
    frexpf(cmax, &e);  // cmax in [2^(e-1), 2^e)
    const float ulp = ldexpf(1.0f, e - 24);
    return 1.0f + 1.75f * dEnd * (0.5f * ulp + 0x1p-23f) + 8.0f * ulp;
}

// Pixel tiles of the bounds (2^SHIFT pixels square; kBoundsTileShift and
// kBatchBoundsTileShift, vm_common.h), row-major over the image; a lane
// takes the tile of its own pixel
template <int SHIFT = kBoundsTileShift>
__host__ __device__ inline int bounds_tiles_x(int W) { return (W + (1 << SHIFT) - 1) >> SHIFT; }
template <int SHIFT = kBoundsTileShift>
__host__ __device__ inline int bounds_tile(int x, int y, int W) {
    return (y >> SHIFT) * bounds_tiles_x<SHIFT>(W) + (x >> SHIFT);
}

// [nearLim, farLim] of tile t for raycast_pixel; without keys (a pose whose
// rotation is not orthonormal) the whole ray
__device__ inline void ray_bounds_limits(const RayBounds& rb, int t, float margin, float& nearLim, float& farLim) {
    if (!rb.keys) {
        nearLim = 0.0f;
        farLim = INFINITY;
        return;
    }
    const unsigned long long* k = rb.keys;
    const int n1 = rb.nTiles + 1;
    const unsigned long long nt = k[t], ng = k[rb.nTiles], ft = k[n1 + t], fg = k[n1 + rb.nTiles];
    const unsigned e = rb.epoch, ne = ~rb.epoch;
    float nr = INFINITY, fr = 0.0f;
    if ((unsigned)(nt >> 32) == ne) nr = __uint_as_float((unsigned)nt);
    if ((unsigned)(ng >> 32) == ne) nr = fminf(nr, __uint_as_float((unsigned)ng));
    if ((unsigned)(ft >> 32) == e) fr = __uint_as_float((unsigned)ft);
    if ((unsigned)(fg >> 32) == e) fr = fmaxf(fr, __uint_as_float((unsigned)fg));
    nearLim = nr - margin;
    farLim = fr + margin;
}

struct TileRect {
    int tx0, ty0, w, n;  // first tile, width in tiles, tile count
    unsigned nearBits, farBits;
};

// A rectangle of at most this many tiles is written by the thread that found
// it; larger ones (blocks near the camera) are shared by the workgroup.
constexpr int kBoundsLightTiles = 32;

template <int CH>
struct BoundsShared {
    TileRect heavy[CH];
    int nHeavy;
};

// Reach of one allocated block (b.x != kEmptyKey) from a camera at Tinv
// (centre c, voxel units) for a W x H image: the distance range [near, far]
// of its box of march positions from c (float bits), and where rays can
// reach it: kReachNone (behind z = 5 cm or beyond z = 5.05 m), kReachAll (the
// box reaches z < 1 cm: any ray), kReachRect (the pixels [x0, x1] x [y0, y1],
// the hull of its corners' projections grown by a pixel, clipped to the
// image; not written when the rectangle misses the image: kReachNone).
enum BlockReach : int { kReachNone = 0, kReachAll = 1, kReachRect = 2 };
struct Reach {
    int kind;
    unsigned nearBits, farBits;
    int x0, x1, y0, y1;
};
// squared distance range [near2, far2] of the box of march positions of
// block b (grown by `grow`) from the point c (voxel units)
struct BoxDist2 {
    float near2, far2;
};
__device__ __forceinline__ BoxDist2 block_dist2(const int4 b, float3 c, float grow) {
    const float lx = (float)(b.x * 8) - 1.0f - grow, hx = (float)(b.x * 8 + 8) + grow;
    const float ly = (float)(b.y * 8) - 1.0f - grow, hy = (float)(b.y * 8 + 8) + grow;
    const float lz = (float)(b.z * 8) - 1.0f - grow, hz = (float)(b.z * 8 + 8) + grow;
    const float nx = fmaxf(fmaxf(lx - c.x, c.x - hx), 0.0f);
    const float ny = fmaxf(fmaxf(ly - c.y, c.y - hy), 0.0f);
    const float nz = fmaxf(fmaxf(lz - c.z, c.z - hz), 0.0f);
    const float ex = fmaxf(fabsf(lx - c.x), fabsf(hx - c.x));
    const float ey = fmaxf(fabsf(ly - c.y), fabsf(hy - c.y));
    const float ez = fmaxf(fabsf(lz - c.z), fabsf(hz - c.z));
    return BoxDist2{nx * nx + ny * ny + nz * nz, ex * ex + ey * ey + ez * ez};
}

// where rays can reach the box of block b (Reach::kind and the pixel
// rectangle; the distance fields are left as they are)
__device__ __forceinline__ void block_reach_kind(const int4 b, int W, int H, const Intrinsics& K, const Affine& Tinv,
                                                 float vs, float grow, Reach& r) {
    const float lx = (float)(b.x * 8) - 1.0f - grow, hx = (float)(b.x * 8 + 8) + grow;
    const float ly = (float)(b.y * 8) - 1.0f - grow;
    const float lz = (float)(b.z * 8) - 1.0f - grow;
    float zmin = INFINITY, zmax = -INFINITY, umin = INFINITY, umax = -INFINITY, vmin = INFINITY, vmax = -INFINITY;
    // the corners in the camera frame: the low corner's transform plus the
    // box's edges along the rotated world axes (a few ulps from transforming
    // each corner, far inside the pixel the rectangle is grown by and the
    // centimetres of the depth tests)
    const float3 p0 = xform(Tinv, make_float3(lx * vs, ly * vs, lz * vs));
    const float E = (hx - lx) * vs;  // the same edge on every axis
    const float3 dx = make_float3(Tinv.r00 * E, Tinv.r10 * E, Tinv.r20 * E);
    const float3 dy = make_float3(Tinv.r01 * E, Tinv.r11 * E, Tinv.r21 * E);
    const float3 dz = make_float3(Tinv.r02 * E, Tinv.r12 * E, Tinv.r22 * E);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float3 p = p0;
        if (k & 1) p = make_float3(p.x + dx.x, p.y + dx.y, p.z + dx.z);
        if (k & 2) p = make_float3(p.x + dy.x, p.y + dy.y, p.z + dy.z);
        if (k & 4) p = make_float3(p.x + dz.x, p.y + dz.y, p.z + dz.z);
        zmin = fminf(zmin, p.z);
        zmax = fmaxf(zmax, p.z);
        // hardware reciprocal (1 ulp): u and v move by a few 2^-23 of their
        // size, far below the one pixel the rectangle is grown by wherever
        // they fall near the image (farther out the rectangle is clipped to
        // the image anyway)
        const float iz = __builtin_amdgcn_rcpf(p.z);
        const float u = K.fx * p.x * iz + K.cx, v = K.fy * p.y * iz + K.cy;
        umin = fminf(umin, u); umax = fmaxf(umax, u);
        vmin = fminf(vmin, v); vmax = fmaxf(vmax, v);
    }
    r.kind = kReachNone;
    if (!(zmax >= 0.05f && zmin <= 5.05f)) return;
    if (zmin < 0.01f) {
        r.kind = kReachAll;
        return;
    }
    const float x0 = fmaxf(floorf(umin) - 1.0f, 0.0f), x1 = fminf(ceilf(umax) + 1.0f, (float)(W - 1));
    const float y0 = fmaxf(floorf(vmin) - 1.0f, 0.0f), y1 = fminf(ceilf(vmax) + 1.0f, (float)(H - 1));
    if (x0 <= x1 && y0 <= y1) {
        r.kind = kReachRect;
        r.x0 = (int)x0; r.x1 = (int)x1; r.y0 = (int)y0; r.y1 = (int)y1;
    }
}

__device__ __forceinline__ Reach block_reach(const int4 b, int W, int H, const Intrinsics& K, const Affine& Tinv,
                                             float3 c, float vs, float grow) {
    Reach r{kReachNone, 0u, 0u, 0, -1, 0, -1};
    const BoxDist2 d = block_dist2(b, c, grow);
    r.nearBits = __float_as_uint(sqrtf(d.near2));
    r.farBits = __float_as_uint(sqrtf(d.far2));
    block_reach_kind(b, W, H, K, Tinv, vs, grow, r);
    return r;
}

// Bounds of the allocated blocks [base, base + CH) of map m, thread i < CH
// taking block base + i, by a workgroup of at least CH threads (all of which
// share the large rectangles). A block's box of march positions whose sample can read one
// of its voxels is [8b - 1, 8b + 8) per axis (the -1 for the +1 corners of
// the trilinear read), grown by `grow` voxels (ray_drift_margin). With all
// corners in front of the camera the box projects into the hull of its
// corners' projections, so every tile meeting that rectangle (plus one
// pixel) gets the box's distance range from the camera centre c (voxel
// units). A box reaching z < 1 cm goes into the all-tiles slot. Boxes wholly
// behind z = 5 cm or beyond z = 5.05 m are out of every ray's reach (the
// march samples z in [0.1, 5] m). Most rectangles are a few tiles and their
// thread writes them; the few larger ones are appended to an LDS list and
// the workgroup's threads take their tiles in turn.
// The same for blocks whose positions the caller has loaded: b is this
// thread's block (x = kEmptyKey for none).
template <int CH, int SHIFT = kBoundsTileShift>
__device__ __forceinline__ void ray_bounds_blocks(const int4 b, const RayBounds& rb, int W, int H, const Intrinsics& K,
                                         const Affine& Tinv, float3 c, float vs, float grow, BoundsShared<CH>& sh) {
    const int tid = threadIdx.x;
    const int tilesX = bounds_tiles_x<SHIFT>(W);
    const unsigned long long ne = (unsigned long long)(~rb.epoch) << 32, fe = (unsigned long long)rb.epoch << 32;
    unsigned long long* nearK = rb.keys;
    unsigned long long* farK = rb.keys + rb.nTiles + 1;
    if (tid == 0) sh.nHeavy = 0;
    __syncthreads();
    if (b.x != kEmptyKey) {
        const Reach r = block_reach(b, W, H, K, Tinv, c, vs, grow);
        const unsigned nb = r.nearBits, fb = r.farBits;
        if (r.kind == kReachAll) {
            atomicMin(&nearK[rb.nTiles], ne | nb);
            atomicMax(&farK[rb.nTiles], fe | fb);
        } else if (r.kind == kReachRect) {
            const int tx0 = r.x0 >> SHIFT, tx1 = r.x1 >> SHIFT;
            const int ty0 = r.y0 >> SHIFT, ty1 = r.y1 >> SHIFT;
            const int n = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
            if (n <= kBoundsLightTiles) {
                for (int ty = ty0; ty <= ty1; ++ty)
                    for (int tx = tx0; tx <= tx1; ++tx) {
                        atomicMin(&nearK[ty * tilesX + tx], ne | nb);
                        atomicMax(&farK[ty * tilesX + tx], fe | fb);
                    }
            } else {
                sh.heavy[atomicAdd(&sh.nHeavy, 1)] = TileRect{tx0, ty0, tx1 - tx0 + 1, n, nb, fb};
            }
        }
    }
    __syncthreads();
    const int nh = sh.nHeavy;
    for (int r = 0; r < nh; ++r) {
        const TileRect t = sh.heavy[r];
        for (int k = tid; k < t.n; k += blockDim.x) {
            const int ty = k / t.w;
            const int tile = (t.ty0 + ty) * tilesX + t.tx0 + (k - ty * t.w);
            atomicMin(&nearK[tile], ne | t.nearBits);
            atomicMax(&farK[tile], fe | t.farBits);
        }
    }
    __syncthreads();
}

template <int CH>
__device__ inline void ray_bounds_chunk(const MapView& m, const RayBounds& rb, int base, int A, int W, int H,
                                        const Intrinsics& K, const Affine& Tinv, float3 c, float vs, float grow,
                                        BoundsShared<CH>& sh) {
    const int i = base + (int)threadIdx.x;
    const int4 b = (int)threadIdx.x < CH && i < A ? m.blockPos[i] : make_int4(kEmptyKey, 0, 0, 0);
    ray_bounds_blocks<CH>(b, rb, W, H, K, Tinv, c, vs, grow, sh);
}

// computeNormalKernel body, image_proc.cu:74-93, for pixel (x, y) with
// x < W - 1 and y < H - 1
__device__ inline float3 normalized3(float3 a) {
    const float n2 = sum3(a.x * a.x, a.y * a.y, a.z * a.z);
    if (n2 > 0.f) {
        const float n = sqrtf(n2);
        return make_float3(a.x / n, a.y / n, a.z / n);
    }
    return a;
}
__device__ inline float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ void normal_pixel(const float* vmap, int64_t vpitch, float* nmap, int64_t npitch, int W, int H,
                                    int x, int y) {
    const int x10 = max(x - 1, 0), x01 = min(x + 1, W);
    const int y10 = max(y - 1, 0), y01 = min(y + 1, H);
    auto at = [&](int yy, int xx) {
        const float4 v = *((const float4*)((const uint8_t*)vmap + (int64_t)yy * vpitch) + xx);
        return make_float3(v.x, v.y, v.z);
    };
    const float3 v00 = at(y, x10), v01 = at(y, x01), v10 = at(y10, x), v11 = at(y01, x);
    const float3 n = normalized3(cross3(make_float3(v01.x - v00.x, v01.y - v00.y, v01.z - v00.z),
                                        make_float3(v11.x - v10.x, v11.y - v10.y, v11.z - v10.z)));
    *((float4*)((uint8_t*)nmap + (int64_t)y * npitch) + x) = make_float4(n.x, n.y, n.z, 1.f);
}

}  // namespace vmamd
