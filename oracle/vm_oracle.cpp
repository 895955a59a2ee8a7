/*
 * CPU oracle: a serial restatement of the reference's fusion pipeline.
 * TEST INFRASTRUCTURE ONLY -- see vm_oracle.h. Citations are
 * reference-relative file:line.
 *
 * Functions without an independent pin are listed in the header comment of
 * tests/test_oracle.py; every function here is pinned there by a closed form,
 * brute force, invariant or the MC table text.
 */
#include "vm_oracle.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const int kBlock = 8;       /* BLOCK_SIZE, map_struct.h:6 */
const int kBlock3 = 512;    /* BLOCK_SIZE_3, map_struct.h:7 */
const int kBlockM1 = 7;     /* BLOCK_SIZE_M1, map_struct.h:8 */

int g_threads = 1;

struct V3 { float x, y, z; };
struct I3 { int x, y, z; };

/* Eigen >= 3.3 reduces a fixed-size 3-vector as a0 + (a1 + a2)
 * (redux_novec_unroller splits Length 3 into 1 + 2). */
inline float sum3(float a, float b, float c) { return a + (b + c); }
inline float dot3(V3 a, V3 b) { return sum3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline float norm3(V3 a) { return std::sqrt(dot3(a, a)); }
/* Eigen normalized(): v / sqrt(|v|^2) when |v|^2 > 0, else v unchanged */
inline V3 normalized(V3 a) {
    float n2 = dot3(a, a);
    if (n2 > 0.f) {
        float n = std::sqrt(n2);
        return V3{a.x / n, a.y / n, a.z / n};
    }
    return a;
}
/* Eigen cross (Geometry/OrthoMethods.h) */
inline V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }

/* Affine transform from a column-major 4x4: R = top-left 3x3, t = column 3.
 * Eigen's Transform * Vector3 evaluates t + R*v, each row of R*v reduced
 * as r0*v0 + (r1*v1 + r2*v2). */
struct Affine {
    float r[3][3];
    float t[3];
};
Affine affine_from(const float* T) {
    Affine a;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) a.r[i][j] = T[j * 4 + i];
        a.t[i] = T[12 + i];
    }
    return a;
}
inline V3 xform(const Affine& a, V3 v) {
    V3 o;
    o.x = a.t[0] + sum3(a.r[0][0] * v.x, a.r[0][1] * v.y, a.r[0][2] * v.z);
    o.y = a.t[1] + sum3(a.r[1][0] * v.x, a.r[1][1] * v.y, a.r[1][2] * v.z);
    o.z = a.t[2] + sum3(a.r[2][0] * v.x, a.r[2][1] * v.y, a.r[2][2] * v.z);
    return o;
}

/* float -> int with the GPU conversion semantics: NaN -> 0, saturating. */
inline int f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT_MAX;
    if (x < -2147483648.0f) return INT_MIN;
    return (int)x;
}
/* __float2int_rd (depth_fusion.cu:165): round toward -inf, saturating */
inline int f2i_rd(float x) { return f2i_sat(std::floor(x)); }

/* hash, map_struct.h:104-109: int32 wrap-around products, C remainder */
inline int hash3(int x, int y, int z, int n) {
    uint32_t r = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349669u) ^
                 ((uint32_t)z * 83492791u);
    int res = (int)r;
    res %= n;
    return res < 0 ? res + n : res;
}

/* floor, map_struct.h:111-114 */
inline I3 floor3(V3 p) {
    return I3{f2i_sat(std::floor(p.x)), f2i_sat(std::floor(p.y)), f2i_sat(std::floor(p.z))};
}
/* worldPtToVoxelPos, map_struct.h:116-120 (IEEE divide, then floor) */
inline I3 world_to_voxel(V3 p, float vs) {
    return floor3(V3{p.x / vs, p.y / vs, p.z / vs});
}
/* voxelPosToBlockPos, map_struct.h:131-141 */
inline int vox_to_blk1(int v) {
    if (v < 0) v -= kBlockM1;
    return v / kBlock;
}
inline I3 voxel_to_block(I3 v) { return I3{vox_to_blk1(v.x), vox_to_blk1(v.y), vox_to_blk1(v.z)}; }
/* voxelPosToLocalPos, map_struct.h:148-162 */
inline int vox_to_loc1(int v) {
    int r = v % kBlock;
    return r < 0 ? r + kBlock : r;
}
inline I3 voxel_to_local(I3 v) { return I3{vox_to_loc1(v.x), vox_to_loc1(v.y), vox_to_loc1(v.z)}; }
/* localPosToLocalIdx, map_struct.h:164-167 */
inline int local_idx(I3 l) { return l.z * kBlock * kBlock + l.y * kBlock + l.x; }

/* unproject, map_struct.h:189-194 */
inline V3 unproject(int x, int y, float z, float ifx, float ify, float cx, float cy) {
    return V3{ifx * ((float)x - cx) * z, ify * ((float)y - cy) * z, z};
}

/* pose inverse: general 4x4 inverse by cofactors in double, rounded to float
 * (the reference calls Eigen's Matrix4f::inverse(), depth_fusion.cu:251,
 * ray_tracing.cu:414; its float rounding is not reproducible here, so the
 * library and this oracle both use the double-precision inverse). */
void pose_inverse(const float* Tf, float* out) {
    double m[16], inv[16];
    for (int i = 0; i < 16; ++i) m[i] = Tf[i];
    /* m is column-major: element (r, c) = m[c*4 + r] */
    auto A = [&](int r, int c) { return m[c * 4 + r]; };
    double det3[4][4];
    for (int r = 0; r < 4; ++r) {
        for (int c = 0; c < 4; ++c) {
            int rr[3], cc[3], k = 0, l = 0;
            for (int i = 0; i < 4; ++i) if (i != r) rr[k++] = i;
            for (int j = 0; j < 4; ++j) if (j != c) cc[l++] = j;
            double d = A(rr[0], cc[0]) * (A(rr[1], cc[1]) * A(rr[2], cc[2]) - A(rr[1], cc[2]) * A(rr[2], cc[1]))
                     - A(rr[0], cc[1]) * (A(rr[1], cc[0]) * A(rr[2], cc[2]) - A(rr[1], cc[2]) * A(rr[2], cc[0]))
                     + A(rr[0], cc[2]) * (A(rr[1], cc[0]) * A(rr[2], cc[1]) - A(rr[1], cc[1]) * A(rr[2], cc[0]));
            det3[r][c] = ((r + c) & 1) ? -d : d; /* cofactor C(r,c) */
        }
    }
    double det = 0;
    for (int c = 0; c < 4; ++c) det += A(0, c) * det3[0][c];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
            inv[c * 4 + r] = det3[c][r] / det; /* inverse(r,c) = C(c,r)/det */
    for (int i = 0; i < 16; ++i) out[i] = (float)inv[i];
}

const char* const kTriStrings[256] = {
#include "mc_tri_strings.inc"
};
/* corner pairs of the 12 cube edges, Bourke numbering */
const int kEdgeCorners[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                 {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
int g_edge_table[256];
int g_tri_table[256][16];
struct TablesInit {
    TablesInit() {
        for (int c = 0; c < 256; ++c) {
            int e = 0;
            for (int k = 0; k < 12; ++k) {
                int a = (c >> kEdgeCorners[k][0]) & 1, b = (c >> kEdgeCorners[k][1]) & 1;
                if (a != b) e |= 1 << k;
            }
            g_edge_table[c] = e;
            for (int i = 0; i < 16; ++i) g_tri_table[c][i] = -1;
            const char* s = kTriStrings[c];
            for (int i = 0; s[i]; ++i) {
                char ch = s[i];
                g_tri_table[c][i] = (ch <= '9') ? ch - '0' : ch - 'a' + 10;
            }
        }
    }
} g_tables_init;

struct Entry { int32_t ptr, offset, x, y, z; }; /* HashEntry, map_struct.h:30-35 */

}  // namespace

struct or_map {
    int nEntry, nBucket, nVBlock;
    float voxelSize, trunc, dmin, dmax, maxW;
    int fmt, vbytes;
    std::vector<Entry> table, visible;
    std::vector<int32_t> heap;
    std::vector<uint8_t> voxels;
    int heapPtr, excessPtr, visibleCount;
    int aliasEvents, allocFailures, chainBoundHits;
};

namespace {

/* voxel access per storage format; returns (sdf, weight) as floats */
inline void vox_read(const or_map* m, int64_t idx, float* sdf, float* w) {
    const uint8_t* p = m->voxels.data() + idx * m->vbytes;
    if (m->fmt == 2) {
        std::memcpy(sdf, p, 4);
        std::memcpy(w, p + 4, 4);
    } else {
        int16_t s;
        std::memcpy(&s, p, 2);
        *sdf = (float)s / 32767.0f; /* unpackFloat, map_struct.h:72-75 */
        *w = (float)p[2];
    }
}
inline bool vox_valid(const or_map* m, int64_t idx) {
    const uint8_t* p = m->voxels.data() + idx * m->vbytes;
    if (m->fmt == 2) {
        float w;
        std::memcpy(&w, p + 4, 4);
        return w != 0.f;
    }
    return p[2] != 0;
}
inline float vox_sdf(const or_map* m, int64_t idx) {
    float s, w;
    vox_read(m, idx, &s, &w);
    return s;
}

/* findEntry, map_struct.h:356-374, with a walk bound so a corrupted chain
 * (see create_block's alias note) cannot loop forever */
inline int find_entry(or_map* m, I3 b) {
    int idx = hash3(b.x, b.y, b.z, m->nBucket);
    const Entry* e = &m->table[idx];
    if (e->ptr != -1 && e->x == b.x && e->y == b.y && e->z == b.z) return idx;
    int guard = m->nEntry;
    while (e->offset >= 0) {
        if (--guard < 0) { m->chainBoundHits++; return -1; }
        idx = m->nBucket + e->offset - 1;
        e = &m->table[idx];
        if (e->ptr != -1 && e->x == b.x && e->y == b.y && e->z == b.z) return idx;
    }
    return -1;
}

/* AllocateHashEntry, map_struct.h:275-296 (serial: pop the stack) */
inline bool allocate_entry(or_map* m, I3 b, int offset, int slot) {
    int cur = m->heapPtr--;
    if (cur >= 0) {
        int ptr = m->heap[cur];
        if (ptr != -1) {
            Entry& e = m->table[slot];
            e.x = b.x; e.y = b.y; e.z = b.z;
            e.ptr = ptr * kBlock3;
            e.offset = offset;
            return true;
        }
    }
    m->heapPtr++;
    m->allocFailures++;
    return false;
}

/* createBlock, map_struct.h:304-354, serial schedule (the try-lock always
 * succeeds). The excess slot of offset o is nBucket + o - 1, so offset 0
 * lands on the last bucket head (map_struct.h:344); when that slot is
 * occupied the reference overwrites it. That is counted in aliasEvents. */
void create_block(or_map* m, I3 b) {
    int idx = hash3(b.x, b.y, b.z, m->nBucket);
    int cur = idx;
    int empty = -1;
    const Entry* e = &m->table[cur];
    if (e->x == b.x && e->y == b.y && e->z == b.z && e->ptr != -1) return;
    if (e->ptr == -1) empty = cur;
    int guard = m->nEntry;
    while (m->table[cur].offset >= 0) {
        if (--guard < 0) { m->chainBoundHits++; return; }
        cur = m->nBucket + m->table[cur].offset - 1;
        e = &m->table[cur];
        if (e->x == b.x && e->y == b.y && e->z == b.z && e->ptr != -1) return;
        if (e->ptr == -1 && empty < 0) empty = cur;
    }
    if (empty >= 0) {
        allocate_entry(m, b, m->table[cur].offset, empty);
    } else {
        int offset = m->excessPtr++;
        if (offset + m->nBucket < m->nEntry) {
            int slot = m->nBucket + offset - 1;
            if (m->table[slot].ptr != -1) m->aliasEvents++;
            if (allocate_entry(m, b, -1, slot)) m->table[cur].offset = offset;
        } else {
            m->excessPtr--;
        }
    }
}

struct Intr { float fx, fy, cx, cy, ifx, ify; };
Intr intr_from(const float* K) {
    Intr k;
    k.fx = K[0]; k.fy = K[1]; k.cx = K[2]; k.cy = K[3];
    k.ifx = (float)(1.0 / (double)K[0]); /* depth_fusion.cu:209-210 */
    k.ify = (float)(1.0 / (double)K[1]);
    return k;
}

/* CreateBlockLineTracingFunctor::operator(), depth_fusion.cu:32-59 */
void allocate_pixels(or_map* m, const float* depth, int W, int H, const Intr& k,
                     const Affine& T) {
    const float truncHalf = (float)((double)m->trunc * 0.5); /* :226 */
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            float dist = depth[(size_t)y * W + x];
            if (std::isnan(dist) || dist < m->dmin || dist > m->dmax) continue;
            float dNear = std::fmax(m->dmin, dist - truncHalf);
            float dFar = std::fmax(m->dmin, dist + truncHalf);
            if (dNear >= dFar) continue;
            I3 pn = world_to_voxel(xform(T, unproject(x, y, dNear, k.ifx, k.ify, k.cx, k.cy)), m->voxelSize);
            I3 pf = world_to_voxel(xform(T, unproject(x, y, dFar, k.ifx, k.ify, k.cx, k.cy)), m->voxelSize);
            V3 dir{(float)(pf.x - pn.x), (float)(pf.y - pn.y), (float)(pf.z - pn.z)};
            float len = norm3(dir);
            int steps = (int)std::ceil(2.0 * (double)len); /* :53 */
            float den = (float)(steps - 1);
            dir = V3{dir.x / den, dir.y / den, dir.z / den};
            V3 pt{(float)pn.x, (float)pn.y, (float)pn.z};
            for (int s = 0; s < steps; ++s) {
                I3 v{f2i_sat(pt.x), f2i_sat(pt.y), f2i_sat(pt.z)}; /* cast<int>: toward zero */
                create_block(m, voxel_to_block(v));
                pt.x += dir.x; pt.y += dir.y; pt.z += dir.z;
            }
        }
    }
}

/* checkVertexVisible / checkBlockVisible, map_struct.h:204-244 */
inline bool block_visible(const or_map* m, I3 b, const Affine& Tinv, int cols, int rows,
                          const Intr& k) {
    float scale = m->voxelSize * (float)kBlock;
    for (int c = 0; c < 8; ++c) {
        V3 w{(float)(b.x + (c & 1)) * scale, (float)(b.y + ((c >> 1) & 1)) * scale,
             (float)(b.z + ((c >> 2) & 1)) * scale};
        V3 p = xform(Tinv, w);
        float u = k.fx * p.x / p.z + k.cx;
        float v = k.fy * p.y / p.z + k.cy;
        bool out = u < 0 || v < 0 || u > (float)(cols - 1) || v > (float)(rows - 1) ||
                   p.z < m->dmin || p.z > m->dmax;
        if (!out) return true;
    }
    return false;
}

/* CheckEntryVisibilityFunctor, depth_fusion.cu:82-126: visible entries are
 * compacted in table order (PrefixSum keeps index order inside a workgroup) */
int check_visible(or_map* m, int W, int H, const Intr& k, const Affine& Tinv) {
    m->visibleCount = 0;
    std::vector<char> vis(m->nEntry, 0);
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int i = 0; i < m->nEntry; ++i) {
        const Entry& e = m->table[i];
        if (e.ptr >= 0) vis[i] = block_visible(m, I3{e.x, e.y, e.z}, Tinv, W, H, k);
    }
    int n = 0;
    for (int i = 0; i < m->nEntry; ++i)
        if (vis[i]) m->visible[n++] = m->table[i];
    m->visibleCount = n;
    return n;
}

/* DepthFusionFunctor / ImageFusionFunctor, depth_fusion.cu:148-196, :328-380 */
void integrate(or_map* m, const float* depth, const uint8_t* rgb, int W, int H,
               const Intr& k, const Affine& Tinv) {
    const float vs = m->voxelSize, tr = m->trunc;
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads)
    for (int bi = 0; bi < m->visibleCount; ++bi) {
        const Entry& cur = m->visible[bi];
        if (cur.ptr == -1) continue;
        I3 base{cur.x * kBlock, cur.y * kBlock, cur.z * kBlock};
        for (int lz = 0; lz < kBlock; ++lz)
        for (int ly = 0; ly < kBlock; ++ly)
        for (int lx = 0; lx < kBlock; ++lx) {
            V3 wp{(float)(base.x + lx) * vs, (float)(base.y + ly) * vs, (float)(base.z + lz) * vs};
            V3 p = xform(Tinv, wp);
            int u = f2i_rd((float)((double)(k.fx * p.x / p.z + k.cx) + 0.5));
            int v = f2i_rd((float)((double)(k.fy * p.y / p.z + k.cy) + 0.5));
            if (u < 0 || v < 0 || u > W - 1 || v > H - 1) continue;
            float dist = depth[(size_t)v * W + u];
            if (std::isnan(dist) || dist > m->dmax || dist < m->dmin) continue;
            float sdf = dist - p.z;
            if (sdf < -tr) continue;
            sdf = std::fmin(1.0f, sdf / tr);
            int64_t vidx = (int64_t)cur.ptr + local_idx(I3{lx, ly, lz});
            uint8_t* q = m->voxels.data() + vidx * m->vbytes;
            const uint8_t* c = rgb ? rgb + ((size_t)v * W + u) * 3 : nullptr;
            if (m->fmt == 2) {
                float os, ow;
                std::memcpy(&os, q, 4);
                std::memcpy(&ow, q + 4, 4);
                float ns, nw;
                if (ow == 0.f) { ns = sdf; nw = 1.f; }
                else {
                    ns = (os * ow + sdf * 1.0f) / (ow + 1.0f);
                    nw = std::fmin(m->maxW, ow + 1.0f);
                }
                std::memcpy(q, &ns, 4);
                std::memcpy(q + 4, &nw, 4);
                if (c) {
                    if (ow == 0.f) { q[8] = c[0]; q[9] = c[1]; q[10] = c[2]; }
                    else for (int ch = 0; ch < 3; ++ch)
                        q[8 + ch] = (uint8_t)(0.7f * (float)q[8 + ch] + 0.3f * (float)c[ch]);
                }
            } else {
                int16_t s;
                std::memcpy(&s, q, 2);
                float os = (float)s / 32767.0f;
                int ow = q[2];
                float ns;
                int nw;
                if (ow == 0) { ns = sdf; nw = 1; }
                else {
                    /* (oldSDF * oldWT + sdf * 1) / (oldWT + 1): float / int -> float */
                    ns = (os * (float)ow + sdf * 1.0f) / (float)(ow + 1);
                    nw = std::min(255, ow + 1);
                }
                int16_t ps = (int16_t)f2i_sat(ns * 32767.0f); /* packFloat, map_struct.h:77-80 */
                std::memcpy(q, &ps, 2);
                q[2] = (uint8_t)nw;
                if (c && m->fmt == 1) {
                    if (ow == 0) { q[3] = c[0]; q[4] = c[1]; q[5] = c[2]; }
                    else for (int ch = 0; ch < 3; ++ch)
                        q[3 + ch] = (uint8_t)(0.7f * (float)q[3 + ch] + 0.3f * (float)c[ch]);
                }
            }
        }
    }
}

/* MapRenderingDelegate::read_sdf, ray_tracing.cu:272-286 */
inline float read_sdf(or_map* m, V3 p, bool* valid) {
    I3 v = floor3(p);
    int e = find_entry(m, voxel_to_block(v));
    if (e >= 0) {
        int64_t idx = (int64_t)m->table[e].ptr + local_idx(voxel_to_local(v));
        if (vox_valid(m, idx)) {
            *valid = true;
            return vox_sdf(m, idx);
        }
    }
    *valid = false;
    return 1.0f;
}
/* read_sdf_interped, ray_tracing.cu:288-321 */
inline float read_sdf_interp(or_map* m, V3 p, bool* valid) {
    V3 f{p.x - std::floor(p.x), p.y - std::floor(p.y), p.z - std::floor(p.z)};
    float s0, s1, r0, r1, r2, r3;
    bool vp, v;
    s0 = read_sdf(m, p, &vp);
    s1 = read_sdf(m, V3{p.x + 1.f, p.y, p.z}, &v); vp &= v;
    r0 = (1.0f - f.x) * s0 + f.x * s1;
    s0 = read_sdf(m, V3{p.x, p.y + 1.f, p.z}, &v); vp &= v;
    s1 = read_sdf(m, V3{p.x + 1.f, p.y + 1.f, p.z}, &v); vp &= v;
    r1 = (1.0f - f.x) * s0 + f.x * s1;
    r2 = (1.0f - f.y) * r0 + f.y * r1;
    s0 = read_sdf(m, V3{p.x, p.y, p.z + 1.f}, &v); vp &= v;
    s1 = read_sdf(m, V3{p.x + 1.f, p.y, p.z + 1.f}, &v); vp &= v;
    r0 = (1.0f - f.x) * s0 + f.x * s1;
    s0 = read_sdf(m, V3{p.x, p.y + 1.f, p.z + 1.f}, &v); vp &= v;
    s1 = read_sdf(m, V3{p.x + 1.f, p.y + 1.f, p.z + 1.f}, &v); vp &= v;
    r1 = (1.0f - f.x) * s0 + f.x * s1;
    r3 = (1.0f - f.y) * r0 + f.y * r1;
    *valid = vp;
    return (1.0f - f.z) * r2 + f.z * r3;
}

/* MapRenderingDelegate::operator(), ray_tracing.cu:323-400 */
void raycast(or_map* m, float* vmap, int W, int H, const float* K, const float* T) {
    Affine pose = affine_from(T);
    float Tinvf[16];
    pose_inverse(T, Tinvf);
    Affine Tinv = affine_from(Tinvf);
    const float ifx = (float)(1.0 / (double)K[0]), ify = (float)(1.0 / (double)K[1]);
    const float cx = K[2], cy = K[3];
    const float vsInv = (float)(1.0 / (double)m->voxelSize); /* :430 */
    const float rtStep = m->trunc / m->voxelSize;             /* :431 */
    const float zNear = 0.1f, zFar = 5.f;                     /* MIN/MAX_DEPTH :340-341 */
    uint32_t nanBits = 0x7fffffff;
    float qnan;
    std::memcpy(&qnan, &nanBits, 4);
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            float* out = vmap + ((size_t)y * W + x) * 4;
            out[0] = qnan;
            float sdf = 1.0f, lastReadSDF;
            V3 pt = unproject(x, y, zNear, ifx, ify, cx, cy);
            float dist_s = norm3(pt) * vsInv;
            V3 bs = xform(pose, pt);
            bs = V3{bs.x * vsInv, bs.y * vsInv, bs.z * vsInv};
            pt = unproject(x, y, zFar, ifx, ify, cx, cy);
            float distEnd = norm3(pt) * vsInv;
            V3 be = xform(pose, pt);
            be = V3{be.x * vsInv, be.y * vsInv, be.z * vsInv};
            V3 dir = normalized(sub(be, bs));
            V3 res = bs;
            bool valid = false, found = false, early = false;
            float step;
            while (dist_s < distEnd) {
                lastReadSDF = sdf;
                sdf = read_sdf(m, res, &valid);
                if (sdf <= 0.5f && sdf >= -0.5f) sdf = read_sdf_interp(m, res, &valid);
                if (sdf <= 0.0f) break;
                if (sdf >= 0.f && lastReadSDF < 0.f) { early = true; break; }
                step = valid ? std::fmax(sdf * rtStep, 1.0f) : 1.0f;
                res = V3{res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z};
                dist_s += step;
            }
            if (early) continue;
            if (sdf <= 0.0f) {
                step = sdf * rtStep;
                res = V3{res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z};
                sdf = read_sdf_interp(m, res, &valid);
                step = sdf * rtStep;
                res = V3{res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z};
                if (valid) found = true;
            }
            if (found) {
                V3 w{res.x / vsInv, res.y / vsInv, res.z / vsInv};
                V3 c = xform(Tinv, w);
                out[0] = c.x; out[1] = c.y; out[2] = c.z; out[3] = 1.0f;
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * The raycast march of ray_tracing.cu:323-400 templated on the scalar type,
 * for finite-difference checks of the raycast adjoint: S = double evaluates
 * the march in FP64 with TSDF values (sdf of the float map, plus sparse
 * perturbations) and a double pose; S = float follows the float raycast
 * above step for step. Both return a signature of the march's discrete
 * decisions (iteration count, cells visited, interpolation flags, break) so a
 * perturbation that changes the path can be detected. */
template <typename S> struct V3T { S x, y, z; };

struct Perturb {
    const int64_t* idx;
    const double* val;
    int n;
    double at(int64_t i) const {
        for (int k = 0; k < n; ++k)
            if (idx[k] == i) return val[k];
        return 0.0;
    }
};

template <typename S>
struct MarchT {
    or_map* m;
    Perturb pert;
    uint64_t sig;
    void mix(uint64_t v) { sig = (sig ^ v) * 1099511628211ull; }
    S read(V3T<S> p, bool* valid) {
        I3 v{(int)std::floor(p.x), (int)std::floor(p.y), (int)std::floor(p.z)};
        int e = find_entry(m, voxel_to_block(v));
        if (e >= 0) {
            int64_t idx = (int64_t)m->table[e].ptr + local_idx(voxel_to_local(v));
            if (vox_valid(m, idx)) {
                *valid = true;
                return (S)vox_sdf(m, idx) + (S)pert.at(idx);
            }
        }
        *valid = false;
        return (S)1;
    }
    S interp(V3T<S> p, bool* valid) {
        V3T<S> f{p.x - std::floor(p.x), p.y - std::floor(p.y), p.z - std::floor(p.z)};
        bool vp, v;
        S s0 = read(p, &vp), s1 = read(V3T<S>{p.x + 1, p.y, p.z}, &v); vp &= v;
        S r0 = ((S)1 - f.x) * s0 + f.x * s1;
        s0 = read(V3T<S>{p.x, p.y + 1, p.z}, &v); vp &= v;
        s1 = read(V3T<S>{p.x + 1, p.y + 1, p.z}, &v); vp &= v;
        S r1 = ((S)1 - f.x) * s0 + f.x * s1;
        S r2 = ((S)1 - f.y) * r0 + f.y * r1;
        s0 = read(V3T<S>{p.x, p.y, p.z + 1}, &v); vp &= v;
        s1 = read(V3T<S>{p.x + 1, p.y, p.z + 1}, &v); vp &= v;
        r0 = ((S)1 - f.x) * s0 + f.x * s1;
        s0 = read(V3T<S>{p.x, p.y + 1, p.z + 1}, &v); vp &= v;
        s1 = read(V3T<S>{p.x + 1, p.y + 1, p.z + 1}, &v); vp &= v;
        r1 = ((S)1 - f.x) * s0 + f.x * s1;
        S r3 = ((S)1 - f.y) * r0 + f.y * r1;
        *valid = vp;
        return ((S)1 - f.z) * r2 + f.z * r3;
    }
    /* the same trilinear value with its spatial gradient d value / d p */
    S interp_grad(V3T<S> p, bool* valid, S g[3]) {
        V3T<S> f{p.x - std::floor(p.x), p.y - std::floor(p.y), p.z - std::floor(p.z)};
        bool vp = true, v;
        S c[8];
        for (int k = 0; k < 8; ++k) {
            c[k] = read(V3T<S>{p.x + (S)(k & 1), p.y + (S)((k >> 1) & 1), p.z + (S)((k >> 2) & 1)}, &v);
            vp = vp && v;
        }
        const S gx = (S)1 - f.x, gy = (S)1 - f.y, gz = (S)1 - f.z;
        const S r0 = gx * c[0] + f.x * c[1], r1 = gx * c[2] + f.x * c[3];
        const S q0 = gx * c[4] + f.x * c[5], q1 = gx * c[6] + f.x * c[7];
        const S r2 = gy * r0 + f.y * r1, r3 = gy * q0 + f.y * q1;
        g[0] = gz * (gy * (c[1] - c[0]) + f.y * (c[3] - c[2])) + f.z * (gy * (c[5] - c[4]) + f.y * (c[7] - c[6]));
        g[1] = gz * (r1 - r0) + f.z * (q1 - q0);
        g[2] = r3 - r2;
        *valid = vp;
        return gz * r2 + f.z * r3;
    }
};

template <typename S>
inline V3T<S> xformT(const S* T, V3T<S> v) {  /* column-major 4x4, t + R v */
    return V3T<S>{T[12] + (T[0] * v.x + (T[4] * v.y + T[8] * v.z)),
                  T[13] + (T[1] * v.x + (T[5] * v.y + T[9] * v.z)),
                  T[14] + (T[2] * v.x + (T[6] * v.y + T[10] * v.z))};
}

/* one pixel; returns hit, writes the camera-frame vertex */
template <typename S>
bool march_pixel(MarchT<S>& mt, int x, int y, const S* K, const S* T, const S* Ti, S vsInv, S rtStep, S* out) {
    /* 1/f in double, rounded to S (the float form is depth_fusion.cu:209's) */
    const S ifx = (S)(1.0 / (double)K[0]), ify = (S)(1.0 / (double)K[1]), cx = K[2], cy = K[3];
    const S zNear = (S)0.1f, zFar = (S)5.f;
    V3T<S> pt{ifx * ((S)x - cx) * zNear, ify * ((S)y - cy) * zNear, zNear};
    S dist_s = std::sqrt(pt.x * pt.x + (pt.y * pt.y + pt.z * pt.z)) * vsInv;
    V3T<S> bs = xformT(T, pt);
    bs = V3T<S>{bs.x * vsInv, bs.y * vsInv, bs.z * vsInv};
    pt = V3T<S>{ifx * ((S)x - cx) * zFar, ify * ((S)y - cy) * zFar, zFar};
    const S distEnd = std::sqrt(pt.x * pt.x + (pt.y * pt.y + pt.z * pt.z)) * vsInv;
    V3T<S> be = xformT(T, pt);
    be = V3T<S>{be.x * vsInv, be.y * vsInv, be.z * vsInv};
    V3T<S> dir{be.x - bs.x, be.y - bs.y, be.z - bs.z};
    const S n2 = dir.x * dir.x + (dir.y * dir.y + dir.z * dir.z);
    if (n2 > 0) {
        const S n = std::sqrt(n2);
        dir = V3T<S>{dir.x / n, dir.y / n, dir.z / n};
    }
    V3T<S> res = bs;
    S sdf = 1, last, step;
    bool valid = false;
    int iters = 0;
    while (dist_s < distEnd) {
        last = sdf;
        sdf = mt.read(res, &valid);
        bool in = false;
        if (sdf <= (S)0.5 && sdf >= (S)-0.5) {
            sdf = mt.interp(res, &valid);
            in = true;
        }
        ++iters;
        if (valid)  /* value-carrying samples: their cells and branches */
            mt.mix(((uint64_t)(uint32_t)(int)std::floor(res.x) << 32) ^ (uint64_t)(uint32_t)(int)std::floor(res.y) ^
                   ((uint64_t)(uint32_t)(int)std::floor(res.z) << 16) ^ (in ? 7ull : 3ull) ^
                   (sdf * rtStep > 1 ? 11ull : 0ull));
        if (sdf <= 0) break;
        if (sdf >= 0 && last < 0) { mt.mix(0xdeadull); return false; }
        step = valid ? std::max(sdf * rtStep, (S)1) : (S)1;
        res = V3T<S>{res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z};
        dist_s += step;
    }
    mt.mix((uint64_t)iters);
    if (!(sdf <= 0)) return false;
    step = sdf * rtStep;
    res = V3T<S>{res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z};
    mt.mix(((uint64_t)(uint32_t)(int)std::floor(res.x) << 32) ^ (uint64_t)(uint32_t)(int)std::floor(res.y) ^
           ((uint64_t)(uint32_t)(int)std::floor(res.z) << 16));
    sdf = mt.interp(res, &valid);
    step = sdf * rtStep;
    res = V3T<S>{res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z};
    if (!valid) { mt.mix(0xbeefull); return false; }
    V3T<S> w{res.x / vsInv, res.y / vsInv, res.z / vsInv};
    V3T<S> c = xformT(Ti, w);
    out[0] = c.x; out[1] = c.y; out[2] = c.z; out[3] = 1;
    return true;
}

/* The pose refinement's depth of a pixel (no counterpart in the reference;
 * it builds on the raycast, ray_tracing.cu:330-399): the march and its first
 * correction, the second correction's step, then kZeroSteps Newton steps
 * p <- p - s(p) / (grad s . dir) dir onto the zero level of the trilinear
 * TSDF; the depth is the camera z of p. Fails where the march misses, an
 * interpolation is invalid, grad s . dir is within 1e-6 of zero, or the
 * final |s| is not below 1e-3. T, Ti: column-major camToWorld and inverse. */
template <typename S>
bool surface_pixel(MarchT<S>& mt, int x, int y, const S* K, const S* T, const S* Ti, S vsInv, S rtStep, S* z) {
    const S ifx = (S)(1.0 / (double)K[0]), ify = (S)(1.0 / (double)K[1]), cx = K[2], cy = K[3];
    const S zNear = (S)0.1f, zFar = (S)5.f;
    V3T<S> pt{ifx * ((S)x - cx) * zNear, ify * ((S)y - cy) * zNear, zNear};
    S dist_s = std::sqrt(pt.x * pt.x + (pt.y * pt.y + pt.z * pt.z)) * vsInv;
    V3T<S> bs = xformT(T, pt);
    bs = V3T<S>{bs.x * vsInv, bs.y * vsInv, bs.z * vsInv};
    pt = V3T<S>{ifx * ((S)x - cx) * zFar, ify * ((S)y - cy) * zFar, zFar};
    const S distEnd = std::sqrt(pt.x * pt.x + (pt.y * pt.y + pt.z * pt.z)) * vsInv;
    V3T<S> be = xformT(T, pt);
    be = V3T<S>{be.x * vsInv, be.y * vsInv, be.z * vsInv};
    V3T<S> dir{be.x - bs.x, be.y - bs.y, be.z - bs.z};
    const S n2 = dir.x * dir.x + (dir.y * dir.y + dir.z * dir.z);
    if (n2 > 0) {
        const S n = std::sqrt(n2);
        dir = V3T<S>{dir.x / n, dir.y / n, dir.z / n};
    }
    V3T<S> res = bs;
    S sdf = 1, last, step;
    bool valid = false;
    while (dist_s < distEnd) {
        last = sdf;
        sdf = mt.read(res, &valid);
        if (sdf <= (S)0.5 && sdf >= (S)-0.5) sdf = mt.interp(res, &valid);
        if (sdf <= 0) break;
        if (sdf >= 0 && last < 0) return false;
        step = valid ? std::max(sdf * rtStep, (S)1) : (S)1;
        res = V3T<S>{res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z};
        dist_s += step;
    }
    if (!(sdf <= 0)) return false;
    step = sdf * rtStep;
    res = V3T<S>{res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z};
    S g[3];
    S v = mt.interp_grad(res, &valid, g);
    if (!valid) return false;
    step = v * rtStep;
    res = V3T<S>{res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z};
    for (int k = 0; k <= 3; ++k) {  /* three Newton steps, then the final check */
        v = mt.interp_grad(res, &valid, g);
        const S den = g[0] * dir.x + g[1] * dir.y + g[2] * dir.z;
        if (!valid || !(std::fabs(den) > (S)1e-6)) return false;
        if (k == 3) {
            if (!(std::fabs(v) < (S)1e-3)) return false;
            break;
        }
        const S t = -v / den;
        res = V3T<S>{res.x + dir.x * t, res.y + dir.y * t, res.z + dir.z * t};
    }
    V3T<S> w{res.x / vsInv, res.y / vsInv, res.z / vsInv};
    *z = xformT(Ti, w).z;
    return true;
}

/* BuildVertexArray::interpolateLinear, marching_cubes.cu:106-115 */
inline float interp_lin(float v1, float v2) {
    if ((double)std::fabs(0.0f - v1) < 1e-6) return 0;
    if ((double)std::fabs(0.0f - v2) < 1e-6) return 1;
    if ((double)std::fabs(v1 - v2) < 1e-6) return 0;
    return (0.0f - v1) / (v2 - v1);
}

}  // namespace

extern "C" {

void or_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

int or_voxel_bytes(int f) { return f == 0 ? 4 : f == 1 ? 6 : 12; }

or_map* or_map_create(int n_entry, int n_bucket, int n_block, float voxel_size,
                      float trunc_dist, float depth_min, float depth_max,
                      int voxel_format, float max_weight) {
    if (n_entry <= 0 || n_bucket <= 0 || n_bucket > n_entry || n_block <= 0 ||
        voxel_format < 0 || voxel_format > 2)
        return nullptr;
    or_map* m = new or_map;
    m->nEntry = n_entry; m->nBucket = n_bucket; m->nVBlock = n_block;
    m->voxelSize = voxel_size; m->trunc = trunc_dist;
    m->dmin = depth_min; m->dmax = depth_max; m->maxW = max_weight;
    m->fmt = voxel_format; m->vbytes = or_voxel_bytes(voxel_format);
    m->table.resize(n_entry);
    m->visible.resize(n_entry);
    m->heap.resize(n_block);
    m->voxels.resize((size_t)n_block * kBlock3 * m->vbytes);
    or_map_reset(m);
    return m;
}

void or_map_destroy(or_map* m) { delete m; }

/* MapStruct::reset, map_struct.cu:29-42 (positions are set to INT_MIN, a key
 * no block can have, where the reference leaves them uninitialised) */
void or_map_reset(or_map* m) {
    for (auto& e : m->table) { e.ptr = -1; e.offset = -1; e.x = e.y = e.z = INT_MIN; }
    for (int i = 0; i < m->nVBlock; ++i) m->heap[i] = m->nVBlock - i - 1;
    m->heapPtr = m->nVBlock - 1;
    m->excessPtr = 0;
    m->visibleCount = 0;
    m->aliasEvents = m->allocFailures = m->chainBoundHits = 0;
    std::fill(m->voxels.begin(), m->voxels.end(), 0);
}

void or_allocate(or_map* m, const float* depth, int W, int H, const float* K, const float* T) {
    allocate_pixels(m, depth, W, H, intr_from(K), affine_from(T));
}

int or_check_visible(or_map* m, int W, int H, const float* K, const float* T) {
    float Ti[16];
    pose_inverse(T, Ti);
    return check_visible(m, W, H, intr_from(K), affine_from(Ti));
}

/* FuseImage / FuseDepthAndImage, depth_fusion.cu:199-301, :383-488 */
int or_fuse(or_map* m, const float* depth, const uint8_t* rgb, int W, int H,
            const float* K, const float* T) {
    Intr k = intr_from(K);
    allocate_pixels(m, depth, W, H, k, affine_from(T));
    float Ti[16];
    pose_inverse(T, Ti);
    Affine Tinv = affine_from(Ti);
    int n = check_visible(m, W, H, k, Tinv);
    if (n == 0) return 0;
    integrate(m, depth, rgb, W, H, k, Tinv);
    return n;
}

void or_raycast(or_map* m, float* vmap, int W, int H, const float* K, const float* T) {
    raycast(m, vmap, W, H, K, T);
}

/* VoxelizationImpl::render_scene, vmapping.cpp:83-104. A visible block always
 * yields at least one rendering block (a corner inside the image gives a
 * non-empty rectangle with zFar >= depthMin), so the rendering-block gate at
 * vmapping.cpp:101 equals the visible-count gate. */
int or_render_scene(or_map* m, float* vmap, int W, int H, const float* K, const float* T) {
    int n = or_check_visible(m, W, H, K, T);
    if (n == 0) return 0;
    raycast(m, vmap, W, H, K, T);
    return n;
}

/* computeNormalKernel, image_proc.cu:74-93 */
void or_compute_normals(const float* vmap, float* nmap, int W, int H) {
    for (int y = 0; y < H - 1; ++y) {
        for (int x = 0; x < W - 1; ++x) {
            int x10 = std::max(x - 1, 0), x01 = std::min(x + 1, W);
            int y10 = std::max(y - 1, 0), y01 = std::min(y + 1, H);
            auto at = [&](int yy, int xx) {
                const float* p = vmap + ((size_t)yy * W + xx) * 4;
                return V3{p[0], p[1], p[2]};
            };
            V3 n = normalized(cross(sub(at(y, x01), at(y, x10)), sub(at(y01, x), at(y10, x))));
            float* o = nmap + ((size_t)y * W + x) * 4;
            o[0] = n.x; o[1] = n.y; o[2] = n.z; o[3] = 1.f;
        }
    }
}

/* renderPoint / renderSceneKernel, image_proc.cu:8-61, light at (5,5,5) */
void or_render_shading(const float* vmap, const float* nmap, uint8_t* rgba, int W, int H) {
    const V3 light{5.f, 5.f, 5.f};
    for (int i = 0; i < W * H; ++i) {
        V3 P{vmap[i * 4], vmap[i * 4 + 1], vmap[i * 4 + 2]};
        V3 N{nmap[i * 4], nmap[i * 4 + 1], nmap[i * 4 + 2]};
        V3 col{4.f / 255.f, 2.f / 255.f, 2.f / 255.f};
        if (!std::isnan(P.x)) {
            const float Ka = 0.3f, Kd = 0.5f, Ks = 0.2f, n = 20.f;
            const float Ax = 1.f, Dx = 1.f, Sx = 1.f, Lx = 1.f;
            V3 L = normalized(sub(light, P));
            V3 Vv = normalized(sub(V3{0.f, 0.f, 0.f}, P));
            float nl = dot3(N, L);
            V3 t{2.f * N.x * nl, 2.f * N.y * nl, 2.f * N.z * nl};
            V3 R = normalized(sub(t, L));
            float I = Ax * Ka * Dx + Lx * Kd * Dx * std::fmax(0.f, nl) +
                      Lx * Ks * Sx * std::pow(std::fmax(0.f, dot3(R, Vv)), n);
            col = V3{I, I, I};
        }
        auto sat = [](float v) { return v != v ? 0.f : std::fmin(1.f, std::fmax(0.f, v)); };
        rgba[i * 4 + 0] = (uint8_t)(sat(col.x) * 255.f);
        rgba[i * 4 + 1] = (uint8_t)(sat(col.y) * 255.f);
        rgba[i * 4 + 2] = (uint8_t)(sat(col.z) * 255.f);
        rgba[i * 4 + 3] = 255;
    }
}

/* Polygonize / BuildVertexArray, marching_cubes.cu:9-305, for one block:
 * emits its triangles at verts/norms + 9 * first (skipping indices >= cap)
 * and returns how many it has. With verts == NULL it only counts. */
static int64_t mc_block(or_map* m, const Entry& e, float* verts, float* norms, int64_t first, int64_t cap) {
    const float vs = m->voxelSize;
    I3 base{e.x * kBlock, e.y * kBlock, e.z * kBlock};
    int64_t count = 0;
    for (int lz = 0; lz < kBlock; ++lz)
    for (int ly = 0; ly < kBlock; ++ly)
    for (int lx = 0; lx < kBlock; ++lx) {
        V3 pos{(float)(base.x + lx), (float)(base.y + ly), (float)(base.z + lz)};
        /* read_sdf_list :68-104, corners in Bourke order */
        static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                      {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
        float s[8];
        bool ok = true;
        for (int c = 0; c < 8 && ok; ++c) {
            bool v;
            s[c] = read_sdf(m, V3{pos.x + (float)off[c][0], pos.y + (float)off[c][1],
                                  pos.z + (float)off[c][2]}, &v);
            ok = v;
        }
        if (!ok) continue;
        int cube = 0;
        for (int c = 0; c < 8; ++c) if (s[c] < 0) cube |= 1 << c;
        int et = g_edge_table[cube];
        if (et == 0) continue;
        if (!verts) {
            for (int i = 0; g_tri_table[cube][i] != -1; i += 3) ++count;
            continue;
        }
        V3 vt[12];
        /* make_vertex :117-207 */
        if (et & 1)    { float t = interp_lin(s[0], s[1]); vt[0]  = V3{pos.x + t, pos.y + 0.f, pos.z + 0.f}; }
        if (et & 2)    { float t = interp_lin(s[1], s[2]); vt[1]  = V3{pos.x + 1.f, pos.y + t, pos.z + 0.f}; }
        if (et & 4)    { float t = interp_lin(s[2], s[3]); vt[2]  = V3{pos.x + (1.f - t), pos.y + 1.f, pos.z + 0.f}; }
        if (et & 8)    { float t = interp_lin(s[3], s[0]); vt[3]  = V3{pos.x + 0.f, pos.y + (1.f - t), pos.z + 0.f}; }
        if (et & 16)   { float t = interp_lin(s[4], s[5]); vt[4]  = V3{pos.x + t, pos.y + 0.f, pos.z + 1.f}; }
        if (et & 32)   { float t = interp_lin(s[5], s[6]); vt[5]  = V3{pos.x + 1.f, pos.y + t, pos.z + 1.f}; }
        if (et & 64)   { float t = interp_lin(s[6], s[7]); vt[6]  = V3{pos.x + (1.f - t), pos.y + 1.f, pos.z + 1.f}; }
        if (et & 128)  { float t = interp_lin(s[7], s[4]); vt[7]  = V3{pos.x + 0.f, pos.y + (1.f - t), pos.z + 1.f}; }
        if (et & 256)  { float t = interp_lin(s[0], s[4]); vt[8]  = V3{pos.x + 0.f, pos.y + 0.f, pos.z + t}; }
        if (et & 512)  { float t = interp_lin(s[1], s[5]); vt[9]  = V3{pos.x + 1.f, pos.y + 0.f, pos.z + t}; }
        if (et & 1024) { float t = interp_lin(s[2], s[6]); vt[10] = V3{pos.x + 1.f, pos.y + 1.f, pos.z + t}; }
        if (et & 2048) { float t = interp_lin(s[3], s[7]); vt[11] = V3{pos.x + 0.f, pos.y + 1.f, pos.z + t}; }
        /* emit :225-240: triangle (a,b,c) is stored as (c,b,a) */
        for (int i = 0; g_tri_table[cube][i] != -1; i += 3) {
            int64_t t = first + count++;
            if (t >= cap) continue;
            V3 a = vt[g_tri_table[cube][i]], b = vt[g_tri_table[cube][i + 1]],
               c = vt[g_tri_table[cube][i + 2]];
            V3 p0{c.x * vs, c.y * vs, c.z * vs};
            V3 p1{b.x * vs, b.y * vs, b.z * vs};
            V3 p2{a.x * vs, a.y * vs, a.z * vs};
            V3 n = normalized(cross(sub(p1, p0), sub(p2, p0)));
            float* o = verts + t * 9;
            o[0] = p0.x; o[1] = p0.y; o[2] = p0.z;
            o[3] = p1.x; o[4] = p1.y; o[5] = p1.z;
            o[6] = p2.x; o[7] = p2.y; o[8] = p2.z;
            float* q = norms + t * 9;
            for (int j = 0; j < 3; ++j) { q[j * 3] = n.x; q[j * 3 + 1] = n.y; q[j * 3 + 2] = n.z; }
        }
    }
    return count;
}

/* Polygonize over every allocated entry in table order (select_blocks
 * :25-50). The reference appends triangles through an atomic counter; here
 * the order is table order, block by block, which a count pass, a prefix sum
 * and a write pass reproduce with the blocks spread over threads. */
int64_t or_polygonize(or_map* m, float* verts, float* norms, int64_t cap, int* block_count) {
    std::vector<int> blocks;
    for (int i = 0; i < m->nEntry; ++i)
        if (m->table[i].ptr >= 0) blocks.push_back(i);
    if (block_count) *block_count = (int)blocks.size();
    const int64_t nb = (int64_t)blocks.size();
    std::vector<int64_t> first(nb + 1, 0);
#pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads)
    for (int64_t k = 0; k < nb; ++k) first[k + 1] = mc_block(m, m->table[blocks[k]], nullptr, nullptr, 0, 0);
    for (int64_t k = 0; k < nb; ++k) first[k + 1] += first[k];
#pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads)
    for (int64_t k = 0; k < nb; ++k)
        if (first[k] < cap && first[k + 1] > first[k])
            mc_block(m, m->table[blocks[k]], verts, norms, first[k], cap);
    return std::min(first[nb], cap);
}

/* compute_surface_point_kernel, marching_cubes.cu:314-328 */
void or_surface_points(const float* verts, float* points, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        const float* v = verts + i * 9;
        for (int j = 0; j < 3; ++j) points[i * 3 + j] = ((v[j] + v[3 + j]) + v[6 + j]) / 3.0f;
    }
}

/* RenderingBlockDelegate + FillRenderingBlockFunctor, ray_tracing.cu:13-169,
 * over the current visible table. Tiles beyond the 1e6 cap are dropped. */
int or_project_rendering_blocks(or_map* m, float* zx, float* zy, int cols8, int rows8,
                                const float* K, const float* Twc) {
    for (int i = 0; i < cols8 * rows8; ++i) { zx[i] = 100.f; zy[i] = 0.f; }
    Affine Tinv = affine_from(Twc);
    const float fx = K[0], fy = K[1], cx = K[2], cy = K[3];
    const float scale = m->voxelSize * (float)kBlock;
    int total = 0;
    for (int bi = 0; bi < m->visibleCount; ++bi) {
        const Entry& e = m->visible[bi];
        if (e.ptr == -1) continue;
        int ulx = cols8, uly = rows8, lrx = -1, lry = -1;
        float z0 = m->dmax, z1 = m->dmin;
        for (int c = 0; c < 8; ++c) {
            V3 w{(float)(e.x + (c & 1)) * scale, (float)(e.y + ((c >> 1) & 1)) * scale,
                 (float)(e.z + ((c >> 2) & 1)) * scale};
            V3 p = xform(Tinv, w);
            float px = (fx * p.x / p.z + cx) / 8.0f, py = (fy * p.y / p.z + cy) / 8.0f;
            if ((float)ulx > std::floor(px)) ulx = (short)f2i_sat(std::floor(px));
            if ((float)lrx < std::ceil(px)) lrx = (short)f2i_sat(std::ceil(px));
            if ((float)uly > std::floor(py)) uly = (short)f2i_sat(std::floor(py));
            if ((float)lry < std::ceil(py)) lry = (short)f2i_sat(std::ceil(py));
            if (z0 > p.z) z0 = p.z;
            if (z1 < p.z) z1 = p.z;
        }
        if (ulx < 0) ulx = 0;
        if (uly < 0) uly = 0;
        if (lrx >= cols8) lrx = cols8 - 1;
        if (lry >= rows8) lry = rows8 - 1;
        if (ulx > lrx || uly > lry) continue;
        if (z0 < m->dmin) z0 = m->dmin;
        if (z1 < m->dmin) continue;
        int nx = (int)std::ceil((float)(lrx - ulx + 1) / 16.f);
        int ny = (int)std::ceil((float)(lry - uly + 1) / 16.f);
        if (total + nx * ny >= 1000000) continue;
        total += nx * ny;
        for (int yy = uly; yy <= lry; ++yy)
            for (int xx = ulx; xx <= lrx; ++xx) {
                float& a = zx[yy * cols8 + xx];
                float& b = zy[yy * cols8 + xx];
                a = std::fmin(a, z0);
                b = std::fmax(b, z1);
            }
    }
    return total;
}

/* createBlock for one key (serial); returns the entry's voxel ptr or -1 */
int or_insert_block(or_map* m, int x, int y, int z) {
    create_block(m, I3{x, y, z});
    int e = find_entry(m, I3{x, y, z});
    return e >= 0 ? m->table[e].ptr : -1;
}

/* The raycast of a pixel list in FP32 (the float raycast's decisions, with
 * their signature) or FP64 (TSDF values perturbed at pidx by pval). */
void or_raycast_pixels_f32(or_map* m, int n, const int* px, const int* py, const float* K, const float* T,
                           float* out_v, uint64_t* sig, int* hit) {
    float Ti[16];
    pose_inverse(T, Ti);
    const float vsInv = (float)(1.0 / (double)m->voxelSize), rtStep = m->trunc / m->voxelSize;
#pragma omp parallel for schedule(dynamic, 8) num_threads(g_threads)
    for (int i = 0; i < n; ++i) {
        MarchT<float> mt{m, Perturb{nullptr, nullptr, 0}, 1469598103934665603ull};
        float* o = out_v + (size_t)i * 4;
        o[0] = NAN; o[1] = o[2] = o[3] = 0;
        hit[i] = march_pixel<float>(mt, px[i], py[i], K, T, Ti, vsInv, rtStep, o);
        sig[i] = mt.sig;
    }
}

void or_raycast_pixels_f64(or_map* m, int n, const int* px, const int* py, const double* K, const double* T,
                           const int64_t* pidx, const double* pval, int npert, double* out_v, uint64_t* sig,
                           int* hit) {
    /* rigid inverse in double: [R^T | -R^T t] */
    double Ti[16] = {0};
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) Ti[c * 4 + r] = T[r * 4 + c];
        Ti[12 + r] = -(T[r * 4 + 0] * T[12] + T[r * 4 + 1] * T[13] + T[r * 4 + 2] * T[14]);
    }
    Ti[15] = 1;
    const double vsInv = 1.0 / (double)m->voxelSize, rtStep = (double)m->trunc / (double)m->voxelSize;
#pragma omp parallel for schedule(dynamic, 8) num_threads(g_threads)
    for (int i = 0; i < n; ++i) {
        MarchT<double> mt{m, Perturb{pidx, pval, npert}, 1469598103934665603ull};
        double* o = out_v + (size_t)i * 4;
        o[0] = NAN; o[1] = o[2] = o[3] = 0;
        hit[i] = march_pixel<double>(mt, px[i], py[i], K, T, Ti, vsInv, rtStep, o);
        sig[i] = mt.sig;
    }
}

void or_surface_depth_f64(or_map* m, int n, const int* px, const int* py, const double* K, const double* T,
                          double* out_z, int* ok) {
    double Ti[16] = {0};
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) Ti[c * 4 + r] = T[r * 4 + c];
        Ti[12 + r] = -(T[r * 4 + 0] * T[12] + T[r * 4 + 1] * T[13] + T[r * 4 + 2] * T[14]);
    }
    Ti[15] = 1;
    const double vsInv = 1.0 / (double)m->voxelSize, rtStep = (double)m->trunc / (double)m->voxelSize;
#pragma omp parallel for schedule(dynamic, 8) num_threads(g_threads)
    for (int i = 0; i < n; ++i) {
        MarchT<double> mt{m, Perturb{nullptr, nullptr, 0}, 0ull};
        out_z[i] = NAN;
        ok[i] = surface_pixel<double>(mt, px[i], py[i], K, T, Ti, vsInv, rtStep, out_z + i);
    }
}

int or_map_counters(or_map* m, int* out) {
    out[0] = m->heapPtr; out[1] = m->excessPtr; out[2] = m->visibleCount;
    out[3] = m->aliasEvents; out[4] = m->allocFailures; out[5] = m->chainBoundHits;
    return 6;
}
int32_t* or_map_hash_table(or_map* m) { return &m->table[0].ptr; }
int32_t* or_map_heap(or_map* m) { return m->heap.data(); }
uint8_t* or_map_voxels(or_map* m) { return m->voxels.data(); }
int32_t* or_map_visible_table(or_map* m) { return &m->visible[0].ptr; }

int or_hash(int x, int y, int z, int n) { return hash3(x, y, z, n); }
void or_voxel_to_block(int vx, int vy, int vz, int* o) {
    I3 b = voxel_to_block(I3{vx, vy, vz});
    o[0] = b.x; o[1] = b.y; o[2] = b.z;
}
void or_voxel_to_local(int vx, int vy, int vz, int* o) {
    I3 l = voxel_to_local(I3{vx, vy, vz});
    o[0] = l.x; o[1] = l.y; o[2] = l.z;
}
int or_float2int_rd(float x) { return f2i_rd(x); }
void or_pose_inverse(const float* T, float* out) { pose_inverse(T, out); }
int or_mc_edge_table(int c) { return g_edge_table[c & 255]; }
int or_mc_tri_table(int c, int i) { return g_tri_table[c & 255][i & 15]; }

}  // extern "C"
