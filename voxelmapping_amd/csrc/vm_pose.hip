// Batched pose refinement: Gauss-Newton on the depth residual
// z(T) - depth over a pixel grid, for many camera-pose hypotheses in one
// Gauss-Newton launch per iteration (after a small launch that bounds each
// hypothesis's rays). No counterpart in the reference; it builds on its
// raycast (MapRenderingDelegate, ray_tracing.cu:323-400).
//
// z(T) of a pixel is the depth of the map's zero level along the pixel's ray:
// the raycast's hit (the same march and corrections) refined by Newton steps
// on the trilinear TSDF, p <- p - s(p) / (grad s . dir) dir. For camToWorld =
// T exp(xi), xi = (rho, phi), the hit p = o + t dir (o the ray start) keeps
// s(p) = 0, so dt = -grad s . (do + t ddir) / (grad s . dir) and
// dp = do + t ddir + dt dir (implicit differentiation); the camera point is
// v = exp(-xi) T^-1 p, dv = R^T dp - rho - phi x v. The raycast's own hit is
// not used as z: its two fixed corrections stop short of the zero level on
// oblique rays by a pose-dependent amount, which makes its exact derivative
// (the raycast adjoint's) a poor Gauss-Newton model.
//
// Each iteration accumulates J^T J, J^T r, sum r^2, the residual count and
// the count of pixels with a valid depth per hypothesis (per-workgroup
// partials, reduced in double), then takes a Levenberg-Marquardt step: the
// pose is kept only if it lowers the truncated cost E = (sum r^2 over
// residuals below max_residual + max_residual^2 per other valid pixel) /
// valid pixels (else the previous pose and its normal equations come back and
// the damping grows tenfold; an accepted step divides it by ten), and the
// damped 6x6 normal equations are solved by Cholesky in double and the pose
// right-multiplied by exp(delta).
#include <algorithm>

#include "vm_common.h"
#include "vm_dev_fusion.h"
#include "vm_dev_ray.h"
#include "vm_launch.h"

namespace vmamd {

namespace {

// 21 of J^T J (upper triangle), 6 of J^T r, sum r^2, residual count, valid-depth count
constexpr int kGnTerms = 30;
constexpr int kLmState = kGnTerms + 2;  // accepted terms, their cost E, damping
constexpr int kGnThreads = 256;

__device__ inline float3 f3add(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline float3 f3mul(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ inline float f3dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float3 rmul(const Affine& T, float3 v) {
    return make_float3(T.r00 * v.x + T.r01 * v.y + T.r02 * v.z, T.r10 * v.x + T.r11 * v.y + T.r12 * v.z,
                       T.r20 * v.x + T.r21 * v.y + T.r22 * v.z);
}

// d(T exp(xi) p)/d xi_j for a camera-frame point p: R e_j (j < 3), R (e_j x p)
__device__ inline float3 point_tangent(const Affine& T, float3 p, int j) {
    float3 e = make_float3(0.f, 0.f, 0.f);
    if (j == 0) e.x = 1.f;
    else if (j == 1) e.y = 1.f;
    else if (j == 2) e.z = 1.f;
    else if (j == 3) e = make_float3(0.f, -p.z, p.y);
    else if (j == 4) e = make_float3(p.z, 0.f, -p.x);
    else e = make_float3(-p.y, p.x, 0.f);
    return rmul(T, e);
}

// Newton steps from the raycast's hit onto the zero level of the trilinear
// TSDF along the ray
constexpr int kZeroSteps = 3;

// Depth of pixel (x, y) at pose T and its derivative by the pose twist.
// Returns false when the raycast misses or the zero level is not found.
// [nearLim, farLim] brackets every sample of any ray of the image that can
// read a voxel of an allocated block (k_pose_bounds), as the raycast's tile
// bounds do (raycast_pixel): the samples before nearLim are the same
// additions without reads, and a ray reaching farLim misses.
template <int FMT>
__device__ inline bool surface_depth(RayReader<FMT>& rd, int x, int y, const Intrinsics& K, const Affine& T,
                                     const Affine& Tinv, float vsInv, float rtStep, float nearLim, float farLim,
                                     float& z, float J[6]) {
    Ray r;
    ray_begin(r, x, y, K, T, vsInv);
    const float3 o = r.res;
    r.distEnd = fminf(r.distEnd, farLim);
    ray_skip(r, nearLim);
    // the raycast's march loop: walks through empty blocks one box per pass
    // (raycast_pixel, vm_dev_ray.h)
    int st = kRayGoing;
    Walk wk{0, 0, 0, 0, false};
    do {
        if (wk.on) wk.on = walk_box(rd, r.res, r.dir, r.inv, r.dist_s, r.distEnd, wk.bx, wk.by, wk.bz, wk.steps);
        if (!wk.on) st = ray_step(r, rd, rtStep, wk);
    } while (st == kRayGoing);
    if (st != kRaySurface) return false;
    float step = r.sdf * rtStep;
    float3 p = make_float3(r.res.x + step * r.dir.x, r.res.y + step * r.dir.y, r.res.z + step * r.dir.z);
    Trilinear t;
    trilinear(rd, p, t);
    if (!t.valid) return false;
    step = t.value * rtStep;
    p = make_float3(p.x + step * r.dir.x, p.y + step * r.dir.y, p.z + step * r.dir.z);
    // zero level along the ray
    for (int k = 0; k < kZeroSteps; ++k) {
        trilinear(rd, p, t);
        const float den = f3dot(t.grad, r.dir);
        if (!t.valid || !(fabsf(den) > 1e-6f)) return false;
        p = f3add(p, f3mul(r.dir, -t.value / den));
    }
    trilinear(rd, p, t);
    const float den = f3dot(t.grad, r.dir);
    if (!t.valid || !(fabsf(den) > 1e-6f) || !(fabsf(t.value) < 1e-3f)) return false;
    const float3 w = make_float3(p.x / vsInv, p.y / vsInv, p.z / vsInv);
    const float3 v = xform(Tinv, w);
    z = v.z;
    // tangents of the ray start o and direction (voxel units)
    const float zNear = 0.1f, zFar = 5.0f;
    const float ux = K.ifx * ((float)x - K.cx), uy = K.ify * ((float)y - K.cy);
    const float3 pn = make_float3(ux * zNear, uy * zNear, zNear), pf = make_float3(ux * zFar, uy * zFar, zFar);
    float3 be = xform(T, pf);
    be = make_float3(be.x * vsInv, be.y * vsInv, be.z * vsInv);
    const float3 dvec = make_float3(be.x - o.x, be.y - o.y, be.z - o.z);
    const float dn = sqrtf(sum3(dvec.x * dvec.x, dvec.y * dvec.y, dvec.z * dvec.z));
    const float tt = f3dot(make_float3(p.x - o.x, p.y - o.y, p.z - o.z), r.dir);
    const float3 c2 = make_float3(T.r02, T.r12, T.r22);  // row 2 of R^T
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const float3 os = f3mul(point_tangent(T, pn, j), vsInv);
        const float3 oe = f3mul(point_tangent(T, pf, j), vsInv);
        const float3 tv = make_float3(oe.x - os.x, oe.y - os.y, oe.z - os.z);
        const float3 dd = dn > 0.f ? f3mul(f3add(tv, f3mul(r.dir, -f3dot(r.dir, tv))), 1.0f / dn)
                                   : make_float3(0.f, 0.f, 0.f);
        const float3 q = f3add(os, f3mul(dd, tt));
        const float3 pd = f3add(q, f3mul(r.dir, -f3dot(t.grad, q) / den));
        float dz = f3dot(c2, pd) / vsInv;
        if (j == 2) dz -= 1.0f;
        if (j == 3) dz -= v.y;
        if (j == 4) dz += v.x;
        J[j] = dz;
    }
    return true;
}

// Per hypothesis, the range of ray distance (voxels along the ray, the unit
// of the march's dist_s) in which any ray of the W x H image can read a
// voxel of an allocated block: the union of the raycast's per-tile ranges
// (block_reach over every allocated block, boxes grown by ray_drift_margin).
// Workgroup (w, h) takes blocks w * 256 + i, w * 256 + i + 256 * gridDim.x,
// ... for hypothesis h and writes its (near, far) float bits to
// bounds[2 (h gridDim.x + w)]; the march takes the minimum and maximum over
// the workgroups. A hypothesis whose rotation is not orthonormal (an entry
// of R^T R - I above 1e-5) gets no bounds (0, +inf).
constexpr int kPoseBoundsThreads = 256;
constexpr int kPoseBoundsGroups = 64;
__global__ __launch_bounds__(kPoseBoundsThreads) void k_pose_bounds(MapView m, int W, int H, Intrinsics K,
                                                                    const Affine* poses, const Affine* posesInv,
                                                                    unsigned* bounds) {
    __shared__ unsigned s_near[kPoseBoundsThreads / 64], s_far[kPoseBoundsThreads / 64];
    const int h = blockIdx.y;
    const Affine T = poses[h], Ti = posesInv[h];
    const float vs = m.voxelSize, vsInv = (float)(1.0 / (double)m.voxelSize);
    const float3 c = make_float3(T.t0 * vsInv, T.t1 * vsInv, T.t2 * vsInv);
    const float grow = ray_drift_margin(W, H, K, T.t0, T.t1, T.t2, vsInv);
    const int A = allocated_count(m.ctr, m.nBlock);
    // squared distances, kept only for blocks the rays reach; a block whose
    // range lies inside the range kept so far cannot change it, so its
    // projection is skipped
    float n2 = INFINITY, f2 = 0.0f;
    const int stride = gridDim.x * kPoseBoundsThreads;
    // four blocks per pass, their positions loaded together; the passes are
    // counted per wave (from its first lane's block), so that every lane
    // takes part in the shuffles below
    const int wave0 = blockIdx.x * kPoseBoundsThreads + (threadIdx.x & ~63);
    for (int p0 = 0; wave0 + p0 < A; p0 += 4 * stride) {
        const int i0 = wave0 + p0 + (threadIdx.x & 63);
        int4 bq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            bq[q] = i0 + q * stride < A ? m.blockPos[i0 + q * stride] : make_int4(kEmptyKey, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int4 b = bq[q];
            if (b.x == kEmptyKey) continue;
            const BoxDist2 d = block_dist2(b, c, grow);
            if (d.near2 < n2 || d.far2 > f2) {
                Reach r;
                block_reach_kind(b, W, H, K, Ti, vs, grow, r);
                if (r.kind != kReachNone) {
                    n2 = fminf(n2, d.near2);
                    f2 = fmaxf(f2, d.far2);
                }
            }
        }
        // the wave's range so far in every lane: a lane then projects only
        // the blocks that could widen the wave's range, and the wave skips
        // the projection when none of its lanes has such a block
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            n2 = fminf(n2, __shfl_xor(n2, o, 64));
            f2 = fmaxf(f2, __shfl_xor(f2, o, 64));
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        n2 = fminf(n2, __shfl_xor(n2, o, 64));
        f2 = fmaxf(f2, __shfl_xor(f2, o, 64));
    }
    // the rounded square roots of the extremes are the extremes of the
    // rounded square roots (sqrt is monotone): the raycast's per-block values
    unsigned nb = __float_as_uint(sqrtf(n2)), fb = __float_as_uint(sqrtf(f2));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_near[wave] = nb;
        s_far[wave] = fb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kPoseBoundsThreads / 64; ++w) {
            nb = min(nb, s_near[w]);
            fb = max(fb, s_far[w]);
        }
        const float rr[3][3] = {{T.r00, T.r01, T.r02}, {T.r10, T.r11, T.r12}, {T.r20, T.r21, T.r22}};
        bool rigid = true;
        for (int a = 0; a < 3; ++a)
            for (int e = 0; e < 3; ++e) {
                const double dot = (double)rr[0][a] * rr[0][e] + (double)rr[1][a] * rr[1][e] + (double)rr[2][a] * rr[2][e];
                rigid = rigid && fabs(dot - (a == e ? 1.0 : 0.0)) <= 1e-5;
            }
        if (!rigid) {
            nb = 0u;
            fb = __float_as_uint(INFINITY);
        }
        bounds[2 * ((int64_t)h * gridDim.x + blockIdx.x)] = nb;
        bounds[2 * ((int64_t)h * gridDim.x + blockIdx.x) + 1] = fb;
    }
}

// one workgroup: kGnThreads sampled pixels of hypothesis blockIdx.y
template <int FMT, int DF>
__global__ __launch_bounds__(kGnThreads) void k_pose_gn(MapView m, DepthImage d, Intrinsics K, const Affine* poses,
                                                        const Affine* posesInv, int stride, int sw, int sh,
                                                        float vsInv, float rtStep, float maxRes,
                                                        const unsigned* bounds, float* partials) {
    __shared__ float s_part[kGnThreads / 64][kGnTerms];
    __shared__ unsigned s_lim[2];
    const int h = blockIdx.y;
    // the hypothesis's ray-distance range, from k_pose_bounds's partials
    static_assert(kPoseBoundsGroups <= 64 && kGnThreads >= 64, "one wave reduces the bounds partials");
    if (threadIdx.x < 64) {
        unsigned nb = __float_as_uint(INFINITY), fb = 0u;
        if (threadIdx.x < kPoseBoundsGroups) {
            nb = bounds[2 * ((int64_t)h * kPoseBoundsGroups + threadIdx.x)];
            fb = bounds[2 * ((int64_t)h * kPoseBoundsGroups + threadIdx.x) + 1];
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            nb = min(nb, (unsigned)__shfl_xor((int)nb, o, 64));
            fb = max(fb, (unsigned)__shfl_xor((int)fb, o, 64));
        }
        if (threadIdx.x == 0) {
            s_lim[0] = nb;
            s_lim[1] = fb;
        }
    }
    __syncthreads();
    const float margin = ray_drift_margin(d.width, d.height, K, poses[h].t0, poses[h].t1, poses[h].t2, vsInv);
    const float nearLim = __uint_as_float(s_lim[0]) - margin, farLim = __uint_as_float(s_lim[1]) + margin;
    const int i = blockIdx.x * kGnThreads + threadIdx.x;
    float acc[kGnTerms];
#pragma unroll
    for (int k = 0; k < kGnTerms; ++k) acc[k] = 0.f;
    if (i < sw * sh) {
        const int x = (i % sw) * stride, y = (i / sw) * stride;
        const float dep = load_depth<DF>(d, x, y);
        if (dep == dep && dep >= m.depthMin && dep <= m.depthMax) {
            acc[29] = 1.f;
            RayReader<FMT> rd(m);
            float z, J[6];
            if (surface_depth(rd, x, y, K, poses[h], posesInv[h], vsInv, rtStep, nearLim, farLim, z, J)) {
                const float res = z - dep;
                if (fabsf(res) < maxRes) {
                    int k = 0;
#pragma unroll
                    for (int a = 0; a < 6; ++a)
#pragma unroll
                        for (int b = a; b < 6; ++b) acc[k++] = J[a] * J[b];
#pragma unroll
                    for (int a = 0; a < 6; ++a) acc[21 + a] = J[a] * res;
                    acc[27] = res * res;
                    acc[28] = 1.f;
                }
            }
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kGnTerms; ++k) {
        float v = acc[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) s_part[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < kGnTerms) {
        float v = 0.f;
        for (int w = 0; w < kGnThreads / 64; ++w) v += s_part[w][threadIdx.x];
        partials[((int64_t)h * gridDim.x + blockIdx.x) * kGnTerms + threadIdx.x] = v;
    }
}

// exp of a twist (rho, phi) as a 4x4 row-major double matrix
__device__ inline void se3_exp(const double xi[6], double E[16]) {
    const double wx = xi[3], wy = xi[4], wz = xi[5];
    const double th2 = wx * wx + wy * wy + wz * wz, th = sqrt(th2);
    double A, B, C;  // sin(th)/th, (1-cos)/th^2, (th-sin)/th^3
    if (th < 1e-6) {
        A = 1.0 - th2 / 6.0;
        B = 0.5 - th2 / 24.0;
        C = 1.0 / 6.0 - th2 / 120.0;
    } else {
        A = sin(th) / th;
        B = (1.0 - cos(th)) / th2;
        C = (th - sin(th)) / (th2 * th);
    }
    const double W[9] = {0, -wz, wy, wz, 0, -wx, -wy, wx, 0};
    double W2[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double s = 0;
            for (int c = 0; c < 3; ++c) s += W[a * 3 + c] * W[c * 3 + b];
            W2[a * 3 + b] = s;
        }
    for (int a = 0; a < 3; ++a) {
        double t = 0;
        for (int b = 0; b < 3; ++b) {
            const double I = a == b ? 1.0 : 0.0;
            E[a * 4 + b] = I + A * W[a * 3 + b] + B * W2[a * 3 + b];
            t += (I + B * W[a * 3 + b] + C * W2[a * 3 + b]) * xi[b];
        }
        E[a * 4 + 3] = t;
    }
    E[12] = E[13] = E[14] = 0.0;
    E[15] = 1.0;
}

__device__ inline Affine affine_from(const double T[16]) {
    Affine a;
    a.r00 = (float)T[0]; a.r01 = (float)T[1]; a.r02 = (float)T[2]; a.t0 = (float)T[3];
    a.r10 = (float)T[4]; a.r11 = (float)T[5]; a.r12 = (float)T[6]; a.t1 = (float)T[7];
    a.r20 = (float)T[8]; a.r21 = (float)T[9]; a.r22 = (float)T[10]; a.t2 = (float)T[11];
    return a;
}

// Rigid inverse in double: [R^T, -R^T t]
__device__ inline Affine rigid_inverse(const double T[16]) {
    double I[16];
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) I[a * 4 + b] = T[b * 4 + a];
        I[a * 4 + 3] = -(T[0 * 4 + a] * T[3] + T[1 * 4 + a] * T[7] + T[2 * 4 + a] * T[11]);
    }
    return affine_from(I);
}

// One workgroup of kSolveThreads per hypothesis: the partials are summed in
// double by kSolveThreads / 32 groups of 32 threads (one term per thread,
// every (kSolveThreads / 32)-th block per group) and the groups' sums added in
// LDS; thread 0 then decides on the last step (Levenberg-Marquardt, above),
// records the cost and normal equations of the pose kept, and unless this is
// the final evaluation solves (J^T J + lambda diag(J^T J)) delta = -J^T r and
// sets T <- T exp(delta). lm (kLmState doubles per hypothesis) and posesKept
// (16 doubles) hold the accepted state between launches.
constexpr int kSolveThreads = 256;
__global__ __launch_bounds__(kSolveThreads) void k_pose_solve(const float* partials, int blocksPerHyp, int nHyp,
                                                              double* posesD, Affine* poses, Affine* posesInv,
                                                              double* posesKept, double* lm, float* cost,
                                                              double* system, int iter, int iters, int update,
                                                              double lambda0, double maxRes) {
    constexpr int kGroups = kSolveThreads / 32;
    __shared__ double s_sum[kGroups][32];
    const int h = blockIdx.x;
    const int k = threadIdx.x & 31, grp = threadIdx.x >> 5;
    double v = 0.0;
    if (k < kGnTerms)
        for (int b = grp; b < blocksPerHyp; b += kGroups)
            v += (double)partials[((int64_t)h * blocksPerHyp + b) * kGnTerms + k];
    s_sum[grp][k] = v;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double a[kGnTerms];
    for (int q = 0; q < kGnTerms; ++q) {
        double t = 0.0;
        for (int g = 0; g < kGroups; ++g) t += s_sum[g][q];
        a[q] = t;
    }
    double* st = lm + (int64_t)h * kLmState;
    double* T = posesD + (int64_t)h * 16;
    double* Tk = posesKept + (int64_t)h * 16;
    const double nValid = a[29];
    const double E = nValid > 0 ? (a[27] + (nValid - a[28]) * maxRes * maxRes) / nValid : 0.0;
    double lambda;
    if (iter == 0 || E <= st[kGnTerms]) {
        // accept: this pose and its terms become the kept state
        lambda = iter == 0 ? lambda0 : fmax(st[kGnTerms + 1] * 0.1, 1e-9);
        for (int q = 0; q < kGnTerms; ++q) st[q] = a[q];
        st[kGnTerms] = E;
        for (int q = 0; q < 16; ++q) Tk[q] = T[q];
    } else {
        // reject: back to the kept pose, more damping
        lambda = st[kGnTerms + 1] * 10.0;
        for (int q = 0; q < kGnTerms; ++q) a[q] = st[q];
        for (int q = 0; q < 16; ++q) T[q] = Tk[q];
        poses[h] = affine_from(Tk);
        posesInv[h] = rigid_inverse(Tk);
    }
    st[kGnTerms + 1] = lambda;
    const double n = a[28];
    for (int q = 0; q < 27; ++q) system[((int64_t)h * (iters + 1) + iter) * 27 + q] = a[q];
    cost[((int64_t)h * (iters + 1) + iter) * 2 + 0] = n > 0 ? (float)(a[27] / n) : -1.f;
    cost[((int64_t)h * (iters + 1) + iter) * 2 + 1] = (float)n;
    if (!update || n < 6) return;
    double H[36];
    int e = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 6; ++j) {
            H[i * 6 + j] = H[j * 6 + i] = a[e++];
        }
    for (int i = 0; i < 6; ++i) H[i * 6 + i] *= 1.0 + lambda;
    // Cholesky H = L L^T
    double L[36] = {0};
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = H[i * 6 + j];
            for (int q = 0; q < j; ++q) s -= L[i * 6 + q] * L[j * 6 + q];
            if (i == j) {
                if (!(s > 0.0)) return;  // not positive definite: keep the pose
                L[i * 6 + i] = sqrt(s);
            } else {
                L[i * 6 + j] = s / L[j * 6 + j];
            }
        }
    double y[6], delta[6];
    for (int i = 0; i < 6; ++i) {
        double s = -a[21 + i];
        for (int q = 0; q < i; ++q) s -= L[i * 6 + q] * y[q];
        y[i] = s / L[i * 6 + i];
    }
    for (int i = 5; i >= 0; --i) {
        double s = y[i];
        for (int q = i + 1; q < 6; ++q) s -= L[q * 6 + i] * delta[q];
        delta[i] = s / L[i * 6 + i];
    }
    double Ex[16], Tn[16];
    se3_exp(delta, Ex);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0;
            for (int q = 0; q < 4; ++q) s += T[i * 4 + q] * Ex[q * 4 + j];
            Tn[i * 4 + j] = s;
        }
    for (int i = 0; i < 16; ++i) T[i] = Tn[i];
    poses[h] = affine_from(Tn);
    posesInv[h] = rigid_inverse(Tn);
}

// row-major double poses -> float Affine pairs
__global__ void k_pose_load(const double* posesD, Affine* poses, Affine* posesInv, int nHyp) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nHyp) return;
    poses[h] = affine_from(posesD + (int64_t)h * 16);
    posesInv[h] = rigid_inverse(posesD + (int64_t)h * 16);
}

template <int FMT>
hipError_t launch_gn(const MapView& m, const DepthImage& d, const Intrinsics& K, const Affine* poses,
                     const Affine* posesInv, int nHyp, int stride, int sw, int sh, float maxRes, unsigned* bounds,
                     float* partials, int blocksPerHyp, hipStream_t s) {
    const float vsInv = (float)(1.0 / (double)m.voxelSize);
    const float rtStep = m.trunc / m.voxelSize;
    hipLaunchKernelGGL(k_pose_bounds, dim3(kPoseBoundsGroups, nHyp), dim3(kPoseBoundsThreads), 0, s, m, d.width,
                       d.height, K, poses, posesInv, bounds);
    const dim3 grid(blocksPerHyp, nHyp);
    if (d.format == kDepthU16)
        hipLaunchKernelGGL((k_pose_gn<FMT, kDepthU16>), grid, dim3(kGnThreads), 0, s, m, d, K, poses, posesInv,
                           stride, sw, sh, vsInv, rtStep, maxRes, bounds, partials);
    else
        hipLaunchKernelGGL((k_pose_gn<FMT, kDepthF32>), grid, dim3(kGnThreads), 0, s, m, d, K, poses, posesInv,
                           stride, sw, sh, vsInv, rtStep, maxRes, bounds, partials);
    return hipGetLastError();
}

}  // namespace

size_t refine_poses_scratch_bytes(int W, int H, int nHyp, int iters, int stride) {
    const int sw = (W + stride - 1) / stride, sh = (H + stride - 1) / stride;
    const int bph = (sw * sh + kGnThreads - 1) / kGnThreads;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return up((size_t)nHyp * bph * kGnTerms * sizeof(float)) + 2 * up((size_t)nHyp * 16 * sizeof(double)) +
           up((size_t)nHyp * kLmState * sizeof(double)) + 2 * up((size_t)nHyp * sizeof(Affine)) +
           up((size_t)nHyp * (iters + 1) * 2 * sizeof(float)) + up((size_t)nHyp * (iters + 1) * 27 * sizeof(double)) +
           up((size_t)nHyp * kPoseBoundsGroups * 2 * sizeof(unsigned));
}

// posesIn: nHyp row-major 4x4 doubles (host). scratch: refine_poses_scratch_bytes.
// On return (asynchronous on s) the scratch holds the refined row-major
// doubles at `posesOut`, the costs at `costOut` and the normal equations of
// every evaluation at `systemOut`.
hipError_t launch_refine_poses(const MapView& m, const DepthImage& d, const Intrinsics& K, const double* posesIn,
                               int nHyp, int iters, int stride, float maxRes, double lambda, void* scratch,
                               double** posesOut, float** costOut, double** systemOut, hipStream_t s) {
    const int sw = (d.width + stride - 1) / stride, sh = (d.height + stride - 1) / stride;
    const int bph = (sw * sh + kGnThreads - 1) / kGnThreads;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    uint8_t* p = (uint8_t*)scratch;
    float* partials = (float*)p;
    p += up((size_t)nHyp * bph * kGnTerms * sizeof(float));
    double* posesD = (double*)p;
    p += up((size_t)nHyp * 16 * sizeof(double));
    double* posesKept = (double*)p;
    p += up((size_t)nHyp * 16 * sizeof(double));
    double* lm = (double*)p;
    p += up((size_t)nHyp * kLmState * sizeof(double));
    Affine* poses = (Affine*)p;
    p += up((size_t)nHyp * sizeof(Affine));
    Affine* posesInv = (Affine*)p;
    p += up((size_t)nHyp * sizeof(Affine));
    float* cost = (float*)p;
    p += up((size_t)nHyp * (iters + 1) * 2 * sizeof(float));
    double* system = (double*)p;
    p += up((size_t)nHyp * (iters + 1) * 27 * sizeof(double));
    unsigned* bounds = (unsigned*)p;
    hipError_t e = hipMemcpyAsync(posesD, posesIn, (size_t)nHyp * 16 * sizeof(double), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    const int tb = 64, nb = (nHyp + tb - 1) / tb;
    hipLaunchKernelGGL(k_pose_load, dim3(nb), dim3(tb), 0, s, posesD, poses, posesInv, nHyp);
    for (int it = 0; it <= iters; ++it) {
        switch (m.fmt) {
            case kFmtRef: e = launch_gn<kFmtRef>(m, d, K, poses, posesInv, nHyp, stride, sw, sh, maxRes, bounds, partials, bph, s); break;
            case kFmtRefRGB: e = launch_gn<kFmtRefRGB>(m, d, K, poses, posesInv, nHyp, stride, sw, sh, maxRes, bounds, partials, bph, s); break;
            default: e = launch_gn<kFmtF32>(m, d, K, poses, posesInv, nHyp, stride, sw, sh, maxRes, bounds, partials, bph, s);
        }
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_pose_solve, dim3(nHyp), dim3(kSolveThreads), 0, s, partials, bph, nHyp, posesD, poses,
                           posesInv, posesKept, lm, cost, system, it, iters, (int)(it < iters), lambda,
                           (double)maxRes);
    }
    *posesOut = posesD;
    *costOut = cost;
    *systemOut = system;
    return hipGetLastError();
}

}  // namespace vmamd
