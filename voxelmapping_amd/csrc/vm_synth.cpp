// Seeded synthetic RGB-D scenes for tests and benchmarks: an inward-facing
// room box with spheres and axis-aligned boxes, rendered analytically (FP64
// ray casting) to camera-z depth, with depth-dependent Gaussian noise,
// random dropout and per-object colours. Random numbers are counter-based
// (splitmix64 of (seed, index)), so every frame and pixel is reproducible on
// any machine and with any thread count.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
// uniform in [0, 1) with 53 random bits
inline double u01(uint64_t seed, uint64_t idx) {
    return (double)(splitmix64(seed ^ splitmix64(idx)) >> 11) * (1.0 / 9007199254740992.0);
}
// standard normal by Box-Muller
inline double gauss(uint64_t seed, uint64_t idx) {
    double a = u01(seed, 2 * idx), b = u01(seed, 2 * idx + 1);
    if (a < 1e-300) a = 1e-300;
    return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
}

struct Sphere { double c[3], r; uint8_t rgb[3]; };
struct Box { double lo[3], hi[3]; uint8_t rgb[3]; };

struct Scene {
    double room[3];
    uint8_t roomRgb[3];
    std::vector<Sphere> spheres;
    std::vector<Box> boxes;
};

void colour(uint64_t seed, uint64_t idx, uint8_t* rgb) {
    for (int i = 0; i < 3; ++i) rgb[i] = (uint8_t)(40 + (int)(u01(seed, idx * 3 + i) * 200.0));
}

// nearest hit along o + t d, t > 0; returns t (inf if none) and the colour
double cast(const Scene& s, const double* o, const double* d, const uint8_t** rgb) {
    double best = INFINITY;
    // room interior: exit of the slab box [0, room]
    {
        double tExit = INFINITY;
        for (int a = 0; a < 3; ++a) {
            if (d[a] > 1e-12) tExit = std::fmin(tExit, (s.room[a] - o[a]) / d[a]);
            else if (d[a] < -1e-12) tExit = std::fmin(tExit, (0.0 - o[a]) / d[a]);
        }
        if (tExit > 0 && tExit < best) { best = tExit; *rgb = s.roomRgb; }
    }
    for (const Sphere& sp : s.spheres) {
        double oc[3] = {o[0] - sp.c[0], o[1] - sp.c[1], o[2] - sp.c[2]};
        double b = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
        double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double c = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - sp.r * sp.r;
        double disc = b * b - dd * c;
        if (disc < 0) continue;
        double q = std::sqrt(disc);
        double t = (-b - q) / dd;
        if (t <= 1e-9) t = (-b + q) / dd;
        if (t > 1e-9 && t < best) { best = t; *rgb = sp.rgb; }
    }
    for (const Box& bx : s.boxes) {
        double t0 = -INFINITY, t1 = INFINITY;
        bool miss = false;
        for (int a = 0; a < 3 && !miss; ++a) {
            if (std::fabs(d[a]) < 1e-12) {
                if (o[a] < bx.lo[a] || o[a] > bx.hi[a]) miss = true;
                continue;
            }
            double ta = (bx.lo[a] - o[a]) / d[a], tb = (bx.hi[a] - o[a]) / d[a];
            if (ta > tb) std::swap(ta, tb);
            t0 = std::fmax(t0, ta);
            t1 = std::fmin(t1, tb);
        }
        if (miss || t0 > t1) continue;
        double t = t0 > 1e-9 ? t0 : t1;
        if (t > 1e-9 && t < best) { best = t; *rgb = bx.rgb; }
    }
    return best;
}

// camera-to-world, column-major: columns are the camera x (right), y (down),
// z (forward) axes and the position
void look_at(const double* eye, const double* target, float* T) {
    double f[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
    double n = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& v : f) v /= n;
    const double up[3] = {0, 0, 1};
    double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    n = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (double& v : r) v /= n;
    double dn[3] = {f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2], f[0] * r[1] - f[1] * r[0]};
    for (int i = 0; i < 3; ++i) {
        T[0 + i] = (float)r[i];
        T[4 + i] = (float)dn[i];
        T[8 + i] = (float)f[i];
        T[12 + i] = (float)eye[i];
    }
    T[3] = T[7] = T[11] = 0.f;
    T[15] = 1.f;
}

// Rodrigues rotation of column-major 3x3 R by axis-angle w (left multiply)
void rotate(float* T, const double* w) {
    double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    double Rw[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (th > 0) {
        double k[3] = {w[0] / th, w[1] / th, w[2] / th};
        double c = std::cos(th), s = std::sin(th), v = 1 - c;
        // row-major Rw
        Rw[0] = c + k[0] * k[0] * v;        Rw[1] = k[0] * k[1] * v - k[2] * s; Rw[2] = k[0] * k[2] * v + k[1] * s;
        Rw[3] = k[1] * k[0] * v + k[2] * s; Rw[4] = c + k[1] * k[1] * v;        Rw[5] = k[1] * k[2] * v - k[0] * s;
        Rw[6] = k[2] * k[0] * v - k[1] * s; Rw[7] = k[2] * k[1] * v + k[0] * s; Rw[8] = c + k[2] * k[2] * v;
    }
    double R[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R[i * 3 + j] = T[j * 4 + i];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double a = 0;
            for (int k = 0; k < 3; ++k) a += Rw[i * 3 + k] * R[k * 3 + j];
            T[j * 4 + i] = (float)a;
        }
}

}  // namespace

extern "C" {

typedef struct vm_synth vm_synth;

// room of size (rx, ry, rz) metres with corner at the origin, n_spheres
// spheres of radius U[0.2, 0.6] and n_boxes boxes of side U[0.2, 0.8],
// uniformly placed inside
vm_synth* vm_synth_create(uint64_t seed, int n_spheres, int n_boxes, double rx, double ry, double rz) {
    Scene* s = new Scene;
    s->room[0] = rx; s->room[1] = ry; s->room[2] = rz;
    colour(seed, 0, s->roomRgb);
    uint64_t k = 1000;
    for (int i = 0; i < n_spheres; ++i) {
        Sphere sp;
        sp.r = 0.2 + 0.4 * u01(seed, k++);
        for (int a = 0; a < 3; ++a) sp.c[a] = sp.r + (s->room[a] - 2 * sp.r) * u01(seed, k++);
        colour(seed, 100 + i, sp.rgb);
        s->spheres.push_back(sp);
    }
    for (int i = 0; i < n_boxes; ++i) {
        Box b;
        for (int a = 0; a < 3; ++a) {
            double side = 0.2 + 0.6 * u01(seed, k++);
            double lo = (s->room[a] - side) * u01(seed, k++);
            b.lo[a] = lo;
            b.hi[a] = lo + side;
        }
        colour(seed, 200 + i, b.rgb);
        s->boxes.push_back(b);
    }
    return (vm_synth*)s;
}

void vm_synth_destroy(vm_synth* s) { delete (Scene*)s; }

// kind 0: smooth circular trajectory around the room centre, looking at a
// point that circles the opposite way; kind 1: seeded random walk with
// translation steps <= max_step metres and rotation steps <= max_rot radians,
// kept inside the room with a 0.5 m margin. poses: n_frames * 16 floats.
int vm_synth_trajectory(vm_synth* sp, int kind, int n_frames, uint64_t seed, double max_step, double max_rot,
                        float* poses) {
    const Scene& s = *(const Scene*)sp;
    const double cx = s.room[0] * 0.5, cy = s.room[1] * 0.5, cz = s.room[2] * 0.5;
    if (kind == 0) {
        const double ph = 6.283185307179586 * u01(seed, 1);
        for (int f = 0; f < n_frames; ++f) {
            double a = ph + 1.2 * (double)f / (double)(n_frames > 1 ? n_frames : 1);
            double eye[3] = {cx + 0.6 * std::cos(a), cy + 0.6 * std::sin(a), cz + 0.1 * std::sin(2 * a)};
            double tgt[3] = {cx - 1.5 * std::cos(a + 0.7), cy - 1.5 * std::sin(a + 0.7), cz - 0.3};
            look_at(eye, tgt, poses + 16 * f);
        }
        return 0;
    }
    double eye[3] = {cx, cy, cz};
    double tgt[3] = {cx + std::cos(6.283185307179586 * u01(seed, 2)), cy + std::sin(6.283185307179586 * u01(seed, 2)),
                     cz - 0.2};
    float T[16];
    look_at(eye, tgt, T);
    uint64_t k = 100;
    for (int f = 0; f < n_frames; ++f) {
        if (f > 0) {
            double dt[3], w[3];
            for (int a = 0; a < 3; ++a) dt[a] = (2 * u01(seed, k++) - 1) * max_step / std::sqrt(3.0);
            for (int a = 0; a < 3; ++a) w[a] = (2 * u01(seed, k++) - 1) * max_rot / std::sqrt(3.0);
            for (int a = 0; a < 3; ++a) {
                double np = T[12 + a] + dt[a];
                if (np < 0.5 || np > s.room[a] - 0.5) np = T[12 + a] - dt[a];
                T[12 + a] = (float)np;
            }
            rotate(T, w);
        }
        std::memcpy(poses + 16 * f, T, sizeof(T));
    }
    return 0;
}

// Render one frame. depth_f32 (metres) and/or depth_u16 (round(z * u16_per_m),
// 0 when z > max_depth or dropped) and/or rgb (3 bytes/pixel) may be NULL.
// Noise: z += N(0, (noise_coef * z^2)^2); dropout: depth 0 with probability p.
int vm_synth_render(vm_synth* sp, const float* pose, int W, int H, const float* K, double noise_coef,
                    double dropout, uint64_t frame_seed, double max_depth, float* depth_f32, uint16_t* depth_u16,
                    double u16_per_m, uint8_t* rgb) {
    const Scene& s = *(const Scene*)sp;
    double R[9], t[3];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) R[i * 3 + j] = pose[j * 4 + i];
        t[i] = pose[12 + i];
    }
    const double fx = K[0], fy = K[1], cx = K[2], cy = K[3];
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            double dc[3] = {(x - cx) / fx, (y - cy) / fy, 1.0};
            double dw[3];
            for (int i = 0; i < 3; ++i) dw[i] = R[i * 3 + 0] * dc[0] + R[i * 3 + 1] * dc[1] + R[i * 3 + 2] * dc[2];
            const uint8_t* col = s.roomRgb;
            double z = cast(s, t, dw, &col);  // dc has unit z, so t is camera z
            const uint64_t pix = (uint64_t)y * W + x;
            if (std::isfinite(z)) {
                double sigma = noise_coef * z * z;
                z += sigma * gauss(frame_seed, pix);
                if (u01(frame_seed ^ 0x5bd1e995ull, pix) < dropout) z = 0;
            } else {
                z = 0;
            }
            if (z < 0) z = 0;
            if (depth_f32) depth_f32[pix] = (float)z;
            if (depth_u16) {
                double q = (z > max_depth) ? 0.0 : std::floor(z * u16_per_m + 0.5);
                depth_u16[pix] = (uint16_t)(q > 65535 ? 0 : q);
            }
            if (rgb) {
                rgb[pix * 3 + 0] = col[0];
                rgb[pix * 3 + 1] = col[1];
                rgb[pix * 3 + 2] = col[2];
            }
        }
    }
    return 0;
}

}  // extern "C"
