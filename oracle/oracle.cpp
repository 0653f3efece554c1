// oracle.cpp -- CPU restatement of SeeSharp's unidirectional PathTracer (fp32,
// reference operation order, no contraction, no fast-math).
//
// TEST INFRASTRUCTURE ONLY: the checker for the GPU path and the CPU baseline
// timed by bench.py. The product never links, loads or calls this file.
//
// Rules this restatement fixes where the reference delegates to packages that
// are not in the reference tree (TinyEmbree 1.2.0 / Embree, SimpleImageIO 1.8.3,
// .NET System.Numerics); the GPU library follows the same rules:
//   * vector sums are evaluated left to right, x*x + y*y + z*z, without FMA;
//   * Vector3.Normalize(v) = v / sqrt(dot(v, v)) per component;
//   * Vector3.Transform(v, M) = ((x*M1 + y*M2) + z*M3) + M4 (row vectors);
//   * ray/triangle: watertight test (Woop, Benthin, Wald, JCGT 2013) with the
//     double-precision fallback on zero edge functions; ties in t go to the
//     lower global triangle id;
//   * hit position / shading normal / uv = u*a1 + v*a2 + (1-u-v)*a0, the order
//     of Mesh.ComputeErrorOffset (SeeSharp/Geometry/Mesh.cs:50-60);
//   * SpawnRay: origin = P + sign(dot(d, Ng)) * ErrorOffset * Ng, tmin = ErrorOffset;
//     IsOccluded: both end points offset the same way towards each other,
//     d = p1 - p0, t in (1e-5, 1 - 1e-5); LeavesScene: SpawnRay to t = inf;
//   * RgbColor.Luminance = 0.2126 r + 0.7152 g + 0.0722 b; image lookups clamp
//     the pixel index into the image;
//   * MathF.ReciprocalSqrtEstimate is replaced by 1 / sqrt.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

constexpr float kPi = 3.14159265358979323846f;

// ---------------------------------------------------------------- vectors --
struct V3 {
    float x, y, z;
};
struct V2 {
    float x, y;
};

inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline bool operator==(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline bool operator!=(V3 a, V3 b) { return !(a == b); }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return a / length(a); }
inline V3 vabs(V3 a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
inline float get(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// RgbColor (SimpleImageIO): component-wise float3.
using Rgb = V3;
inline Rgb rgb(float v) { return {v, v, v}; }
inline float average(Rgb c) { return (c.x + c.y + c.z) / 3.0f; }
inline float luminance(Rgb c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
inline Rgb lerp(float w, Rgb a, Rgb b) { return (1.0f - w) * a + w * b; }
inline bool is_black(Rgb c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }

// C# (int)float on x86: truncation, out-of-range and NaN -> int.MinValue.
inline int to_int(float f) {
    if (!(f > -2147483904.0f && f < 2147483648.0f)) return INT_MIN;
    return (int)f;
}

// ---------------------------------------------------------------- RNG ------
// SeeSharp/Sampling/RNG.cs:90-118
inline uint32_t pcg_hash(uint32_t input) {
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28) + 4)) ^ state) * 277803737u;
    return (word >> 22) ^ word;
}
inline uint32_t fnv_hash(uint32_t hash, uint32_t data) {
    const uint32_t prime = 16777619u;
    hash = (hash * prime) ^ (data & 0xFF);
    hash = (hash * prime) ^ ((data >> 8) & 0xFF);
    hash = (hash * prime) ^ ((data >> 16) & 0xFF);
    hash = (hash * prime) ^ ((data >> 24) & 0xFF);
    return hash;
}
inline uint32_t hash_seed(uint32_t base, uint32_t chain, uint32_t sample) {
    uint32_t h = fnv_hash(2166136261u, pcg_hash(base));
    h = fnv_hash(h, pcg_hash(chain));
    h = fnv_hash(h, pcg_hash(sample));
    return pcg_hash(h);
}

struct Rng {
    uint32_t state;
    // RNG.Next (RNG.cs:90-94)
    uint32_t next() {
        uint32_t word = ((state >> ((state >> 28) + 4)) ^ state) * 277803737u;
        state = state * 747796405u + 2891336453u;
        return (word >> 22) ^ word;
    }
    // RNG.NextFloat (RNG.cs:43-45): u32 -> f32 (round to nearest) / 2^32.
    float next_float() { return (float)next() / 4294967296.0f; }
    V2 next_float2() {
        float a = next_float();
        float b = next_float();
        return {a, b};
    }
    // RNG.NextInt(uint) (RNG.cs:56-70), Lemire's multiply-shift with rejection.
    uint32_t next_int(uint32_t max) {
        uint32_t x = next();
        uint64_t m = (uint64_t)x * max;
        uint32_t l = (uint32_t)m;
        if (l < max) {
            uint32_t t = (0u - max) % max;
            while (l < t) {
                x = next();
                m = (uint64_t)x * max;
                l = (uint32_t)m;
            }
        }
        return (uint32_t)(m >> 32);
    }
};

// ---------------------------------------------------------------- warps ----
// SampleWarp.ToUniformTriangle (SampleWarp.cs:9-14)
inline V2 to_uniform_triangle(V2 p) {
    float s = std::sqrt(p.x);
    return {1.0f - s, p.y * s};
}
// SampleWarp.SphericalToCartesian (SampleWarp.cs:89-96)
inline V3 spherical_to_cartesian(float sintheta, float costheta, float phi) {
    float sinphi = std::sin(phi), cosphi = std::cos(phi);
    return {sintheta * cosphi, sintheta * sinphi, costheta};
}
struct DirSample {
    V3 dir;
    float pdf;
};
// SampleWarp.ToCosHemisphere (SampleWarp.cs:130-137)
inline DirSample to_cos_hemisphere(V2 p) {
    V3 d = spherical_to_cartesian(std::sqrt(1.0f - p.y), std::sqrt(p.y), 2.0f * kPi * p.x);
    return {d, d.z / kPi};
}
// SampleWarp.ToCosineLobe (SampleWarp.cs:171-182)
inline DirSample to_cosine_lobe(float power, V2 p) {
    float phi = kPi * 2.0f * p.x;
    float cos_t = std::pow(p.y, 1.0f / (power + 1.0f));
    float sin_t = std::sqrt(1.0f - cos_t * cos_t);
    V3 d = spherical_to_cartesian(sin_t, cos_t, phi);
    return {d, (power + 1.0f) * std::pow(cos_t, power) * 1.0f / (2.0f * kPi)};
}

// ------------------------------------------------------- piecewise const ---
// PiecewiseConstantPDF (SeeSharp/Sampling/PiecewiseConstant.cs:12-70)
struct Pwc {
    std::vector<float> cdf;
    void build(const float* w, int n) {
        cdf.resize(n);
        float sum = 0.0f;
        for (int i = 0; i < n; ++i) cdf[i] = (sum += w[i]);
        float total = cdf[n - 1];
        for (int i = 0; i < n && total > 0.0f; ++i) cdf[i] /= total;
        cdf[n - 1] = 1.0f;
    }
    // List.BinarySearch + walk back to the first equal element == lower_bound.
    void sample(float u, int* idx_out, float* rel_out) const {
        int idx = (int)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        float lo = idx == 0 ? 0.0f : cdf[idx - 1];
        float delta = cdf[idx] - lo;
        *idx_out = idx;
        *rel_out = (u - lo) / delta;
    }
    float probability(int idx) const { return idx > 0 ? cdf[idx] - cdf[idx - 1] : cdf[idx]; }
};

// RegularGrid2d (SeeSharp/Sampling/RegularGrid2d.cs:20-81), Normalize path.
struct Grid2d {
    int cols = 0, rows = 0;
    std::vector<float> density, marginals;
    Pwc row_dist;
    std::vector<Pwc> col_dist;
    std::vector<char> col_valid;
    void init(int c, int r) {
        cols = c;
        rows = r;
        density.assign((size_t)c * r, 0.0f);
        marginals.assign(r, 0.0f);
    }
    void splat(int col, int row, float v) {
        density[(size_t)row * cols + col] += v;
        marginals[row] += v;
    }
    void normalize() {
        row_dist.build(marginals.data(), rows);
        col_dist.resize(rows);
        col_valid.assign(rows, 0);
        for (int i = 0; i < rows; ++i) {
            if (marginals[i] == 0.0f) continue;
            col_dist[i].build(&density[(size_t)i * cols], cols);
            col_valid[i] = 1;
        }
    }
    V2 sample(V2 p) const {
        int ri, ci;
        float rr, rc;
        row_dist.sample(p.y, &ri, &rr);
        col_dist[ri].sample(p.x, &ci, &rc);
        return {((float)ci + rc) / (float)cols, ((float)ri + rr) / (float)rows};
    }
    float pdf(V2 pos) const {
        int row = std::min(std::max(to_int(pos.y * (float)rows), 0), rows - 1);
        int col = std::min(std::max(to_int(pos.x * (float)cols), 0), cols - 1);
        float p = row_dist.probability(row);
        if (p == 0.0f) return 0.0f;
        p *= col_dist[row].probability(col);
        return p * (float)(rows * cols);
    }
};

// ---------------------------------------------------------------- camera ---
// System.Numerics.Matrix4x4, row-major, row vectors.
struct M4 {
    float m[16];
    float at(int r, int c) const { return m[r * 4 + c]; }
};

// General inverse by cofactors in double, rounded to float. The reference uses
// Matrix4x4.Invert (Camera.cs:38, PerspectiveCamera.cs:48), whose last-ulp
// behaviour is not pinned; the GPU library uses this same rule.
bool invert(const M4& a, M4* out) {
    double m[16], inv[16];
    for (int i = 0; i < 16; ++i) m[i] = a.m[i];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] +
             m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] -
             m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] +
             m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] -
              m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] -
             m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] +
             m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] -
             m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] +
              m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] +
             m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] -
             m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] +
              m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] -
              m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] -
             m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] +
             m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] -
              m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] +
              m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0.0) return false;
    for (int i = 0; i < 16; ++i) out->m[i] = (float)(inv[i] / det);
    return true;
}

// Vector3.Transform(position, M): row vector (x, y, z, 1).
inline V3 transform_point(V3 p, const M4& M) {
    return {p.x * M.at(0, 0) + p.y * M.at(1, 0) + p.z * M.at(2, 0) + M.at(3, 0),
            p.x * M.at(0, 1) + p.y * M.at(1, 1) + p.z * M.at(2, 1) + M.at(3, 1),
            p.x * M.at(0, 2) + p.y * M.at(1, 2) + p.z * M.at(2, 2) + M.at(3, 2)};
}
// Vector4.Transform((d, 0), M).xyz
inline V3 transform_dir(V3 d, const M4& M) {
    return {d.x * M.at(0, 0) + d.y * M.at(1, 0) + d.z * M.at(2, 0),
            d.x * M.at(0, 1) + d.y * M.at(1, 1) + d.z * M.at(2, 1),
            d.x * M.at(0, 2) + d.y * M.at(1, 2) + d.z * M.at(2, 2)};
}

// Matrix4x4.CreatePerspectiveFieldOfView (right-handed), values in fp32.
M4 perspective_fov(float fov, float aspect, float n, float f) {
    M4 r;
    std::memset(r.m, 0, sizeof(r.m));
    float y_scale = 1.0f / std::tan(fov * 0.5f);
    float x_scale = y_scale / aspect;
    r.m[0] = x_scale;
    r.m[5] = y_scale;
    float range = f / (n - f);
    r.m[10] = range;
    r.m[11] = -1.0f;
    r.m[14] = range * n;
    return r;
}

struct Camera {
    M4 world_to_camera, camera_to_world, view_to_camera, camera_to_view;
    int width = 0, height = 0;
    V3 position;
    float image_plane_distance = 0.0f;
    // Camera.WorldToCamera setter (Camera.cs:34-42) + PerspectiveCamera ctor and
    // UpdateResolution (PerspectiveCamera.cs:29-52).
    bool init(const float* w2c, float fov_deg, int w, int h) {
        std::memcpy(world_to_camera.m, w2c, sizeof(float) * 16);
        if (!invert(world_to_camera, &camera_to_world)) return false;
        float fov_rad = fov_deg * kPi / 180.0f;
        width = w;
        height = h;
        float aspect = (float)width / (float)height;
        camera_to_view = perspective_fov(fov_rad, aspect, 0.001f, 1000.0f);
        if (!invert(camera_to_view, &view_to_camera)) return false;
        float tan_half = std::tan(fov_rad * 0.5f);
        image_plane_distance = (float)height / (2.0f * tan_half);
        position = transform_point(v3(0, 0, 0), camera_to_world);
        return true;
    }
    // PerspectiveCamera.SampleResponse (PerspectiveCamera.cs:111-146) without a
    // lens: the pixel that sees `pos` (WorldToFilm, :148-166) and the weight
    // SolidAngleToPixelJacobian(pos) (:177-191) * SurfaceAreaToSolidAngleJacobian
    // (Camera.cs:91-98) of the surface normal. False if the point is behind the
    // camera or outside the frustum.
    bool sample_response(V3 pos, V3 normal, int* px, int* py, float* weight) const {
        V3 local = transform_point(pos, world_to_camera);
        if (local.z > 0.0f) return false;
        // Vector4.Transform(local, cameraToView) with w = 1
        const M4& P = camera_to_view;
        float vx = local.x * P.at(0, 0) + local.y * P.at(1, 0) + local.z * P.at(2, 0) + P.at(3, 0);
        float vy = local.x * P.at(0, 1) + local.y * P.at(1, 1) + local.z * P.at(2, 1) + P.at(3, 1);
        float vw = local.x * P.at(0, 3) + local.y * P.at(1, 3) + local.z * P.at(2, 3) + P.at(3, 3);
        float fx = (vx / vw + 1.0f) / 2.0f * (float)width;
        float fy = (1.0f - vy / vw) / 2.0f * (float)height;
        if (fx < 0.0f || fx > (float)width || fy < 0.0f || fy > (float)height) return false;
        float cosine = local.z / length(local);
        float d = image_plane_distance / cosine;
        float jac = d * d / std::fabs(cosine);
        V3 to_cam = position - pos;
        float dist = length(to_cam);
        float cos_cam = std::fabs(dot(normal, to_cam)) / dist;
        jac *= dist == 0.0f ? 0.0f : cos_cam / (dist * dist);
        *px = std::min(std::max(to_int(fx), 0), width - 1);
        *py = std::min(std::max(to_int(fy), 0), height - 1);
        *weight = jac;
        return true;
    }
    // PerspectiveCamera.GenerateRay (PerspectiveCamera.cs:64-102), lensRadius = 0.
    // `raw` (optional) receives the unnormalised world direction.
    void generate_ray(float fx, float fy, V3* org, V3* dir, V3* raw = nullptr) const {
        V3 view = v3(2.0f * fx / (float)width - 1.0f, 1.0f - 2.0f * fy / (float)height, 0.0f);
        V3 local = transform_point(view, view_to_camera);
        V3 d = transform_dir(local, camera_to_world);
        *org = position;
        *dir = normalize(d);
        if (raw) *raw = d;
    }
    // PerspectiveCamera.SolidAngleToPixelJacobian (PerspectiveCamera.cs:177-191)
    float pixel_jacobian(V3 pos) const {
        V3 local = transform_point(pos, world_to_camera);
        float cosine = local.z / length(local);
        float d = image_plane_distance / cosine;
        return d * d / std::fabs(cosine);
    }
    // Camera.Direction (Camera.cs:16-21)
    V3 direction() const { return transform_dir(v3(0, 0, -1), camera_to_world); }
};

// ----------------------------------------------------------- shading space -
// ShadingSpace.ComputeBasisVectors (ShadingSpace.cs:14-23)
inline void basis(V3 n, V3* t, V3* b) {
    if (std::fabs(n.x) > std::fabs(n.y)) {
        float denom = std::sqrt(n.x * n.x + n.z * n.z);
        *t = v3(-n.z, 0.0f, n.x) / denom;
    } else {
        float denom = std::sqrt(n.y * n.y + n.z * n.z);
        *t = v3(0.0f, n.z, -n.y) / denom;
    }
    *b = cross(n, *t);
}
struct Frame {
    V3 n, t, b;
    // ShadingSpace.WorldToShading (ShadingSpace.cs:40-49)
    V3 to_local(V3 d) const { return {dot(t, d), dot(b, d), dot(n, d)}; }
    // ShadingSpace.ShadingToWorld (ShadingSpace.cs:62-71)
    V3 to_world(V3 d) const { return d.z * n + d.x * t + d.y * b; }
};
inline float cos_theta(V3 d) { return d.z; }
inline float cos2_theta(V3 d) { return d.z * d.z; }
inline float abs_cos_theta(V3 d) { return std::fabs(d.z); }
inline float sin2_theta(V3 d) { return std::max(0.0f, 1.0f - cos2_theta(d)); }
inline float sin_theta(V3 d) { return std::sqrt(sin2_theta(d)); }
inline float tan_theta(V3 d) { return sin_theta(d) / cos_theta(d); }
inline float tan2_theta(V3 d) { return sin2_theta(d) / cos2_theta(d); }
inline float cos_phi(V3 d) {
    float s = sin_theta(d);
    return s == 0.0f ? 1.0f : std::min(std::max(d.x / s, -1.0f), 1.0f);
}
inline float sin_phi(V3 d) {
    float s = sin_theta(d);
    return s == 0.0f ? 0.0f : std::min(std::max(d.y / s, -1.0f), 1.0f);
}
inline float cos2_phi(V3 d) {
    float c = cos_phi(d);
    return c * c;
}
inline float sin2_phi(V3 d) {
    float s = sin_phi(d);
    return s * s;
}
// ShadingSpace.Reflect / Refract / SameHemisphere (ShadingSpace.cs:109-140)
inline V3 reflect(V3 o, V3 n) { return -o + 2.0f * dot(o, n) * n; }
inline bool refract(V3 i, V3 n, float eta, V3* out) {
    float cos_i = dot(n, i);
    float sin2_i = std::max(0.0f, 1.0f - cos_i * cos_i);
    float sin2_t = eta * eta * sin2_i;
    if (sin2_t >= 1.0f) return false;
    float cos_t = std::sqrt(1.0f - sin2_t);
    *out = eta * -i + (eta * cos_i - cos_t) * n;
    return true;
}
inline bool same_hemisphere(V3 a, V3 b) { return a.z * b.z > 0.0f; }

// ------------------------------------------------------------- microfacet --
// TrowbridgeReitzDistribution (SeeSharp/Shading/TrowbridgeReitzDistribution.cs)
struct Ggx {
    float ax, ay;
    float ndf(V3 n) const {  // :22-34
        float tan2 = tan2_theta(n);
        if (std::isinf(tan2)) return 0.0f;
        float cos4 = cos2_theta(n) * cos2_theta(n);
        float e = tan2 * (cos2_phi(n) / (ax * ax) + sin2_phi(n) / (ay * ay));
        return 1.0f / (kPi * ax * ay * cos4 * (1.0f + e) * (1.0f + e));
    }
    float masking_ratio(V3 o) const {  // :104-113
        float abs_tan = std::fabs(tan_theta(o));
        if (std::isinf(abs_tan)) return 0.0f;
        float alpha = std::sqrt(cos2_phi(o) * ax * ax + sin2_phi(o) * ay * ay);
        float a2t2 = alpha * abs_tan * (alpha * abs_tan);
        return (-1.0f + std::sqrt(1.0f + a2t2)) / 2.0f;
    }
    float g1(V3 o) const { return 1.0f / (1.0f + masking_ratio(o)); }
    float g(V3 o, V3 i) const { return 1.0f / (1.0f + masking_ratio(o) + masking_ratio(i)); }
    float pdf(V3 o, V3 n) const {  // :59-63
        if (!same_hemisphere(o, n)) return 0.0f;
        return g1(o) * std::max(0.0f, dot(o, n)) * ndf(n) / std::fabs(o.z);
    }
    V3 sample(V3 o, V2 p) const {  // :70-100, Heitz 2018 VNDF
        bool flip = false;
        if (o.z < 0.0f) {
            flip = true;
            o = -o;
        }
        V3 vh = normalize(v3(ax * o.x, ay * o.y, o.z));
        float lensq = vh.x * vh.x + vh.y * vh.y;
        V3 t1 = lensq > 0.0f ? v3(-vh.y, vh.x, 0.0f) * (1.0f / std::sqrt(lensq)) : v3(1, 0, 0);
        V3 t2 = cross(vh, t1);
        float r = std::sqrt(p.x);
        float phi = 2.0f * kPi * p.y;
        float a1 = r * std::cos(phi);
        float a2 = r * std::sin(phi);
        float s = 0.5f * (1.0f + vh.z);
        a2 = (1.0f - s) * std::sqrt(1.0f - a1 * a1) + s * a2;
        V3 nh = a1 * t1 + a2 * t2 + std::sqrt(std::max(0.0f, 1.0f - a1 * a1 - a2 * a2)) * vh;
        V3 ne = normalize(v3(ax * nh.x, ay * nh.y, std::max(0.0f, nh.z)));
        return flip ? -ne : ne;
    }
};

// Fresnel (SeeSharp/Shading/Fresnel.cs)
inline float fresnel_dielectric(float cos_i, float eta_i, float eta_t) {  // :4-25
    cos_i = std::min(std::max(cos_i, -1.0f), 1.0f);
    bool entering = cos_i > 0.0f;
    if (!entering) {
        std::swap(eta_i, eta_t);
        cos_i = std::fabs(cos_i);
    }
    float sin_i = std::sqrt(std::max(0.0f, 1.0f - cos_i * cos_i));
    float sin_t = eta_i / eta_t * sin_i;
    if (sin_t >= 1.0f) return 1.0f;
    float cos_t = std::sqrt(std::max(0.0f, 1.0f - sin_t * sin_t));
    float rparl = (eta_t * cos_i - eta_i * cos_t) / (eta_t * cos_i + eta_i * cos_t);
    float rperp = (eta_i * cos_i - eta_t * cos_t) / (eta_i * cos_i + eta_t * cos_t);
    return (rparl * rparl + rperp * rperp) / 2.0f;
}
inline float schlick_weight(float c) {  // :35-38
    float m = std::min(std::max(1.0f - c, 0.0f), 1.0f);
    return (m * m) * (m * m) * m;
}
inline Rgb schlick_fresnel(Rgb r0, float c) { return lerp(schlick_weight(c), r0, rgb(1.0f)); }
inline float schlick_r0(float eta) {
    float r = (eta - 1.0f) / (eta + 1.0f);
    return r * r;
}

// ---------------------------------------------------------------- scene ----
struct Texture {
    int w = 0, h = 0, ch = 0, border = 0;
    std::vector<float> texels;
    // ImageTexture.ComputeTexel (ImageTexture.cs:27-44)
    int texel(V2 uv) const {
        int col = to_int(uv.x * (float)w);
        int row = to_int(uv.y * (float)h);
        if (border == 0) {
            row = (row % h + h) % h;
            col = (col % w + w) % w;
        } else {
            row = std::min(std::max(row, 0), h - 1);
            col = std::min(std::max(col, 0), w - 1);
        }
        return row * w + col;
    }
};

struct Material {
    int kind;
    Rgb base_color;
    int base_tex;
    float roughness;
    int rough_tex;
    float metallic, spec_tint, anisotropic, spec_trans, ior;
    bool transmitter;
};

struct Tri {
    V3 p0, p1, p2;
    V3 n0, n1, n2;
    V2 t0, t1, t2;
    V3 face_normal;
    int mesh, prim, material;
    bool has_normals, has_uvs;
};

struct SurfacePoint {
    V3 position, normal, shading_normal;
    V2 bary, uv;
    int tri = -1;  // global triangle id
    float error_offset = 0.0f, distance = 0.0f;
};

struct Emitter {
    int tri;
    Rgb radiance;
    float inv_area;
    int kind;        // sspt_emitter_kind
    float exponent;  // GlossyEmitter lobe exponent
    float norm;      // GlossyEmitter ctor (GlossyEmitter.cs:14): (exponent + 1) / (2 pi)
};

struct EnvMap {
    int w = 0, h = 0;
    std::vector<float> rgb;
    Grid2d grid;
    Rgb pixel(int col, int row) const {
        col = std::min(std::max(col, 0), w - 1);
        row = std::min(std::max(row, 0), h - 1);
        const float* p = &rgb[((size_t)row * w + col) * 3];
        return v3(p[0], p[1], p[2]);
    }
};

struct Bvh {
    struct Node {
        float lo[3], hi[3];
        int left, count;  // count > 0: leaf, left = first index into order
    };
    std::vector<Node> nodes;
    std::vector<int> order;
};

}  // namespace

struct or_scene {
    std::vector<Tri> tris;
    std::vector<Material> materials;
    std::vector<Texture> textures;
    std::vector<Emitter> emitters;
    std::vector<int> tri_emitter;  // global tri -> emitter index or -1
    bool has_env = false;
    EnvMap env;
    V3 scene_center{};         // Background.SceneCenter (Scene.cs:114-141)
    float scene_radius = 0.0f;  // Background.SceneRadius
    Camera camera;
    int width = 0, height = 0;
    float abs_max = 0.0f;  // largest vertex coordinate magnitude
    Bvh bvh;
};


namespace {

// ------------------------------------------------------- surface queries ---
// Mesh.ComputeErrorOffset (Mesh.cs:50-60)
inline float error_offset(const Tri& t, V2 b) {
    V3 diag = vabs(b.x * t.p1) + vabs(b.y * t.p2) + vabs((1.0f - b.x - b.y) * t.p0);
    return length(diag) * 32.0f * 1.19209e-07f;
}

SurfacePoint make_point(const or_scene& s, int tri, V2 b, float dist) {
    const Tri& t = s.tris[tri];
    SurfacePoint p;
    p.tri = tri;
    p.bary = b;
    p.distance = dist;
    float w0 = 1.0f - b.x - b.y;
    p.position = b.x * t.p1 + b.y * t.p2 + w0 * t.p0;
    p.normal = t.face_normal;
    p.shading_normal = t.has_normals ? normalize(b.x * t.n1 + b.y * t.n2 + w0 * t.n0) : t.face_normal;
    if (t.has_uvs) {
        p.uv = {b.x * t.t1.x + b.y * t.t2.x + w0 * t.t0.x, b.x * t.t1.y + b.y * t.t2.y + w0 * t.t0.y};
    } else {
        p.uv = {0.0f, 0.0f};
    }
    p.error_offset = error_offset(t, b);
    return p;
}

using TextureSet = std::vector<Texture>;

inline Rgb lookup_rgb(const TextureSet& texs, int tex, Rgb constant, V2 uv) {
    if (tex < 0) return constant;
    const Texture& t = texs[tex];
    const float* p = &t.texels[(size_t)t.texel(uv) * 3];
    return v3(p[0], p[1], p[2]);
}
inline float lookup_mono(const TextureSet& texs, int tex, float constant, V2 uv) {
    if (tex < 0) return constant;
    const Texture& t = texs[tex];
    return t.texels[t.texel(uv)];
}

// ------------------------------------------------------------ intersection -
struct Ray {
    V3 o, d;
    float tmin, tmax;
};

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013).
// Returns t and the barycentrics (u, v) of vertices 1 and 2.
bool intersect_tri(const Ray& r, int kx, int ky, int kz, float sx, float sy, float sz, V3 p0, V3 p1,
                   V3 p2, float* t_out, float* u_out, float* v_out) {
    V3 a = p0 - r.o, b = p1 - r.o, c = p2 - r.o;
    float az_ = get(a, kz), bz_ = get(b, kz), cz_ = get(c, kz);
    float ax = get(a, kx) - sx * az_;
    float ay = get(a, ky) - sy * az_;
    float bx = get(b, kx) - sx * bz_;
    float by = get(b, ky) - sy * bz_;
    float cx = get(c, kx) - sx * cz_;
    float cy = get(c, ky) - sy * cz_;
    float U = cx * by - cy * bx;
    float V = ax * cy - ay * cx;
    float W = bx * ay - by * ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double cxby = (double)cx * (double)by, cybx = (double)cy * (double)bx;
        U = (float)(cxby - cybx);
        double axcy = (double)ax * (double)cy, aycx = (double)ay * (double)cx;
        V = (float)(axcy - aycx);
        double bxay = (double)bx * (double)ay, byax = (double)by * (double)ax;
        W = (float)(bxay - byax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float az = sz * az_, bz = sz * bz_, cz = sz * cz_;
    float T = U * az + V * bz + W * cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= r.tmin && t <= r.tmax)) return false;
    *t_out = t;
    *u_out = V * rcp;
    *v_out = W * rcp;
    return true;
}

struct RayPre {
    int kx, ky, kz;
    float sx, sy, sz;
    V3 inv;
};
inline RayPre precompute(const Ray& r) {
    RayPre p;
    float ax = std::fabs(r.d.x), ay = std::fabs(r.d.y), az = std::fabs(r.d.z);
    p.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    p.kx = p.kz + 1 == 3 ? 0 : p.kz + 1;
    p.ky = p.kx + 1 == 3 ? 0 : p.kx + 1;
    if (get(r.d, p.kz) < 0.0f) std::swap(p.kx, p.ky);
    float dz = get(r.d, p.kz);
    p.sx = get(r.d, p.kx) / dz;
    p.sy = get(r.d, p.ky) / dz;
    p.sz = 1.0f / dz;
    p.inv = v3(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    return p;
}

struct HitRec {
    float t;
    float u, v;
    int tri;
};

inline void consider(const or_scene& s, const Ray& r, const RayPre& p, int tri, HitRec* best,
                     float* tmax) {
    const Tri& T = s.tris[tri];
    float t, u, v;
    Ray rr = r;
    rr.tmax = *tmax;
    if (!intersect_tri(rr, p.kx, p.ky, p.kz, p.sx, p.sy, p.sz, T.p0, T.p1, T.p2, &t, &u, &v)) return;
    if (t < best->t || (t == best->t && tri < best->tri)) {
        best->t = t;
        best->u = u;
        best->v = v;
        best->tri = tri;
        *tmax = t;
    }
}

// Robust slab test (Ize, JCGT 2013): the far distance is widened by 2 gamma(3).
// [t0, t1] is the query interval, already widened by the ray slack.
inline bool box_hit(const Bvh::Node& n, const Ray& r, const RayPre& p, float t0, float t1, float* tnear) {
    for (int k = 0; k < 3; ++k) {
        float o = get(r.o, k), inv = get(p.inv, k);
        float a = (n.lo[k] - o) * inv, b = (n.hi[k] - o) * inv;
        float lo = std::fmin(a, b), hi = std::fmax(a, b);
        hi *= 1.0000004f;
        t0 = std::fmax(t0, lo);
        t1 = std::fmin(t1, hi);
    }
    *tnear = t0;
    return t0 <= t1;
}

// The watertight t carries an absolute error of a few ulps of the coordinate
// magnitudes divided by |d|; box culling uses [tmin - slack, tmax + slack] with
// slack = 2^-16 * max(|o|, scene extent) / |d| so that it never changes which
// triangle is closest.
inline float ray_slack(const or_scene& s, const Ray& r) {
    float m = std::max(s.abs_max, std::max(std::max(std::fabs(r.o.x), std::fabs(r.o.y)), std::fabs(r.o.z)));
    return m * 1.52587890625e-05f / length(r.d);
}

HitRec trace_closest(const or_scene& s, const Ray& r, bool brute) {
    HitRec best{INFINITY, 0.0f, 0.0f, -1};
    RayPre p = precompute(r);
    float tmax = r.tmax;
    if (brute || s.bvh.nodes.empty()) {
        for (int i = 0; i < (int)s.tris.size(); ++i) consider(s, r, p, i, &best, &tmax);
        return best;
    }
    const float slack = ray_slack(s, r);
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Bvh::Node& n = s.bvh.nodes[stack[--sp]];
        float tn;
        if (!box_hit(n, r, p, r.tmin - slack, tmax + slack, &tn)) continue;
        if (n.count > 0) {
            for (int i = 0; i < n.count; ++i) consider(s, r, p, s.bvh.order[n.left + i], &best, &tmax);
        } else {
            float ta, tb;
            bool ha = box_hit(s.bvh.nodes[n.left], r, p, r.tmin - slack, tmax + slack, &ta);
            bool hb = box_hit(s.bvh.nodes[n.left + 1], r, p, r.tmin - slack, tmax + slack, &tb);
            if (ha && hb) {
                if (ta <= tb) {
                    stack[sp++] = n.left + 1;
                    stack[sp++] = n.left;
                } else {
                    stack[sp++] = n.left;
                    stack[sp++] = n.left + 1;
                }
            } else if (ha) {
                stack[sp++] = n.left;
            } else if (hb) {
                stack[sp++] = n.left + 1;
            }
        }
    }
    return best;
}

bool trace_any(const or_scene& s, const Ray& r, bool brute) {
    RayPre p = precompute(r);
    const float slack = ray_slack(s, r);
    float t, u, v;
    if (brute || s.bvh.nodes.empty()) {
        for (const Tri& T : s.tris)
            if (intersect_tri(r, p.kx, p.ky, p.kz, p.sx, p.sy, p.sz, T.p0, T.p1, T.p2, &t, &u, &v))
                return true;
        return false;
    }
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Bvh::Node& n = s.bvh.nodes[stack[--sp]];
        float tn;
        if (!box_hit(n, r, p, r.tmin - slack, r.tmax + slack, &tn)) continue;
        if (n.count > 0) {
            for (int i = 0; i < n.count; ++i) {
                const Tri& T = s.tris[s.bvh.order[n.left + i]];
                if (intersect_tri(r, p.kx, p.ky, p.kz, p.sx, p.sy, p.sz, T.p0, T.p1, T.p2, &t, &u, &v))
                    return true;
            }
        } else {
            stack[sp++] = n.left;
            stack[sp++] = n.left + 1;
        }
    }
    return false;
}

// Binned-SAH BVH over triangle centroids, 16 bins, leaves of <= 4 triangles.
void build_bvh(or_scene& s) {
    int n = (int)s.tris.size();
    Bvh& b = s.bvh;
    b.nodes.clear();
    b.order.resize(n);
    if (n == 0) return;
    std::vector<float> lo((size_t)n * 3), hi((size_t)n * 3), ce((size_t)n * 3);
    for (int i = 0; i < n; ++i) {
        b.order[i] = i;
        const Tri& t = s.tris[i];
        for (int k = 0; k < 3; ++k) {
            float a = get(t.p0, k), c = get(t.p1, k), d = get(t.p2, k);
            lo[i * 3 + k] = std::min(a, std::min(c, d));
            hi[i * 3 + k] = std::max(a, std::max(c, d));
            ce[i * 3 + k] = 0.5f * (lo[i * 3 + k] + hi[i * 3 + k]);
        }
    }
    struct Task {
        int node, begin, end;
    };
    b.nodes.reserve(2 * (size_t)n);
    b.nodes.push_back({});
    std::vector<Task> stack{{0, 0, n}};
    auto area = [](const float* l, const float* h) {
        float dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
        if (dx < 0) return 0.0f;
        return dx * dy + dy * dz + dz * dx;
    };
    while (!stack.empty()) {
        Task tk = stack.back();
        stack.pop_back();
        float nl[3] = {INFINITY, INFINITY, INFINITY}, nh[3] = {-INFINITY, -INFINITY, -INFINITY};
        float cl[3] = {INFINITY, INFINITY, INFINITY}, chh[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = tk.begin; i < tk.end; ++i) {
            int t = b.order[i];
            for (int k = 0; k < 3; ++k) {
                nl[k] = std::min(nl[k], lo[t * 3 + k]);
                nh[k] = std::max(nh[k], hi[t * 3 + k]);
                cl[k] = std::min(cl[k], ce[t * 3 + k]);
                chh[k] = std::max(chh[k], ce[t * 3 + k]);
            }
        }
        Bvh::Node& node = b.nodes[tk.node];
        for (int k = 0; k < 3; ++k) {
            node.lo[k] = nl[k];
            node.hi[k] = nh[k];
        }
        int count = tk.end - tk.begin;
        if (count <= 4) {
            node.left = tk.begin;
            node.count = count;
            continue;
        }
        const int B = 16;
        float best_cost = INFINITY;
        int best_axis = -1, best_split = -1;
        for (int k = 0; k < 3; ++k) {
            float extent = chh[k] - cl[k];
            if (extent <= 0.0f) continue;
            int cnt[B] = {0};
            float bl[B][3], bh[B][3];
            for (int j = 0; j < B; ++j)
                for (int q = 0; q < 3; ++q) bl[j][q] = INFINITY, bh[j][q] = -INFINITY;
            for (int i = tk.begin; i < tk.end; ++i) {
                int t = b.order[i];
                int bin = std::min(B - 1, (int)((ce[t * 3 + k] - cl[k]) / extent * B));
                cnt[bin]++;
                for (int q = 0; q < 3; ++q) {
                    bl[bin][q] = std::min(bl[bin][q], lo[t * 3 + q]);
                    bh[bin][q] = std::max(bh[bin][q], hi[t * 3 + q]);
                }
            }
            float ra[B], rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
            int rc[B], rcount = 0;
            for (int j = B - 1; j > 0; --j) {
                rcount += cnt[j];
                for (int q = 0; q < 3; ++q) rl[q] = std::min(rl[q], bl[j][q]), rh[q] = std::max(rh[q], bh[j][q]);
                ra[j] = area(rl, rh);
                rc[j] = rcount;
            }
            float ll[3] = {INFINITY, INFINITY, INFINITY}, lh[3] = {-INFINITY, -INFINITY, -INFINITY};
            int lcount = 0;
            for (int j = 0; j < B - 1; ++j) {
                lcount += cnt[j];
                for (int q = 0; q < 3; ++q) ll[q] = std::min(ll[q], bl[j][q]), lh[q] = std::max(lh[q], bh[j][q]);
                if (lcount == 0 || rc[j + 1] == 0) continue;
                float cost = area(ll, lh) * lcount + ra[j + 1] * rc[j + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = k;
                    best_split = j;
                }
            }
        }
        int mid;
        if (best_axis < 0) {
            mid = tk.begin + count / 2;
        } else {
            int k = best_axis;
            float extent = chh[k] - cl[k];
            auto it = std::partition(b.order.begin() + tk.begin, b.order.begin() + tk.end, [&](int t) {
                int bin = std::min(B - 1, (int)((ce[t * 3 + k] - cl[k]) / extent * B));
                return bin <= best_split;
            });
            mid = (int)(it - b.order.begin());
            if (mid == tk.begin || mid == tk.end) mid = tk.begin + count / 2;
        }
        int left = (int)b.nodes.size();
        b.nodes.push_back({});
        b.nodes.push_back({});
        b.nodes[tk.node].left = left;
        b.nodes[tk.node].count = 0;
        stack.push_back({left + 1, mid, tk.end});
        stack.push_back({left, tk.begin, mid});
    }
}

// TinyEmbree Raytracer.SpawnRay
inline Ray spawn_ray(const SurfacePoint& from, V3 dir) {
    float sign = dot(dir, from.normal) < 0.0f ? -1.0f : 1.0f;
    Ray r;
    r.o = from.position + sign * from.error_offset * from.normal;
    r.d = dir;
    r.tmin = from.error_offset;
    r.tmax = INFINITY;
    return r;
}
inline V3 offset_point(const SurfacePoint& p, V3 dir) {
    float sign = dot(dir, p.normal) < 0.0f ? -1.0f : 1.0f;
    return p.position + sign * p.error_offset * p.normal;
}
// TinyEmbree Raytracer.IsOccluded(from, to)
inline Ray shadow_ray(const SurfacePoint& from, const SurfacePoint& to) {
    V3 d = to.position - from.position;
    V3 p0 = offset_point(from, d);
    V3 p1 = offset_point(to, -d);
    Ray r;
    r.o = p0;
    r.d = p1 - p0;
    r.tmin = 1e-5f;
    r.tmax = 1.0f - 1e-5f;
    return r;
}

// ----------------------------------------------------------- environment ---
// EnvironmentMap (SeeSharp/Shading/Background/EnvironmentMap.cs)
inline V2 world_to_spherical(V3 dir) {  // :158-166
    dir = normalize(dir);
    V2 sp{std::atan2(dir.z, dir.x), std::atan2(std::sqrt(dir.x * dir.x + dir.z * dir.z), dir.y)};
    if (sp.x < 0.0f) sp.x += kPi * 2.0f;
    return sp;
}
inline V3 spherical_to_world(V2 s) {  // :168-175
    float sin_t = std::sin(s.y);
    return v3(sin_t * std::cos(s.x), std::cos(s.y), sin_t * std::sin(s.x));
}
inline V2 spherical_to_pixel(V2 s) { return {s.x / (2.0f * kPi), s.y / kPi}; }
inline V2 pixel_to_spherical(V2 p) { return {p.x * 2.0f * kPi, p.y * kPi}; }

Rgb env_radiance(const EnvMap& e, V3 dir) {  // :19-26
    V2 px = spherical_to_pixel(world_to_spherical(dir));
    float col = px.x * (float)e.w, row = px.y * (float)e.h;
    return e.pixel(to_int(col), to_int(row));
}
float env_pdf(const EnvMap& e, V3 dir) {  // :81-85
    V2 sph = world_to_spherical(dir);
    V2 px = spherical_to_pixel(sph);
    return e.grid.pdf(px) / (std::sin(sph.y) * kPi * kPi * 2.0f);
}
struct EnvSample {
    V3 dir;
    float pdf;
    Rgb weight;
};
EnvSample env_sample(const EnvMap& e, V2 primary) {  // :40-73
    V2 pp = e.grid.sample(primary);
    float pdf = e.grid.pdf(pp);
    V2 sph = pixel_to_spherical(pp);
    V3 dir = spherical_to_world(sph);
    float jac = std::sin(sph.y) * kPi * kPi * 2.0f;
    if (jac == 0.0f) return {dir, 0.0f, rgb(0.0f)};
    pdf /= jac;
    Rgb w = e.pixel(to_int(pp.x * (float)e.w), to_int(pp.y * (float)e.h)) / pdf;
    return {dir, pdf, w};
}

// SampleWarp.ToConcentricDisc (SampleWarp.cs:219-247): Shirley and Chiu's
// concentric map of the unit square onto the unit disc.
inline V2 to_concentric_disc(V2 p) {
    float phi, r;
    const float a = 2.0f * p.x - 1.0f, b = 2.0f * p.y - 1.0f;
    if (a > -b) {
        if (a > b) {
            r = a;
            phi = (kPi * 0.25f) * (b / a);
        } else {
            r = b;
            phi = (kPi * 0.25f) * (2.0f - (a / b));
        }
    } else if (a < b) {
        r = -a;
        phi = (kPi * 0.25f) * (4.0f + (b / a));
    } else {
        r = -b;
        phi = b != 0.0f ? (kPi * 0.25f) * (6.0f - (a / b)) : 0.0f;
    }
    return {r * std::cos(phi), r * std::sin(phi)};
}

// ------------------------------------------------------------- materials ---
struct Shader {
    const or_scene* s;
    const TextureSet* tex;  // the texels the material reads (the scene's, or a perturbed copy)
    const Material* m;
    SurfacePoint p;
    Frame f;
    V3 out_local, out_world;
    bool light = false;  // ShadingContext.IsOnLightSubpath
};

inline Shader make_shader(const or_scene& s, const SurfacePoint& p, V3 out_dir) {
    Shader sh;
    sh.s = &s;
    sh.tex = &s.textures;
    sh.m = &s.materials[s.tris[p.tri].material];
    sh.p = p;
    sh.f.n = p.shading_normal;
    basis(sh.f.n, &sh.f.t, &sh.f.b);
    sh.out_local = sh.f.to_local(out_dir);
    sh.out_world = out_dir;
    return sh;
}

// Material.ShouldReflect (Material.cs:127-134)
inline bool should_reflect(const SurfacePoint& p, V3 out_dir, V3 in_dir) {
    float a = dot(out_dir, p.normal), b = dot(in_dir, p.normal);
    return a * b >= 0.0f;
}

struct GenericLocal {  // GenericMaterial.LocalParams (GenericMaterial.cs:340-381)
    Rgb color_tint, spec_tint;
    float roughness, ax, ay;
    Rgb diffuse_refl, retro_refl, spec_trans, spec_r0;
};
GenericLocal generic_local(const Shader& sh) {
    const Material& m = *sh.m;
    GenericLocal r;
    Rgb base = lookup_rgb(*sh.tex, m.base_tex, m.base_color, sh.p.uv);
    float lum = luminance(base);
    r.color_tint = lum > 0.0f ? base / lum : rgb(1.0f);
    r.spec_tint = lerp(m.spec_tint, rgb(1.0f), r.color_tint);
    r.roughness = lookup_mono(*sh.tex, m.rough_tex, m.roughness, sh.p.uv);
    float aspect = std::sqrt(1.0f - m.anisotropic * 0.9f);
    r.ax = std::max(0.001f, r.roughness * r.roughness / aspect);
    r.ay = std::max(0.001f, r.roughness * r.roughness * aspect);
    r.spec_r0 = lerp(m.metallic, schlick_r0(m.ior) * r.spec_tint, base);
    float dw = (1.0f - m.metallic) * (1.0f - m.spec_trans);
    r.diffuse_refl = base * dw;
    r.spec_trans = m.spec_trans * base;
    r.retro_refl = base * dw;
    return r;
}

// GenericMaterial.ComputeValueAndPdf (GenericMaterial.cs:152-338): the value, the
// forward pdf and, if `pdf_rev_out` is given, the reverse pdf.
void generic_value_pdf(const Shader& sh, V3 in, const GenericLocal& lp, Rgb* value, float* pdf_out,
                       float* pdf_rev_out = nullptr) {
    const Material& m = *sh.m;
    Ggx ggx{lp.ax, lp.ay};
    V3 out = sh.out_local;
    float cos_o = std::fabs(out.z), cos_i = std::fabs(in.z);
    bool same_geo = should_reflect(sh.p, sh.out_world, sh.f.to_world(in));
    Rgb val = rgb(0.0f);
    float pdf0 = 0.0f, pdf1 = 0.0f, pdf2 = 0.0f;
    float rev0 = 0.0f, rev1 = 0.0f, rev2 = 0.0f;  // pdfsRev

    float f_out = schlick_weight(cos_o), f_in = schlick_weight(cos_i);
    if (same_hemisphere(out, in)) {
        if (same_geo) {
            Rgb diffuse = lp.diffuse_refl / kPi * (1.0f - f_out * 0.5f) * (1.0f - f_in * 0.5f);
            val = val + diffuse;
        }
        pdf0 = cos_i / kPi;
        rev0 = cos_o / kPi;
    }

    V3 h = out + in;
    if (h != v3(0, 0, 0)) {
        h = normalize(h);
        float c_in = dot(in, h);
        float rr = 2.0f * lp.roughness * c_in * c_in;
        if (same_hemisphere(out, in) && same_geo) {
            Rgb retro = lp.retro_refl / kPi * rr * (f_out + f_in + f_out * f_in * (rr - 1.0f));
            val = val + retro;
        }
        if (cos_i != 0.0f && cos_o != 0.0f) {
            float c_out = dot(out, h);
            if (same_hemisphere(out, in) && same_geo && c_in * c_out > 0.0f) {
                float c_tir = dot(in, h.z < 0.0f ? -h : h);
                Rgb diel = rgb(fresnel_dielectric(c_tir, 1.0f, m.ior));
                Rgb schl = schlick_fresnel(lp.spec_r0, c_tir);
                Rgb fres = lerp(m.metallic, diel, schl);
                Rgb refl = lp.spec_tint * ggx.ndf(h) * ggx.g(out, in) * fres / (4.0f * cos_i * cos_o);
                val = val + refl;
            }
            float jf = std::fabs(4.0f * c_out), jr = std::fabs(4.0f * c_in);
            if (jf != 0.0f && jr != 0.0f) {
                pdf1 = ggx.pdf(out, h) / jf;
                rev1 = ggx.pdf(in, h) / jr;
            }
        }
    }

    float eta = out.z > 0.0f ? m.ior : (1.0f / m.ior);
    V3 ht = out + in * eta;
    if (cos_o != 0.0f && cos_i != 0.0f && ht != v3(0, 0, 0) && m.spec_trans > 0.0f) {
        ht = normalize(ht);
        ht = ht.z < 0.0f ? -ht : ht;
        float c_out = dot(out, ht), c_in = dot(in, ht);
        float sqrt_denom = c_out + eta * c_in;
        if (!same_hemisphere(out, in) && c_out * c_in < 0.0f) {
            Rgb F = rgb(fresnel_dielectric(c_out, 1.0f, m.ior));
            float factor = sh.light ? (1.0f / eta) : 1.0f;  // GenericMaterial.cs:255
            V3 wh = !same_hemisphere(out, ht) ? -ht : ht;
            float num = ggx.ndf(wh) * ggx.g(out, in);
            num *= eta * eta * std::max(0.0f, dot(in, -wh)) * std::max(0.0f, dot(out, wh));
            num *= factor * factor;
            float denom = in.z * out.z * sqrt_denom * sqrt_denom;
            Rgb tr = (rgb(1.0f) - F) * lp.spec_trans * std::fabs(num / denom);
            val = val + tr;
        }
        if (sqrt_denom != 0.0f) {
            V3 wh = !same_hemisphere(out, ht) ? -ht : ht;
            float jac = eta * eta * std::max(0.0f, dot(in, -wh)) / (sqrt_denom * sqrt_denom);
            pdf2 += ggx.pdf(out, wh) * jac;
        }
        // the reverse transmission pdf, from the half vector of the swapped pair (:277-295)
        float eta_in = in.z > 0.0f ? m.ior : (1.0f / m.ior);
        V3 hin = normalize(out * eta_in + in);
        hin = hin.z < 0.0f ? -hin : hin;
        if (hin != v3(0, 0, 0)) {
            V3 wh = !same_hemisphere(in, hin) ? -hin : hin;
            float sqrt_denom_in = dot(hin, in) + eta_in * dot(out, hin);
            if (sqrt_denom_in != 0.0f) {
                float jac = eta_in * eta_in * std::max(0.0f, dot(out, -wh)) / (sqrt_denom_in * sqrt_denom_in);
                rev2 += ggx.pdf(in, wh) * jac;
            }
        }
    }

    float schl = average(schlick_fresnel(lp.spec_r0, std::fabs(out.z)));
    float bias = std::max(1.0f - schl, 0.75f);
    float dw = bias * (1.0f - m.metallic) * (1.0f - m.spec_trans);
    float fr = fresnel_dielectric(dot(out, h.z < 0.0f ? -h : h), 1.0f, m.ior);
    float sel_refl = (1.0f - dw) * (1.0f - (1.0f - fr) * m.spec_trans);
    float ft = fresnel_dielectric(dot(out, ht), 1.0f, m.ior);
    float sel_trans = (1.0f - dw) * (1.0f - ft) * m.spec_trans;
    *value = val;
    *pdf_out = pdf0 * dw + pdf1 * sel_refl + pdf2 * sel_trans;
    // selectTransmitIn (:311) repeats selectTransmit: both use fresnelT of the out direction
    if (pdf_rev_out) *pdf_rev_out = rev0 * dw + rev1 * sel_refl + rev2 * sel_trans;
}

struct BsdfSample {
    V3 dir;
    float pdf;
    Rgb weight;
    float pdf_rev = 0.0f;  // BsdfSample.PdfReverse
};

// GenericMaterial.Sample (GenericMaterial.cs:92-150)
BsdfSample generic_sample(const Shader& sh, float pc, V2 pd) {
    const Material& m = *sh.m;
    V3 out = sh.out_local;
    if (out.z == 0.0f) return {v3(0, 0, 0), 0.0f, rgb(0.0f)};
    GenericLocal lp = generic_local(sh);
    Ggx ggx{lp.ax, lp.ay};
    float eta = out.z > 0.0f ? (1.0f / m.ior) : m.ior;
    float schl = average(schlick_fresnel(lp.spec_r0, std::fabs(out.z)));
    float bias = std::max(1.0f - schl, 0.75f);
    float dw = bias * (1.0f - m.metallic) * (1.0f - m.spec_trans);
    V3 in;
    if (pc <= dw) {
        in = to_cos_hemisphere(pd).dir;
        if (out.z < 0.0f) in.z *= -1.0f;
    } else {
        V3 h = ggx.sample(out, pd);
        float c_out = dot(out, h.z < 0.0f ? -h : h);
        float fres = fresnel_dielectric(c_out, 1.0f, m.ior);
        float sel_t = (1.0f - fres) * m.spec_trans;
        if ((pc - dw) / (1.0f - dw) < sel_t) {
            if (dot(out, h) < 0.0f) return {v3(0, 0, 0), 0.0f, rgb(0.0f)};
            V3 r;
            if (!refract(out, h, eta, &r)) return {v3(0, 0, 0), 0.0f, rgb(0.0f)};
            in = r;
        } else {
            in = reflect(out, h);
        }
    }
    Rgb val;
    float pdf, pdf_rev;
    generic_value_pdf(sh, in, lp, &val, &pdf, &pdf_rev);
    BsdfSample bs;
    bs.pdf = pdf;
    bs.pdf_rev = pdf_rev;
    bs.weight = val * std::fabs(in.z) / pdf;
    bs.dir = sh.f.to_world(in);
    return bs;
}

// DiffuseMaterial.Evaluate (DiffuseMaterial.cs:127-139)
Rgb diffuse_eval(const Shader& sh, V3 in_world) {
    bool refl = should_reflect(sh.p, sh.out_world, in_world);
    V3 in = sh.f.to_local(in_world);
    Rgb base = lookup_rgb(*sh.tex, sh.m->base_tex, sh.m->base_color, sh.p.uv);
    if (sh.m->transmitter && !refl) {
        if (same_hemisphere(sh.out_local, in)) return rgb(0.0f);
        return base / kPi;
    } else if (refl) {
        if (!same_hemisphere(sh.out_local, in)) return rgb(0.0f);
        return base / kPi;
    }
    return rgb(0.0f);
}
// DiffuseMaterial.Pdf (DiffuseMaterial.cs:167-188) with DiffuseBsdf.Pdf /
// DiffuseTransmission.Pdf (:63-71,95-103): the forward pdf and, if `rev` is
// given, the reverse pdf.
float diffuse_pdf(const Shader& sh, V3 in_world, float* rev = nullptr) {
    V3 in = sh.f.to_local(in_world);
    bool same = same_hemisphere(sh.out_local, in);
    float refl = same ? abs_cos_theta(in) / kPi : 0.0f;
    float refl_rev = same ? abs_cos_theta(sh.out_local) / kPi : 0.0f;
    if (sh.m->transmitter) {
        float trans = !same ? abs_cos_theta(in) / kPi : 0.0f;
        float trans_rev = !same ? abs_cos_theta(sh.out_local) / kPi : 0.0f;
        if (rev) *rev = (refl_rev + trans_rev) * 0.5f;
        return (refl + trans) * 0.5f;
    }
    if (rev) *rev = refl_rev;
    return refl;
}

// SurfaceShader.EvaluateWithCosine (SurfaceShader.cs, Material.cs:41-44, GenericMaterial.cs:72-77)
Rgb eval_cos(const Shader& sh, V3 in_world) {
    if (sh.m->kind == SSPT_MATERIAL_GENERIC) {
        V3 in = sh.f.to_local(in_world);
        Rgb v;
        float pdf;
        generic_value_pdf(sh, in, generic_local(sh), &v, &pdf);
        return v * std::fabs(in.z);
    }
    Rgb v = diffuse_eval(sh, in_world);
    return v * abs_cos_theta(sh.f.to_local(in_world));
}
Rgb eval_nocos(const Shader& sh, V3 in_world) {
    if (sh.m->kind == SSPT_MATERIAL_GENERIC) {
        V3 in = sh.f.to_local(in_world);
        Rgb v;
        float pdf;
        generic_value_pdf(sh, in, generic_local(sh), &v, &pdf);
        return v;
    }
    return diffuse_eval(sh, in_world);
}
// SurfaceShader.Pdf(inDir).Pdf
float pdf_fwd(const Shader& sh, V3 in_world) {
    if (sh.m->kind == SSPT_MATERIAL_GENERIC) {
        V3 in = sh.f.to_local(in_world);
        Rgb v;
        float pdf;
        generic_value_pdf(sh, in, generic_local(sh), &v, &pdf);
        return pdf;
    }
    return diffuse_pdf(sh, in_world);
}
// SurfaceShader.Sample (DiffuseMaterial.cs:141-165, GenericMaterial.cs:92-150)
BsdfSample sample_bsdf(const Shader& sh, float pc, V2 pd) {
    if (sh.m->kind == SSPT_MATERIAL_GENERIC) return generic_sample(sh, pc, pd);
    DirSample loc = to_cos_hemisphere(pd);
    V3 d = loc.dir;
    if (sh.m->transmitter && pc < 0.5f) {
        if (cos_theta(sh.out_local) > 0.0f) d.z *= -1.0f;
    } else {
        if (cos_theta(sh.out_local) < 0.0f) d.z *= -1.0f;
    }
    // ShadingToWorld(context.Normal, sample) recomputes the same basis.
    V3 dir = sh.f.to_world(d);
    Rgb value = eval_cos(sh, dir);
    float rev;
    float pdf = diffuse_pdf(sh, dir, &rev);
    if (pdf == 0.0f) return {v3(0, 0, 0), 0.0f, rgb(0.0f)};
    return {dir, pdf, value / pdf, rev};
}

// SurfaceShader.Pdf(inDir) (SurfaceShader.cs:64-67): forward and reverse solid-angle pdfs.
void pdf_pair(const Shader& sh, V3 in_world, float* fwd, float* rev) {
    if (sh.m->kind == SSPT_MATERIAL_GENERIC) {
        V3 in = sh.f.to_local(in_world);
        Rgb v;
        generic_value_pdf(sh, in, generic_local(sh), &v, fwd, rev);
        return;
    }
    *fwd = diffuse_pdf(sh, in_world, rev);
}

// ----------------------------------------------------------- path tracer ---
struct Counters {
    uint64_t rays = 0, shadow = 0, occluded = 0, hits = 0;
};

// SampleWarp.SurfaceAreaToSolidAngle (SampleWarp.cs:290-295)
inline float area_to_solid_angle(const SurfacePoint& from, const SurfacePoint& to) {
    V3 d = to.position - from.position;
    float d2 = dot(d, d);
    return std::fabs(dot(to.normal, -d)) / (d2 * std::sqrt(d2));
}

// DiffuseEmitter.EmittedRadiance (DiffuseEmitter.cs:33-37) and
// GlossyEmitter.EmittedRadiance (GlossyEmitter.cs:36-40)
inline Rgb emitted(const Emitter& e, const SurfacePoint& p, V3 dir) {
    if (e.kind == SSPT_EMITTER_GLOSSY) {
        float cosine = dot(p.shading_normal, dir) / std::sqrt(dot(dir, dir));
        if (cosine < 0.0f) return rgb(0.0f);
        return e.radiance * std::pow(cosine, e.exponent) * e.norm;
    }
    if (dot(p.shading_normal, dir) < 0.0f) return rgb(0.0f);
    return e.radiance;
}

struct PathCtx {
    const or_scene* s;
    const sspt_pt_settings* cfg;
    bool brute;
    Counters* c;
};

// PathGraph (PathGraph.cs) of one replayed path: nodes in the order
// RenderPixel / EstimateIncidentRadiance add them, each with its ancestor.
struct Graph {
    std::vector<sspt_path_node> nodes;
    int add(int kind, int ancestor, V3 pos) {
        sspt_path_node n{};
        n.kind = kind;
        n.ancestor = ancestor;
        n.position[0] = pos.x;
        n.position[1] = pos.y;
        n.position[2] = pos.z;
        nodes.push_back(n);
        return (int)nodes.size() - 1;
    }
    static void put(float* dst, V3 v) {
        dst[0] = v.x;
        dst[1] = v.y;
        dst[2] = v.z;
    }
};

// The estimators below run K texture sets along one path: every sample, pdf,
// MIS weight and Russian-roulette decision comes from set 0 (the scene's own
// texels); sets 1.. only change the BSDF values the estimate multiplies in.
// K = 1 is the reference path tracer itself.

// PathTracerBase.PerformBackgroundNextEvent (PathTracer.cs:423-452)
template <int K>
void background_nee(const PathCtx& pc, const Shader* sh, Rng& rng, Rgb* out, Graph* graph = nullptr, int vertex = -1,
                    Rgb prefix = {}) {
    const or_scene& s = *pc.s;
    for (int k = 0; k < K; ++k) out[k] = rgb(0.0f);
    if (!s.has_env) return;
    EnvSample smp = env_sample(s.env, rng.next_float2());
    Ray r = spawn_ray(sh[0].p, smp.dir);
    pc.c->shadow++;
    if (trace_any(s, r, pc.brute)) {
        pc.c->occluded++;
        return;
    }
    float pdf_bsdf = pdf_fwd(sh[0], smp.dir);
    if (pdf_bsdf == 0.0f || smp.pdf == 0.0f) return;
    float nsr = (float)pc.cfg->num_shadow_rays;
    float mis = pc.cfg->enable_bsdf_di ? 1.0f / (1.0f + pdf_bsdf / (smp.pdf * nsr)) : 1.0f;
    for (int k = 0; k < K; ++k) {
        Rgb bsdf_cos = eval_cos(sh[k], smp.dir);
        Rgb contrib = smp.weight * bsdf_cos / nsr;
        out[k] = mis * contrib;
        if (graph && k == 0 && !is_black(contrib)) {  // PathTracer.cs:446-447
            const float* p = graph->nodes[vertex].position;
            const V3 at = v3(p[0], p[1], p[2]) + smp.dir;  // NextEventNode(direction, ancestor, ...)
            sspt_path_node& n = graph->nodes[graph->add(SSPT_NODE_NEXT_EVENT, vertex, at)];
            Graph::put(n.emission, smp.weight * smp.pdf);
            n.pdf = smp.pdf;
            Graph::put(n.bsdf_cos, bsdf_cos);
            n.mis_weight = mis;
            Graph::put(n.prefix_weight, prefix);
            n.flags = SSPT_NODE_TO_BACKGROUND;
        }
    }
}

// PathTracerBase.PerformNextEventEstimation (PathTracer.cs:454-501)
template <int K>
void light_nee(const PathCtx& pc, const Shader* sh, Rng& rng, Rgb* out, Graph* graph = nullptr, int vertex = -1,
               Rgb prefix = {}) {
    const or_scene& s = *pc.s;
    for (int k = 0; k < K; ++k) out[k] = rgb(0.0f);
    if (s.emitters.empty()) return;
    int count = (int)s.emitters.size();
    int idx = (int)rng.next_int((uint32_t)count);
    const Emitter& e = s.emitters[idx];
    float sel = 1.0f / (float)count;
    // Triangle.SampleUniformArea (Triangle.cs:42-56)
    V2 b = to_uniform_triangle(rng.next_float2());
    SurfacePoint lp = make_point(s, e.tri, b, 0.0f);
    float lpdf = e.inv_area;
    V3 l2s = normalize(sh[0].p.position - lp.position);
    Ray r = shadow_ray(sh[0].p, lp);
    pc.c->shadow++;
    if (trace_any(s, r, pc.brute)) {
        pc.c->occluded++;
        return;
    }
    Rgb em = emitted(e, lp, l2s);
    float jac = area_to_solid_angle(sh[0].p, lp);
    float nsr = (float)pc.cfg->num_shadow_rays;
    float pdf_ne = lpdf * sel * nsr;
    float pdf_bsdf_sa = pdf_fwd(sh[0], -l2s);
    float pdf_bsdf = pdf_bsdf_sa * jac;
    if (jac == 0.0f) return;
    float ratio = pdf_bsdf / pdf_ne;
    float mis = pc.cfg->enable_bsdf_di ? 1.0f / (ratio + 1.0f) : 1.0f;
    float pdf = lpdf / jac * sel * nsr;
    for (int k = 0; k < K; ++k) {
        Rgb bsdf_cos = eval_cos(sh[k], -l2s);
        Rgb contrib = em / pdf * bsdf_cos;
        out[k] = mis * contrib;
        if (graph && k == 0 && !is_black(contrib)) {  // PathTracer.cs:495-496
            sspt_path_node& n = graph->nodes[graph->add(SSPT_NODE_NEXT_EVENT, vertex, lp.position)];
            Graph::put(n.emission, em);
            n.pdf = pdf;
            Graph::put(n.bsdf_cos, bsdf_cos);
            n.mis_weight = mis;
            Graph::put(n.prefix_weight, prefix);
        }
    }
}

// PathTracerBase.RenderPixel + EstimateIncidentRadiance (PathTracer.cs:287-382)
template <int K>
void render_pixel_k(const PathCtx& pc, const TextureSet* const* sets, uint32_t row, uint32_t col,
                    uint32_t iteration, Rgb* radiance, Graph* graph = nullptr) {
    const or_scene& s = *pc.s;
    const sspt_pt_settings& cfg = *pc.cfg;
    Rng rng{hash_seed(cfg.base_seed, row * (uint32_t)s.width + col, iteration)};
    V2 off = rng.next_float2();
    Ray ray;
    s.camera.generate_ray((float)col + off.x, (float)row + off.y, &ray.o, &ray.d);
    ray.tmin = 0.0f;
    ray.tmax = INFINITY;

    Rgb prefix[K];
    for (int k = 0; k < K; ++k) {
        radiance[k] = rgb(0.0f);
        prefix[k] = rgb(1.0f);
    }
    Rgb approx = rgb(1.0f);
    int depth = 1;
    SurfacePoint prev;
    float prev_pdf = 0.0f;
    int count = (int)s.emitters.size();
    float nsr = (float)cfg.num_shadow_rays;
    // ReplayPixel's graph (PathTracer.cs:303): the root, then the current vertex
    int vertex = graph ? graph->add(SSPT_NODE_CAMERA, -1, ray.o) : -1;
    Rgb prev_scatter = rgb(1.0f);  // PathState.PreviousScatterWeight / PreviousSurvivalProbability
    float prev_survival = 1.0f;

    while (depth <= cfg.max_depth) {
        HitRec h = trace_closest(s, ray, pc.brute);
        pc.c->rays++;
        if (h.tri < 0) {
            if (depth >= cfg.min_depth) {
                // OnBackgroundHit (PathTracer.cs:384-400)
                float mis = 0.0f;
                Rgb contrib = rgb(0.0f);
                if (s.has_env && cfg.enable_bsdf_di) {
                    mis = 1.0f;
                    if (depth > 1) {
                        float pdf_ne = env_pdf(s.env, ray.d) * nsr;
                        mis = 1.0f / (1.0f + pdf_ne / prev_pdf);
                    }
                    contrib = env_radiance(s.env, ray.d);
                }
                for (int k = 0; k < K; ++k) radiance[k] = radiance[k] + prefix[k] * mis * contrib;
                if (graph) {  // PathTracer.cs:326
                    const float* p = graph->nodes[vertex].position;
                    sspt_path_node& n = graph->nodes[graph->add(SSPT_NODE_BACKGROUND, vertex, v3(p[0], p[1], p[2]) + ray.d)];
                    Graph::put(n.emission, contrib);
                    n.mis_weight = mis;
                }
            }
            break;
        }
        pc.c->hits++;
        SurfacePoint hit = make_point(s, h.tri, {h.u, h.v}, h.t);
        Shader sh[K];
        sh[0] = make_shader(s, hit, -ray.d);
        for (int k = 1; k < K; ++k) {
            sh[k] = sh[0];
            sh[k].tex = sets[k];
        }

        int ei = s.tri_emitter[h.tri];
        if (ei >= 0 && depth >= cfg.min_depth) {
            // OnLightHit (PathTracer.cs:402-421)
            const Emitter& e = s.emitters[ei];
            float mis = 1.0f;
            if (depth > 1) {
                float jac = area_to_solid_angle(prev, hit);
                float pdf_ne = e.inv_area / (float)count * nsr / jac;
                float ratio = pdf_ne / prev_pdf;
                mis = 1.0f / (ratio + 1.0f);
                if (!cfg.enable_bsdf_di) mis = 0.0f;
            }
            Rgb em = emitted(e, hit, -ray.d);
            for (int k = 0; k < K; ++k) radiance[k] = radiance[k] + prefix[k] * mis * em;
            if (graph) {  // PathTracer.cs:346
                vertex = graph->add(SSPT_NODE_BSDF_SAMPLE, vertex, hit.position);
                sspt_path_node& n = graph->nodes[vertex];
                Graph::put(n.scatter_weight, prev_scatter);
                n.survival_probability = prev_survival;
                Graph::put(n.emission, em);
                n.mis_weight = mis;
                n.flags = SSPT_NODE_HIT_EMITTER;
            }
        } else if (graph) {  // PathTracer.cs:348
            vertex = graph->add(SSPT_NODE_BSDF_SAMPLE, vertex, hit.position);
            Graph::put(graph->nodes[vertex].scatter_weight, prev_scatter);
            graph->nodes[vertex].survival_probability = prev_survival;
        }

        // ComputeSurvivalProbability (PathTracer.cs:277-282)
        float q = depth > 4 ? std::min(std::max(average(approx), 0.05f), 0.95f) : 1.0f;
        if (rng.next_float() > q || depth == cfg.max_depth) break;

        if (depth + 1 >= cfg.min_depth) {
            Rgb nee[K], part[K];
            for (int k = 0; k < K; ++k) nee[k] = rgb(0.0f);
            for (int i = 0; i < cfg.num_shadow_rays; ++i) {
                background_nee<K>(pc, sh, rng, part, graph, vertex, prefix[0]);
                for (int k = 0; k < K; ++k) nee[k] = nee[k] + part[k];
                light_nee<K>(pc, sh, rng, part, graph, vertex, prefix[0]);
                for (int k = 0; k < K; ++k) nee[k] = nee[k] + part[k];
            }
            for (int k = 0; k < K; ++k) radiance[k] = radiance[k] + prefix[k] * nee[k] / q;
        }

        // SampleDirection (PathTracer.cs:526-530)
        float u_comp = rng.next_float();
        V2 u_dir = rng.next_float2();
        BsdfSample bs = sample_bsdf(sh[0], u_comp, u_dir);
        if (bs.pdf == 0.0f || is_black(bs.weight)) break;
        ray = spawn_ray(hit, bs.dir);

        prefix[0] = prefix[0] * (bs.weight / q);
        for (int k = 1; k < K; ++k) prefix[k] = prefix[k] * (eval_cos(sh[k], bs.dir) / bs.pdf / q);
        approx = approx * (bs.weight / q);
        depth++;
        prev = hit;
        prev_pdf = bs.pdf * q;
        prev_scatter = bs.weight / q;
        prev_survival = q;
    }
}

Rgb render_pixel(const PathCtx& pc, uint32_t row, uint32_t col, uint32_t iteration) {
    const TextureSet* sets[1] = {&pc.s->textures};
    Rgb L;
    render_pixel_k<1>(pc, sets, row, col, iteration, &L);
    return L;
}

template <class F>
void parallel_for(int64_t n, int threads, F&& f) {
    // default: the hardware threads, at most 16 (one GPU's share of a shared host)
    if (threads <= 0) threads = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        for (;;) {
            int64_t i = next.fetch_add(1);
            if (i >= n) break;
            f(i);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
}

void set_err(char* err, const char* msg) {
    if (err) std::snprintf(err, 256, "%s", msg);
}

}  // namespace

// ======================================================================= C API

extern "C" {

uint32_t or_hash_seed(uint32_t b, uint32_t c, uint32_t s) { return hash_seed(b, c, s); }
uint32_t or_pcg_hash(uint32_t x) { return pcg_hash(x); }
uint32_t or_fnv_hash(uint32_t h, uint32_t d) { return fnv_hash(h, d); }

void or_rng_stream(uint32_t seed, uint32_t n, uint32_t* out_u32, float* out_float) {
    Rng a{seed}, b{seed};
    for (uint32_t i = 0; i < n; ++i) {
        if (out_u32) out_u32[i] = a.next();
        if (out_float) out_float[i] = b.next_float();
    }
}

void or_rng_next_int(uint32_t seed, uint32_t max, uint32_t n, uint32_t* out) {
    Rng r{seed};
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next_int(max);
}

void or_warp_uniform_triangle(const float* p, float* out) {
    V2 r = to_uniform_triangle({p[0], p[1]});
    out[0] = r.x;
    out[1] = r.y;
}

void or_warp_cos_hemisphere(const float* p, float* out, float* pdf) {
    DirSample d = to_cos_hemisphere({p[0], p[1]});
    out[0] = d.dir.x;
    out[1] = d.dir.y;
    out[2] = d.dir.z;
    *pdf = d.pdf;
}

void or_pwc_sample(const float* w, int32_t n, float u, int32_t* idx, float* rel) {
    Pwc p;
    p.build(w, n);
    int i;
    p.sample(u, &i, rel);
    *idx = i;
}

float or_pwc_probability(const float* w, int32_t n, int32_t idx) {
    Pwc p;
    p.build(w, n);
    return p.probability(idx);
}

void or_camera_ray(const float* w2c, float fov, int32_t w, int32_t h, float fx, float fy, float* out) {
    Camera c;
    c.init(w2c, fov, w, h);
    V3 o, d;
    c.generate_ray(fx, fy, &o, &d);
    out[0] = o.x, out[1] = o.y, out[2] = o.z;
    out[3] = d.x, out[4] = d.y, out[5] = d.z;
}

// Matrix4x4.CreateLookAt (right-handed), fp32.
void or_look_at(const float* pos, const float* target, const float* up, float* out) {
    V3 p = v3(pos[0], pos[1], pos[2]), t = v3(target[0], target[1], target[2]);
    V3 u = v3(up[0], up[1], up[2]);
    V3 z = normalize(p - t);
    V3 x = normalize(cross(u, z));
    V3 y = cross(z, x);
    float m[16] = {x.x, y.x, z.x, 0, x.y, y.y, z.y, 0, x.z, y.z, z.z, 0,
                   -dot(x, p), -dot(y, p), -dot(z, p), 1};
    std::memcpy(out, m, sizeof(m));
}

float or_ggx_ndf(float ax, float ay, const float* n) { return Ggx{ax, ay}.ndf(v3(n[0], n[1], n[2])); }
float or_ggx_pdf(float ax, float ay, const float* o, const float* n) {
    return Ggx{ax, ay}.pdf(v3(o[0], o[1], o[2]), v3(n[0], n[1], n[2]));
}
void or_ggx_sample(float ax, float ay, const float* o, const float* p, float* out) {
    V3 r = Ggx{ax, ay}.sample(v3(o[0], o[1], o[2]), {p[0], p[1]});
    out[0] = r.x, out[1] = r.y, out[2] = r.z;
}
float or_fresnel_dielectric(float c, float a, float b) { return fresnel_dielectric(c, a, b); }

static or_scene* flat_surface(const sspt_material_desc* mat, const sspt_texture_desc* textures,
                              const float* normal, const float* uv, SurfacePoint* p) {
    or_scene* s = new or_scene();
    Material m{mat->kind,
               v3(mat->base_color[0], mat->base_color[1], mat->base_color[2]),
               mat->base_color_texture,
               mat->roughness,
               mat->roughness_texture,
               mat->metallic,
               mat->specular_tint_strength,
               mat->anisotropic,
               mat->specular_transmittance,
               mat->index_of_refraction,
               mat->transmitter != 0};
    s->materials.push_back(m);
    int max_tex = std::max(mat->base_color_texture, mat->roughness_texture);
    for (int i = 0; i <= max_tex; ++i) {
        Texture t;
        t.w = textures[i].width, t.h = textures[i].height, t.ch = textures[i].channels;
        t.border = textures[i].border;
        t.texels.assign(textures[i].texels, textures[i].texels + (size_t)t.w * t.h * t.ch);
        s->textures.push_back(t);
    }
    Tri t{};
    t.material = 0;
    t.face_normal = v3(normal[0], normal[1], normal[2]);
    s->tris.push_back(t);
    p->tri = 0;
    p->position = v3(0, 0, 0);
    p->normal = t.face_normal;
    p->shading_normal = t.face_normal;
    p->uv = {uv[0], uv[1]};
    return s;
}

void or_material_eval(const sspt_material_desc* mat, const sspt_texture_desc* textures,
                      const float* normal, const float* uv, const float* out_dir, const float* in_dir,
                      float* value, float* value_cos, float* pdf) {
    SurfacePoint p;
    or_scene* s = flat_surface(mat, textures, normal, uv, &p);
    Shader sh = make_shader(*s, p, v3(out_dir[0], out_dir[1], out_dir[2]));
    V3 in = v3(in_dir[0], in_dir[1], in_dir[2]);
    Rgb v = eval_nocos(sh, in), vc = eval_cos(sh, in);
    value[0] = v.x, value[1] = v.y, value[2] = v.z;
    value_cos[0] = vc.x, value_cos[1] = vc.y, value_cos[2] = vc.z;
    *pdf = pdf_fwd(sh, in);
    delete s;
}

void or_material_pdf_pair(const sspt_material_desc* mat, const sspt_texture_desc* textures, const float* normal,
                          const float* uv, const float* out_dir, const float* in_dir, float* pdf, float* pdf_rev) {
    SurfacePoint p;
    or_scene* s = flat_surface(mat, textures, normal, uv, &p);
    Shader sh = make_shader(*s, p, v3(out_dir[0], out_dir[1], out_dir[2]));
    pdf_pair(sh, v3(in_dir[0], in_dir[1], in_dir[2]), pdf, pdf_rev);
    delete s;
}

void or_material_eval_many(const sspt_material_desc* mat, const sspt_texture_desc* textures,
                           const float* normal, const float* uv, const float* out_dir, int64_t n,
                           const float* in_dirs, float* value_cos, float* pdf) {
    SurfacePoint p;
    or_scene* s = flat_surface(mat, textures, normal, uv, &p);
    Shader sh = make_shader(*s, p, v3(out_dir[0], out_dir[1], out_dir[2]));
    for (int64_t i = 0; i < n; ++i) {
        V3 in = v3(in_dirs[3 * i], in_dirs[3 * i + 1], in_dirs[3 * i + 2]);
        Rgb vc = eval_cos(sh, in);
        value_cos[3 * i] = vc.x, value_cos[3 * i + 1] = vc.y, value_cos[3 * i + 2] = vc.z;
        if (pdf) pdf[i] = pdf_fwd(sh, in);
    }
    delete s;
}

void or_material_sample(const sspt_material_desc* mat, const sspt_texture_desc* textures,
                        const float* normal, const float* uv, const float* out_dir, float pc,
                        const float* p2, float* out) {
    SurfacePoint p;
    or_scene* s = flat_surface(mat, textures, normal, uv, &p);
    Shader sh = make_shader(*s, p, v3(out_dir[0], out_dir[1], out_dir[2]));
    BsdfSample bs = sample_bsdf(sh, pc, {p2[0], p2[1]});
    out[0] = bs.dir.x, out[1] = bs.dir.y, out[2] = bs.dir.z;
    out[3] = bs.pdf;
    out[4] = bs.weight.x, out[5] = bs.weight.y, out[6] = bs.weight.z;
    out[7] = bs.pdf_rev;
    delete s;
}

// Monte Carlo integrals of Material_SampleTransform.cs:26-117 over n samples:
// mode 0: pdf with uniform sphere directions, 1: average of f*|cos| with uniform
// sphere directions, 2: average BSDF-sample weight. Sample i draws from
// RNG(HashSeed(seed, i, 0)); uniform sphere = SampleWarp.ToUniformSphere (SampleWarp.cs:230-237).
void or_material_integrate(const sspt_material_desc* mat, const sspt_texture_desc* textures,
                           const float* normal, const float* uv, const float* out_dir, int64_t n,
                           uint32_t seed, int32_t mode, double* mean, double* var) {
    SurfacePoint p;
    or_scene* s = flat_surface(mat, textures, normal, uv, &p);
    Shader sh = make_shader(*s, p, v3(out_dir[0], out_dir[1], out_dir[2]));
    double sum = 0.0, sq = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        Rng rng{hash_seed(seed, (uint32_t)i, 0)};
        float val = 0.0f;
        if (mode == 2) {
            float pc = rng.next_float();
            V2 pd = rng.next_float2();
            BsdfSample bs = sample_bsdf(sh, pc, pd);
            if (bs.pdf > 0.0f) val = average(bs.weight);
        } else {
            V2 u = rng.next_float2();
            float a = 2.0f * kPi * u.y;
            float b = 2.0f * std::sqrt(u.x - u.x * u.x);
            V3 d = spherical_to_cartesian(b, 1.0f - 2.0f * u.x, a);
            float pdf_u = 1.0f / (4.0f * kPi);
            val = (mode == 0 ? pdf_fwd(sh, d) : average(eval_cos(sh, d))) / pdf_u;
        }
        sum += val;
        sq += (double)val * val;
    }
    *mean = sum / (double)n;
    *var = sq / (double)n - (*mean) * (*mean);
    delete s;
}

// Scene.Prepare (SeeSharp/Scene.cs:103-164) + IsValid (:170-197)
or_scene* or_scene_create(const sspt_scene_desc* d, char* err) {
    if (!d) {
        set_err(err, "null scene description");
        return nullptr;
    }
    if (d->width == 0 || d->height == 0) {
        set_err(err, "Framebuffer not set.");
        return nullptr;
    }
    if (d->num_meshes == 0) {
        set_err(err, "No meshes in the scene.");
        return nullptr;
    }
    if (d->num_emitters == 0 && !d->background) {
        set_err(err, "No emitters and no background in the scene.");
        return nullptr;
    }
    or_scene* s = new or_scene();
    s->width = (int)d->width;
    s->height = (int)d->height;
    for (uint32_t i = 0; i < d->num_textures; ++i) {
        const sspt_texture_desc& td = d->textures[i];
        Texture t;
        t.w = (int)td.width, t.h = (int)td.height, t.ch = (int)td.channels, t.border = td.border;
        t.texels.assign(td.texels, td.texels + (size_t)t.w * t.h * t.ch);
        s->textures.push_back(std::move(t));
    }
    for (uint32_t i = 0; i < d->num_materials; ++i) {
        const sspt_material_desc& m = d->materials[i];
        s->materials.push_back(Material{m.kind, v3(m.base_color[0], m.base_color[1], m.base_color[2]),
                                        m.base_color_texture, m.roughness, m.roughness_texture,
                                        m.metallic, m.specular_tint_strength, m.anisotropic,
                                        m.specular_transmittance, m.index_of_refraction,
                                        m.transmitter != 0});
    }
    std::vector<int> mesh_offset(d->num_meshes);
    for (uint32_t mi = 0; mi < d->num_meshes; ++mi) {
        const sspt_mesh_desc& m = d->meshes[mi];
        if (m.material < 0 || (uint32_t)m.material >= d->num_materials) {
            char buf[128];
            std::snprintf(buf, sizeof(buf), "Mesh[%u] does not have a material.", mi);
            set_err(err, buf);
            delete s;
            return nullptr;
        }
        mesh_offset[mi] = (int)s->tris.size();
        auto P = [&](int v) { return v3(m.positions[v * 3], m.positions[v * 3 + 1], m.positions[v * 3 + 2]); };
        for (uint32_t f = 0; f < m.num_triangles; ++f) {
            Tri t{};
            int i0 = m.indices[f * 3], i1 = m.indices[f * 3 + 1], i2 = m.indices[f * 3 + 2];
            t.p0 = P(i0), t.p1 = P(i1), t.p2 = P(i2);
            t.face_normal = normalize(cross(t.p1 - t.p0, t.p2 - t.p0));
            t.has_normals = m.normals != nullptr;
            if (t.has_normals) {
                auto N = [&](int v) { return v3(m.normals[v * 3], m.normals[v * 3 + 1], m.normals[v * 3 + 2]); };
                t.n0 = N(i0), t.n1 = N(i1), t.n2 = N(i2);
            }
            t.has_uvs = m.uvs != nullptr;
            if (t.has_uvs) {
                t.t0 = {m.uvs[i0 * 2], m.uvs[i0 * 2 + 1]};
                t.t1 = {m.uvs[i1 * 2], m.uvs[i1 * 2 + 1]};
                t.t2 = {m.uvs[i2 * 2], m.uvs[i2 * 2 + 1]};
            }
            t.mesh = (int)mi;
            t.prim = (int)f;
            t.material = m.material;
            s->tris.push_back(t);
        }
    }
    s->tri_emitter.assign(s->tris.size(), -1);
    for (uint32_t i = 0; i < d->num_emitters; ++i) {
        const sspt_emitter_desc& e = d->emitters[i];
        int tri = mesh_offset[e.mesh] + (int)e.face;
        const Tri& t = s->tris[tri];
        V3 n = cross(t.p1 - t.p0, t.p2 - t.p0);
        float area = length(n) * 0.5f;
        if (area == 0.0f) {
            set_err(err, "Zero-area triangles cannot be emitters");
            delete s;
            return nullptr;
        }
        if (s->tri_emitter[tri] >= 0) {
            set_err(err, "duplicate emitter for one triangle");
            delete s;
            return nullptr;
        }
        s->tri_emitter[tri] = (int)i;
        if (e.kind != SSPT_EMITTER_DIFFUSE && e.kind != SSPT_EMITTER_GLOSSY) {
            set_err(err, "unknown emitter kind");
            delete s;
            return nullptr;
        }
        const bool glossy = e.kind == SSPT_EMITTER_GLOSSY;
        if (glossy && !(std::isfinite(e.exponent) && e.exponent > -1.0f)) {
            set_err(err, "glossy exponent must be finite and > -1");
            delete s;
            return nullptr;
        }
        const float kPiF = 3.14159265358979323846f;
        s->emitters.push_back({tri, v3(e.radiance[0], e.radiance[1], e.radiance[2]), 1.0f / area, e.kind,
                               glossy ? e.exponent : 0.0f, glossy ? (e.exponent + 1.0f) / (2.0f * kPiF) : 1.0f});
    }
    if (d->background) {
        s->has_env = true;
        EnvMap& e = s->env;
        e.w = (int)d->background->width;
        e.h = (int)d->background->height;
        e.rgb.assign(d->background->rgb, d->background->rgb + (size_t)e.w * e.h * 3);
        // EnvironmentMap.BuildSamplingGrid (EnvironmentMap.cs:139-156)
        e.grid.init(e.w, e.h);
        for (int r = 0; r < e.h; ++r)
            for (int c = 0; c < e.w; ++c) e.grid.splat(c, r, luminance(e.pixel(c, r)));
        e.grid.normalize();
        // Scene.Prepare (Scene.cs:114-141): the centre is the mean of every mesh vertex
        // (summed in mesh and vertex order), the radius the distance to the furthest one
        V3 sum = v3(0.0f, 0.0f, 0.0f);
        uint64_t total = 0;
        for (uint32_t mi = 0; mi < d->num_meshes; ++mi) {
            const sspt_mesh_desc& m = d->meshes[mi];
            for (uint32_t v = 0; v < m.num_vertices; ++v)
                sum = sum + v3(m.positions[v * 3], m.positions[v * 3 + 1], m.positions[v * 3 + 2]);
            total += m.num_vertices;
        }
        const float count = (float)total;
        s->scene_center = v3(sum.x / count, sum.y / count, sum.z / count);
        float r2 = 0.0f;
        for (uint32_t mi = 0; mi < d->num_meshes; ++mi) {
            const sspt_mesh_desc& m = d->meshes[mi];
            for (uint32_t v = 0; v < m.num_vertices; ++v) {
                V3 o = v3(m.positions[v * 3], m.positions[v * 3 + 1], m.positions[v * 3 + 2]) - s->scene_center;
                r2 = std::max(dot(o, o), r2);
            }
        }
        s->scene_radius = std::sqrt(r2);
    }
    if (!s->camera.init(d->camera.world_to_camera, d->camera.vertical_fov_degrees, s->width, s->height)) {
        set_err(err, "World to camera transform must be invertible.");
        delete s;
        return nullptr;
    }
    for (const Tri& t : s->tris)
        for (V3 v : {t.p0, t.p1, t.p2})
            s->abs_max = std::max(s->abs_max, std::max(std::fabs(v.x), std::max(std::fabs(v.y), std::fabs(v.z))));
    build_bvh(*s);
    return s;
}

void or_scene_destroy(or_scene* s) { delete s; }

void or_env_eval(const or_scene* s, const float* d, float* out) {
    Rgb c = env_radiance(s->env, v3(d[0], d[1], d[2]));
    out[0] = c.x, out[1] = c.y, out[2] = c.z;
}
float or_env_pdf(const or_scene* s, const float* d) { return env_pdf(s->env, v3(d[0], d[1], d[2])); }
void or_env_sample(const or_scene* s, const float* p, float* out) {
    EnvSample e = env_sample(s->env, {p[0], p[1]});
    out[0] = e.dir.x, out[1] = e.dir.y, out[2] = e.dir.z, out[3] = e.pdf;
    out[4] = e.weight.x, out[5] = e.weight.y, out[6] = e.weight.z;
}

static Ray read_ray(const float* r) {
    Ray ray;
    ray.o = v3(r[0], r[1], r[2]);
    ray.tmin = r[3];
    ray.d = v3(r[4], r[5], r[6]);
    ray.tmax = r[7];
    return ray;
}

void or_trace_closest(const or_scene* s, const float* rays, uint64_t n, sspt_hit* hits, int32_t brute) {
    parallel_for((int64_t)((n + 255) / 256), 0, [&](int64_t blk) {
        uint64_t end = std::min<uint64_t>(n, (uint64_t)(blk + 1) * 256);
        for (uint64_t i = (uint64_t)blk * 256; i < end; ++i) {
            HitRec h = trace_closest(*s, read_ray(rays + i * 8), brute != 0);
            hits[i] = sspt_hit{h.t, h.u, h.v, h.tri};
        }
    });
}

void or_trace_occluded(const or_scene* s, const float* rays, uint64_t n, uint8_t* occ, int32_t brute) {
    parallel_for((int64_t)((n + 255) / 256), 0, [&](int64_t blk) {
        uint64_t end = std::min<uint64_t>(n, (uint64_t)(blk + 1) * 256);
        for (uint64_t i = (uint64_t)blk * 256; i < end; ++i)
            occ[i] = trace_any(*s, read_ray(rays + i * 8), brute != 0) ? 1 : 0;
    });
}

void or_primary_hits(const or_scene* s, uint32_t base_seed, uint32_t iteration, float* rays_out,
                     sspt_hit* hits_out, int32_t brute) {
    int W = s->width, H = s->height;
    parallel_for(H, 0, [&](int64_t row) {
        for (int col = 0; col < W; ++col) {
            size_t pix = (size_t)row * W + col;
            Rng rng{hash_seed(base_seed, (uint32_t)pix, iteration)};
            V2 off = rng.next_float2();
            Ray r;
            s->camera.generate_ray((float)col + off.x, (float)row + off.y, &r.o, &r.d);
            r.tmin = 0.0f;
            r.tmax = INFINITY;
            if (rays_out) {
                float* o = rays_out + pix * 8;
                o[0] = r.o.x, o[1] = r.o.y, o[2] = r.o.z, o[3] = r.tmin;
                o[4] = r.d.x, o[5] = r.d.y, o[6] = r.d.z, o[7] = r.tmax;
            }
            if (hits_out) {
                HitRec h = trace_closest(*s, r, brute != 0);
                hits_out[pix] = sspt_hit{h.t, h.u, h.v, h.tri};
            }
        }
    });
}

uint64_t or_render(const or_scene* s, const sspt_pt_settings* cfg, uint32_t first, uint32_t num,
                   float* image, sspt_ray_stats* stats, const int32_t* pixels, int64_t num_pixels,
                   int32_t threads) {
    int W = s->width, H = s->height;
    int64_t total = pixels ? num_pixels : (int64_t)W * H;
    std::atomic<uint64_t> rays{0}, shadow{0}, occl{0}, hits{0}, nans{0};
    for (uint32_t it = first; it < first + num; ++it) {
        // FrameBuffer.StartIteration / Normalize (FrameBuffer.cs:146-160)
        float n = (float)(it + 1);
        if (it + 1 > 1) {
            float scale = (n - 1.0f) / n;
            for (int64_t k = 0; k < total; ++k) {
                int64_t pix = pixels ? pixels[k] : k;
                for (int c = 0; c < 3; ++c) image[pix * 3 + c] *= scale;
            }
        }
        int64_t chunks = (total + 63) / 64;
        parallel_for(chunks, threads, [&](int64_t ch) {
            Counters c;
            PathCtx pc{s, cfg, false, &c};
            int64_t end = std::min<int64_t>(total, (ch + 1) * 64);
            for (int64_t k = ch * 64; k < end; ++k) {
                int64_t pix = pixels ? pixels[k] : k;
                uint32_t row = (uint32_t)(pix / W), col = (uint32_t)(pix % W);
                Rgb L = render_pixel(pc, row, col, it);
                // FrameBuffer.Splat (FrameBuffer.cs:89-114), IgnoreNanAndInf.
                if (!std::isfinite(average(L))) {
                    nans++;
                    continue;
                }
                Rgb add = L / n;
                image[pix * 3 + 0] += add.x;
                image[pix * 3 + 1] += add.y;
                image[pix * 3 + 2] += add.z;
            }
            rays += c.rays, shadow += c.shadow, occl += c.occluded, hits += c.hits;
        });
    }
    if (stats) {
        stats->num_rays = rays;
        stats->num_shadow_rays = shadow;
        stats->num_occluded = occl;
        stats->num_ray_hits = hits;
    }
    return nans;
}

// ------------------------------------------------------------ light tracer -
namespace {

// TinyEmbree Raytracer.IsOccluded(from, Vector3 to) with a free end point (the
// camera): only the surface end is offset, d = to - p0, t in (1e-5, 1 - 1e-5).
inline Ray shadow_ray_to(const SurfacePoint& from, V3 to) {
    V3 p0 = offset_point(from, to - from.position);
    Ray r;
    r.o = p0;
    r.d = to - p0;
    r.tmin = 1e-5f;
    r.tmax = 1.0f - 1e-5f;
    return r;
}

struct EmitterRay {
    SurfacePoint p;
    V3 dir;
    float pdf;
    Rgb weight;
};
// Emitter.SampleRay: DiffuseEmitter (DiffuseEmitter.cs:44-64), GlossyEmitter
// (GlossyEmitter.cs:47-70); the point from Triangle.SampleUniformArea (Triangle.cs:42-56).
EmitterRay emitter_sample_ray(const or_scene& s, const Emitter& e, V2 primary_pos, V2 primary_dir) {
    EmitterRay r;
    r.p = make_point(s, e.tri, to_uniform_triangle(primary_pos), 0.0f);
    float pos_pdf = e.inv_area;
    V3 n = r.p.shading_normal, t, b;
    basis(n, &t, &b);
    bool glossy = e.kind == SSPT_EMITTER_GLOSSY;
    DirSample local = glossy ? to_cosine_lobe(e.exponent + 1.0f, primary_dir) : to_cos_hemisphere(primary_dir);
    r.dir = local.dir.z * n + local.dir.x * t + local.dir.y * b;
    r.pdf = local.pdf * pos_pdf;
    if (glossy) {
        Rgb w = e.radiance * std::pow(local.dir.z, e.exponent + 1.0f) * e.norm;
        r.weight = w / pos_pdf / local.pdf;
    } else {
        r.weight = e.radiance / pos_pdf * kPi;
    }
    return r;
}

struct Splat {
    int64_t pix;
    Rgb value;
};

struct LightCtx {
    const or_scene* s;
    const sspt_lt_settings* cfg;
    float num_paths;
    Counters* c;
};

// BidirBase.ConnectLightVertexToCamera (BidirBase.Light.cs:35-92) with MIS weight 1.
void connect_to_camera(const LightCtx& lc, const SurfacePoint& v, V3 ancestor, Rgb weight, int depth,
                       std::vector<Splat>* out) {
    const or_scene& s = *lc.s;
    if (depth + 1 < lc.cfg->min_depth) return;
    int px, py;
    float jac;
    if (!s.camera.sample_response(v.position, v.normal, &px, &py, &jac)) return;
    lc.c->shadow++;
    if (trace_any(s, shadow_ray_to(v, s.camera.position), false)) {
        lc.c->occluded++;
        return;
    }
    V3 to_cam = s.camera.position - v.position;
    float dist = length(to_cam);
    to_cam = to_cam / dist;
    V3 to_anc = normalize(ancestor - v.position);
    Shader sh = make_shader(s, v, to_cam);  // the camera side: IsOnLightSubpath = false
    Rgb f = eval_nocos(sh, to_anc);
    if (is_black(f)) return;
    f = f * (std::fabs(dot(v.shading_normal, to_anc)) / std::fabs(dot(v.normal, to_anc)));
    Rgb w = weight * f * rgb(jac) / lc.num_paths;
    out->push_back({(int64_t)py * s.width + px, w});
}

// BidirBase.TraceLightPath / TraceEmitterPath (BidirBase.Light.cs:212-277) and
// RandomWalk.StartFromEmitter / ContinueWalk (RandomWalk.cs:82-90,150-217) for
// one light path, each vertex connected to the camera as LightPathWalk.OnHit
// stores it (the connection draws no random numbers, so connecting during the
// walk equals connecting the cached vertices afterwards).
void light_path(const LightCtx& lc, uint32_t idx, uint32_t iteration, std::vector<Splat>* out) {
    const or_scene& s = *lc.s;
    const sspt_lt_settings& cfg = *lc.cfg;
    Rng rng{hash_seed(cfg.base_seed_light, idx, iteration)};
    // SelectLight (:103-110): without a background BackgroundProbability is 0 and draws nothing
    int count = (int)s.emitters.size();
    const Emitter& e = s.emitters[rng.next_int((uint32_t)count)];
    float sel = (1.0f - 0.0f) / (float)count;
    V2 ppos = rng.next_float2();
    V2 pdir = rng.next_float2();
    EmitterRay er = emitter_sample_ray(s, e, ppos, pdir);
    er.pdf *= sel;
    if (er.pdf == 0.0f || is_black(er.weight)) return;
    Rgb prefix = er.weight / sel;
    Rgb approx = rgb(1.0f);
    Ray ray = spawn_ray(er.p, er.dir);
    SurfacePoint prev = er.p;
    float pdf_dir = er.pdf;
    for (int depth = 1; depth < cfg.max_depth;) {
        HitRec h = trace_closest(s, ray, false);
        lc.c->rays++;
        if (h.tri < 0) break;
        lc.c->hits++;
        SurfacePoint hit = make_point(s, h.tri, {h.u, h.v}, h.t);
        float pdf_from = pdf_dir * area_to_solid_angle(prev, hit);
        if (pdf_from == 0.0f) break;
        connect_to_camera(lc, hit, prev.position, prefix, depth, out);
        if (depth + 1 >= cfg.max_depth) break;
        float q = depth > 4 ? std::min(std::max(average(approx), 0.05f), 0.95f) : 1.0f;
        if (rng.next_float() > q) break;
        Shader sh = make_shader(s, hit, -ray.d);
        sh.light = true;
        float pc = rng.next_float();
        V2 pd = rng.next_float2();
        BsdfSample bs = sample_bsdf(sh, pc, pd);
        approx = approx * (bs.weight / q);
        if (bs.pdf == 0.0f || is_black(bs.weight)) break;
        // RandomWalk.cs:193-204: geometric instead of shading cosine, and the
        // rendering equation's cosine ratio at the light subpath's vertex
        Rgb w = bs.weight * (std::fabs(dot(hit.normal, bs.dir)) / std::fabs(dot(hit.shading_normal, bs.dir)));
        w = w * (std::fabs(dot(hit.shading_normal, -ray.d)) / std::fabs(dot(hit.normal, -ray.d)));
        prefix = prefix * (w / q);
        ++depth;
        pdf_dir = bs.pdf;
        prev = hit;
        ray = spawn_ray(hit, bs.dir);
    }
}

}  // namespace

uint64_t or_render_light_tracer(const or_scene* s, const sspt_lt_settings* cfg, uint32_t first, uint32_t num,
                                float* image, sspt_ray_stats* stats, int32_t threads) {
    if (s->has_env) {
        // VertexCacheBidir with no connections, next events or emitter hits: with a
        // background its camera paths (BidirBase's default BaseSeedCamera) hit it under
        // EmitterHitMis (BidirBase.Camera.cs:532-561)
        sspt_vcm_settings v{};
        v.algorithm = SSPT_VERTEX_CACHE_BIDIR;
        v.base_seed_camera = 0xC030114u;
        v.base_seed_light = cfg->base_seed_light;
        v.max_depth = cfg->max_depth;
        v.min_depth = cfg->min_depth;
        v.num_light_paths = cfg->num_light_paths;
        v.enable_light_tracer = 1;
        v.max_num_photons = 8;
        return or_render_vcm(s, &v, 1.0f, first, num, image, stats, threads);
    }
    if (s->emitters.empty()) return UINT64_MAX;
    const int64_t W = s->width, H = s->height;
    const uint32_t paths = cfg->num_light_paths < 0 ? (uint32_t)(W * H) : (uint32_t)cfg->num_light_paths;
    std::atomic<uint64_t> rays{0}, shadow{0}, occl{0}, hits{0};
    uint64_t nans = 0;
    std::vector<std::vector<Splat>> splats(paths);
    for (uint32_t it = first; it < first + num; ++it) {
        // FrameBuffer.StartIteration / Normalize (FrameBuffer.cs:232-249)
        float n = (float)(it + 1);
        if (it + 1 > 1) {
            float scale = (n - 1.0f) / n;
            for (int64_t k = 0; k < W * H * 3; ++k) image[k] *= scale;
        }
        int64_t chunks = ((int64_t)paths + 63) / 64;
        parallel_for(chunks, threads, [&](int64_t ch) {
            Counters c;
            LightCtx lc{s, cfg, (float)paths, &c};
            int64_t end = std::min<int64_t>(paths, (ch + 1) * 64);
            for (int64_t k = ch * 64; k < end; ++k) {
                splats[k].clear();
                light_path(lc, (uint32_t)k, it, &splats[k]);
            }
            rays += c.rays, shadow += c.shadow, occl += c.occluded, hits += c.hits;
        });
        // FrameBuffer.Splat (FrameBuffer.cs:160-191), IgnoreNanAndInf, in light-path order
        for (uint32_t k = 0; k < paths; ++k)
            for (const Splat& sp : splats[k]) {
                if (!std::isfinite(average(sp.value))) {
                    ++nans;
                    continue;
                }
                Rgb add = sp.value / n;
                image[sp.pix * 3 + 0] += add.x;
                image[sp.pix * 3 + 1] += add.y;
                image[sp.pix * 3 + 2] += add.z;
            }
    }
    if (stats) {
        stats->num_rays = rays;
        stats->num_shadow_rays = shadow;
        stats->num_occluded = occl;
        stats->num_ray_hits = hits;
        stats->num_shadow_rays_traced = shadow;
        stats->num_iterations = num;
    }
    return nans;
}

// --------------------------------------------- bidirectional path tracing -
namespace {

// Emitter.PdfRay: DiffuseEmitter.cs:39-42, GlossyEmitter.cs:42-45 with
// SampleWarp.ToCosineLobeJacobian (SampleWarp.cs:192-194).
float emitter_pdf_ray(const Emitter& e, const SurfacePoint& p, V3 dir) {
    float cosine = dot(p.shading_normal, dir) / length(dir);
    if (e.kind == SSPT_EMITTER_GLOSSY) {
        float power = e.exponent + 1.0f;
        float jac = cosine > 0.0f ? ((power + 1.0f) * std::pow(cosine, power) * 1.0f / (2.0f * kPi)) : 0.0f;
        return e.inv_area * jac;
    }
    return e.inv_area * std::max(cosine, 0.0f) / kPi;
}

// One stored light-path vertex: the fields of PathVertex (PathVertex.cs:6-66)
// that ClassicBidir reads. The vertex at index k of a path has Depth k.
struct CachedVertex {
    SurfacePoint p;
    float pdf_from = 0.0f;  // PdfFromAncestor
    float pdf_rev = 0.0f;   // PdfReverseAncestor
    float pdf_ne = 0.0f;    // PdfNextEventAncestor
    Rgb weight = rgb(0.0f);
    int depth = 0;
    bool from_background = false;  // FromBackground; vertex 0 of such a path is the ray origin, on no mesh
};

// CameraPath.Vertices (BidirBase.Camera.cs:7-18)
struct PdfPair {
    float from = 0.0f, to = 0.0f;
};

// BidirPathPdfs (BidirPathPdfs.cs:9-114): per-vertex area pdfs of both sampling
// directions, reused across the calls of one thread.
struct Pdfs {
    std::vector<float> c2l, l2c;  // PdfsCameraToLight, PdfsLightToCamera
    float ne = 0.0f;              // PdfNextEvent
    void reset(int n) {
        c2l.assign((size_t)n, 0.0f);
        l2c.assign((size_t)n, 0.0f);
        ne = 0.0f;
    }
    int n() const { return (int)c2l.size(); }
    // GatherCameraPdfs (:46-53)
    void gather_camera(const std::vector<PdfPair>& v, int last) {
        for (int i = 0; i < last; ++i) {
            c2l[(size_t)i] = v[(size_t)i].from;
            if (i < last - 1) l2c[(size_t)i] = v[(size_t)i + 1].to;
        }
    }
    // GatherLightPdfs (:82-92): from `vertex` back along its path
    void gather_light(const std::vector<CachedVertex>& path, const CachedVertex& vertex, int last) {
        const CachedVertex* next = &vertex;
        for (int i = last + 1; i < n() - 2; ++i) {
            l2c[(size_t)i] = next->pdf_from;
            c2l[(size_t)i + 2] = next->pdf_rev;
            ne += next->pdf_ne;
            next = &path[(size_t)next->depth - 1];
        }
        l2c[(size_t)n() - 2] = next->pdf_from;
        l2c[(size_t)n() - 1] = 1.0f;
    }
};

struct BidirCtx {
    const or_scene* s;
    const sspt_bidir_settings* cfg;
    float num_paths;  // NumLightPaths
    const std::vector<std::vector<CachedVertex>>* cache;
    Counters* c;
    Pdfs pdfs;
};

// ClassicBidir.CameraPathReciprocals (ClassicBidir.cs:164-176)
float camera_reciprocals(const BidirCtx& bc, const Pdfs& p, int last) {
    float sum = 0.0f, next = 1.0f;
    for (int i = last; i > 0; --i) {
        next *= p.l2c[(size_t)i] / p.c2l[(size_t)i];
        if (bc.cfg->enable_connections) sum += next;
    }
    if (bc.cfg->enable_light_tracer) sum += next * p.l2c[0] / p.c2l[0] * bc.num_paths;
    return sum;
}
// ClassicBidir.LightPathReciprocals (ClassicBidir.cs:187-198)
float light_reciprocals(const BidirCtx& bc, const Pdfs& p, int last) {
    float sum = 0.0f, next = 1.0f;
    const int n = p.n();
    for (int i = last + 1; i < n; ++i) {
        if (i == n - 1) sum += next * p.ne / p.l2c[(size_t)i];
        next *= p.c2l[(size_t)i] / p.l2c[(size_t)i];
        if (bc.cfg->enable_connections && i < n - 2) sum += next;
    }
    sum += next;
    return sum;
}

// BidirBase.BackgroundProbability (BidirBase.Light.cs:132-133), which is also
// ComputeNextEventBackgroundProbability (BidirBase.Camera.cs:354-355)
float background_probability(const or_scene& s) {
    return s.has_env ? 1.0f / (1.0f + (float)s.emitters.size()) : 0.0f;
}

// ClassicBidir.NextEventPdf (ClassicBidir.cs:42-44; VertexCacheBidir.cs:51-53) over
// BidirBase.NextEventPdf (BidirBase.Camera.cs:338-347) of the vertex `to` seen from
// `from`: the background's direction pdf towards the (unnormalised) offset, or the
// emitter's area pdf.
float next_event_pdf(const BidirCtx& bc, V3 from, const SurfacePoint& to) {
    const or_scene& s = *bc.s;
    const float bg = background_probability(s);
    if (to.tri < 0) return env_pdf(s.env, to.position - from) * bg * (float)bc.cfg->num_shadow_rays;
    const Emitter& e = s.emitters[(size_t)s.tri_emitter[(size_t)to.tri]];
    float pmf = 1.0f / (float)s.emitters.size();
    return e.inv_area * pmf * (1.0f - bg) * (float)bc.cfg->num_shadow_rays;
}
// BidirBase.ComputeEmitterPdf (BidirBase.Light.cs:160-166), SelectLightPmf (:116-122)
float emitter_pdf(const BidirCtx& bc, const Emitter& e, const SurfacePoint& p, V3 light_to_surface,
                  float jacobian) {
    float pdf = emitter_pdf_ray(e, p, light_to_surface);
    pdf *= jacobian;
    pdf *= (1.0f - background_probability(*bc.s)) / (float)bc.s->emitters.size();
    return pdf;
}

// EnvironmentMap.SampleRay (EnvironmentMap.cs:87-103): a direction of the map and a
// point on the disc of radius SceneRadius about SceneCenter + SceneRadius * direction,
// orthogonal to it; the ray runs back along the direction.
struct EnvRay {
    Ray ray;
    Rgb weight;
    float pdf;
};
EnvRay env_sample_ray(const or_scene& s, V2 primary_pos, V2 primary_dir) {
    EnvSample ds = env_sample(s.env, primary_dir);
    V2 disc = to_concentric_disc(primary_pos);
    V3 t, b;
    basis(ds.dir, &t, &b);
    V3 pos = s.scene_center + s.scene_radius * (ds.dir + t * disc.x + b * disc.y);
    float pos_pdf = (1.0f / kPi) / (s.scene_radius * s.scene_radius);  // ToConcentricDiscJacobian / R^2
    EnvRay r;
    r.ray.o = pos;
    r.ray.d = -ds.dir;
    r.ray.tmin = 0.0f;
    r.ray.tmax = INFINITY;
    r.weight = ds.weight / pos_pdf;
    r.pdf = pos_pdf * ds.pdf;
    return r;
}
// BidirBase.ComputeBackgroundPdf (BidirBase.Light.cs:174-178) over EnvironmentMap.RayPdf
// (EnvironmentMap.cs:127-130): the pdf of the background ray along `light_to_surface`
float background_pdf(const or_scene& s, V3 light_to_surface) {
    float dir_pdf = env_pdf(s.env, -light_to_surface);
    float pos_pdf = (1.0f / kPi) / (s.scene_radius * s.scene_radius);
    float pdf = pos_pdf * dir_pdf;
    return pdf * background_probability(s);
}

// TraceLightPath (BidirBase.Light.cs:212-219): SelectLight (:103-111) picks the
// background with BackgroundProbability, else an emitter; TraceEmitterPath /
// TraceBackgroundPath (:268-299) with RandomWalk.StartFromEmitter /
// StartFromBackground / ContinueWalk (RandomWalk.cs:82-128,150-217) and LightPathWalk
// (BidirBase.Light.cs:324-410): the vertices of light path `idx` (none if the
// emitter or background sample is invalid).
void bidir_light_path(const BidirCtx& bc, uint32_t idx, uint32_t iteration, std::vector<CachedVertex>* out) {
    const or_scene& s = *bc.s;
    const sspt_bidir_settings& cfg = *bc.cfg;
    out->clear();
    Rng rng{hash_seed(cfg.base_seed_light, idx, iteration)};
    const float bg = background_probability(s);
    const bool from_bg = bg > 0.0f && rng.next_float() <= bg;
    Rgb prefix, approx = rgb(1.0f);
    Ray ray;
    SurfacePoint prev;
    float pdf_dir, next_rev = 0.0f;
    int depth = 1;
    if (from_bg) {
        // SampleBackground: primaryPos, then primaryDir
        V2 ppos = rng.next_float2();
        V2 pdir = rng.next_float2();
        EnvRay er = env_sample_ray(s, ppos, pdir);
        er.pdf *= bg;
        er.weight = er.weight / bg;
        if (er.pdf == 0.0f || is_black(er.weight)) return;
        CachedVertex v0;
        v0.p.position = er.ray.o;
        v0.from_background = true;
        out->push_back(v0);
        // StartFromBackground: the first hit samples its direction before it is
        // stored; its pdfs stay per solid angle and its weight is the initial one
        HitRec h = trace_closest(s, er.ray, false);
        bc.c->rays++;
        if (h.tri < 0) return;
        bc.c->hits++;
        SurfacePoint hit = make_point(s, h.tri, {h.u, h.v}, h.t);
        Shader sh = make_shader(s, hit, -er.ray.d);
        sh.light = true;
        float pc = rng.next_float();
        V2 pd = rng.next_float2();
        BsdfSample bs = sample_bsdf(sh, pc, pd);
        approx = approx * bs.weight;
        CachedVertex v;
        v.p = hit;
        v.pdf_from = er.pdf;
        v.pdf_rev = 0.0f;
        v.weight = er.weight;
        v.depth = 1;
        v.from_background = true;
        out->push_back(v);
        next_rev = bs.pdf_rev;  // OnContinue
        if (cfg.max_depth <= 1) return;
        if (bs.pdf == 0.0f || is_black(bs.weight)) return;
        prefix = er.weight * bs.weight;
        depth = 2;
        pdf_dir = bs.pdf;
        prev = hit;
        ray = spawn_ray(hit, bs.dir);
    } else {
        int count = (int)s.emitters.size();
        const Emitter& e = s.emitters[rng.next_int((uint32_t)count)];
        float sel = (1.0f - bg) / (float)count;
        V2 ppos = rng.next_float2();
        V2 pdir = rng.next_float2();
        EmitterRay er = emitter_sample_ray(s, e, ppos, pdir);
        er.pdf *= sel;
        if (er.pdf == 0.0f || is_black(er.weight)) return;
        prefix = er.weight / sel;
        CachedVertex v0;
        v0.p = er.p;
        out->push_back(v0);
        ray = spawn_ray(er.p, er.dir);
        prev = er.p;
        pdf_dir = er.pdf;
    }
    for (; depth < cfg.max_depth;) {
        HitRec h = trace_closest(s, ray, false);
        bc.c->rays++;
        if (h.tri < 0) break;
        bc.c->hits++;
        SurfacePoint hit = make_point(s, h.tri, {h.u, h.v}, h.t);
        float pdf_from = pdf_dir * area_to_solid_angle(prev, hit);
        if (pdf_from == 0.0f) break;
        // LightPathWalk.OnHit: the next-event pdf of the edge to the light is
        // stored with the vertex of depth 2
        CachedVertex v;
        v.p = hit;
        v.pdf_from = pdf_from;
        v.pdf_rev = next_rev;
        v.pdf_ne = depth == 2 ? next_event_pdf(bc, (*out)[1].p.position, (*out)[0].p) : 0.0f;
        v.weight = prefix;
        v.depth = depth;
        v.from_background = from_bg;
        out->push_back(v);
        if (depth + 1 >= cfg.max_depth) break;
        float q = depth > 4 ? std::min(std::max(average(approx), 0.05f), 0.95f) : 1.0f;
        if (rng.next_float() > q) break;
        Shader sh = make_shader(s, hit, -ray.d);
        sh.light = true;
        float pc = rng.next_float();
        V2 pd = rng.next_float2();
        BsdfSample bs = sample_bsdf(sh, pc, pd);
        approx = approx * (bs.weight / q);
        next_rev = bs.pdf_rev * area_to_solid_angle(hit, prev);  // OnContinue
        if (bs.pdf == 0.0f || is_black(bs.weight)) break;
        Rgb w = bs.weight * (std::fabs(dot(hit.normal, bs.dir)) / std::fabs(dot(hit.shading_normal, bs.dir)));
        w = w * (std::fabs(dot(hit.shading_normal, -ray.d)) / std::fabs(dot(hit.normal, -ray.d)));
        prefix = prefix * (w / q);
        ++depth;
        pdf_dir = bs.pdf;
        prev = hit;
        ray = spawn_ray(hit, bs.dir);
    }
}

// ConnectLightVertexToCamera (BidirBase.Light.cs:35-92) with
// ClassicBidir.LightTracerMis (ClassicBidir.cs:131-134).
void bidir_splat(BidirCtx& bc, const std::vector<CachedVertex>& path, size_t k, std::vector<Splat>* out) {
    const or_scene& s = *bc.s;
    const CachedVertex& v = path[k];
    const CachedVertex& anc = path[k - 1];
    V3 to_anc = normalize(anc.p.position - v.p.position);
    if (v.depth + 1 < bc.cfg->min_depth) return;
    int px, py;
    float jac;
    if (!s.camera.sample_response(v.p.position, v.p.normal, &px, &py, &jac)) return;
    bc.c->shadow++;
    if (trace_any(s, shadow_ray_to(v.p, s.camera.position), false)) {
        bc.c->occluded++;
        return;
    }
    V3 to_cam = s.camera.position - v.p.position;
    to_cam = to_cam / length(to_cam);
    Shader sh = make_shader(s, v.p, to_cam);
    Rgb f = eval_nocos(sh, to_anc);
    if (is_black(f)) return;
    f = f * (std::fabs(dot(v.p.shading_normal, to_anc)) / std::fabs(dot(v.p.normal, to_anc)));
    float pdf_rev, unused;
    pdf_pair(sh, to_anc, &pdf_rev, &unused);
    if (anc.p.tri >= 0) pdf_rev *= area_to_solid_angle(v.p, anc.p);  // not from the background
    Pdfs& p = bc.pdfs;
    p.reset(v.depth + 1);
    p.gather_light(path, v, -1);
    p.c2l[0] = 1.0f * jac;  // response.PdfEmit = invLensArea * jacobian
    p.c2l[1] = pdf_rev;
    if (v.depth == 1) p.ne = next_event_pdf(bc, v.p.position, anc.p);
    float mis = 1.0f / (1.0f + light_reciprocals(bc, p, -1) / bc.num_paths);
    Rgb w = v.weight * f * rgb(jac) / bc.num_paths;
    out->push_back({(int64_t)py * s.width + px, mis * w});
}

// BidirBase.OnEmitterHit (BidirBase.Camera.cs:500-524) with ClassicBidir.EmitterHitMis (ClassicBidir.cs:119-128)
Rgb bidir_emitter_hit(BidirCtx& bc, const Emitter& e, const SurfacePoint& hit, V3 out_dir,
                      const std::vector<PdfPair>& v, float rev_jac) {
    Rgb emission = emitted(e, hit, out_dir);
    float pdf_emit = emitter_pdf(bc, e, hit, out_dir, rev_jac);
    float pdf_ne = next_event_pdf(bc, V3{}, hit);  // NextEventPdf(new SurfacePoint(), hit)
    const int n = (int)v.size();
    Pdfs& p = bc.pdfs;
    p.reset(n);
    p.gather_camera(v, n - 1);
    if (n > 1) p.l2c[(size_t)n - 2] = pdf_emit;
    p.ne = pdf_ne;
    p.c2l[(size_t)n - 1] = v.back().from;
    float mis = 1.0f;
    if (n > 1) {
        float pdf_this = v.back().from;
        float sum = 1.0f;
        sum += p.ne / pdf_this;
        sum += camera_reciprocals(bc, p, n - 2) / pdf_this;
        mis = 1.0f / sum;
    }
    return mis * emission;
}

// BidirBase.OnBackgroundHit (BidirBase.Camera.cs:532-561) with ClassicBidir.EmitterHitMis:
// `v` ends with the escaped ray's entry (OnInvalidHit, :604-612: PdfFromAncestor is
// the direction's solid-angle pdf). No MinDepth test, as in the reference.
Rgb bidir_background_hit(BidirCtx& bc, V3 dir, const std::vector<PdfPair>& v) {
    const or_scene& s = *bc.s;
    float pdf_emit = background_pdf(s, -dir);
    float pdf_ne = env_pdf(s.env, dir) * background_probability(s);
    const int n = (int)v.size();
    Pdfs& p = bc.pdfs;
    p.reset(n);
    p.gather_camera(v, n - 1);
    if (n > 1) p.l2c[(size_t)n - 2] = pdf_emit;
    p.ne = pdf_ne;
    p.c2l[(size_t)n - 1] = v.back().from;
    float mis = 1.0f;
    if (n > 1) {
        float pdf_this = v.back().from;
        float sum = 1.0f;
        sum += p.ne / pdf_this;
        sum += camera_reciprocals(bc, p, n - 2) / pdf_this;
        mis = 1.0f / sum;
    }
    return mis * env_radiance(s.env, dir);
}

// PerformNextEventEstimation's background branch (BidirBase.Camera.cs:381-418): a
// direction of the map (its pdf without the NumShadowRays factor of the emitter
// branch), an any-hit ray to infinity (Raytracer.LeavesScene), the solid-angle pdfs
// of both directions. Fills the technique's pdfs in `p`; returns the unweighted
// sample (black if it contributes nothing).
Rgb background_next_event(const or_scene& s, Counters* c, const Shader& sh, Rng& rng, float bg, float rev_jac,
                          Pdfs& p, int n) {
    EnvSample smp = env_sample(s.env, rng.next_float2());
    smp.pdf *= bg;
    smp.weight = smp.weight / bg;
    if (smp.pdf == 0.0f) return rgb(0.0f);
    c->shadow++;
    if (trace_any(s, spawn_ray(sh.p, smp.dir), false)) {
        c->occluded++;
        return rgb(0.0f);
    }
    Rgb bsdf_cos = eval_cos(sh, smp.dir);
    float fwd, rev;
    pdf_pair(sh, smp.dir, &fwd, &rev);
    rev *= rev_jac;
    if (fwd == 0.0f || rev == 0.0f || is_black(bsdf_cos)) return rgb(0.0f);
    p.l2c[(size_t)n - 2] = background_pdf(s, -smp.dir);
    if (n > 2) p.l2c[(size_t)n - 3] = rev;
    p.ne = smp.pdf;
    p.c2l[(size_t)n - 1] = fwd;
    return smp.weight * bsdf_cos;
}

// BidirBase.PerformNextEventEstimation (BidirBase.Camera.cs:368-484), emitter
// branch, with ClassicBidir.SampleNextEvent / NextEventMis (ClassicBidir.cs:47-51,147-153).
Rgb bidir_next_event(BidirCtx& bc, const Shader& sh, Rng& rng, const std::vector<PdfPair>& v, float rev_jac) {
    const or_scene& s = *bc.s;
    const int n = (int)v.size() + 1;
    Pdfs& p = bc.pdfs;
    p.reset(n);
    p.gather_camera(v, n - 2);
    p.c2l[(size_t)n - 2] = v.back().from;
    const float bg = background_probability(s);  // ComputeNextEventBackgroundProbability
    if (rng.next_float() < bg) {
        Rgb w = background_next_event(s, bc.c, sh, rng, bg, rev_jac, p, n);
        if (is_black(w)) return rgb(0.0f);
        float sum = 1.0f;  // ClassicBidir.NextEventMis (ClassicBidir.cs:147-153)
        sum += p.c2l[(size_t)n - 1] / p.ne;
        sum += camera_reciprocals(bc, p, (int)v.size() - 1) / p.ne;
        return (1.0f / sum) * w;
    }
    if (s.emitters.empty()) return rgb(0.0f);
    // SampleNextEvent: SelectLight (BidirBase.Camera.cs:70-73), SampleUniformArea
    int count = (int)s.emitters.size();
    const Emitter& e = s.emitters[rng.next_int((uint32_t)count)];
    float prob = 1.0f / (float)count;
    SurfacePoint lp = make_point(s, e.tri, to_uniform_triangle(rng.next_float2()), 0.0f);
    float lpdf = e.inv_area;
    lpdf *= prob;
    lpdf *= (float)bc.cfg->num_shadow_rays;
    lpdf *= (1.0f - bg);
    if (lpdf == 0.0f) return rgb(0.0f);
    bc.c->shadow++;
    if (trace_any(s, shadow_ray(sh.p, lp), false)) {
        bc.c->occluded++;
        return rgb(0.0f);
    }
    V3 l2s = normalize(sh.p.position - lp.position);
    Rgb emission = emitted(e, lp, l2s);
    if (is_black(emission)) return rgb(0.0f);
    Rgb bsdf_cos = eval_cos(sh, -l2s);
    if (is_black(bsdf_cos)) return rgb(0.0f);
    float jac = area_to_solid_angle(sh.p, lp);
    if (jac == 0.0f) return rgb(0.0f);
    float fwd, rev;
    pdf_pair(sh, -l2s, &fwd, &rev);
    fwd *= area_to_solid_angle(sh.p, lp);
    rev *= rev_jac;
    if (fwd == 0.0f || rev == 0.0f) return rgb(0.0f);
    float pdf_emit = emitter_pdf(bc, e, lp, l2s, area_to_solid_angle(lp, sh.p));
    p.l2c[(size_t)n - 2] = pdf_emit;
    if (n > 2) p.l2c[(size_t)n - 3] = rev;
    p.ne = lpdf;
    p.c2l[(size_t)n - 1] = fwd;
    float sum = 1.0f;
    sum += p.c2l[(size_t)n - 1] / p.ne;
    sum += camera_reciprocals(bc, p, (int)v.size() - 1) / p.ne;
    float mis = 1.0f / sum;
    Rgb w = emission * bsdf_cos * (jac / lpdf);
    return mis * w;
}

// BidirBase.Connect (BidirBase.Camera.cs:192-270) with ClassicBidir.BidirConnectMis (ClassicBidir.cs:137-144)
Rgb bidir_connect(BidirCtx& bc, const Shader& sh, const std::vector<CachedVertex>& path, size_t k,
                  const std::vector<PdfPair>& v, float rev_jac) {
    const or_scene& s = *bc.s;
    const CachedVertex& lv = path[k];
    const CachedVertex& anc = path[k - 1];
    V3 to_anc = normalize(anc.p.position - lv.p.position);
    int depth = lv.depth + (int)v.size() + 1;
    if (depth > bc.cfg->max_depth || depth < bc.cfg->min_depth) return rgb(0.0f);
    bc.c->shadow++;
    if (trace_any(s, shadow_ray(lv.p, sh.p), false)) {
        bc.c->occluded++;
        return rgb(0.0f);
    }
    V3 dir = normalize(lv.p.position - sh.p.position);  // camera vertex -> light vertex
    Shader ls = make_shader(s, lv.p, to_anc);
    ls.light = true;
    Rgb f_light = eval_nocos(ls, -dir) * std::fabs(dot(lv.p.normal, -dir));
    f_light = f_light * (std::fabs(dot(lv.p.shading_normal, to_anc)) / std::fabs(dot(lv.p.normal, to_anc)));
    Rgb f_cam = eval_cos(sh, dir);
    if (is_black(f_cam) || is_black(f_light)) return rgb(0.0f);
    float pdf_cam_to_light, pdf_cam_rev;
    pdf_pair(sh, dir, &pdf_cam_to_light, &pdf_cam_rev);
    pdf_cam_rev *= rev_jac;
    pdf_cam_to_light *= area_to_solid_angle(sh.p, lv.p);
    if (pdf_cam_to_light == 0.0f) return rgb(0.0f);
    float pdf_light_to_cam, pdf_light_rev;
    pdf_pair(ls, -dir, &pdf_light_to_cam, &pdf_light_rev);
    if (anc.p.tri >= 0) pdf_light_rev *= area_to_solid_angle(lv.p, anc.p);  // not from the background
    pdf_light_to_cam *= area_to_solid_angle(lv.p, sh.p);
    const int last = (int)v.size() - 1;
    Pdfs& p = bc.pdfs;
    p.reset((int)v.size() + lv.depth + 1);
    p.gather_camera(v, last);
    p.gather_light(path, lv, last);
    if (lv.depth == 1) p.ne = next_event_pdf(bc, lv.p.position, anc.p);
    if (last > 0) p.l2c[(size_t)last - 1] = pdf_cam_rev;
    p.c2l[(size_t)last] = v.back().from;
    p.l2c[(size_t)last] = pdf_light_to_cam;
    p.c2l[(size_t)last + 1] = pdf_cam_to_light;
    p.c2l[(size_t)last + 2] = pdf_light_rev;
    float sum = 1.0f;
    sum += camera_reciprocals(bc, p, last);
    sum += light_reciprocals(bc, p, last);
    float mis = 1.0f / sum;
    V3 d = sh.p.position - lv.p.position;
    float dist2 = dot(d, d);
    if (dist2 == 0.0f) return rgb(0.0f);
    Rgb w = lv.weight * f_light * f_cam / dist2 / 1.0f;
    return mis * w;
}

// BidirBase.RenderPixel (BidirBase.Camera.cs:97-106) and the camera's RandomWalk
// (RandomWalk.StartFromCamera / ContinueWalk, RandomWalk.cs:73-80,150-217;
// CameraRandomWalk, BidirBase.Camera.cs:588-653) with ClassicBidir.OnCameraHit
// (ClassicBidir.cs:94-116): the pixel value of one camera path.
Rgb bidir_camera_path(BidirCtx& bc, uint32_t row, uint32_t col, uint32_t iteration) {
    const or_scene& s = *bc.s;
    const sspt_bidir_settings& cfg = *bc.cfg;
    const uint32_t pixel = row * (uint32_t)s.width + col;
    Rng rng{hash_seed(cfg.base_seed_camera, pixel, iteration)};
    V2 off = rng.next_float2();
    Ray ray;
    V3 raw;
    s.camera.generate_ray((float)col + off.x, (float)row + off.y, &ray.o, &ray.d, &raw);
    ray.tmin = 0.0f;
    ray.tmax = INFINITY;
    // CameraRaySample: Point (position, normal = view direction), PdfRay
    SurfacePoint prev;
    prev.position = s.camera.position;
    prev.normal = s.camera.direction();
    float pdf_dir = s.camera.pixel_jacobian(ray.o + raw);
    const std::vector<CachedVertex>& light = (*bc.cache)[pixel];  // SelectBidirPath (:179-184)
    std::vector<PdfPair> v;
    Rgb prefix = rgb(1.0f), approx = rgb(1.0f), L = rgb(0.0f);
    for (int depth = 1; depth < cfg.max_depth + 1;) {
        HitRec h = trace_closest(s, ray, false);
        bc.c->rays++;
        if (h.tri < 0) {
            if (s.has_env) {  // CameraRandomWalk.OnInvalidHit (BidirBase.Camera.cs:604-612)
                v.push_back({pdf_dir, 0.0f});
                L = L + prefix * bidir_background_hit(bc, ray.d, v);
            }
            break;
        }
        bc.c->hits++;
        SurfacePoint hit = make_point(s, h.tri, {h.u, h.v}, h.t);
        float pdf_from = pdf_dir * area_to_solid_angle(prev, hit);
        if (pdf_from == 0.0f) break;
        float rev_jac = area_to_solid_angle(hit, prev);
        Shader sh = make_shader(s, hit, -ray.d);
        v.push_back({pdf_from, 0.0f});
        Rgb value = rgb(0.0f);
        int ei = s.tri_emitter[(size_t)h.tri];
        if (ei >= 0 && depth >= cfg.min_depth)
            value = value + prefix * bidir_emitter_hit(bc, s.emitters[(size_t)ei], hit, -ray.d, v, rev_jac);
        if (depth < cfg.max_depth && cfg.enable_connections) {
            Rgb conn = rgb(0.0f);
            for (size_t k = 1; k < light.size(); ++k) conn = conn + bidir_connect(bc, sh, light, k, v, rev_jac);
            value = value + prefix * conn;
        }
        if (depth < cfg.max_depth && depth + 1 >= cfg.min_depth)
            for (int i = 0; i < cfg.num_shadow_rays; ++i)
                value = value + prefix * bidir_next_event(bc, sh, rng, v, rev_jac);
        L = L + value;
        if (depth + 1 >= cfg.max_depth + 1) break;
        float q = depth > 4 ? std::min(std::max(average(approx), 0.05f), 0.95f) : 1.0f;
        if (rng.next_float() > q) break;
        float pc = rng.next_float();
        V2 pd = rng.next_float2();
        BsdfSample bs = sample_bsdf(sh, pc, pd);
        approx = approx * (bs.weight / q);
        v.back().to = bs.pdf_rev * rev_jac;  // OnContinue
        if (bs.pdf == 0.0f || is_black(bs.weight)) break;
        prefix = prefix * (bs.weight / q);
        ++depth;
        pdf_dir = bs.pdf;
        prev = hit;
        ray = spawn_ray(hit, bs.dir);
    }
    return L;
}

}  // namespace

uint64_t or_render_bidir(const or_scene* s, const sspt_bidir_settings* cfg, uint32_t first, uint32_t num,
                         float* image, sspt_ray_stats* stats, int32_t threads) {
    const int64_t W = s->width, H = s->height;
    if (cfg->num_light_paths >= 0 && (int64_t)cfg->num_light_paths != W * H) return UINT64_MAX;
    const uint32_t paths = (uint32_t)(W * H);
    std::atomic<uint64_t> rays{0}, shadow{0}, occl{0}, hits{0};
    uint64_t nans = 0;
    std::vector<std::vector<CachedVertex>> cache(paths);
    std::vector<std::vector<Splat>> splats(paths);
    std::vector<Rgb> values((size_t)(W * H));
    const int64_t chunks = ((int64_t)paths + 63) / 64;
    for (uint32_t it = first; it < first + num; ++it) {
        // FrameBuffer.StartIteration / Normalize (FrameBuffer.cs:232-249)
        float n = (float)(it + 1);
        if (it + 1 > 1) {
            float scale = (n - 1.0f) / n;
            for (int64_t k = 0; k < W * H * 3; ++k) image[k] *= scale;
        }
        auto run = [&](int phase) {
            parallel_for(chunks, threads, [&](int64_t ch) {
                Counters c;
                BidirCtx bc{s, cfg, (float)paths, &cache, &c, {}};
                int64_t end = std::min<int64_t>(paths, (ch + 1) * 64);
                for (int64_t k = ch * 64; k < end; ++k) {
                    if (phase == 0) {
                        bidir_light_path(bc, (uint32_t)k, it, &cache[(size_t)k]);
                    } else if (phase == 1) {
                        splats[(size_t)k].clear();
                        for (size_t i = 1; i < cache[(size_t)k].size(); ++i)
                            bidir_splat(bc, cache[(size_t)k], i, &splats[(size_t)k]);
                    } else {
                        values[(size_t)k] = bidir_camera_path(bc, (uint32_t)(k / W), (uint32_t)(k % W), it);
                    }
                }
                rays += c.rays, shadow += c.shadow, occl += c.occluded, hits += c.hits;
            });
        };
        run(0);  // TraceLightPaths
        if (cfg->enable_light_tracer) run(1);  // ProcessPathCache: SplatLightVertices
        run(2);  // TraceAllCameraPaths
        // FrameBuffer.Splat (FrameBuffer.cs:160-191), IgnoreNanAndInf
        auto add = [&](int64_t pix, Rgb value) {
            if (!std::isfinite(average(value))) {
                ++nans;
                return;
            }
            Rgb a = value / n;
            image[pix * 3 + 0] += a.x;
            image[pix * 3 + 1] += a.y;
            image[pix * 3 + 2] += a.z;
        };
        if (cfg->enable_light_tracer)
            for (uint32_t k = 0; k < paths; ++k)
                for (const Splat& sp : splats[k]) add(sp.pix, sp.value);
        for (int64_t k = 0; k < W * H; ++k) add(k, values[(size_t)k]);
    }
    if (stats) {
        stats->num_rays = rays;
        stats->num_shadow_rays = shadow;
        stats->num_occluded = occl;
        stats->num_ray_hits = hits;
        stats->num_shadow_rays_traced = shadow;
        stats->num_iterations = num;
    }
    return nans;
}

// ----------------------------- light vertex cache and vertex merging (VCM) -
namespace {

struct VcmCtx {
    const or_scene* s;
    const sspt_vcm_settings* cfg;
    sspt_bidir_settings light_cfg;  // the fields the shared light-path and next-event helpers read
    bool vcm;                       // VertexConnectionAndMerging semantics
    float num_paths;                // NumLightPaths
    const std::vector<std::vector<CachedVertex>>* cache;
    const std::vector<int64_t>* cum;  // PathCache.cumPathLen: inclusive prefix sums of the path lengths
    int64_t num_vertices;             // PathCache.NumVertices
    Counters* c;
    Pdfs pdfs;
    // the photon map (NearestNeighborSearch over the positions of the light vertices of
    // depth >= 1 with non-zero weight, in path order) as a uniform grid
    const std::vector<V3>* photon_pos;
    const std::vector<std::pair<int32_t, int32_t>>* photon_ref;  // (path, vertex)
    const std::unordered_map<int64_t, std::vector<int32_t>>* grid;
    float cell;
    BidirCtx light() const { return BidirCtx{s, &light_cfg, num_paths, cache, c, {}}; }
};

// VertexCacheBidir.BidirSelectDensity (VertexCacheBidir.cs:71-82)
float select_density(const VcmCtx& vc) {
    if (vc.num_vertices == 0) return 0.0f;
    float select_prob = 1.0f / (float)vc.num_vertices;
    float num_samples = (float)(vc.cfg->num_connections * (int32_t)vc.num_paths);
    return select_prob * num_samples;
}

// VertexConnectionAndMerging.CorrelAwareRatios (VertexConnectionAndMerging.cs:580-646)
struct CorrelRatios {
    std::vector<float> c2l, l2c;
    void init(const Pdfs& p, float dist_to_cam, bool from_background) {
        float radius = dist_to_cam * 0.0174550649f;  // distance * tan(1 degree)
        float accept_area = radius * radius * kPi;
        int nsv = p.n() - 1;
        c2l.assign((size_t)std::max(nsv, 0), 0.0f);
        l2c.assign((size_t)std::max(nsv, 0), 0.0f);
        float product = 1.0f;
        for (int i = 0; i < nsv; ++i) {
            product *= std::min(p.c2l[(size_t)i] * accept_area, 1.0f);
            c2l[(size_t)i] = product;
        }
        product = 1.0f;
        for (int i = nsv - 1; i >= 0; --i) {
            float next = p.l2c[(size_t)i] * accept_area;
            if (i == nsv - 1 && !from_background) next *= accept_area;
            product *= std::min(next, 1.0f);
            l2c[(size_t)i] = product;
        }
    }
    float operator[](int idx) const {
        if (idx == 0) return 1.0f;  // primary merges are not correlated
        float light = l2c[(size_t)idx], cam = c2l[(size_t)idx];
        if (cam == 0.0f && light == 0.0f) return 1.0f;
        return cam / (cam + light - cam * light);
    }
};

// CameraPathReciprocals: VertexCacheBidir.cs:210-223, VertexConnectionAndMerging.cs:839-889
float vcm_camera_reciprocals(const VcmCtx& vc, const Pdfs& p, int last, float radius, const CorrelRatios& cr) {
    const sspt_vcm_settings& cfg = *vc.cfg;
    const bool merging = vc.vcm && cfg.enable_merging;
    float sum = 0.0f, next = 1.0f;
    for (int i = last; i > 0; --i) {
        if (merging) {
            float accept = p.l2c[(size_t)i] * kPi * radius * radius;
            if (!cfg.disable_correl_aware_mis) accept *= cr[i];
            sum += next * vc.num_paths * accept;
        }
        next *= p.l2c[(size_t)i] / p.c2l[(size_t)i];
        if (cfg.num_connections > 0) sum += next * select_density(vc);
    }
    if (cfg.enable_light_tracer) sum += next * p.l2c[0] / p.c2l[0] * vc.num_paths;
    if (merging && cfg.merge_primary) sum += next * vc.num_paths * p.l2c[0] * kPi * radius * radius;
    return sum;
}

// LightPathReciprocals: VertexCacheBidir.cs:225-237, VertexConnectionAndMerging.cs:891-929
float vcm_light_reciprocals(const VcmCtx& vc, const Pdfs& p, int last, float radius, const CorrelRatios& cr) {
    const sspt_vcm_settings& cfg = *vc.cfg;
    const bool merging = vc.vcm && cfg.enable_merging;
    const int n = p.n();
    float sum = 0.0f, next = 1.0f;
    for (int i = last + 1; i < n; ++i) {
        if (i == n - 1) sum += next * p.ne / p.l2c[(size_t)i];
        if (merging && i < n - 1 && (cfg.merge_primary || i > 0)) {
            float accept = p.c2l[(size_t)i] * kPi * radius * radius;
            if (!cfg.disable_correl_aware_mis) accept *= cr[i];
            sum += next * vc.num_paths * accept;
        }
        next *= p.c2l[(size_t)i] / p.l2c[(size_t)i];
        if (i < n - 2 && cfg.num_connections > 0) sum += next * select_density(vc);
    }
    if (cfg.enable_hitting) sum += next;
    return sum;
}

// ConnectLightVertexToCamera (BidirBase.Light.cs:35-92) with LightTracerMis
// (VertexCacheBidir.cs:183-186, VertexConnectionAndMerging.cs:735-766)
void vcm_splat(VcmCtx& vc, const std::vector<CachedVertex>& path, size_t k, std::vector<Splat>* out) {
    const or_scene& s = *vc.s;
    const CachedVertex& v = path[k];
    const CachedVertex& anc = path[k - 1];
    V3 to_anc = normalize(anc.p.position - v.p.position);
    if (v.depth + 1 < vc.cfg->min_depth) return;
    int px, py;
    float jac;
    if (!s.camera.sample_response(v.p.position, v.p.normal, &px, &py, &jac)) return;
    vc.c->shadow++;
    if (trace_any(s, shadow_ray_to(v.p, s.camera.position), false)) {
        vc.c->occluded++;
        return;
    }
    V3 to_cam = s.camera.position - v.p.position;
    float dist = length(to_cam);
    to_cam = to_cam / dist;
    Shader sh = make_shader(s, v.p, to_cam);
    Rgb f = eval_nocos(sh, to_anc);
    if (is_black(f)) return;
    f = f * (std::fabs(dot(v.p.shading_normal, to_anc)) / std::fabs(dot(v.p.normal, to_anc)));
    float pdf_rev, unused;
    pdf_pair(sh, to_anc, &pdf_rev, &unused);
    if (anc.p.tri >= 0) pdf_rev *= area_to_solid_angle(v.p, anc.p);  // not from the background
    Pdfs& p = vc.pdfs;
    p.reset(v.depth + 1);
    p.gather_light(path, v, -1);
    p.c2l[0] = 1.0f * jac;
    p.c2l[1] = pdf_rev;
    if (v.depth == 1) p.ne = next_event_pdf(vc.light(), v.p.position, anc.p);
    float mis;
    if (vc.vcm) {
        CorrelRatios cr;
        cr.init(p, dist, v.from_background);
        float footprint = std::sqrt(1.0f / p.c2l[0]);
        float sum = vcm_light_reciprocals(vc, p, -1, footprint, cr);
        sum /= vc.num_paths;
        sum += 1.0f;
        mis = 1.0f / sum;
    } else {
        CorrelRatios none;
        mis = 1.0f / (1.0f + vcm_light_reciprocals(vc, p, -1, 0.0f, none) / vc.num_paths);
    }
    Rgb w = v.weight * f * rgb(jac) / vc.num_paths;
    out->push_back({(int64_t)py * s.width + px, mis * w});
}

// The camera path's state the techniques read (CameraPath, BidirBase.Camera.cs:24-68)
struct VcmPath {
    std::vector<PdfPair> v;
    float footprint = 0.0f;  // FootprintRadius
    float dist0 = 0.0f;      // Distances[0]
};

// OnEmitterHit (BidirBase.Camera.cs:500-524) with EmitterHitMis (VertexCacheBidir.cs:171-180,
// VertexConnectionAndMerging.cs:700-732)
Rgb vcm_emitter_hit(VcmCtx& vc, const Emitter& e, const SurfacePoint& hit, V3 out_dir, const VcmPath& cp,
                    float rev_jac) {
    Rgb emission = emitted(e, hit, out_dir);
    BidirCtx bc = vc.light();
    float pdf_emit = emitter_pdf(bc, e, hit, out_dir, rev_jac);
    float pdf_ne = next_event_pdf(bc, V3{}, hit);  // NextEventPdf(new SurfacePoint(), hit)
    const int n = (int)cp.v.size();
    Pdfs& p = vc.pdfs;
    p.reset(n);
    p.gather_camera(cp.v, n - 1);
    if (n > 1) p.l2c[(size_t)n - 2] = pdf_emit;
    p.ne = pdf_ne;
    p.c2l[(size_t)n - 1] = cp.v.back().from;
    float mis = 1.0f;
    if (n > 1) {
        CorrelRatios cr;
        if (vc.vcm) cr.init(p, cp.dist0, false);
        float pdf_this = p.c2l[(size_t)n - 1];
        float sum = 1.0f;
        sum += p.ne / pdf_this;
        sum += vcm_camera_reciprocals(vc, p, n - 2, cp.footprint, cr) / pdf_this;
        mis = 1.0f / sum;
    }
    return mis * emission;
}

// OnBackgroundHit (BidirBase.Camera.cs:532-561) with EmitterHitMis (isBackground)
Rgb vcm_background_hit(VcmCtx& vc, V3 dir, const VcmPath& cp) {
    const or_scene& s = *vc.s;
    float pdf_emit = background_pdf(s, -dir);
    float pdf_ne = env_pdf(s.env, dir) * background_probability(s);
    const int n = (int)cp.v.size();
    Pdfs& p = vc.pdfs;
    p.reset(n);
    p.gather_camera(cp.v, n - 1);
    if (n > 1) p.l2c[(size_t)n - 2] = pdf_emit;
    p.ne = pdf_ne;
    p.c2l[(size_t)n - 1] = cp.v.back().from;
    float mis = 1.0f;
    if (n > 1) {
        CorrelRatios cr;
        if (vc.vcm) cr.init(p, cp.dist0, true);
        float pdf_this = p.c2l[(size_t)n - 1];
        float sum = 1.0f;
        sum += p.ne / pdf_this;
        sum += vcm_camera_reciprocals(vc, p, n - 2, cp.footprint, cr) / pdf_this;
        mis = 1.0f / sum;
    }
    return mis * env_radiance(s.env, dir);
}

// PerformNextEventEstimation (BidirBase.Camera.cs:368-484) with NextEventMis
// (VertexCacheBidir.cs:200-207, VertexConnectionAndMerging.cs:802-836)
Rgb vcm_next_event(VcmCtx& vc, const Shader& sh, Rng& rng, const VcmPath& cp, float rev_jac) {
    const or_scene& s = *vc.s;
    const int n = (int)cp.v.size() + 1;
    Pdfs& p = vc.pdfs;
    p.reset(n);
    p.gather_camera(cp.v, n - 2);
    p.c2l[(size_t)n - 2] = cp.v.back().from;
    const float bg = background_probability(s);
    if (rng.next_float() < bg) {
        Rgb w = background_next_event(s, vc.c, sh, rng, bg, rev_jac, p, n);
        if (is_black(w)) return rgb(0.0f);
        CorrelRatios cr;
        if (vc.vcm) cr.init(p, cp.dist0, true);
        float sum = 1.0f;
        if (vc.cfg->enable_hitting) sum += p.c2l[(size_t)n - 1] / p.ne;
        sum += vcm_camera_reciprocals(vc, p, (int)cp.v.size() - 1, cp.footprint, cr) / p.ne;
        return (1.0f / sum) * w;
    }
    if (s.emitters.empty()) return rgb(0.0f);
    int count = (int)s.emitters.size();
    const Emitter& e = s.emitters[rng.next_int((uint32_t)count)];
    float prob = 1.0f / (float)count;
    SurfacePoint lp = make_point(s, e.tri, to_uniform_triangle(rng.next_float2()), 0.0f);
    float lpdf = e.inv_area;
    lpdf *= prob;
    lpdf *= (float)vc.cfg->num_shadow_rays;
    lpdf *= (1.0f - bg);
    if (lpdf == 0.0f) return rgb(0.0f);
    vc.c->shadow++;
    if (trace_any(s, shadow_ray(sh.p, lp), false)) {
        vc.c->occluded++;
        return rgb(0.0f);
    }
    V3 l2s = normalize(sh.p.position - lp.position);
    Rgb emission = emitted(e, lp, l2s);
    if (is_black(emission)) return rgb(0.0f);
    Rgb bsdf_cos = eval_cos(sh, -l2s);
    if (is_black(bsdf_cos)) return rgb(0.0f);
    float jac = area_to_solid_angle(sh.p, lp);
    if (jac == 0.0f) return rgb(0.0f);
    float fwd, rev;
    pdf_pair(sh, -l2s, &fwd, &rev);
    fwd *= area_to_solid_angle(sh.p, lp);
    rev *= rev_jac;
    if (fwd == 0.0f || rev == 0.0f) return rgb(0.0f);
    float pdf_emit = emitter_pdf(vc.light(), e, lp, l2s, area_to_solid_angle(lp, sh.p));
    p.l2c[(size_t)n - 2] = pdf_emit;
    if (n > 2) p.l2c[(size_t)n - 3] = rev;
    p.ne = lpdf;
    p.c2l[(size_t)n - 1] = fwd;
    CorrelRatios cr;
    if (vc.vcm) cr.init(p, cp.dist0, false);
    float sum = 1.0f;
    if (vc.cfg->enable_hitting) sum += p.c2l[(size_t)n - 1] / p.ne;
    sum += vcm_camera_reciprocals(vc, p, (int)cp.v.size() - 1, cp.footprint, cr) / p.ne;
    float mis = 1.0f / sum;
    Rgb w = emission * bsdf_cos * (jac / lpdf);
    return mis * w;
}

// PathCache.GetVertex (PathCache.cs:40-56): the global index in path order
std::pair<size_t, size_t> global_vertex(const VcmCtx& vc, int64_t g) {
    const std::vector<int64_t>& cum = *vc.cum;
    size_t path = (size_t)(std::upper_bound(cum.begin(), cum.end(), g) - cum.begin());
    int64_t first = path == 0 ? 0 : cum[path - 1];
    return {path, (size_t)(g - first)};
}

// BidirConnections with a vertex from VertexSelector (BidirBase.Camera.cs:279-295) and
// Connect (:192-270) with BidirConnectMis (VertexCacheBidir.cs:189-197, VertexConnectionAndMerging.cs:768-800)
Rgb vcm_connect(VcmCtx& vc, const Shader& sh, Rng& rng, const VcmPath& cp, float rev_jac) {
    const or_scene& s = *vc.s;
    int64_t g = (int64_t)rng.next_int((uint32_t)vc.num_vertices);
    float density = select_density(vc);
    if (g <= 0) return rgb(0.0f);
    auto [pi, k] = global_vertex(vc, g);
    const std::vector<CachedVertex>& path = (*vc.cache)[pi];
    const CachedVertex& lv = path[k];
    if (lv.depth < 1) return rgb(0.0f);
    const CachedVertex& anc = path[k - 1];
    V3 to_anc = normalize(anc.p.position - lv.p.position);
    int depth = lv.depth + (int)cp.v.size() + 1;
    if (depth > vc.cfg->max_depth || depth < vc.cfg->min_depth) return rgb(0.0f);
    vc.c->shadow++;
    if (trace_any(s, shadow_ray(lv.p, sh.p), false)) {
        vc.c->occluded++;
        return rgb(0.0f);
    }
    V3 dir = normalize(lv.p.position - sh.p.position);
    Shader ls = make_shader(s, lv.p, to_anc);
    ls.light = true;
    Rgb f_light = eval_nocos(ls, -dir) * std::fabs(dot(lv.p.normal, -dir));
    f_light = f_light * (std::fabs(dot(lv.p.shading_normal, to_anc)) / std::fabs(dot(lv.p.normal, to_anc)));
    Rgb f_cam = eval_cos(sh, dir);
    if (is_black(f_cam) || is_black(f_light)) return rgb(0.0f);
    float pcl, pcr;
    pdf_pair(sh, dir, &pcl, &pcr);
    pcr *= rev_jac;
    pcl *= area_to_solid_angle(sh.p, lv.p);
    if (pcl == 0.0f) return rgb(0.0f);
    float plc, plr;
    pdf_pair(ls, -dir, &plc, &plr);
    if (anc.p.tri >= 0) plr *= area_to_solid_angle(lv.p, anc.p);  // not from the background
    plc *= area_to_solid_angle(lv.p, sh.p);
    const int last = (int)cp.v.size() - 1;
    Pdfs& p = vc.pdfs;
    p.reset((int)cp.v.size() + lv.depth + 1);
    p.gather_camera(cp.v, last);
    p.gather_light(path, lv, last);
    if (lv.depth == 1) p.ne = next_event_pdf(vc.light(), lv.p.position, anc.p);
    if (last > 0) p.l2c[(size_t)last - 1] = pcr;
    p.c2l[(size_t)last] = cp.v.back().from;
    p.l2c[(size_t)last] = plc;
    p.c2l[(size_t)last + 1] = pcl;
    p.c2l[(size_t)last + 2] = plr;
    CorrelRatios cr;
    if (vc.vcm) cr.init(p, cp.dist0, lv.from_background);
    float sum = 1.0f;
    sum += vcm_camera_reciprocals(vc, p, last, cp.footprint, cr) / density;
    sum += vcm_light_reciprocals(vc, p, last, cp.footprint, cr) / density;
    float mis = 1.0f / sum;
    V3 d = sh.p.position - lv.p.position;
    float dist2 = dot(d, d);
    if (dist2 == 0.0f) return rgb(0.0f);
    Rgb w = lv.weight * f_light * f_cam / dist2 / density;
    return mis * w;
}

// VertexConnectionAndMerging.Merge (VertexConnectionAndMerging.cs:329-425) with MergeMis (:648-697)
Rgb vcm_merge(VcmCtx& vc, const Shader& sh, const VcmPath& cp, float rev_jac, int32_t photon, float dist_sqr,
              float radius_sqr) {
    auto [pi, k] = (*vc.photon_ref)[(size_t)photon];
    const std::vector<CachedVertex>& path = (*vc.cache)[(size_t)pi];
    const CachedVertex& ph = path[(size_t)k];
    int depth = (int)cp.v.size() + ph.depth;
    if (depth > vc.cfg->max_depth || depth < vc.cfg->min_depth) return rgb(0.0f);
    if (std::fabs(dot(sh.p.normal, ph.p.normal)) < 0.4f) return rgb(0.0f);
    const CachedVertex& anc = path[(size_t)k - 1];
    V3 to_anc = normalize(anc.p.position - sh.p.position);
    Rgb bsdf = eval_nocos(sh, to_anc);
    bsdf = bsdf * (std::fabs(dot(sh.p.shading_normal, to_anc)) / std::fabs(dot(ph.p.normal, to_anc)));
    Rgb contrib = ph.weight * bsdf / vc.num_paths;
    if (is_black(contrib)) return rgb(0.0f);
    if (std::fabs(dot(to_anc, sh.p.normal)) < 1e-4f) return rgb(0.0f);
    float pdf_light_rev, pdf_cam_rev;
    pdf_pair(sh, to_anc, &pdf_light_rev, &pdf_cam_rev);
    pdf_cam_rev *= rev_jac;
    // at the first hit from the background the pdf stays per solid angle
    if (!(ph.depth == 1 && ph.from_background)) pdf_light_rev *= area_to_solid_angle(sh.p, anc.p);
    const int last = (int)cp.v.size() - 1;
    Pdfs& p = vc.pdfs;
    p.reset((int)cp.v.size() + ph.depth);
    p.gather_camera(cp.v, last);
    p.gather_light(path, ph, last - 1);
    if (last > 0) p.l2c[(size_t)last - 1] = pdf_cam_rev;
    p.l2c[(size_t)last] = ph.pdf_from;
    p.c2l[(size_t)last] = cp.v.back().from;
    p.c2l[(size_t)last + 1] = pdf_light_rev;
    if (ph.depth == 1) p.ne = next_event_pdf(vc.light(), sh.p.position, anc.p);
    // MergeMis
    float mis;
    {
        const sspt_vcm_settings& cfg = *vc.cfg;
        float radius = cp.footprint;
        float approx = p.l2c[(size_t)last] * kPi * radius * radius * vc.num_paths;
        CorrelRatios cr;
        cr.init(p, cp.dist0, ph.from_background);
        if (!cfg.disable_correl_aware_mis) approx *= cr[last];
        if (approx == 0.0f) {
            mis = 0.0f;
        } else {
            float sum = 0.0f;
            sum += vcm_camera_reciprocals(vc, p, last, radius, cr) / approx;
            sum += vcm_light_reciprocals(vc, p, last, radius, cr) / approx;
            if (ph.depth > 1 && cfg.num_connections > 0) sum += select_density(vc) / approx;
            mis = 1.0f / sum;
        }
    }
    if (pdf_cam_rev == 0.0f || pdf_light_rev == 0.0f) return rgb(0.0f);
    float kernel = 2.0f * (radius_sqr - dist_sqr) / (kPi * radius_sqr * radius_sqr);
    return contrib * kernel * mis;
}

// PerformMerging (VertexConnectionAndMerging.cs:469-516): the merges with the
// max_num_photons nearest photons within the footprint radius.
Rgb vcm_merging(VcmCtx& vc, const Shader& sh, const VcmPath& cp, float rev_jac) {
    const sspt_vcm_settings& cfg = *vc.cfg;
    if (!cfg.enable_merging) return rgb(0.0f);
    if (!cfg.merge_primary && cp.v.size() == 1) return rgb(0.0f);
    float radius = cp.footprint;  // ComputeLocalMergeRadius
    float r2 = radius * radius;
    V3 q = sh.p.position;
    // NearestNeighborSearch.ForAllNearest(q, MaxNumPhotons, radius): the photons with the
    // smallest squared distances below radius^2, ties to the earlier photon
    std::vector<std::pair<float, int32_t>> found;
    const float h = vc.cell;
    const double span = 2.0 * (double)radius / (double)h + 2.0;
    if (!(span * span * span <= (double)std::max<size_t>(vc.grid->size(), 1))) {
        // a query box of more cells than there are occupied cells (or a non-finite radius): every photon
        for (size_t ph = 0; ph < vc.photon_pos->size(); ++ph) {
            V3 d = (*vc.photon_pos)[ph] - q;
            float d2 = d.x * d.x + d.y * d.y + d.z * d.z;
            if (d2 < r2) found.push_back({d2, (int32_t)ph});
        }
    }
    int64_t lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
    if (found.empty() && span * span * span <= (double)std::max<size_t>(vc.grid->size(), 1))
        for (int a = 0; a < 3; ++a) {
            lo[a] = (int64_t)std::floor((get(q, a) - radius) / h);
            hi[a] = (int64_t)std::floor((get(q, a) + radius) / h);
        }
    for (int64_t x = lo[0]; x <= hi[0]; ++x)
        for (int64_t y = lo[1]; y <= hi[1]; ++y)
            for (int64_t z = lo[2]; z <= hi[2]; ++z) {
                auto it = vc.grid->find((x * 73856093) ^ (y * 19349663) ^ (z * 83492791));
                if (it == vc.grid->end()) continue;
                for (int32_t ph : it->second) {
                    V3 pp = (*vc.photon_pos)[(size_t)ph];
                    if ((int64_t)std::floor(pp.x / h) != x || (int64_t)std::floor(pp.y / h) != y ||
                        (int64_t)std::floor(pp.z / h) != z)
                        continue;
                    V3 d = pp - q;
                    float d2 = d.x * d.x + d.y * d.y + d.z * d.z;
                    if (d2 < r2) found.push_back({d2, ph});
                }
            }
    std::sort(found.begin(), found.end());
    const int k = cfg.max_num_photons;
    if ((int)found.size() > k) found.resize((size_t)k);
    const int num_found = (int)found.size();
    Rgb est = rgb(0.0f);
    if (num_found == 0) return est;
    float furthest = std::sqrt(found.back().first);
    for (const auto& f : found) {
        float distance = std::sqrt(f.first);
        float radius_sqr = num_found == k ? furthest * furthest : r2;  // MergeHelper (:494-516)
        est = est + vcm_merge(vc, sh, cp, rev_jac, f.second, distance * distance, radius_sqr);
    }
    return est;
}

// BidirBase.RenderPixel (BidirBase.Camera.cs:97-106), the camera's RandomWalk and
// OnCameraHit of VertexCacheBidir (VertexCacheBidir.cs:143-168) /
// VertexConnectionAndMerging (VertexConnectionAndMerging.cs:525-572).
Rgb vcm_camera_path(VcmCtx& vc, uint32_t row, uint32_t col, uint32_t iteration) {
    const or_scene& s = *vc.s;
    const sspt_vcm_settings& cfg = *vc.cfg;
    const uint32_t pixel = row * (uint32_t)s.width + col;
    Rng rng{hash_seed(cfg.base_seed_camera, pixel, iteration)};
    V2 off = rng.next_float2();
    Ray ray;
    V3 raw;
    s.camera.generate_ray((float)col + off.x, (float)row + off.y, &ray.o, &ray.d, &raw);
    ray.tmin = 0.0f;
    ray.tmax = INFINITY;
    SurfacePoint prev;
    prev.position = s.camera.position;
    prev.normal = s.camera.direction();
    float pdf_dir = s.camera.pixel_jacobian(ray.o + raw);
    VcmPath cp;
    Rgb prefix = rgb(1.0f), approx = rgb(1.0f), L = rgb(0.0f);
    for (int depth = 1; depth < cfg.max_depth + 1;) {
        HitRec h = trace_closest(s, ray, false);
        vc.c->rays++;
        if (h.tri < 0) {
            if (s.has_env) {  // CameraRandomWalk.OnInvalidHit (BidirBase.Camera.cs:604-612)
                cp.v.push_back({pdf_dir, 0.0f});
                L = L + prefix * vcm_background_hit(vc, ray.d, cp);
            }
            break;
        }
        vc.c->hits++;
        SurfacePoint hit = make_point(s, h.tri, {h.u, h.v}, h.t);
        float pdf_from = pdf_dir * area_to_solid_angle(prev, hit);
        if (pdf_from == 0.0f) break;
        float rev_jac = area_to_solid_angle(hit, prev);
        Shader sh = make_shader(s, hit, -ray.d);
        if (depth == 1) {
            cp.footprint = std::sqrt(1.0f / pdf_from);
            cp.dist0 = hit.distance;
        }
        cp.v.push_back({pdf_from, 0.0f});
        Rgb value = rgb(0.0f);
        int ei = s.tri_emitter[(size_t)h.tri];
        bool hitting = vc.vcm ? (cfg.enable_hitting || depth == 1) : (bool)cfg.enable_hitting;
        if (ei >= 0 && hitting && depth >= cfg.min_depth)
            value = value + prefix * vcm_emitter_hit(vc, s.emitters[(size_t)ei], hit, -ray.d, cp, rev_jac);
        if (depth < cfg.max_depth) {
            for (int i = 0; i < cfg.num_connections; ++i) value = value + prefix * vcm_connect(vc, sh, rng, cp, rev_jac);
            if (vc.vcm) value = value + prefix * vcm_merging(vc, sh, cp, rev_jac);
        }
        if (depth < cfg.max_depth && depth + 1 >= cfg.min_depth)
            for (int i = 0; i < cfg.num_shadow_rays; ++i) value = value + prefix * vcm_next_event(vc, sh, rng, cp, rev_jac);
        L = L + value;
        if (depth + 1 >= cfg.max_depth + 1) break;
        float q = depth > 4 ? std::min(std::max(average(approx), 0.05f), 0.95f) : 1.0f;
        if (rng.next_float() > q) break;
        float pc = rng.next_float();
        V2 pd = rng.next_float2();
        BsdfSample bs = sample_bsdf(sh, pc, pd);
        approx = approx * (bs.weight / q);
        cp.v.back().to = bs.pdf_rev * rev_jac;
        if (bs.pdf == 0.0f || is_black(bs.weight)) break;
        prefix = prefix * (bs.weight / q);
        ++depth;
        pdf_dir = bs.pdf;
        prev = hit;
        ray = spawn_ray(hit, bs.dir);
    }
    return L;
}

}  // namespace

uint64_t or_render_vcm(const or_scene* s, const sspt_vcm_settings* cfg, float cell_size, uint32_t first,
                       uint32_t num, float* image, sspt_ray_stats* stats, int32_t threads) {
    const int64_t W = s->width, H = s->height;
    if (cfg->algorithm != SSPT_VERTEX_CACHE_BIDIR && cfg->algorithm != SSPT_VCM) return UINT64_MAX;
    if (cfg->algorithm == SSPT_VCM && (cfg->max_num_photons < 1 || cfg->max_num_photons > 32)) return UINT64_MAX;
    const uint32_t paths = cfg->num_light_paths < 0 ? (uint32_t)(W * H) : (uint32_t)cfg->num_light_paths;
    sspt_bidir_settings lcfg{};
    lcfg.base_seed_light = cfg->base_seed_light;
    lcfg.max_depth = cfg->max_depth;
    lcfg.num_shadow_rays = cfg->num_shadow_rays;
    std::atomic<uint64_t> rays{0}, shadow{0}, occl{0}, hits{0};
    uint64_t nans = 0;
    std::vector<std::vector<CachedVertex>> cache(paths);
    std::vector<std::vector<Splat>> splats(paths);
    std::vector<Rgb> values((size_t)(W * H));
    std::vector<int64_t> cum(paths);
    std::vector<V3> photon_pos;
    std::vector<std::pair<int32_t, int32_t>> photon_ref;
    std::unordered_map<int64_t, std::vector<int32_t>> grid;
    const bool vcm = cfg->algorithm == SSPT_VCM;
    for (uint32_t it = first; it < first + num; ++it) {
        float n = (float)(it + 1);
        if (it + 1 > 1) {
            float scale = (n - 1.0f) / n;
            for (int64_t k = 0; k < W * H * 3; ++k) image[k] *= scale;
        }
        int64_t num_vertices = 0;
        auto run = [&](int phase, int64_t count) {
            parallel_for((count + 63) / 64, threads, [&](int64_t ch) {
                Counters c;
                VcmCtx vc{s, cfg, lcfg, vcm, (float)paths, &cache, &cum, num_vertices, &c, {},
                          &photon_pos, &photon_ref, &grid, cell_size};
                BidirCtx bc = vc.light();
                int64_t end = std::min<int64_t>(count, (ch + 1) * 64);
                for (int64_t k = ch * 64; k < end; ++k) {
                    if (phase == 0) {
                        bidir_light_path(bc, (uint32_t)k, it, &cache[(size_t)k]);
                    } else if (phase == 1) {
                        splats[(size_t)k].clear();
                        for (size_t i = 1; i < cache[(size_t)k].size(); ++i)
                            vcm_splat(vc, cache[(size_t)k], i, &splats[(size_t)k]);
                    } else {
                        values[(size_t)k] = vcm_camera_path(vc, (uint32_t)(k / W), (uint32_t)(k % W), it);
                    }
                }
                rays += c.rays, shadow += c.shadow, occl += c.occluded, hits += c.hits;
            });
        };
        run(0, paths);  // TraceLightPaths
        // PathCache.Prepare (PathCache.cs:87-94)
        int64_t sum = 0;
        for (uint32_t k = 0; k < paths; ++k) {
            sum += (int64_t)cache[k].size();
            cum[k] = sum;
        }
        num_vertices = sum;
        // ProcessPathCache: the photon map (VertexConnectionAndMerging.cs:297-316)
        photon_pos.clear();
        photon_ref.clear();
        grid.clear();
        if (vcm && cfg->enable_merging) {
            for (uint32_t k = 0; k < paths; ++k)
                for (size_t i = 1; i < cache[k].size(); ++i) {
                    if (is_black(cache[k][i].weight)) continue;
                    V3 pp = cache[k][i].p.position;
                    int32_t id = (int32_t)photon_pos.size();
                    photon_pos.push_back(pp);
                    photon_ref.push_back({(int32_t)k, (int32_t)i});
                    int64_t x = (int64_t)std::floor(pp.x / cell_size), y = (int64_t)std::floor(pp.y / cell_size),
                            z = (int64_t)std::floor(pp.z / cell_size);
                    grid[(x * 73856093) ^ (y * 19349663) ^ (z * 83492791)].push_back(id);
                }
        }
        if (cfg->enable_light_tracer) run(1, paths);
        run(2, W * H);
        auto add = [&](int64_t pix, Rgb value) {
            if (!std::isfinite(average(value))) {
                ++nans;
                return;
            }
            Rgb a = value / n;
            image[pix * 3 + 0] += a.x;
            image[pix * 3 + 1] += a.y;
            image[pix * 3 + 2] += a.z;
        };
        if (cfg->enable_light_tracer)
            for (uint32_t k = 0; k < paths; ++k)
                for (const Splat& sp : splats[k]) add(sp.pix, sp.value);
        for (int64_t k = 0; k < W * H; ++k) add(k, values[(size_t)k]);
    }
    if (stats) {
        stats->num_rays = rays;
        stats->num_shadow_rays = shadow;
        stats->num_occluded = occl;
        stats->num_ray_hits = hits;
        stats->num_shadow_rays_traced = shadow;
        stats->num_iterations = num;
    }
    return nans;
}

// System.Collections.Generic.PriorityQueue<TElement, float> of .NET 6+ (the
// runtime SeeSharp targets; not part of the reference's sources): an array
// quaternary min-heap. Enqueue appends and moves the entry up while its
// priority is strictly smaller than its parent's ((i - 1) / 4);
// EnqueueDequeue returns the new entry at once unless its priority is strictly
// larger than the root's, otherwise it replaces the root and moves down: into
// the first strictly smallest of the children 4i+1..4i+4 while that child's
// priority is strictly smaller. UnorderedItems enumerates the array in order.
struct QuaternaryHeap {
    std::vector<std::pair<float, int32_t>> a;  // (priority, element)
    void enqueue(float p, int32_t e) {
        size_t i = a.size();
        a.push_back({p, e});
        while (i > 0) {
            size_t parent = (i - 1) >> 2;
            if (!(p < a[parent].first)) break;
            a[i] = a[parent];
            i = parent;
        }
        a[i] = {p, e};
    }
    void enqueue_dequeue(float p, int32_t e) {
        if (a.empty() || !(p > a[0].first)) return;
        size_t i = 0;
        for (;;) {
            size_t c = (i << 2) + 1;
            if (c >= a.size()) break;
            size_t mc = c;
            for (size_t k = c + 1; k < std::min(c + 4, a.size()); ++k)
                if (a[k].first < a[mc].first) mc = k;
            if (p <= a[mc].first) break;
            a[i] = a[mc];
            i = mc;
        }
        a[i] = {p, e};
    }
};

// FrameBuffer's OutlierReplayCache (OutlierReplayCache.cs:9-22), fed by every
// finite splat (FrameBuffer.cs:166-189) with Weight = the sample and
// Iteration = CurIteration - 1, priority Weight.Average: Enqueue while the
// queue holds fewer than n entries, else EnqueueDequeue. Writes per pixel the
// queue's entries in WriteToFile's order (UnorderedItems stably sorted by
// descending priority, FrameBuffer.cs:321-325), weight 0 / iteration -1 past
// the queue's size.
void or_render_outliers(const or_scene* s, const sspt_pt_settings* cfg, uint32_t first, uint32_t num, int32_t n,
                        float* weights, int32_t* iterations, int32_t threads) {
    const int64_t total = (int64_t)s->width * s->height;
    parallel_for((total + 63) / 64, threads, [&](int64_t ch) {
        Counters c;
        PathCtx pc{s, cfg, false, &c};
        const int64_t end = std::min<int64_t>(total, (ch + 1) * 64);
        for (int64_t pix = ch * 64; pix < end; ++pix) {
            QuaternaryHeap q;
            for (uint32_t it = first; it < first + num; ++it) {
                Rgb L = render_pixel(pc, (uint32_t)(pix / s->width), (uint32_t)(pix % s->width), it);
                float w = average(L);
                if (!std::isfinite(w)) continue;  // Splat returns before Notify (IgnoreNanAndInf)
                if ((int32_t)q.a.size() < n) q.enqueue(w, (int32_t)it);
                else q.enqueue_dequeue(w, (int32_t)it);
            }
            std::vector<std::pair<float, int32_t>> out = q.a;
            std::stable_sort(out.begin(), out.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
            for (int32_t j = 0; j < n; ++j) {
                const bool have = j < (int32_t)out.size();
                weights[pix * n + j] = have ? out[j].first : 0.0f;
                iterations[pix * n + j] = have ? out[j].second : -1;
            }
        }
    });
}

// DebugVisualizer.Render (DebugVisualizer.cs:20-32), RenderPixel (:47-59) and
// ComputeColor (:37-45): per iteration and pixel the jittered primary ray's
// hit, |dot(face normal, -dir)| / |normal| / |dir| (0 on a miss), splatted
// into the running-mean image (FrameBuffer.StartIteration / Splat).
void or_render_debug(const or_scene* s, uint32_t base_seed, uint32_t first, uint32_t num, float* image,
                     int32_t threads) {
    const int64_t total = (int64_t)s->width * s->height;
    for (uint32_t it = first; it < first + num; ++it) {
        const float n = (float)(it + 1);
        parallel_for((total + 63) / 64, threads, [&](int64_t ch) {
            const int64_t end = std::min<int64_t>(total, (ch + 1) * 64);
            for (int64_t pix = ch * 64; pix < end; ++pix) {
                Rng rng{hash_seed(base_seed, (uint32_t)pix, it)};
                V2 off = rng.next_float2();
                Ray ray;
                s->camera.generate_ray((float)(pix % s->width) + off.x, (float)(pix / s->width) + off.y, &ray.o,
                                       &ray.d);
                ray.tmin = 0.0f;
                ray.tmax = INFINITY;
                HitRec h = trace_closest(*s, ray, false);
                float value = 0.0f;
                if (h.tri >= 0) {
                    SurfacePoint hit = make_point(*s, h.tri, {h.u, h.v}, h.t);
                    V3 from = -ray.d;
                    value = std::fabs(dot(hit.normal, from));
                    value /= std::sqrt(dot(hit.normal, hit.normal));
                    value /= std::sqrt(dot(from, from));
                }
                for (int c = 0; c < 3; ++c) {
                    if (it + 1 > 1) image[pix * 3 + c] *= (n - 1.0f) / n;
                    image[pix * 3 + c] += value / n;
                }
            }
        });
    }
}

void or_render_pixel(const or_scene* s, const sspt_pt_settings* cfg, uint32_t row, uint32_t col,
                     uint32_t iteration, float* out) {
    Counters c;
    PathCtx pc{s, cfg, false, &c};
    Rgb L = render_pixel(pc, row, col, iteration);
    out[0] = L.x, out[1] = L.y, out[2] = L.z;
}

// DenoiseBuffers.LogPrimaryHit (PathTracer.cs:336-339, DenoiseBuffers.cs:20-23): at the
// primary hit, Material.GetScatterStrength (the base colour lookup, GenericMaterial.cs:65,
// DiffuseMaterial.cs:37-38) and the shading normal, splatted into running-mean layers
// (RgbLayer.Splat: value / n; Layer.OnStartIteration: scale by (n - 1) / n). A primary
// miss splats nothing. Only the RNG's jitter draws and the primary ray are involved.
void or_primary_aov(const or_scene* s, const sspt_pt_settings* cfg, uint32_t first, uint32_t num, float* albedo,
                    float* normal, int32_t threads) {
    int W = s->width, H = s->height;
    int64_t total = (int64_t)W * H;
    for (uint32_t it = first; it < first + num; ++it) {
        float n = (float)(it + 1);
        int64_t chunks = (total + 63) / 64;
        parallel_for(chunks, threads, [&](int64_t ch) {
            int64_t end = std::min<int64_t>(total, (ch + 1) * 64);
            for (int64_t pix = ch * 64; pix < end; ++pix) {
                uint32_t row = (uint32_t)(pix / W), col = (uint32_t)(pix % W);
                float* a = albedo + pix * 3;
                float* nm = normal + pix * 3;
                if (it + 1 > 1) {
                    float scale = (n - 1.0f) / n;
                    for (int c = 0; c < 3; ++c) a[c] *= scale, nm[c] *= scale;
                }
                Rng rng{hash_seed(cfg->base_seed, row * (uint32_t)W + col, it)};
                V2 off = rng.next_float2();
                Ray ray;
                s->camera.generate_ray((float)col + off.x, (float)row + off.y, &ray.o, &ray.d);
                ray.tmin = 0.0f;
                ray.tmax = INFINITY;
                HitRec h = trace_closest(*s, ray, false);
                if (h.tri < 0) continue;
                SurfacePoint hit = make_point(*s, h.tri, {h.u, h.v}, h.t);
                const Material& m = s->materials[s->tris[h.tri].material];
                Rgb alb = lookup_rgb(s->textures, m.base_tex, m.base_color, hit.uv);
                V3 sn = hit.shading_normal;
                Rgb av = alb / n;
                V3 nv = sn / n;
                a[0] += av.x, a[1] += av.y, a[2] += av.z;
                nm[0] += nv.x, nm[1] += nv.y, nm[2] += nv.z;
            }
        });
    }
}

int32_t or_set_texture(or_scene* s, uint32_t index, const float* texels) {
    if (index >= s->textures.size()) return 1;
    Texture& t = s->textures[index];
    std::copy(texels, texels + t.texels.size(), t.texels.begin());
    return 0;
}

// PathTracerBase.ReplayPixel (PathTracer.cs:61-69): the PathGraph and the
// estimate of one path. Writes at most `capacity` nodes; returns the count.
uint32_t or_replay_pixel(const or_scene* s, const sspt_pt_settings* cfg, uint32_t col, uint32_t row,
                         uint32_t iteration, sspt_path_node* nodes, uint32_t capacity, float* estimate) {
    Counters c;
    PathCtx pc{s, cfg, false, &c};
    const TextureSet* sets[1] = {&s->textures};
    Graph g;
    Rgb L;
    render_pixel_k<1>(pc, sets, row, col, iteration, &L, &g);
    for (uint32_t i = 0; i < std::min<uint32_t>(capacity, (uint32_t)g.nodes.size()); ++i) nodes[i] = g.nodes[i];
    estimate[0] = L.x;
    estimate[1] = L.y;
    estimate[2] = L.z;
    return (uint32_t)g.nodes.size();
}

uint64_t or_render_tangent(const or_scene* s, const sspt_pt_settings* cfg, uint32_t first, uint32_t num,
                           uint32_t texture_index, const float* tangent, float eps, float* dimage,
                           const int32_t* pixels, int64_t num_pixels, int32_t threads) {
    if (texture_index >= s->textures.size() || num == 0) return 0;
    // texels + eps * tangent and texels - eps * tangent
    TextureSet plus = s->textures, minus = s->textures;
    std::vector<float>& tp = plus[texture_index].texels;
    std::vector<float>& tm = minus[texture_index].texels;
    for (size_t k = 0; k < tp.size(); ++k) {
        tp[k] = tp[k] + eps * tangent[k];
        tm[k] = tm[k] - eps * tangent[k];
    }
    const TextureSet* sets[3] = {&s->textures, &plus, &minus};
    int W = s->width, H = s->height;
    int64_t total = pixels ? num_pixels : (int64_t)W * H;
    std::atomic<uint64_t> nans{0};
    int64_t chunks = (total + 15) / 16;
    parallel_for(chunks, threads, [&](int64_t ch) {
        Counters c;
        PathCtx pc{s, cfg, false, &c};
        int64_t end = std::min<int64_t>(total, (ch + 1) * 16);
        for (int64_t k = ch * 16; k < end; ++k) {
            int64_t pix = pixels ? pixels[k] : k;
            uint32_t row = (uint32_t)(pix / W), col = (uint32_t)(pix % W);
            double acc[3] = {0.0, 0.0, 0.0};
            for (uint32_t it = first; it < first + num; ++it) {
                Rgb L[3];
                render_pixel_k<3>(pc, sets, row, col, it, L);
                if (!std::isfinite(average(L[0]))) {  // skipped by the FrameBuffer
                    nans++;
                    continue;
                }
                acc[0] += ((double)L[1].x - (double)L[2].x) / (2.0 * eps);
                acc[1] += ((double)L[1].y - (double)L[2].y) / (2.0 * eps);
                acc[2] += ((double)L[1].z - (double)L[2].z) / (2.0 * eps);
            }
            for (int c3 = 0; c3 < 3; ++c3) dimage[pix * 3 + c3] = (float)(acc[c3] / (double)num);
        }
    });
    return nans;
}

}  // extern "C"
