// api.hip -- the C ABI of include/sspt.h: contexts, scenes (Scene.Prepare),
// framebuffers and the render loop of PathTracerBase.Render
// (SeeSharp/Integrators/PathTracer.cs:199-267).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sspt.h"
#include "bvh_build.h"
#include "kernels.h"
#include "scene_types.h"

using namespace sspt;

namespace {

struct Timing {
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    bool enabled = false;
    std::vector<hipEvent_t> pool;
    std::vector<Pending> pending;
    std::map<std::string, std::pair<uint64_t, double>> totals;

    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        hipEventCreate(&e);
        return e;
    }
    // Folds finished launches into the totals (stream synchronized before).
    void collect() {
        for (auto& p : pending) {
            float ms = 0.0f;
            hipEventElapsedTime(&ms, p.a, p.b);
            auto& t = totals[p.name];
            t.first += 1;
            t.second += ms;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    ~Timing() {
        for (auto& p : pending) {
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        for (auto e : pool) hipEventDestroy(e);
    }
};

}  // namespace

struct sspt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string error;
    Timing timing;
    // wavefront buffers, grown on demand
    Wave wave{};
    uint32_t wave_capacity = 0;
    std::vector<void*> wave_allocs;
    uint32_t* host_counters = nullptr;  // pinned
    size_t shadow_capacity = 0;         // shadow queue entries allocated
};

struct sspt_scene {
    sspt_ctx* ctx = nullptr;
    DevScene dev{};
    std::vector<void*> allocs;
    float* image = nullptr;  // W*H*3 running mean
    uint64_t num_tris = 0, num_emitters = 0, num_nodes = 0, bytes = 0;
    uint64_t nan_count = 0;
    int bvh_depth = 0;
};

namespace {

sspt_status fail(sspt_ctx* ctx, sspt_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->error = buf;
    return st;
}

#define HIP_TRY(ctx, expr)                                                                       \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess)                                                                    \
            return fail((ctx), _e == hipErrorOutOfMemory ? SSPT_ERR_OUT_OF_MEMORY : SSPT_ERR_DEVICE, \
                        "%s failed: %s", #expr, hipGetErrorString(_e));                          \
    } while (0)

template <class T>
sspt_status upload(sspt_ctx* ctx, sspt_scene* sc, const T* host, size_t count, const T** out) {
    if (count == 0) {
        *out = nullptr;
        return SSPT_OK;
    }
    void* p = nullptr;
    HIP_TRY(ctx, hipMalloc(&p, count * sizeof(T)));
    sc->allocs.push_back(p);
    sc->bytes += count * sizeof(T);
    HIP_TRY(ctx, hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice));
    *out = static_cast<const T*>(p);
    return SSPT_OK;
}

#define TRY(expr)                        \
    do {                                 \
        sspt_status _s = (expr);         \
        if (_s != SSPT_OK) return _s;    \
    } while (0)

// General 4x4 inverse by cofactors in double, rounded to float (the rule the
// test oracle follows too). Returns false for a singular matrix.
bool invert4(const float* a, float* out) {
    double m[16], inv[16];
    for (int i = 0; i < 16; ++i) m[i] = a[i];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
             m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
             m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
             m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
              m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
             m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
             m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
             m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
              m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
             m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
             m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
              m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
              m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
             m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
             m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
              m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
              m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0.0) return false;
    for (int i = 0; i < 16; ++i) out[i] = (float)(inv[i] / det);
    return true;
}

// PerspectiveCamera ctor + UpdateResolution (SeeSharp/Cameras/PerspectiveCamera.cs:29-52)
// and Camera.WorldToCamera (SeeSharp/Cameras/Camera.cs:34-42).
bool setup_camera(const sspt_camera_desc& cd, int w, int h, DevCamera* cam) {
    if (!invert4(cd.world_to_camera, cam->c2w)) return false;
    const float pi = 3.14159265358979323846f;
    float fov = cd.vertical_fov_degrees * pi / 180.0f;
    float aspect = (float)w / (float)h;
    // Matrix4x4.CreatePerspectiveFieldOfView(fov, aspect, 0.001, 1000), right-handed
    float p[16] = {0};
    float ys = 1.0f / std::tan(fov * 0.5f);
    p[0] = ys / aspect;
    p[5] = ys;
    float range = 1000.0f / (0.001f - 1000.0f);
    p[10] = range;
    p[11] = -1.0f;
    p[14] = range * 0.001f;
    if (!invert4(p, cam->v2c)) return false;
    const float* W = cam->c2w;
    // Transform((0,0,0), cameraToWorld) = translation row
    cam->pos[0] = 0.0f * W[0] + 0.0f * W[4] + 0.0f * W[8] + W[12];
    cam->pos[1] = 0.0f * W[1] + 0.0f * W[5] + 0.0f * W[9] + W[13];
    cam->pos[2] = 0.0f * W[2] + 0.0f * W[6] + 0.0f * W[10] + W[14];
    cam->width = w;
    cam->height = h;
    return true;
}

sspt_status ensure_wave(sspt_ctx* ctx, uint32_t capacity) {
    if (ctx->wave_capacity >= capacity) return SSPT_OK;
    for (void* p : ctx->wave_allocs) hipFree(p);
    ctx->wave_allocs.clear();
    ctx->wave_capacity = 0;
    ctx->shadow_capacity = 0;
    ctx->wave.s_o = ctx->wave.s_d = ctx->wave.s_c = nullptr;
    auto alloc = [&](void** p, size_t bytes) -> sspt_status {
        HIP_TRY(ctx, hipMalloc(p, bytes));
        ctx->wave_allocs.push_back(*p);
        return SSPT_OK;
    };
    Wave& w = ctx->wave;
    size_t P = capacity;
    // every queue holds at most one entry per slot; the shadow queue two per
    // slot and shadow ray (environment + area light)
    for (int b = 0; b < 2; ++b) {
        TRY(alloc((void**)&w.q_o[b], P * 16));
        TRY(alloc((void**)&w.q_d[b], P * 16));
        TRY(alloc((void**)&w.q_slot[b], P * 4));
    }
    TRY(alloc((void**)&w.hits, P * 16));
    TRY(alloc((void**)&w.st_rng, P * 4));
    TRY(alloc((void**)&w.st_thr, P * 16));
    TRY(alloc((void**)&w.st_prev, P * 16));
    TRY(alloc((void**)&w.st_L, P * 16));
    TRY(alloc((void**)&w.st_Le, P * 16));
    TRY(alloc((void**)&w.st_La, P * 16));
    TRY(alloc((void**)&w.counters, 64));
    TRY(alloc((void**)&w.stats, 64));
    w.capacity = capacity;
    w.s_o = w.s_d = w.s_c = nullptr;
    ctx->wave_capacity = capacity;
    return SSPT_OK;
}

sspt_status ensure_shadow(sspt_ctx* ctx, size_t entries, size_t* have) {
    if (*have >= entries) return SSPT_OK;
    Wave& w = ctx->wave;
    for (void* p : {(void*)w.s_o, (void*)w.s_d, (void*)w.s_c}) {
        if (!p) continue;
        hipFree(p);
        ctx->wave_allocs.erase(std::remove(ctx->wave_allocs.begin(), ctx->wave_allocs.end(), p),
                               ctx->wave_allocs.end());
    }
    for (float4** p : {&w.s_o, &w.s_d, &w.s_c}) {
        HIP_TRY(ctx, hipMalloc((void**)p, entries * 16));
        ctx->wave_allocs.push_back(*p);
    }
    *have = entries;
    return SSPT_OK;
}

template <class F>
void timed(sspt_ctx* ctx, const char* name, F&& launch) {
    if (!ctx->timing.enabled) {
        launch();
        return;
    }
    hipEvent_t a = ctx->timing.get(), b = ctx->timing.get();
    hipEventRecord(a, ctx->stream);
    launch();
    hipEventRecord(b, ctx->stream);
    ctx->timing.pending.push_back({name, a, b});
}

}  // namespace

extern "C" {

int32_t sspt_abi_version(void) { return SSPT_ABI_VERSION; }

sspt_status sspt_ctx_create(int32_t device, sspt_ctx** out) {
    if (!out) return SSPT_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return SSPT_ERR_DEVICE;
    if (device < 0 || device >= n) return SSPT_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return SSPT_ERR_DEVICE;
    sspt_ctx* ctx = new (std::nothrow) sspt_ctx();
    if (!ctx) return SSPT_ERR_OUT_OF_MEMORY;
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return SSPT_ERR_DEVICE;
    }
    if (hipHostMalloc((void**)&ctx->host_counters, 64, hipHostMallocDefault) != hipSuccess) {
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return SSPT_ERR_DEVICE;
    }
    *out = ctx;
    return SSPT_OK;
}

void sspt_ctx_destroy(sspt_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    ctx->timing.collect();
    for (void* p : ctx->wave_allocs) hipFree(p);
    if (ctx->host_counters) hipHostFree(ctx->host_counters);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* sspt_last_error(const sspt_ctx* ctx) { return ctx ? ctx->error.c_str() : "null context"; }

sspt_status sspt_scene_create(sspt_ctx* ctx, const sspt_scene_desc* d, sspt_scene** out) {
    if (!ctx || !d || !out) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    hipSetDevice(ctx->device);
    // Scene.IsValid (SeeSharp/Scene.cs:170-197)
    if (d->width == 0 || d->height == 0) return fail(ctx, SSPT_ERR_INVALID_SCENE, "Framebuffer not set.");
    if (d->num_meshes == 0 || !d->meshes) return fail(ctx, SSPT_ERR_INVALID_SCENE, "No meshes in the scene.");
    for (uint32_t i = 0; i < d->num_meshes; ++i) {
        const sspt_mesh_desc& m = d->meshes[i];
        if (m.material < 0 || (uint32_t)m.material >= d->num_materials)
            return fail(ctx, SSPT_ERR_INVALID_SCENE, "Mesh[%u] does not have a material.", i);
        if (m.num_triangles && (!m.positions || !m.indices))
            return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "Mesh[%u] has no vertex or index data", i);
        for (uint64_t k = 0; k < 3ull * m.num_triangles; ++k)
            if (m.indices[k] < 0 || (uint32_t)m.indices[k] >= m.num_vertices)
                return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "Mesh[%u] index %llu out of range", i,
                            (unsigned long long)k);
    }
    if (d->num_emitters == 0 && !d->background)
        return fail(ctx, SSPT_ERR_INVALID_SCENE, "No emitters and no background in the scene.");
    for (uint32_t i = 0; i < d->num_materials; ++i) {
        const sspt_material_desc& m = d->materials[i];
        if (m.kind != SSPT_MATERIAL_DIFFUSE && m.kind != SSPT_MATERIAL_GENERIC)
            return fail(ctx, SSPT_ERR_UNSUPPORTED, "material %u: unknown kind %d", i, m.kind);
        for (int t : {m.base_color_texture, m.roughness_texture})
            if (t >= (int)d->num_textures) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "material %u: bad texture", i);
        if (m.base_color_texture >= 0 && d->textures[m.base_color_texture].channels != 3)
            return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "material %u: base color texture must be RGB", i);
        if (m.roughness_texture >= 0 && d->textures[m.roughness_texture].channels != 1)
            return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "material %u: roughness texture must be mono", i);
    }

    std::unique_ptr<sspt_scene> sc(new sspt_scene());
    sc->ctx = ctx;

    // flatten triangles in global order (mesh offset + PrimId)
    std::vector<uint32_t> mesh_offset(d->num_meshes);
    uint64_t ntri = 0;
    for (uint32_t i = 0; i < d->num_meshes; ++i) {
        mesh_offset[i] = (uint32_t)ntri;
        ntri += d->meshes[i].num_triangles;
    }
    if (ntri >= (1u << 30)) return fail(ctx, SSPT_ERR_UNSUPPORTED, "too many triangles");
    std::vector<float> verts(ntri * 9);
    std::vector<uint32_t> tri_mesh(ntri);
    for (uint32_t i = 0, t = 0; i < d->num_meshes; ++i) {
        const sspt_mesh_desc& m = d->meshes[i];
        for (uint32_t f = 0; f < m.num_triangles; ++f, ++t) {
            for (int k = 0; k < 3; ++k) {
                int v = m.indices[f * 3 + k];
                for (int c = 0; c < 3; ++c) verts[(size_t)t * 9 + k * 3 + c] = m.positions[(size_t)v * 3 + c];
            }
            tri_mesh[t] = i;
        }
    }

    // emitters: Scene.Prepare's mesh -> face -> emitter map (Scene.cs:147-163)
    std::vector<int> tri_emitter_global(ntri, -1);
    std::vector<DevEmitter> emitters(d->num_emitters);
    for (uint32_t i = 0; i < d->num_emitters; ++i) {
        const sspt_emitter_desc& e = d->emitters[i];
        if (e.mesh >= d->num_meshes || e.face >= d->meshes[e.mesh].num_triangles)
            return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "emitter %u references a missing triangle", i);
        uint32_t t = mesh_offset[e.mesh] + e.face;
        if (tri_emitter_global[t] >= 0)
            return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "duplicate emitter for one triangle");
        tri_emitter_global[t] = (int)i;
        // Triangle ctor (SeeSharp/Geometry/Triangle.cs:23-35): area from the cross product
        const float* v = &verts[(size_t)t * 9];
        float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
        float e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
        float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        float area = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) * 0.5f;
        if (area == 0.0f)
            return fail(ctx, SSPT_ERR_ZERO_AREA_EMITTER, "Zero-area triangles cannot be emitters");
        emitters[i] = DevEmitter{(int)t, e.radiance[0], e.radiance[1], e.radiance[2], 1.0f / area, {0, 0, 0}};
    }

    // BVH over all triangles
    BvhOutput bvh;
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    build_bvh(BvhInput{verts.data(), (int)ntri}, 4, (int)std::min(hw, 16u), &bvh);
    sc->bvh_depth = bvh.depth;
    std::vector<int> global_to_leaf(ntri);
    for (uint64_t li = 0; li < ntri; ++li) global_to_leaf[bvh.leaf_order[li]] = (int)li;
    for (auto& e : emitters) e.tri = global_to_leaf[e.tri];

    // leaf-order triangle data
    std::vector<float4> tri_pos(ntri * 3), tri_aux(ntri * 4);
    std::vector<int> tri_emitter(ntri);
    std::vector<uint32_t> local_face(ntri);
    for (uint32_t i = 0, t = 0; i < d->num_meshes; ++i)
        for (uint32_t f = 0; f < d->meshes[i].num_triangles; ++f, ++t) local_face[t] = f;
    for (uint64_t li = 0; li < ntri; ++li) {
        int g = bvh.leaf_order[li];
        const sspt_mesh_desc& m = d->meshes[tri_mesh[g]];
        uint32_t f = local_face[g];
        const float* v = &verts[(size_t)g * 9];
        float gbits;
        std::memcpy(&gbits, &g, 4);
        tri_pos[li * 3 + 0] = make_float4(v[0], v[1], v[2], gbits);
        tri_pos[li * 3 + 1] = make_float4(v[3], v[4], v[5], 0.0f);
        tri_pos[li * 3 + 2] = make_float4(v[6], v[7], v[8], 0.0f);
        int idx[3] = {m.indices[f * 3], m.indices[f * 3 + 1], m.indices[f * 3 + 2]};
        float n[3][3] = {{0}}, uv[3][2] = {{0}};
        if (m.normals)
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) n[k][c] = m.normals[(size_t)idx[k] * 3 + c];
        if (m.uvs)
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 2; ++c) uv[k][c] = m.uvs[(size_t)idx[k] * 2 + c];
        int info = (m.material & kInfoMaterialMask) | (m.normals ? kInfoHasNormals : 0) | (m.uvs ? kInfoHasUvs : 0);
        float ibits;
        std::memcpy(&ibits, &info, 4);
        tri_aux[li * 4 + 0] = make_float4(n[0][0], n[0][1], n[0][2], uv[0][0]);
        tri_aux[li * 4 + 1] = make_float4(n[1][0], n[1][1], n[1][2], uv[0][1]);
        tri_aux[li * 4 + 2] = make_float4(n[2][0], n[2][1], n[2][2], uv[1][0]);
        tri_aux[li * 4 + 3] = make_float4(uv[1][1], uv[2][0], uv[2][1], ibits);
        tri_emitter[li] = tri_emitter_global[g];
    }

    // materials and textures
    std::vector<DevMaterial> mats(d->num_materials);
    for (uint32_t i = 0; i < d->num_materials; ++i) {
        const sspt_material_desc& m = d->materials[i];
        DevMaterial& o = mats[i];
        std::memset(&o, 0, sizeof(o));
        o.kind = m.kind;
        for (int c = 0; c < 3; ++c) o.base[c] = m.base_color[c];
        o.base_tex = m.base_color_texture;
        o.roughness = m.roughness;
        o.rough_tex = m.kind == SSPT_MATERIAL_GENERIC ? m.roughness_texture : -1;
        o.metallic = m.metallic;
        o.spec_tint = m.specular_tint_strength;
        o.anisotropic = m.anisotropic;
        o.spec_trans = m.specular_transmittance;
        o.ior = m.index_of_refraction;
        o.transmitter = m.transmitter;
    }
    std::vector<DevTexture> texs(d->num_textures);
    std::vector<float> texels;
    for (uint32_t i = 0; i < d->num_textures; ++i) {
        const sspt_texture_desc& t = d->textures[i];
        if (!t.texels || t.width == 0 || t.height == 0 || (t.channels != 1 && t.channels != 3))
            return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "texture %u is malformed", i);
        texs[i] = DevTexture{texels.size(), (int)t.width, (int)t.height, (int)t.channels, t.border};
        texels.insert(texels.end(), t.texels, t.texels + (size_t)t.width * t.height * t.channels);
    }

    // environment map: EnvironmentMap.BuildSamplingGrid (EnvironmentMap.cs:139-156),
    // RegularGrid2d.Normalize + PiecewiseConstantPDF ctor (RegularGrid2d.cs:56-71, PiecewiseConstant.cs:12-31)
    std::vector<float4> env_rgb;
    std::vector<float> row_cdf, col_cdf;
    int ew = 0, eh = 0;
    if (d->background) {
        ew = (int)d->background->width;
        eh = (int)d->background->height;
        if (ew <= 0 || eh <= 0 || !d->background->rgb)
            return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "environment map is malformed");
        env_rgb.resize((size_t)ew * eh);
        std::vector<float> density((size_t)ew * eh, 0.0f), marg(eh, 0.0f);
        for (int r = 0; r < eh; ++r)
            for (int c = 0; c < ew; ++c) {
                const float* p = d->background->rgb + ((size_t)r * ew + c) * 3;
                env_rgb[(size_t)r * ew + c] = make_float4(p[0], p[1], p[2], 0.0f);
                float lum = 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
                density[(size_t)r * ew + c] += lum;
                marg[r] += lum;
            }
        auto build_cdf = [](const float* w, int n, float* cdf) {
            float sum = 0.0f;
            for (int i = 0; i < n; ++i) cdf[i] = (sum += w[i]);
            float total = cdf[n - 1];
            for (int i = 0; i < n && total > 0.0f; ++i) cdf[i] /= total;
            cdf[n - 1] = 1.0f;
        };
        row_cdf.resize(eh);
        build_cdf(marg.data(), eh, row_cdf.data());
        col_cdf.resize((size_t)ew * eh);
        for (int r = 0; r < eh; ++r) {
            if (marg[r] == 0.0f) {
                // zero-probability row, never selected: keep a uniform cdf so sampling stays finite
                for (int c = 0; c < ew; ++c) col_cdf[(size_t)r * ew + c] = (float)(c + 1) / (float)ew;
                continue;
            }
            build_cdf(&density[(size_t)r * ew], ew, &col_cdf[(size_t)r * ew]);
        }
    }

    DevScene& S = sc->dev;
    const float4* p4;
    TRY(upload(ctx, sc.get(), reinterpret_cast<const float4*>(bvh.nodes.data()), bvh.nodes.size() / 4, &p4));
    S.nodes = p4;
    TRY(upload(ctx, sc.get(), tri_pos.data(), tri_pos.size(), &S.tri_pos));
    TRY(upload(ctx, sc.get(), tri_aux.data(), tri_aux.size(), &S.tri_aux));
    TRY(upload(ctx, sc.get(), tri_emitter.data(), tri_emitter.size(), &S.tri_emitter));
    TRY(upload(ctx, sc.get(), mats.data(), mats.size(), &S.materials));
    TRY(upload(ctx, sc.get(), texs.data(), texs.size(), &S.textures));
    TRY(upload(ctx, sc.get(), texels.data(), texels.size(), &S.texels));
    TRY(upload(ctx, sc.get(), emitters.data(), emitters.size(), &S.emitters));
    S.num_emitters = (int)emitters.size();
    S.num_tris = (int)ntri;
    float abs_max = 0.0f;
    for (float v : verts) abs_max = std::max(abs_max, std::fabs(v));
    S.abs_max = abs_max;
    S.has_env = d->background ? 1 : 0;
    S.env_w = ew;
    S.env_h = eh;
    TRY(upload(ctx, sc.get(), env_rgb.data(), env_rgb.size(), &S.env_rgb));
    TRY(upload(ctx, sc.get(), row_cdf.data(), row_cdf.size(), &S.env_row_cdf));
    TRY(upload(ctx, sc.get(), col_cdf.data(), col_cdf.size(), &S.env_col_cdf));
    if (!setup_camera(d->camera, (int)d->width, (int)d->height, &S.cam))
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "World to camera transform must be invertible.");
    S.width = (int)d->width;
    S.height = (int)d->height;

    size_t img_bytes = (size_t)d->width * d->height * 3 * sizeof(float);
    HIP_TRY(ctx, hipMalloc((void**)&sc->image, img_bytes));
    sc->allocs.push_back(sc->image);
    sc->bytes += img_bytes;
    HIP_TRY(ctx, hipMemset(sc->image, 0, img_bytes));
    sc->num_tris = ntri;
    sc->num_emitters = emitters.size();
    sc->num_nodes = bvh.nodes.size() / 16;
    *out = sc.release();
    return SSPT_OK;
}

void sspt_scene_destroy(sspt_scene* sc) {
    if (!sc) return;
    hipSetDevice(sc->ctx->device);
    hipStreamSynchronize(sc->ctx->stream);
    for (void* p : sc->allocs) hipFree(p);
    delete sc;
}

sspt_status sspt_scene_info(const sspt_scene* sc, uint64_t* ntri, uint64_t* nem, uint64_t* nnodes,
                            uint64_t* bytes) {
    if (!sc) return SSPT_ERR_INVALID_ARGUMENT;
    if (ntri) *ntri = sc->num_tris;
    if (nem) *nem = sc->num_emitters;
    if (nnodes) *nnodes = sc->num_nodes;
    if (bytes) *bytes = sc->bytes;
    return SSPT_OK;
}

sspt_status sspt_framebuffer_clear(sspt_ctx* ctx, sspt_scene* sc) {
    if (!ctx || !sc) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    hipSetDevice(ctx->device);
    HIP_TRY(ctx, hipMemsetAsync(sc->image, 0, (size_t)sc->dev.width * sc->dev.height * 12, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    sc->nan_count = 0;
    return SSPT_OK;
}

sspt_status sspt_framebuffer_read(sspt_ctx* ctx, sspt_scene* sc, float* host) {
    if (!ctx || !sc || !host) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    hipSetDevice(ctx->device);
    HIP_TRY(ctx, hipMemcpyAsync(host, sc->image, (size_t)sc->dev.width * sc->dev.height * 12,
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SSPT_OK;
}

sspt_status sspt_framebuffer_nan_count(sspt_ctx* ctx, sspt_scene* sc, uint64_t* count) {
    if (!ctx || !sc || !count) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    *count = sc->nan_count;
    return SSPT_OK;
}

sspt_status sspt_render(sspt_ctx* ctx, sspt_scene* sc, const sspt_pt_settings* cfg, uint32_t first_iteration,
                        uint32_t num_iterations, sspt_ray_stats* stats) {
    if (!ctx || !sc || !cfg) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (cfg->max_depth < 1 || cfg->num_shadow_rays < 0 || cfg->min_depth < 0)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "invalid settings");
    hipSetDevice(ctx->device);
    const DevScene& S = sc->dev;
    uint32_t np = (uint32_t)(S.width * S.height);
    // wave: as many whole iterations as fit in ~4M path slots
    const uint32_t kTargetSlots = 1u << 22;
    uint32_t per_wave = std::max(1u, std::min(num_iterations, kTargetSlots / std::max(1u, np)));
    if ((uint64_t)np * per_wave >= (1ull << 30)) per_wave = 1;
    TRY(ensure_wave(ctx, np * per_wave));
    size_t shadow_need = (size_t)np * per_wave * 2 * std::max(1, cfg->num_shadow_rays);
    TRY(ensure_shadow(ctx, shadow_need, &ctx->shadow_capacity));
    Wave& w = ctx->wave;
    HIP_TRY(ctx, hipMemsetAsync(w.stats, 0, 64, ctx->stream));
    DevSettings ds{cfg->base_seed, cfg->max_depth, cfg->min_depth, cfg->num_shadow_rays, cfg->enable_bsdf_di};

    auto t0 = std::chrono::steady_clock::now();
    uint32_t done = 0;
    while (done < num_iterations) {
        // MaximumRenderTimeMs (PathTracer.cs:221-227), checked per wave
        if (cfg->max_render_time_ms >= 0 && done > 0) {
            double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            double per = ms / done;
            if (ms + per * std::min(per_wave, num_iterations - done) > (double)cfg->max_render_time_ms) break;
        }
        uint32_t iters = std::min(per_wave, num_iterations - done);
        uint32_t first = first_iteration + done;
        timed(ctx, "generate", [&] { launch_generate(ctx->stream, S, w, cfg->base_seed, first, iters); });
        uint32_t count = np * iters;
        int buf = 0;
        for (int depth = 1; depth <= cfg->max_depth && count > 0; ++depth) {
            HIP_TRY(ctx, hipMemsetAsync(w.counters, 0, 8, ctx->stream));
            timed(ctx, "extend", [&] { launch_extend(ctx->stream, S, w, buf, count); });
            timed(ctx, "shade", [&] { launch_shade(ctx->stream, S, w, ds, buf, count); });
            HIP_TRY(ctx, hipMemcpyAsync(ctx->host_counters, w.counters, 8, hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            uint32_t next = ctx->host_counters[kCntNext];
            uint32_t nshadow = ctx->host_counters[kCntShadow];
            timed(ctx, "shadow", [&] { launch_shadow(ctx->stream, S, w, nshadow); });
            count = next;
            buf ^= 1;
        }
        timed(ctx, "accumulate", [&] { launch_accumulate(ctx->stream, w, sc->image, np, first, iters); });
        done += iters;
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        ctx->timing.collect();
    }
    HIP_TRY(ctx, hipGetLastError());
    unsigned long long hs[kStatNum];
    HIP_TRY(ctx, hipMemcpyAsync(hs, w.stats, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    sc->nan_count += hs[kStatNan];
    if (stats) {
        stats->num_rays = hs[kStatRays];
        stats->num_shadow_rays = hs[kStatShadow];
        stats->num_occluded = hs[kStatOccluded];
        stats->num_ray_hits = hs[kStatHits];
    }
    return SSPT_OK;
}

sspt_status sspt_trace_closest(sspt_ctx* ctx, sspt_scene* sc, const float* rays, uint64_t n, sspt_hit* hits) {
    if (!ctx || !sc || (n && (!rays || !hits))) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (!n) return SSPT_OK;
    hipSetDevice(ctx->device);
    float* dr = nullptr;
    float4* dh = nullptr;
    HIP_TRY(ctx, hipMalloc(&dr, n * 32));
    if (hipMalloc(&dh, n * 16) != hipSuccess) {
        hipFree(dr);
        return fail(ctx, SSPT_ERR_OUT_OF_MEMORY, "hipMalloc failed");
    }
    hipError_t e = hipMemcpyAsync(dr, rays, n * 32, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        launch_trace_closest(ctx->stream, sc->dev, dr, n, dh);
        e = hipMemcpyAsync(hits, dh, n * 16, hipMemcpyDeviceToHost, ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(dr);
    hipFree(dh);
    if (e != hipSuccess) return fail(ctx, SSPT_ERR_DEVICE, "trace failed: %s", hipGetErrorString(e));
    return SSPT_OK;
}

sspt_status sspt_trace_occluded(sspt_ctx* ctx, sspt_scene* sc, const float* rays, uint64_t n, uint8_t* occ) {
    if (!ctx || !sc || (n && (!rays || !occ))) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (!n) return SSPT_OK;
    hipSetDevice(ctx->device);
    float* dr = nullptr;
    uint8_t* dh = nullptr;
    HIP_TRY(ctx, hipMalloc(&dr, n * 32));
    if (hipMalloc(&dh, n) != hipSuccess) {
        hipFree(dr);
        return fail(ctx, SSPT_ERR_OUT_OF_MEMORY, "hipMalloc failed");
    }
    hipError_t e = hipMemcpyAsync(dr, rays, n * 32, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        launch_trace_occluded(ctx->stream, sc->dev, dr, n, dh);
        e = hipMemcpyAsync(occ, dh, n, hipMemcpyDeviceToHost, ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(dr);
    hipFree(dh);
    if (e != hipSuccess) return fail(ctx, SSPT_ERR_DEVICE, "trace failed: %s", hipGetErrorString(e));
    return SSPT_OK;
}

sspt_status sspt_primary_hits(sspt_ctx* ctx, sspt_scene* sc, uint32_t base_seed, uint32_t iteration,
                              float* rays_out, sspt_hit* hits_out) {
    if (!ctx || !sc) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    hipSetDevice(ctx->device);
    uint64_t n = (uint64_t)sc->dev.width * sc->dev.height;
    float* dr = nullptr;
    float4* dh = nullptr;
    HIP_TRY(ctx, hipMalloc(&dr, n * 32));
    if (hipMalloc(&dh, n * 16) != hipSuccess) {
        hipFree(dr);
        return fail(ctx, SSPT_ERR_OUT_OF_MEMORY, "hipMalloc failed");
    }
    launch_primary_rays(ctx->stream, sc->dev, base_seed, iteration, dr);
    launch_trace_closest(ctx->stream, sc->dev, dr, n, dh);
    hipError_t e = hipSuccess;
    if (rays_out) e = hipMemcpyAsync(rays_out, dr, n * 32, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && hits_out) e = hipMemcpyAsync(hits_out, dh, n * 16, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(dr);
    hipFree(dh);
    if (e != hipSuccess) return fail(ctx, SSPT_ERR_DEVICE, "primary hits failed: %s", hipGetErrorString(e));
    return SSPT_OK;
}

sspt_status sspt_set_kernel_timing(sspt_ctx* ctx, int32_t enable) {
    if (!ctx) return SSPT_ERR_INVALID_ARGUMENT;
    ctx->timing.enabled = enable != 0;
    return SSPT_OK;
}

sspt_status sspt_get_kernel_timing(sspt_ctx* ctx, const char* name, uint64_t* launches, double* total_ms) {
    if (!ctx || !name) return SSPT_ERR_INVALID_ARGUMENT;
    hipStreamSynchronize(ctx->stream);
    ctx->timing.collect();
    auto it = ctx->timing.totals.find(name);
    if (launches) *launches = it == ctx->timing.totals.end() ? 0 : it->second.first;
    if (total_ms) *total_ms = it == ctx->timing.totals.end() ? 0.0 : it->second.second;
    return SSPT_OK;
}

sspt_status sspt_reset_kernel_timing(sspt_ctx* ctx) {
    if (!ctx) return SSPT_ERR_INVALID_ARGUMENT;
    hipStreamSynchronize(ctx->stream);
    ctx->timing.collect();
    ctx->timing.totals.clear();
    return SSPT_OK;
}

}  // extern "C"
