// api.hip -- the C ABI of include/sspt.h: contexts, scenes and batches of
// scenes (Scene.Prepare), framebuffers and the render loop of
// PathTracerBase.Render (SeeSharp/Integrators/PathTracer.cs:199-267).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
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
    int num_cus = 256;
    int trace_per_cu = 0, shade_per_cu = 0;  // resident blocks per CU (occupancy query)
    hipStream_t stream = nullptr;
    std::string error;
    Timing timing;
    // wavefront buffers, grown on demand
    Wave wave{};
    uint32_t wave_capacity = 0;
    size_t shadow_capacity = 0;
    std::vector<void*> wave_allocs;
    WaveDesc* wave_descs = nullptr;  // device
    size_t wave_descs_cap = 0;
    // execution knobs (0 = default)
    uint32_t want_capacity = 0;
    uint32_t trace_blocks = 0, shade_blocks = 0;
    uint32_t trace_block = 0, shade_block = 0;  // threads per block (0 = the default sizes)
    bool use_graphs = true;
    int32_t exec_mode = SSPT_EXEC_WAVEFRONT;
    int32_t bvh_builder = SSPT_BVH_DEVICE;
    int32_t bvh_leaf_size = kMaxLeafTriangles;
    unsigned long long* lt_stats = nullptr;  // the light tracer's striped ray statistics
    // adjoint buffers, grown on demand
    std::vector<void*> adj_allocs;
    uint64_t adj_records = 0;   // capacity * depth
    size_t adj_shadow = 0;
    // this device's copy of the file-scope scratch array (sspt_render_noalloc), or null
    unsigned char* scratch = nullptr;
    hipError_t scratch_status = hipSuccess;
};

struct sspt_scene {
    sspt_ctx* ctx = nullptr;
    DevScene dev{};
    std::vector<SceneRec> recs;  // host copy
    SceneRec* d_recs = nullptr;
    uint64_t* d_path_offsets = nullptr;
    std::vector<void*> allocs;
    float* image = nullptr;  // pooled framebuffers
    uint64_t image_floats = 0;
    uint64_t num_tris = 0, num_emitters = 0, num_nodes = 0, bytes = 0;
    uint64_t nan_count = 0;
    int stack_need = 0;
    int32_t bvh_builder = SSPT_BVH_DEVICE;
    double bvh_kernel_ms = 0.0, bvh_wall_ms = 0.0;
    // adjoint: d loss / d texel (texel-pool layout) and d loss / d image (framebuffer layout)
    std::vector<DevTexture> textures;  // host copy of the texture table
    uint64_t texel_floats = 0;
    float* grad = nullptr;
    float* adj_image = nullptr;
    float* layers = nullptr;  // albedo pool then normal pool, framebuffer layout each
    // outlier tracking: image_floats / 3 * dev.num_outliers keys, or null (off)
    unsigned long long* outliers = nullptr;
    // bidirectional rendering: the light-path cache and camera pdf levels, grown on demand
    char* bidir = nullptr;
    uint64_t bidir_bytes = 0;
};

// Scratch memory of sspt_render_noalloc: every wave buffer of an allocation-free
// render lives here. Each GPU the library runs on holds its own copy until the
// process ends; in host code `scratch` only names the symbol, and its device
// address (per device) comes from hipGetSymbolAddress.
static_assert(SSPT_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size is a multiple of 256 bytes");
static __device__ __align__(256) unsigned char scratch[SSPT_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kMaxScratchDevices] = {};
hipError_t g_scratch_status[kMaxScratchDevices] = {};

// Looks up this device's scratch address once per device (the device current).
hipError_t scratch_address(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    if (!g_scratch_ptr[device] && g_scratch_status[device] == hipSuccess) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) g_scratch_status[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
        else g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[device];
    return g_scratch_status[device];
}

sspt_status fail(sspt_ctx* ctx, sspt_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->error = buf;
    return st;
}

#define HIP_TRY(ctx, expr)                                                                            \
    do {                                                                                              \
        hipError_t _e = (expr);                                                                       \
        if (_e != hipSuccess)                                                                         \
            return fail((ctx), _e == hipErrorOutOfMemory ? SSPT_ERR_OUT_OF_MEMORY : SSPT_ERR_DEVICE, \
                        "%s failed: %s", #expr, hipGetErrorString(_e));                               \
    } while (0)

#define TRY(expr)                     \
    do {                              \
        sspt_status _s = (expr);      \
        if (_s != SSPT_OK) return _s; \
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
    std::memcpy(cam->w2c, cd.world_to_camera, sizeof(cam->w2c));
    std::memcpy(cam->c2v, p, sizeof(cam->c2v));
    // PerspectiveCamera.UpdateResolution (PerspectiveCamera.cs:48-51)
    cam->image_plane = (float)h / (2.0f * std::tan(fov * 0.5f));
    const float* W = cam->c2w;
    // Transform((0,0,0), cameraToWorld) = translation row
    cam->pos[0] = 0.0f * W[0] + 0.0f * W[4] + 0.0f * W[8] + W[12];
    cam->pos[1] = 0.0f * W[1] + 0.0f * W[5] + 0.0f * W[9] + W[13];
    cam->pos[2] = 0.0f * W[2] + 0.0f * W[6] + 0.0f * W[10] + W[14];
    cam->width = w;
    cam->height = h;
    return true;
}


// Host-side pools of a world under construction.
struct WorldBuilder {
    std::vector<float> nodes;   // 32 floats per node
    std::vector<float4> tri_pos, tri_aux;
    std::vector<int> tri_emitter;
    std::vector<DevMaterial> materials;
    std::vector<DevTexture> textures;
    std::vector<float> texels;
    std::vector<DevEmitter> emitters;
    std::vector<float4> env_rgb;
    std::vector<float> env_cdf;
    std::vector<int> env_guide;
    std::vector<SceneRec> recs;
    uint64_t fb_floats = 0;
    int stack_need = 0;
    double bvh_kernel_ms = 0.0, bvh_wall_ms = 0.0;
};

// Scene.IsValid (SeeSharp/Scene.cs:170-197) and structural checks of a description.
sspt_status validate(sspt_ctx* ctx, const sspt_scene_desc* d) {
    if (d->width == 0 || d->height == 0) return fail(ctx, SSPT_ERR_INVALID_SCENE, "Framebuffer not set.");
    if ((uint64_t)d->width * d->height > kSlotMask)
        return fail(ctx, SSPT_ERR_UNSUPPORTED, "framebuffer of %ux%u is too large", d->width, d->height);
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
        if (m.kind == SSPT_MATERIAL_GENERIC && m.roughness_texture >= 0 &&
            d->textures[m.roughness_texture].channels != 1)
            return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "material %u: roughness texture must be mono", i);
    }
    for (uint32_t i = 0; i < d->num_textures; ++i) {
        const sspt_texture_desc& t = d->textures[i];
        if (!t.texels || t.width == 0 || t.height == 0 || (t.channels != 1 && t.channels != 3))
            return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "texture %u is malformed", i);
    }
    if (d->background && (d->background->width == 0 || d->background->height == 0 || !d->background->rgb))
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "environment map is malformed");
    return SSPT_OK;
}

// Scene.Prepare for one scene appended to the world: BVH build (TinyEmbree
// CommitScene, Scene.cs:108-113), emitter map (:147-163), environment grid
// (EnvironmentMap.cs:139-156), camera (PerspectiveCamera.UpdateResolution).
sspt_status add_scene(sspt_ctx* ctx, WorldBuilder& W, const sspt_scene_desc* d) {
    TRY(validate(ctx, d));
    const uint32_t tri_base = (uint32_t)(W.tri_pos.size() / 3);
    const int node_base = (int)(W.nodes.size() / 32);
    const int mat_base = (int)W.materials.size();
    const int tex_base = (int)W.textures.size();
    const int em_base = (int)W.emitters.size();

    std::vector<uint32_t> mesh_offset(d->num_meshes);
    uint64_t ntri = 0;
    for (uint32_t i = 0; i < d->num_meshes; ++i) {
        mesh_offset[i] = (uint32_t)ntri;
        ntri += d->meshes[i].num_triangles;
    }
    if (ntri + tri_base >= (1u << 30)) return fail(ctx, SSPT_ERR_UNSUPPORTED, "too many triangles");
    std::vector<float> verts(ntri * 9);
    std::vector<uint32_t> tri_mesh(ntri), local_face(ntri);
    for (uint32_t i = 0, t = 0; i < d->num_meshes; ++i) {
        const sspt_mesh_desc& m = d->meshes[i];
        for (uint32_t f = 0; f < m.num_triangles; ++f, ++t) {
            for (int k = 0; k < 3; ++k) {
                int v = m.indices[f * 3 + k];
                for (int c = 0; c < 3; ++c) verts[(size_t)t * 9 + k * 3 + c] = m.positions[(size_t)v * 3 + c];
            }
            tri_mesh[t] = i;
            local_face[t] = f;
        }
    }

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
        if (area == 0.0f) return fail(ctx, SSPT_ERR_ZERO_AREA_EMITTER, "Zero-area triangles cannot be emitters");
        if (e.kind != SSPT_EMITTER_DIFFUSE && e.kind != SSPT_EMITTER_GLOSSY)
            return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "emitter %u has unknown kind %d", i, e.kind);
        if (e.kind == SSPT_EMITTER_GLOSSY && !(std::isfinite(e.exponent) && e.exponent > -1.0f))
            return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "emitter %u: glossy exponent must be finite and > -1", i);
        DevEmitter de{};
        de.tri = (int)t;
        de.r = e.radiance[0];
        de.g = e.radiance[1];
        de.b = e.radiance[2];
        de.inv_area = 1.0f / area;
        de.kind = e.kind;
        de.exponent = e.kind == SSPT_EMITTER_GLOSSY ? e.exponent : 0.0f;
        // GlossyEmitter ctor (GlossyEmitter.cs:14): the lobe keeps a diffuse emitter's power
        de.norm = e.kind == SSPT_EMITTER_GLOSSY ? (e.exponent + 1.0f) / (2.0f * 3.14159265358979323846f) : 1.0f;
        emitters[i] = de;
    }

    BvhOutput bvh;
    const auto build_start = std::chrono::steady_clock::now();
    if (ctx->bvh_builder == SSPT_BVH_DEVICE) {
        std::string err;
        double kernel_ms = 0.0;
        if (!build_bvh_device(ctx->stream, BvhInput{verts.data(), (int)ntri}, ctx->bvh_leaf_size, &bvh, &kernel_ms,
                              &err))
            return fail(ctx, err.find("out of memory") != std::string::npos ? SSPT_ERR_OUT_OF_MEMORY : SSPT_ERR_DEVICE,
                        "device BVH build: %s", err.c_str());
        W.bvh_kernel_ms += kernel_ms;
    } else {
        unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        build_bvh(BvhInput{verts.data(), (int)ntri}, ctx->bvh_leaf_size, (int)std::min(hw, 16u), &bvh);
    }
    W.bvh_wall_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - build_start).count();
    if (bvh.stack_need > kStackTotal)
        return fail(ctx, SSPT_ERR_UNSUPPORTED, "BVH needs a traversal stack of %d entries (limit %d)", bvh.stack_need,
                    kStackTotal);
    W.stack_need = std::max(W.stack_need, bvh.stack_need);
    std::vector<int> global_to_leaf(ntri);
    for (uint64_t li = 0; li < ntri; ++li) global_to_leaf[bvh.leaf_order[li]] = (int)li;
    for (auto& e : emitters) e.tri = (int)tri_base + global_to_leaf[e.tri];

    // nodes: shift inner-child indices by the node base and leaf ranges by the triangle base
    size_t nn = bvh.nodes.size() / 32;
    for (size_t n = 0; n < nn; ++n) {
        float* node = &bvh.nodes[n * 32];
        int* idx = reinterpret_cast<int*>(node + 24);
        int* cnt = reinterpret_cast<int*>(node + 28);
        uint32_t leaf = 0, inner = 0;
        for (int k = 0; k < 4; ++k) {
            if (cnt[k] > 0) idx[k] += (int)tri_base;
            else if (cnt[k] == 0) idx[k] += node_base;
            if (cnt[k] >= (1 << 24)) return fail(ctx, SSPT_ERR_UNSUPPORTED, "BVH leaf of 2^24 or more triangles");
            leaf |= (uint32_t)(cnt[k] > 0) << k;
            inner |= (uint32_t)(cnt[k] == 0) << k;
        }
        // the node's leaf and inner slot masks ride in the top byte of slot 0's count
        cnt[0] = (int)(((uint32_t)cnt[0] & 0x00FFFFFFu) | ((leaf | (inner << 4)) << 24));
    }
    W.nodes.insert(W.nodes.end(), bvh.nodes.begin(), bvh.nodes.end());

    // leaf-order triangle data
    for (uint64_t li = 0; li < ntri; ++li) {
        int g = bvh.leaf_order[li];
        const sspt_mesh_desc& m = d->meshes[tri_mesh[g]];
        uint32_t f = local_face[g];
        const float* v = &verts[(size_t)g * 9];
        int wid = (int)tri_base + g;
        float gbits;
        std::memcpy(&gbits, &wid, 4);
        W.tri_pos.push_back(make_float4(v[0], v[1], v[2], gbits));
        W.tri_pos.push_back(make_float4(v[3], v[4], v[5], 0.0f));
        W.tri_pos.push_back(make_float4(v[6], v[7], v[8], 0.0f));
        int idx[3] = {m.indices[f * 3], m.indices[f * 3 + 1], m.indices[f * 3 + 2]};
        float n[3][3] = {{0}}, uv[3][2] = {{0}};
        if (m.normals)
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) n[k][c] = m.normals[(size_t)idx[k] * 3 + c];
        if (m.uvs)
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 2; ++c) uv[k][c] = m.uvs[(size_t)idx[k] * 2 + c];
        int info = ((mat_base + m.material) & kInfoMaterialMask) | (m.normals ? kInfoHasNormals : 0) |
                   (m.uvs ? kInfoHasUvs : 0);
        float ibits;
        std::memcpy(&ibits, &info, 4);
        W.tri_aux.push_back(make_float4(n[0][0], n[0][1], n[0][2], uv[0][0]));
        W.tri_aux.push_back(make_float4(n[1][0], n[1][1], n[1][2], uv[0][1]));
        W.tri_aux.push_back(make_float4(n[2][0], n[2][1], n[2][2], uv[1][0]));
        W.tri_aux.push_back(make_float4(uv[1][1], uv[2][0], uv[2][1], ibits));
        W.tri_emitter.push_back(tri_emitter_global[g]);
    }

    for (uint32_t i = 0; i < d->num_materials; ++i) {
        const sspt_material_desc& m = d->materials[i];
        DevMaterial o;
        std::memset(&o, 0, sizeof(o));
        o.kind = m.kind;
        for (int c = 0; c < 3; ++c) o.base[c] = m.base_color[c];
        o.base_tex = m.base_color_texture >= 0 ? tex_base + m.base_color_texture : -1;
        o.roughness = m.roughness;
        o.rough_tex = (m.kind == SSPT_MATERIAL_GENERIC && m.roughness_texture >= 0) ? tex_base + m.roughness_texture : -1;
        o.metallic = m.metallic;
        o.spec_tint = m.specular_tint_strength;
        o.anisotropic = m.anisotropic;
        o.spec_trans = m.specular_transmittance;
        o.ior = m.index_of_refraction;
        o.transmitter = m.transmitter;
        W.materials.push_back(o);
    }
    for (uint32_t i = 0; i < d->num_textures; ++i) {
        const sspt_texture_desc& t = d->textures[i];
        W.textures.push_back(DevTexture{W.texels.size(), (int)t.width, (int)t.height, (int)t.channels, t.border});
        W.texels.insert(W.texels.end(), t.texels, t.texels + (size_t)t.width * t.height * t.channels);
    }
    W.emitters.insert(W.emitters.end(), emitters.begin(), emitters.end());

    SceneRec R;
    std::memset(&R, 0, sizeof(R));
    if (!setup_camera(d->camera, (int)d->width, (int)d->height, &R.cam))
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "World to camera transform must be invertible.");
    R.root = node_base;
    R.emitter_off = em_base;
    R.num_emitters = (int)d->num_emitters;
    float abs_max = 0.0f;
    for (float v : verts) abs_max = std::max(abs_max, std::fabs(v));
    R.abs_max = abs_max;
    R.num_pixels = d->width * d->height;
    R.fb_off = W.fb_floats;
    W.fb_floats += (uint64_t)R.num_pixels * 3;

    // EnvironmentMap.BuildSamplingGrid (EnvironmentMap.cs:139-156): luminance grid,
    // RegularGrid2d.Normalize + PiecewiseConstantPDF ctor (RegularGrid2d.cs:56-71,
    // PiecewiseConstant.cs:12-31)
    if (d->background) {
        // Scene.Prepare (Scene.cs:114-141): the centre is the mean of every mesh vertex
        // (summed in mesh and vertex order), the radius the distance to the furthest one
        float sum[3] = {0.0f, 0.0f, 0.0f};
        uint64_t total = 0;
        for (uint32_t mi = 0; mi < d->num_meshes; ++mi) {
            const sspt_mesh_desc& m = d->meshes[mi];
            for (uint32_t v = 0; v < m.num_vertices; ++v)
                for (int a = 0; a < 3; ++a) sum[a] = sum[a] + m.positions[v * 3 + a];
            total += m.num_vertices;
        }
        for (int a = 0; a < 3; ++a) R.scene_center[a] = sum[a] / (float)total;
        float r2 = 0.0f;
        for (uint32_t mi = 0; mi < d->num_meshes; ++mi) {
            const sspt_mesh_desc& m = d->meshes[mi];
            for (uint32_t v = 0; v < m.num_vertices; ++v) {
                const float* q = m.positions + (size_t)v * 3;
                const float x = q[0] - R.scene_center[0], y = q[1] - R.scene_center[1], z = q[2] - R.scene_center[2];
                r2 = std::max(x * x + y * y + z * z, r2);
            }
        }
        R.scene_radius = std::sqrt(r2);
        int ew = (int)d->background->width, eh = (int)d->background->height;
        R.has_env = 1;
        R.env_w = ew;
        R.env_h = eh;
        R.env_rgb_off = (uint32_t)W.env_rgb.size();
        std::vector<float> density((size_t)ew * eh, 0.0f), marg(eh, 0.0f);
        for (int r = 0; r < eh; ++r)
            for (int c = 0; c < ew; ++c) {
                const float* p = d->background->rgb + ((size_t)r * ew + c) * 3;
                W.env_rgb.push_back(make_float4(p[0], p[1], p[2], 0.0f));
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
        R.env_row_off = (uint32_t)W.env_cdf.size();
        W.env_cdf.resize(W.env_cdf.size() + eh);
        build_cdf(marg.data(), eh, &W.env_cdf[R.env_row_off]);
        R.env_col_off = (uint32_t)W.env_cdf.size();
        W.env_cdf.resize(W.env_cdf.size() + (size_t)ew * eh);
        for (int r = 0; r < eh; ++r) {
            float* cdf = &W.env_cdf[R.env_col_off + (size_t)r * ew];
            if (marg[r] == 0.0f) {
                // zero-probability row, never selected: a uniform cdf keeps sampling finite
                for (int c = 0; c < ew; ++c) cdf[c] = (float)(c + 1) / (float)ew;
                continue;
            }
            build_cdf(&density[(size_t)r * ew], ew, cdf);
        }
        // guide tables: guide[j] = lower bound of j / g (the device searches
        // between guide[j] and guide[j + 1] only)
        auto pow2_at_least = [](int n) {
            int g = 1;
            while (g < n) g <<= 1;
            return g;
        };
        auto build_guide = [](const float* cdf, int n, int g, int* guide) {
            int i = 0;
            for (int j = 0; j <= g; ++j) {
                const float v = (float)j / (float)g;
                while (i < n && cdf[i] < v) ++i;
                guide[j] = i;
            }
        };
        R.env_guide_row_g = pow2_at_least(eh);
        R.env_guide_col_g = pow2_at_least(ew);
        R.env_guide_row_off = (uint32_t)W.env_guide.size();
        W.env_guide.resize(W.env_guide.size() + R.env_guide_row_g + 1);
        build_guide(&W.env_cdf[R.env_row_off], eh, R.env_guide_row_g, &W.env_guide[R.env_guide_row_off]);
        R.env_guide_col_off = (uint32_t)W.env_guide.size();
        W.env_guide.resize(W.env_guide.size() + (size_t)eh * (R.env_guide_col_g + 1));
        for (int r = 0; r < eh; ++r)
            build_guide(&W.env_cdf[R.env_col_off + (size_t)r * ew], ew, R.env_guide_col_g,
                        &W.env_guide[R.env_guide_col_off + (size_t)r * (R.env_guide_col_g + 1)]);
    }
    W.recs.push_back(R);
    return SSPT_OK;
}

// Builds a world from n scene descriptions. With `views`, n is 1 and the world
// holds num_views records of that one scene (shared geometry, BVH, materials,
// textures, emitters and environment), each with its own camera and framebuffer.
sspt_status create_world(sspt_ctx* ctx, const sspt_scene_desc* descs, uint32_t n, const sspt_camera_desc* views,
                         uint32_t num_views, sspt_scene** out) {
    if (!ctx || !descs || !out) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (n == 0 || n > (uint32_t)kMaxScenes)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "a batch holds 1..%d scenes", kMaxScenes);
    if (views && (n != 1 || num_views == 0 || num_views > (uint32_t)kMaxScenes))
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "a view set holds 1..%d views of one scene", kMaxScenes);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    WorldBuilder W;
    for (uint32_t i = 0; i < n; ++i) TRY(add_scene(ctx, W, &descs[i]));
    if (views) {
        const SceneRec base = W.recs[0];
        W.recs.clear();
        W.fb_floats = 0;
        for (uint32_t v = 0; v < num_views; ++v) {
            SceneRec R = base;
            if (!setup_camera(views[v], (int)descs[0].width, (int)descs[0].height, &R.cam))
                return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "view %u: world to camera transform must be invertible",
                            v);
            R.fb_off = W.fb_floats;
            W.fb_floats += (uint64_t)R.num_pixels * 3;
            W.recs.push_back(R);
        }
    }
    if (W.materials.size() > (size_t)kInfoMaterialMask)
        return fail(ctx, SSPT_ERR_UNSUPPORTED, "too many materials");

    // texel offsets travel as 32-bit ints through the shading and adjoint kernels
    if (W.texels.size() > (size_t)INT32_MAX)
        return fail(ctx, SSPT_ERR_UNSUPPORTED, "texture data exceeds 2^31 floats");
    // the traversal addresses nodes by 32-bit byte offsets (80 B per traversal node)
    if (W.nodes.size() / 32 > (size_t)kMaxNodes)
        return fail(ctx, SSPT_ERR_UNSUPPORTED, "BVH exceeds 2^25 nodes");
    std::unique_ptr<sspt_scene> sc(new sspt_scene());
    sc->ctx = ctx;
    DevScene& S = sc->dev;
    TRY(upload(ctx, sc.get(), reinterpret_cast<const float4*>(W.nodes.data()), W.nodes.size() / 4, &S.nodes));
    {
        // the traversal's copy of the tree (fp16 child planes, 80 B per node); its
        // encoding counts as BVH build time
        const auto enc_start = std::chrono::steady_clock::now();
        std::vector<uint32_t> qwords;
        std::vector<int> big;
        std::string err;
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        if (!encode_traversal_nodes(W.nodes, (int)std::min(hw, 16u), &qwords, &big, &err))
            return fail(ctx, SSPT_ERR_UNSUPPORTED, "%s", err.c_str());
        if (big.empty()) big.assign(2, 0);
        TRY(upload(ctx, sc.get(), qwords.data(), qwords.size(), &S.qnodes));
        TRY(upload(ctx, sc.get(), big.data(), big.size(), &S.big_leaves));
        W.bvh_wall_ms +=
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - enc_start).count();
    }
    TRY(upload(ctx, sc.get(), W.tri_pos.data(), W.tri_pos.size(), &S.tri_pos));
    TRY(upload(ctx, sc.get(), W.tri_aux.data(), W.tri_aux.size(), &S.tri_aux));
    TRY(upload(ctx, sc.get(), W.tri_emitter.data(), W.tri_emitter.size(), &S.tri_emitter));
    TRY(upload(ctx, sc.get(), W.materials.data(), W.materials.size(), &S.materials));
    TRY(upload(ctx, sc.get(), W.textures.data(), W.textures.size(), &S.textures));
    TRY(upload(ctx, sc.get(), W.texels.data(), W.texels.size(), &S.texels));
    TRY(upload(ctx, sc.get(), W.emitters.data(), W.emitters.size(), &S.emitters));
    TRY(upload(ctx, sc.get(), W.env_rgb.data(), W.env_rgb.size(), &S.env_rgb));
    TRY(upload(ctx, sc.get(), W.env_cdf.data(), W.env_cdf.size(), &S.env_cdf));
    TRY(upload(ctx, sc.get(), W.env_guide.data(), W.env_guide.size(), &S.env_guide));
    HIP_TRY(ctx, hipMalloc((void**)&sc->d_recs, W.recs.size() * sizeof(SceneRec)));
    sc->allocs.push_back(sc->d_recs);
    HIP_TRY(ctx, hipMalloc((void**)&sc->d_path_offsets, (W.recs.size() + 1) * sizeof(uint64_t)));
    sc->allocs.push_back(sc->d_path_offsets);
    HIP_TRY(ctx, hipMemcpy(sc->d_recs, W.recs.data(), W.recs.size() * sizeof(SceneRec), hipMemcpyHostToDevice));
    S.recs = sc->d_recs;
    S.path_offsets = sc->d_path_offsets;
    S.num_scenes = (int)W.recs.size();
    S.num_nodes = (int)(W.nodes.size() / 32);
    sc->recs = W.recs;
    sc->image_floats = W.fb_floats;
    HIP_TRY(ctx, hipMalloc((void**)&sc->image, W.fb_floats * sizeof(float)));
    sc->allocs.push_back(sc->image);
    sc->bytes += W.fb_floats * sizeof(float);
    HIP_TRY(ctx, hipMemset(sc->image, 0, W.fb_floats * sizeof(float)));
    sc->textures = W.textures;
    sc->texel_floats = W.texels.size();
    if (sc->texel_floats) {
        HIP_TRY(ctx, hipMalloc((void**)&sc->grad, sc->texel_floats * sizeof(float)));
        sc->allocs.push_back(sc->grad);
        HIP_TRY(ctx, hipMemset(sc->grad, 0, sc->texel_floats * sizeof(float)));
    }
    HIP_TRY(ctx, hipMalloc((void**)&sc->adj_image, W.fb_floats * sizeof(float)));
    sc->allocs.push_back(sc->adj_image);
    HIP_TRY(ctx, hipMalloc((void**)&sc->layers, 2 * W.fb_floats * sizeof(float)));
    sc->allocs.push_back(sc->layers);
    sc->bytes += 2 * W.fb_floats * sizeof(float);
    HIP_TRY(ctx, hipMemset(sc->layers, 0, 2 * W.fb_floats * sizeof(float)));
    S.layers = sc->layers;
    S.layer_stride = W.fb_floats;
    S.grad = sc->grad;
    S.adj_image = sc->adj_image;
    sc->num_tris = W.tri_pos.size() / 3;
    sc->num_emitters = W.emitters.size();
    sc->num_nodes = W.nodes.size() / 32;
    sc->stack_need = W.stack_need;
    sc->bvh_builder = ctx->bvh_builder;
    sc->bvh_kernel_ms = W.bvh_kernel_ms;
    sc->bvh_wall_ms = W.bvh_wall_ms;
    *out = sc.release();
    return SSPT_OK;
}

sspt_status ensure_wave(sspt_ctx* ctx, uint32_t capacity) {
    if (ctx->wave_capacity >= capacity) return SSPT_OK;
    for (void* p : ctx->wave_allocs) (void)hipFree(p);
    ctx->wave_allocs.clear();
    for (void* p : ctx->adj_allocs) (void)hipFree(p);
    ctx->adj_allocs.clear();
    ctx->adj_records = 0;
    ctx->adj_shadow = 0;
    ctx->wave.ad_dlog = nullptr;
    ctx->wave.ad_A = nullptr;
    ctx->wave.ad_tex = nullptr;
    ctx->wave.s_g = nullptr;
    ctx->wave.s_t = nullptr;
    ctx->wave.ad_depth = 0;
    ctx->wave_capacity = 0;
    ctx->shadow_capacity = 0;
    auto alloc = [&](void** p, size_t bytes) -> sspt_status {
        HIP_TRY(ctx, hipMalloc(p, bytes));
        ctx->wave_allocs.push_back(*p);
        return SSPT_OK;
    };
    Wave& w = ctx->wave;
    size_t P = capacity;
    for (int b = 0; b < 2; ++b) {
        TRY(alloc((void**)&w.q_o[b], P * 16));
        TRY(alloc((void**)&w.q_d[b], P * 16));
        TRY(alloc((void**)&w.q_slot[b], P * 4));
        TRY(alloc((void**)&w.q_rng[b], P * 4));
        TRY(alloc((void**)&w.q_thr[b], P * 16));
        TRY(alloc((void**)&w.q_prev[b], P * 16));
        TRY(alloc((void**)&w.q_L[b], P * 16));
    }
    TRY(alloc((void**)&w.hits, P * 16));
    TRY(alloc((void**)&w.st_pix, P * 4));
    TRY(alloc((void**)&w.st_prev, P * 16));
    TRY(alloc((void**)&w.st_L, P * 16));
    TRY(alloc((void**)&w.st_Le, P * 16));
    TRY(alloc((void**)&w.st_La, P * 16));
    TRY(alloc((void**)&w.st_alb, P * 16));
    TRY(alloc((void**)&w.st_nrm, P * 16));
    TRY(alloc((void**)&w.counters, 64));
    TRY(alloc((void**)&w.stats, kStatBytes));
    w.capacity = capacity;
    w.s_o = w.s_d = w.s_c = nullptr;
    ctx->wave_capacity = capacity;
    return SSPT_OK;
}

sspt_status ensure_shadow(sspt_ctx* ctx, size_t entries) {
    if (ctx->shadow_capacity >= entries) return SSPT_OK;
    Wave& w = ctx->wave;
    for (void* p : {(void*)w.s_o, (void*)w.s_d, (void*)w.s_c}) {
        if (!p) continue;
        (void)hipFree(p);
        ctx->wave_allocs.erase(std::remove(ctx->wave_allocs.begin(), ctx->wave_allocs.end(), p),
                               ctx->wave_allocs.end());
    }
    w.s_o = w.s_d = w.s_c = nullptr;
    ctx->shadow_capacity = 0;
    for (float4** p : {&w.s_o, &w.s_d, &w.s_c}) {
        HIP_TRY(ctx, hipMalloc((void**)p, entries * 16));
        ctx->wave_allocs.push_back(*p);
    }
    ctx->shadow_capacity = entries;
    return SSPT_OK;
}

sspt_status ensure_wave_descs(sspt_ctx* ctx, size_t n) {
    if (ctx->wave_descs_cap >= n) return SSPT_OK;
    if (ctx->wave_descs) (void)hipFree(ctx->wave_descs);
    ctx->wave_descs = nullptr;
    ctx->wave_descs_cap = 0;
    HIP_TRY(ctx, hipMalloc((void**)&ctx->wave_descs, n * sizeof(WaveDesc)));
    ctx->wave_descs_cap = n;
    return SSPT_OK;
}

// Per-vertex adjoint records (capacity x depth) and per-shadow-ray gradients.
sspt_status ensure_adjoint(sspt_ctx* ctx, uint32_t depth) {
    Wave& w = ctx->wave;
    uint64_t records = (uint64_t)ctx->wave_capacity * depth;
    if (ctx->adj_records < records || ctx->adj_shadow < ctx->shadow_capacity) {
        for (void* p : ctx->adj_allocs) (void)hipFree(p);
        ctx->adj_allocs.clear();
        ctx->adj_records = 0;
        ctx->adj_shadow = 0;
        auto alloc = [&](void** p, size_t bytes) -> sspt_status {
            HIP_TRY(ctx, hipMalloc(p, bytes));
            ctx->adj_allocs.push_back(*p);
            return SSPT_OK;
        };
        TRY(alloc((void**)&w.ad_dlog, records * 48));
        TRY(alloc((void**)&w.ad_A, records * 16));
        TRY(alloc((void**)&w.ad_tex, records * 8));
        TRY(alloc((void**)&w.s_g, ctx->shadow_capacity * 16));
        TRY(alloc((void**)&w.s_t, ctx->shadow_capacity * 8));
        ctx->adj_records = records;
        ctx->adj_shadow = ctx->shadow_capacity;
    }
    w.ad_depth = (uint32_t)(ctx->adj_records / std::max(1u, ctx->wave_capacity));
    return SSPT_OK;
}

template <class F>
void timed(sspt_ctx* ctx, const char* name, F&& launch) {
    if (!ctx->timing.enabled) {
        launch();
        return;
    }
    hipEvent_t a = ctx->timing.get(), b = ctx->timing.get();
    (void)hipEventRecord(a, ctx->stream);
    launch();
    (void)hipEventRecord(b, ctx->stream);
    ctx->timing.pending.push_back({name, a, b});
}

// Grid sizes for a render whose largest wave has `max_paths` paths and whose
// scenes trace up to `max_nsr` shadow rays per kind and vertex.
LaunchConfig launch_config(const sspt_ctx* ctx, uint32_t max_paths, uint32_t max_nsr) {
    LaunchConfig lc;
    lc.trace_blocks = ctx->trace_blocks;
    lc.shade_blocks = ctx->shade_blocks;
    lc.trace_block = ctx->trace_block;
    lc.shade_block = ctx->shade_block;
    lc.max_paths = max_paths;
    lc.max_nsr = (int)std::max(1u, max_nsr);
    const uint64_t shadow = (uint64_t)max_paths * 2u * std::max(1u, max_nsr);
    lc.max_shadow = (uint32_t)std::min<uint64_t>(shadow, ctx->shadow_capacity);
    return lc;
}

// Whether a bounce's shadow rays are traced in one launch with the next
// bounce's continuation rays (1) or in launches of their own (0).
#ifndef SSPT_FUSE_TRACE
#define SSPT_FUSE_TRACE 1
#endif

// Enqueues the kernels of one wave; the wave's work is read on the device from
// the descriptor the wave counter selects.
enum class Mode { Mean, Sum, Adjoint };

// Whether every scene record agrees on the fields the shade kernel reads
// (PathVertex, env_of, the denoiser flag), so that it may read record 0 for all.
bool records_alike(const std::vector<SceneRec>& recs) {
    for (size_t i = 1; i < recs.size(); ++i) {
        const SceneRec &a = recs[0], &b = recs[i];
        if (a.num_shadow_rays != b.num_shadow_rays || a.enable_bsdf_di != b.enable_bsdf_di ||
            a.min_depth != b.min_depth || a.max_depth != b.max_depth || a.enable_denoiser != b.enable_denoiser ||
            a.num_emitters != b.num_emitters || a.emitter_off != b.emitter_off || a.has_env != b.has_env ||
            a.env_w != b.env_w || a.env_h != b.env_h || a.env_rgb_off != b.env_rgb_off ||
            a.env_row_off != b.env_row_off || a.env_col_off != b.env_col_off ||
            a.env_guide_row_off != b.env_guide_row_off || a.env_guide_col_off != b.env_guide_col_off ||
            a.env_guide_row_g != b.env_guide_row_g || a.env_guide_col_g != b.env_guide_col_g)
            return false;
    }
    return true;
}

void enqueue_wave(sspt_ctx* ctx, sspt_scene* sc, const Wave& w, const LaunchConfig& lc, int max_depth, Mode mode,
                  float* mean_image, uint32_t mean_pixels, float* mean_albedo) {
    float* mean_normal = mean_albedo ? mean_albedo + sc->dev.layer_stride : nullptr;
    // running-mean renders are single-scene: the scene's keys start at the pool's start
    unsigned long long* mean_outl = sc->dev.outliers;
    const int n_outl = sc->dev.num_outliers;
    const DevScene& S = sc->dev;
    hipStream_t st = ctx->stream;
    const bool adj = mode == Mode::Adjoint;
    launch_wave_begin(st, w);
    if (ctx->exec_mode == SSPT_EXEC_SINGLE_KERNEL && !adj) {
        timed(ctx, "path", [&] { launch_path(st, S, w, lc); });
        if (mode == Mode::Mean)
            timed(ctx, "accumulate",
                  [&] {
                      launch_accumulate_mean(st, w, mean_image, mean_pixels, mean_albedo, mean_normal, mean_outl,
                                             n_outl);
                  });
        else
            timed(ctx, "accumulate", [&] { launch_accumulate_sum(st, S, w, lc, sc->image); });
        launch_wave_end(st, w);
        return;
    }
    int buf = 0;
    if (SSPT_FUSE_TRACE) {
        // the paths start and their camera rays are traced in one launch; bounce
        // d's shade is followed by one launch tracing the continuation rays of
        // bounce d + 1 and the shadow rays of bounce d
        launch_bounce_reset(st, w, buf);
        timed(ctx, "extend", [&] { launch_extend_primary(st, S, w, lc); });
        for (int depth = 1; depth <= max_depth; ++depth) {
            if (adj) {
                timed(ctx, "shade_adjoint", [&] { launch_shade_adjoint(st, S, w, lc, buf); });
                timed(ctx, "shade_grad", [&] { launch_shade_grad(st, S, w, lc, buf); });
            } else {
                timed(ctx, "shade", [&] { launch_shade(st, S, w, lc, buf); });
            }
            // the last bounce samples no next events (every path still alive is at its
            // scene's MaxDepth, PathTracer.cs:367-368), so its shadow queue stays empty
            if (depth < max_depth) {
                timed(ctx, adj ? "trace_adjoint" : "trace", [&] { launch_trace(st, S, w, lc, buf ^ 1, adj); });
                launch_bounce_reset(st, w, buf ^ 1);
            }
            buf ^= 1;
        }
    } else {
        timed(ctx, "generate", [&] { launch_generate(st, S, w, lc); });
        for (int depth = 1; depth <= max_depth; ++depth) {
            launch_bounce_reset(st, w, buf);
            timed(ctx, "extend", [&] { launch_extend(st, S, w, lc, buf); });
            if (adj) {
                timed(ctx, "shade_adjoint", [&] { launch_shade_adjoint(st, S, w, lc, buf); });
                timed(ctx, "shade_grad", [&] { launch_shade_grad(st, S, w, lc, buf); });
                timed(ctx, "shadow_adjoint", [&] { launch_shadow_adjoint(st, S, w, lc, buf ^ 1); });
            } else {
                timed(ctx, "shade", [&] { launch_shade(st, S, w, lc, buf); });
                timed(ctx, "shadow", [&] { launch_shadow(st, S, w, lc, buf ^ 1); });
            }
            buf ^= 1;
        }
    }
    if (mode == Mode::Mean)
        timed(ctx, "accumulate",
              [&] {
                  launch_accumulate_mean(st, w, mean_image, mean_pixels, mean_albedo, mean_normal, mean_outl, n_outl);
              });
    else if (mode == Mode::Sum)
        timed(ctx, "accumulate", [&] { launch_accumulate_sum(st, S, w, lc, sc->image); });
    else
        timed(ctx, "adjoint_finalize", [&] { launch_adjoint_finalize(st, S, w, lc); });
    launch_wave_end(st, w);
}

// Runs the given waves: either every wave enqueued directly (kernel timing on,
// or graphs off), or one wave captured as a graph and replayed.
sspt_status run_waves(sspt_ctx* ctx, sspt_scene* sc, const std::vector<WaveDesc>& waves, int max_depth,
                      int max_nsr, Mode mode, float* mean_image, uint32_t mean_pixels, float* mean_albedo,
                      int64_t budget_ms, uint32_t* waves_run = nullptr) {
    uint32_t max_paths = 0;
    for (const WaveDesc& d : waves) max_paths = std::max(max_paths, d.count);
    const LaunchConfig lc = launch_config(ctx, max_paths, (uint32_t)std::max(1, max_nsr));
    sc->dev.recs_alike = records_alike(sc->recs) ? 1 : 0;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->wave_descs, waves.data(), waves.size() * sizeof(WaveDesc),
                                hipMemcpyHostToDevice, ctx->stream));
    // counters: queues and wave index 0 (issued before any capture: the
    // captured graph holds kernels only)
    HIP_TRY(ctx, hipMemsetAsync(ctx->wave.counters, 0, 64, ctx->stream));
    ctx->wave.waves = ctx->wave_descs;
    const bool graphs = ctx->use_graphs && !ctx->timing.enabled && budget_ms < 0;
    hipGraphExec_t exec = nullptr;
    if (graphs) {
        hipGraph_t graph = nullptr;
        HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        enqueue_wave(ctx, sc, ctx->wave, lc, max_depth, mode, mean_image, mean_pixels, mean_albedo);
        HIP_TRY(ctx, hipStreamEndCapture(ctx->stream, &graph));
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(ctx, SSPT_ERR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    // MaximumRenderTimeMs (PathTracer.cs:221-227) with RenderTimer's arithmetic
    // (SeeSharp/Integrators/Util/RenderTimer.cs:35-60): RenderTime sums each
    // iteration's whole elapsed milliseconds, PerIterationCost = RenderTime / done,
    // and the next wave runs only while RenderTime + PerIterationCost <= budget.
    // Budgeted renders pass waves of one iteration, and each is waited for.
    int64_t render_ms = 0;
    uint32_t done = 0;
    sspt_status st = SSPT_OK;
    for (size_t i = 0; i < waves.size(); ++i) {
        std::chrono::steady_clock::time_point t0;
        if (budget_ms >= 0) {
            const int64_t per_iteration = done ? render_ms / (int64_t)done : 0;
            if (render_ms + per_iteration > budget_ms) break;
            t0 = std::chrono::steady_clock::now();
        }
        if (graphs) {
            hipError_t e = hipGraphLaunch(exec, ctx->stream);
            if (e != hipSuccess) {
                st = fail(ctx, SSPT_ERR_DEVICE, "hipGraphLaunch failed: %s", hipGetErrorString(e));
                break;
            }
        } else {
            enqueue_wave(ctx, sc, ctx->wave, lc, max_depth, mode, mean_image, mean_pixels, mean_albedo);
        }
        ++done;
        if (budget_ms >= 0) {
            hipError_t e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) {
                st = fail(ctx, SSPT_ERR_DEVICE, "render failed: %s", hipGetErrorString(e));
                break;
            }
            render_ms += std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0)
                             .count();
        }
        if (ctx->timing.enabled && (i % 4 == 3)) {
            (void)hipStreamSynchronize(ctx->stream);
            ctx->timing.collect();
        }
    }
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (waves_run) *waves_run = done;
    if (st != SSPT_OK) return st;
    if (e != hipSuccess) return fail(ctx, SSPT_ERR_DEVICE, "render failed: %s", hipGetErrorString(e));
    HIP_TRY(ctx, hipGetLastError());
    ctx->timing.collect();
    return SSPT_OK;
}

sspt_status read_stats(sspt_ctx* ctx, sspt_scene* sc, sspt_ray_stats* stats,
                       const unsigned long long* device_stats = nullptr) {
    unsigned long long striped[(size_t)kStatStripes * kStatStride];
    HIP_TRY(ctx, hipMemcpyAsync(striped, device_stats ? device_stats : ctx->wave.stats, kStatBytes,
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    unsigned long long hs[kStatNum] = {};
    for (int k = 0; k < kStatStripes; ++k)
        for (int j = 0; j < kStatNum; ++j) hs[j] += striped[(size_t)k * kStatStride + j];
    sc->nan_count += hs[kStatNan];
    if (stats) {
        stats->num_rays = hs[kStatRays];
        stats->num_shadow_rays = hs[kStatShadow];
        stats->num_occluded = hs[kStatOccluded];
        stats->num_ray_hits = hs[kStatHits];
        stats->num_shadow_rays_traced = hs[kStatShadowTraced];
        stats->num_iterations = 0;
    }
    return SSPT_OK;
}

bool valid_settings(const sspt_pt_settings& s) {
    return s.max_depth >= 1 && s.num_shadow_rays >= 0 && s.min_depth >= 0 && s.total_spp >= 0;
}

// Empties outlier tracking (an all-zero key is "no sample").
sspt_status clear_outliers(sspt_ctx* ctx, sspt_scene* sc) {
    if (!sc->outliers) return SSPT_OK;
    HIP_TRY(ctx, hipMemsetAsync(sc->outliers, 0,
                                (size_t)(sc->image_floats / 3) * sc->dev.num_outliers * sizeof(unsigned long long),
                                ctx->stream));
    return SSPT_OK;
}

void apply_settings(SceneRec& R, const sspt_pt_settings& s) {
    R.base_seed = s.base_seed;
    R.max_depth = s.max_depth;
    R.min_depth = s.min_depth;
    R.num_shadow_rays = s.num_shadow_rays;
    R.enable_bsdf_di = s.enable_bsdf_di;
    R.enable_denoiser = s.enable_denoiser;
}

uint32_t wave_capacity(const sspt_ctx* ctx) {
    uint32_t c = ctx->want_capacity ? ctx->want_capacity : kMaxSlots;
    return std::min(c, kMaxSlots);
}

// Wave buffers of sspt_render_noalloc carved out of the scratch array: the
// counters, the striped ray statistics, a block of wave descriptors, then the
// per-slot arrays of the largest wave (a multiple of 256 slots) that fits next
// to a shadow queue of 2 * NumShadowRays entries per slot.
constexpr uint32_t kScratchWaveDescs = 1024;
struct ScratchWaves {
    Wave wave{};
    WaveDesc* descs = nullptr;
    uint32_t capacity = 0;
    size_t shadow_entries = 0;
};
bool carve_scratch(unsigned char* base, int num_shadow_rays, ScratchWaves* out) {
    const size_t nsr = (size_t)std::max(1, num_shadow_rays);
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t fixed = up(kCntNum * sizeof(uint32_t)) + up(kStatBytes) + up(kScratchWaveDescs * sizeof(WaveDesc));
    // per slot: two extension queues (origin, direction, tag, RNG, prefix weight, previous
    // hit, radiance), the hit record, pixel, last vertex, three radiance sums, two
    // denoiser layers; the shadow queue's three float4 per entry
    const size_t per_slot = 2 * (16 + 16 + 4 + 4 + 16 + 16 + 16) + 16 + 4 + 6 * 16 + 2 * nsr * 48;
    const size_t arrays = 2 * 7 + 8 + 3;  // 256-byte alignment slack per array
    const size_t avail = (size_t)SSPT_NOALLOC_SCRATCH_BYTES - fixed - arrays * 256;
    size_t cap = avail / per_slot;
    cap -= cap % 256;
    cap = std::min<size_t>(cap, kMaxSlots);
    if (cap < 256) return false;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base + off;
        off += up(bytes);
        return p;
    };
    Wave& w = out->wave;
    w.counters = reinterpret_cast<uint32_t*>(take(kCntNum * sizeof(uint32_t)));
    w.stats = reinterpret_cast<unsigned long long*>(take(kStatBytes));
    out->descs = reinterpret_cast<WaveDesc*>(take(kScratchWaveDescs * sizeof(WaveDesc)));
    w.waves = out->descs;
    for (int b = 0; b < 2; ++b) {
        w.q_o[b] = reinterpret_cast<float4*>(take(cap * 16));
        w.q_d[b] = reinterpret_cast<float4*>(take(cap * 16));
        w.q_slot[b] = reinterpret_cast<uint32_t*>(take(cap * 4));
        w.q_rng[b] = reinterpret_cast<uint32_t*>(take(cap * 4));
        w.q_thr[b] = reinterpret_cast<float4*>(take(cap * 16));
        w.q_prev[b] = reinterpret_cast<float4*>(take(cap * 16));
        w.q_L[b] = reinterpret_cast<float4*>(take(cap * 16));
    }
    w.hits = reinterpret_cast<float4*>(take(cap * 16));
    w.st_pix = reinterpret_cast<uint32_t*>(take(cap * 4));
    w.st_prev = reinterpret_cast<float4*>(take(cap * 16));
    w.st_L = reinterpret_cast<float4*>(take(cap * 16));
    w.st_Le = reinterpret_cast<float4*>(take(cap * 16));
    w.st_La = reinterpret_cast<float4*>(take(cap * 16));
    w.st_alb = reinterpret_cast<float4*>(take(cap * 16));
    w.st_nrm = reinterpret_cast<float4*>(take(cap * 16));
    const size_t entries = cap * 2 * nsr;
    w.s_o = reinterpret_cast<float4*>(take(entries * 16));
    w.s_d = reinterpret_cast<float4*>(take(entries * 16));
    w.s_c = reinterpret_cast<float4*>(take(entries * 16));
    if (off > (size_t)SSPT_NOALLOC_SCRATCH_BYTES) return false;
    w.capacity = (uint32_t)cap;
    out->capacity = (uint32_t)cap;
    out->shadow_entries = entries;
    return true;
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
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        ctx->num_cus = cus;
    resident_blocks_per_cu(0, 0, &ctx->trace_per_cu, &ctx->shade_per_cu);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return SSPT_ERR_DEVICE;
    }
    // the allocation-free render's scratch on this device (a failure only
    // disables sspt_render_noalloc, which reports it)
    ctx->scratch_status = scratch_address(device, &ctx->scratch);
    *out = ctx;
    return SSPT_OK;
}

void sspt_ctx_destroy(sspt_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->timing.collect();
    for (void* p : ctx->wave_allocs) (void)hipFree(p);
    for (void* p : ctx->adj_allocs) (void)hipFree(p);
    if (ctx->wave_descs) (void)hipFree(ctx->wave_descs);
    if (ctx->lt_stats) (void)hipFree(ctx->lt_stats);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* sspt_last_error(const sspt_ctx* ctx) { return ctx ? ctx->error.c_str() : "null context"; }

sspt_status sspt_scene_create(sspt_ctx* ctx, const sspt_scene_desc* d, sspt_scene** out) {
    return create_world(ctx, d, 1, nullptr, 0, out);
}

sspt_status sspt_batch_create(sspt_ctx* ctx, const sspt_scene_desc* descs, uint32_t n, sspt_scene** out) {
    return create_world(ctx, descs, n, nullptr, 0, out);
}

sspt_status sspt_views_create(sspt_ctx* ctx, const sspt_scene_desc* desc, const sspt_camera_desc* cameras,
                              uint32_t num_views, sspt_scene** out) {
    if (!cameras) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    return create_world(ctx, desc, 1, cameras, num_views, out);
}

void sspt_scene_destroy(sspt_scene* sc) {
    if (!sc) return;
    (void)hipSetDevice(sc->ctx->device);
    (void)hipStreamSynchronize(sc->ctx->stream);
    for (void* p : sc->allocs) (void)hipFree(p);
    if (sc->outliers) (void)hipFree(sc->outliers);
    if (sc->bidir) (void)hipFree(sc->bidir);
    delete sc;
}

sspt_status sspt_scene_info(const sspt_scene* sc, uint64_t* ntri, uint64_t* nem, uint64_t* nnodes, uint64_t* bytes) {
    if (!sc) return SSPT_ERR_INVALID_ARGUMENT;
    if (ntri) *ntri = sc->num_tris;
    if (nem) *nem = sc->num_emitters;
    if (nnodes) *nnodes = sc->num_nodes;
    if (bytes) *bytes = sc->bytes;
    return SSPT_OK;
}

sspt_status sspt_set_bvh_builder(sspt_ctx* ctx, int32_t builder) {
    if (!ctx) return SSPT_ERR_INVALID_ARGUMENT;
    if (builder != SSPT_BVH_DEVICE && builder != SSPT_BVH_HOST)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "unknown BVH builder %d", builder);
    ctx->bvh_builder = builder;
    return SSPT_OK;
}

sspt_status sspt_get_bvh_builder(const sspt_ctx* ctx, int32_t* builder) {
    if (!ctx || !builder) return SSPT_ERR_INVALID_ARGUMENT;
    *builder = ctx->bvh_builder;
    return SSPT_OK;
}

sspt_status sspt_set_bvh_leaf_size(sspt_ctx* ctx, int32_t max_triangles) {
    if (!ctx) return SSPT_ERR_INVALID_ARGUMENT;
    if (max_triangles < 1 || max_triangles > 16)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "BVH leaf size %d outside 1..16", max_triangles);
    ctx->bvh_leaf_size = max_triangles;
    return SSPT_OK;
}

sspt_status sspt_get_bvh_leaf_size(const sspt_ctx* ctx, int32_t* max_triangles) {
    if (!ctx || !max_triangles) return SSPT_ERR_INVALID_ARGUMENT;
    *max_triangles = ctx->bvh_leaf_size;
    return SSPT_OK;
}

sspt_status sspt_scene_bvh_build_time(const sspt_scene* sc, int32_t* builder, double* kernel_ms, double* wall_ms) {
    if (!sc) return SSPT_ERR_INVALID_ARGUMENT;
    if (builder) *builder = sc->bvh_builder;
    if (kernel_ms) *kernel_ms = sc->bvh_kernel_ms;
    if (wall_ms) *wall_ms = sc->bvh_wall_ms;
    return SSPT_OK;
}

sspt_status sspt_scene_read_bvh(sspt_ctx* ctx, sspt_scene* sc, float* nodes, uint64_t max_nodes,
                                int32_t* leaf_tris, uint64_t max_tris) {
    if (!ctx || !sc || !nodes || !leaf_tris) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (max_nodes < sc->num_nodes || max_tris < sc->num_tris)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "buffers hold %llu nodes and %llu triangles, the scene has %llu and %llu",
                    (unsigned long long)max_nodes, (unsigned long long)max_tris, (unsigned long long)sc->num_nodes,
                    (unsigned long long)sc->num_tris);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::vector<float4> pos(sc->num_tris * 3);
    HIP_TRY(ctx, hipMemcpyAsync(nodes, sc->dev.nodes, sc->num_nodes * 32 * sizeof(float), hipMemcpyDeviceToHost,
                                ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(pos.data(), sc->dev.tri_pos, pos.size() * sizeof(float4), hipMemcpyDeviceToHost,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    // the first vertex record of each leaf triangle carries its global id in w
    for (uint64_t i = 0; i < sc->num_tris; ++i) std::memcpy(&leaf_tris[i], &pos[i * 3].w, 4);
    return SSPT_OK;
}

sspt_status sspt_scene_count(const sspt_scene* sc, uint32_t* n) {
    if (!sc || !n) return SSPT_ERR_INVALID_ARGUMENT;
    *n = (uint32_t)sc->recs.size();
    return SSPT_OK;
}

sspt_status sspt_framebuffer_clear(sspt_ctx* ctx, sspt_scene* sc) {
    if (!ctx || !sc) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemsetAsync(sc->image, 0, sc->image_floats * sizeof(float), ctx->stream));
    HIP_TRY(ctx, hipMemsetAsync(sc->layers, 0, 2 * sc->image_floats * sizeof(float), ctx->stream));
    TRY(clear_outliers(ctx, sc));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    sc->nan_count = 0;
    return SSPT_OK;
}

sspt_status sspt_framebuffer_track_outliers(sspt_ctx* ctx, sspt_scene* sc, uint32_t n) {
    if (!ctx || !sc) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (n > SSPT_MAX_OUTLIERS)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "at most %d outliers per pixel are tracked", SSPT_MAX_OUTLIERS);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (sc->outliers) {
        HIP_TRY(ctx, hipFree(sc->outliers));
        sc->outliers = nullptr;
    }
    sc->dev.outliers = nullptr;
    sc->dev.num_outliers = 0;
    if (n == 0) return SSPT_OK;
    const size_t bytes = (size_t)(sc->image_floats / 3) * n * sizeof(unsigned long long);
    if (hipMalloc((void**)&sc->outliers, bytes) != hipSuccess) {
        sc->outliers = nullptr;
        return fail(ctx, SSPT_ERR_OUT_OF_MEMORY, "hipMalloc of %zu bytes for outlier tracking failed", bytes);
    }
    sc->dev.outliers = sc->outliers;
    sc->dev.num_outliers = (int)n;
    TRY(clear_outliers(ctx, sc));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SSPT_OK;
}

sspt_status sspt_framebuffer_read_outliers(sspt_ctx* ctx, sspt_scene* sc, uint32_t index, float* weights,
                                           int32_t* iterations) {
    if (!ctx || !sc) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (index >= sc->recs.size()) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "scene index out of range");
    const int n = sc->dev.num_outliers;
    if (n == 0) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "outlier tracking is off");
    const SceneRec& R = sc->recs[index];
    const size_t count = (size_t)R.num_pixels * n;
    std::vector<unsigned long long> keys(count);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpyAsync(keys.data(), sc->outliers + (R.fb_off / 3) * n, count * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    // FrameBuffer.WriteToFile's order: the queue's entries in heap-array order
    // (UnorderedItems), stably sorted by descending weight (FrameBuffer.cs:321-325)
    std::vector<std::pair<float, int32_t>> q;
    for (uint32_t p = 0; p < R.num_pixels; ++p) {
        q.clear();
        for (int j = 0; j < n; ++j) {
            const unsigned long long e = keys[(size_t)p * n + j];
            if ((uint32_t)e == 0) continue;  // empty
            float w;
            const uint32_t b = (uint32_t)(e >> 32);
            std::memcpy(&w, &b, sizeof(w));
            q.push_back({w, (int32_t)((uint32_t)e - 1u)});
        }
        std::stable_sort(q.begin(), q.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        for (int j = 0; j < n; ++j) {
            const size_t o = (size_t)p * n + j;
            if (weights) weights[o] = j < (int)q.size() ? q[j].first : 0.0f;
            if (iterations) iterations[o] = j < (int)q.size() ? q[j].second : -1;
        }
    }
    return SSPT_OK;
}

sspt_status sspt_framebuffer_read_layer(sspt_ctx* ctx, sspt_scene* sc, uint32_t index, int32_t layer, float* host) {
    if (!ctx || !sc || !host) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (index >= sc->recs.size()) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "scene index out of range");
    if (layer != SSPT_LAYER_ALBEDO && layer != SSPT_LAYER_NORMAL)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "unknown layer %d", layer);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const SceneRec& R = sc->recs[index];
    const float* src = sc->layers + (layer == SSPT_LAYER_NORMAL ? sc->image_floats : 0) + R.fb_off;
    HIP_TRY(ctx, hipMemcpyAsync(host, src, (size_t)R.num_pixels * 3 * sizeof(float), hipMemcpyDeviceToHost,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SSPT_OK;
}

sspt_status sspt_framebuffer_read_scene(sspt_ctx* ctx, sspt_scene* sc, uint32_t index, float* host) {
    if (!ctx || !sc || !host) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (index >= sc->recs.size()) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "scene index out of range");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const SceneRec& R = sc->recs[index];
    HIP_TRY(ctx, hipMemcpyAsync(host, sc->image + R.fb_off, (size_t)R.num_pixels * 12, hipMemcpyDeviceToHost,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SSPT_OK;
}

sspt_status sspt_framebuffer_read(sspt_ctx* ctx, sspt_scene* sc, float* host) {
    return sspt_framebuffer_read_scene(ctx, sc, 0, host);
}

sspt_status sspt_framebuffer_nan_count(sspt_ctx* ctx, sspt_scene* sc, uint64_t* count) {
    if (!ctx || !sc || !count) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    *count = sc->nan_count;
    return SSPT_OK;
}

sspt_status sspt_render(sspt_ctx* ctx, sspt_scene* sc, const sspt_pt_settings* cfg, uint32_t first_iteration,
                        uint32_t num_iterations, sspt_ray_stats* stats) {
    if (!ctx || !sc || !cfg) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (!valid_settings(*cfg)) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "invalid settings");
    if (sc->recs.size() != 1) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "sspt_render renders a single scene");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    SceneRec& R = sc->recs[0];
    const uint32_t np = R.num_pixels;
    // waves of whole iterations (running mean), as many as fit the wave capacity;
    // one iteration per wave under a time budget, which is checked between waves
    uint32_t per_wave = std::max(1u, std::min(num_iterations, wave_capacity(ctx) / std::max(1u, np)));
    if (cfg->max_render_time_ms >= 0) per_wave = 1;
    uint32_t cap = np * per_wave;
    TRY(ensure_wave(ctx, cap));
    TRY(ensure_shadow(ctx, (size_t)cap * 2 * std::max(1, cfg->num_shadow_rays)));
    std::vector<WaveDesc> waves;
    for (uint32_t done = 0; done < num_iterations;) {
        uint32_t it = std::min(per_wave, num_iterations - done);
        waves.push_back(WaveDesc{(uint64_t)done * np, it * np, first_iteration + done, it, 0});
        done += it;
    }
    HIP_TRY(ctx, hipMemsetAsync(ctx->wave.stats, 0, kStatBytes, ctx->stream));
    uint32_t waves_run = 0;
    if (!waves.empty()) {
        TRY(ensure_wave_descs(ctx, waves.size()));
        apply_settings(R, *cfg);
        R.first_iter = first_iteration;
        R.path_off = 0;
        uint64_t offs[2] = {0, (uint64_t)num_iterations * np};
        HIP_TRY(ctx, hipMemcpyAsync(sc->d_recs, &R, sizeof(SceneRec), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(sc->d_path_offsets, offs, sizeof(offs), hipMemcpyHostToDevice, ctx->stream));
        float* mean_albedo = cfg->enable_denoiser ? sc->layers + R.fb_off : nullptr;
        TRY(run_waves(ctx, sc, waves, cfg->max_depth, cfg->num_shadow_rays, Mode::Mean, sc->image + R.fb_off, np,
                      mean_albedo, cfg->max_render_time_ms, &waves_run));
    }
    TRY(read_stats(ctx, sc, stats));
    if (stats) stats->num_iterations = (uint64_t)std::min(num_iterations, waves_run * per_wave);
    return SSPT_OK;
}

sspt_status sspt_render_noalloc(sspt_ctx* ctx, sspt_scene* sc, const sspt_pt_settings* cfg,
                                uint32_t first_iteration, uint32_t num_iterations, sspt_ray_stats* stats) {
    if (!ctx || !sc || !cfg) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (!valid_settings(*cfg)) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "invalid settings");
    if (sc->recs.size() != 1)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "sspt_render_noalloc renders a single scene");
    if (cfg->max_render_time_ms >= 0)
        return fail(ctx, SSPT_ERR_UNSUPPORTED, "sspt_render_noalloc takes no time budget (max_render_time_ms < 0)");
    if (!ctx->scratch)
        return fail(ctx, SSPT_ERR_DEVICE, "the scratch array of device %d is unavailable: %s", ctx->device,
                    hipGetErrorString(ctx->scratch_status));
    ScratchWaves sw;
    if (!carve_scratch(ctx->scratch, cfg->num_shadow_rays, &sw))
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT,
                    "num_shadow_rays %d leaves less than 256 path slots in the %u-byte scratch array",
                    cfg->num_shadow_rays, (unsigned)SSPT_NOALLOC_SCRATCH_BYTES);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    SceneRec& R = sc->recs[0];
    const uint32_t np = R.num_pixels;
    const uint64_t total = (uint64_t)num_iterations * np;
    // waves of whole iterations while one fits, else of at most one iteration's paths
    uint32_t per_wave, whole;
    if (np <= sw.capacity) {
        const uint32_t iters = std::max(1u, std::min(num_iterations, sw.capacity / np));
        per_wave = iters * np;
        whole = np;
    } else {
        const uint32_t parts = (np + sw.capacity - 1) / sw.capacity;
        per_wave = (np + parts - 1) / parts;
        whole = 0;
    }
    const uint64_t num_waves = total ? (total + per_wave - 1) / per_wave : 0;
    const bool timing = ctx->timing.enabled;
    ctx->timing.enabled = false;  // no event objects: kernels are launched directly
    const sspt_status st = [&]() -> sspt_status {
        HIP_TRY(ctx, hipMemsetAsync(sw.wave.stats, 0, kStatBytes, ctx->stream));
        if (!num_waves) return SSPT_OK;
        apply_settings(R, *cfg);
        R.first_iter = first_iteration;
        R.path_off = 0;
        sc->dev.recs_alike = 1;  // every path of this render is record 0's
        const uint64_t offs[2] = {0, total};
        HIP_TRY(ctx, hipMemcpyAsync(sc->d_recs, &R, sizeof(SceneRec), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(sc->d_path_offsets, offs, sizeof(offs), hipMemcpyHostToDevice, ctx->stream));
        LaunchConfig lc;
        lc.trace_blocks = ctx->trace_blocks;
        lc.shade_blocks = ctx->shade_blocks;
        lc.trace_block = ctx->trace_block;
        lc.shade_block = ctx->shade_block;
        lc.max_paths = per_wave;
        lc.max_nsr = std::max(1, cfg->num_shadow_rays);
        lc.max_shadow = (uint32_t)std::min<uint64_t>((uint64_t)per_wave * 2u * lc.max_nsr, sw.shadow_entries);
        float* mean_albedo = cfg->enable_denoiser ? sc->layers + R.fb_off : nullptr;
        for (uint64_t w0 = 0; w0 < num_waves; w0 += kScratchWaveDescs) {
            const uint32_t n = (uint32_t)std::min<uint64_t>(kScratchWaveDescs, num_waves - w0);
            // the descriptors of this block of waves, written on the device behind the previous
            // block's waves; the counters (queues and wave index) start at zero
            launch_fill_waves(ctx->stream, sw.descs, n, w0 * per_wave, per_wave, total,
                              first_iteration, whole);
            HIP_TRY(ctx, hipMemsetAsync(sw.wave.counters, 0, kCntNum * sizeof(uint32_t), ctx->stream));
            for (uint32_t i = 0; i < n; ++i)
                enqueue_wave(ctx, sc, sw.wave, lc, cfg->max_depth, Mode::Mean, sc->image + R.fb_off, np, mean_albedo);
        }
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        HIP_TRY(ctx, hipGetLastError());
        return SSPT_OK;
    }();
    ctx->timing.enabled = timing;
    if (st != SSPT_OK) return st;
    TRY(read_stats(ctx, sc, stats, sw.wave.stats));
    if (stats) stats->num_iterations = num_iterations;
    return SSPT_OK;
}

// Lays the paths of every scene of the world end to end (scene, then iteration,
// then pixel), cuts them into waves of the wave capacity and uploads the
// records and path offsets. Scene i renders iterations first_iteration ..
// first_iteration + total_spp_i - 1.
struct BatchPlan {
    std::vector<WaveDesc> waves;
    int max_depth = 1, max_nsr = 1;
};
sspt_status plan_batch(sspt_ctx* ctx, sspt_scene* sc, const sspt_pt_settings* cfgs, uint32_t first_iteration,
                       bool adjoint, BatchPlan* plan) {
    const size_t n = sc->recs.size();
    std::vector<uint64_t> offs(n + 1, 0);
    for (size_t i = 0; i < n; ++i) {
        if (!valid_settings(cfgs[i])) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "invalid settings for scene %zu", i);
        SceneRec& R = sc->recs[i];
        apply_settings(R, cfgs[i]);
        R.first_iter = first_iteration;
        R.path_off = offs[i];
        R.adj_scale = cfgs[i].total_spp > 0 ? 1.0f / (float)cfgs[i].total_spp : 0.0f;
        offs[i + 1] = offs[i] + (uint64_t)R.num_pixels * (uint64_t)cfgs[i].total_spp;
        plan->max_depth = std::max(plan->max_depth, cfgs[i].max_depth);
        plan->max_nsr = std::max(plan->max_nsr, cfgs[i].num_shadow_rays);
    }
    const uint64_t total = offs[n];
    uint32_t cap = wave_capacity(ctx);
    if (total < cap) cap = (uint32_t)std::max<uint64_t>(total, 1);
    TRY(ensure_wave(ctx, cap));
    TRY(ensure_shadow(ctx, (size_t)cap * 2 * plan->max_nsr));
    if (adjoint) TRY(ensure_adjoint(ctx, (uint32_t)plan->max_depth));
    for (uint64_t start = 0; start < total; start += cap)
        plan->waves.push_back(WaveDesc{start, (uint32_t)std::min<uint64_t>(cap, total - start), 0, 0, 0});
    if (!plan->waves.empty()) {
        TRY(ensure_wave_descs(ctx, plan->waves.size()));
        HIP_TRY(ctx, hipMemcpyAsync(sc->d_recs, sc->recs.data(), n * sizeof(SceneRec), hipMemcpyHostToDevice,
                                    ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(sc->d_path_offsets, offs.data(), offs.size() * sizeof(uint64_t),
                                    hipMemcpyHostToDevice, ctx->stream));
    }
    return SSPT_OK;
}

sspt_status sspt_render_batch(sspt_ctx* ctx, sspt_scene* sc, const sspt_pt_settings* cfgs, sspt_ray_stats* stats) {
    if (!ctx || !sc || !cfgs) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    BatchPlan plan;
    TRY(plan_batch(ctx, sc, cfgs, 0, false, &plan));
    HIP_TRY(ctx, hipMemsetAsync(ctx->wave.stats, 0, kStatBytes, ctx->stream));
    HIP_TRY(ctx, hipMemsetAsync(sc->image, 0, sc->image_floats * sizeof(float), ctx->stream));
    HIP_TRY(ctx, hipMemsetAsync(sc->layers, 0, 2 * sc->image_floats * sizeof(float), ctx->stream));
    TRY(clear_outliers(ctx, sc));
    sc->nan_count = 0;
    if (!plan.waves.empty()) {
        TRY(run_waves(ctx, sc, plan.waves, plan.max_depth, plan.max_nsr, Mode::Sum, nullptr, 0, nullptr, -1));
        // sums -> means
        for (size_t i = 0; i < sc->recs.size(); ++i) {
            if (cfgs[i].total_spp <= 0) continue;
            const uint64_t off = sc->recs[i].fb_off, n = (uint64_t)sc->recs[i].num_pixels * 3;
            const float inv = 1.0f / (float)cfgs[i].total_spp;
            launch_scale(ctx->stream, sc->image + off, n, inv);
            if (cfgs[i].enable_denoiser) {
                launch_scale(ctx->stream, sc->layers + off, n, inv);
                launch_scale(ctx->stream, sc->layers + sc->image_floats + off, n, inv);
            }
        }
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    TRY(read_stats(ctx, sc, stats));
    if (stats)
        for (size_t i = 0; i < sc->recs.size(); ++i)
            stats->num_iterations = std::max<uint64_t>(stats->num_iterations, (uint64_t)std::max(0, cfgs[i].total_spp));
    return SSPT_OK;
}

sspt_status sspt_render_adjoint_batch(sspt_ctx* ctx, sspt_scene* sc, const sspt_pt_settings* cfgs,
                                      const float* dloss_dimages, uint32_t first_iteration,
                                      sspt_ray_stats* stats) {
    if (!ctx || !sc || !cfgs || !dloss_dimages) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (ctx->exec_mode != SSPT_EXEC_WAVEFRONT)
        return fail(ctx, SSPT_ERR_UNSUPPORTED, "adjoint renders run on the wavefront tracer only");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    BatchPlan plan;
    TRY(plan_batch(ctx, sc, cfgs, first_iteration, true, &plan));
    HIP_TRY(ctx, hipMemsetAsync(ctx->wave.stats, 0, kStatBytes, ctx->stream));
    // host or device memory (unified addressing)
    HIP_TRY(ctx, hipMemcpyAsync(sc->adj_image, dloss_dimages, sc->image_floats * sizeof(float),
                                hipMemcpyDefault, ctx->stream));
    if (!plan.waves.empty())
        TRY(run_waves(ctx, sc, plan.waves, plan.max_depth, plan.max_nsr, Mode::Adjoint, nullptr, 0, nullptr, -1));
    TRY(read_stats(ctx, sc, stats));
    if (stats)
        for (size_t i = 0; i < sc->recs.size(); ++i)
            stats->num_iterations = std::max<uint64_t>(stats->num_iterations, (uint64_t)std::max(0, cfgs[i].total_spp));
    return SSPT_OK;
}

sspt_status sspt_render_adjoint(sspt_ctx* ctx, sspt_scene* sc, const sspt_pt_settings* cfg,
                                const float* dloss_dimage, uint32_t first_iteration, uint32_t num_iterations,
                                sspt_ray_stats* stats) {
    if (!ctx || !sc || !cfg || !dloss_dimage) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (!valid_settings(*cfg)) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "invalid settings");
    if (sc->recs.size() != 1) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "sspt_render_adjoint renders a single scene");
    if (ctx->exec_mode != SSPT_EXEC_WAVEFRONT)
        return fail(ctx, SSPT_ERR_UNSUPPORTED, "adjoint renders run on the wavefront tracer only");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    SceneRec& R = sc->recs[0];
    const uint32_t np = R.num_pixels;
    uint32_t per_wave = std::max(1u, std::min(num_iterations, wave_capacity(ctx) / std::max(1u, np)));
    uint32_t cap = np * per_wave;
    TRY(ensure_wave(ctx, cap));
    TRY(ensure_shadow(ctx, (size_t)cap * 2 * std::max(1, cfg->num_shadow_rays)));
    TRY(ensure_adjoint(ctx, (uint32_t)cfg->max_depth));
    std::vector<WaveDesc> waves;
    for (uint32_t done = 0; done < num_iterations;) {
        uint32_t it = std::min(per_wave, num_iterations - done);
        waves.push_back(WaveDesc{(uint64_t)done * np, it * np, first_iteration + done, it, 0});
        done += it;
    }
    HIP_TRY(ctx, hipMemsetAsync(ctx->wave.stats, 0, kStatBytes, ctx->stream));
    // host or device memory (unified addressing)
    HIP_TRY(ctx, hipMemcpyAsync(sc->adj_image + R.fb_off, dloss_dimage, (size_t)np * 12, hipMemcpyDefault,
                                ctx->stream));
    if (!waves.empty()) {
        TRY(ensure_wave_descs(ctx, waves.size()));
        apply_settings(R, *cfg);
        R.first_iter = first_iteration;
        R.path_off = 0;
        R.adj_scale = 1.0f / (float)num_iterations;
        uint64_t offs[2] = {0, (uint64_t)num_iterations * np};
        HIP_TRY(ctx, hipMemcpyAsync(sc->d_recs, &R, sizeof(SceneRec), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(sc->d_path_offsets, offs, sizeof(offs), hipMemcpyHostToDevice, ctx->stream));
        TRY(run_waves(ctx, sc, waves, cfg->max_depth, cfg->num_shadow_rays, Mode::Adjoint, nullptr, 0, nullptr, -1));
    }
    TRY(read_stats(ctx, sc, stats));
    if (stats) stats->num_iterations = num_iterations;
    return SSPT_OK;
}

sspt_status sspt_gradient_clear(sspt_ctx* ctx, sspt_scene* sc) {
    if (!ctx || !sc) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (sc->grad) HIP_TRY(ctx, hipMemsetAsync(sc->grad, 0, sc->texel_floats * sizeof(float), ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SSPT_OK;
}

static sspt_status texture_range(sspt_ctx* ctx, sspt_scene* sc, uint32_t index, uint64_t* off, uint64_t* n) {
    if (index >= sc->textures.size()) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "texture index out of range");
    const DevTexture& t = sc->textures[index];
    *off = t.offset;
    *n = (uint64_t)t.w * t.h * t.ch;
    return SSPT_OK;
}

sspt_status sspt_gradient_read(sspt_ctx* ctx, sspt_scene* sc, uint32_t texture_index, float* host) {
    if (!ctx || !sc || !host) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    uint64_t off, n;
    TRY(texture_range(ctx, sc, texture_index, &off, &n));
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpyAsync(host, sc->grad + off, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SSPT_OK;
}

sspt_status sspt_texture_write(sspt_ctx* ctx, sspt_scene* sc, uint32_t texture_index, const float* host) {
    if (!ctx || !sc || !host) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    uint64_t off, n;
    TRY(texture_range(ctx, sc, texture_index, &off, &n));
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpyAsync(const_cast<float*>(sc->dev.texels) + off, host, n * sizeof(float),
                                hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SSPT_OK;
}

sspt_status sspt_trace_closest(sspt_ctx* ctx, sspt_scene* sc, const float* rays, uint64_t n, sspt_hit* hits) {
    if (!ctx || !sc || (n && (!rays || !hits))) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (!n) return SSPT_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    float* dr = nullptr;
    float4* dh = nullptr;
    HIP_TRY(ctx, hipMalloc(&dr, n * 32));
    if (hipMalloc(&dh, n * 16) != hipSuccess) {
        (void)hipFree(dr);
        return fail(ctx, SSPT_ERR_OUT_OF_MEMORY, "hipMalloc failed");
    }
    hipError_t e = hipMemcpyAsync(dr, rays, n * 32, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        launch_trace_closest(ctx->stream, sc->dev, 0, dr, n, dh);
        e = hipMemcpyAsync(hits, dh, n * 16, hipMemcpyDeviceToHost, ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dr);
    (void)hipFree(dh);
    if (e != hipSuccess) return fail(ctx, SSPT_ERR_DEVICE, "trace failed: %s", hipGetErrorString(e));
    return SSPT_OK;
}

sspt_status sspt_trace_occluded(sspt_ctx* ctx, sspt_scene* sc, const float* rays, uint64_t n, uint8_t* occ) {
    if (!ctx || !sc || (n && (!rays || !occ))) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (!n) return SSPT_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    float* dr = nullptr;
    uint8_t* dh = nullptr;
    HIP_TRY(ctx, hipMalloc(&dr, n * 32));
    if (hipMalloc(&dh, n) != hipSuccess) {
        (void)hipFree(dr);
        return fail(ctx, SSPT_ERR_OUT_OF_MEMORY, "hipMalloc failed");
    }
    hipError_t e = hipMemcpyAsync(dr, rays, n * 32, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        launch_trace_occluded(ctx->stream, sc->dev, 0, dr, n, dh);
        e = hipMemcpyAsync(occ, dh, n, hipMemcpyDeviceToHost, ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dr);
    (void)hipFree(dh);
    if (e != hipSuccess) return fail(ctx, SSPT_ERR_DEVICE, "trace failed: %s", hipGetErrorString(e));
    return SSPT_OK;
}

sspt_status sspt_primary_hits(sspt_ctx* ctx, sspt_scene* sc, uint32_t base_seed, uint32_t iteration,
                              float* rays_out, sspt_hit* hits_out) {
    if (!ctx || !sc) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    uint64_t n = sc->recs[0].num_pixels;
    float* dr = nullptr;
    float4* dh = nullptr;
    HIP_TRY(ctx, hipMalloc(&dr, n * 32));
    if (hipMalloc(&dh, n * 16) != hipSuccess) {
        (void)hipFree(dr);
        return fail(ctx, SSPT_ERR_OUT_OF_MEMORY, "hipMalloc failed");
    }
    // the camera of the scene record; settings do not matter for primary rays
    hipError_t e = hipMemcpyAsync(sc->d_recs, sc->recs.data(), sizeof(SceneRec), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        launch_primary_rays(ctx->stream, sc->dev, 0, (uint32_t)n, base_seed, iteration, dr);
        launch_trace_closest(ctx->stream, sc->dev, 0, dr, n, dh);
        if (rays_out) e = hipMemcpyAsync(rays_out, dr, n * 32, hipMemcpyDeviceToHost, ctx->stream);
    }
    if (e == hipSuccess && hits_out) e = hipMemcpyAsync(hits_out, dh, n * 16, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dr);
    (void)hipFree(dh);
    if (e != hipSuccess) return fail(ctx, SSPT_ERR_DEVICE, "primary hits failed: %s", hipGetErrorString(e));
    return SSPT_OK;
}

sspt_status sspt_render_debug(sspt_ctx* ctx, sspt_scene* sc, uint32_t index, uint32_t base_seed,
                              uint32_t first_iteration, uint32_t num_iterations) {
    if (!ctx || !sc) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (index >= sc->recs.size()) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "scene index %u out of range", index);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const SceneRec& R = sc->recs[index];
    if (num_iterations == 0) return SSPT_OK;
    HIP_TRY(ctx, hipMemcpyAsync(sc->d_recs + index, &R, sizeof(SceneRec), hipMemcpyHostToDevice, ctx->stream));
    launch_debug_visualize(ctx->stream, sc->dev, (int)index, R.num_pixels, base_seed, first_iteration,
                           num_iterations, sc->image + R.fb_off);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SSPT_OK;
}

sspt_status sspt_render_light_tracer(sspt_ctx* ctx, sspt_scene* sc, const sspt_lt_settings* cfg,
                                     uint32_t first_iteration, uint32_t num_iterations, sspt_ray_stats* stats) {
    if (!ctx || !sc || !cfg) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (sc->recs.size() != 1)
        return fail(ctx, SSPT_ERR_UNSUPPORTED, "the light tracer renders a single scene, not a batch or view set");
    const SceneRec& R = sc->recs[0];
    if (R.has_env) {
        // With a background this configuration of VertexCacheBidir also traces camera
        // paths, which hit the background under EmitterHitMis (BidirBase.Camera.cs:532-561):
        // it is sspt_render_vcm's VertexCacheBidir with no connections, next events or
        // emitter hits, its camera paths seeded with BidirBase's default BaseSeedCamera.
        sspt_vcm_settings v{};
        v.algorithm = SSPT_VERTEX_CACHE_BIDIR;
        v.base_seed_camera = 0xC030114u;
        v.base_seed_light = cfg->base_seed_light;
        v.max_depth = cfg->max_depth;
        v.min_depth = cfg->min_depth;
        v.num_light_paths = cfg->num_light_paths;
        v.num_connections = 0;
        v.num_shadow_rays = 0;
        v.enable_light_tracer = 1;
        v.enable_hitting = 0;
        v.enable_merging = 0;
        v.max_num_photons = 8;
        return sspt_render_vcm(ctx, sc, &v, first_iteration, num_iterations, stats);
    }
    if (R.num_emitters == 0) return fail(ctx, SSPT_ERR_UNSUPPORTED, "the light tracer needs emitters");
    if (cfg->max_depth < 0 || cfg->min_depth < 0)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "max_depth and min_depth must be non-negative");
    const uint64_t paths = cfg->num_light_paths < 0 ? (uint64_t)R.num_pixels : (uint64_t)cfg->num_light_paths;
    if (paths > 0xFFFFFFFFull) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "too many light paths");
    if ((uint64_t)first_iteration + num_iterations > 0xFFFFFFFFull)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "iteration index overflows");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (!ctx->lt_stats) HIP_TRY(ctx, hipMalloc((void**)&ctx->lt_stats, kStatBytes));
    HIP_TRY(ctx, hipMemsetAsync(ctx->lt_stats, 0, kStatBytes, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(sc->d_recs, &R, sizeof(SceneRec), hipMemcpyHostToDevice, ctx->stream));
    float* image = sc->image + R.fb_off;
    for (uint32_t it = first_iteration; it < first_iteration + num_iterations; ++it) {
        // FrameBuffer.StartIteration / Normalize (FrameBuffer.cs:232-249)
        const float n = (float)(it + 1);
        if (it + 1 > 1) launch_scale(ctx->stream, image, 3ull * R.num_pixels, (n - 1.0f) / n);
        if (paths > 0) {
            LightParams P{cfg->base_seed_light, it, (uint32_t)paths, cfg->max_depth, cfg->min_depth, (float)paths, n};
            timed(ctx, "light_paths", [&] { launch_light_paths(ctx->stream, sc->dev, 0, P, image, ctx->lt_stats); });
        }
        HIP_TRY(ctx, hipGetLastError());
    }
    TRY(read_stats(ctx, sc, stats, ctx->lt_stats));
    if (stats) stats->num_iterations = num_iterations;
    return SSPT_OK;
}

sspt_status sspt_render_bidir(sspt_ctx* ctx, sspt_scene* sc, const sspt_bidir_settings* cfg,
                              uint32_t first_iteration, uint32_t num_iterations, sspt_ray_stats* stats) {
    if (!ctx || !sc || !cfg) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (sc->recs.size() != 1)
        return fail(ctx, SSPT_ERR_UNSUPPORTED, "bidirectional rendering takes a single scene, not a batch or view set");
    const SceneRec& R = sc->recs[0];
    if (cfg->max_depth < 0 || cfg->min_depth < 0 || cfg->num_shadow_rays < 0)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "max_depth, min_depth and num_shadow_rays must be non-negative");
    const uint64_t paths = (uint64_t)R.num_pixels;
    // ClassicBidir.Render (ClassicBidir.cs:72-75): one light path per camera path
    if (cfg->num_light_paths >= 0 && (uint64_t)cfg->num_light_paths != paths)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "ClassicBidir requires exactly one light path per pixel (%llu)",
                    (unsigned long long)paths);
    if ((uint64_t)first_iteration + num_iterations > 0xFFFFFFFFull)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "iteration index overflows");
    // light cache: max_depth levels of two float4 per path, the lengths and next-event pdfs;
    // camera pdfs: max_depth levels of a float2 per pixel. A path from the background
    // stores its first hit whatever MaxDepth (RandomWalk.StartFromBackground): two levels.
    const uint64_t levels = (uint64_t)std::max(cfg->max_depth, R.has_env ? 2 : 1);
    // the camera wavefront: per pixel its state (10 float4 and the RNG word), two list
    // entries, and per bounce max_depth - 1 connection and num_shadow_rays next-event
    // records of three float4 and a list entry each
    const uint64_t recs = (uint64_t)std::max(cfg->max_depth - 1, 0) + (uint64_t)cfg->num_shadow_rays;
    const uint64_t wave_bytes = paths * (10 * sizeof(float4) + 3 * 4) + paths * std::max<uint64_t>(recs, 1) * 52 +
                                32 * 256;  // and each of the 18 carves' alignment
    const uint64_t need = levels * paths * (2 * sizeof(float4) + sizeof(float2)) + paths * 8 + 256 * 4 + wave_bytes;
    if (need > (64ull << 30)) return fail(ctx, SSPT_ERR_OUT_OF_MEMORY, "bidirectional buffers exceed 64 GiB");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (sc->bidir_bytes < need) {
        if (sc->bidir) {
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            (void)hipFree(sc->bidir);
            sc->bidir = nullptr;
            sc->bidir_bytes = 0;
        }
        if (hipMalloc((void**)&sc->bidir, need) != hipSuccess) {
            sc->bidir = nullptr;
            return fail(ctx, SSPT_ERR_OUT_OF_MEMORY, "bidirectional buffers (%llu bytes)", (unsigned long long)need);
        }
        sc->bidir_bytes = need;
    }
    auto carve = [&](uint64_t* off, uint64_t bytes) {
        char* p = sc->bidir + *off;
        *off += (bytes + 255) & ~255ull;
        return p;
    };
    uint64_t off = 0;
    BidirParams P{};
    P.lv_a = reinterpret_cast<float4*>(carve(&off, levels * paths * sizeof(float4)));
    P.lv_b = reinterpret_cast<float4*>(carve(&off, levels * paths * sizeof(float4)));
    P.cam = reinterpret_cast<float2*>(carve(&off, levels * paths * sizeof(float2)));
    P.lv_len = reinterpret_cast<int*>(carve(&off, paths * sizeof(int)));
    P.lv_ne = reinterpret_cast<float*>(carve(&off, paths * sizeof(float)));
    const uint64_t nrec = paths * std::max<uint64_t>(recs, 1);
    P.bc_rng = reinterpret_cast<uint32_t*>(carve(&off, paths * 4));
    float4** state[] = {&P.bc_a, &P.bc_b, &P.bc_o, &P.bc_d, &P.bc_pp, &P.bc_pn, &P.bc_L, &P.bc_hit, &P.bc_v0, &P.bc_v1};
    for (float4** f : state) *f = reinterpret_cast<float4*>(carve(&off, paths * sizeof(float4)));
    P.bc_r0 = reinterpret_cast<float4*>(carve(&off, nrec * sizeof(float4)));
    P.bc_r1 = reinterpret_cast<float4*>(carve(&off, nrec * sizeof(float4)));
    P.bc_rc = reinterpret_cast<float4*>(carve(&off, nrec * sizeof(float4)));
    P.bc_list0 = reinterpret_cast<uint32_t*>(carve(&off, paths * 4));
    P.bc_list1 = reinterpret_cast<uint32_t*>(carve(&off, paths * 4));
    P.bc_shadow = reinterpret_cast<uint32_t*>(carve(&off, nrec * 4));
    P.bc_count = reinterpret_cast<uint32_t*>(carve(&off, 16));
    P.bc_rec = (int)recs;
    P.seed_camera = cfg->base_seed_camera;
    P.seed_light = cfg->base_seed_light;
    P.num_paths = (uint32_t)paths;
    P.num_pixels = (uint32_t)paths;
    P.mode = kModeClassic;
    P.env = R.has_env;
    P.max_depth = cfg->max_depth;
    P.min_depth = cfg->min_depth;
    P.num_shadow_rays = cfg->num_shadow_rays;
    P.connections = cfg->enable_connections ? 1 : 0;
    P.light_tracer = cfg->enable_light_tracer ? 1 : 0;
    P.num_paths_f = (float)paths;
    if (!ctx->lt_stats) HIP_TRY(ctx, hipMalloc((void**)&ctx->lt_stats, kStatBytes));
    HIP_TRY(ctx, hipMemsetAsync(ctx->lt_stats, 0, kStatBytes, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(sc->d_recs, &R, sizeof(SceneRec), hipMemcpyHostToDevice, ctx->stream));
    float* image = sc->image + R.fb_off;
    for (uint32_t it = first_iteration; it < first_iteration + num_iterations; ++it) {
        // FrameBuffer.StartIteration / Normalize (FrameBuffer.cs:232-249)
        const float n = (float)(it + 1);
        if (it + 1 > 1) launch_scale(ctx->stream, image, 3ull * R.num_pixels, (n - 1.0f) / n);
        P.iteration = it;
        P.n = n;
        if (paths > 0) {
            timed(ctx, "bidir_light", [&] { launch_bidir_light(ctx->stream, sc->dev, 0, P, image, ctx->lt_stats); });
            timed(ctx, "bidir_camera",
                  [&] { launch_bidir_camera_wave(ctx->stream, sc->dev, 0, P, image, ctx->lt_stats); });
        }
        HIP_TRY(ctx, hipGetLastError());
    }
    TRY(read_stats(ctx, sc, stats, ctx->lt_stats));
    if (stats) stats->num_iterations = num_iterations;
    return SSPT_OK;
}

sspt_status sspt_render_vcm(sspt_ctx* ctx, sspt_scene* sc, const sspt_vcm_settings* cfg, uint32_t first_iteration,
                            uint32_t num_iterations, sspt_ray_stats* stats) {
    if (!ctx || !sc || !cfg) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (sc->recs.size() != 1)
        return fail(ctx, SSPT_ERR_UNSUPPORTED, "bidirectional rendering takes a single scene, not a batch or view set");
    const SceneRec& R = sc->recs[0];
    if (cfg->algorithm != SSPT_VERTEX_CACHE_BIDIR && cfg->algorithm != SSPT_VCM)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "unknown algorithm %d", cfg->algorithm);
    if (cfg->max_depth < 0 || cfg->min_depth < 0 || cfg->num_shadow_rays < 0 || cfg->num_connections < 0)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT,
                    "max_depth, min_depth, num_shadow_rays and num_connections must be non-negative");
    const bool vcm = cfg->algorithm == SSPT_VCM;
    if (vcm && (cfg->max_num_photons < 1 || cfg->max_num_photons > kMaxPhotons))
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "max_num_photons must be in 1..%d", kMaxPhotons);
    const uint64_t pixels = (uint64_t)R.num_pixels;
    const uint64_t paths = cfg->num_light_paths < 0 ? pixels : (uint64_t)cfg->num_light_paths;
    const uint64_t levels = (uint64_t)std::max(cfg->max_depth, R.has_env ? 2 : 1);  // as in sspt_render_bidir
    if (paths * levels > 0x7FFFFFFFull)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "too many light path vertices (paths x max_depth >= 2^31)");
    if ((uint64_t)first_iteration + num_iterations > 0xFFFFFFFFull)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "iteration index overflows");
    const bool merging = vcm && cfg->enable_merging && cfg->max_depth >= 2;
    uint64_t table = 1024;
    while (merging && table < 2 * paths) table <<= 1;
    const uint64_t max_photons = merging ? (levels - 1) * paths : 0;
    const uint64_t scan_words = scan_tmp_words((uint32_t)std::max<uint64_t>(paths, table));
    auto align = [](uint64_t b) { return (b + 255) & ~255ull; };
    const uint64_t need = 2 * align(levels * paths * sizeof(float4)) + align(levels * pixels * sizeof(float2)) +
                          3 * align(paths * 4) + align(scan_words * 4) + 2 * align((table + 1) * 4) +
                          align(max_photons * sizeof(float4)) + 256;
    if (need > (64ull << 30)) return fail(ctx, SSPT_ERR_OUT_OF_MEMORY, "bidirectional buffers exceed 64 GiB");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (sc->bidir_bytes < need) {
        if (sc->bidir) {
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            (void)hipFree(sc->bidir);
            sc->bidir = nullptr;
            sc->bidir_bytes = 0;
        }
        if (hipMalloc((void**)&sc->bidir, need) != hipSuccess) {
            sc->bidir = nullptr;
            return fail(ctx, SSPT_ERR_OUT_OF_MEMORY, "bidirectional buffers (%llu bytes)", (unsigned long long)need);
        }
        sc->bidir_bytes = need;
    }
    uint64_t off = 0;
    auto carve = [&](uint64_t bytes) {
        char* p = sc->bidir + off;
        off += align(bytes);
        return p;
    };
    BidirParams P{};
    P.lv_a = reinterpret_cast<float4*>(carve(levels * paths * sizeof(float4)));
    P.lv_b = reinterpret_cast<float4*>(carve(levels * paths * sizeof(float4)));
    P.cam = reinterpret_cast<float2*>(carve(levels * pixels * sizeof(float2)));
    P.lv_len = reinterpret_cast<int*>(carve(paths * 4));
    P.lv_ne = reinterpret_cast<float*>(carve(paths * 4));
    P.cum = reinterpret_cast<int*>(carve(paths * 4));
    int* scan_tmp = reinterpret_cast<int*>(carve(scan_words * 4));
    P.bucket_start = reinterpret_cast<uint32_t*>(carve((table + 1) * 4));
    P.bucket_fill = reinterpret_cast<uint32_t*>(carve((table + 1) * 4));
    P.photons = reinterpret_cast<float4*>(carve(max_photons * sizeof(float4)));
    P.seed_camera = cfg->base_seed_camera;
    P.seed_light = cfg->base_seed_light;
    P.num_paths = (uint32_t)paths;
    P.num_pixels = (uint32_t)pixels;
    P.max_depth = cfg->max_depth;
    P.min_depth = cfg->min_depth;
    P.num_shadow_rays = cfg->num_shadow_rays;
    P.light_tracer = cfg->enable_light_tracer ? 1 : 0;
    P.num_paths_f = (float)paths;
    P.mode = vcm ? kModeVcm : kModeVertexCache;
    P.env = R.has_env;
    P.num_connections = cfg->num_connections;
    P.hitting = cfg->enable_hitting ? 1 : 0;
    P.merging = merging ? 1 : 0;
    P.merge_primary = cfg->merge_primary ? 1 : 0;
    P.max_photons = vcm ? cfg->max_num_photons : 1;
    P.correl = cfg->disable_correl_aware_mis ? 0 : 1;
    // grid cells of about a pixel's footprint at the scene's scale; any size finds the same photons
    P.cell = std::max(1e-6f, 1.7320508f * R.abs_max / (float)std::max(R.cam.width, R.cam.height));
    P.table_mask = (uint32_t)(table - 1);
    if (!ctx->lt_stats) HIP_TRY(ctx, hipMalloc((void**)&ctx->lt_stats, kStatBytes));
    HIP_TRY(ctx, hipMemsetAsync(ctx->lt_stats, 0, kStatBytes, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(sc->d_recs, &R, sizeof(SceneRec), hipMemcpyHostToDevice, ctx->stream));
    float* image = sc->image + R.fb_off;
    for (uint32_t it = first_iteration; it < first_iteration + num_iterations; ++it) {
        const float n = (float)(it + 1);
        if (it + 1 > 1) launch_scale(ctx->stream, image, 3ull * R.num_pixels, (n - 1.0f) / n);
        P.iteration = it;
        P.n = n;
        if (paths > 0)
            timed(ctx, "vcm_light", [&] { launch_bidir_light(ctx->stream, sc->dev, 0, P, image, ctx->lt_stats); });
        else
            HIP_TRY(ctx, hipMemsetAsync(P.lv_len, 0, 4, ctx->stream));
        timed(ctx, "vcm_prepare", [&] { launch_vcm_prepare(ctx->stream, sc->dev, 0, P, scan_tmp); });
        timed(ctx, "vcm_camera", [&] { launch_vcm_camera(ctx->stream, sc->dev, 0, P, image, ctx->lt_stats); });
        HIP_TRY(ctx, hipGetLastError());
    }
    TRY(read_stats(ctx, sc, stats, ctx->lt_stats));
    if (stats) stats->num_iterations = num_iterations;
    return SSPT_OK;
}

static_assert(sizeof(PathNode) == sizeof(sspt_path_node), "PathNode mirrors sspt_path_node");

sspt_status sspt_replay_pixel(sspt_ctx* ctx, sspt_scene* sc, uint32_t index, const sspt_pt_settings* cfg,
                              uint32_t col, uint32_t row, uint32_t iteration, sspt_path_node* host_nodes,
                              uint32_t capacity, uint32_t* num_nodes, float* estimate_rgb) {
    if (!ctx || !sc || !cfg || !num_nodes) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    if (capacity > 0 && !host_nodes) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null node array");
    if (!valid_settings(*cfg)) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "invalid settings");
    if (index >= sc->recs.size()) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "scene index %u out of range", index);
    SceneRec& R = sc->recs[index];
    if (col >= (uint32_t)R.cam.width || row >= (uint32_t)R.cam.height)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "pixel (%u, %u) outside the %dx%d framebuffer", col, row,
                    R.cam.width, R.cam.height);
    // the most nodes a path can add: root, per vertex a BSDF sample and two next events per shadow ray, background
    const uint64_t need = 2 + (uint64_t)cfg->max_depth * (1 + 2 * (uint64_t)cfg->num_shadow_rays);
    const uint32_t cap = (uint32_t)std::min<uint64_t>(capacity, need);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    apply_settings(R, *cfg);
    char* buf = nullptr;
    const size_t bytes = (size_t)std::max(1u, cap) * sizeof(PathNode) + 64;
    HIP_TRY(ctx, hipMalloc((void**)&buf, bytes));
    PathNode* d_nodes = reinterpret_cast<PathNode*>(buf + 64);
    uint32_t* d_count = reinterpret_cast<uint32_t*>(buf);
    float* d_est = reinterpret_cast<float*>(buf + 16);
    uint32_t count = 0;
    float est[3] = {0.0f, 0.0f, 0.0f};
    hipError_t e = hipMemcpyAsync(sc->d_recs + index, &R, sizeof(SceneRec), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        launch_replay(ctx->stream, sc->dev, (int)index, row * (uint32_t)R.cam.width + col, iteration, d_nodes, cap,
                      d_count, d_est);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&count, d_count, sizeof(count), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(est, d_est, sizeof(est), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && cap > 0)
        e = hipMemcpy(host_nodes, d_nodes, std::min(count, cap) * sizeof(PathNode), hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(ctx, SSPT_ERR_DEVICE, "replay failed: %s", hipGetErrorString(e));
    *num_nodes = count;
    if (estimate_rgb) std::memcpy(estimate_rgb, est, sizeof(est));
    return SSPT_OK;
}

sspt_status sspt_set_wave_capacity(sspt_ctx* ctx, uint32_t slots) {
    if (!ctx) return SSPT_ERR_INVALID_ARGUMENT;
    if (slots > kMaxSlots) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "wave capacity above %u", kMaxSlots);
    ctx->want_capacity = slots;
    return SSPT_OK;
}

sspt_status sspt_set_launch_config(sspt_ctx* ctx, uint32_t trace_blocks, uint32_t shade_blocks, int32_t use_graphs) {
    if (!ctx) return SSPT_ERR_INVALID_ARGUMENT;
    ctx->trace_blocks = trace_blocks;
    ctx->shade_blocks = shade_blocks;
    ctx->use_graphs = use_graphs != 0;
    return SSPT_OK;
}

sspt_status sspt_get_launch_config(sspt_ctx* ctx, uint32_t* trace_blocks, uint32_t* shade_blocks,
                                   uint32_t* trace_per_cu, uint32_t* shade_per_cu) {
    if (!ctx || !trace_blocks || !shade_blocks || !trace_per_cu || !shade_per_cu)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    *trace_blocks = ctx->trace_blocks;
    *shade_blocks = ctx->shade_blocks;
    *trace_per_cu = (uint32_t)std::max(0, ctx->trace_per_cu);
    *shade_per_cu = (uint32_t)std::max(0, ctx->shade_per_cu);
    return SSPT_OK;
}

sspt_status sspt_set_block_shape(sspt_ctx* ctx, uint32_t trace_block, uint32_t shade_block) {
    if (!ctx) return SSPT_ERR_INVALID_ARGUMENT;
    if (!valid_block_shape(trace_block, shade_block))
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT,
                    "block shape %u x %u: traversal blocks of 64, 128 or 256 threads, shade blocks of 128, 256 or 512",
                    trace_block, shade_block);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    ctx->trace_block = trace_block;
    ctx->shade_block = shade_block;
    resident_blocks_per_cu(trace_block, shade_block, &ctx->trace_per_cu, &ctx->shade_per_cu);
    return SSPT_OK;
}

sspt_status sspt_get_block_shape(sspt_ctx* ctx, uint32_t* trace_block, uint32_t* shade_block) {
    if (!ctx || !trace_block || !shade_block) return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "null argument");
    *trace_block = ctx->trace_block ? ctx->trace_block : (uint32_t)kTraceBlock;
    *shade_block = ctx->shade_block ? ctx->shade_block : 256u;
    return SSPT_OK;
}

sspt_status sspt_set_execution_mode(sspt_ctx* ctx, int32_t mode) {
    if (!ctx) return SSPT_ERR_INVALID_ARGUMENT;
    if (mode != SSPT_EXEC_WAVEFRONT && mode != SSPT_EXEC_SINGLE_KERNEL)
        return fail(ctx, SSPT_ERR_INVALID_ARGUMENT, "unknown execution mode %d", mode);
    ctx->exec_mode = mode;
    return SSPT_OK;
}

sspt_status sspt_set_kernel_timing(sspt_ctx* ctx, int32_t enable) {
    if (!ctx) return SSPT_ERR_INVALID_ARGUMENT;
    ctx->timing.enabled = enable != 0;
    return SSPT_OK;
}

sspt_status sspt_get_kernel_timing(sspt_ctx* ctx, const char* name, uint64_t* launches, double* total_ms) {
    if (!ctx || !name) return SSPT_ERR_INVALID_ARGUMENT;
    (void)hipStreamSynchronize(ctx->stream);
    ctx->timing.collect();
    auto it = ctx->timing.totals.find(name);
    if (launches) *launches = it == ctx->timing.totals.end() ? 0 : it->second.first;
    if (total_ms) *total_ms = it == ctx->timing.totals.end() ? 0.0 : it->second.second;
    return SSPT_OK;
}

sspt_status sspt_reset_kernel_timing(sspt_ctx* ctx) {
    if (!ctx) return SSPT_ERR_INVALID_ARGUMENT;
    (void)hipStreamSynchronize(ctx->stream);
    ctx->timing.collect();
    ctx->timing.totals.clear();
    return SSPT_OK;
}

}  // extern "C"
