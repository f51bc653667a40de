// api.cpp -- C ABI: runtime state, scene assembly, BVH / trace entry points and
// the seeded generators of the benchmark workloads.
//
// Scene assembly restates Scene::Scene / createDeviceData
// (src/renderer/scene.cpp:4-51): triangles are numbered in mesh order then
// face order, lights are the emissive triangles in that order, and every
// array the kernels read is uploaded once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "gpr_internal.h"

namespace gpr {

namespace {
// Last error, stream and device are per host thread, like HIP's current device.
thread_local std::string t_err;
thread_local hipStream_t g_stream = nullptr;
thread_local int g_device = 0;
}  // namespace

void set_error(const std::string& m) { t_err = m; }
int fail(int code, const std::string& m) {
    t_err = m;
    return code;
}
hipStream_t stream() { return g_stream; }
int device_id() { return g_device; }

int trace_overflow_count(unsigned long long* out);
int trace_overflow_accumulate(unsigned long long* acc, hipStream_t st);

// ------------------------------------------------------ host transforms --
// Matrix4x4 / Transform semantics of src/renderer/transform.{h,cpp}: float
// products summed left to right, Gauss-Jordan inverse with full pivoting.
namespace {
struct M4 {
    float m[4][4];
};
M4 m_ident() {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = i == j ? 1.f : 0.f;
    return r;
}
M4 m_mul(const M4& a, const M4& b) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] +
                        a.m[i][3] * b.m[3][j];
    return r;
}
M4 m_inverse(const M4& a) {
    int col_of[4], row_of[4];
    int used[4] = {0, 0, 0, 0};
    float w[4][4];
    std::memcpy(w, a.m, sizeof(w));
    for (int it = 0; it < 4; ++it) {
        int pr = 0, pc = 0;
        float best = 0.f;
        for (int r = 0; r < 4; ++r) {
            if (used[r] == 1) continue;
            for (int c = 0; c < 4; ++c) {
                if (used[c] != 0) continue;
                if (std::fabs(w[r][c]) >= best) {
                    best = std::fabs(w[r][c]);
                    pr = r;
                    pc = c;
                }
            }
        }
        ++used[pc];
        if (pr != pc)
            for (int k = 0; k < 4; ++k) std::swap(w[pr][k], w[pc][k]);
        row_of[it] = pr;
        col_of[it] = pc;
        const float inv = 1.f / w[pc][pc];
        w[pc][pc] = 1.f;
        for (int k = 0; k < 4; ++k) w[pc][k] *= inv;
        for (int r = 0; r < 4; ++r) {
            if (r == pc) continue;
            const float f = w[r][pc];
            w[r][pc] = 0;
            for (int k = 0; k < 4; ++k) w[r][k] -= w[pc][k] * f;
        }
    }
    for (int it = 3; it >= 0; --it)
        if (row_of[it] != col_of[it])
            for (int k = 0; k < 4; ++k) std::swap(w[k][row_of[it]], w[k][col_of[it]]);
    M4 r;
    std::memcpy(r.m, w, sizeof(w));
    return r;
}
struct Xf {
    M4 m, inv;
};
}  // namespace
// Inverse of a row-major 4x4; false when it is not finite (singular input).
bool invert4(const float a[16], float out[16]) {
    M4 m;
    std::memcpy(m.m, a, sizeof(m.m));
    const M4 r = m_inverse(m);
    std::memcpy(out, r.m, sizeof(r.m));
    for (int k = 0; k < 16; ++k)
        if (!std::isfinite(out[k])) return false;
    return true;
}
namespace {
Xf x_mul(const Xf& a, const Xf& b) { return Xf{m_mul(a.m, b.m), m_mul(b.inv, a.inv)}; }
Xf x_scale(float x, float y, float z) {
    Xf t{m_ident(), m_ident()};
    t.m.m[0][0] = x;
    t.m.m[1][1] = y;
    t.m.m[2][2] = z;
    t.inv.m[0][0] = (float)(1. / x);
    t.inv.m[1][1] = (float)(1. / y);
    t.inv.m[2][2] = (float)(1. / z);
    return t;
}
Xf x_translate(float x, float y, float z) {
    Xf t{m_ident(), m_ident()};
    t.m.m[0][3] = x;
    t.m.m[1][3] = y;
    t.m.m[2][3] = z;
    t.inv.m[0][3] = -x;
    t.inv.m[1][3] = -y;
    t.inv.m[2][3] = -z;
    return t;
}
Xf x_perspective(float fov_rad, float n, float f) {
    const float cot = (float)(1. / (double)std::tan(fov_rad * 0.5f));
    M4 p = m_ident();
    p.m[0][0] = cot;
    p.m[1][1] = cot;
    p.m[2][2] = f / (f - n);
    p.m[2][3] = -f * n / (f - n);
    p.m[3][2] = 1;
    p.m[3][3] = 0;
    return Xf{p, m_inverse(p)};
}
struct F3 {
    float x, y, z;
};
F3 f3sub(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
F3 f3cross(F3 a, F3 b) { return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
F3 f3norm(F3 v) {
    const float inv = 1.0f / std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return F3{v.x * inv, v.y * inv, v.z * inv};
}
}  // namespace

// Camera::setup (src/renderer/camera.h:21-37) non-free mode:
// rasterToCamera = Inverse(Scale(W,H,1) * Scale(-0.5, 0.5a, 1) * Translate(-1, 1/a, 0) * Perspective).
void camera_raster_to_camera(uint32_t w, uint32_t h, float fov_deg, float out[16]) {
    const float fov = fov_deg / 180 * 3.14159265358979323846f;
    const float aspect = float(w) / h;
    const Xf c2r = x_mul(x_mul(x_mul(x_scale((float)w, (float)h, 1), x_scale(-0.5f, (float)(0.5 * aspect), 1)),
                               x_translate(-1, (float)(1. / aspect), 0)),
                         x_perspective(fov, 1e-4f, 10000.f));
    std::memcpy(out, c2r.inv.m, sizeof(float) * 16);
}

// LookAt (src/renderer/transform.cpp:220-248)
void lookat_matrix(const float target[3], const float origin[3], const float up[3], float m[16]) {
    M4 c = m_ident();
    const F3 from{origin[0], origin[1], origin[2]}, to{target[0], target[1], target[2]},
        u{up[0], up[1], up[2]};
    c.m[0][3] = from.x;
    c.m[1][3] = from.y;
    c.m[2][3] = from.z;
    c.m[3][3] = 1;
    const F3 fwd = f3norm(f3sub(to, from));
    const F3 left = f3norm(f3cross(u, fwd));
    const F3 rup = f3cross(fwd, left);
    c.m[0][0] = left.x; c.m[1][0] = left.y; c.m[2][0] = left.z; c.m[3][0] = 0;
    c.m[0][1] = rup.x; c.m[1][1] = rup.y; c.m[2][1] = rup.z; c.m[3][1] = 0;
    c.m[0][2] = fwd.x; c.m[1][2] = fwd.y; c.m[2][2] = fwd.z; c.m[3][2] = 0;
    std::memcpy(m, c.m, sizeof(float) * 16);
}

// ------------------------------------------------------------- PCG32 ----
struct Pcg {
    uint64_t state, inc;
    Pcg(uint64_t seed, uint64_t stream) {
        state = 0;
        inc = (stream << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    float unit() { return float(next() >> 8) * (1.0f / 16777216.0f); }  // [0,1)
    double open01() { return (double(next() >> 8) + 0.5) * (1.0 / 16777216.0); }  // (0,1)
    float normal() {  // Box-Muller, one value per call
        const double u1 = open01(), u2 = open01();
        return (float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    }
};

}  // namespace gpr

using namespace gpr;

template <class T>
static int upload(DevBuf& b, const std::vector<T>& v) {
    GPR_TRY(b.alloc(std::max<size_t>(v.size() * sizeof(T), 256)));
    if (!v.empty())
        GPR_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream()));
    return GPR_OK;
}

// Camera::Camera (camera.h:9-19): eye (0,2,7), direction (0,-0.2,-1), up
// (0,1,0), 40-degree film, yaw / pitch of the direction.
gpr_scene::gpr_scene() {
    yaw = std::atan2(dir[0], dir[2]);
    pitch = std::asin(dir[1]);
    fovx = fovy = 40.f / 180 * 3.14159265358979323846f;
}

namespace {
F3 f3mul(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
F3 f3add(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
// film frame of the free camera (Camera::generateRay, camera.h:89-107)
void free_frame(const gpr_scene* s, SceneView& v) {
    const F3 d = f3norm(F3{s->dir[0], s->dir[1], s->dir[2]});
    const F3 up{s->up[0], s->up[1], s->up[2]};
    const F3 h = f3norm(f3cross(d, up));
    const F3 vv = f3norm(f3cross(h, d));
    const F3 eye{s->eye[0], s->eye[1], s->eye[2]};
    const F3 mid = f3add(eye, d);
    const F3 hs = f3mul(h, std::tan(s->fovx * 0.5f));
    const F3 vs = f3mul(vv, std::tan(s->fovy * 0.5f));
    const float* src[4] = {&eye.x, &mid.x, &hs.x, &vs.x};
    float* dst[4] = {v.cam_eye, v.cam_mid, v.cam_h, v.cam_v};
    for (int k = 0; k < 4; ++k) std::memcpy(dst[k], src[k], sizeof(float) * 3);
}
}  // namespace

gpr::SceneView gpr_scene::view() const {
    SceneView v{};
    v.ntri = (uint32_t)mat.size();
    v.nlights = (uint32_t)lights.size();
    v.pos = d_pos.as<float>();
    v.nrm = d_nrm.as<float>();
    v.flags = d_flags.as<uint8_t>();
    v.mat = d_mat.as<uint32_t>();
    v.mats = d_mats.as<gpr_material>();
    v.lights = d_lights.as<uint32_t>();
    v.ref = bvh.ref;
    v.bvh4 = bvh.bvh4;
    v.tris = bvh.tris;
    v.shade = d_shade.as<TriShade>();
    v.root4 = bvh.root4;
    v.nbvh4 = bvh.nbvh4;
    v.nref = bvh.n ? 2 * bvh.n - 1 : 0;
    v.has_env = has_env;
    v.env[0] = env[0];
    v.env[1] = env[1];
    v.env[2] = env[2];
    std::memcpy(v.r2c, r2c, sizeof(r2c));
    std::memcpy(v.c2w, c2w, sizeof(c2w));
    v.width = width;
    v.height = height;
    v.cam_free = cam_free ? 1u : 0u;
    if (cam_free) free_frame(this, v);
    v.uv = d_uv.as<float>();
    v.tex = d_tex.as<gpr_texture>();
    v.texels = d_texels.as<float4>();
    v.ntex = (uint32_t)textures.size();
    v.env_tex = has_env ? env_tex : 0u;
    std::memcpy(v.env_w2o, env_w2o, sizeof(env_w2o));
    std::memcpy(v.env_center, env_center, sizeof(env_center));
    return v;
}

extern "C" {

const char* gpr_last_error(void) { return t_err.c_str(); }

int gpr_set_device(int device) {
    // per-device state (traversal resources, sort scratch) is kept in tables of 16
    if (device < 0 || device >= 16) return fail(GPR_ERR_INVALID, "gpr_set_device: device index outside 0..15");
    GPR_HIP(hipSetDevice(device));
    g_device = device;
    return GPR_OK;
}
int gpr_set_stream(void* s) {
    g_stream = reinterpret_cast<hipStream_t>(s);
    return GPR_OK;
}
int gpr_synchronize(void) {
    GPR_HIP(hipStreamSynchronize(g_stream));
    return GPR_OK;
}

int gpr_scene_create(gpr_scene** out) {
    if (!out) return fail(GPR_ERR_INVALID, "gpr_scene_create: null out");
    gpr_scene* s = new (std::nothrow) gpr_scene();
    if (!s) return fail(GPR_ERR_OOM, "gpr_scene_create: host allocation failed");
    s->device = g_device;
    float id[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    std::memcpy(s->c2w, id, sizeof(id));
    *out = s;
    return GPR_OK;
}

int gpr_scene_add_mesh(gpr_scene* s, const float* p, uint32_t np, const float* n, uint32_t nn,
                       const float* uv, uint32_t nuv, const int32_t* corners, uint32_t ntri,
                       const gpr_material* mat) {
    if (!s || !mat || (ntri && (!p || !corners))) return fail(GPR_ERR_INVALID, "gpr_scene_add_mesh: null argument");
    if (s->committed) return fail(GPR_ERR_INVALID, "gpr_scene_add_mesh: scene already committed");
    if (mat->type > GPR_MAT_ROUGH_PLASTIC) return fail(GPR_ERR_UNSUPPORTED, "unknown material type");
    for (int k = 0; k < GPR_TEX_SLOTS; ++k)
        if (mat->tex[k] > s->textures.size())
            return fail(GPR_ERR_INVALID, "material texture handle " + std::to_string(mat->tex[k]) + " not in the scene");
    for (uint64_t t = 0; t < ntri; ++t) {
        const int32_t* c = corners + 9 * t;
        const bool hn = c[2] != -1, huv = c[1] != -1;
        for (int k = 0; k < 3; ++k) {
            const int32_t v = c[3 * k], vt = c[3 * k + 1], vn = c[3 * k + 2];
            if (v < 0 || (uint32_t)v >= np) return fail(GPR_ERR_INVALID, "corner vertex index out of range");
            if (huv ? (vt < 0 || (uint32_t)vt >= nuv) : vt != -1)
                return fail(GPR_ERR_INVALID, "corner uv index out of range / inconsistent");
            if (hn ? (vn < 0 || (uint32_t)vn >= nn) : vn != -1)
                return fail(GPR_ERR_INVALID, "corner normal index out of range / inconsistent");
        }
    }
    Mesh m;
    m.p.assign(p, p + 3ull * np);
    if (n && nn) m.n.assign(n, n + 3ull * nn);
    if (uv && nuv) m.uv.assign(uv, uv + 2ull * nuv);
    m.corners.assign(corners, corners + 9ull * ntri);
    m.ntri = ntri;
    m.mat = *mat;
    s->meshes.push_back(std::move(m));
    s->flattened = false;
    return GPR_OK;
}

int gpr_scene_set_camera(gpr_scene* s, const float c2w[16], float fov_deg, uint32_t w, uint32_t h) {
    if (!s || !c2w || w == 0 || h == 0 || !(fov_deg > 0 && fov_deg < 180))
        return fail(GPR_ERR_INVALID, "gpr_scene_set_camera: bad argument");
    std::memcpy(s->c2w, c2w, sizeof(float) * 16);
    s->fov = fov_deg;
    s->width = w;
    s->height = h;
    s->cam_set = true;
    s->cam_free = false;
    s->fovx = s->fovy = fov_deg / 180 * 3.14159265358979323846f;  // Camera::setup: make_float2(degToRad(fov))
    camera_raster_to_camera(w, h, fov_deg, s->r2c);
    return GPR_OK;
}

int gpr_scene_set_camera_free(gpr_scene* s, const float eye[3], const float dir[3], const float up[3],
                              float fov_deg, uint32_t w, uint32_t h) {
    if (!s || !eye || !dir || !up || w == 0 || h == 0 || !(fov_deg > 0 && fov_deg < 180))
        return fail(GPR_ERR_INVALID, "gpr_scene_set_camera_free: bad argument");
    if (!(dir[0] != 0 || dir[1] != 0 || dir[2] != 0)) return fail(GPR_ERR_INVALID, "gpr_scene_set_camera_free: zero direction");
    float ident[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    GPR_TRY(gpr_scene_set_camera(s, ident, fov_deg, w, h));
    std::memcpy(s->eye, eye, sizeof(float) * 3);
    std::memcpy(s->dir, dir, sizeof(float) * 3);
    std::memcpy(s->up, up, sizeof(float) * 3);
    s->yaw = std::atan2(dir[0], dir[2]);
    s->pitch = std::asin(dir[1] < -1.f ? -1.f : (dir[1] > 1.f ? 1.f : dir[1]));
    s->cam_free = true;
    return GPR_OK;
}

int gpr_scene_set_camera_lookat(gpr_scene* s, const float target[3], const float origin[3],
                                const float up[3], float fov_deg, uint32_t w, uint32_t h) {
    if (!target || !origin || !up) return fail(GPR_ERR_INVALID, "gpr_scene_set_camera_lookat: null vector");
    float m[16];
    lookat_matrix(target, origin, up, m);
    return gpr_scene_set_camera(s, m, fov_deg, w, h);
}

int gpr_scene_set_env(gpr_scene* s, const float rgb[3]) {
    if (!s || !rgb) return fail(GPR_ERR_INVALID, "gpr_scene_set_env: null argument");
    s->has_env = 1;
    s->env_tex = 0;
    std::memcpy(s->env, rgb, sizeof(float) * 3);
    return GPR_OK;
}

int gpr_scene_add_texture(gpr_scene* s, const float* rgb, uint32_t w, uint32_t h, const float uv_offset[2],
                          const float uv_scale[2], uint32_t* out_id) {
    if (!s || !rgb || !out_id || w == 0 || h == 0) return fail(GPR_ERR_INVALID, "gpr_scene_add_texture: bad argument");
    if (s->committed) return fail(GPR_ERR_INVALID, "gpr_scene_add_texture: scene already committed");
    const uint64_t n = (uint64_t)w * h, off = s->texels.size() / 4;
    if (off + n >= (1ull << 31)) return fail(GPR_ERR_UNSUPPORTED, "texel pool above 2^31 texels");
    gpr_texture t{};
    t.offset = (uint32_t)off;
    t.width = w;
    t.height = h;
    t.uv_offset[0] = uv_offset ? uv_offset[0] : 0.f;
    t.uv_offset[1] = uv_offset ? uv_offset[1] : 0.f;
    t.uv_scale[0] = uv_scale ? uv_scale[0] : 1.f;
    t.uv_scale[1] = uv_scale ? uv_scale[1] : 1.f;
    s->texels.resize(4 * (off + n), 0.f);
    for (uint64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) s->texels[4 * (off + i) + c] = rgb[3 * i + c];
    s->textures.push_back(t);
    *out_id = (uint32_t)s->textures.size();
    return GPR_OK;
}

int gpr_scene_add_checkerboard(gpr_scene* s, const float c0[3], const float c1[3], const float uv_offset[2],
                               const float uv_scale[2], uint32_t* out_id) {
    if (!c0 || !c1) return fail(GPR_ERR_INVALID, "gpr_scene_add_checkerboard: null colour");
    float px[12];
    for (int c = 0; c < 3; ++c) {
        px[c] = px[9 + c] = c0[c];
        px[3 + c] = px[6 + c] = c1[c];
    }
    return gpr_scene_add_texture(s, px, 2, 2, uv_offset, uv_scale, out_id);
}

int gpr_scene_set_env_map(gpr_scene* s, uint32_t tex_id, const float o2w[16]) {
    if (!s || tex_id == 0 || tex_id > s->textures.size())
        return fail(GPR_ERR_INVALID, "gpr_scene_set_env_map: bad scene or texture handle");
    // worldToObject = Inverse(o2w) (environment.cpp:11-15); transformVector
    // reads its upper-left 3x3 (transform.cpp:177-183)
    float inv[16];
    if (o2w) {
        if (!invert4(o2w, inv)) return fail(GPR_ERR_INVALID, "gpr_scene_set_env_map: singular toWorld");
    } else {
        const float id[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        std::memcpy(inv, id, sizeof(id));
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) s->env_w2o[3 * r + c] = inv[4 * r + c];
    s->has_env = 1;
    s->env_tex = tex_id;
    return GPR_OK;
}

int gpr_scene_export_textures(const gpr_scene* s, gpr_texture* recs, float* texels) {
    if (!s) return fail(GPR_ERR_INVALID, "gpr_scene_export_textures: null scene");
    if (recs && !s->textures.empty()) std::memcpy(recs, s->textures.data(), s->textures.size() * sizeof(gpr_texture));
    if (texels && !s->texels.empty()) std::memcpy(texels, s->texels.data(), s->texels.size() * sizeof(float));
    return GPR_OK;
}

// Host half of Scene::Scene (src/renderer/scene.cpp:7-20): triangles in mesh
// order then face order, per-triangle corner data gathered through the
// (v, vt, vn) corner table, emissive triangles collected as the light list.
int gpr_scene_flatten(gpr_scene* s) {
    if (!s) return fail(GPR_ERR_INVALID, "gpr_scene_flatten: null scene");
    if (s->flattened) return GPR_OK;
    uint64_t total = 0;
    for (const Mesh& m : s->meshes) total += m.ntri;
    if (total >= (1ull << 31)) return fail(GPR_ERR_UNSUPPORTED, "more than 2^31-1 triangles");
    s->pos.assign(9 * total, 0.f);
    s->nrm.assign(9 * total, 0.f);
    s->uv.assign(6 * total, 0.f);
    s->flags.assign(total, 0);
    s->mat.assign(total, 0);
    s->mats.clear();
    s->lights.clear();
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint64_t t = 0;
    for (size_t mi = 0; mi < s->meshes.size(); ++mi) {
        const Mesh& m = s->meshes[mi];
        s->mats.push_back(m.mat);
        for (uint32_t f = 0; f < m.ntri; ++f, ++t) {
            const int32_t* c = m.corners.data() + 9ull * f;
            uint8_t fl = 0;
            if (c[2] != -1) fl |= 1;
            if (c[1] != -1) fl |= 2;
            for (int k = 0; k < 3; ++k) {
                for (int a = 0; a < 3; ++a) s->pos[9 * t + 3 * k + a] = m.p[3ull * c[3 * k] + a];
                if (fl & 1)
                    for (int a = 0; a < 3; ++a) s->nrm[9 * t + 3 * k + a] = m.n[3ull * c[3 * k + 2] + a];
                if (fl & 2)
                    for (int a = 0; a < 2; ++a) s->uv[6 * t + 2 * k + a] = m.uv[2ull * c[3 * k + 1] + a];
            }
            s->flags[t] = fl;
            s->mat[t] = (uint32_t)mi;
            for (int k = 0; k < 9; ++k) {
                const float x = s->pos[9 * t + k];
                lo[k % 3] = std::fmin(lo[k % 3], x);
                hi[k % 3] = std::fmax(hi[k % 3], x);
            }
            const float* e = m.mat.emission;
            if (!(e[0] == 0 && e[1] == 0 && e[2] == 0)) s->lights.push_back((uint32_t)t);
        }
    }
    // world box centroid (scene.cpp:26-27, AABB::centroid aabb.h:23)
    for (int a = 0; a < 3; ++a) s->env_center[a] = total ? (lo[a] + hi[a]) * 0.5f : 0.f;
    s->flattened = true;
    return GPR_OK;
}

int gpr_scene_commit(gpr_scene* s) {
    if (!s) return fail(GPR_ERR_INVALID, "gpr_scene_commit: null scene");
    if (s->committed) return GPR_OK;
    GPR_TRY(gpr_scene_flatten(s));
    GPR_HIP(hipSetDevice(s->device));
    GPR_TRY(upload(s->d_pos, s->pos));
    GPR_TRY(upload(s->d_nrm, s->nrm));
    GPR_TRY(upload(s->d_flags, s->flags));
    GPR_TRY(upload(s->d_mat, s->mat));
    GPR_TRY(upload(s->d_mats, s->mats));
    GPR_TRY(upload(s->d_lights, s->lights));
    if (!s->textures.empty()) {
        GPR_TRY(upload(s->d_uv, s->uv));
        GPR_TRY(upload(s->d_tex, s->textures));
        GPR_TRY(upload(s->d_texels, s->texels));
    }
    GPR_TRY(pack_shade(s));
    GPR_TRY(s->d_stats.alloc(256));
    if (!s->cam_set) camera_raster_to_camera(s->width, s->height, s->fov, s->r2c);
    GPR_TRY(build_bvh(s));
    s->committed = true;
    return GPR_OK;
}

int gpr_scene_free(gpr_scene* s) {
    if (s && s->renderers.load() != 0)
        return fail(GPR_ERR_INVALID, "gpr_scene_free: " + std::to_string(s->renderers.load()) +
                                         " renderer(s) still use this scene; free them first");
    if (s) {
        (void)hipStreamSynchronize(stream());
        delete s;
    }
    return GPR_OK;
}

int gpr_scene_get_info(const gpr_scene* s, gpr_scene_info* o) {
    if (!s || !o) return fail(GPR_ERR_INVALID, "gpr_scene_get_info: null argument");
    std::memset(o, 0, sizeof(*o));
    uint64_t nt = 0;
    for (const Mesh& m : s->meshes) nt += m.ntri;
    o->ntri = (uint32_t)nt;
    o->nlights = (uint32_t)s->lights.size();
    o->nmat = (uint32_t)s->meshes.size();
    o->width = s->width;
    o->height = s->height;
    o->fov_deg = s->fov;
    std::memcpy(o->c2w, s->c2w, sizeof(o->c2w));
    o->has_env = s->has_env;
    std::memcpy(o->env, s->env, sizeof(o->env));
    o->bvh4_nodes = s->bvh.nbvh4;
    o->committed = s->committed ? 1u : 0u;
    o->cam_free = s->cam_free ? 1u : 0u;
    std::memcpy(o->eye, s->eye, sizeof(o->eye));
    std::memcpy(o->dir, s->dir, sizeof(o->dir));
    std::memcpy(o->up, s->up, sizeof(o->up));
    o->fov_x = s->fovx;
    o->fov_y = s->fovy;
    o->yaw = s->yaw;
    o->pitch = s->pitch;
    o->ntex = (uint32_t)s->textures.size();
    o->ntexel = (uint32_t)(s->texels.size() / 4);
    o->env_tex = s->has_env ? s->env_tex : 0u;
    std::memcpy(o->env_w2o, s->env_w2o, sizeof(o->env_w2o));
    std::memcpy(o->env_center, s->env_center, sizeof(o->env_center));
    return GPR_OK;
}

int gpr_scene_export_flat(const gpr_scene* s, float* pos, float* nrm, float* uv, uint8_t* flags,
                          uint32_t* mat, gpr_material* mats) {
    if (!s || !s->flattened) return fail(GPR_ERR_INVALID, "gpr_scene_export_flat: scene not flattened or committed");
    if (pos) std::memcpy(pos, s->pos.data(), s->pos.size() * 4);
    if (nrm) std::memcpy(nrm, s->nrm.data(), s->nrm.size() * 4);
    if (uv) std::memcpy(uv, s->uv.data(), s->uv.size() * 4);
    if (flags) std::memcpy(flags, s->flags.data(), s->flags.size());
    if (mat) std::memcpy(mat, s->mat.data(), s->mat.size() * 4);
    if (mats) std::memcpy(mats, s->mats.data(), s->mats.size() * sizeof(gpr_material));
    return GPR_OK;
}

int gpr_bvh_build(gpr_scene* s) {
    if (!s || !s->committed) return fail(GPR_ERR_INVALID, "gpr_bvh_build: scene not committed");
    GPR_HIP(hipSetDevice(s->device));
    return build_bvh(s);
}

int gpr_bvh_export_keys(const gpr_scene* s, uint32_t* keys, uint32_t* sk, uint32_t* si) {
    if (!s || !s->committed) return fail(GPR_ERR_INVALID, "gpr_bvh_export_keys: scene not committed");
    const size_t n = s->bvh.n;
    if (n == 0) return GPR_OK;
    if (keys) GPR_HIP(hipMemcpyAsync(keys, s->bvh.keys, 4 * n, hipMemcpyDeviceToHost, stream()));
    if (sk) GPR_HIP(hipMemcpyAsync(sk, s->bvh.skeys, 4 * n, hipMemcpyDeviceToHost, stream()));
    if (si) GPR_HIP(hipMemcpyAsync(si, s->bvh.sidx, 4 * n, hipMemcpyDeviceToHost, stream()));
    GPR_HIP(hipStreamSynchronize(stream()));
    return GPR_OK;
}

int gpr_bvh_export_ref_layout(const gpr_scene* s, gpr_ref_node* out, size_t cap) {
    if (!s || !s->committed || !out) return fail(GPR_ERR_INVALID, "gpr_bvh_export_ref_layout: bad argument");
    const size_t nn = s->bvh.n ? 2ull * s->bvh.n - 1 : 0;
    if (cap < nn) return fail(GPR_ERR_INVALID, "gpr_bvh_export_ref_layout: capacity below 2*ntri-1");
    if (nn) GPR_HIP(hipMemcpyAsync(out, s->bvh.ref, nn * sizeof(gpr_ref_node), hipMemcpyDeviceToHost, stream()));
    GPR_HIP(hipStreamSynchronize(stream()));
    return GPR_OK;
}

int gpr_bvh_last_timing(const gpr_scene* s, float ms[5]) {
    if (!s || !ms) return fail(GPR_ERR_INVALID, "gpr_bvh_last_timing: null argument");
    std::memcpy(ms, s->bvh.timing, sizeof(float) * 5);
    return GPR_OK;
}

namespace {
// Scratch of gpr_sort_pairs_device, one per (device, stream): sorts on one
// stream run in order and share it; sorts on different streams may overlap.
struct SortScratch {
    DevBuf b;
};
std::mutex g_sort_mu;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<SortScratch>> g_sort_scratch;
}  // namespace

int gpr_sort_pairs_device(const uint32_t* keys, uint32_t* sk, uint32_t* si, uint32_t n) {
    if (n == 0) return GPR_OK;
    if (!keys || !sk || !si) return fail(GPR_ERR_INVALID, "gpr_sort_pairs_device: null pointer");
    SortScratch* sc = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_sort_mu);
        std::unique_ptr<SortScratch>& p = g_sort_scratch[{g_device, stream()}];
        if (!p) p = std::make_unique<SortScratch>();
        sc = p.get();
    }
    const size_t need = sort_scratch_bytes(n);
    if (!sc->b.p || sc->b.bytes < need) {
        GPR_TRY(sc->b.alloc(need));
        GPR_TRY(sort_prepare(sc->b.p, sc->b.bytes, stream()));
    }
    return sort_pairs(keys, sk, si, n, sc->b.p, sc->b.bytes, stream());
}

namespace {
// Host-buffer traces above this many rays move in chunks, so that one chunk's
// copies overlap another's traversal.
constexpr uint64_t kHostChunk = 1ull << 20;

// Rays go in on one copy stream and hits come out on another; the chunks'
// traversals run in order on the caller's stream, each waiting for its rays
// and each copy-out waiting for its traversal.  The copy-out of chunk c-1 is
// issued after chunk c's copy-in and launch: a copy to or from pageable host
// memory can block the calling thread until it completes, and this order keeps
// the device busy meanwhile.  Results are those of one launch over the batch.
int trace_host_pipelined(gpr_scene* s, const SceneView& sv, const gpr_ray* rays, gpr_hit* hits, uint64_t n,
                         uint32_t flags, gpr_ray* dr, gpr_hit* dh, DevBuf& dov) {
    hipStream_t st = stream();
    const uint64_t nch = (n + kHostChunk - 1) / kHostChunk;
    GPR_TRY(dov.alloc(sizeof(unsigned long long)));
    hipStream_t cin = nullptr, cout = nullptr;
    std::vector<hipEvent_t> ev;
    auto cleanup = [&]() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        if (cin) (void)hipStreamDestroy(cin);
        if (cout) (void)hipStreamDestroy(cout);
    };
    auto run = [&]() -> int {
        GPR_HIP(hipStreamCreateWithFlags(&cin, hipStreamNonBlocking));
        GPR_HIP(hipStreamCreateWithFlags(&cout, hipStreamNonBlocking));
        ev.assign(2 * nch, nullptr);
        for (uint64_t k = 0; k < 2 * nch; ++k) GPR_HIP(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
        GPR_HIP(hipMemsetAsync(dov.p, 0, sizeof(unsigned long long), st));
        auto copy_out = [&](uint64_t c) -> int {
            const uint64_t off = c * kHostChunk, m = std::min(kHostChunk, n - off);
            GPR_HIP(hipStreamWaitEvent(cout, ev[nch + c], 0));
            GPR_HIP(hipMemcpyAsync(hits + off, dh + off, m * sizeof(gpr_hit), hipMemcpyDeviceToHost, cout));
            return GPR_OK;
        };
        for (uint64_t c = 0; c < nch; ++c) {
            const uint64_t off = c * kHostChunk, m = std::min(kHostChunk, n - off);
            GPR_HIP(hipMemcpyAsync(dr + off, rays + off, m * sizeof(gpr_ray), hipMemcpyHostToDevice, cin));
            GPR_HIP(hipEventRecord(ev[c], cin));
            GPR_HIP(hipStreamWaitEvent(st, ev[c], 0));
            GPR_TRY(launch_trace(sv, dr + off, dh + off, m, flags, nullptr, st, s->trace_nt, s->trace_sl,
                                 s->trace_chunk, s->trace_refill));
            GPR_TRY(trace_overflow_accumulate(dov.as<unsigned long long>(), st));
            GPR_HIP(hipEventRecord(ev[nch + c], st));
            if (c > 0) GPR_TRY(copy_out(c - 1));
        }
        GPR_TRY(copy_out(nch - 1));
        unsigned long long ov = 0;
        GPR_HIP(hipMemcpyAsync(&ov, dov.p, sizeof ov, hipMemcpyDeviceToHost, st));
        GPR_HIP(hipStreamSynchronize(cout));
        GPR_HIP(hipStreamSynchronize(st));
        if (ov) return fail(GPR_ERR_INVALID, "traversal stack capacity exceeded on " + std::to_string(ov) + " rays");
        return GPR_OK;
    };
    const int rc = run();
    if (rc != GPR_OK) {
        // nothing may still read or write the staging buffers when they are freed
        if (cin) (void)hipStreamSynchronize(cin);
        if (cout) (void)hipStreamSynchronize(cout);
        (void)hipStreamSynchronize(st);
    }
    cleanup();
    return rc;
}

// Device staging of host-buffer traces, one set per (device, stream), grown as
// needed and kept until the process ends.  A call holds its set's lock until
// its copies are done, so host-buffer traces on one stream from several
// threads take turns.
struct TraceStaging {
    std::mutex mu;
    DevBuf rays, hits, ov;
};
std::mutex g_stage_mu;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<TraceStaging>> g_stage;

TraceStaging* staging_for_stream() {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    std::unique_ptr<TraceStaging>& p = g_stage[{g_device, stream()}];
    if (!p) p = std::make_unique<TraceStaging>();
    return p.get();
}
}  // namespace

int gpr_trace(gpr_scene* s, const gpr_ray* rays, gpr_hit* hits, uint64_t n, uint32_t flags) {
    if (!s || !s->committed) return fail(GPR_ERR_INVALID, "gpr_trace: scene not committed");
    if (n == 0) return GPR_OK;
    if (!rays || !hits) return fail(GPR_ERR_INVALID, "gpr_trace: null rays / hits");
    const SceneView sv = s->view();
    unsigned long long* stats = s->d_stats.as<unsigned long long>();
    if (sv.ntri == 0) {
        // empty scene: every ray misses
        if (flags & GPR_TRACE_DEVICE_PTRS) {
            std::vector<gpr_hit> miss(n, gpr_hit{INFINITY, -1, 0.f, 0.f});
            GPR_HIP(hipMemcpyAsync(hits, miss.data(), n * sizeof(gpr_hit), hipMemcpyHostToDevice, stream()));
            GPR_HIP(hipStreamSynchronize(stream()));
        } else {
            for (uint64_t i = 0; i < n; ++i) hits[i] = gpr_hit{INFINITY, -1, 0.f, 0.f};
        }
        return GPR_OK;
    }
    if (flags & GPR_TRACE_DEVICE_PTRS)
        return launch_trace(sv, rays, hits, n, flags, stats, stream(), s->trace_nt, s->trace_sl, s->trace_chunk,
                            s->trace_refill);
    TraceStaging* sg = staging_for_stream();
    std::lock_guard<std::mutex> lk(sg->mu);
    DevBuf& dr = sg->rays;
    DevBuf& dh = sg->hits;
    GPR_TRY(dr.alloc(n * sizeof(gpr_ray)));
    GPR_TRY(dh.alloc(n * sizeof(gpr_hit)));
    if (!(flags & GPR_TRACE_REF_ORDER) && n > kHostChunk)
        return trace_host_pipelined(s, sv, rays, hits, n, flags, dr.as<gpr_ray>(), dh.as<gpr_hit>(), sg->ov);
    GPR_HIP(hipMemcpyAsync(dr.p, rays, n * sizeof(gpr_ray), hipMemcpyHostToDevice, stream()));
    GPR_TRY(launch_trace(sv, dr.as<gpr_ray>(), dh.as<gpr_hit>(), n, flags, stats, stream(), s->trace_nt,
                         s->trace_sl, s->trace_chunk, s->trace_refill));
    GPR_HIP(hipMemcpyAsync(hits, dh.p, n * sizeof(gpr_hit), hipMemcpyDeviceToHost, stream()));
    GPR_HIP(hipStreamSynchronize(stream()));
    unsigned long long ov = 0;
    if (!(flags & GPR_TRACE_REF_ORDER)) {
        GPR_TRY(trace_overflow_count(&ov));
        if (ov) return fail(GPR_ERR_INVALID, "traversal stack capacity exceeded on " + std::to_string(ov) + " rays");
    }
    return GPR_OK;
}

int gpr_trace_set_cfg(gpr_scene* s, const gpr_trace_cfg* c) {
    if (!s || !c) return fail(GPR_ERR_INVALID, "gpr_trace_set_cfg: null argument");
    if (c->layout != 0 || !trace_cfg_valid((int)c->block, (int)c->stack, (int)c->chunk))
        return fail(GPR_ERR_INVALID, "gpr_trace_set_cfg: configuration outside the tuned set");
    s->trace_nt = (int)c->block;
    s->trace_sl = (int)c->stack;
    s->trace_chunk = (int)c->chunk;
    s->trace_refill = c->refill ? 1 : 0;
    return GPR_OK;
}

int gpr_trace_tune(gpr_scene* s, const gpr_ray* rays, uint64_t n, uint32_t flags, gpr_trace_cfg* out, uint32_t cap,
                   uint32_t* nout) {
    if (!s || !s->committed || !rays || n == 0) return fail(GPR_ERR_INVALID, "gpr_trace_tune: bad argument");
    if (!(flags & GPR_TRACE_DEVICE_PTRS) || (flags & GPR_TRACE_REF_ORDER))
        return fail(GPR_ERR_INVALID, "gpr_trace_tune: needs device rays and the BVH4 traversal");
    if (s->mat.empty()) return fail(GPR_ERR_INVALID, "gpr_trace_tune: empty scene");
    const SceneView sv = s->view();
    DevBuf hits;
    GPR_TRY(hits.alloc(n * sizeof(gpr_hit)));
    hipStream_t st = stream();
    hipEvent_t e0, e1;
    GPR_HIP(hipEventCreate(&e0));
    GPR_HIP(hipEventCreate(&e1));
    std::vector<gpr_trace_cfg> res;
    int rc = GPR_OK;
    for (int refill : {0, 1})
        for (int nt : {64, 128, 256, 512})
            for (int sl : {8, 16, 32})
                for (int chunk : {64, 128, 256}) {
                    if (rc != GPR_OK) break;
                    rc = launch_trace(sv, rays, hits.as<gpr_hit>(), n, flags, nullptr, st, nt, sl, chunk, refill);
                    if (rc == GPR_OK && hipEventRecord(e0, st) != hipSuccess) rc = fail(GPR_ERR_HIP, "tune: event");
                    for (int r = 0; r < 2 && rc == GPR_OK; ++r)
                        rc = launch_trace(sv, rays, hits.as<gpr_hit>(), n, flags, nullptr, st, nt, sl, chunk, refill);
                    float ms = 0;
                    if (rc == GPR_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                         hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                        rc = fail(GPR_ERR_HIP, "tune: event timing");
                    if (rc == GPR_OK)
                        res.push_back(gpr_trace_cfg{(uint32_t)nt, (uint32_t)sl, (uint32_t)chunk, 0, (uint32_t)refill, ms / 2});
                }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    GPR_TRY(rc);
    unsigned long long ov = 0;
    GPR_TRY(trace_overflow_count(&ov));
    if (ov) return fail(GPR_ERR_INVALID, "gpr_trace_tune: traversal stack capacity exceeded");
    std::stable_sort(res.begin(), res.end(), [](const gpr_trace_cfg& a, const gpr_trace_cfg& b) { return a.ms < b.ms; });
    s->trace_nt = (int)res[0].block;
    s->trace_sl = (int)res[0].stack;
    s->trace_chunk = (int)res[0].chunk;
    s->trace_refill = (int)res[0].refill;
    if (nout) *nout = (uint32_t)std::min<size_t>(cap, res.size());
    if (out)
        for (uint32_t k = 0; k < cap && k < res.size(); ++k) out[k] = res[k];
    return GPR_OK;
}

int gpr_trace_ref_stats(const gpr_scene* s, uint64_t* nv, uint64_t* nt) {
    if (!s || !s->committed) return fail(GPR_ERR_INVALID, "gpr_trace_ref_stats: scene not committed");
    unsigned long long h[2] = {0, 0};
    GPR_HIP(hipMemcpyAsync(h, s->d_stats.p, sizeof(h), hipMemcpyDeviceToHost, stream()));
    GPR_HIP(hipStreamSynchronize(stream()));
    if (nv) *nv = h[0];
    if (nt) *nt = h[1];
    return GPR_OK;
}

int gpr_trace_bvh4_stats(gpr_scene* s, const gpr_ray* rays, uint64_t n, uint32_t flags, uint64_t* node_fetches,
                         uint64_t* tri_tests) {
    if (!s || !s->committed) return fail(GPR_ERR_INVALID, "gpr_trace_bvh4_stats: scene not committed");
    if (n && !rays) return fail(GPR_ERR_INVALID, "gpr_trace_bvh4_stats: null rays");
    if (!(flags & GPR_TRACE_DEVICE_PTRS) || (flags & GPR_TRACE_REF_ORDER))
        return fail(GPR_ERR_INVALID, "gpr_trace_bvh4_stats: needs device rays and the BVH4 traversal");
    unsigned long long h[3] = {0, 0, 0};
    if (s->view().ntri > 0 && n > 0) {
        unsigned long long* d = s->d_stats.as<unsigned long long>() + 8;
        GPR_TRY(launch_trace_stats(s->view(), rays, n, (flags & GPR_TRACE_ANY) != 0, d, stream()));
        GPR_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, stream()));
        GPR_HIP(hipStreamSynchronize(stream()));
    }
    if (h[2]) return fail(GPR_ERR_INVALID, "traversal stack capacity exceeded on " + std::to_string(h[2]) + " rays");
    if (node_fetches) *node_fetches = h[0];
    if (tri_tests) *tri_tests = h[1];
    return GPR_OK;
}

int gpr_trace_overflow_count(uint64_t* count) {
    if (!count) return fail(GPR_ERR_INVALID, "gpr_trace_overflow_count: null argument");
    unsigned long long ov = 0;
    GPR_TRY(trace_overflow_count(&ov));
    *count = ov;
    return GPR_OK;
}

int gpr_release_stream(void* hip_stream) {
    hipStream_t st = reinterpret_cast<hipStream_t>(hip_stream);
    GPR_HIP(hipStreamSynchronize(st));
    std::unique_ptr<TraceStaging> stage;
    std::unique_ptr<SortScratch> sort;
    {
        std::lock_guard<std::mutex> lk(g_stage_mu);
        auto it = g_stage.find({g_device, st});
        if (it != g_stage.end()) {
            // a host-buffer trace of another thread may still hold the set
            std::lock_guard<std::mutex> busy(it->second->mu);
            stage = std::move(it->second);
            g_stage.erase(it);
        }
    }
    {
        std::lock_guard<std::mutex> lk(g_sort_mu);
        auto it = g_sort_scratch.find({g_device, st});
        if (it != g_sort_scratch.end()) {
            sort = std::move(it->second);
            g_sort_scratch.erase(it);
        }
    }
    return trace_release_stream(st);
}

// ------------------------------------------------------- generators ----
int gpr_gen_soup(uint64_t seed, uint64_t stream_id, uint32_t n, float sigma, float* pos) {
    if (!pos && n) return fail(GPR_ERR_INVALID, "gpr_gen_soup: null output");
    Pcg g(seed, stream_id);
    for (uint64_t i = 0; i < n; ++i) {
        const float c[3] = {g.unit(), g.unit(), g.unit()};
        double z[10];
        for (int k = 0; k < 10; k += 2) {  // Box-Muller pairs
            const double u1 = g.open01(), u2 = g.open01();
            const double r = std::sqrt(-2.0 * std::log(u1));
            z[k] = r * std::cos(6.283185307179586 * u2);
            z[k + 1] = r * std::sin(6.283185307179586 * u2);
        }
        for (int v = 0; v < 3; ++v)
            for (int a = 0; a < 3; ++a) pos[9 * i + 3 * v + a] = c[a] + (float)(sigma * z[3 * v + a]);
    }
    return GPR_OK;
}

int gpr_gen_rays(uint64_t seed, uint64_t stream_id, uint64_t n, gpr_ray* rays) {
    if (!rays && n) return fail(GPR_ERR_INVALID, "gpr_gen_rays: null output");
    Pcg g(seed, stream_id);
    for (uint64_t i = 0; i < n; ++i) {
        gpr_ray& r = rays[i];
        r.o[0] = g.unit();
        r.o[1] = g.unit();
        r.o[2] = g.unit();
        const double z = 1.0 - 2.0 * g.unit();
        const double phi = 6.283185307179586 * g.unit();
        const double rr = std::sqrt(std::max(0.0, 1.0 - z * z));
        r.d[0] = (float)(rr * std::cos(phi));
        r.d[1] = (float)(rr * std::sin(phi));
        r.d[2] = (float)z;
        r.tmin = 0.f;
        r.tmax = INFINITY;
    }
    return GPR_OK;
}

int gpr_gen_shadow_rays(uint64_t seed, uint64_t stream_id, const gpr_ray* rays, uint64_t n, gpr_ray* out) {
    if ((!rays || !out) && n) return fail(GPR_ERR_INVALID, "gpr_gen_shadow_rays: null pointer");
    Pcg g(seed, stream_id);
    for (uint64_t i = 0; i < n; ++i) {
        const float q[3] = {g.unit(), g.unit(), g.unit()};
        gpr_ray r;
        for (int a = 0; a < 3; ++a) {
            r.o[a] = rays[i].o[a];
            r.d[a] = q[a] - rays[i].o[a];
        }
        r.tmin = 0.f;
        r.tmax = 0.9999f;  // shadow-ray convention (pathtracer_kernel.h:580-584)
        out[i] = r;
    }
    return GPR_OK;
}

}  // extern "C"

// ------------------------------------------------- procedural scenes ----
namespace {
struct MeshBuilder {
    std::vector<float> p, n;
    std::vector<int32_t> c;
    uint32_t ntri = 0;
    void quad(const float a[3], const float b[3], const float cc[3], const float d[3]) {
        const uint32_t base = (uint32_t)(p.size() / 3);
        for (const float* v : {a, b, cc, d}) p.insert(p.end(), v, v + 3);
        const uint32_t t[6] = {0, 1, 2, 0, 2, 3};
        for (uint32_t k : t) {
            c.push_back((int32_t)(base + k));
            c.push_back(-1);
            c.push_back(-1);
        }
        ntri += 2;
    }
    int add(gpr_scene* s, const gpr_material& m) {
        return gpr_scene_add_mesh(s, p.data(), (uint32_t)(p.size() / 3), n.empty() ? nullptr : n.data(),
                                  (uint32_t)(n.size() / 3), nullptr, 0, c.data(), ntri, &m);
    }
};
gpr_material lambert(float r, float g, float b) {
    gpr_material m{};
    m.type = GPR_MAT_DIFFUSE;
    m.color[0] = r;
    m.color[1] = g;
    m.color[2] = b;
    return m;
}
// axis-aligned box of half extents (hx, hy, hz) rotated about y and moved to c
void box_quads(MeshBuilder& mb, float cx, float cy, float cz, float hx, float hy, float hz, float deg) {
    const float a = deg / 180.f * 3.14159265358979323846f;
    const float ca = std::cos(a), sa = std::sin(a);
    auto P = [&](float x, float y, float z, float* o) {
        o[0] = cx + ca * x + sa * z;
        o[1] = cy + y;
        o[2] = cz - sa * x + ca * z;
    };
    float v[8][3];
    for (int i = 0; i < 8; ++i) P(i & 1 ? hx : -hx, i & 2 ? hy : -hy, i & 4 ? hz : -hz, v[i]);
    // faces wound counter-clockwise seen from outside
    mb.quad(v[4], v[5], v[7], v[6]);  // +z
    mb.quad(v[1], v[0], v[2], v[3]);  // -z
    mb.quad(v[5], v[1], v[3], v[7]);  // +x
    mb.quad(v[0], v[4], v[6], v[2]);  // -x
    mb.quad(v[6], v[7], v[3], v[2]);  // +y
    mb.quad(v[0], v[1], v[5], v[4]);  // -y
}
}  // namespace

extern "C" int gpr_scene_cornell(uint32_t width, uint32_t height, gpr_scene** out) {
    gpr_scene* s = nullptr;
    GPR_TRY(gpr_scene_create(&s));
    // room: 5 walls of the cube [-1,1]^3, open at +z
    const float A[3] = {-1, -1, -1}, B[3] = {1, -1, -1}, Cc[3] = {1, -1, 1}, D[3] = {-1, -1, 1};
    const float E[3] = {-1, 1, -1}, F[3] = {1, 1, -1}, G[3] = {1, 1, 1}, H[3] = {-1, 1, 1};
    MeshBuilder white, red, green, boxes, light;
    white.quad(D, Cc, B, A);  // floor
    white.quad(E, F, G, H);   // ceiling
    white.quad(A, B, F, E);   // back
    red.quad(D, A, E, H);     // left  (x = -1)
    green.quad(B, Cc, G, F);  // right (x = +1)
    box_quads(boxes, 0.35f, -1.f + 0.3f, 0.35f, 0.3f, 0.3f, 0.3f, -17.f);   // short box, 0.6 tall
    box_quads(boxes, -0.35f, -1.f + 0.6f, -0.3f, 0.3f, 0.6f, 0.3f, 17.f);   // tall box, 1.2 tall
    const float la[3] = {-0.25f, 0.99f, -0.25f}, lb[3] = {0.25f, 0.99f, -0.25f},
                lc[3] = {0.25f, 0.99f, 0.25f}, ld[3] = {-0.25f, 0.99f, 0.25f};
    light.quad(la, lb, lc, ld);  // normal -y (emits downwards)
    const gpr_material mw = lambert(0.725f, 0.71f, 0.68f), mr = lambert(0.63f, 0.065f, 0.05f),
                       mg = lambert(0.14f, 0.45f, 0.091f);
    gpr_material ml = lambert(0.78f, 0.78f, 0.78f);
    ml.emission[0] = ml.emission[1] = ml.emission[2] = 17.f;
    int rc = GPR_OK;
    if (rc == GPR_OK) rc = white.add(s, mw);
    if (rc == GPR_OK) rc = red.add(s, mr);
    if (rc == GPR_OK) rc = green.add(s, mg);
    if (rc == GPR_OK) rc = boxes.add(s, mw);
    if (rc == GPR_OK) rc = light.add(s, ml);
    const float tgt[3] = {0, 0, 0}, org[3] = {0, 0, 3.9f}, up[3] = {0, 1, 0};
    if (rc == GPR_OK) rc = gpr_scene_set_camera_lookat(s, tgt, org, up, 40.f, width, height);
    if (rc == GPR_OK) rc = gpr_scene_commit(s);
    if (rc != GPR_OK) {
        gpr_scene_free(s);
        return rc;
    }
    *out = s;
    return GPR_OK;
}

// UV sphere of create_sphere_triangles (src/renderer/loaders/mitsubaloader.cpp:153-221):
// nLat = subdiv rings, nLong = 2*subdiv segments, poles as single vertices,
// per-vertex normals, (nLat-1)*nLong*2 triangles.
static void uv_sphere(MeshBuilder& mb, float cx, float cy, float cz, float radius, int subdiv) {
    const int nlat = subdiv, nlong = 2 * subdiv;
    const float dlat = 3.14159265358979323846f / nlat, dlong = 2 * 3.14159265358979323846f / nlong;
    const uint32_t base = (uint32_t)(mb.p.size() / 3);
    auto vert = [&](float x, float y, float z) {
        mb.p.push_back(cx + x * radius);
        mb.p.push_back(cy + y * radius);
        mb.p.push_back(cz + z * radius);
        mb.n.push_back(x);
        mb.n.push_back(y);
        mb.n.push_back(z);
    };
    vert(0, 0, 1);
    for (int i = 1; i < nlat; ++i) {
        const float th = i * dlat;
        for (int j = 0; j < nlong; ++j) {
            const float ph = j * dlong;
            vert(std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th));
        }
    }
    vert(0, 0, -1);
    auto tri = [&](int a, int b, int c) {
        for (int v : {a, b, c}) {
            mb.c.push_back((int32_t)(base + v));
            mb.c.push_back(-1);
            mb.c.push_back((int32_t)(base + v));
        }
        ++mb.ntri;
    };
    for (int i = 0; i < nlong; ++i) {
        const int a = i + 1, b = i == nlong - 1 ? 1 : a + 1;
        tri(a, b, 0);
    }
    for (int i = 2; i < nlat; ++i)
        for (int j = 0; j < nlong; ++j) {
            const int a = (i - 1) * nlong + j + 1;
            int b = a + 1, c = a - nlong, d = c + 1;
            if (j == nlong - 1) {
                b = (i - 1) * nlong + 1;
                d = b - nlong;
            }
            tri(a, b, c);
            tri(c, b, d);
        }
    const int bottom = nlong * (nlat - 1) + 1;
    for (int i = 0; i < nlong; ++i) {
        const int a = (nlat - 2) * nlong + i + 1;
        const int b = i == nlong - 1 ? (nlat - 2) * nlong + 1 : a + 1;
        tri(a, bottom, b);
    }
}

// 8x8 UV spheres with the 8 material classes, ground plane and two quad lights.
static int add_sphere_grid(gpr_scene* s, uint32_t subdiv) {
    // 8 material classes: GGX gold / copper at four roughnesses, smooth and
    // rough glass (IOR 1.5), two Lambert colours.
    const float gold_eta[3] = {0.200438f, 0.924033f, 1.10221f}, gold_k[3] = {3.91295f, 2.45285f, 2.14219f};
    const float cu_eta[3] = {0.27105f, 0.67693f, 1.31640f}, cu_k[3] = {3.60920f, 2.62480f, 2.29210f};
    gpr_material mats[8];
    const float rough[4] = {0.01f, 0.1f, 0.25f, 0.5f};
    for (int i = 0; i < 4; ++i) {
        gpr_material m{};
        m.type = GPR_MAT_ROUGH_CONDUCTOR;
        m.color[0] = m.color[1] = m.color[2] = 1.f;
        m.alpha_x = m.alpha_y = rough[i];
        const float* e = (i & 1) ? cu_eta : gold_eta;
        const float* k = (i & 1) ? cu_k : gold_k;
        for (int a = 0; a < 3; ++a) {
            m.eta_i[a] = 1.f;
            m.eta_t[a] = e[a];
            m.k[a] = k[a];
        }
        mats[i] = m;
    }
    gpr_material glass{};
    glass.type = GPR_MAT_SPECULAR;
    glass.color[0] = glass.color[1] = glass.color[2] = 1.f;
    glass.ior = 1.5f;
    mats[4] = glass;
    gpr_material rglass = glass;
    rglass.type = GPR_MAT_ROUGH_DIELECTRIC;
    rglass.alpha_x = rglass.alpha_y = 0.1f;
    mats[5] = rglass;
    mats[6] = lambert(0.8f, 0.3f, 0.2f);
    mats[7] = lambert(0.2f, 0.5f, 0.8f);
    int rc = GPR_OK;
    for (int cls = 0; cls < 8 && rc == GPR_OK; ++cls) {
        MeshBuilder mb;
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 8; ++j)
                if ((i + j) % 8 == cls) uv_sphere(mb, -3.5f + i, 0.45f, -3.5f + j, 0.45f, (int)subdiv);
        rc = mb.add(s, mats[cls]);
    }
    MeshBuilder ground, l1, l2;
    const float g0[3] = {-10, 0, -10}, g1[3] = {-10, 0, 10}, g2[3] = {10, 0, 10}, g3[3] = {10, 0, -10};
    ground.quad(g0, g1, g2, g3);  // normal +y
    const float a0[3] = {-4, 6, -2}, a1[3] = {0, 6, -2}, a2[3] = {0, 6, 2}, a3[3] = {-4, 6, 2};
    l1.quad(a0, a1, a2, a3);  // normal -y
    const float b0[3] = {2.5f, 5, 1.5f}, b1[3] = {3.5f, 5, 1.5f}, b2[3] = {3.5f, 5, 2.5f}, b3[3] = {2.5f, 5, 2.5f};
    l2.quad(b0, b1, b2, b3);
    gpr_material lw = lambert(0.5f, 0.5f, 0.5f);
    gpr_material e1 = lambert(0.5f, 0.5f, 0.5f), e2 = lambert(0.5f, 0.5f, 0.5f);
    e1.emission[0] = e1.emission[1] = e1.emission[2] = 10.f;
    e2.emission[0] = e2.emission[1] = e2.emission[2] = 50.f;
    if (rc == GPR_OK) rc = ground.add(s, lw);
    if (rc == GPR_OK) rc = l1.add(s, e1);
    if (rc == GPR_OK) rc = l2.add(s, e2);
    return rc;
}

extern "C" int gpr_scene_sphere_grid(uint32_t width, uint32_t height, uint32_t subdiv, gpr_scene** out) {
    if (subdiv < 2) return fail(GPR_ERR_INVALID, "gpr_scene_sphere_grid: subdiv < 2");
    gpr_scene* s = nullptr;
    GPR_TRY(gpr_scene_create(&s));
    int rc = add_sphere_grid(s, subdiv);
    const float tgt[3] = {0, 0, 0}, org[3] = {0, 6, 10}, up[3] = {0, 1, 0};
    if (rc == GPR_OK) rc = gpr_scene_set_camera_lookat(s, tgt, org, up, 40.f, width, height);
    if (rc == GPR_OK) rc = gpr_scene_commit(s);
    if (rc != GPR_OK) {
        gpr_scene_free(s);
        return rc;
    }
    *out = s;
    return GPR_OK;
}

// Sphere grid sized to a triangle budget plus a random triangle soup: spheres
// take (1 - soup_fraction) of ntri (64 spheres of 4 s (s-1) triangles), the
// soup the rest, as small Lambert triangles with centres uniform in
// [-4,4] x [0,2] x [-4,4] and N(0, 0.03^2) vertex offsets (PCG32 seed, stream 7).
extern "C" int gpr_scene_mix(uint64_t seed, uint32_t ntri, float soup_fraction, gpr_scene** out) {
    if (!out || ntri < 1000 || !(soup_fraction >= 0.f && soup_fraction <= 1.f))
        return fail(GPR_ERR_INVALID, "gpr_scene_mix: ntri < 1000 or soup_fraction outside [0, 1]");
    const double sphere_tris = (1.0 - soup_fraction) * ntri;
    uint32_t sub = (uint32_t)std::lround(0.5 + std::sqrt(0.25 + sphere_tris / 256.0));
    if (sub < 2) sub = 2;
    const uint64_t grid_tris = 256ull * sub * (sub - 1) + 6;
    const uint32_t soup = grid_tris >= ntri ? 0u : (uint32_t)(ntri - grid_tris);
    gpr_scene* s = nullptr;
    GPR_TRY(gpr_scene_create(&s));
    int rc = add_sphere_grid(s, sub);
    if (rc == GPR_OK && soup > 0) {
        Pcg g(seed, 7);
        std::vector<float> p(9ull * soup);
        for (uint32_t t = 0; t < soup; ++t) {
            const float cx = -4.f + 8.f * g.unit(), cy = 2.f * g.unit(), cz = -4.f + 8.f * g.unit();
            for (int v = 0; v < 3; ++v) {
                p[9ull * t + 3 * v + 0] = cx + 0.03f * g.normal();
                p[9ull * t + 3 * v + 1] = cy + 0.03f * g.normal();
                p[9ull * t + 3 * v + 2] = cz + 0.03f * g.normal();
            }
        }
        std::vector<int32_t> c(9ull * soup, -1);
        for (uint64_t k = 0; k < 3ull * soup; ++k) c[3 * k] = (int32_t)k;
        const gpr_material m = lambert(0.6f, 0.6f, 0.6f);
        rc = gpr_scene_add_mesh(s, p.data(), 3 * soup, nullptr, 0, nullptr, 0, c.data(), soup, &m);
    }
    const float tgt[3] = {0, 0, 0}, org[3] = {0, 6, 10}, up[3] = {0, 1, 0};
    if (rc == GPR_OK) rc = gpr_scene_set_camera_lookat(s, tgt, org, up, 40.f, 256, 256);
    if (rc == GPR_OK) rc = gpr_scene_commit(s);
    if (rc != GPR_OK) {
        gpr_scene_free(s);
        return rc;
    }
    *out = s;
    return GPR_OK;
}
