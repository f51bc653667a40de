// gpr_internal.h -- host-side state shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gpr.h"

namespace gpr {

// ------------------------------------------------------------ errors ---
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// Runs an entry point's body and turns a C++ exception escaping it (an
// allocation failure while parsing a file, say) into a status code, so that
// no exception crosses the C ABI.
template <class F>
int guarded(const char* what, F&& f) noexcept {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(GPR_ERR_OOM, std::string(what) + ": host allocation failed");
    } catch (const std::exception& e) {
        return fail(GPR_ERR_INVALID, std::string(what) + ": " + e.what());
    } catch (...) {
        return fail(GPR_ERR_INVALID, std::string(what) + ": unexpected exception");
    }
}
#define GPR_HIP(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return ::gpr::fail(GPR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define GPR_TRY(expr)              \
    do {                           \
        int r_ = (expr);           \
        if (r_ != GPR_OK) return r_; \
    } while (0)

hipStream_t stream();
int device_id();

// Device allocation owned by RAII.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int alloc(size_t n) {
        if (n <= bytes && p) return GPR_OK;
        release();
        if (n == 0) return GPR_OK;
        if (hipMalloc(&p, n) != hipSuccess) {
            p = nullptr;
            return fail(GPR_ERR_OOM, "hipMalloc of " + std::to_string(n) + " bytes failed");
        }
        bytes = n;
        return GPR_OK;
    }
    template <class T>
    T* as() const { return reinterpret_cast<T*>(p); }
};

// 128-B BVH4 node: child boxes as SoA per axis, child references.
//   child >= 0             : BVH4 node index
//   child <  0, != EMPTY   : triangle record ~child (sorted-leaf order)
//   child == kEmptyChild   : unused slot (box is all +inf, never hit)
struct alignas(128) Bvh4Node {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(Bvh4Node) == 128, "BVH4 node must be one 128-B line");
constexpr int32_t kEmptyChild = (int32_t)0x80000000;

// Triangle record in sorted-leaf order, 48 B: p0 p1 p2 (as in the mesh) and the
// original triangle id.
struct alignas(16) TriRec {
    float4 a;  // p0.x p0.y p0.z p1.x
    float4 b;  // p1.y p1.z p2.x p2.y
    float4 c;  // p2.z id(bits) 0 0
};

// Shading record of one triangle in triangle-id order, one 128-B line: the
// corner positions and normals, the flags (1: corner normals, 2: corner uvs),
// the material index and the corner uvs.  Shading reads a hit or light
// triangle with a few 16-byte loads from one line instead of ~20 scattered
// 4-byte loads from four arrays; the values are copies, so every result
// computed from them is unchanged.
struct alignas(16) TriShade {
    float4 a;  // p0.x p0.y p0.z p1.x
    float4 b;  // p1.y p1.z p2.x p2.y
    float4 c;  // p2.z n0.x n0.y n0.z
    float4 d;  // n1.x n1.y n1.z n2.x
    float4 e;  // n2.y n2.z flags(bits) material(bits)
    float4 f;  // uv0.u uv0.v uv1.u uv1.v
    float4 g;  // uv2.u uv2.v 0 0
    float4 h;  // 0
};
static_assert(sizeof(TriShade) == 128, "shading record must be one 128-B line");

// Device-side view of a committed scene, passed by value to kernels.
struct SceneView {
    uint32_t ntri;
    uint32_t nlights;
    const float* pos;        // ntri*9, triangle-id order
    const float* nrm;        // ntri*9
    const uint8_t* flags;    // ntri
    const uint32_t* mat;     // ntri
    const gpr_material* mats;
    const uint32_t* lights;  // nlights triangle ids
    const gpr_ref_node* ref; // 2n-1, reference DFS layout
    const Bvh4Node* bvh4;
    const TriRec* tris;      // sorted-leaf order
    const TriShade* shade;   // ntri, triangle-id order
    int32_t root4;           // BVH4 root reference (node index or ~leaf)
    uint32_t nbvh4;          // BVH4 node count (bounds every node fetch)
    uint32_t nref;           // reference-layout node count (2n-1)
    uint32_t has_env;
    float env[3];
    // camera: rasterToCamera and cameraToWorld (row-major 4x4)
    float r2c[16];
    float c2w[16];
    uint32_t width, height;
    // free camera (camera.h:89-107): per pixel o = eye,
    // d = normalize(mid + (2 px/W - 1) h + (2 py/H - 1) v - eye)
    uint32_t cam_free;
    float cam_eye[3], cam_mid[3], cam_h[3], cam_v[3];
    // textures (texture.h:37-52): per-triangle corner uvs (read where flags&2
    // and the material has a texture), texture table, RGBA texel pool
    const float* uv;
    const gpr_texture* tex;
    const float4* texels;
    uint32_t ntex;
    // latitude-longitude environment (environment.h:16-28); 0 = constant env[]
    uint32_t env_tex;
    float env_w2o[9];
    float env_center[3];
};

struct BvhDev {
    uint32_t n = 0;
    uint32_t nbvh4 = 0;
    int32_t root4 = 0;
    DevBuf arena;      // all per-build scratch and results
    DevBuf sort_scratch;
    uint32_t* keys = nullptr;
    uint32_t* skeys = nullptr;
    uint32_t* sidx = nullptr;
    gpr_ref_node* ref = nullptr;
    Bvh4Node* bvh4 = nullptr;
    TriRec* tris = nullptr;
    float timing[5] = {0, 0, 0, 0, 0};
};

struct Mesh {
    std::vector<float> p, n, uv;
    std::vector<int32_t> corners;
    uint32_t ntri = 0;
    gpr_material mat{};
};

}  // namespace gpr

struct gpr_scene {
    int device = 0;
    std::atomic<uint32_t> renderers{0};  // live gpr_renderer handles on this scene: gpr_scene_free refuses while > 0
    std::vector<gpr::Mesh> meshes;
    float c2w[16];
    float fov = 40.f;
    uint32_t width = 256, height = 256;
    bool cam_set = false;
    // free camera state; defaults of Camera::Camera (camera.h:9-19)
    bool cam_free = false;
    float eye[3] = {0.f, 2.f, 7.f}, dir[3] = {0.f, -0.2f, -1.f}, up[3] = {0.f, 1.f, 0.f};
    float yaw = 0.f, pitch = 0.f, fovx = 0.f, fovy = 0.f;  // set by the constructor below
    gpr_scene();
    uint32_t has_env = 0;
    float env[3] = {0, 0, 0};
    uint32_t env_tex = 0;                                  // lat-long map texture handle
    float env_w2o[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float env_center[3] = {0, 0, 0};                       // world-box centroid (flatten)
    std::vector<gpr_texture> textures;
    std::vector<float> texels;                             // RGBA
    // flattened host copies
    std::vector<float> pos, nrm, uv;
    std::vector<uint8_t> flags;
    std::vector<uint32_t> mat, lights;
    std::vector<gpr_material> mats;
    // device copies
    gpr::DevBuf d_pos, d_nrm, d_flags, d_mat, d_mats, d_lights, d_stats, d_uv, d_tex, d_texels, d_shade;
    gpr::BvhDev bvh;
    bool flattened = false;
    bool committed = false;
    float r2c[16];
    uint64_t last_nv = 0, last_nt = 0;
    int trace_nt = 256, trace_sl = 16, trace_chunk = 64, trace_refill = 0;  // BVH4 traversal launch configuration
    gpr::SceneView view() const;
};

namespace gpr {
// bvh_build.hip
int build_bvh(gpr_scene* s);
// render.hip: packs the scene's TriShade records (d_shade) from its uploaded arrays
int pack_shade(gpr_scene* s);
// sort.hip
int sort_pairs(const uint32_t* keys, uint32_t* skeys, uint32_t* sidx, uint32_t n, void* scratch,
               size_t scratch_bytes, hipStream_t st);
size_t sort_scratch_bytes(uint32_t n);
int sort_prepare(void* scratch, size_t bytes, hipStream_t st);
// trace.hip
// nt / sl / chunk: block size, LDS stack entries, rays claimed per wave (the autotuner's space)
int launch_trace(const SceneView& sv, const gpr_ray* rays, gpr_hit* hits, uint64_t n, uint32_t flags,
                 unsigned long long* stats, hipStream_t st, int nt = 256, int sl = 16, int chunk = 64,
                 int refill = 0);
bool trace_cfg_valid(int nt, int sl, int chunk);
int launch_trace_stats(const SceneView& sv, const gpr_ray* rays, uint64_t n, bool any, unsigned long long* stats,
                       hipStream_t st);
int trace_release_stream(hipStream_t st);
// imageio.cpp: ReadImage -- width*height*3 linear floats, bottom row first
int read_image(const std::string& path, uint32_t* w, uint32_t* h, std::vector<float>* rgb);
// Gauss-Jordan inverse of a row-major 4x4 (transform.cpp:30-90); false when
// the result is not finite
bool invert4(const float a[16], float out[16]);
// host math shared by api.cpp and the renderer
void camera_raster_to_camera(uint32_t w, uint32_t h, float fov_deg, float out[16]);
void lookat_matrix(const float target[3], const float origin[3], const float up[3], float m[16]);
}  // namespace gpr
