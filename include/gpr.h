/*
 * gpr.h -- C ABI of the MI355X-native wavefront path-tracing core.
 *
 * Drop-in boundary for the path that slongle/GPU-Renderer exposes through
 * in-process C++ types.  Each entry point names the reference interface it
 * replaces (file:line in the reference tree).  Conventions:
 *   - every call returns an int status: GPR_OK (0) or a negative GPR_ERR_*;
 *     gpr_last_error() gives the message of the last failing call on this
 *     thread (the reference prints and exit(1)s: src/utility/buffer.h:3-12);
 *   - handles are opaque; device memory is owned by the library; host
 *     buffers are owned by the caller;
 *   - one handle is used from one host thread at a time;
 *   - work is issued on the stream set with gpr_set_stream (default: the
 *     null stream) and every call that returns host data synchronizes it.
 */
#ifndef GPR_H
#define GPR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPR_OK 0
#define GPR_ERR_INVALID -1     /* bad argument / state */
#define GPR_ERR_OOM -2         /* device or host allocation failed */
#define GPR_ERR_HIP -3         /* HIP runtime error */
#define GPR_ERR_IO -4          /* file could not be read / written */
#define GPR_ERR_PARSE -5       /* scene file malformed */
#define GPR_ERR_UNSUPPORTED -6 /* feature outside the supported subset */
#define GPR_ERR_BUDGET -7      /* request exceeds the device-memory budget */
#define GPR_ERR_CANCELLED -8   /* served request cancelled before it finished */

/* Ray: origin, direction, [tmin, tmax] -- Ray (src/renderer/ray.h:5-11), 32 B. */
typedef struct {
    float o[3];
    float d[3];
    float tmin;
    float tmax;
} gpr_ray;

/* Hit: t, triangle id (-1 = miss), barycentrics -- Hit (src/renderer/ray.h:13-19), 16 B.
 * Misses are returned as {t=+inf, id=-1, u=0, v=0}. */
typedef struct {
    float t;
    int32_t id;
    float u;
    float v;
} gpr_hit;

/* Reference BVH node layout -- BVHLinearNode (src/renderer/bvh.h:34-43), 32 B:
 * AABB, union{triangle index (leaf), right child index (interior)}, leaf flag
 * in the low byte of `leaf` (upper bytes zero).  Left child = own index + 1. */
typedef struct {
    float bmin[3];
    float bmax[3];
    uint32_t idx;
    uint32_t leaf;
} gpr_ref_node;

/* Material -- Material (src/renderer/material.h:6-35), factories in
 * src/renderer/loaders/mitsubaloader.cpp:289-441.  Every parameter is a
 * TextureView in the reference; here the constant fields hold the value and
 * tex[slot] (a handle from gpr_scene_add_texture, 0 = constant) replaces a
 * field by a nearest-texel lookup at the hit's uv.  Float parameters (ior,
 * alpha_x, alpha_y) read the red channel (TextureView::evalFloat). */
#define GPR_MAT_DIFFUSE 0          /* Lambert */
#define GPR_MAT_SPECULAR 1         /* smooth dielectric / mirror (FresnelSpecular) */
#define GPR_MAT_ROUGH_CONDUCTOR 2  /* GGX reflection, conductor Fresnel */
#define GPR_MAT_ROUGH_DIELECTRIC 3 /* GGX reflection + transmission (Walter 2007) */
#define GPR_MAT_ROUGH_PLASTIC 4    /* Lambert + GGX, dielectric Fresnel */
typedef struct {
    uint32_t type;
    float color[3];    /* reflectance / specularReflectance / diffuseReflectance */
    float color1[3];   /* plastic: specularReflectance */
    float emission[3]; /* area-light radiance (0 = not a light) */
    float ior;
    float alpha_x, alpha_y;
    float eta_i[3], eta_t[3], k[3];
    uint32_t tex[8];   /* GPR_TEX_* slot -> texture handle (0 = constant field) */
} gpr_material;
#define GPR_TEX_COLOR 0
#define GPR_TEX_COLOR1 1
#define GPR_TEX_IOR 2
#define GPR_TEX_ALPHA_X 3
#define GPR_TEX_ALPHA_Y 4
#define GPR_TEX_ETA_I 5
#define GPR_TEX_ETA_T 6
#define GPR_TEX_K 7
#define GPR_TEX_SLOTS 8

/* Texture record (TextureView, src/renderer/texture.h:14-35): texels
 * [offset, offset + width*height) of the scene's RGBA float pool, row-major,
 * row 0 at v = 0 (the reference's images are flipped vertically on load). */
typedef struct {
    uint32_t offset;
    uint32_t width, height;
    uint32_t pad;
    float uv_offset[2];
    float uv_scale[2];
} gpr_texture;

typedef struct gpr_scene gpr_scene;
typedef struct gpr_renderer gpr_renderer;

/* ------------------------------------------------------------ runtime --- */
const char* gpr_last_error(void);
/* Select the HIP device used by handles created afterwards on the calling
 * host thread (default 0; devices 0..15, GPR_ERR_INVALID beyond).  Like
 * hipSetDevice, the setting is per host thread. */
int gpr_set_device(int device);
/* Stream (hipStream_t) for all subsequent work issued from the calling host
 * thread; NULL = null stream.  Per host thread. */
int gpr_set_stream(void* hip_stream);
int gpr_synchronize(void);

/* -------------------------------------------------------------- scene --- */
/* Scene(filename) (src/renderer/scene.cpp:4-28) is split into create / add /
 * commit so that procedural and loaded scenes share one path. */
int gpr_scene_create(gpr_scene** out);
/* One TriangleMesh (src/renderer/triangle.h:28-49): positions, optional
 * normals and uvs, corner table of (v, vt, vn) int triples (-1 = absent),
 * per-mesh material.  Triangles are appended in mesh order, then face
 * order, which defines the triangle id (scene.cpp:7-14). */
int gpr_scene_add_mesh(gpr_scene* s, const float* p, uint32_t np, const float* n, uint32_t nn,
                       const float* uv, uint32_t nuv, const int32_t* corners, uint32_t ntri,
                       const gpr_material* mat);
/* Camera::setup (src/renderer/camera.h:21-37), Mitsuba raster mode. c2w is
 * the row-major cameraToWorld matrix. */
int gpr_scene_set_camera(gpr_scene* s, const float c2w[16], float fov_deg, uint32_t width,
                         uint32_t height);
/* Free (GUI) camera: Camera::setup with free = true and the eye / direction /
 * up state of Camera (camera.h:9-19, 89-107); rays through eye + dir with
 * tan(fov/2)-scaled film axes.  fov_deg sets both film axes, as setup does;
 * gpr_renderer_resize rescales the vertical one. */
int gpr_scene_set_camera_free(gpr_scene* s, const float eye[3], const float dir[3], const float up[3],
                              float fov_deg, uint32_t width, uint32_t height);
/* LookAt (src/renderer/transform.cpp:220-248) + Camera::setup. */
int gpr_scene_set_camera_lookat(gpr_scene* s, const float target[3], const float origin[3],
                                const float up[3], float fov_deg, uint32_t width, uint32_t height);
/* Constant environment light (ParseRecord::createLight "constant",
 * mitsubaloader.cpp:467-471). */
int gpr_scene_set_env(gpr_scene* s, const float rgb[3]);
/* Texture (src/renderer/texture.cpp:16-62): width*height RGB texels,
 * row-major, row 0 at v = 0.  uv_offset / uv_scale may be NULL (0 / 1).
 * *out_id receives the handle (>= 1) for gpr_material.tex[].  Textures must be
 * added before the meshes that use them. */
int gpr_scene_add_texture(gpr_scene* s, const float* rgb, uint32_t width, uint32_t height,
                          const float uv_offset[2], const float uv_scale[2], uint32_t* out_id);
/* Checkerboard texture: 2x2 texels c0 c1 / c1 c0 (texture.cpp:50-62). */
int gpr_scene_add_checkerboard(gpr_scene* s, const float c0[3], const float c1[3], const float uv_offset[2],
                               const float uv_scale[2], uint32_t* out_id);
/* Latitude-longitude environment light (EnvironmentLight, environment.h:10-53,
 * ParseRecord::createLight "envmap", mitsubaloader.cpp:472-478): texture
 * tex_id, o2w the row-major toWorld matrix (NULL = identity).  Replaces a
 * constant environment light and vice versa. */
int gpr_scene_set_env_map(gpr_scene* s, uint32_t tex_id, const float o2w[16]);
/* Host copies of the texture table (ntex records) and the texel pool
 * (ntexel RGBA floats), sizes in gpr_scene_info.  Either pointer may be NULL. */
int gpr_scene_export_textures(const gpr_scene* s, gpr_texture* recs, float* texels_rgba);
/* ReadImage (src/renderer/imageio.cpp:45-57, stbi_loadf with a vertical
 * flip): PNG (8/16-bit, non-interlaced), Radiance .hdr and binary PNM (P5/P6)
 * -> *rgb, width*height*3 linear floats, row 0 = bottom row of the file.  8-bit
 * data is linearised as (c/255)^2.2, 16-bit PNG samples keep their high byte
 * first.  Free with gpr_image_free. */
int gpr_image_read(const char* path, float** rgb, uint32_t* width, uint32_t* height);
void gpr_image_free(float* rgb);
/* Flatten meshes, upload, build the light list (scene.cpp:16-20) and the
 * BVH on the device (replaces LBVH_build, src/renderer/bvh.h:186-221). */
int gpr_scene_commit(gpr_scene* s);
/* Host half of gpr_scene_commit: flatten the meshes into per-triangle arrays
 * and collect the emissive triangles (scene.cpp:7-20), without device work. */
int gpr_scene_flatten(gpr_scene* s);
/* Releases the scene; GPR_ERR_INVALID (nothing freed) while renderers created
 * on it are still alive. */
int gpr_scene_free(gpr_scene* s);

typedef struct {
    uint32_t ntri;
    uint32_t nlights;
    uint32_t nmat;
    uint32_t width, height;
    float fov_deg;
    float c2w[16];
    uint32_t has_env;
    float env[3];
    uint32_t bvh4_nodes; /* packed 128-B traversal nodes */
    uint32_t committed;
    uint32_t cam_free;   /* 1 = free camera (eye / dir / up below), 0 = cameraToWorld */
    float eye[3], dir[3], up[3];
    float fov_x, fov_y;  /* radians (film axes of the free camera) */
    float yaw, pitch;    /* free-camera orientation state (Camera::rotate) */
    uint32_t ntex;       /* texture records */
    uint32_t ntexel;     /* RGBA texels in the pool */
    uint32_t env_tex;    /* latitude-longitude environment texture (0 = none) */
    float env_w2o[9];    /* its worldToObject rotation part (row-major 3x3) */
    float env_center[3]; /* world-box centroid (scene.cpp:26-27), set by flatten */
} gpr_scene_info;
int gpr_scene_get_info(const gpr_scene* s, gpr_scene_info* out);
/* Flattened per-triangle data in triangle-id order (what the kernels read):
 * pos ntri*9, nrm ntri*9, uv ntri*6, flags ntri (bit0 normals, bit1 uvs),
 * mat ntri; mats nmat.  Any pointer may be NULL. */
int gpr_scene_export_flat(const gpr_scene* s, float* pos, float* nrm, float* uv, uint8_t* flags,
                          uint32_t* mat, gpr_material* mats);

/* Mitsuba-XML subset / .obj / .gprs scene file -> committed scene
 * (load_scene, src/renderer/loaders/loader.h:11-29; XML semantics of
 * ParseRecord, src/renderer/loaders/mitsubaloader.cpp:262-522; OBJ corners as
 * load_obj_file, objloader.cpp:12-121).  The reference's top-level .obj path
 * is dead code (loader.h:17-20) and it has no custom format: here a .obj file
 * becomes one Lambert mesh (reflectance 0.5) seen by a 40-degree 256x256
 * camera on +z framing its bounds, and .gprs is the binary format written by
 * gpr_scene_save. */
int gpr_scene_load(const char* path, gpr_scene** out);
/* gpr_scene_load without the commit: parses and assembles the scene on the
 * host only (no device work), for inspection with gpr_scene_get_info /
 * gpr_scene_export_flat after a commit, or for conversion with gpr_scene_save. */
int gpr_scene_parse(const char* path, gpr_scene** out);
/* Write the meshes, materials, camera, film size and constant environment of
 * a scene to a .gprs file (little-endian: "GPRS", u32 version 1, u32 meshes,
 * u32 width, u32 height, f32 fov, f32 c2w[16], u32 has_env, f32 env[3]; per
 * mesh: gpr_material, u32 np, nn, nuv, ntri, f32 p[3np], n[3nn], uv[2nuv],
 * i32 corners[9ntri]). */
int gpr_scene_save(const gpr_scene* s, const char* path);

/* ---------------------------------------------------------------- BVH --- */
/* Rebuild the BVH of a committed scene on the device (LBVH_build,
 * bvh.h:186-221): centroid bounds, 30-bit Morton keys, radix sort,
 * radix-tree split with the equal-key midpoint rule, bottom-up refit,
 * reference DFS layout and packed 128-B BVH4 nodes. */
int gpr_bvh_build(gpr_scene* s);
/* Morton keys per triangle id, and the sorted (key, triangle id) pairs
 * (bvh.h:199-208; ties in triangle-id order). Any pointer may be NULL. */
int gpr_bvh_export_keys(const gpr_scene* s, uint32_t* keys, uint32_t* sorted_keys,
                        uint32_t* sorted_idx);
/* Reference DFS layout (bvh.h:121-141, 2*ntri-1 nodes). */
int gpr_bvh_export_ref_layout(const gpr_scene* s, gpr_ref_node* out, size_t cap);
/* Device time (ms) of each stage of the last build:
 * [0] bounds+morton, [1] sort, [2] tree+refit, [3] layouts, [4] total. */
int gpr_bvh_last_timing(const gpr_scene* s, float ms[5]);

/* Standalone LDS-staged radix sort of n Morton keys into (key, index) pairs,
 * stable (ties in index order), the sort of LBVH_build (bvh.h:207-208).
 * Device pointers; keys must be below 2^30 (30-bit Morton codes).  Its
 * scratch belongs to the calling thread's (device, stream) and grows on the
 * first call with a larger n: to capture it into a hipGraph, call it once on
 * the capture stream with that n beforehand, outside the capture. */
int gpr_sort_pairs_device(const uint32_t* keys, uint32_t* sorted_keys, uint32_t* sorted_idx,
                          uint32_t n);
/* gpr_sort_pairs_device without allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * hipGraph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as gpr_sort_pairs_device; its scratch is a fixed
 * device array of 18.9 MB per GPU, held until the process ends, so n is at
 * most 1,048,576 (larger n fails with GPR_ERR_INVALID).  Every call on one GPU
 * shares that array: calls that overlap on one GPU (different streams or
 * threads without ordering between them) race and give wrong results. */
int gpr_sort_pairs_device_noalloc(const uint32_t* keys, uint32_t* sorted_keys, uint32_t* sorted_idx,
                                  uint32_t n);

/* -------------------------------------------------------------- trace --- */
/* Traversal launch configuration picked by the autotuners (gpr_trace_tune,
 * gpr_batch_tune, gpr_server_tune). */
typedef struct {
    uint32_t block;     /* threads per block of the traversal kernels: 64, 128, 256, 512 */
    uint32_t stack;     /* LDS stack entries per lane: 8, 16, 32 */
    uint32_t chunk;     /* rays / queue entries claimed per wave per atomic */
    uint32_t layout;    /* 0 BVH4 128-B nodes, 1 reference DFS layout (batched / served only) */
    uint32_t refill;    /* gpr_trace: 1 = a lane takes a new ray as soon as its ray is done */
    float ms;           /* time of the probe with this configuration */
} gpr_trace_cfg;
#define GPR_TRACE_ANY 1u         /* any-hit: stop at the first accepted hit */
#define GPR_TRACE_REF_ORDER 2u   /* reference visit order on the DFS layout */
#define GPR_TRACE_DEVICE_PTRS 4u /* rays / hits are device pointers */
/* trace / trace_shadow (src/renderer/pathtracer_kernel.h:235-246, 310-321).
 * The BVH4 traversal's work buffers (a ray cursor and a stack spill area of
 * 768 B per resident lane) belong to the calling thread's (device, stream):
 * traces on different streams may run concurrently.
 * Host buffers (no GPR_TRACE_DEVICE_PTRS): the call stages the rays in and the
 * hits out itself and returns when the hits are written; batches above 2^20
 * rays move in chunks whose copies overlap the traversal of other chunks.  The
 * device staging (64 B per ray) is kept per (device, stream) until the process
 * ends; host-buffer calls on one stream from several threads take turns. */
int gpr_trace(gpr_scene* s, const gpr_ray* rays, gpr_hit* hits, uint64_t n, uint32_t flags);
/* Autotune the BVH4 traversal of gpr_trace for this scene on a batch of device
 * rays (flags: GPR_TRACE_DEVICE_PTRS, optionally GPR_TRACE_ANY): block size
 * {64,128,256,512} x LDS stack {8,16,32} x chunk {64,128,256} x per-lane ray
 * refill {off, on}; the fastest is
 * kept for the scene's later gpr_trace calls.  Results do not depend on the
 * configuration.  No counterpart in the reference (its traversal launch shape
 * is fixed, pathtracer_kernel.h:235-246). */
int gpr_trace_tune(gpr_scene* s, const gpr_ray* rays, uint64_t n, uint32_t flags, gpr_trace_cfg* out,
                   uint32_t cap, uint32_t* n_out);
int gpr_trace_set_cfg(gpr_scene* s, const gpr_trace_cfg* cfg);
/* Node visits and triangle tests of the last GPR_TRACE_REF_ORDER call. */
int gpr_trace_ref_stats(const gpr_scene* s, uint64_t* node_visits, uint64_t* tri_tests);
/* BVH4 node fetches (128 B each) and triangle-record tests (48 B each) summed
 * over n device rays (flags: GPR_TRACE_DEVICE_PTRS, optionally GPR_TRACE_ANY),
 * for the walk the fast closest-hit kernels take (any hit: with leaves tested
 * as soon as their node is visited).  No hits are written.  No counterpart in
 * the reference (its traversal keeps no counters, pathtracer_kernel.h:175-321). */
int gpr_trace_bvh4_stats(gpr_scene* s, const gpr_ray* rays, uint64_t n, uint32_t flags, uint64_t* node_fetches,
                         uint64_t* tri_tests);
/* Rays whose traversal ran out of stack in the last fast gpr_trace on the
 * calling thread's (device, stream); 0 = every ray completed.  Host-buffer
 * calls report overflow as an error themselves; device-pointer calls return
 * without waiting, so their callers read it here (this call synchronizes the
 * stream).  The reference's fixed int[64] stack has no check at all
 * (pathtracer_kernel.h:183-233). */
int gpr_trace_overflow_count(uint64_t* count);
/* Frees the work buffers that gpr_trace, gpr_sort_pairs_device and
 * host-buffer traces keep for the calling thread's device and hip_stream
 * (NULL = null stream), after synchronizing that stream; the next call on the
 * stream allocates them again.  For applications that create streams per
 * request or per thread; no other call may use the stream meanwhile.  No
 * counterpart in the reference (one stream). */
int gpr_release_stream(void* hip_stream);

/* ----------------------------------------------------------- renderer --- */
#define GPR_RNG_REFERENCE 0u /* TEA seed + LCG, slot-indexed seeds (pathtracer_kernel.h:92, rayqueue.h) */
#define GPR_RNG_PCG 1u       /* PCG32 state carried with the path, unbiased roulette */
#define GPR_TRAVERSAL_FAST 0u
#define GPR_TRAVERSAL_REF 1u
typedef struct {
    uint32_t max_depth; /* PTOptions::m_max_path_length (pathtracer.h:7-14) */
    uint32_t rng_mode;  /* GPR_RNG_* */
    uint32_t rr_start;  /* pcg: roulette when bounce >= rr_start */
    uint64_t seed;      /* pcg seed */
    uint32_t traversal; /* GPR_TRAVERSAL_* */
} gpr_render_opts;
/* PathTracer(filename) + init() (pathtracer.cu:8-39). */
int gpr_renderer_create(gpr_scene* s, const gpr_render_opts* opts, gpr_renderer** out);
/* spp x PathTracer::render() (pathtracer.cu:41-150). */
int gpr_render(gpr_renderer* r, uint32_t spp);
/* Linear radiance sum and per-pixel sample count (FrameBuffer, framebuffer.h:33-70). */
int gpr_renderer_read(gpr_renderer* r, float* rgb_sum, uint32_t* sample_num);
/* Continue a progressive render from a checkpoint: rgb_sum (width*height*3
 * linear radiance sums, as gpr_renderer_read returns them) holding `spp`
 * samples.  The next gpr_render renders sample `spp` onward; in PCG mode the
 * result is bit-identical to an uninterrupted render, since every sample's
 * random numbers depend on the seed, the sample index and the pixel only.
 * Reference RNG mode keeps slot-indexed seeds across samples and is refused.
 * No counterpart in the reference (it writes its buffer once, pathtracer.cu:168);
 * for progressive and served renders that are saved and resumed. */
int gpr_renderer_restore(gpr_renderer* r, const float* rgb_sum, uint32_t spp);
/* FrameBuffer::output (framebuffer.cpp:11-26): mean, sRGB gamma, 8-bit PNG. */
int gpr_renderer_write_png(gpr_renderer* r, const char* path);
/* WriteImage (imageio.cpp:59-82) by extension: ".png" as gpr_renderer_write_png,
 * ".hdr" a Radiance RGBE file of the mean linear radiance, rows flipped like the
 * PNG.  (The reference's ".hdr" branch hands its 8-bit buffer to stbi_write_hdr
 * as floats; this writes the radiance that buffer stands for.) */
int gpr_renderer_write_image(gpr_renderer* r, const char* path);
#define GPR_RESOLVE_DEVICE_PTRS 1u
/* filter / filter_kernel (pathtracer_kernel.h:792-821): per pixel mean -> sRGB
 * -> RGBA8 (alpha 255) into out, width*height uint32 in frame-buffer order (no
 * flip), R in the low byte.  With GPR_RESOLVE_DEVICE_PTRS out is device memory
 * (a display pixel buffer) and the call returns without synchronising; else out
 * is host memory. */
int gpr_renderer_resolve_rgba8(gpr_renderer* r, uint32_t* out, uint32_t flags);
/* PathTracer::reset (pathtracer.cu:195-200). */
int gpr_renderer_reset(gpr_renderer* r);
/* New options (seed, depth, roulette start, RNG mode, traversal) for the next
 * accumulation; resets it.  The reference fixes PTOptions at construction
 * (pathtracer.h:7-14, pathtracer.cu:8-25): this lets one renderer serve
 * requests that differ only in their options. */
int gpr_renderer_set_opts(gpr_renderer* r, const gpr_render_opts* opts);
/* Interactive camera and frame operations of PathTracer (pathtracer.cu:177-209):
 * they change the scene's free camera (Camera::zoom / translate / rotate,
 * camera.h:40-63) or the image size, and the next gpr_render starts a new
 * accumulation, as the reference's m_reset does. */
int gpr_renderer_zoom(gpr_renderer* r, float d);
int gpr_renderer_translate(gpr_renderer* r, float x, float y);
int gpr_renderer_rotate(gpr_renderer* r, float yaw, float pitch);
/* PathTracer::resize + Camera::updateAspectRation (camera.h:65-84): reallocates
 * the frame buffer and queues. */
int gpr_renderer_resize(gpr_renderer* r, uint32_t width, uint32_t height);
int gpr_renderer_get_resolution(const gpr_renderer* r, uint32_t* width, uint32_t* height);
/* Queue sizes of the last sample: in/shadow/scatter per bounce (up to cap
 * bounces); the reference's m_sum_bounce (pathtracer.cu:96). */
int gpr_renderer_queue_sizes(gpr_renderer* r, uint32_t* shadow, uint32_t* scatter, uint32_t cap);
int gpr_renderer_free(gpr_renderer* r);

/* ------------------------------------------------- batched and served --- */
/* Both forms render many jobs (scene, camera, resolution, spp) through one
 * multi-job wavefront: scenes are copied into one device arena, every pass
 * renders the next sample of each job it holds with all their paths in shared
 * ray queues, and the pass is replayed from a hipGraph.  They run in PCG mode
 * with BVH4 traversal, and a job's image is bit-identical to the single-scene
 * renderer's (gpr_renderer_create with GPR_RNG_PCG, GPR_TRAVERSAL_FAST, the
 * same seed / depth / roulette start, the same camera) whatever other jobs it
 * shares passes with.  No counterpart in the reference: these forms are new
 * components built around its path. */
typedef struct {
    uint32_t scene;     /* scene id returned by gpr_batch_add_scene / gpr_server_add_scene */
    uint32_t width, height;
    uint32_t spp;
    uint32_t max_depth; /* 1..64 */
    uint32_t rr_start;  /* roulette from this bounce on (unbiased) */
    uint64_t seed;      /* PCG seed */
    float fov_deg;      /* Mitsuba perspective, as gpr_scene_set_camera */
    float c2w[16];      /* row-major cameraToWorld */
} gpr_job;


typedef struct gpr_batch gpr_batch;
typedef struct {
    uint64_t arena_bytes; /* device bytes for the scene arena (see gpr_scene_arena_bytes) */
    uint32_t capacity;    /* paths per pass (queue entries); >= the largest job's pixel count */
    uint32_t max_jobs;    /* jobs per pass, <= 1024 */
} gpr_batch_opts;
/* Device bytes a committed scene takes in an arena. */
int gpr_scene_arena_bytes(const gpr_scene* s, uint64_t* bytes);
int gpr_batch_create(const gpr_batch_opts* opts, gpr_batch** out);
/* Copy a committed scene into the arena (the scene may be freed afterwards). */
int gpr_batch_add_scene(gpr_batch* b, const gpr_scene* s, uint32_t* id);
int gpr_batch_add_job(gpr_batch* b, const gpr_job* job, uint32_t* id);
/* Autotune the traversal configuration on the batch's own first pass (block
 * size {64,128,256,512} x LDS stack {8,16,32} x chunk {64..512}, plus the
 * reference layout); best first in out[0..n), n <= cap. */
int gpr_batch_tune(gpr_batch* b, gpr_trace_cfg* out, uint32_t cap, uint32_t* n);
int gpr_batch_set_cfg(gpr_batch* b, const gpr_trace_cfg* cfg);
/* Render every job to completion (synchronous). */
int gpr_batch_run(gpr_batch* b);
/* Linear radiance sum of a job (width*height*3 floats) after gpr_batch_run. */
int gpr_batch_read(gpr_batch* b, uint32_t job, float* rgb_sum);
typedef struct {
    uint64_t passes;
    uint64_t paths;     /* camera paths rendered */
    double ms;          /* wall time of the last gpr_batch_run */
} gpr_batch_stats;
int gpr_batch_get_stats(const gpr_batch* b, gpr_batch_stats* out);
int gpr_batch_free(gpr_batch* b);

typedef struct gpr_server gpr_server;
typedef struct {
    uint64_t hbm_budget;  /* device bytes for arena + queues + frame buffers */
    uint64_t arena_bytes; /* part of the budget for scenes */
    uint32_t capacity;    /* paths per pass */
    uint32_t max_jobs;    /* jobs in flight, <= 1024 */
} gpr_server_opts;
/* Starts the server's loop thread.  Frame buffers of admitted requests come
 * out of hbm_budget - arena_bytes - the engine's queues; a request waits until
 * its frame buffer fits and a job slot is free. */
int gpr_server_create(const gpr_server_opts* opts, gpr_server** out);
int gpr_server_add_scene(gpr_server* s, const gpr_scene* scene, uint32_t* id);
/* Queue a request; returns at once with a ticket.  GPR_ERR_BUDGET if the
 * request alone can never fit. */
int gpr_server_submit(gpr_server* s, const gpr_job* job, uint64_t* ticket);
/* Block until the request is done; copy its linear radiance sum
 * (width*height*3 floats, may be NULL) and its latency (submit -> result
 * readable, ms, may be NULL).  Each ticket is waited once. */
int gpr_server_wait(gpr_server* s, uint64_t ticket, float* rgb_sum, double* latency_ms);
/* Cancel a request: a queued one leaves the queue; an admitted one receives no
 * further samples and its frame-buffer range returns to the pool once the passes
 * already launched with it have completed.  gpr_server_wait then returns
 * GPR_ERR_CANCELLED.  A request that has already finished keeps its result.
 * No counterpart in the reference (it has no request loop). */
int gpr_server_cancel(gpr_server* s, uint64_t ticket);
/* Autotune the server's traversal configuration on probe jobs (one pass of
 * their first sample per configuration, as gpr_batch_tune); only while no
 * request is queued or in flight.  Best first in out[0..n), n <= cap. */
int gpr_server_tune(gpr_server* s, const gpr_job* jobs, uint32_t njobs, gpr_trace_cfg* out, uint32_t cap,
                    uint32_t* n);
int gpr_server_set_cfg(gpr_server* s, const gpr_trace_cfg* cfg);
typedef struct {
    uint64_t submitted, completed;
    uint64_t passes;
    double renders_per_s;  /* completed / (last completion - first submit) */
    double p50_ms, p99_ms; /* latency percentiles of the completed requests */
    double mean_jobs_per_pass;
    uint64_t cancelled;    /* requests cancelled before they finished */
} gpr_server_stats;
int gpr_server_get_stats(gpr_server* s, gpr_server_stats* out);
/* Stops the loop after the requests in flight are done, frees everything. */
int gpr_server_free(gpr_server* s);

/* --------------------------------------------------- synthetic inputs --- */
/* Sphere grid (as gpr_scene_sphere_grid) with its subdivision chosen so the
 * spheres take (1 - soup_fraction) of ntri triangles, plus a random soup of
 * small Lambert triangles for the rest (seeded PCG32); committed. */
int gpr_scene_mix(uint64_t seed, uint32_t ntri, float soup_fraction, gpr_scene** out);
/* Seeded generators of the benchmark workloads (PCG32 streams). */
int gpr_gen_soup(uint64_t seed, uint64_t stream, uint32_t n, float sigma, float* pos);
int gpr_gen_rays(uint64_t seed, uint64_t stream, uint64_t n, gpr_ray* rays);
/* Shadow rays from rays[i].o to a uniform in-cube target, t in [0, 0.9999]. */
int gpr_gen_shadow_rays(uint64_t seed, uint64_t stream, const gpr_ray* rays, uint64_t n,
                        gpr_ray* out);
/* Procedural Cornell box (36 triangles, Lambert, one 0.5x0.5 light of radiance 17). */
int gpr_scene_cornell(uint32_t width, uint32_t height, gpr_scene** out);
/* 8x8 UV spheres (64x32 segments) on a ground plane, two quad lights, mixed
 * GGX conductor / dielectric / Lambert materials. */
int gpr_scene_sphere_grid(uint32_t width, uint32_t height, uint32_t subdiv, gpr_scene** out);

#ifdef __cplusplus
}
#endif
#endif
