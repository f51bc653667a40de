/*
 * oracle.h -- C ABI of the CPU restatement of the reference wavefront
 * path-tracing core (slongle/GPU-Renderer).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load liboracle.so, and only as the checker (or as the
 * timed CPU baseline), never as the thing measured on the GPU.
 *
 * Every function cites the reference file:line it restates.  Float op order
 * follows the reference source left to right; the library is compiled with
 * -ffp-contract=off so no FMA is formed.
 */
#ifndef GPR_ORACLE_H
#define GPR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference-layout BVH node: BVHLinearNode (src/renderer/bvh.h:34-43).
 * 24 B AABB + union{tri_idx, right_child_idx} + u8 leaf, padded to 32 B. */
typedef struct {
    float bmin[3];
    float bmax[3];
    uint32_t idx;      /* leaf: triangle index; interior: right child index */
    uint32_t leaf;     /* low byte = m_leaf, upper bytes zero */
} orc_node;

/* Ray / Hit (src/renderer/ray.h:5-19). */
typedef struct { float o[3]; float d[3]; float tmin; float tmax; } orc_ray;
typedef struct { float t; int32_t id; float u; float v; } orc_hit;

/* Material (src/renderer/material.h:6-35): constant fields, and per
 * parameter a texture handle (1-based, 0 = the constant) in the slot order
 * color, color1, ior, alpha_x, alpha_y, eta_i, eta_t, k. */
typedef struct {
    uint32_t type;        /* 0 diffuse, 1 specular, 2 rough conductor, 3 rough dielectric, 4 rough plastic */
    float color[3];
    float color1[3];
    float emission[3];
    float ior;
    float alpha_x, alpha_y;
    float eta_i[3], eta_t[3], k[3];
    uint32_t tex[8];
} orc_material;

/* TextureView (src/renderer/texture.h:14-35): texels [offset, offset+w*h) of
 * an RGBA float pool, row-major, row 0 at v = 0. */
typedef struct {
    uint32_t offset, width, height, pad;
    float uv_offset[2], uv_scale[2];
} orc_texture;

/* Flat triangle scene: the global triangle order of Scene::Scene
 * (src/renderer/scene.cpp:4-14): mesh load order, then face order. */
typedef struct {
    uint32_t ntri;
    const float* pos;       /* ntri*9: p0,p1,p2 */
    const float* nrm;       /* ntri*9 corner normals (read only where flags&1) */
    const float* uv;        /* ntri*6 corner uvs (read only where flags&2) */
    const uint8_t* flags;   /* ntri: bit0 has normals, bit1 has uvs */
    const uint32_t* mat;    /* ntri material ids */
    const orc_material* mats;
    uint32_t nmat;
    /* camera (Mitsuba perspective, camera.h:21-37) */
    float c2w[16];          /* row-major cameraToWorld matrix m[4][4] */
    float c2w_inv[16];      /* its inverse (as built by LookAt / Transform) */
    float fov_deg;
    uint32_t width, height;
    uint32_t has_env;       /* constant environment light */
    float env[3];
    /* free camera (camera.h:89-107): eye, direction, up, film half-angles (radians) */
    uint32_t cam_free;
    float eye[3], dir[3], up[3];
    float fov_x, fov_y;
    /* textures and the latitude-longitude environment (environment.h:10-35) */
    const orc_texture* tex;
    const float* texels;    /* RGBA */
    uint32_t ntex;
    uint32_t env_tex;       /* 1-based texture handle of the lat-long map, 0 = constant env[] */
    float env_w2o[9];       /* worldToObject, upper-left 3x3, row-major */
    float env_center[3];    /* world-box centroid (scene.cpp:26-27) */
} orc_scene;

typedef struct {
    uint32_t spp;
    uint32_t max_depth;     /* PTOptions::m_max_path_length: bounce iterations */
    uint32_t rng_mode;      /* 0 = reference (TEA+LCG, slot seeds), 1 = pcg (state per path) */
    uint32_t rr_start;      /* pcg mode: unbiased roulette when bounce >= rr_start */
    uint64_t seed;          /* pcg mode seed */
    uint32_t first_sample;  /* sample index of the first sample rendered */
} orc_render_opts;

/* --- math helpers (restated so GPU and CPU can be compared bit for bit) --- */
float orc_sinf(float x);
float orc_cosf(float x);

/* --- Morton / LBVH (bvh.h:12-32, 186-221) --- */
uint32_t orc_left_shift3(uint32_t x);
uint32_t orc_encode_morton3(float x, float y, float z);
void orc_centroid_bounds(const float* pos, uint32_t ntri, float bmin[3], float bmax[3]);
void orc_morton_keys(const float* pos, uint32_t ntri, uint32_t* keys);
/* (key, index) pairs sorted by key, equal keys in index order.  Parity
 * unpinned for ties: the reference's std::sort (bvh.h:207-208) is not stable
 * and its order among equal keys is the standard library's own; the stable
 * order here is a definition, checked only as sorted + permutation. */
void orc_sort_pairs(const uint32_t* keys, uint32_t n, uint32_t* skeys, uint32_t* sidx);
/* Full reference build; returns number of nodes written (2n-1) or 0 for n==0. */
uint32_t orc_lbvh_build(const float* pos, uint32_t ntri, orc_node* nodes,
                        uint32_t* skeys, uint32_t* sidx);
/* Tree topology from sorted keys only: writes, for every node in DFS order,
 * the [lo, hi) range of sorted positions it covers (2n-1 entries). */
void orc_lbvh_ranges(const uint32_t* skeys, uint32_t n, uint32_t* lo, uint32_t* hi);

/* --- traversal (pathtracer_kernel.h:119-321) --- */
int orc_intersect_ref(const float* p0, const float* p1, const float* p2,
                      const orc_ray* ray, float* t, float* u, float* v);
void orc_trace(const orc_node* nodes, const float* pos, uint32_t ntri,
               const orc_ray* rays, uint64_t n, int any, orc_hit* hits,
               uint64_t* node_visits, uint64_t* tri_tests, int nthreads);
void orc_brute(const float* pos, uint32_t ntri, const orc_ray* rays, uint64_t n,
               int any, int fp64, orc_hit* hits, double* t64, int nthreads);

/* --- RNG (sampling.h:31-54) and PCG32 --- */
void orc_tea_keyed(uint32_t* v0, uint32_t* v1, const uint32_t key[4], uint32_t rounds);
uint32_t orc_init_random(uint32_t v0, uint32_t v1, uint32_t backoff);
float orc_next_random(uint32_t* seed);
void orc_pcg32_seed(uint64_t* state, uint64_t* inc, uint64_t initstate, uint64_t initseq);
uint32_t orc_pcg32_next(uint64_t* state, uint64_t inc);

/* --- transforms / camera (transform.cpp, camera.h) --- */
void orc_lookat(const float target[3], const float origin[3], const float up[3],
                float m[16], float minv[16]);
void orc_matrix_inverse(const float m[16], float out[16]);
void orc_camera_ray(const orc_scene* s, float px, float py, orc_ray* ray);

/* --- BSDF (bsdf.h and the bsdfs directory); frame-free entry points for tests --- */
/* eval f*|cos|, pdf, sample in the shading frame of normal n; u2 is the lobe-choice
 * uniform of the rough dielectric (samples[4] in the integrator) */
/* TextureView::evalSpectrum (texture.h:37-44) of texture `handle` at uv. */
void orc_tex_eval(const orc_scene* s, uint32_t handle, float u, float v, float out[3]);
/* EnvironmentLightView::Le (environment.h:16-28) for a ray o + t d. */
void orc_env_le(const orc_scene* s, const float o[3], const float d[3], float out[3]);
void orc_bsdf_eval(const orc_material* m, const float n[3], const float wo[3],
                   const float wi[3], float f[3], float* pdf);
int  orc_bsdf_sample(const orc_material* m, const float n[3], const float wo[3],
                     float u0, float u1, float u2, float wi[3], float f[3], float* pdf, int* specular);
float orc_fr_dielectric(float cos_i, float eta_i, float eta_t);
void orc_fr_conductor(float cos_i, const float eta_i[3], const float eta_t[3],
                      const float k[3], float out[3]);
float orc_ggx_d(const float wh[3], float ax, float ay);
float orc_ggx_g(const float wo[3], const float wi[3], float ax, float ay);
/* Smith Lambda of the Trowbridge-Reitz distribution (microfacet.h:7-19). */
float orc_ggx_lambda(const float w[3], float ax, float ay);
/* D at n half vectors (wh: 3n floats). */
void orc_ggx_d_n(const float* wh, uint32_t n, float ax, float ay, float* out);
/* Sample_wh (microfacet.h:41-130, pbrt-v3's visible-normal sampler) for n
 * uniform pairs u[2i], u[2i+1]; wh receives 3n floats. */
void orc_ggx_sample_wh(const float wo[3], const float* u, uint32_t n, float ax, float ay, float* wh);

/* --- integrator (pathtracer.cu:41-150, pathtracer_kernel.h:59-790) --- */
int orc_render(const orc_scene* scene, const orc_node* nodes,
               const orc_render_opts* opts, float* rgb_sum, uint32_t* sample_num,
               int nthreads);
/* orc_render in pcg mode that also sets tie_pixels[p] = 1 when a closest hit on
 * one of pixel p's paths had another triangle within 1e-6 relative of its t
 * (a hit whose triangle depends on the traversal's visit order). */
int orc_render_ties(const orc_scene* scene, const orc_node* nodes,
                    const orc_render_opts* opts, float* rgb_sum, uint32_t* sample_num,
                    uint8_t* tie_pixels, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
