// render.hip -- wavefront path tracer on gfx950 (PathTracer::render,
// src/renderer/pathtracer.cu:41-150, kernels of src/renderer/pathtracer_kernel.h).
//
// One sample of every pixel is a fixed launch sequence with no host round
// trip:  gen -> extend(primary) -> max_depth x { shade -> shadow -> extend }
// -> finish.  Queue sizes live on the device; each kernel reads the size it
// consumes and early-exits, so the whole sample can be replayed as a hipGraph.
//
//   gen     primary rays (InitRandom/NextRandom jitter or PCG32), queue A
//   extend  closest hit of a queue; for scattered rays also the BSDF-sampling
//           MIS contribution (accumulate_BSDF_sampling_contribution) and, in
//           reference RNG mode, its roulette (pathtracer_kernel.h:684-749)
//   shade   shade_hit_with_MIS: emission, NEE shadow append, BSDF scatter
//           append.  Appends are compacted in queue order: a wave ballot and
//           mbcnt give lane offsets, LDS gives wave offsets, and a decoupled
//           look-back over tiles (one 64-bit status word per tile, no other
//           cross-XCD data) gives tile offsets -- so slot order equals the
//           serial order of the reference's host loop.
//   shadow  any-hit of the shadow queue + accumulate_light_sampling_contribution
//
// Per pixel the frame buffer receives emission, NEE and BSDF terms in the
// same order as the reference kernels (one queue entry per pixel per kernel,
// so the read-modify-write needs no atomics).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <memory>

#include "dev_bsdf.h"
#include "dev_texture.h"
#include "engine.h"
#include "traverse.h"

namespace gpr {
namespace {

constexpr int kShadeBlock = 256;
constexpr uint32_t kMaxBounce = 64;

// Queue of path vertices / rays (SoA, capacity = paths per sample or pass).
struct QueueView {
    float4* o_tmin;   // origin, tmin
    float4* d_tmax;   // direction, tmax
    float4* hit;      // t, id bits, u, v
    float4* w_pdf;    // throughput rgb, BSDF pdf
    uint32_t* pix;    // frame-buffer pixel | specular << 31
    unsigned long long* rng;  // PCG32 state (pcg mode)
    uint32_t* seed;   // slot-indexed LCG seed (reference mode, persists across samples)
    uint32_t* job;    // multi-job passes: the entry's scene << kJobBits | pass job
};

// Per-sample control block (counters cleared by k_gen).
struct Ctrl {
    uint32_t sample;         // current sample index (single-scene; advanced by k_finish)
    uint32_t error;          // bit0: traversal stack overflow, bit1: look-back timeout
    uint32_t shadow_size[kMaxBounce];
    uint32_t scatter_size[kMaxBounce];
    uint32_t tile_counter[kMaxBounce];
    uint32_t ext_cursor[kMaxBounce + 1];  // [0] primary, [b+1] scatter queue of bounce b
    uint32_t sh_cursor[kMaxBounce];
};

// Jobs of one multi-job pass (device copy of EngineJob[] plus the path offsets).
struct PassJob {
    EngineJob j;
    uint32_t path_off;  // first path of the job in the pass
    uint32_t pad[3];
};
struct PassBuf {
    uint32_t npaths;
    uint32_t njobs;
    uint32_t pad[62];
    PassJob jobs[kMaxPassJobs];
};

struct RenderArgs {
    SceneView sv;              // single-scene renders
    const SceneView* scenes;   // multi-job passes: scene table
    const PassBuf* pass;       // multi-job passes: jobs of this pass
    QueueView q[3];            // A, B (in/scatter alternate), S (shadow)
    float4* fb;
    Ctrl* ctrl;
    unsigned long long* status;  // [kMaxBounce][ntiles] look-back words
    int2* spill;
    uint32_t npix;             // single-scene: paths per sample
    uint32_t ntiles;           // look-back words per bounce (capacity / kShadeBlock)
    uint32_t depth;            // bounces of the sample / pass
    uint32_t chunk;            // entries a wave claims per atomic in extend / shadow
    uint32_t rng_mode;
    uint32_t rr_start;
    unsigned long long seed;
};

template <bool MULTI>
GPR_D uint32_t npaths_of(const RenderArgs& a) {
    if constexpr (MULTI) return a.pass->npaths;
    else return a.npix;
}
// Multi-job queue entries carry scene << kJobBits | pass job, so the kernels
// that only need the scene (traversal) index the scene table directly.
constexpr uint32_t kJobBits = 10;
static_assert(kMaxPassJobs <= (1u << kJobBits), "pass job index must fit its bits");
template <bool MULTI>
GPR_D SceneView scene_of(const RenderArgs& a, const QueueView& q, uint32_t i) {
    if constexpr (MULTI) return ldg(a.scenes + (q.job[i] >> kJobBits));
    else return a.sv;
}

GPR_D void fb_add(float4* fb, uint32_t pix, c3 v) {
    float4 x = fb[pix];
    x.x = x.x + v.r;
    x.y = x.y + v.g;
    x.z = x.z + v.b;
    fb[pix] = x;
}

// Camera::generateRay (camera.h:86-119), Mitsuba raster mode.
GPR_D void camera_ray(const float* m, const float* c, float px, float py, v3& o, v3& d) {
    float x = m[0] * px + m[1] * py + m[2] * 0.f + m[3];
    float y = m[4] * px + m[5] * py + m[6] * 0.f + m[7];
    float z = m[8] * px + m[9] * py + m[10] * 0.f + m[11];
    const float w = m[12] * px + m[13] * py + m[14] * 0.f + m[15];
    v3 pc = V(x, y, z);
    if (!(w == 1)) pc = pc / w;
    x = c[0] * 0.f + c[1] * 0.f + c[2] * 0.f + c[3];
    y = c[4] * 0.f + c[5] * 0.f + c[6] * 0.f + c[7];
    z = c[8] * 0.f + c[9] * 0.f + c[10] * 0.f + c[11];
    const float ow = c[12] * 0.f + c[13] * 0.f + c[14] * 0.f + c[15];
    o = V(x, y, z);
    if (!(ow == 1)) o = o / ow;
    const v3 n = norm3(pc);
    d = norm3(V(c[0] * n.x + c[1] * n.y + c[2] * n.z, c[4] * n.x + c[5] * n.y + c[6] * n.z,
                c[8] * n.x + c[9] * n.y + c[10] * n.z));
}

// Camera::generateRay (camera.h:89-107), free mode: film frame precomputed on the host.
GPR_D void camera_ray_free(const SceneView& sv, float px, float py, v3& o, v3& d) {
    const float sx = 2.f * (px / (float)sv.width) - 1.f;
    const float sy = 2.f * (py / (float)sv.height) - 1.f;
    const v3 p = load3(sv.cam_mid) + load3(sv.cam_h) * sx + load3(sv.cam_v) * sy;
    o = load3(sv.cam_eye);
    d = norm3(p - o);
}

template <bool MULTI>
__global__ __launch_bounds__(256) void k_gen(RenderArgs a) {
    const uint32_t tid = blockIdx.x * 256 + threadIdx.x;
    const uint32_t gsz = gridDim.x * 256;
    const uint32_t npix = npaths_of<MULTI>(a);
    // clear the per-sample counters, cursors and the look-back words this sample uses
    Ctrl* c = a.ctrl;
    for (uint32_t i = tid; i < kMaxBounce; i += gsz) {
        c->shadow_size[i] = 0u;
        c->scatter_size[i] = 0u;
        c->tile_counter[i] = 0u;
        c->ext_cursor[i + 1] = 0u;
        c->sh_cursor[i] = 0u;
    }
    if (tid == 0) c->ext_cursor[0] = 0u;
    const uint32_t used_tiles = (npix + kShadeBlock - 1) / kShadeBlock;
    const uint64_t nstat = (uint64_t)a.depth * used_tiles;
    for (uint64_t i = tid; i < nstat; i += gsz) a.status[(i / used_tiles) * a.ntiles + i % used_tiles] = 0ull;
    if (tid >= npix) return;
    const QueueView& q = a.q[0];
    if constexpr (MULTI) {
        // the job holding path tid: last job whose path_off <= tid
        const PassBuf* pb = a.pass;
        uint32_t lo = 0, hi = pb->njobs;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (pb->jobs[mid].path_off <= tid) lo = mid;
            else hi = mid;
        }
        const EngineJob& jb = pb->jobs[lo].j;
        const uint32_t lp = tid - pb->jobs[lo].path_off;
        const uint32_t fbi = jb.fb_off + lp;
        if (jb.sample == 0) a.fb[fbi] = make_float4(0.f, 0.f, 0.f, 0.f);
        const uint64_t inc = pcg_inc(jb.sample, lp);
        uint64_t st = pcg_init(jb.seed, inc);
        const float jx = pcg_unit(st, inc);
        const float jy = pcg_unit(st, inc);
        v3 o, d;
        camera_ray(jb.r2c, jb.c2w, (float)(lp % jb.width) + jx, (float)(lp / jb.width) + jy, o, d);
        q.rng[tid] = st;
        q.o_tmin[tid] = make_float4(o.x, o.y, o.z, 0.f);
        q.d_tmax[tid] = make_float4(d.x, d.y, d.z, 1e34f);
        q.w_pdf[tid] = make_float4(1.f, 1.f, 1.f, 0.f);
        q.pix[tid] = fbi;
        q.job[tid] = (jb.scene << kJobBits) | lo;
    } else {
        const uint32_t sample = c->sample;
        const uint32_t W = a.sv.width;
        const uint32_t x = tid % W, y = tid / W;
        float jx, jy;
        if (a.rng_mode == GPR_RNG_REFERENCE) {
            uint32_t s = tea16(tid, sample);  // InitRandom(idx, sample) (pathtracer_kernel.h:92)
            jx = lcg_next(s);
            jy = lcg_next(s);
            q.seed[tid] = s;
        } else {
            const uint64_t inc = pcg_inc(sample, tid);
            uint64_t st = pcg_init(a.seed, inc);
            jx = pcg_unit(st, inc);
            jy = pcg_unit(st, inc);
            q.rng[tid] = st;
        }
        v3 o, d;
        if (a.sv.cam_free) camera_ray_free(a.sv, (float)x + jx, (float)y + jy, o, d);
        else camera_ray(a.sv.r2c, a.sv.c2w, (float)x + jx, (float)y + jy, o, d);
        q.o_tmin[tid] = make_float4(o.x, o.y, o.z, 0.f);
        q.d_tmax[tid] = make_float4(d.x, d.y, d.z, 1e34f);
        q.w_pdf[tid] = make_float4(1.f, 1.f, 1.f, 0.f);
        q.pix[tid] = tid;
    }
}

template <class Stack>
GPR_D gpr_hit trace_one(const RenderArgs& a, const SceneView& sv, const Ray3& r, Stack& st, bool any,
                        bool ref_order) {
    gpr_hit h;
    if (ref_order) {
        uint32_t nv = 0, nt = 0;
        h = any ? ref_trace<true>(sv, r, &nv, &nt) : ref_trace<false>(sv, r, &nv, &nt);
    } else {
        const bool ok = any ? fast_trace<true>(sv, r, st, h) : fast_trace<false>(sv, r, st, h);
        if (!ok) atomicOr(&a.ctrl->error, 1u);
    }
    return h;
}

// One triangle's shading record, read as five 16-byte loads from its line
// (its corner uvs, needed only for textured hits, are read by resolve_textures).
struct Shade {
    v3 p0, p1, p2, n0, n1, n2;
    uint32_t flags, mat;
};
GPR_D Shade load_shade(const SceneView& sv, uint32_t t) {
    const float4* r = reinterpret_cast<const float4*>(sv.shade + t);
    const float4 a = ldg(r), b = ldg(r + 1), c = ldg(r + 2), d = ldg(r + 3), e = ldg(r + 4);
    Shade s;
    s.p0 = V(a.x, a.y, a.z);
    s.p1 = V(a.w, b.x, b.y);
    s.p2 = V(b.z, b.w, c.x);
    s.n0 = V(c.y, c.z, c.w);
    s.n1 = V(d.x, d.y, d.z);
    s.n2 = V(d.w, e.x, e.y);
    s.flags = __float_as_uint(e.z);
    s.mat = __float_as_uint(e.w);
    return s;
}
// Light pdf (Triangle::pdf + pdf_lights, triangle.h:156-175, pathtracer_kernel.h:338-348)
GPR_D float tri_area(const Shade& s) { return len3(cross3(s.p1 - s.p0, s.p2 - s.p0)) * 0.5f; }
struct Geom {
    v3 p, ns;
};
// Triangle::setupDifferential (triangle.h:104-138)
GPR_D Geom setup_geom(const Shade& s, float u, float v) {
    Geom g;
    g.p = s.p0 * (1 - u - v) + s.p1 * u + s.p2 * v;
    if (s.flags & 1u) g.ns = norm3(s.n0 * (1 - u - v) + s.n1 * u + s.n2 * v);
    else g.ns = norm3(cross3(s.p0 - s.p2, s.p1 - s.p2));
    return g;
}
GPR_D c3 emission_of(const SceneView& sv, const Shade& s) {
    const float* e = sv.mats[s.mat].emission;
    return C(ldg(e), ldg(e + 1), ldg(e + 2));
}
GPR_D float power_heur(float a, float b) {
    a *= a;
    b *= b;
    return a / (a + b);
}

// Claim `chunk` consecutive queue entries per wave from a cursor; calls
// f(i) for every claimed entry below n.
template <class F>
GPR_D void for_each_claimed(uint32_t* cursor, uint32_t n, uint32_t chunk, F&& f) {
    const int lane = threadIdx.x & 63;
    while (true) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(cursor, chunk);
        base = __shfl(base, 0, 64);
        if (base >= n) break;
        for (uint32_t k = 0; k < chunk; k += 64) {
            const uint32_t i = base + k + lane;
            if (i < n) f(i);
        }
    }
}

// Per-lane ray refill over queue entries: a lane whose ray is done takes the
// next entry from a wave-local pool of `chunk` entries (one atomic per chunk).
// begin(i) sets entry i up and returns true if it finished at once; step()
// advances the lane's ray one node and returns true when it is done.
template <class Begin, class Step>
GPR_D void refill_loop(uint32_t* cursor, uint32_t n, uint32_t chunk, Begin&& begin, Step&& step) {
    const int lane = threadIdx.x & 63;
    uint32_t pb = 0, pe = 0;  // wave-local pool of entries (wave-uniform)
    bool drained = false;
    bool active = false;
    while (true) {
        const unsigned long long need = __ballot(!active);
        if (need != 0ull && !(drained && pb == pe)) {
            const uint32_t cnt = (uint32_t)__popcll(need);
            const uint32_t rank =
                __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
            const uint32_t avail = pe - pb;
            uint32_t nb = pe, ne = pe;
            if (avail < cnt && !drained) {
                uint32_t b = 0;
                if (lane == 0) b = atomicAdd(cursor, chunk);
                b = __shfl(b, 0, 64);
                nb = b < n ? b : n;
                ne = (uint64_t)b + chunk < n ? b + chunk : n;
                if ((uint64_t)b + chunk >= n) drained = true;
            }
            if (!active) {
                const bool from_old = rank < avail;
                const uint32_t idx = from_old ? pb + rank : nb + (rank - avail);
                if (from_old || idx < ne) active = !begin(idx);
            }
            if (avail >= cnt) {
                pb += cnt;
            } else {
                const uint32_t taken = nb + (cnt - avail);
                pb = taken < ne ? taken : ne;
                pe = ne;
            }
        }
        if (__ballot(active) == 0ull) {
            if (drained && pb == pe) break;
            continue;
        }
        if (active) active = !step();
    }
}

GPR_D Ray3 queue_ray(const QueueView& q, uint32_t i) {
    const float4 ot = q.o_tmin[i], dt = q.d_tmax[i];
    Ray3 r;
    r.o = V(ot.x, ot.y, ot.z);
    r.d = V(dt.x, dt.y, dt.z);
    r.tmin = ot.w;
    r.tmax = dt.w;
    return r;
}

// After the closest hit of entry i: for scattered rays the BSDF-sampling MIS
// term and, in reference RNG mode, its roulette (pathtracer_kernel.h:684-749).
template <bool PRIMARY, bool MULTI>
GPR_D void extend_finish(const RenderArgs& a, const QueueView& q, uint32_t i, const SceneView& sv, const Ray3& r,
                         gpr_hit h, uint32_t bounce) {
    if (!PRIMARY) {
        const float4 wp = q.w_pdf[i];
        const uint32_t pw = q.pix[i];
        const uint32_t pix = pw & 0x7fffffffu;
        const bool spec = (pw >> 31) != 0;
        const c3 w = C(wp.x, wp.y, wp.z);
        float u_rr = 0.f;
        if (!MULTI && a.rng_mode == GPR_RNG_REFERENCE) u_rr = lcg_next(q.seed[i]);
        if (h.id != -1 && !spec) {
            const Shade ts = load_shade(sv, (uint32_t)h.id);
            const c3 le = emission_of(sv, ts);
            if (!black(le)) {
                const Geom lg = setup_geom(ts, h.u, h.v);
                float lpdf = 1.f / tri_area(ts);
                lpdf /= sv.nlights;
                const float d = h.t;
                lpdf *= d * d / fabsf(dot3(r.d, lg.ns));
                const c3 fl = dot3(-r.d, lg.ns) > 0 ? le : C(0.f);
                const float wgt = power_heur(wp.w, lpdf);
                fb_add(a.fb, pix, w * fl * wgt);
            }
        }
        if (!MULTI && a.rng_mode == GPR_RNG_REFERENCE && cmax(w) < 1 && bounce > 3) {
            // roulette without reweighting; a terminated path is a miss (pathtracer_kernel.h:738-748)
            const float qq = fmax_h(0.05f, 1 - cmax(w));
            if (u_rr < qq) h.id = -1;
        }
    }
    q.hit[i] = make_float4(h.t, __int_as_float(h.id), h.u, h.v);
}

// After the any-hit test of shadow entry i (pathtracer_kernel.h:642-664).
GPR_D void shadow_finish(const RenderArgs& a, const QueueView& q, uint32_t i, const gpr_hit& h) {
    if (h.id == -1) {
        const float4 wp = q.w_pdf[i];
        fb_add(a.fb, q.pix[i] & 0x7fffffffu, C(wp.x, wp.y, wp.z));
    }
}

// extend: closest hit; for scattered rays also the BSDF-sampling MIS term.
// REFILL: per-lane ray refill (BVH4 traversal only).
template <bool PRIMARY, bool REF_ORDER, bool MULTI, int NT, int SL, bool REFILL = false>
__global__ __launch_bounds__(NT) void k_extend(RenderArgs a, uint32_t bounce) {
    __shared__ int2 lds[SL * NT];
    LaneStackT<NT, SL> st;
    st.lds = lds + threadIdx.x;
    st.spill = a.spill + (uint64_t)(blockIdx.x * NT + threadIdx.x) * kStackGlobal;
    st.sp = 0;
    // queue: primary = A; at bounce b the scatter queue is q[(b+1)&1]
    const QueueView& q = a.q[PRIMARY ? 0 : ((bounce + 1) & 1)];
    const uint32_t n = PRIMARY ? npaths_of<MULTI>(a) : a.ctrl->scatter_size[bounce];
    uint32_t* cur = &a.ctrl->ext_cursor[PRIMARY ? 0 : bounce + 1];
    if constexpr (REFILL && !REF_ORDER) {
        uint32_t ei = 0;
        SceneView sv;
        Trav t;
        refill_loop(
            cur, n, a.chunk,
            [&](uint32_t i) {
                ei = i;
                sv = scene_of<MULTI>(a, q, i);
                if (!trav_begin<false>(sv, queue_ray(q, i), st, t)) return false;
                extend_finish<PRIMARY, MULTI>(a, q, i, sv, t.r, t.hit, bounce);
                return true;
            },
            [&]() {
                const int rc = trav_step<false>(sv, st, t);
                if (rc == kTravMore) return false;
                if (rc == kTravOverflow) atomicOr(&a.ctrl->error, 1u);
                extend_finish<PRIMARY, MULTI>(a, q, ei, sv, t.r, t.hit, bounce);
                return true;
            });
    } else {
        for_each_claimed(cur, n, a.chunk, [&](uint32_t i) {
            const SceneView sv = scene_of<MULTI>(a, q, i);
            const Ray3 r = queue_ray(q, i);
            const gpr_hit h = trace_one(a, sv, r, st, false, REF_ORDER);
            extend_finish<PRIMARY, MULTI>(a, q, i, sv, r, h, bounce);
        });
    }
}

// shadow: any hit + accumulate_light_sampling_contribution (pathtracer_kernel.h:642-664)
template <bool REF_ORDER, bool MULTI, int NT, int SL, bool REFILL = false>
__global__ __launch_bounds__(NT) void k_shadow(RenderArgs a, uint32_t bounce) {
    __shared__ int2 lds[SL * NT];
    LaneStackT<NT, SL> st;
    st.lds = lds + threadIdx.x;
    st.spill = a.spill + (uint64_t)(blockIdx.x * NT + threadIdx.x) * kStackGlobal;
    st.sp = 0;
    const QueueView& q = a.q[2];
    const uint32_t n = a.ctrl->shadow_size[bounce];
    uint32_t* cur = &a.ctrl->sh_cursor[bounce];
    if constexpr (REFILL && !REF_ORDER) {
        uint32_t ei = 0;
        SceneView sv;
        Trav t;
        refill_loop(
            cur, n, a.chunk,
            [&](uint32_t i) {
                ei = i;
                sv = scene_of<MULTI>(a, q, i);
                if (!trav_begin<true>(sv, queue_ray(q, i), st, t)) return false;
                shadow_finish(a, q, i, t.hit);
                return true;
            },
            [&]() {
                const int rc = trav_step<true>(sv, st, t);
                if (rc == kTravMore) return false;
                if (rc == kTravOverflow) atomicOr(&a.ctrl->error, 1u);
                shadow_finish(a, q, ei, t.hit);
                return true;
            });
    } else {
        for_each_claimed(cur, n, a.chunk, [&](uint32_t i) {
            const SceneView sv = scene_of<MULTI>(a, q, i);
            const gpr_hit h = trace_one(a, sv, queue_ray(q, i), st, true, REF_ORDER);
            shadow_finish(a, q, i, h);
        });
    }
}

// ------------------------------------------------------------- shade ----
struct Out {
    v3 o, d;
    float tmin, tmax;
    c3 w;
    float pdf;
    bool spec;
};

// 64-bit look-back word: [63:62] flag (1 aggregate, 2 inclusive) [61:31] shadow [30:0] scatter
GPR_D unsigned long long pack_status(uint32_t flag, uint32_t sh, uint32_t sc) {
    return ((unsigned long long)flag << 62) | ((unsigned long long)sh << 31) | (unsigned long long)sc;
}

template <bool MULTI>
__global__ __launch_bounds__(kShadeBlock) void k_shade(RenderArgs a, uint32_t bounce) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave[2][kShadeBlock / 64];
    __shared__ uint32_t s_base[2];
    const uint32_t n_in = bounce == 0 ? npaths_of<MULTI>(a) : a.ctrl->scatter_size[bounce - 1];
    if (threadIdx.x == 0) s_tile = atomicAdd(&a.ctrl->tile_counter[bounce], 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t ntiles_in = (n_in + kShadeBlock - 1) / kShadeBlock;
    if (tile >= ntiles_in) return;
    const uint32_t i = tile * kShadeBlock + threadIdx.x;
    const QueueView& qi = a.q[bounce & 1];
    const QueueView& qs = a.q[(bounce + 1) & 1];
    const QueueView& qsh = a.q[2];
    bool valid = i < n_in;
    uint32_t job = 0, rr_start = a.rr_start, sample = 0, lpix = 0;
    SceneView sv;
    if constexpr (MULTI) {
        if (valid) {
            job = qi.job[i];  // scene << kJobBits | pass job, copied to the appended entries
            const PassJob pj = ldg(a.pass->jobs + (job & ((1u << kJobBits) - 1)));
            sv = ldg(a.scenes + (job >> kJobBits));
            // a job's paths stop at its own depth
            valid = bounce < pj.j.max_depth;
            rr_start = pj.j.rr_start;
            sample = pj.j.sample;
            lpix = (qi.pix[i] & 0x7fffffffu) - pj.j.fb_off;
        }
    } else {
        sv = a.sv;
        sample = a.ctrl->sample;
    }
    bool has_sh = false, has_sc = false;
    Out osh, osc;
    uint32_t pix = 0;
    unsigned long long rng_out = 0;
    if (valid) {
        const float4 dt = qi.d_tmax[i], hh = qi.hit[i], wp = qi.w_pdf[i];
        const uint32_t pw = qi.pix[i];
        pix = pw & 0x7fffffffu;
        if constexpr (!MULTI) lpix = pix;
        const bool spec = (pw >> 31) != 0;
        const v3 rd = V(dt.x, dt.y, dt.z);
        const c3 w = C(wp.x, wp.y, wp.z);
        const int32_t hid = __float_as_int(hh.y);
        float smp[8];
        if (!MULTI && a.rng_mode == GPR_RNG_REFERENCE) {
            uint32_t s = qi.seed[i];
#pragma unroll
            for (int k = 0; k < 8; ++k) smp[k] = lcg_next(s);
            qi.seed[i] = s;
        } else {
            const uint64_t inc = pcg_inc(sample, lpix);
            uint64_t s = qi.rng[i];
#pragma unroll
            for (int k = 0; k < 8; ++k) smp[k] = pcg_unit(s, inc);
            rng_out = s;
        }
        if (hid != -1) {
            const uint32_t t = (uint32_t)hid;
            const Shade ts = load_shade(sv, t);
            const Geom g = setup_geom(ts, hh.z, hh.w);
            const v3 wo = -norm3(rd);
            gpr_material m = ldg(sv.mats + ts.mat);
            if (sv.ntex != 0 && has_textures(m)) resolve_textures(sv, m, t, ts.flags, hh.z, hh.w);
            const Frame fr = make_frame(g.ns);
            if (bounce == 0 || spec) {
                const c3 le = loadc(m.emission);
                if (!black(le)) {
                    const c3 fl = dot3(wo, g.ns) > 0 ? le : C(0.f);
                    fb_add(a.fb, pix, w * fl);
                }
            }
            if (!bsdf_is_delta(m) && sv.nlights > 0) {
                // sample_lights (pathtracer_kernel.h:323-336): uniform by count, uniform on the triangle
                const uint32_t ln = sv.nlights;
                uint32_t lid = (uint32_t)(smp[0] * ln);
                if (lid > ln - 1) lid = ln - 1;
                const uint32_t lt = ldg(sv.lights + lid);
                const float sa = sqrtf(smp[1]);
                const Shade ls = load_shade(sv, lt);
                const Geom lg = setup_geom(ls, 1 - sa, smp[2] * sa);
                float lpdf = 1.f / tri_area(ls);
                lpdf /= ln;
                v3 wi = lg.p - g.p;
                const float d2 = dot3(wi, wi);
                const float dd = sqrtf(d2);
                wi = wi / dd;
                lpdf *= d2 / fabsf(dot3(wi, lg.ns));
                if (isinf(lpdf)) lpdf = 0.f;
                const c3 fl = dot3(-wi, lg.ns) > 0 ? emission_of(sv, ls) : C(0.f);
                c3 fb;
                float bpdf;
                bsdf_eval_local(m, to_local(fr, wo), to_local(fr, wi), fb, bpdf);
                const float wgt = power_heur(lpdf, bpdf);
                if (lpdf > 0.f) {
                    const c3 ow = w * fl * fb * wgt / lpdf;
                    if (!black(ow)) {
                        osh.o = g.p - rd * 1.0e-4f;
                        osh.d = lg.p - osh.o;
                        osh.tmin = 0.f;
                        osh.tmax = 0.9999f;
                        osh.w = ow;
                        osh.pdf = 0.f;
                        osh.spec = true;
                        has_sh = true;
                    }
                }
            }
            {
                const BSample bs = bsdf_sample(m, fr, wo, smp[5], smp[6], smp[4]);
                c3 ow = bs.pdf == 0.f ? C(0.f) : w * bs.f / bs.pdf;
                if ((MULTI || a.rng_mode != GPR_RNG_REFERENCE) && bounce >= rr_start && !black(ow)) {
                    const float qq = fmax_h(0.05f, 1 - cmax(ow));
                    if (smp[7] < qq) ow = C(0.f);
                    else ow = ow / (1 - qq);
                }
                if (!black(ow)) {
                    osc.o = g.p;
                    osc.d = bs.wi;
                    osc.tmin = 1e-4f;
                    osc.tmax = 1e38f;
                    osc.w = ow;
                    osc.pdf = bs.pdf;
                    osc.spec = bs.specular;
                    has_sc = true;
                }
            }
        } else if (sv.has_env) {
            c3 le = C(sv.env[0], sv.env[1], sv.env[2]);
            if (sv.env_tex) {
                const float4 ot = qi.o_tmin[i];
                le = env_le(sv, V(ot.x, ot.y, ot.z), rd);
            }
            fb_add(a.fb, pix, w * le);
        }
    }
    // ---- order-preserving compaction of both appends
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long bsh = __ballot(has_sh), bsc = __ballot(has_sc);
    const uint32_t lsh = __builtin_amdgcn_mbcnt_hi((uint32_t)(bsh >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bsh, 0u));
    const uint32_t lsc = __builtin_amdgcn_mbcnt_hi((uint32_t)(bsc >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bsc, 0u));
    if (lane == 0) {
        s_wave[0][wave] = (uint32_t)__popcll(bsh);
        s_wave[1][wave] = (uint32_t)__popcll(bsc);
    }
    __syncthreads();
    uint32_t wsh = 0, wsc = 0, tsh = 0, tsc = 0;
#pragma unroll
    for (int k = 0; k < kShadeBlock / 64; ++k) {
        if (k < wave) {
            wsh += s_wave[0][k];
            wsc += s_wave[1][k];
        }
        tsh += s_wave[0][k];
        tsc += s_wave[1][k];
    }
    if (threadIdx.x == 0) {
        unsigned long long* st = a.status + (uint64_t)bounce * a.ntiles;
        uint32_t psh = 0, psc = 0;
        if (tile == 0) {
            __hip_atomic_store(st, pack_status(2, tsh, tsc), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(st + tile, pack_status(1, tsh, tsc), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int j = (int)tile - 1;
            uint32_t spins = 0;
            while (j >= 0) {
                const unsigned long long v = __hip_atomic_load(st + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t f = (uint32_t)(v >> 62);
                if (f == 0) {
                    // predecessor tiles took their ids earlier and are resident: this wait ends
                    if (++spins > (1u << 26)) {
                        atomicOr(&a.ctrl->error, 2u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                psh += (uint32_t)((v >> 31) & 0x7fffffffu);
                psc += (uint32_t)(v & 0x7fffffffu);
                if (f == 2) break;
                --j;
            }
            __hip_atomic_store(st + tile, pack_status(2, psh + tsh, psc + tsc), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        s_base[0] = psh;
        s_base[1] = psc;
        if (tile == ntiles_in - 1) {
            a.ctrl->shadow_size[bounce] = psh + tsh;
            a.ctrl->scatter_size[bounce] = psc + tsc;
        }
    }
    __syncthreads();
    if (has_sh) {
        const uint32_t s = s_base[0] + wsh + lsh;
        qsh.o_tmin[s] = make_float4(osh.o.x, osh.o.y, osh.o.z, osh.tmin);
        qsh.d_tmax[s] = make_float4(osh.d.x, osh.d.y, osh.d.z, osh.tmax);
        qsh.w_pdf[s] = make_float4(osh.w.r, osh.w.g, osh.w.b, osh.pdf);
        qsh.pix[s] = pix | 0x80000000u;
        if constexpr (MULTI) qsh.job[s] = job;
    }
    if (has_sc) {
        const uint32_t s = s_base[1] + wsc + lsc;
        qs.o_tmin[s] = make_float4(osc.o.x, osc.o.y, osc.o.z, osc.tmin);
        qs.d_tmax[s] = make_float4(osc.d.x, osc.d.y, osc.d.z, osc.tmax);
        qs.w_pdf[s] = make_float4(osc.w.r, osc.w.g, osc.w.b, osc.pdf);
        qs.pix[s] = pix | (osc.spec ? 0x80000000u : 0u);
        qs.rng[s] = rng_out;
        if constexpr (MULTI) qs.job[s] = job;
    }
}

__global__ void k_finish(Ctrl* c) {
    if (threadIdx.x == 0) c->sample = c->sample + 1;
}

__global__ void k_clear_fb(float4* fb, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) fb[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_set_sample(Ctrl* c, uint32_t s) {
    if (threadIdx.x == 0) {
        c->sample = s;
        c->error = 0;
    }
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------- engine ----
// Kernel set of one traversal configuration of the multi-job pass.
namespace {
struct PassKernels {
    void (*ext_primary)(RenderArgs, uint32_t);
    void (*ext_scatter)(RenderArgs, uint32_t);
    void (*shadow)(RenderArgs, uint32_t);
    int nt, sl, layout;
};
template <int NT, int SL, bool R>
PassKernels pass_kernels() {
    return {k_extend<true, false, true, NT, SL, R>, k_extend<false, false, true, NT, SL, R>,
            k_shadow<false, true, NT, SL, R>, NT, SL, 0};
}
const PassKernels& pass_kernel_table(int idx) {
    static const PassKernels t[] = {
        pass_kernels<64, 8, false>(),   pass_kernels<64, 16, false>(),  pass_kernels<64, 32, false>(),
        pass_kernels<128, 8, false>(),  pass_kernels<128, 16, false>(), pass_kernels<128, 32, false>(),
        pass_kernels<256, 8, false>(),  pass_kernels<256, 16, false>(), pass_kernels<256, 32, false>(),
        pass_kernels<512, 8, false>(),  pass_kernels<512, 16, false>(), pass_kernels<512, 32, false>(),
        pass_kernels<64, 8, true>(),    pass_kernels<64, 16, true>(),   pass_kernels<64, 32, true>(),
        pass_kernels<128, 8, true>(),   pass_kernels<128, 16, true>(),  pass_kernels<128, 32, true>(),
        pass_kernels<256, 8, true>(),   pass_kernels<256, 16, true>(),  pass_kernels<256, 32, true>(),
        pass_kernels<512, 8, true>(),   pass_kernels<512, 16, true>(),  pass_kernels<512, 32, true>(),
        {k_extend<true, true, true, 256, 16>, k_extend<false, true, true, 256, 16>, k_shadow<true, true, 256, 16>,
         256, 16, 1},
    };
    return t[idx];
}
constexpr int kPassKernelCount = 25;
int pass_kernel_index(const TraceCfg& c) {
    if (c.layout == 1) return c.refill ? -1 : 24;
    const int a = c.nt == 64 ? 0 : c.nt == 128 ? 1 : c.nt == 256 ? 2 : c.nt == 512 ? 3 : -1;
    const int b = c.sl == 8 ? 0 : c.sl == 16 ? 1 : c.sl == 32 ? 2 : -1;
    return (a < 0 || b < 0) ? -1 : (c.refill ? 12 : 0) + 3 * a + b;
}
}  // namespace

std::vector<TraceCfg> engine_tune_space() {
    std::vector<TraceCfg> v;
    for (int nt : {64, 128, 256, 512})
        for (int sl : {8, 16, 32}) {
            for (int chunk : {64, 128, 256, 512}) v.push_back(TraceCfg{nt, sl, chunk, 0, 0});
            for (int chunk : {64, 128}) v.push_back(TraceCfg{nt, sl, chunk, 0, 1});
        }
    v.push_back(TraceCfg{256, 16, 64, 1, 0});
    return v;
}

namespace {
int alloc_queue(DevBuf& mem, QueueView& q, uint32_t n, DevBuf* seed, bool with_job) {
    const size_t per = 16 * 4 + 4 + 8 + (with_job ? 4 : 0);
    GPR_TRY(mem.alloc((size_t)n * per + 8 * 256));
    char* p = mem.as<char>();
    size_t off = 0;
    auto take = [&](size_t bytes) {
        off = (off + 255) & ~(size_t)255;
        char* r = p + off;
        off += bytes;
        return r;
    };
    q.o_tmin = reinterpret_cast<float4*>(take(16ull * n));
    q.d_tmax = reinterpret_cast<float4*>(take(16ull * n));
    q.hit = reinterpret_cast<float4*>(take(16ull * n));
    q.w_pdf = reinterpret_cast<float4*>(take(16ull * n));
    q.pix = reinterpret_cast<uint32_t*>(take(4ull * n));
    q.rng = reinterpret_cast<unsigned long long*>(take(8ull * n));
    q.job = with_job ? reinterpret_cast<uint32_t*>(take(4ull * n)) : nullptr;
    q.seed = nullptr;
    if (seed) {
        GPR_TRY(seed->alloc(4ull * n));
        GPR_HIP(hipMemsetAsync(seed->p, 0, 4ull * n, stream()));  // DEVICE_BUFFER allocations start zeroed (buffer.h:62)
        q.seed = seed->as<uint32_t>();
    }
    return GPR_OK;
}
}  // namespace

struct Engine {
    int device = 0;
    uint32_t capacity = 0, max_jobs = 0, ntiles = 0;
    DevBuf qmem[3], ctrl, status, spill, pass, scenes;
    QueueView q[3];
    int grid[kPassKernelCount] = {};
    hipStream_t cap_stream = nullptr;
    static constexpr int kStage = 4;
    PassBuf* h_pass[kStage] = {};
    hipEvent_t ev[kStage] = {};
    int slot = 0;
    struct GraphKey {
        uint32_t depth;
        int kidx;
        uint32_t chunk;
        uint32_t size;  // launch size class: paths rounded up to a power of two
        float4* fb;
        bool operator==(const GraphKey& o) const {
            return depth == o.depth && kidx == o.kidx && chunk == o.chunk && size == o.size && fb == o.fb;
        }
    };
    std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs;
    ~Engine() {
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
        for (int k = 0; k < kStage; ++k) {
            if (ev[k]) (void)hipEventDestroy(ev[k]);
            if (h_pass[k]) (void)hipHostFree(h_pass[k]);
        }
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
    }
};

int engine_create(int device, uint32_t capacity, uint32_t max_jobs, Engine** out) {
    if (capacity == 0 || capacity >= (1u << 30) || max_jobs == 0 || max_jobs > kMaxPassJobs)
        return fail(GPR_ERR_INVALID, "engine: capacity or job count out of range");
    GPR_HIP(hipSetDevice(device));
    Engine* e = new (std::nothrow) Engine();
    if (!e) return fail(GPR_ERR_OOM, "engine: host allocation failed");
    std::unique_ptr<Engine> guard(e);
    e->device = device;
    e->capacity = capacity;
    e->max_jobs = max_jobs;
    e->ntiles = div_up(capacity, kShadeBlock);
    for (int k = 0; k < 3; ++k) GPR_TRY(alloc_queue(e->qmem[k], e->q[k], capacity, nullptr, true));
    GPR_TRY(e->ctrl.alloc(sizeof(Ctrl)));
    GPR_HIP(hipMemsetAsync(e->ctrl.p, 0, sizeof(Ctrl), stream()));
    GPR_TRY(e->status.alloc(8ull * kMaxBounce * e->ntiles));
    GPR_TRY(e->pass.alloc(sizeof(PassBuf)));
    int cus = 0;
    GPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    size_t threads = 0;
    for (int k = 0; k < kPassKernelCount; ++k) {
        const PassKernels& pk = pass_kernel_table(k);
        int per_cu = 0;
        GPR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pk.ext_scatter, pk.nt, 0));
        // more persistent threads than queue entries would only idle
        e->grid[k] = std::max(1, std::min<int>(cus * std::max(1, per_cu), (int)div_up(capacity, pk.nt)));
        threads = std::max(threads, (size_t)e->grid[k] * pk.nt);
    }
    GPR_TRY(e->spill.alloc(threads * kStackGlobal * sizeof(int2)));
    for (int k = 0; k < Engine::kStage; ++k) {
        GPR_HIP(hipHostMalloc((void**)&e->h_pass[k], sizeof(PassBuf), hipHostMallocDefault));
        GPR_HIP(hipEventCreateWithFlags(&e->ev[k], hipEventDisableTiming));
    }
    GPR_HIP(hipStreamCreateWithFlags(&e->cap_stream, hipStreamNonBlocking));
    GPR_HIP(hipStreamSynchronize(stream()));
    *out = guard.release();
    return GPR_OK;
}

void engine_destroy(Engine* e) { delete e; }

size_t engine_device_bytes(const Engine* e) {
    return e->qmem[0].bytes + e->qmem[1].bytes + e->qmem[2].bytes + e->ctrl.bytes + e->status.bytes +
           e->spill.bytes + e->pass.bytes + e->scenes.bytes;
}

int engine_set_scenes(Engine* e, const std::vector<SceneView>& views, hipStream_t st) {
    if (views.empty()) return GPR_OK;
    const size_t bytes = views.size() * sizeof(SceneView);
    if (bytes > e->scenes.bytes) {
        GPR_HIP(hipStreamSynchronize(st));
        GPR_TRY(e->scenes.alloc(bytes));
        for (auto& g : e->graphs) (void)hipGraphExecDestroy(g.second);
        e->graphs.clear();
    }
    GPR_HIP(hipMemcpyAsync(e->scenes.p, views.data(), bytes, hipMemcpyHostToDevice, st));
    GPR_HIP(hipStreamSynchronize(st));
    return GPR_OK;
}

namespace {
// Launch grids cover `size` paths (the pass's size class), not the capacity:
// a small pass launches small grids.
int record_pass(Engine* e, uint32_t depth, int kidx, uint32_t chunk, uint32_t size, float4* fb, hipStream_t st) {
    RenderArgs a{};
    a.scenes = e->scenes.as<SceneView>();
    a.pass = e->pass.as<PassBuf>();
    for (int k = 0; k < 3; ++k) a.q[k] = e->q[k];
    a.fb = fb;
    a.ctrl = e->ctrl.as<Ctrl>();
    a.status = e->status.as<unsigned long long>();
    a.spill = e->spill.as<int2>();
    a.ntiles = e->ntiles;
    a.depth = depth;
    a.chunk = chunk;
    a.rng_mode = GPR_RNG_PCG;
    const PassKernels& pk = pass_kernel_table(kidx);
    const dim3 tg(std::min<uint32_t>(e->grid[kidx], div_up(size, pk.nt))), tb(pk.nt);
    hipLaunchKernelGGL(k_gen<true>, dim3(div_up(size, 256)), dim3(256), 0, st, a);
    hipLaunchKernelGGL(pk.ext_primary, tg, tb, 0, st, a, 0u);
    for (uint32_t b = 0; b < depth; ++b) {
        hipLaunchKernelGGL(k_shade<true>, dim3(div_up(size, kShadeBlock)), dim3(kShadeBlock), 0, st, a, b);
        hipLaunchKernelGGL(pk.shadow, tg, tb, 0, st, a, b);
        hipLaunchKernelGGL(pk.ext_scatter, tg, tb, 0, st, a, b);
    }
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}
}  // namespace

int engine_run_pass(Engine* e, const EngineJob* jobs, uint32_t njobs, float4* fb, const TraceCfg& cfg,
                    hipStream_t st) {
    if (njobs == 0) return GPR_OK;
    if (njobs > e->max_jobs) return fail(GPR_ERR_INVALID, "engine: too many jobs in one pass");
    const int kidx = pass_kernel_index(cfg);
    if (kidx < 0 || cfg.chunk < 64 || cfg.chunk % 64 != 0 || cfg.chunk > 4096)
        return fail(GPR_ERR_INVALID, "engine: traversal configuration outside the tuned set");
    const int s = e->slot;
    e->slot = (s + 1) % Engine::kStage;
    GPR_HIP(hipEventSynchronize(e->ev[s]));  // the staging slot's previous upload is done
    PassBuf* h = e->h_pass[s];
    uint64_t off = 0;
    uint32_t depth = 0;
    for (uint32_t k = 0; k < njobs; ++k) {
        const EngineJob& j = jobs[k];
        if (j.npix == 0 || j.width == 0 || j.max_depth > kMaxBounce)
            return fail(GPR_ERR_INVALID, "engine: malformed job");
        h->jobs[k].j = j;
        h->jobs[k].path_off = (uint32_t)off;
        off += j.npix;
        depth = std::max(depth, j.max_depth);
    }
    if (off > e->capacity) return fail(GPR_ERR_INVALID, "engine: pass exceeds the queue capacity");
    h->npaths = (uint32_t)off;
    h->njobs = njobs;
    const size_t bytes = offsetof(PassBuf, jobs) + sizeof(PassJob) * njobs;
    GPR_HIP(hipMemcpyAsync(e->pass.p, h, bytes, hipMemcpyHostToDevice, st));
    GPR_HIP(hipEventRecord(e->ev[s], st));
    if (depth == 0) depth = 1;
    uint32_t size = 1u << 16;
    while (size < off) size <<= 1;
    size = std::min(size, e->capacity);
    const Engine::GraphKey key{depth, kidx, (uint32_t)cfg.chunk, size, fb};
    hipGraphExec_t exec = nullptr;
    for (auto& g : e->graphs)
        if (g.first == key) exec = g.second;
    if (!exec) {
        hipGraph_t graph;
        GPR_HIP(hipStreamBeginCapture(e->cap_stream, hipStreamCaptureModeThreadLocal));
        const int rc = record_pass(e, depth, kidx, (uint32_t)cfg.chunk, size, fb, e->cap_stream);
        const hipError_t ec = hipStreamEndCapture(e->cap_stream, &graph);
        GPR_TRY(rc);
        if (ec != hipSuccess) return fail(GPR_ERR_HIP, std::string("pass capture: ") + hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return fail(GPR_ERR_HIP, std::string("pass instantiate: ") + hipGetErrorString(ei));
        e->graphs.push_back({key, exec});
    }
    GPR_HIP(hipGraphLaunch(exec, st));
    return GPR_OK;
}

int engine_errors(Engine* e, hipStream_t st, uint32_t* flags) {
    uint32_t err = 0;
    GPR_HIP(hipMemcpyAsync(&err, reinterpret_cast<char*>(e->ctrl.p) + offsetof(Ctrl, error), 4,
                           hipMemcpyDeviceToHost, st));
    GPR_HIP(hipStreamSynchronize(st));
    *flags = err;
    return GPR_OK;
}

namespace {
__global__ __launch_bounds__(256) void k_pack_shade(const float* __restrict__ pos, const float* __restrict__ nrm,
                                                    const uint8_t* __restrict__ flags,
                                                    const uint32_t* __restrict__ mat,
                                                    const float* __restrict__ uv, uint32_t n,
                                                    TriShade* __restrict__ out) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const float* p = pos + 9ull * t;
    const float* q = nrm + 9ull * t;
    const uint32_t f = flags[t];
    float w[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (uv) {
#pragma unroll
        for (int k = 0; k < 6; ++k) w[k] = uv[6ull * t + k];
    }
    TriShade r;
    r.a = make_float4(p[0], p[1], p[2], p[3]);
    r.b = make_float4(p[4], p[5], p[6], p[7]);
    r.c = make_float4(p[8], q[0], q[1], q[2]);
    r.d = make_float4(q[3], q[4], q[5], q[6]);
    r.e = make_float4(q[7], q[8], __uint_as_float(f), __uint_as_float(mat[t]));
    r.f = make_float4(w[0], w[1], w[2], w[3]);
    r.g = make_float4(w[4], w[5], 0.f, 0.f);
    r.h = make_float4(0.f, 0.f, 0.f, 0.f);
    out[t] = r;
}
}  // namespace

int pack_shade(gpr_scene* s) {
    const uint32_t n = (uint32_t)s->mat.size();
    GPR_TRY(s->d_shade.alloc(std::max<size_t>(sizeof(TriShade) * n, 256)));
    if (n == 0) return GPR_OK;
    const float* uv = s->uv.empty() || s->textures.empty() ? nullptr : s->d_uv.as<float>();
    hipLaunchKernelGGL(k_pack_shade, dim3((n + 255) / 256), dim3(256), 0, stream(), s->d_pos.as<float>(),
                       s->d_nrm.as<float>(), s->d_flags.as<uint8_t>(), s->d_mat.as<uint32_t>(), uv, n,
                       s->d_shade.as<TriShade>());
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

}  // namespace gpr

using namespace gpr;

struct gpr_renderer {
    gpr_scene* scene = nullptr;
    gpr_render_opts opts{};
    uint32_t npix = 0, ntiles = 0;
    uint32_t spp_done = 0;
    DevBuf fb, ctrl, status, spill;
    DevBuf qmem[3];
    DevBuf seeds[2];
    QueueView q[3];
    int grid_trace = 0;
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t graph = nullptr;
    SceneView baked{};  // scene view (camera included) the graph was captured with
    SceneView accum{};  // scene view the accumulated samples were taken with
    bool reset_pending = false;  // PathTracer::m_reset: the next render starts a new accumulation
    ~gpr_renderer() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
    }
};

namespace {
RenderArgs make_args(gpr_renderer* r) {
    RenderArgs a{};
    a.sv = r->scene->view();
    for (int k = 0; k < 3; ++k) a.q[k] = r->q[k];
    a.fb = r->fb.as<float4>();
    a.ctrl = r->ctrl.as<Ctrl>();
    a.status = r->status.as<unsigned long long>();
    a.spill = r->spill.as<int2>();
    a.npix = r->npix;
    a.ntiles = r->ntiles;
    a.depth = r->opts.max_depth;
    a.chunk = 64;
    a.rng_mode = r->opts.rng_mode;
    a.rr_start = r->opts.rr_start;
    a.seed = r->opts.seed;
    return a;
}

constexpr int kRenderNT = 256, kRenderSL = 8;  // block size and LDS stack entries per lane of the single-scene renderer's traversal

// One sample: gen -> extend(primary) -> max_depth x {shade, shadow, extend} -> finish.
int launch_sample(gpr_renderer* r, hipStream_t st) {
    const RenderArgs a = make_args(r);
    const bool ref = r->opts.traversal == GPR_TRAVERSAL_REF;
    const dim3 tg(r->grid_trace), tb(kRenderNT);
    constexpr int NT = kRenderNT, SL = kRenderSL;
    hipLaunchKernelGGL(k_gen<false>, dim3(div_up(r->npix, 256)), dim3(256), 0, st, a);
    if (ref) hipLaunchKernelGGL((k_extend<true, true, false, NT, SL>), tg, tb, 0, st, a, 0u);
    else hipLaunchKernelGGL((k_extend<true, false, false, NT, SL>), tg, tb, 0, st, a, 0u);
    for (uint32_t b = 0; b < r->opts.max_depth; ++b) {
        hipLaunchKernelGGL(k_shade<false>, dim3(r->ntiles), dim3(kShadeBlock), 0, st, a, b);
        if (ref) {
            hipLaunchKernelGGL((k_shadow<true, false, NT, SL>), tg, tb, 0, st, a, b);
            hipLaunchKernelGGL((k_extend<false, true, false, NT, SL>), tg, tb, 0, st, a, b);
        } else {
            hipLaunchKernelGGL((k_shadow<false, false, NT, SL>), tg, tb, 0, st, a, b);
            hipLaunchKernelGGL((k_extend<false, false, false, NT, SL>), tg, tb, 0, st, a, b);
        }
    }
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, st, a.ctrl);
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

// The sample sequence captured once into a hipGraph and replayed per sample;
// captured again when the scene view (camera included) has changed.
// Every field of the view is baked into the captured graph's kernel
// arguments; gpr_scene::view() value-initialises the struct, so padding is
// zero and a byte comparison covers all of them.
bool same_view(const SceneView& a, const SceneView& b) { return std::memcmp(&a, &b, sizeof(SceneView)) == 0; }

int sample_graph(gpr_renderer* r, hipGraphExec_t* out) {
    const SceneView now = r->scene->view();
    if (r->graph && !same_view(now, r->baked)) {
        GPR_HIP(hipStreamSynchronize(stream()));
        (void)hipGraphExecDestroy(r->graph);
        r->graph = nullptr;
    }
    if (!r->graph) {
        r->baked = now;
        if (!r->cap_stream) GPR_HIP(hipStreamCreateWithFlags(&r->cap_stream, hipStreamNonBlocking));
        hipGraph_t g;
        GPR_HIP(hipStreamBeginCapture(r->cap_stream, hipStreamCaptureModeThreadLocal));
        const int rc = launch_sample(r, r->cap_stream);
        const hipError_t ec = hipStreamEndCapture(r->cap_stream, &g);
        GPR_TRY(rc);
        if (ec != hipSuccess) return fail(GPR_ERR_HIP, std::string("sample capture: ") + hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&r->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            r->graph = nullptr;
            return fail(GPR_ERR_HIP, std::string("sample instantiate: ") + hipGetErrorString(ei));
        }
    }
    *out = r->graph;
    return GPR_OK;
}
}  // namespace

extern "C" {

}  // extern "C"

namespace {
// Frame buffer, queues and traversal resources for the scene's current image size.
int renderer_alloc(gpr_renderer* r) {
    gpr_scene* s = r->scene;
    const uint64_t npix = (uint64_t)s->width * s->height;
    if (npix == 0 || npix >= (1ull << 31)) return fail(GPR_ERR_INVALID, "image size out of range");
    r->npix = (uint32_t)npix;
    r->ntiles = div_up(npix, kShadeBlock);
    GPR_HIP(hipSetDevice(s->device));
    GPR_TRY(r->fb.alloc(16ull * npix));
    GPR_TRY(r->ctrl.alloc(sizeof(Ctrl)));
    GPR_TRY(r->status.alloc(8ull * kMaxBounce * r->ntiles));
    for (int k = 0; k < 3; ++k)
        GPR_TRY(alloc_queue(r->qmem[k], r->q[k], r->npix, k < 2 ? &r->seeds[k] : nullptr, false));
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_extend<false, false, false, kRenderNT, kRenderSL>,
                                                     kRenderNT, 0) != hipSuccess)
        return fail(GPR_ERR_HIP, "occupancy query failed");
    r->grid_trace = std::max(1, cus * std::max(1, per_cu));
    r->grid_trace = std::min<int>(r->grid_trace, (int)div_up(npix, kRenderNT));
    GPR_TRY(r->spill.alloc((size_t)r->grid_trace * kRenderNT * kStackGlobal * sizeof(int2)));
    return GPR_OK;
}
}  // namespace

extern "C" {

int gpr_renderer_create(gpr_scene* s, const gpr_render_opts* o, gpr_renderer** out) {
    if (!s || !o || !out) return fail(GPR_ERR_INVALID, "gpr_renderer_create: null argument");
    if (!s->committed) return fail(GPR_ERR_INVALID, "gpr_renderer_create: scene not committed");
    if (o->max_depth > kMaxBounce) return fail(GPR_ERR_INVALID, "max_depth above 64");
    if (o->rng_mode > GPR_RNG_PCG || o->traversal > GPR_TRAVERSAL_REF)
        return fail(GPR_ERR_INVALID, "gpr_renderer_create: unknown mode");
    gpr_renderer* r = new (std::nothrow) gpr_renderer();
    if (!r) return fail(GPR_ERR_OOM, "gpr_renderer_create: host allocation failed");
    r->scene = s;
    r->opts = *o;
    int rc = renderer_alloc(r);
    if (rc == GPR_OK) rc = gpr_renderer_reset(r);
    if (rc != GPR_OK) {
        delete r;
        return rc;
    }
    ++s->renderers;
    *out = r;
    return GPR_OK;
}

int gpr_renderer_zoom(gpr_renderer* r, float d) {
    if (!r) return fail(GPR_ERR_INVALID, "gpr_renderer_zoom: null renderer");
    float* e = r->scene->eye;
    const float* v = r->scene->dir;
    for (int k = 0; k < 3; ++k) e[k] = e[k] + v[k] * d;  // m_eye += m_dir * d
    r->reset_pending = true;
    return GPR_OK;
}

int gpr_renderer_translate(gpr_renderer* r, float x, float y) {
    if (!r) return fail(GPR_ERR_INVALID, "gpr_renderer_translate: null renderer");
    gpr_scene* s = r->scene;
    const v3 d = norm3(V(s->dir[0], s->dir[1], s->dir[2]));
    const v3 h = norm3(cross3(d, V(s->up[0], s->up[1], s->up[2])));
    const v3 vv = norm3(cross3(h, d));
    const v3 m = h * x + vv * y;
    s->eye[0] = s->eye[0] + m.x;
    s->eye[1] = s->eye[1] + m.y;
    s->eye[2] = s->eye[2] + m.z;
    r->reset_pending = true;
    return GPR_OK;
}

int gpr_renderer_rotate(gpr_renderer* r, float yaw, float pitch) {
    if (!r) return fail(GPR_ERR_INVALID, "gpr_renderer_rotate: null renderer");
    gpr_scene* s = r->scene;
    s->yaw += yaw;
    s->pitch += pitch;
    s->dir[0] = std::sin(s->yaw) * std::cos(s->pitch);
    s->dir[1] = std::sin(s->pitch);
    s->dir[2] = std::cos(s->yaw) * std::cos(s->pitch);
    r->reset_pending = true;
    return GPR_OK;
}

int gpr_renderer_resize(gpr_renderer* r, uint32_t w, uint32_t h) {
    if (!r || w == 0 || h == 0) return fail(GPR_ERR_INVALID, "gpr_renderer_resize: bad argument");
    gpr_scene* s = r->scene;
    GPR_HIP(hipStreamSynchronize(stream()));
    s->width = w;
    s->height = h;
    if (!s->cam_free) camera_raster_to_camera(w, h, s->fov, s->r2c);
    // Camera::updateAspectRation: fov.y = atan(tan(fov.x * 0.5) / aspect) * 2 (in double)
    const float aspect = (float)w / (float)h;
    s->fovy = (float)(std::atan(std::tan(s->fovx * 0.5) / aspect) * 2.0);
    if (r->graph) {
        (void)hipGraphExecDestroy(r->graph);
        r->graph = nullptr;
    }
    GPR_TRY(renderer_alloc(r));
    return gpr_renderer_reset(r);
}

int gpr_renderer_get_resolution(const gpr_renderer* r, uint32_t* w, uint32_t* h) {
    if (!r || !w || !h) return fail(GPR_ERR_INVALID, "gpr_renderer_get_resolution: null argument");
    *w = r->scene->width;
    *h = r->scene->height;
    return GPR_OK;
}

int gpr_renderer_reset(gpr_renderer* r) {
    if (!r) return fail(GPR_ERR_INVALID, "gpr_renderer_reset: null renderer");
    hipLaunchKernelGGL(k_clear_fb, dim3(div_up(r->npix, 256)), dim3(256), 0, stream(), r->fb.as<float4>(), r->npix);
    hipLaunchKernelGGL(k_set_sample, dim3(1), dim3(64), 0, stream(), r->ctrl.as<Ctrl>(), 0u);
    GPR_HIP(hipGetLastError());
    if (r->graph) {
        GPR_HIP(hipStreamSynchronize(stream()));
        (void)hipGraphExecDestroy(r->graph);
        r->graph = nullptr;
    }
    r->spp_done = 0;
    r->reset_pending = false;
    r->accum = r->scene->view();
    return GPR_OK;
}

int gpr_renderer_set_opts(gpr_renderer* r, const gpr_render_opts* o) {
    if (!r || !o) return fail(GPR_ERR_INVALID, "gpr_renderer_set_opts: null argument");
    if (o->max_depth > kMaxBounce) return fail(GPR_ERR_INVALID, "max_depth above 64");
    if (o->rng_mode > GPR_RNG_PCG || o->traversal > GPR_TRAVERSAL_REF)
        return fail(GPR_ERR_INVALID, "gpr_renderer_set_opts: unknown mode");
    r->opts = *o;  // baked into the sample graph: reset drops it
    return gpr_renderer_reset(r);
}

int gpr_renderer_restore(gpr_renderer* r, const float* rgb, uint32_t spp) {
    if (!r || !rgb) return fail(GPR_ERR_INVALID, "gpr_renderer_restore: null argument");
    if (r->opts.rng_mode != GPR_RNG_PCG)
        return fail(GPR_ERR_UNSUPPORTED,
                    "gpr_renderer_restore: reference RNG mode keeps slot-indexed seeds across samples; "
                    "only PCG renders can be continued");
    std::vector<float4> tmp(r->npix);
    for (uint64_t i = 0; i < r->npix; ++i) tmp[i] = make_float4(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], 0.f);
    GPR_HIP(hipMemcpyAsync(r->fb.p, tmp.data(), 16ull * r->npix, hipMemcpyHostToDevice, stream()));
    hipLaunchKernelGGL(k_set_sample, dim3(1), dim3(64), 0, stream(), r->ctrl.as<Ctrl>(), spp);
    GPR_HIP(hipGetLastError());
    GPR_HIP(hipStreamSynchronize(stream()));  // tmp is released on return
    r->spp_done = spp;
    r->reset_pending = false;
    r->accum = r->scene->view();
    return GPR_OK;
}

int gpr_render(gpr_renderer* r, uint32_t spp) {
    if (!r) return fail(GPR_ERR_INVALID, "gpr_render: null renderer");
    if (r->scene->bvh.n == 0) return fail(GPR_ERR_UNSUPPORTED, "gpr_render: scene has no triangles");
    if ((uint64_t)r->scene->width * r->scene->height != r->npix)
        return fail(GPR_ERR_INVALID, "gpr_render: the scene's image size changed; use gpr_renderer_resize");
    if (r->reset_pending) GPR_TRY(gpr_renderer_reset(r));
    // The camera lives on the scene: another renderer's zoom / translate /
    // rotate on the same scene, or a new camera set on it, changes the view
    // under this renderer.  Samples of two views must not share one sum, so
    // the accumulation starts again, as PathTracer::m_reset makes it
    // (pathtracer.cu:177-200).
    if (r->spp_done > 0 && !same_view(r->scene->view(), r->accum)) GPR_TRY(gpr_renderer_reset(r));
    hipStream_t st = stream();
    hipGraphExec_t g = nullptr;
    GPR_TRY(sample_graph(r, &g));
    for (uint32_t s = 0; s < spp; ++s) GPR_HIP(hipGraphLaunch(g, st));
    r->spp_done += spp;
    uint32_t err = 0;
    GPR_HIP(hipMemcpyAsync(&err, reinterpret_cast<char*>(r->ctrl.p) + offsetof(Ctrl, error), 4,
                           hipMemcpyDeviceToHost, st));
    GPR_HIP(hipStreamSynchronize(st));
    if (err) return fail(GPR_ERR_INVALID, "render error flags " + std::to_string(err));
    return GPR_OK;
}

int gpr_renderer_read(gpr_renderer* r, float* rgb, uint32_t* cnt) {
    if (!r) return fail(GPR_ERR_INVALID, "gpr_renderer_read: null renderer");
    std::vector<float> tmp(4ull * r->npix);
    GPR_HIP(hipMemcpyAsync(tmp.data(), r->fb.p, 16ull * r->npix, hipMemcpyDeviceToHost, stream()));
    GPR_HIP(hipStreamSynchronize(stream()));
    for (uint64_t i = 0; i < r->npix; ++i) {
        if (rgb) {
            rgb[3 * i + 0] = tmp[4 * i + 0];
            rgb[3 * i + 1] = tmp[4 * i + 1];
            rgb[3 * i + 2] = tmp[4 * i + 2];
        }
        if (cnt) cnt[i] = r->spp_done;
    }
    return GPR_OK;
}

int gpr_renderer_queue_sizes(gpr_renderer* r, uint32_t* shadow, uint32_t* scatter, uint32_t cap) {
    if (!r) return fail(GPR_ERR_INVALID, "gpr_renderer_queue_sizes: null renderer");
    Ctrl c;
    GPR_HIP(hipMemcpyAsync(&c, r->ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, stream()));
    GPR_HIP(hipStreamSynchronize(stream()));
    for (uint32_t b = 0; b < cap && b < kMaxBounce; ++b) {
        if (shadow) shadow[b] = c.shadow_size[b];
        if (scatter) scatter[b] = c.scatter_size[b];
    }
    return GPR_OK;
}

int gpr_renderer_free(gpr_renderer* r) {
    if (r) {
        (void)hipStreamSynchronize(stream());
        --r->scene->renderers;
        delete r;
    }
    return GPR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- PNG ----
namespace {
uint32_t crc_table[256];
bool crc_ready = false;
uint32_t crc32(const unsigned char* p, size_t n, uint32_t c = 0xffffffffu) {
    if (!crc_ready) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t v = i;
            for (int k = 0; k < 8; ++k) v = (v & 1) ? 0xedb88320u ^ (v >> 1) : v >> 1;
            crc_table[i] = v;
        }
        crc_ready = true;
    }
    for (size_t i = 0; i < n; ++i) c = crc_table[(c ^ p[i]) & 0xff] ^ (c >> 8);
    return c;
}
void put32(std::vector<unsigned char>& o, uint32_t v) {
    o.push_back((unsigned char)(v >> 24));
    o.push_back((unsigned char)(v >> 16));
    o.push_back((unsigned char)(v >> 8));
    o.push_back((unsigned char)v);
}
void chunk(std::vector<unsigned char>& o, const char* type, const std::vector<unsigned char>& data) {
    put32(o, (uint32_t)data.size());
    std::vector<unsigned char> td(type, type + 4);
    td.insert(td.end(), data.begin(), data.end());
    o.insert(o.end(), td.begin(), td.end());
    put32(o, crc32(td.data(), td.size()) ^ 0xffffffffu);
}
// 8-bit RGB PNG with stored (uncompressed) deflate blocks.
int write_png(const std::string& path, uint32_t w, uint32_t h, const std::vector<unsigned char>& rgb) {
    std::vector<unsigned char> raw;
    raw.reserve((3ull * w + 1) * h);
    for (uint32_t y = 0; y < h; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb.begin() + 3ull * w * y, rgb.begin() + 3ull * w * (y + 1));
    }
    std::vector<unsigned char> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - pos);
        z.push_back(pos + len == raw.size() ? 1 : 0);
        z.push_back((unsigned char)(len & 0xff));
        z.push_back((unsigned char)(len >> 8));
        z.push_back((unsigned char)(~len & 0xff));
        z.push_back((unsigned char)((~len >> 8) & 0xff));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
        pos += len;
    } while (pos < raw.size());
    uint32_t s1 = 1, s2 = 0;
    for (unsigned char c : raw) {
        s1 = (s1 + c) % 65521u;
        s2 = (s2 + s1) % 65521u;
    }
    put32(z, (s2 << 16) | s1);
    std::vector<unsigned char> out = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
    std::vector<unsigned char> ihdr;
    put32(ihdr, w);
    put32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(GPR_ERR_IO, "cannot open " + path);
    const size_t wr = std::fwrite(out.data(), 1, out.size(), f);
    std::fclose(f);
    if (wr != out.size()) return fail(GPR_ERR_IO, "short write to " + path);
    return GPR_OK;
}
// Radiance RGBE of one pixel (Ward's float2rgbe): e from the largest channel,
// mantissas scaled so that channel / 2^(e-8) fits a byte.  Non-finite or
// negative channels are written as 0.
void rgbe_of(const float* c, unsigned char* o) {
    float v[3];
    for (int k = 0; k < 3; ++k) v[k] = std::isfinite(c[k]) && c[k] > 0.f ? c[k] : 0.f;
    const float m = std::max(v[0], std::max(v[1], v[2]));
    if (m < 1e-32f) {
        o[0] = o[1] = o[2] = o[3] = 0;
        return;
    }
    int e;
    const float scale = std::frexp(m, &e) * 256.f / m;
    for (int k = 0; k < 3; ++k) o[k] = (unsigned char)std::min(255.f, v[k] * scale);
    o[3] = (unsigned char)(e + 128);
}
// Radiance .hdr, 32-bit_rle_rgbe: new-style scanlines (each component as
// literal dumps of up to 128 bytes) where the format allows them (8 <= w <
// 32768), flat pixels otherwise; `rgb` rows top first.
int write_hdr(const std::string& path, uint32_t w, uint32_t h, const std::vector<float>& rgb) {
    std::string out = "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y " + std::to_string(h) + " +X " + std::to_string(w) + "\n";
    std::vector<unsigned char> px(4ull * w);
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) rgbe_of(&rgb[3 * ((uint64_t)y * w + x)], &px[4 * x]);
        if (w < 8 || w >= 32768) {
            out.append(reinterpret_cast<const char*>(px.data()), px.size());
            continue;
        }
        out += (char)2;
        out += (char)2;
        out += (char)(w >> 8);
        out += (char)(w & 0xff);
        for (int c = 0; c < 4; ++c)
            for (uint32_t x0 = 0; x0 < w; x0 += 128) {
                const uint32_t cnt = std::min<uint32_t>(128, w - x0);
                out += (char)cnt;
                for (uint32_t x = x0; x < x0 + cnt; ++x) out += (char)px[4 * x + c];
            }
    }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(GPR_ERR_IO, "cannot open " + path);
    const size_t wr = std::fwrite(out.data(), 1, out.size(), f);
    std::fclose(f);
    if (wr != out.size()) return fail(GPR_ERR_IO, "short write to " + path);
    return GPR_OK;
}
}  // namespace

// filter_kernel (pathtracer_kernel.h:792-810): mean, sRGB bytes, alpha from
// GammaCorrect(1); pixel i of the frame buffer to out[i] (R in the low byte).
__global__ __launch_bounds__(256) void k_resolve_rgba8(const float4* __restrict__ fb, uint32_t npix, float n,
                                                       uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= npix) return;
    const float4 s = fb[i];
    out[i] = srgb_byte(s.x / n) | (srgb_byte(s.y / n) << 8) | (srgb_byte(s.z / n) << 16) | (srgb_byte(1.f) << 24);
}

extern "C" int gpr_renderer_resolve_rgba8(gpr_renderer* r, uint32_t* out, uint32_t flags) {
    if (!r || !out) return fail(GPR_ERR_INVALID, "gpr_renderer_resolve_rgba8: null argument");
    if (flags & ~GPR_RESOLVE_DEVICE_PTRS) return fail(GPR_ERR_INVALID, "gpr_renderer_resolve_rgba8: unknown flags");
    const float n = (float)std::max(1u, r->spp_done);
    const uint32_t grid = (uint32_t)((r->npix + 255) / 256);
    if (flags & GPR_RESOLVE_DEVICE_PTRS) {
        hipLaunchKernelGGL(k_resolve_rgba8, dim3(grid), dim3(256), 0, stream(), r->fb.as<float4>(), r->npix, n, out);
        GPR_HIP(hipGetLastError());
        return GPR_OK;
    }
    DevBuf d;
    GPR_TRY(d.alloc(4ull * r->npix));
    hipLaunchKernelGGL(k_resolve_rgba8, dim3(grid), dim3(256), 0, stream(), r->fb.as<float4>(), r->npix, n,
                       d.as<uint32_t>());
    GPR_HIP(hipGetLastError());
    GPR_HIP(hipMemcpyAsync(out, d.p, 4ull * r->npix, hipMemcpyDeviceToHost, stream()));
    GPR_HIP(hipStreamSynchronize(stream()));
    return GPR_OK;
}

extern "C" int gpr_renderer_write_png(gpr_renderer* r, const char* path) {
    if (!r || !path) return fail(GPR_ERR_INVALID, "gpr_renderer_write_png: null argument");
    const uint32_t w = r->scene->width, h = r->scene->height;
    std::vector<float> rgb(3ull * r->npix);
    GPR_TRY(gpr_renderer_read(r, rgb.data(), nullptr));
    std::vector<unsigned char> px(3ull * r->npix);
    const float n = (float)std::max(1u, r->spp_done);
    // FrameBuffer::output (framebuffer.cpp:11-26): mean, gamma, x255 + 0.5,
    // clamp, truncate; rows flipped on write (imageio.cpp:65).
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const uint64_t src = (uint64_t)y * w + x, dst = (uint64_t)(h - 1 - y) * w + x;
            for (int c = 0; c < 3; ++c) px[3 * dst + c] = (unsigned char)srgb_byte(rgb[3 * src + c] / n);
        }
    return write_png(path, w, h, px);
}

// WriteImage (imageio.cpp:59-82) chosen by the file extension.
extern "C" int gpr_renderer_write_image(gpr_renderer* r, const char* path) {
    if (!r || !path) return fail(GPR_ERR_INVALID, "gpr_renderer_write_image: null argument");
    std::string p(path), ext;
    const size_t dot = p.find_last_of('.');
    if (dot != std::string::npos) ext = p.substr(dot);
    for (char& ch : ext) ch = (char)std::tolower((unsigned char)ch);
    if (ext == ".png") return gpr_renderer_write_png(r, path);
    if (ext != ".hdr") return fail(GPR_ERR_UNSUPPORTED, "gpr_renderer_write_image: extension must be .png or .hdr");
    const uint32_t w = r->scene->width, h = r->scene->height;
    std::vector<float> rgb(3ull * r->npix), mean(3ull * r->npix);
    GPR_TRY(gpr_renderer_read(r, rgb.data(), nullptr));
    const float n = (float)std::max(1u, r->spp_done);
    for (uint32_t y = 0; y < h; ++y)  // rows flipped on write, as for the PNG
        for (uint32_t x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c)
                mean[3 * ((uint64_t)(h - 1 - y) * w + x) + c] = rgb[3 * ((uint64_t)y * w + x) + c] / n;
    return write_hdr(path, w, h, mean);
}
