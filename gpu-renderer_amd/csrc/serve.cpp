// serve.cpp -- batched and served forms on top of the multi-job engine:
// the scene arena, pass packing, the autotuner, and the request server with
// HBM-budget admission, one loop thread and latency statistics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "engine.h"
#include "gpr_internal.h"

namespace gpr {
namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// ---------------------------------------------------------------- arena --
// One device allocation holding every scene's triangle data, materials,
// lights, BVH4 nodes, triangle records and reference layout.
struct Arena {
    DevBuf mem;
    size_t used = 0;
    std::vector<SceneView> views;
};

// uvs, texture table and texels travel only with scenes that have textures
size_t texture_bytes(const gpr_scene* s) {
    if (s->textures.empty()) return 0;
    return align256(24 * s->mat.size()) + align256(sizeof(gpr_texture) * s->textures.size()) +
           align256(4 * s->texels.size());
}

size_t scene_bytes(const gpr_scene* s) {
    const size_t n = s->mat.size();
    return align256(36 * n) * 2 + align256(n) + align256(4 * n) + align256(sizeof(gpr_material) * s->mats.size()) +
           align256(4 * std::max<size_t>(1, s->lights.size())) + align256(sizeof(Bvh4Node) * s->bvh.nbvh4) +
           align256(sizeof(TriRec) * n) + align256(sizeof(gpr_ref_node) * (n ? 2 * n - 1 : 0)) +
           align256(sizeof(TriShade) * n) + texture_bytes(s);
}

int arena_add(Arena& A, const gpr_scene* s, hipStream_t st, uint32_t* id) {
    if (!s || !s->committed) return fail(GPR_ERR_INVALID, "add_scene: scene not committed");
    const size_t n = s->mat.size();
    if (n == 0 || s->bvh.n != n) return fail(GPR_ERR_INVALID, "add_scene: scene has no triangles or no BVH");
    // queue entries carry the scene index in 22 bits (scene << 10 | pass job)
    if (A.views.size() >= (1u << 22)) return fail(GPR_ERR_INVALID, "add_scene: more than 4,194,303 scenes");
    const size_t need = scene_bytes(s);
    if (A.used + need > A.mem.bytes)
        return fail(GPR_ERR_BUDGET, "add_scene: arena full (" + std::to_string(A.used + need) + " of " +
                                        std::to_string(A.mem.bytes) + " bytes)");
    char* base = A.mem.as<char>() + A.used;
    size_t off = 0;
    auto put = [&](const void* src, size_t bytes) -> char* {
        char* dst = base + off;
        off += align256(bytes);
        if (bytes && hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess) return nullptr;
        return dst;
    };
    SceneView v = s->view();
    SceneView a = v;
    a.pos = reinterpret_cast<const float*>(put(v.pos, 36 * n));
    a.nrm = reinterpret_cast<const float*>(put(v.nrm, 36 * n));
    a.flags = reinterpret_cast<const uint8_t*>(put(v.flags, n));
    a.mat = reinterpret_cast<const uint32_t*>(put(v.mat, 4 * n));
    a.mats = reinterpret_cast<const gpr_material*>(put(v.mats, sizeof(gpr_material) * s->mats.size()));
    a.lights = reinterpret_cast<const uint32_t*>(put(v.lights, 4 * s->lights.size()));
    if (s->lights.empty()) off += 256 - align256(0);
    a.bvh4 = reinterpret_cast<const Bvh4Node*>(put(v.bvh4, sizeof(Bvh4Node) * s->bvh.nbvh4));
    a.tris = reinterpret_cast<const TriRec*>(put(v.tris, sizeof(TriRec) * n));
    a.ref = reinterpret_cast<const gpr_ref_node*>(put(v.ref, sizeof(gpr_ref_node) * (2 * n - 1)));
    a.shade = reinterpret_cast<const TriShade*>(put(v.shade, sizeof(TriShade) * n));
    if (!a.pos || !a.nrm || !a.flags || !a.mat || !a.mats || !a.lights || !a.bvh4 || !a.tris || !a.ref ||
        !a.shade)
        return fail(GPR_ERR_HIP, "add_scene: arena copy failed");
    if (!s->textures.empty()) {
        a.uv = reinterpret_cast<const float*>(put(v.uv, 24 * n));
        a.tex = reinterpret_cast<const gpr_texture*>(put(v.tex, sizeof(gpr_texture) * s->textures.size()));
        a.texels = reinterpret_cast<const float4*>(put(v.texels, 4 * s->texels.size()));
        if (!a.uv || !a.tex || !a.texels) return fail(GPR_ERR_HIP, "add_scene: arena copy failed");
    }
    if (off > need) return fail(GPR_ERR_INVALID, "add_scene: arena layout mismatch");
    A.used += need;
    *id = (uint32_t)A.views.size();
    A.views.push_back(a);
    return GPR_OK;
}

int make_engine_job(const gpr_job& j, const std::vector<SceneView>& views, uint32_t capacity, EngineJob* out) {
    if (j.scene >= views.size()) return fail(GPR_ERR_INVALID, "job: unknown scene id");
    const uint64_t npix = (uint64_t)j.width * j.height;
    if (npix == 0 || npix > capacity) return fail(GPR_ERR_INVALID, "job: image larger than the pass capacity");
    if (j.spp == 0 || j.max_depth == 0 || j.max_depth > 64) return fail(GPR_ERR_INVALID, "job: spp or depth out of range");
    if (!(j.fov_deg > 0 && j.fov_deg < 180)) return fail(GPR_ERR_INVALID, "job: fov out of range");
    EngineJob e{};
    e.scene = j.scene;
    e.npix = (uint32_t)npix;
    e.width = j.width;
    e.max_depth = j.max_depth;
    e.rr_start = j.rr_start;
    e.seed = j.seed;
    camera_raster_to_camera(j.width, j.height, j.fov_deg, e.r2c);
    std::memcpy(e.c2w, j.c2w, sizeof(e.c2w));
    *out = e;
    return GPR_OK;
}

TraceCfg to_cfg(const gpr_trace_cfg& c) {
    return TraceCfg{(int)c.block, (int)c.stack, (int)c.chunk, (int)c.layout, c.refill ? 1 : 0};
}

bool cfg_valid(const TraceCfg& t) {
    bool ok = false;
    for (const TraceCfg& s : engine_tune_space())
        ok = ok || (s.nt == t.nt && s.sl == t.sl && s.layout == t.layout && s.refill == t.refill);
    return ok && t.chunk >= 64 && t.chunk % 64 == 0 && t.chunk <= 4096;
}

// Time one probe pass per configuration of the search space (after one
// untimed pass that captures its graph); results sorted fastest first.
int tune_engine(Engine* e, const std::vector<EngineJob>& pass, float4* fb, hipStream_t st,
                std::vector<gpr_trace_cfg>& res) {
    if (pass.empty()) return fail(GPR_ERR_INVALID, "tune: no probe jobs");
    hipEvent_t e0, e1;
    GPR_HIP(hipEventCreate(&e0));
    GPR_HIP(hipEventCreate(&e1));
    res.clear();
    int rc = GPR_OK;
    for (const TraceCfg& c : engine_tune_space()) {
        rc = engine_run_pass(e, pass.data(), (uint32_t)pass.size(), fb, c, st);
        if (rc == GPR_OK && hipEventRecord(e0, st) != hipSuccess) rc = fail(GPR_ERR_HIP, "tune: event record");
        for (int r = 0; r < 2 && rc == GPR_OK; ++r) rc = engine_run_pass(e, pass.data(), (uint32_t)pass.size(), fb, c, st);
        float ms = 0;
        if (rc == GPR_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                             hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = fail(GPR_ERR_HIP, "tune: event timing");
        if (rc != GPR_OK) break;
        res.push_back(gpr_trace_cfg{(uint32_t)c.nt, (uint32_t)c.sl, (uint32_t)c.chunk, (uint32_t)c.layout,
                                    (uint32_t)c.refill, ms / 2});
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    GPR_TRY(rc);
    uint32_t err = 0;
    GPR_TRY(engine_errors(e, st, &err));
    if (err) return fail(GPR_ERR_INVALID, "tune: error flags " + std::to_string(err));
    std::stable_sort(res.begin(), res.end(), [](const gpr_trace_cfg& a, const gpr_trace_cfg& c) { return a.ms < c.ms; });
    return GPR_OK;
}

}  // namespace
}  // namespace gpr

using namespace gpr;

// ================================================================= batch ==
struct gpr_batch {
    gpr_batch_opts opts{};
    int device = 0;
    Arena arena;
    Engine* eng = nullptr;
    std::vector<EngineJob> jobs;
    std::vector<uint32_t> spp;
    DevBuf fb;
    uint64_t fb_pixels = 0;
    bool fb_ready = false;
    TraceCfg cfg{256, 16, 64, 0, 0};
    gpr_batch_stats stats{};
    ~gpr_batch() {
        if (eng) engine_destroy(eng);
    }
};

namespace {
int batch_prepare(gpr_batch* b) {
    if (b->jobs.empty()) return fail(GPR_ERR_INVALID, "batch: no jobs");
    if (!b->fb_ready) {
        uint64_t total = 0;
        for (auto& j : b->jobs) {
            j.fb_off = (uint32_t)total;
            total += j.npix;
        }
        if (total >= (1ull << 31)) return fail(GPR_ERR_INVALID, "batch: more than 2^31 pixels in total");
        GPR_TRY(b->fb.alloc(16 * total));
        b->fb_pixels = total;
        GPR_TRY(engine_set_scenes(b->eng, b->arena.views, stream()));
        b->fb_ready = true;
    }
    return GPR_OK;
}

// Pack the next pass: jobs in order, each with samples left, while the
// capacity holds.  Returns the number of jobs packed.
uint32_t batch_pack(gpr_batch* b, std::vector<uint32_t>& issued, std::vector<EngineJob>& pass) {
    pass.clear();
    uint64_t used = 0;
    for (size_t k = 0; k < b->jobs.size() && pass.size() < b->opts.max_jobs; ++k) {
        if (issued[k] >= b->spp[k]) continue;
        if (used + b->jobs[k].npix > b->opts.capacity) continue;
        EngineJob j = b->jobs[k];
        j.sample = issued[k]++;
        used += j.npix;
        pass.push_back(j);
    }
    return (uint32_t)pass.size();
}
}  // namespace

extern "C" {

int gpr_scene_arena_bytes(const gpr_scene* s, uint64_t* bytes) {
    if (!s || !bytes || !s->committed) return fail(GPR_ERR_INVALID, "gpr_scene_arena_bytes: scene not committed");
    *bytes = scene_bytes(s);
    return GPR_OK;
}

int gpr_batch_create(const gpr_batch_opts* o, gpr_batch** out) {
    if (!o || !out) return fail(GPR_ERR_INVALID, "gpr_batch_create: null argument");
    if (o->max_jobs == 0 || o->max_jobs > kMaxPassJobs) return fail(GPR_ERR_INVALID, "gpr_batch_create: max_jobs out of range");
    std::unique_ptr<gpr_batch> b(new (std::nothrow) gpr_batch());
    if (!b) return fail(GPR_ERR_OOM, "gpr_batch_create: host allocation failed");
    b->opts = *o;
    b->device = device_id();
    GPR_HIP(hipSetDevice(b->device));
    GPR_TRY(b->arena.mem.alloc(std::max<uint64_t>(o->arena_bytes, 256)));
    GPR_TRY(engine_create(b->device, o->capacity, o->max_jobs, &b->eng));
    *out = b.release();
    return GPR_OK;
}

int gpr_batch_add_scene(gpr_batch* b, const gpr_scene* s, uint32_t* id) {
    if (!b || !id) return fail(GPR_ERR_INVALID, "gpr_batch_add_scene: null argument");
    if (b->fb_ready) return fail(GPR_ERR_INVALID, "gpr_batch_add_scene: batch already prepared");
    GPR_TRY(arena_add(b->arena, s, stream(), id));
    GPR_HIP(hipStreamSynchronize(stream()));
    return GPR_OK;
}

int gpr_batch_add_job(gpr_batch* b, const gpr_job* j, uint32_t* id) {
    if (!b || !j || !id) return fail(GPR_ERR_INVALID, "gpr_batch_add_job: null argument");
    if (b->fb_ready) return fail(GPR_ERR_INVALID, "gpr_batch_add_job: batch already prepared");
    EngineJob e;
    GPR_TRY(make_engine_job(*j, b->arena.views, b->opts.capacity, &e));
    *id = (uint32_t)b->jobs.size();
    b->jobs.push_back(e);
    b->spp.push_back(j->spp);
    return GPR_OK;
}

int gpr_batch_set_cfg(gpr_batch* b, const gpr_trace_cfg* c) {
    if (!b || !c) return fail(GPR_ERR_INVALID, "gpr_batch_set_cfg: null argument");
    const TraceCfg t = to_cfg(*c);
    if (!cfg_valid(t)) return fail(GPR_ERR_INVALID, "gpr_batch_set_cfg: configuration outside the tuned set");
    b->cfg = t;
    return GPR_OK;
}

int gpr_batch_tune(gpr_batch* b, gpr_trace_cfg* out, uint32_t cap, uint32_t* n) {
    if (!b) return fail(GPR_ERR_INVALID, "gpr_batch_tune: null batch");
    GPR_TRY(batch_prepare(b));
    // probe pass: the first pass of the batch at sample 0 (every job's first
    // real pass clears its pixels again, so the probes leave no trace)
    std::vector<uint32_t> issued(b->jobs.size(), 0);
    std::vector<EngineJob> pass;
    batch_pack(b, issued, pass);
    std::vector<gpr_trace_cfg> res;
    GPR_TRY(tune_engine(b->eng, pass, b->fb.as<float4>(), stream(), res));
    b->cfg = to_cfg(res[0]);
    if (n) *n = (uint32_t)std::min<size_t>(cap, res.size());
    if (out)
        for (uint32_t k = 0; k < cap && k < res.size(); ++k) out[k] = res[k];
    return GPR_OK;
}

int gpr_batch_run(gpr_batch* b) {
    if (!b) return fail(GPR_ERR_INVALID, "gpr_batch_run: null batch");
    GPR_TRY(batch_prepare(b));
    hipStream_t st = stream();
    GPR_HIP(hipStreamSynchronize(st));
    const double t0 = now_ms();
    std::vector<uint32_t> issued(b->jobs.size(), 0);
    std::vector<EngineJob> pass;
    uint64_t passes = 0, paths = 0;
    while (batch_pack(b, issued, pass) > 0) {
        GPR_TRY(engine_run_pass(b->eng, pass.data(), (uint32_t)pass.size(), b->fb.as<float4>(), b->cfg, st));
        ++passes;
        for (auto& j : pass) paths += j.npix;
    }
    uint32_t err = 0;
    GPR_TRY(engine_errors(b->eng, st, &err));
    b->stats.passes = passes;
    b->stats.paths = paths;
    b->stats.ms = now_ms() - t0;
    if (err) return fail(GPR_ERR_INVALID, "gpr_batch_run: error flags " + std::to_string(err));
    return GPR_OK;
}

int gpr_batch_read(gpr_batch* b, uint32_t job, float* rgb) {
    if (!b || !rgb || job >= b->jobs.size() || !b->fb_ready) return fail(GPR_ERR_INVALID, "gpr_batch_read: bad argument");
    const EngineJob& j = b->jobs[job];
    std::vector<float> tmp(4ull * j.npix);
    GPR_HIP(hipMemcpyAsync(tmp.data(), b->fb.as<float4>() + j.fb_off, 16ull * j.npix, hipMemcpyDeviceToHost, stream()));
    GPR_HIP(hipStreamSynchronize(stream()));
    for (uint64_t i = 0; i < j.npix; ++i) {
        rgb[3 * i + 0] = tmp[4 * i + 0];
        rgb[3 * i + 1] = tmp[4 * i + 1];
        rgb[3 * i + 2] = tmp[4 * i + 2];
    }
    return GPR_OK;
}

int gpr_batch_get_stats(const gpr_batch* b, gpr_batch_stats* out) {
    if (!b || !out) return fail(GPR_ERR_INVALID, "gpr_batch_get_stats: null argument");
    *out = b->stats;
    return GPR_OK;
}

int gpr_batch_free(gpr_batch* b) {
    if (b) {
        (void)hipStreamSynchronize(stream());
        delete b;
    }
    return GPR_OK;
}

}  // extern "C"

// ================================================================ server ==
namespace {
// First-fit allocator over the frame-buffer pool (units: pixels).
struct PoolAlloc {
    std::map<uint64_t, uint64_t> free_;  // offset -> length
    void init(uint64_t n) {
        free_.clear();
        if (n) free_[0] = n;
    }
    bool take(uint64_t n, uint64_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const uint64_t rest = it->second - n, at = it->first + n;
                free_.erase(it);
                if (rest) free_[at] = rest;
                return true;
            }
        }
        return false;
    }
    void give(uint64_t off, uint64_t n) {
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Request {
    uint64_t ticket = 0;
    gpr_job job{};
    EngineJob ej{};
    uint32_t issued = 0;
    int state = 0;  // 0 pending, 1 active, 2 done, 3 failed, 4 cancelled
    bool cancel = false;    // cancellation requested while active
    bool retiring = false;  // no more samples; the range is released after the passes in flight
    double t_submit = 0, t_done = 0;
    std::vector<float> rgb;
};
}  // namespace

struct gpr_server {
    gpr_server_opts opts{};
    int device = 0;
    Arena arena;
    Engine* eng = nullptr;
    DevBuf fb;
    PoolAlloc pool;
    hipStream_t st = nullptr, cst = nullptr;
    TraceCfg cfg{256, 16, 64, 0, 0};
    std::mutex mu;
    std::condition_variable cv;       // loop thread: new work / stop
    std::condition_variable done_cv;  // waiters: a request finished
    std::deque<Request*> pending;
    std::vector<Request*> active;
    std::map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_ticket = 1;
    bool stop = false;
    bool scenes_dirty = false;
    bool in_pass = false;  // the loop is launching a pass outside the lock
    bool cancel_req = false;  // an active request asked to be cancelled
    int loop_error = GPR_OK;
    std::string loop_msg;
    std::thread loop;
    // statistics
    uint64_t submitted = 0, completed = 0, passes = 0, jobs_in_passes = 0, cancelled = 0;
    double t_first_submit = 0, t_last_done = 0;
    std::vector<double> latencies;
    ~gpr_server() {
        if (eng) engine_destroy(eng);
        if (st) (void)hipStreamDestroy(st);
        if (cst) (void)hipStreamDestroy(cst);
    }
};

namespace {
struct Finishing {
    hipEvent_t ev;
    std::vector<Request*> reqs;
    bool cancelled = false;  // retire the requests (no result) instead of copying them out
};

// Admission: FIFO, a request enters when its frame buffer fits in the pool
// and a job slot is free.
void admit(gpr_server* s) {
    while (!s->pending.empty() && s->active.size() < s->opts.max_jobs) {
        Request* r = s->pending.front();
        uint64_t off = 0;
        if (!s->pool.take(r->ej.npix, &off)) break;
        r->ej.fb_off = (uint32_t)off;
        r->state = 1;
        s->pending.pop_front();
        s->active.push_back(r);
    }
}

void loop_fail(gpr_server* s, const std::string& msg) {
    std::lock_guard<std::mutex> g(s->mu);
    s->loop_error = GPR_ERR_HIP;
    s->loop_msg = msg;
    for (Request* r : s->active) r->state = 3;
    for (Request* r : s->pending) r->state = 3;
    s->active.clear();
    s->pending.clear();
    s->done_cv.notify_all();
}

// Copy the finished requests' frame buffers out, release their pool ranges.
bool drain(gpr_server* s, Finishing& f) {
    if (hipStreamWaitEvent(s->cst, f.ev, 0) != hipSuccess) return false;
    if (f.cancelled) {
        // the passes that could still write into these ranges have completed
        if (hipStreamSynchronize(s->cst) != hipSuccess) return false;
        std::lock_guard<std::mutex> g(s->mu);
        for (Request* r : f.reqs) {
            s->pool.give(r->ej.fb_off, r->ej.npix);
            r->state = 4;
            s->cancelled++;
            s->active.erase(std::remove(s->active.begin(), s->active.end(), r), s->active.end());
        }
        (void)hipEventDestroy(f.ev);
        s->done_cv.notify_all();
        s->cv.notify_all();
        return true;
    }
    std::vector<float> tmp;
    for (Request* r : f.reqs) {
        tmp.resize(4ull * r->ej.npix);
        if (hipMemcpyAsync(tmp.data(), s->fb.as<float4>() + r->ej.fb_off, 16ull * r->ej.npix, hipMemcpyDeviceToHost,
                           s->cst) != hipSuccess ||
            hipStreamSynchronize(s->cst) != hipSuccess)
            return false;
        r->rgb.resize(3ull * r->ej.npix);
        for (uint64_t i = 0; i < r->ej.npix; ++i) {
            r->rgb[3 * i + 0] = tmp[4 * i + 0];
            r->rgb[3 * i + 1] = tmp[4 * i + 1];
            r->rgb[3 * i + 2] = tmp[4 * i + 2];
        }
        std::lock_guard<std::mutex> g(s->mu);
        s->pool.give(r->ej.fb_off, r->ej.npix);
        r->t_done = now_ms();
        r->state = 2;
        s->completed++;
        s->t_last_done = r->t_done;
        s->latencies.push_back(r->t_done - r->t_submit);
        s->active.erase(std::remove(s->active.begin(), s->active.end(), r), s->active.end());
    }
    (void)hipEventDestroy(f.ev);
    s->done_cv.notify_all();
    s->cv.notify_all();
    return true;
}

void server_loop(gpr_server* s) {
    if (hipSetDevice(s->device) != hipSuccess) {
        loop_fail(s, "server: hipSetDevice failed");
        return;
    }
    std::deque<Finishing> finishing;
    std::vector<EngineJob> pass;
    std::vector<Request*> pass_reqs;
    std::vector<Request*> retire;
    while (true) {
        TraceCfg cfg;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                admit(s);
                bool work = false;
                for (Request* r : s->active) work = work || r->issued < r->job.spp;
                if (work || !finishing.empty() || s->scenes_dirty || s->cancel_req) break;
                if (s->stop && s->pending.empty() && s->active.empty()) return;
                s->cv.wait(lk);
            }
            if (s->scenes_dirty) {
                if (engine_set_scenes(s->eng, s->arena.views, s->st) != GPR_OK) {
                    lk.unlock();
                    loop_fail(s, "server: scene table upload failed");
                    return;
                }
                s->scenes_dirty = false;
            }
            cfg = s->cfg;
            // cancelled requests still owed samples stop here; the ones whose
            // last sample is already issued finish normally
            retire.clear();
            for (Request* r : s->active)
                if (r->cancel && !r->retiring && r->issued < r->job.spp) {
                    r->retiring = true;
                    r->issued = r->job.spp;
                    retire.push_back(r);
                }
            s->cancel_req = false;
            // pack: active requests in admission order, one sample each
            pass.clear();
            pass_reqs.clear();
            uint64_t used = 0;
            for (Request* r : s->active) {
                if (r->issued >= r->job.spp || used + r->ej.npix > s->opts.capacity) continue;
                EngineJob j = r->ej;
                j.sample = r->issued++;
                used += j.npix;
                pass.push_back(j);
                pass_reqs.push_back(r);
            }
            s->in_pass = !pass.empty();
        }
        if (!pass.empty()) {
            const int prc = engine_run_pass(s->eng, pass.data(), (uint32_t)pass.size(), s->fb.as<float4>(), cfg, s->st);
            {
                std::lock_guard<std::mutex> g(s->mu);
                s->in_pass = false;
            }
            s->cv.notify_all();
            if (prc != GPR_OK) {
                loop_fail(s, "server: pass launch failed: " + std::string(gpr_last_error()));
                return;
            }
            Finishing f{};
            for (Request* r : pass_reqs)
                if (r->issued == r->job.spp) f.reqs.push_back(r);
            {
                std::lock_guard<std::mutex> g(s->mu);
                s->passes++;
                s->jobs_in_passes += pass.size();
            }
            if (!f.reqs.empty()) {
                if (hipEventCreateWithFlags(&f.ev, hipEventDisableTiming) != hipSuccess ||
                    hipEventRecord(f.ev, s->st) != hipSuccess) {
                    loop_fail(s, "server: event failed");
                    return;
                }
                finishing.push_back(std::move(f));
            }
        }
        if (!retire.empty()) {
            // released once every pass launched so far (the last one that may
            // have included them) has completed
            Finishing f{};
            f.reqs = retire;
            f.cancelled = true;
            if (hipEventCreateWithFlags(&f.ev, hipEventDisableTiming) != hipSuccess || hipEventRecord(f.ev, s->st) != hipSuccess) {
                loop_fail(s, "server: event failed");
                return;
            }
            finishing.push_back(std::move(f));
        }
        // keep one pass queued ahead of the completions being drained
        while (finishing.size() > (pass.empty() ? 0u : 1u)) {
            if (!drain(s, finishing.front())) {
                loop_fail(s, "server: result copy failed");
                return;
            }
            finishing.pop_front();
        }
    }
}
}  // namespace

extern "C" {

int gpr_server_create(const gpr_server_opts* o, gpr_server** out) {
    if (!o || !out) return fail(GPR_ERR_INVALID, "gpr_server_create: null argument");
    if (o->max_jobs == 0 || o->max_jobs > kMaxPassJobs) return fail(GPR_ERR_INVALID, "gpr_server_create: max_jobs out of range");
    std::unique_ptr<gpr_server> s(new (std::nothrow) gpr_server());
    if (!s) return fail(GPR_ERR_OOM, "gpr_server_create: host allocation failed");
    s->opts = *o;
    s->device = device_id();
    GPR_HIP(hipSetDevice(s->device));
    GPR_HIP(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
    GPR_HIP(hipStreamCreateWithFlags(&s->cst, hipStreamNonBlocking));
    GPR_TRY(engine_create(s->device, o->capacity, o->max_jobs, &s->eng));
    const uint64_t fixed = o->arena_bytes + engine_device_bytes(s->eng);
    if (o->hbm_budget <= fixed + 16ull * o->capacity)
        return fail(GPR_ERR_BUDGET, "gpr_server_create: budget below arena + queues + one full frame buffer");
    const uint64_t fb_pixels = std::min<uint64_t>((o->hbm_budget - fixed) / 16, (1ull << 31) - 1);
    GPR_TRY(s->arena.mem.alloc(std::max<uint64_t>(o->arena_bytes, 256)));
    GPR_TRY(s->fb.alloc(16 * fb_pixels));
    s->pool.init(fb_pixels);
    gpr_server* raw = s.get();
    try {
        raw->loop = std::thread(server_loop, raw);
    } catch (...) {
        return fail(GPR_ERR_OOM, "gpr_server_create: cannot start the loop thread");
    }
    *out = s.release();
    return GPR_OK;
}

int gpr_server_add_scene(gpr_server* s, const gpr_scene* scene, uint32_t* id) {
    if (!s || !id) return fail(GPR_ERR_INVALID, "gpr_server_add_scene: null argument");
    std::lock_guard<std::mutex> g(s->mu);
    hipStream_t st = stream();
    GPR_TRY(arena_add(s->arena, scene, st, id));
    GPR_HIP(hipStreamSynchronize(st));
    s->scenes_dirty = true;
    s->cv.notify_all();
    return GPR_OK;
}

int gpr_server_submit(gpr_server* s, const gpr_job* j, uint64_t* ticket) {
    if (!s || !j || !ticket) return fail(GPR_ERR_INVALID, "gpr_server_submit: null argument");
    std::unique_ptr<Request> r(new (std::nothrow) Request());
    if (!r) return fail(GPR_ERR_OOM, "gpr_server_submit: host allocation failed");
    std::lock_guard<std::mutex> g(s->mu);
    if (s->loop_error != GPR_OK) return fail(s->loop_error, s->loop_msg);
    GPR_TRY(make_engine_job(*j, s->arena.views, s->opts.capacity, &r->ej));
    if (16ull * r->ej.npix > s->fb.bytes) return fail(GPR_ERR_BUDGET, "gpr_server_submit: frame buffer larger than the pool");
    r->job = *j;
    r->ticket = s->next_ticket++;
    r->t_submit = now_ms();
    if (s->submitted == 0) s->t_first_submit = r->t_submit;
    s->submitted++;
    *ticket = r->ticket;
    s->pending.push_back(r.get());
    s->reqs[r->ticket] = std::move(r);
    s->cv.notify_all();
    return GPR_OK;
}

int gpr_server_wait(gpr_server* s, uint64_t ticket, float* rgb, double* latency_ms) {
    if (!s) return fail(GPR_ERR_INVALID, "gpr_server_wait: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(GPR_ERR_INVALID, "gpr_server_wait: unknown ticket");
    Request* r = it->second.get();
    s->done_cv.wait(lk, [&] { return r->state >= 2; });
    if (r->state == 3) {
        s->reqs.erase(it);
        return fail(GPR_ERR_HIP, "gpr_server_wait: " + s->loop_msg);
    }
    if (r->state == 4) {
        s->reqs.erase(it);
        return fail(GPR_ERR_CANCELLED, "gpr_server_wait: request " + std::to_string(ticket) + " was cancelled");
    }
    if (rgb) std::memcpy(rgb, r->rgb.data(), sizeof(float) * r->rgb.size());
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    s->reqs.erase(it);
    return GPR_OK;
}

int gpr_server_cancel(gpr_server* s, uint64_t ticket) {
    if (!s) return fail(GPR_ERR_INVALID, "gpr_server_cancel: null server");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(GPR_ERR_INVALID, "gpr_server_cancel: unknown ticket");
    Request* r = it->second.get();
    if (r->state == 0) {
        s->pending.erase(std::remove(s->pending.begin(), s->pending.end(), r), s->pending.end());
        r->state = 4;
        s->cancelled++;
        s->done_cv.notify_all();
        s->cv.notify_all();  // the queue head may have changed
    } else if (r->state == 1 && !r->cancel) {
        r->cancel = true;
        s->cancel_req = true;
        s->cv.notify_all();
    }
    return GPR_OK;
}

int gpr_server_set_cfg(gpr_server* s, const gpr_trace_cfg* c) {
    if (!s || !c) return fail(GPR_ERR_INVALID, "gpr_server_set_cfg: null argument");
    const TraceCfg t = to_cfg(*c);
    if (!cfg_valid(t)) return fail(GPR_ERR_INVALID, "gpr_server_set_cfg: configuration outside the tuned set");
    std::lock_guard<std::mutex> g(s->mu);
    s->cfg = t;
    return GPR_OK;
}

int gpr_server_tune(gpr_server* s, const gpr_job* jobs, uint32_t njobs, gpr_trace_cfg* out, uint32_t cap,
                    uint32_t* n) {
    if (!s || !jobs || njobs == 0) return fail(GPR_ERR_INVALID, "gpr_server_tune: no probe jobs");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv.wait(lk, [&] { return !s->in_pass; });
    if (!s->pending.empty() || !s->active.empty()) return fail(GPR_ERR_INVALID, "gpr_server_tune: server is busy");
    if (s->scenes_dirty) {
        GPR_TRY(engine_set_scenes(s->eng, s->arena.views, s->st));
        s->scenes_dirty = false;
    }
    std::vector<EngineJob> pass;
    uint64_t used = 0;
    int rc = GPR_OK;
    for (uint32_t k = 0; k < njobs && rc == GPR_OK; ++k) {
        EngineJob e;
        rc = make_engine_job(jobs[k], s->arena.views, s->opts.capacity, &e);
        if (rc != GPR_OK || used + e.npix > s->opts.capacity || pass.size() >= s->opts.max_jobs) break;
        uint64_t off = 0;
        if (!s->pool.take(e.npix, &off)) break;
        e.fb_off = (uint32_t)off;
        used += e.npix;
        pass.push_back(e);
    }
    std::vector<gpr_trace_cfg> res;
    if (rc == GPR_OK) rc = tune_engine(s->eng, pass, s->fb.as<float4>(), s->st, res);
    for (const EngineJob& e : pass) s->pool.give(e.fb_off, e.npix);
    GPR_TRY(rc);
    s->cfg = to_cfg(res[0]);
    if (n) *n = (uint32_t)std::min<size_t>(cap, res.size());
    if (out)
        for (uint32_t k = 0; k < cap && k < res.size(); ++k) out[k] = res[k];
    return GPR_OK;
}

int gpr_server_get_stats(gpr_server* s, gpr_server_stats* out) {
    if (!s || !out) return fail(GPR_ERR_INVALID, "gpr_server_get_stats: null argument");
    std::lock_guard<std::mutex> g(s->mu);
    gpr_server_stats o{};
    o.submitted = s->submitted;
    o.completed = s->completed;
    o.passes = s->passes;
    o.mean_jobs_per_pass = s->passes ? (double)s->jobs_in_passes / s->passes : 0.0;
    o.cancelled = s->cancelled;
    if (s->completed) {
        const double span = s->t_last_done - s->t_first_submit;
        o.renders_per_s = span > 0 ? 1000.0 * s->completed / span : 0.0;
        std::vector<double> l = s->latencies;
        std::sort(l.begin(), l.end());
        auto pct = [&](double p) {
            const size_t k = (size_t)std::min<double>(l.size() - 1, std::ceil(p * l.size()) - 1);
            return l[k];
        };
        o.p50_ms = pct(0.50);
        o.p99_ms = pct(0.99);
    }
    *out = o;
    return GPR_OK;
}

int gpr_server_free(gpr_server* s) {
    if (!s) return GPR_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->loop.joinable()) s->loop.join();
    (void)hipStreamSynchronize(s->st);
    delete s;
    return GPR_OK;
}

}  // extern "C"
