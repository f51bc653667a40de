// engine.h -- multi-job wavefront engine shared by the batched and served
// forms.  One pass renders one sample of every job it holds: the jobs' paths
// are concatenated in the ray queues (each entry carries its job id), scenes
// live in one device arena, and the pass is a fixed launch sequence replayed
// from a hipGraph.  In PCG mode a job's image does not depend on which other
// jobs share its passes: every path keeps its own RNG stream (job seed,
// sample, pixel) and each pixel receives its terms in the same kernel order as
// in the single-scene renderer.
#pragma once

#include <vector>

#include "gpr_internal.h"

namespace gpr {

constexpr uint32_t kMaxPassJobs = 1024;

// One job's slice of a pass (host copy; uploaded before the pass).
struct EngineJob {
    uint32_t scene;      // index into the engine's scene table
    uint32_t npix;       // width * height
    uint32_t width;
    uint32_t fb_off;     // first frame-buffer pixel of the job in the pool
    uint32_t sample;     // sample index this pass renders (0 clears the job's pixels first)
    uint32_t max_depth;
    uint32_t rr_start;
    uint32_t pad;
    unsigned long long seed;
    float r2c[16];
    float c2w[16];
};

// Traversal launch configuration (the autotuner's search space).
struct TraceCfg {
    int nt;      // threads per block of the extend / shadow kernels: 64, 128, 256, 512
    int sl;      // LDS stack entries per lane: 8, 16, 32
    int chunk;   // queue entries a wave claims per atomic: 64 .. 512
    int layout;  // 0 = BVH4 128-B nodes, 1 = reference DFS layout (BVH2)
    int refill;  // 1 = a lane takes a new queue entry as soon as its ray is done (BVH4 only)
};
std::vector<TraceCfg> engine_tune_space();

struct Engine;
// capacity = paths per pass (queue entries); max_jobs <= kMaxPassJobs
int engine_create(int device, uint32_t capacity, uint32_t max_jobs, Engine** out);
void engine_destroy(Engine* e);
size_t engine_device_bytes(const Engine* e);
// Scene table: device views into scene memory owned by the caller.
int engine_set_scenes(Engine* e, const std::vector<SceneView>& views, hipStream_t st);
// Stage the jobs (sum of npix <= capacity) and enqueue one pass on st.  fb is
// the frame-buffer pool the jobs' fb_off index.
int engine_run_pass(Engine* e, const EngineJob* jobs, uint32_t njobs, float4* fb, const TraceCfg& cfg,
                    hipStream_t st);
// Error flags accumulated by the passes so far (bit0 stack overflow, bit1
// look-back timeout); synchronises st.
int engine_errors(Engine* e, hipStream_t st, uint32_t* flags);

}  // namespace gpr
