// trace.hip -- ray-query kernels behind gpr_trace (trace / trace_shadow,
// src/renderer/pathtracer_kernel.h:235-246, 310-321).
//
// k_trace_fast is persistent: every wave pulls 64 consecutive rays from a
// global atomic cursor until the batch is drained, so the grid is sized to
// the device (not to the batch) and the per-lane overflow stacks are indexed
// by the persistent thread id.
#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <type_traits>
#include <utility>

#include "traverse.h"

namespace gpr {
namespace {

__global__ void k_zero_u64(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

__global__ void k_add_overflow(const unsigned long long* cur, unsigned long long* acc) {
    if (threadIdx.x == 0) acc[0] += cur[1];
}

GPR_D Ray3 load_ray(const gpr_ray* rays, uint64_t i) {
    const float4* p = reinterpret_cast<const float4*>(rays + i);
    const float4 a = p[0], b = p[1];
    Ray3 r;
    r.o = V(a.x, a.y, a.z);
    r.d = V(a.w, b.x, b.y);
    r.tmin = b.z;
    r.tmax = b.w;
    return r;
}
GPR_D void store_hit(gpr_hit* hits, uint64_t i, const gpr_hit& h) {
    *reinterpret_cast<float4*>(hits + i) = make_float4(h.t, __int_as_float(h.id), h.u, h.v);
}

template <bool ANY, int NT, int SL>
__global__ __launch_bounds__(NT) void k_trace_fast(SceneView sv, const gpr_ray* __restrict__ rays,
                                                   gpr_hit* __restrict__ hits, uint64_t n,
                                                   unsigned long long* __restrict__ cursor,
                                                   int2* __restrict__ spill, uint32_t chunk) {
    __shared__ int2 lds[SL * NT];
    LaneStackT<NT, SL> st;
    st.lds = lds + threadIdx.x;
    st.spill = spill + (uint64_t)(blockIdx.x * NT + threadIdx.x) * kStackGlobal;
    st.sp = 0;
    const int lane = threadIdx.x & 63;
    while (true) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(cursor, (unsigned long long)chunk);
        base = __shfl(base, 0, 64);
        if (base >= n) break;
        for (uint32_t k = 0; k < chunk; k += 64) {
            const uint64_t i = base + k + lane;
            if (i < n) {
                const Ray3 r = load_ray(rays, i);
                gpr_hit h;
                if (!fast_trace<ANY>(sv, r, st, h)) atomicAdd(cursor + 1, 1ull);  // stack overflow count
                store_hit(hits, i, h);
            }
        }
    }
}

// Per-lane ray refill: the lanes of a wave step their rays one node at a
// time and a lane whose ray is done takes the next one from a wave-local pool
// of `chunk` ray indices (one global atomic per chunk), so no lane idles while
// the slowest ray of the wave finishes.  A closest-hit ray takes the same
// node steps as in k_trace_fast, so the hits are identical.  Any hit parks
// leaf children and tests them in wave-wide leaf rounds (TravD): the
// occlusion answer is the same, the occluder reported may differ.
constexpr int kLeafRound = 32;  // parked lanes of a wave that start an any-hit leaf round

template <bool ANY, int NT, int SL>
__global__ __launch_bounds__(NT) void k_trace_refill(SceneView sv, const gpr_ray* __restrict__ rays,
                                                     gpr_hit* __restrict__ hits, uint64_t n,
                                                     unsigned long long* __restrict__ cursor,
                                                     int2* __restrict__ spill, uint32_t chunk) {
    __shared__ int2 lds[SL * NT];
    LaneStackT<NT, SL> st;
    st.lds = lds + threadIdx.x;
    st.spill = spill + (uint64_t)(blockIdx.x * NT + threadIdx.x) * kStackGlobal;
    st.sp = 0;
    const int lane = threadIdx.x & 63;
    uint64_t pb = 0, pe = 0;  // wave-local pool [pb, pe) of ray indices (wave-uniform)
    bool drained = false;     // the global cursor has passed n (wave-uniform)
    bool active = false;
    uint64_t ri = 0;
    // any hit walks with postponed leaves (leaf rounds once kLeafRound lanes have parked some)
    std::conditional_t<ANY, TravD, Trav> t;
    while (true) {
        const unsigned long long need = __ballot(!active);
        if (need != 0ull && !(drained && pb == pe)) {
            const uint32_t cnt = (uint32_t)__popcll(need);
            const uint32_t rank =
                __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
            const uint64_t avail = pe - pb;
            uint64_t nb = pe, ne = pe;
            if (avail < cnt && !drained) {
                unsigned long long b = 0;
                if (lane == 0) b = atomicAdd(cursor, (unsigned long long)chunk);
                b = __shfl(b, 0, 64);
                nb = b < n ? b : n;
                ne = b + chunk < n ? b + chunk : n;
                if (b + chunk >= n) drained = true;
            }
            if (!active) {
                const bool from_old = rank < avail;
                const uint64_t idx = from_old ? pb + rank : nb + (rank - avail);
                if (from_old || idx < ne) {
                    ri = idx;
                    const Ray3 r = load_ray(rays, ri);
                    bool done;
                    if constexpr (ANY) done = travd_begin<ANY>(sv, r, st, t);
                    else done = trav_begin<ANY>(sv, r, st, t);
                    if (done) store_hit(hits, ri, t.hit);
                    else active = true;
                }
            }
            if (avail >= cnt) {
                pb += cnt;
            } else {
                const uint64_t taken = nb + (cnt - avail);
                pb = taken < ne ? taken : ne;
                pe = ne;
            }
        }
        if (__ballot(active) == 0ull) {
            if (drained && pb == pe) break;
            continue;
        }
        if constexpr (ANY) {
            if (active && t.pmask == 0) {
                const int rc = travd_step<ANY>(sv, st, t);
                if (rc != kTravMore) {
                    if (rc == kTravOverflow) atomicAdd(cursor + 1, 1ull);  // stack overflow count
                    store_hit(hits, ri, t.hit);
                    active = false;
                }
            }
            // leaf round: when enough lanes have parked leaves, or no lane can step
            const unsigned long long parked = __ballot(active && t.pmask != 0);
            if (parked != 0ull && (__popcll(parked) >= kLeafRound || __ballot(active && t.pmask == 0) == 0ull)) {
                if (active && t.pmask != 0 && travd_leaves<ANY>(sv, t)) {
                    store_hit(hits, ri, t.hit);
                    active = false;
                }
            }
        } else {
            if (active) {
                const int rc = trav_step<ANY>(sv, st, t);
                if (rc != kTravMore) {
                    if (rc == kTravOverflow) atomicAdd(cursor + 1, 1ull);  // stack overflow count
                    store_hit(hits, ri, t.hit);
                    active = false;
                }
            }
        }
    }
}

template <bool ANY>
__global__ __launch_bounds__(256) void k_trace_ref(SceneView sv, const gpr_ray* __restrict__ rays,
                                                   gpr_hit* __restrict__ hits, uint64_t n,
                                                   unsigned long long* __restrict__ stats) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t nv = 0, nt = 0;
    if (i < n) {
        const Ray3 r = load_ray(rays, i);
        const gpr_hit h = ref_trace<ANY>(sv, r, &nv, &nt);
        store_hit(hits, i, h);
    }
    // block-reduce the counters, one atomic per wave
    unsigned long long v = nv, t = nt;
    for (int o = 32; o > 0; o >>= 1) {
        v += __shfl_down(v, o, 64);
        t += __shfl_down(t, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(stats + 0, v);
        atomicAdd(stats + 1, t);
    }
}

// Node fetches and triangle tests of the BVH4 walk, summed over a batch of
// rays (the immediate-leaf walk of trav_step, which is what the closest-hit
// kernels run for every ray).  Persistent grid-stride loop over the rays, so
// the spill area of the (256, 16) configuration bounds the lanes.
template <bool ANY>
__global__ __launch_bounds__(256) void k_trace_stats(SceneView sv, const gpr_ray* __restrict__ rays, uint64_t n,
                                                     unsigned long long* __restrict__ stats,
                                                     int2* __restrict__ spill) {
    __shared__ int2 lds[16 * 256];
    LaneStackT<256, 16> st;
    st.lds = lds + threadIdx.x;
    st.spill = spill + (uint64_t)(blockIdx.x * 256 + threadIdx.x) * kStackGlobal;
    st.sp = 0;
    unsigned long long nf = 0, nt = 0, ov = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        Trav t;
        int rc = trav_begin<ANY>(sv, load_ray(rays, i), st, t) ? kTravDone : kTravMore;
        while (rc == kTravMore) rc = trav_step<ANY>(sv, st, t);
        nf += t.fetched;
        nt += t.tested + (sv.root4 < 0 ? 1u : 0u);
        ov += rc == kTravOverflow ? 1u : 0u;
    }
    for (int o = 32; o > 0; o >>= 1) {
        nf += __shfl_down(nf, o, 64);
        nt += __shfl_down(nt, o, 64);
        ov += __shfl_down(ov, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(stats + 0, nf);
        atomicAdd(stats + 1, nt);
        atomicAdd(stats + 2, ov);
    }
}

using TraceKernel = void (*)(SceneView, const gpr_ray*, gpr_hit*, uint64_t, unsigned long long*, int2*, uint32_t);
struct TraceVariant {
    TraceKernel closest, any;          // wave-synchronous rays
    TraceKernel closest_r, any_r;      // per-lane ray refill
    int nt, sl;
};
template <int NT, int SL>
TraceVariant trace_variant() {
    return {k_trace_fast<false, NT, SL>, k_trace_fast<true, NT, SL>, k_trace_refill<false, NT, SL>,
            k_trace_refill<true, NT, SL>, NT, SL};
}
constexpr int kTraceVariants = 12;
const TraceVariant& trace_variant_at(int k) {
    static const TraceVariant t[kTraceVariants] = {
        trace_variant<64, 8>(),  trace_variant<64, 16>(),  trace_variant<64, 32>(),
        trace_variant<128, 8>(), trace_variant<128, 16>(), trace_variant<128, 32>(),
        trace_variant<256, 8>(), trace_variant<256, 16>(), trace_variant<256, 32>(),
        trace_variant<512, 8>(), trace_variant<512, 16>(), trace_variant<512, 32>(),
    };
    return t[k];
}
int trace_variant_index(int nt, int sl) {
    const int a = nt == 64 ? 0 : nt == 128 ? 1 : nt == 256 ? 2 : nt == 512 ? 3 : -1;
    const int b = sl == 8 ? 0 : sl == 16 ? 1 : sl == 32 ? 2 : -1;
    return (a < 0 || b < 0) ? -1 : 3 * a + b;
}

struct TraceResources {
    DevBuf spill;
    DevBuf cursor;
    int grid[kTraceVariants][4] = {};  // [variant][closest, any, closest refill, any refill]
    int dev = -1;
};
// Work buffers of the fast traces, one set per (device, stream): launches on
// one stream run in order and may share a ray cursor and a stack spill area;
// launches on different streams may overlap, so each stream has its own.  The
// sets live until the process ends.
std::mutex g_res_mu;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<TraceResources>> g_res;

TraceResources* find_resources(hipStream_t st) {
    std::lock_guard<std::mutex> lk(g_res_mu);
    auto it = g_res.find({device_id(), st});
    return it == g_res.end() ? nullptr : it->second.get();
}

int ensure_resources(TraceResources& R) {
    const int dev = device_id();
    if (R.dev == dev) return GPR_OK;
    int cus = 0;
    GPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    size_t threads = 0;
    for (int k = 0; k < kTraceVariants; ++k) {
        const TraceVariant& v = trace_variant_at(k);
        const TraceKernel ks[4] = {v.closest, v.any, v.closest_r, v.any_r};
        for (int a = 0; a < 4; ++a) {
            int per_cu = 0;
            GPR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ks[a], v.nt, 0));
            R.grid[k][a] = cus * std::max(1, per_cu);
            threads = std::max(threads, (size_t)R.grid[k][a] * v.nt);
        }
    }
    GPR_TRY(R.spill.alloc(threads * kStackGlobal * sizeof(int2)));
    GPR_TRY(R.cursor.alloc(4 * sizeof(unsigned long long)));
    R.dev = dev;
    return GPR_OK;
}

int resources_for(hipStream_t st, TraceResources** out) {
    std::lock_guard<std::mutex> lk(g_res_mu);
    std::unique_ptr<TraceResources>& p = g_res[{device_id(), st}];
    if (!p) p = std::make_unique<TraceResources>();
    GPR_TRY(ensure_resources(*p));
    *out = p.get();
    return GPR_OK;
}

}  // namespace

bool trace_cfg_valid(int nt, int sl, int chunk) {
    return trace_variant_index(nt, sl) >= 0 && chunk >= 64 && chunk % 64 == 0 && chunk <= 4096;
}

int launch_trace(const SceneView& sv, const gpr_ray* rays, gpr_hit* hits, uint64_t n, uint32_t flags,
                 unsigned long long* stats, hipStream_t st, int nt, int sl, int chunk, int refill) {
    if (n == 0) return GPR_OK;
    const bool any = (flags & GPR_TRACE_ANY) != 0;
    if (flags & GPR_TRACE_REF_ORDER) {
        if (!stats) return fail(GPR_ERR_INVALID, "ref-order trace needs a stats buffer");
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, stats, 2);
        const dim3 grid((unsigned)((n + 255) / 256));
        if (any) hipLaunchKernelGGL(k_trace_ref<true>, grid, dim3(256), 0, st, sv, rays, hits, n, stats);
        else hipLaunchKernelGGL(k_trace_ref<false>, grid, dim3(256), 0, st, sv, rays, hits, n, stats);
        GPR_HIP(hipGetLastError());
        return GPR_OK;
    }
    const int k = trace_variant_index(nt, sl);
    if (k < 0 || !trace_cfg_valid(nt, sl, chunk)) return fail(GPR_ERR_INVALID, "trace: configuration outside the tuned set");
    TraceResources* rp = nullptr;
    GPR_TRY(resources_for(st, &rp));
    TraceResources& R = *rp;
    const TraceVariant& v = trace_variant_at(k);
    unsigned long long* cur = R.cursor.as<unsigned long long>();
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, cur, 4);
    const uint64_t want = (n + v.nt - 1) / v.nt;  // more blocks than rays would only idle
    const int kind = (refill ? 2 : 0) + (any ? 1 : 0);
    const TraceKernel ks[4] = {v.closest, v.any, v.closest_r, v.any_r};
    const dim3 grid((unsigned)std::min<uint64_t>(R.grid[k][kind], want));
    hipLaunchKernelGGL(ks[kind], grid, dim3(v.nt), 0, st, sv, rays, hits, n, cur, R.spill.as<int2>(),
                       (uint32_t)chunk);
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

// BVH4 node fetches, triangle tests and overflows summed over n device rays
// (stats: 3 device words, cleared here); the per-ray node sequence is the one
// every fast closest-hit configuration walks.
int launch_trace_stats(const SceneView& sv, const gpr_ray* rays, uint64_t n, bool any, unsigned long long* stats,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, stats, 3);
    if (n == 0) return GPR_OK;
    TraceResources* rp = nullptr;
    GPR_TRY(resources_for(st, &rp));
    const int k = trace_variant_index(256, 16);
    const uint64_t want = (n + 255) / 256;
    const dim3 grid((unsigned)std::min<uint64_t>(rp->grid[k][0], want));
    if (any) hipLaunchKernelGGL(k_trace_stats<true>, grid, dim3(256), 0, st, sv, rays, n, stats, rp->spill.as<int2>());
    else hipLaunchKernelGGL(k_trace_stats<false>, grid, dim3(256), 0, st, sv, rays, n, stats, rp->spill.as<int2>());
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

// Releases the fast-trace work buffers of (this thread's device, st) after
// the stream's work has finished; the next trace on st allocates them again.
int trace_release_stream(hipStream_t st) {
    std::unique_ptr<TraceResources> victim;
    {
        std::lock_guard<std::mutex> lk(g_res_mu);
        auto it = g_res.find({device_id(), st});
        if (it == g_res.end()) return GPR_OK;
        victim = std::move(it->second);
        g_res.erase(it);
    }
    return GPR_OK;  // the buffers are freed as victim goes out of scope
}

// Overflow count of the last fast trace on this thread's device and stream
// (0 = every ray completed).
int trace_overflow_count(unsigned long long* out) {
    *out = 0;
    const TraceResources* R = find_resources(stream());
    if (!R || R->dev < 0) return GPR_OK;
    GPR_HIP(hipMemcpyAsync(out, R->cursor.as<unsigned long long>() + 1, 8, hipMemcpyDeviceToHost, stream()));
    GPR_HIP(hipStreamSynchronize(stream()));
    return GPR_OK;
}

// Adds the overflow count of the last fast trace launched on st to *acc
// (device memory), in stream order, so that a batch traced in several launches
// keeps the sum over its launches.
int trace_overflow_accumulate(unsigned long long* acc, hipStream_t st) {
    const TraceResources* R = find_resources(st);
    if (!R || R->dev < 0) return fail(GPR_ERR_INVALID, "trace_overflow_accumulate: no fast trace launched");
    hipLaunchKernelGGL(k_add_overflow, dim3(1), dim3(64), 0, st, R->cursor.as<unsigned long long>(), acc);
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

}  // namespace gpr
