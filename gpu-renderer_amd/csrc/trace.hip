// trace.hip -- ray-query kernels behind gpr_trace (trace / trace_shadow,
// src/renderer/pathtracer_kernel.h:235-246, 310-321).
//
// k_trace_fast is persistent: every wave pulls 64 consecutive rays from a
// global atomic cursor until the batch is drained, so the grid is sized to
// the device (not to the batch) and the per-lane overflow stacks are indexed
// by the persistent thread id.
#include "traverse.h"

namespace gpr {
namespace {

__global__ void k_zero_u64(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
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

template <bool ANY>
__global__ __launch_bounds__(kTraceBlock) void k_trace_fast(SceneView sv, const gpr_ray* __restrict__ rays,
                                                            gpr_hit* __restrict__ hits, uint64_t n,
                                                            unsigned long long* __restrict__ cursor,
                                                            int2* __restrict__ spill) {
    __shared__ int2 lds[kStackLds * kTraceBlock];
    LaneStack st;
    st.lds = lds + threadIdx.x;
    st.spill = spill + (uint64_t)(blockIdx.x * kTraceBlock + threadIdx.x) * kStackGlobal;
    st.sp = 0;
    const int lane = threadIdx.x & 63;
    while (true) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(cursor, 64ull);
        base = __shfl(base, 0, 64);
        if (base >= n) break;
        const uint64_t i = base + lane;
        if (i < n) {
            const Ray3 r = load_ray(rays, i);
            gpr_hit h;
            if (!fast_trace<ANY>(sv, r, st, h)) atomicAdd(cursor + 1, 1ull);  // stack overflow count
            store_hit(hits, i, h);
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

struct TraceResources {
    DevBuf spill;
    DevBuf cursor;
    int grid = 0;
    int dev = -1;
};
TraceResources g_res[16];

int ensure_resources(TraceResources& R) {
    const int dev = device_id();
    if (R.dev == dev && R.grid > 0) return GPR_OK;
    int cus = 0;
    GPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    int per_cu = 0;
    GPR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_fast<false>, kTraceBlock, 0));
    if (per_cu < 1) per_cu = 1;
    R.grid = cus * per_cu;
    GPR_TRY(R.spill.alloc((size_t)R.grid * kTraceBlock * kStackGlobal * sizeof(int2)));
    GPR_TRY(R.cursor.alloc(4 * sizeof(unsigned long long)));
    R.dev = dev;
    return GPR_OK;
}

}  // namespace

int launch_trace(const SceneView& sv, const gpr_ray* rays, gpr_hit* hits, uint64_t n, uint32_t flags,
                 unsigned long long* stats, hipStream_t st) {
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
    TraceResources& R = g_res[device_id() & 15];
    GPR_TRY(ensure_resources(R));
    unsigned long long* cur = R.cursor.as<unsigned long long>();
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, cur, 4);
    if (any)
        hipLaunchKernelGGL(k_trace_fast<true>, dim3(R.grid), dim3(kTraceBlock), 0, st, sv, rays, hits, n,
                           cur, R.spill.as<int2>());
    else
        hipLaunchKernelGGL(k_trace_fast<false>, dim3(R.grid), dim3(kTraceBlock), 0, st, sv, rays, hits, n,
                           cur, R.spill.as<int2>());
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

// Overflow count of the last fast trace on this device (0 = every ray completed).
int trace_overflow_count(unsigned long long* out) {
    TraceResources& R = g_res[device_id() & 15];
    *out = 0;
    if (R.grid == 0) return GPR_OK;
    GPR_HIP(hipMemcpyAsync(out, R.cursor.as<unsigned long long>() + 1, 8, hipMemcpyDeviceToHost, stream()));
    GPR_HIP(hipStreamSynchronize(stream()));
    return GPR_OK;
}

}  // namespace gpr
