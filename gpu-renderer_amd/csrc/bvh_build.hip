// bvh_build.hip -- device LBVH construction for gfx950.
//
// Replaces the reference's single-threaded host builder LBVH_build
// (src/renderer/bvh.h:186-221) with an all-device pipeline that produces the
// same tree:
//   1. centroid bounds            (bvh.h:192-196)           two-level block reduction
//   2. 30-bit Morton keys         (bvh.h:12-32, 199-204)    one thread per triangle
//   3. (key, index) sort          (bvh.h:207-208)           LDS-staged MSD radix sort
//   4. split weights              (bvh.h:161-176)           adjacent-pair common prefix;
//      equal-key runs get 32 + the split depth of the reference's midpoint
//      rule ((l+r)>>1)-1, so the radix tree reproduces that rule exactly
//   5. tree + refit               (bvh.h:143-184)           bottom-up agglomerative build
//      (parent = deeper of the two boundary splits) fused with the box union
//   6. layouts                    (bvh.h:121-141)           reference DFS order via
//      pre-order index = 2*first_leaf + number_of_left_steps, and packed 128-B
//      BVH4 nodes (binary nodes at even depth) for the traversal kernels.
#include "dev_math.h"
#include "gpr_internal.h"

#include <algorithm>

namespace gpr {
namespace {

constexpr int kB = 256;              // threads per block
constexpr int kMaxDepth = 256;       // > 30 radix levels + 32 midpoint levels

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Triangle::worldBound (src/renderer/triangle.h:96-101): Union(AABB(p0,p1), p2)
GPR_D void tri_bounds(const float* p, v3& lo, v3& hi) {
    const v3 p0 = load3(p), p1 = load3(p + 3), p2 = load3(p + 6);
    const v3 a = vmin3(p0, p1), b = vmax3(p0, p1);
    lo = vmin3(a, p2);
    hi = vmax3(b, p2);
}

// ------------------------------------------------------------ bounds ----
template <int NT>
GPR_D void block_minmax(v3& lo, v3& hi, float* sh) {
    const int t = threadIdx.x;
    sh[t * 6 + 0] = lo.x; sh[t * 6 + 1] = lo.y; sh[t * 6 + 2] = lo.z;
    sh[t * 6 + 3] = hi.x; sh[t * 6 + 4] = hi.y; sh[t * 6 + 5] = hi.z;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (t < s) {
            for (int k = 0; k < 3; ++k) sh[t * 6 + k] = fmin_h(sh[t * 6 + k], sh[(t + s) * 6 + k]);
            for (int k = 3; k < 6; ++k) sh[t * 6 + k] = fmax_h(sh[t * 6 + k], sh[(t + s) * 6 + k]);
        }
        __syncthreads();
    }
    lo = V(sh[0], sh[1], sh[2]);
    hi = V(sh[3], sh[4], sh[5]);
}

__global__ __launch_bounds__(kB) void k_centroid_partial(const float* __restrict__ pos, uint32_t n,
                                                         float* __restrict__ part) {
    __shared__ float sh[kB * 6];
    v3 lo = V(INFINITY, INFINITY, INFINITY), hi = V(-INFINITY, -INFINITY, -INFINITY);
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        v3 a, b;
        tri_bounds(pos + 9ull * i, a, b);
        const v3 c = (a + b) * 0.5f;  // AABB::centroid (aabb.h:23)
        lo = vmin3(lo, c);
        hi = vmax3(hi, c);
    }
    block_minmax<kB>(lo, hi, sh);
    if (threadIdx.x == 0) {
        float* o = part + 6 * blockIdx.x;
        o[0] = lo.x; o[1] = lo.y; o[2] = lo.z; o[3] = hi.x; o[4] = hi.y; o[5] = hi.z;
    }
}

__global__ __launch_bounds__(kB) void k_centroid_final(const float* __restrict__ part, uint32_t np,
                                                       float* __restrict__ out) {
    __shared__ float sh[kB * 6];
    v3 lo = V(INFINITY, INFINITY, INFINITY), hi = V(-INFINITY, -INFINITY, -INFINITY);
    for (uint32_t i = threadIdx.x; i < np; i += kB) {
        lo = vmin3(lo, load3(part + 6 * i));
        hi = vmax3(hi, load3(part + 6 * i + 3));
    }
    block_minmax<kB>(lo, hi, sh);
    if (threadIdx.x == 0) {
        out[0] = lo.x; out[1] = lo.y; out[2] = lo.z; out[3] = hi.x; out[4] = hi.y; out[5] = hi.z;
    }
}

// ------------------------------------------------------------ Morton ----
// LeftShift3 (bvh.h:12-25): 10 bits to every third bit, 1024 -> 1023.
GPR_D uint32_t spread3(uint32_t x) {
    x = x == 1024u ? 1023u : x;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
// float -> uint truncation; NaN (flat axis, 0/0) -> 0.
GPR_D uint32_t quant(float v) {
    if (!(v == v) || v <= 0.0f) return 0u;
    return v >= 4294967040.0f ? 0xffffffffu : (uint32_t)v;
}

__global__ __launch_bounds__(kB) void k_morton(const float* __restrict__ pos, uint32_t n,
                                               const float* __restrict__ bounds,
                                               uint32_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const v3 lo = load3(bounds), hi = load3(bounds + 3);
    v3 a, b;
    tri_bounds(pos + 9ull * i, a, b);
    const v3 c = (a + b) * 0.5f;
    const v3 q = (c - lo) / (hi - lo) * 1024.f;  // EncodeMorton3 input (bvh.h:203)
    keys[i] = (spread3(quant(q.z)) << 2) | (spread3(quant(q.y)) << 1) | spread3(quant(q.x));
}

// Inclusive block scan (256 threads, wave64) of one value per thread.
GPR_D uint32_t block_incl_scan(uint32_t v, uint32_t* sh /* >= 4 */, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    if (lane == 63) sh[wave] = v;
    __syncthreads();
    uint32_t add = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kB / 64; ++w) {
        const uint32_t s = sh[w];
        if (w < wave) add += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return v + add;
}

// ---------------------------------------------------------- weights -----
// Split depth of adjacent pair j in the reference's midpoint tree over a run
// of m equal keys: root split after floor(m/2) leaves (bvh.h:162-165).
GPR_D uint32_t midpoint_depth(uint32_t j, uint32_t m) {
    uint32_t d = 0;
    while (m > 1) {
        const uint32_t h = m >> 1;
        if (j == h - 1) return d;
        if (j < h - 1) {
            m = h;
        } else {
            j -= h;
            m -= h;
        }
        ++d;
    }
    return d;
}

__global__ __launch_bounds__(kB) void k_weights(const uint32_t* __restrict__ sk, uint32_t n,
                                                int32_t* __restrict__ w) {
    const uint32_t k = blockIdx.x * kB + threadIdx.x;
    if (k + 1 >= n) return;
    const uint32_t a = sk[k], b = sk[k + 1];
    if (a != b) {
        w[k] = __clz(a ^ b);
        return;
    }
    // run [s, e) of key a
    uint32_t lo = 0, hi = k;
    while (lo < hi) {
        const uint32_t m = (lo + hi) >> 1;
        if (sk[m] < a) lo = m + 1;
        else hi = m;
    }
    const uint32_t s = lo;
    lo = k + 1;
    hi = n;
    while (lo < hi) {
        const uint32_t m = (lo + hi) >> 1;
        if (sk[m] <= a) lo = m + 1;
        else hi = m;
    }
    w[k] = 32 + (int32_t)midpoint_depth(k - s, lo - s);
}

// ------------------------------------------------------ tree + refit ----
// Node ids: internal node k (split between sorted leaves k and k+1) = k,
// leaf i = (n-1)+i.  A node covering [l, r] has parent r (as left child) if
// the split after r is deeper than the split before l, else parent l-1 (as
// right child).  The second child to arrive at a parent forms its box.
//
// Everything one thread publishes for another (boxes, child ids, range
// bounds) is written with device-scope relaxed atomic stores, which gfx950
// issues as write-through stores (sc1) to the device-coherent level; the
// writer waits for their completion (vmcnt(0)) before it bumps the arrival
// counter (a device-scope RMW, performed at the coherent level), and the
// reader loads with device-scope atomic loads (sc1, served from the coherent
// level) only after that RMW has returned.  So no L2 write-back (buffer_wbl2)
// or invalidate is needed although the eight XCD L2 caches are not coherent
// with each other; device-scope release / acquire on the RMW would add both
// on every arrival (measured 16x slower at 1M triangles).  Values read back
// are range-checked, and k_export_ref checks every parent box against the
// union of its children's, so a protocol failure raises the error word
// instead of indexing out of bounds or producing a wrong tree silently.
GPR_D void pub(uint32_t* p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
GPR_D uint32_t get(uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
GPR_D void pub2(uint64_t* p, float a, float b) {
    const uint64_t v = (uint64_t)__float_as_uint(a) | ((uint64_t)__float_as_uint(b) << 32);
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
GPR_D void get2(uint64_t* p, float& a, float& b) {
    const uint64_t v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    a = __uint_as_float((uint32_t)v);
    b = __uint_as_float((uint32_t)(v >> 32));
}

__global__ __launch_bounds__(kB) void k_tree(const float* __restrict__ pos, uint32_t n,
                                             const uint32_t* __restrict__ sidx,
                                             const int32_t* __restrict__ w,
                                             uint32_t* __restrict__ parent,
                                             uint32_t* __restrict__ child,
                                             uint32_t* __restrict__ rlo, uint32_t* __restrict__ rhi,
                                             uint32_t* __restrict__ arrive,
                                             float* __restrict__ box,  // (2n-1)*6
                                             uint32_t* __restrict__ root) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    v3 lo, hi;
    tri_bounds(pos + 9ull * sidx[i], lo, hi);
    uint32_t node = (n - 1) + i;
    // a node's box is 24 B at 6*node floats: three 8-byte words
    uint64_t* bx = reinterpret_cast<uint64_t*>(box);
    uint32_t l = i, r = i;
    for (int guard = 0; guard < kMaxDepth; ++guard) {
        uint64_t* b = bx + 3ull * node;
        pub2(b + 0, lo.x, lo.y);
        pub2(b + 1, lo.z, hi.x);
        pub2(b + 2, hi.y, hi.z);
        if (l == 0 && r == n - 1) {
            *root = node;
            break;
        }
        bool as_left;
        if (l == 0) as_left = true;
        else if (r == n - 1) as_left = false;
        else as_left = w[r] > w[l - 1];
        const uint32_t p = as_left ? r : l - 1;
        parent[node] = p;
        if (as_left) {
            pub(&child[2 * p], node);
            pub(&rlo[p], l);
        } else {
            pub(&child[2 * p + 1], node);
            pub(&rhi[p], r);
        }
        // every publish above has completed (s_waitcnt vmcnt(0)) before the
        // arrival RMW is issued; the fence keeps the compiler from moving the
        // publishing stores past the wait
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
        if (__hip_atomic_fetch_add(&arrive[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
        // second arrival: the sibling's publishes completed before its
        // increment; the loads below are issued after this RMW returned
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (as_left) r = get(&rhi[p]);
        else l = get(&rlo[p]);
        const uint32_t sib = get(&child[2 * p + (as_left ? 1 : 0)]);
        if (r >= n || l > p || (as_left && r <= p) || sib >= 2 * n - 1) {
            atomicOr(root + 1, 1u);
            break;
        }
        uint64_t* sb = bx + 3ull * sib;
        v3 slo, shi;
        get2(sb + 0, slo.x, slo.y);
        get2(sb + 1, slo.z, shi.x);
        get2(sb + 2, shi.y, shi.z);
        // Union(l_child, r_child) (aabb.h:37-46, bvh.h:182)
        if (as_left) {
            lo = vmin3(lo, slo);
            hi = vmax3(hi, shi);
        } else {
            lo = vmin3(slo, lo);
            hi = vmax3(shi, hi);
        }
        node = p;
    }
}

// Depth and number of left steps from the root, per node.
__global__ __launch_bounds__(kB) void k_walkup(uint32_t n, const uint32_t* __restrict__ parent,
                                               const uint32_t* __restrict__ child,
                                               const uint32_t* __restrict__ rlo,
                                               const uint32_t* __restrict__ root,
                                               uint32_t* __restrict__ pre,
                                               uint32_t* __restrict__ depth) {
    const uint32_t x = blockIdx.x * kB + threadIdx.x;
    if (x >= 2 * n - 1) return;
    const uint32_t rt = *root;
    uint32_t y = x, dep = 0, left = 0;
    while (y != rt && dep < kMaxDepth) {
        const uint32_t p = parent[y];
        if (p >= n - 1) break;  // only reachable through a malformed tree
        left += child[2 * p] == y ? 1u : 0u;
        ++dep;
        y = p;
    }
    const uint32_t first = x < n - 1 ? rlo[x] : x - (n - 1);
    pre[x] = 2 * first + left;
    depth[x] = dep;
}

__global__ __launch_bounds__(kB) void k_export_ref(uint32_t n, const uint32_t* __restrict__ child,
                                                   const uint32_t* __restrict__ sidx,
                                                   const float* __restrict__ box,
                                                   const uint32_t* __restrict__ pre,
                                                   gpr_ref_node* __restrict__ ref,
                                                   uint32_t* __restrict__ err) {
    const uint32_t x = blockIdx.x * kB + threadIdx.x;
    if (x >= 2 * n - 1) return;
    if (pre[x] >= 2 * n - 1) {
        atomicOr(err, 4u);  // a node the walk from the root did not reach
        return;
    }
    if (x < n - 1) {
        // the parent box must be the union of its children's boxes as they
        // stand now (a stale box read in k_tree would break the equality)
        const uint32_t c0 = child[2 * x], c1 = child[2 * x + 1];
        if (c0 >= 2 * n - 1 || c1 >= 2 * n - 1) {
            atomicOr(err, 2u);
        } else {
            const float* a = box + 6ull * c0;
            const float* b = box + 6ull * c1;
            const float* p = box + 6ull * x;
            const v3 lo = vmin3(V(a[0], a[1], a[2]), V(b[0], b[1], b[2]));
            const v3 hi = vmax3(V(a[3], a[4], a[5]), V(b[3], b[4], b[5]));
            const bool same = __float_as_uint(lo.x) == __float_as_uint(p[0]) &&
                              __float_as_uint(lo.y) == __float_as_uint(p[1]) &&
                              __float_as_uint(lo.z) == __float_as_uint(p[2]) &&
                              __float_as_uint(hi.x) == __float_as_uint(p[3]) &&
                              __float_as_uint(hi.y) == __float_as_uint(p[4]) &&
                              __float_as_uint(hi.z) == __float_as_uint(p[5]);
            if (!same) atomicOr(err, 2u);
        }
    }
    gpr_ref_node o;
    const float* b = box + 6ull * x;
    o.bmin[0] = b[0]; o.bmin[1] = b[1]; o.bmin[2] = b[2];
    o.bmax[0] = b[3]; o.bmax[1] = b[4]; o.bmax[2] = b[5];
    if (x >= n - 1) {
        o.idx = sidx[x - (n - 1)];
        o.leaf = 1;
    } else {
        const uint32_t rc = child[2 * x + 1];
        o.idx = rc < 2 * n - 1 ? pre[rc] : 0u;
        o.leaf = 0;
    }
    ref[pre[x]] = o;
}

// Flags of BVH4 nodes (internal, even depth) in DFS (pre-order) position.
__global__ __launch_bounds__(kB) void k_bvh4_flags(uint32_t n, const uint32_t* __restrict__ pre,
                                                   const uint32_t* __restrict__ depth,
                                                   uint32_t* __restrict__ flag) {
    const uint32_t x = blockIdx.x * kB + threadIdx.x;
    if (x >= 2 * n - 1 || pre[x] >= 2 * n - 1) return;
    flag[pre[x]] = (x < n - 1 && (depth[x] & 1u) == 0) ? 1u : 0u;
}

// Three-pass exclusive scan of u32 (blocks of 1024 elements).
__global__ __launch_bounds__(kB) void k_scan_blocks(const uint32_t* __restrict__ in, uint64_t n,
                                                    uint32_t* __restrict__ out,
                                                    uint32_t* __restrict__ bsum) {
    __shared__ uint32_t sh[4];
    const uint64_t b0 = (uint64_t)blockIdx.x * 1024;
    uint32_t v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint64_t i = b0 + 4 * threadIdx.x + k;
        v[k] = i < n ? in[i] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = block_incl_scan(s, sh, &tot) - s;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint64_t i = b0 + 4 * threadIdx.x + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}
__global__ __launch_bounds__(kB) void k_scan_top(uint32_t* __restrict__ bsum, uint32_t nb,
                                                 uint32_t* __restrict__ total) {
    __shared__ uint32_t sh[4];
    uint32_t carry = 0;
    for (uint32_t b = 0; b < nb; b += kB) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t v = i < nb ? bsum[i] : 0u;
        uint32_t tot;
        const uint32_t inc = block_incl_scan(v, sh, &tot);
        if (i < nb) bsum[i] = carry + inc - v;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}
__global__ __launch_bounds__(kB) void k_scan_add(uint32_t* __restrict__ out, uint64_t n,
                                                 const uint32_t* __restrict__ bsum) {
    const uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x;
    if (i < n) out[i] += bsum[i / 1024];
}

GPR_D void put_child(Bvh4Node& o, int slot, const float* b, int32_t ref) {
    o.lox[slot] = b[0]; o.loy[slot] = b[1]; o.loz[slot] = b[2];
    o.hix[slot] = b[3]; o.hiy[slot] = b[4]; o.hiz[slot] = b[5];
    o.child[slot] = ref;
}

__global__ __launch_bounds__(kB) void k_bvh4(uint32_t n, const uint32_t* __restrict__ child,
                                             const float* __restrict__ box,
                                             const uint32_t* __restrict__ pre,
                                             const uint32_t* __restrict__ depth,
                                             const uint32_t* __restrict__ id4,  // by pre-order pos
                                             Bvh4Node* __restrict__ out) {
    const uint32_t x = blockIdx.x * kB + threadIdx.x;
    if (x + 1 >= n || (depth[x] & 1u)) return;
    const uint32_t nn = 2 * n - 1;
    if (pre[x] >= nn || id4[pre[x]] >= n - 1) return;
    Bvh4Node o;
    int slot = 0;
    for (int side = 0; side < 2; ++side) {
        const uint32_t c = child[2 * x + side];
        if (c >= nn) return;  // malformed tree: leave the node unwritten
        if (c >= n - 1) {
            put_child(o, slot++, box + 6ull * c, ~(int32_t)(c - (n - 1)));
        } else {
            for (int s2 = 0; s2 < 2; ++s2) {
                const uint32_t g = child[2 * c + s2];
                if (g >= nn || pre[g] >= nn) return;
                const int32_t ref = g >= n - 1 ? ~(int32_t)(g - (n - 1)) : (int32_t)id4[pre[g]];
                put_child(o, slot++, box + 6ull * g, ref);
            }
        }
    }
    for (; slot < 4; ++slot) {
        o.lox[slot] = o.loy[slot] = o.loz[slot] = INFINITY;
        o.hix[slot] = o.hiy[slot] = o.hiz[slot] = INFINITY;
        o.child[slot] = kEmptyChild;
    }
    o.pad[0] = x;
    o.pad[1] = o.pad[2] = o.pad[3] = 0;
    out[id4[pre[x]]] = o;
}

__global__ __launch_bounds__(kB) void k_tris(const float* __restrict__ pos, uint32_t n,
                                             const uint32_t* __restrict__ sidx,
                                             TriRec* __restrict__ out) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = sidx[i];
    const float* p = pos + 9ull * t;
    TriRec r;
    r.a = make_float4(p[0], p[1], p[2], p[3]);
    r.b = make_float4(p[4], p[5], p[6], p[7]);
    r.c = make_float4(p[8], __uint_as_float(t), 0.f, 0.f);
    out[i] = r;
}

struct Carve {
    char* base;
    size_t off;
    template <class T>
    T* take(size_t count) {
        off = (off + 255) & ~(size_t)255;
        T* p = reinterpret_cast<T*>(base + off);
        off += count * sizeof(T);
        return p;
    }
};

}  // namespace

int build_bvh(gpr_scene* s) {
    BvhDev& B = s->bvh;
    const uint32_t n = (uint32_t)s->mat.size();
    B.n = n;
    if (n == 0) {
        B.nbvh4 = 0;
        return GPR_OK;
    }
    hipStream_t st = stream();
    const uint32_t nn = 2 * n - 1;
    const uint32_t nscan_blocks = div_up(nn, 1024);
    const uint32_t npart = std::min(1024u, div_up(n, kB));
    // ---- carve the arena
    size_t need = 0;
    for (int pass = 0; pass < 2; ++pass) {
        Carve c{pass ? B.arena.as<char>() : nullptr, 0};
        float* part = c.take<float>(6ull * npart);
        float* bounds = c.take<float>(8);
        uint32_t* keys = c.take<uint32_t>(n);
        uint32_t* skeys = c.take<uint32_t>(n);
        uint32_t* sidx = c.take<uint32_t>(n);
        int32_t* w = c.take<int32_t>(n);
        uint32_t* parent = c.take<uint32_t>(nn);
        uint32_t* child = c.take<uint32_t>(2ull * n);
        uint32_t* rlo = c.take<uint32_t>(n);
        uint32_t* rhi = c.take<uint32_t>(n);
        uint32_t* arrive = c.take<uint32_t>(n);
        float* box = c.take<float>(6ull * nn);
        uint32_t* root = c.take<uint32_t>(4);
        uint32_t* pre = c.take<uint32_t>(nn);
        uint32_t* depth = c.take<uint32_t>(nn);
        uint32_t* flag = c.take<uint32_t>(nn);
        uint32_t* id4 = c.take<uint32_t>(nn);
        uint32_t* bsum = c.take<uint32_t>(nscan_blocks + 4);
        uint32_t* total4 = c.take<uint32_t>(4);
        gpr_ref_node* ref = c.take<gpr_ref_node>(nn);
        Bvh4Node* bvh4 = c.take<Bvh4Node>(std::max(1u, n - 1));
        TriRec* tris = c.take<TriRec>(n);
        if (pass == 0) {
            need = c.off + 256;
            GPR_TRY(B.arena.alloc(need));
            const size_t sb = sort_scratch_bytes(n);
            if (B.sort_scratch.bytes < sb || !B.sort_scratch.p) {
                GPR_TRY(B.sort_scratch.alloc(sb));
                GPR_TRY(sort_prepare(B.sort_scratch.p, B.sort_scratch.bytes, st));
            }
            continue;
        }
        B.keys = keys;
        B.skeys = skeys;
        B.sidx = sidx;
        B.ref = ref;
        B.bvh4 = bvh4;
        B.tris = tris;
        const float* pos = s->d_pos.as<float>();
        hipEvent_t ev[5];
        for (int k = 0; k < 5; ++k) GPR_HIP(hipEventCreate(&ev[k]));
        GPR_HIP(hipEventRecord(ev[0], st));
        hipLaunchKernelGGL(k_centroid_partial, dim3(npart), dim3(kB), 0, st, pos, n, part);
        hipLaunchKernelGGL(k_centroid_final, dim3(1), dim3(kB), 0, st, part, npart, bounds);
        hipLaunchKernelGGL(k_morton, dim3(div_up(n, kB)), dim3(kB), 0, st, pos, n, bounds, keys);
        GPR_HIP(hipEventRecord(ev[1], st));
        GPR_TRY(sort_pairs(keys, skeys, sidx, n, B.sort_scratch.p, B.sort_scratch.bytes, st));
        GPR_HIP(hipEventRecord(ev[2], st));
        hipLaunchKernelGGL(k_fill_u32, dim3(std::min(1024u, div_up(n, kB))), dim3(kB), 0, st, arrive, 0u, (uint64_t)n);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kB), 0, st, root + 1, 0u, (uint64_t)1);
        hipLaunchKernelGGL(k_weights, dim3(div_up(n, kB)), dim3(kB), 0, st, skeys, n, w);
        hipLaunchKernelGGL(k_tree, dim3(div_up(n, kB)), dim3(kB), 0, st, pos, n, sidx, w, parent,
                           child, rlo, rhi, arrive, box, root);
        GPR_HIP(hipEventRecord(ev[3], st));
        hipLaunchKernelGGL(k_walkup, dim3(div_up(nn, kB)), dim3(kB), 0, st, n, parent, child, rlo, root, pre, depth);
        hipLaunchKernelGGL(k_export_ref, dim3(div_up(nn, kB)), dim3(kB), 0, st, n, child, sidx, box, pre, ref,
                           root + 1);
        hipLaunchKernelGGL(k_bvh4_flags, dim3(div_up(nn, kB)), dim3(kB), 0, st, n, pre, depth, flag);
        hipLaunchKernelGGL(k_scan_blocks, dim3(nscan_blocks), dim3(kB), 0, st, flag, (uint64_t)nn, id4, bsum);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kB), 0, st, bsum, nscan_blocks, total4);
        hipLaunchKernelGGL(k_scan_add, dim3(div_up(nn, kB)), dim3(kB), 0, st, id4, (uint64_t)nn, bsum);
        if (n > 1)
            hipLaunchKernelGGL(k_bvh4, dim3(div_up(n - 1, kB)), dim3(kB), 0, st, n, child, box, pre,
                               depth, id4, bvh4);
        hipLaunchKernelGGL(k_tris, dim3(div_up(n, kB)), dim3(kB), 0, st, pos, n, sidx, tris);
        GPR_HIP(hipEventRecord(ev[4], st));
        GPR_HIP(hipGetLastError());
        uint32_t nb4 = 0, tree_err = 0;
        GPR_HIP(hipMemcpyAsync(&nb4, total4, 4, hipMemcpyDeviceToHost, st));
        GPR_HIP(hipMemcpyAsync(&tree_err, root + 1, 4, hipMemcpyDeviceToHost, st));
        GPR_HIP(hipStreamSynchronize(st));
        if (tree_err)
            return fail(GPR_ERR_INVALID, std::string("LBVH build: ") +
                                             ((tree_err & 1u) ? "inconsistent node publication " : "") +
                                             ((tree_err & 2u) ? "parent box differs from the union of its children " : "") +
                                             ((tree_err & 4u) ? "node unreachable from the root" : ""));
        B.nbvh4 = nb4;
        if (n > 1 && (nb4 == 0 || nb4 > n - 1)) return fail(GPR_ERR_INVALID, "LBVH build produced an invalid BVH4 node count");
        B.root4 = n == 1 ? ~(int32_t)0 : 0;
        float ms;
        for (int k = 0; k < 4; ++k) {
            GPR_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            B.timing[k] = ms;
        }
        GPR_HIP(hipEventElapsedTime(&ms, ev[0], ev[4]));
        B.timing[4] = ms;
        for (int k = 0; k < 5; ++k) (void)hipEventDestroy(ev[k]);
    }
    return GPR_OK;
}

}  // namespace gpr
