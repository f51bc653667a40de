// traverse.h -- closest-hit and any-hit traversal device code for gfx950.
//
// Two traversals share one triangle test:
//   * ref_*  : the reference's visit order on its DFS layout (BVHLinearNode),
//              its slab test and its per-thread 64-entry stack
//              (src/renderer/pathtracer_kernel.h:154-308).  Used for exact
//              reference semantics and for the traversal statistics.
//   * fast_* : packed 128-B BVH4 nodes (one cache line per node), a robust slab
//              test on all four children, nearest-first descent, leaf triangles
//              tested as soon as their parent is visited (the persistent any-hit
//              kernel instead parks them for wave-wide leaf rounds, TravD), and a
//              short per-lane stack in LDS that spills its oldest entries to a
//              global overflow area.  Closest-hit results equal the reference's up
//              to ties of t.
#pragma once

#include "dev_math.h"
#include "gpr_internal.h"

namespace gpr {

constexpr int kStackLds = 16;      // LDS stack entries per lane (power of two)
constexpr int kStackGlobal = 96;   // global overflow entries per lane
constexpr int kTraceBlock = 256;

struct Ray3 {
    v3 o, d;
    float tmin, tmax;
};

// triangle_intersect (pathtracer_kernel.h:119-152): float dots widened to
// double, double 1/det; rejects det==0, u<0|u>1, v<0|u+v>1, t<tmin|t>tmax.
GPR_D bool tri_hit_ref(v3 p0, v3 p1, v3 p2, v3 o, v3 d, float tmin, float tmax, float& th,
                       float& uh, float& vh) {
    const v3 e1 = p1 - p0;
    const v3 e2 = p2 - p0;
    const v3 P = cross3(d, e2);
    const double det = (double)dot3(P, e1);
    if (det == 0) return false;
    const double inv_det = 1.0 / det;
    const v3 T = o - p0;
    const double u = (double)dot3(P, T) * inv_det;
    if (u < 0.0 || u > 1.0) return false;
    const v3 Q = cross3(T, e1);
    const double v = (double)dot3(Q, d) * inv_det;
    if (v < 0.0 || u + v > 1.0) return false;
    const double t = (double)dot3(Q, e2) * inv_det;
    if (t < (double)tmin || t > (double)tmax) return false;
    th = (float)t;
    uh = (float)u;
    vh = (float)v;
    return true;
}

// ------------------------------------------------- reference traversal ---
// aabb_intersect (pathtracer_kernel.h:154-173)
GPR_D bool ref_box(const gpr_ref_node& nd, v3 o, float tmin_r, float tmax_r, v3 inv, int nx, int ny,
                   int nz) {
    const float bx0 = nx ? nd.bmax[0] : nd.bmin[0], bx1 = nx ? nd.bmin[0] : nd.bmax[0];
    const float by0 = ny ? nd.bmax[1] : nd.bmin[1], by1 = ny ? nd.bmin[1] : nd.bmax[1];
    const float bz0 = nz ? nd.bmax[2] : nd.bmin[2], bz1 = nz ? nd.bmin[2] : nd.bmax[2];
    float t0 = (bx0 - o.x) * inv.x;
    float t1 = (bx1 - o.x) * inv.x;
    const float ty0 = (by0 - o.y) * inv.y;
    const float ty1 = (by1 - o.y) * inv.y;
    if (t0 > ty1 || ty0 > t1) return false;
    t0 = fmax_h(t0, ty0);
    t1 = fmin_h(t1, ty1);
    const float tz0 = (bz0 - o.z) * inv.z;
    const float tz1 = (bz1 - o.z) * inv.z;
    if (t0 > tz1 || tz0 > t1) return false;
    t0 = fmax_h(t0, tz0);
    t1 = fmin_h(t1, tz1);
    return (t0 < tmax_r) && (t1 > tmin_r);
}

// trace_kernel / trace_shadow_kernel bodies (pathtracer_kernel.h:175-308).
template <bool ANY>
GPR_D gpr_hit ref_trace(const SceneView& sv, Ray3 r, uint32_t* nvisit, uint32_t* ntest) {
    gpr_hit hit{INFINITY, -1, 0.f, 0.f};
    const v3 inv = V(1 / r.d.x, 1 / r.d.y, 1 / r.d.z);
    const int nx = inv.x < 0, ny = inv.y < 0, nz = inv.z < 0;
    int cur = 0, sp = 0;
    int stack[64];
    uint32_t visits = 0, tests = 0;
    while (true) {
        // every node is visited at most once; both checks only trip on a malformed layout
        if ((uint32_t)cur >= sv.nref || visits > sv.nref) break;
        const gpr_ref_node nd = ldg(sv.ref + cur);
        ++visits;
        if (ref_box(nd, r.o, r.tmin, r.tmax, inv, nx, ny, nz)) {
            if (nd.leaf & 0xffu) {
                if (nd.idx >= sv.ntri) break;
                const float* p = sv.pos + 9ull * nd.idx;
                float t, u, v;
                ++tests;
                const v3 p0 = V(ldg(p), ldg(p + 1), ldg(p + 2)), p1 = V(ldg(p + 3), ldg(p + 4), ldg(p + 5));
                const v3 p2 = V(ldg(p + 6), ldg(p + 7), ldg(p + 8));
                if (tri_hit_ref(p0, p1, p2, r.o, r.d, r.tmin, r.tmax, t, u, v)) {
                    r.tmax = t;
                    hit = gpr_hit{t, (int32_t)nd.idx, u, v};
                    if (ANY) break;
                }
                if (sp == 0) break;
                cur = stack[--sp];
            } else {
                if (sp >= 64) break;  // unreachable for 30-bit keys; never write past the stack
                stack[sp++] = cur + 1;
                cur = (int)nd.idx;
            }
        } else {
            if (sp == 0) break;
            cur = stack[--sp];
        }
    }
    *nvisit += visits;
    *ntest += tests;
    return hit;
}

// ------------------------------------------------------ BVH4 traversal ---
// Per-lane traversal stack: the top SL entries in an LDS column of a block of
// NT lanes (entry k at lds[k * NT]), older entries spilled to a per-lane global
// slab of kStackGlobal entries.
template <int NT, int SL>
struct LaneStackT {
    static_assert((SL & (SL - 1)) == 0, "LDS stack depth must be a power of two");
    int2* lds;    // this lane's column
    int2* spill;  // this lane's overflow area (kStackGlobal entries)
    int sp;
    GPR_D bool push(int2 e) {
        if (sp >= SL + kStackGlobal) return false;
        const int slot = sp & (SL - 1);
        if (sp >= SL) spill[sp - SL] = lds[slot * NT];
        lds[slot * NT] = e;
        ++sp;
        return true;
    }
    GPR_D int2 pop() {
        --sp;
        const int slot = sp & (SL - 1);
        const int2 e = lds[slot * NT];
        if (sp >= SL) lds[slot * NT] = spill[sp - SL];
        return e;
    }
};
using LaneStack = LaneStackT<kTraceBlock, kStackLds>;

GPR_D void tri_rec_hit(const TriRec* tris, uint32_t ntri, int32_t leaf, const Ray3& r, float& tmax,
                       gpr_hit& hit, bool& found) {
    if ((uint32_t)leaf >= ntri) return;
    const TriRec* t = tris + leaf;
    const float4 a = ldg(&t->a), b = ldg(&t->b), c = ldg(&t->c);
    float th, uh, vh;
    if (tri_hit_ref(V(a.x, a.y, a.z), V(a.w, b.x, b.y), V(b.z, b.w, c.x), r.o, r.d, r.tmin, tmax, th,
                    uh, vh)) {
        tmax = th;
        hit = gpr_hit{th, (int32_t)__float_as_uint(c.y), uh, vh};
        found = true;
    }
}

// BVH4 traversal as an explicit state machine: trav_begin sets a ray up,
// each trav_step visits one node (its leaf children are tested first, they
// tighten tmax before the descent decision), so a caller can either run a ray
// to completion (fast_trace) or interleave the steps of many rays and give a
// lane a new ray as soon as its ray is done.
struct Trav {
    Ray3 r;
    v3 inv;
    int sx, sy, sz;  // 1 where the direction component is negative: the far plane comes first
    float tmax;
    gpr_hit hit;
    int32_t node;
    uint32_t fetched;
    uint32_t tested;  // triangle tests (read only by the statistics kernel; dead code elsewhere)
    bool found;
};
constexpr float kSlabCons = 1.0f + 2.0f * 1.7881394e-07f;  // 1 + 2*gamma(3)
enum : int { kTravMore = 0, kTravDone = 1, kTravOverflow = 2 };

// Returns true when the ray is already finished (single-triangle scene).
template <bool ANY, class Stack>
GPR_D bool trav_begin(const SceneView& sv, const Ray3& r, Stack& st, Trav& t) {
    t.r = r;
    t.hit = gpr_hit{INFINITY, -1, 0.f, 0.f};
    t.tmax = r.tmax;
    t.found = false;
    t.fetched = 0;
    t.tested = 0;
    st.sp = 0;
    if (sv.root4 < 0) {
        tri_rec_hit(sv.tris, sv.ntri, ~sv.root4, r, t.tmax, t.hit, t.found);
        return true;
    }
    t.inv = V(1 / r.d.x, 1 / r.d.y, 1 / r.d.z);
    t.sx = __builtin_signbit(t.inv.x) ? 1 : 0;
    t.sy = __builtin_signbit(t.inv.y) ? 1 : 0;
    t.sz = __builtin_signbit(t.inv.z) ? 1 : 0;
    t.node = sv.root4;
    return false;
}

template <bool ANY, class Stack>
GPR_D int trav_step(const SceneView& sv, Stack& st, Trav& t) {
    // a depth-first walk fetches every node at most once: the cap only trips on a malformed tree
    if ((uint32_t)t.node >= sv.nbvh4 || t.fetched > sv.nbvh4) return kTravOverflow;
    const Ray3& r = t.r;
    const float4* np = reinterpret_cast<const float4*>(sv.bvh4 + t.node);
    // the ray enters each slab at the plane its direction meets first: near
    // and far planes are loaded by the direction's signs (lx/hx at np[0]/np[1],
    // ly/hy at np[2]/np[3], lz/hz at np[4]/np[5]), so no per-axis min / max.
    // A 0 * inf (ray in a slab plane) gives NaN, which fmaxf / fminf ignore:
    // that axis then does not cull the child.
    const float4 nx = ldg(np + 0 + t.sx), fx = ldg(np + 1 - t.sx), ny = ldg(np + 2 + t.sy), fy = ldg(np + 3 - t.sy);
    const float4 nz = ldg(np + 4 + t.sz), fz = ldg(np + 5 - t.sz);
    const int4 ch = ldg(reinterpret_cast<const int4*>(np + 6));
    const float nxa[4] = {nx.x, nx.y, nx.z, nx.w}, fxa[4] = {fx.x, fx.y, fx.z, fx.w};
    const float nya[4] = {ny.x, ny.y, ny.z, ny.w}, fya[4] = {fy.x, fy.y, fy.z, fy.w};
    const float nza[4] = {nz.x, nz.y, nz.z, nz.w}, fza[4] = {fz.x, fz.y, fz.z, fz.w};
    const int32_t cha[4] = {ch.x, ch.y, ch.z, ch.w};
    float tn[4];
    bool hitc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float x0 = (nxa[c] - r.o.x) * t.inv.x, x1 = (fxa[c] - r.o.x) * t.inv.x;
        const float y0 = (nya[c] - r.o.y) * t.inv.y, y1 = (fya[c] - r.o.y) * t.inv.y;
        const float z0 = (nza[c] - r.o.z) * t.inv.z, z1 = (fza[c] - r.o.z) * t.inv.z;
        const float tnear = fmaxf(fmaxf(x0, y0), fmaxf(z0, r.tmin));
        const float tfar = fminf(fminf(x1, y1), fminf(z1, t.tmax));
        tn[c] = tnear;
        hitc[c] = tnear <= tfar * kSlabCons && cha[c] != kEmptyChild;
    }
    // leaf children in ascending slot order, as a loop over this lane's own
    // leaves: a wave runs the triangle test as often as its busiest lane
    // needs it, not once per slot that any lane has a leaf in
    uint32_t leaves = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) leaves |= (hitc[c] && cha[c] < 0) ? 1u << c : 0u;
    while (leaves != 0) {
        const int c = __builtin_ctz(leaves);
        leaves &= leaves - 1;
        const int32_t leaf = c == 0 ? cha[0] : c == 1 ? cha[1] : c == 2 ? cha[2] : cha[3];
        tri_rec_hit(sv.tris, sv.ntri, ~leaf, r, t.tmax, t.hit, t.found);
        ++t.tested;
        if (ANY && t.found) return kTravDone;
    }
    if constexpr (ANY) {
        // any hit needs no near-to-far order (any occluder ends the walk): the
        // lowest hit slot is visited first and the others pushed, with no
        // sorting network
        uint32_t m = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) m |= (hitc[c] && cha[c] >= 0) ? 1u << c : 0u;
        ++t.fetched;
        if (m != 0) {
            const int c0 = __builtin_ctz(m);
            m &= m - 1;
            while (m != 0) {
                const int c = 31 - __builtin_clz(m);
                m &= ~(1u << c);
                const int32_t ref = c == 1 ? cha[1] : c == 2 ? cha[2] : cha[3];
                if (!st.push(make_int2(ref, 0))) return kTravOverflow;
            }
            t.node = c0 == 0 ? cha[0] : c0 == 1 ? cha[1] : c0 == 2 ? cha[2] : cha[3];
            return kTravMore;
        }
        if (st.sp > 0) {
            t.node = st.pop().x;
            return kTravMore;
        }
        return kTravDone;
    }
    uint32_t key[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const bool go = hitc[c] && cha[c] >= 0 && tn[c] <= t.tmax * kSlabCons;
        key[c] = go ? ((__float_as_uint(fmaxf(tn[c], 0.f)) & ~3u) | (uint32_t)c) : 0xffffffffu;
    }
    // 4-key sorting network, ascending
#define GPR_CSWAP(a, b)                           \
    {                                             \
        const uint32_t lo_ = min(key[a], key[b]); \
        const uint32_t hi_ = max(key[a], key[b]); \
        key[a] = lo_;                             \
        key[b] = hi_;                             \
    }
    GPR_CSWAP(0, 1) GPR_CSWAP(2, 3) GPR_CSWAP(0, 2) GPR_CSWAP(1, 3) GPR_CSWAP(1, 2)
#undef GPR_CSWAP
    const int cnt = (key[0] != 0xffffffffu) + (key[1] != 0xffffffffu) + (key[2] != 0xffffffffu) +
                    (key[3] != 0xffffffffu);
    ++t.fetched;
    if (cnt > 0) {
        for (int k = cnt - 1; k >= 1; --k) {
            const int c = key[k] & 3;
            if (!st.push(make_int2(cha[c], __float_as_int(tn[c])))) return kTravOverflow;
        }
        t.node = cha[key[0] & 3];
        return kTravMore;
    }
    // pop the nearest pending node that can still hold a closer hit
    while (st.sp > 0) {
        const int2 e = st.pop();
        if (ANY || __int_as_float(e.y) <= t.tmax * kSlabCons) {
            t.node = e.x;
            return kTravMore;
        }
    }
    return kTravDone;
}

// Postponed-leaf form of the BVH4 walk, used by the persistent any-hit
// kernel.  A node step does not test its leaf children: it parks them (slot
// mask + references) and the lane pauses; the wave tests the parked leaves of
// its lanes together in one leaf round, so the (double precision) triangle
// test runs with many lanes active instead of the few whose node held
// leaves.  A lane descends before its leaves are tested; pops wait for the
// round.  For an any-hit query the answer (occluded or not) is unchanged; the
// occluder reported may be a different triangle than in the immediate order.
struct TravD : Trav {
    uint32_t pmask;    // parked leaf slots of the last visited node
    int32_t pleaf[4];  // that node's child references
};

template <bool ANY, class Stack>
GPR_D bool travd_begin(const SceneView& sv, const Ray3& r, Stack& st, TravD& t) {
    t.pmask = 0;
    return trav_begin<ANY>(sv, r, st, t);
}

// One node; pops first when the lane has no node to visit (kTravDone when the
// stack is empty).  Leaves hit are parked in t.pmask / t.pleaf.
template <bool ANY, class Stack>
GPR_D int travd_step(const SceneView& sv, Stack& st, TravD& t) {
    if (t.node < 0) {
        bool got = false;
        while (st.sp > 0) {
            const int2 e = st.pop();
            if (ANY || __int_as_float(e.y) <= t.tmax * kSlabCons) {
                t.node = e.x;
                got = true;
                break;
            }
        }
        if (!got) return kTravDone;
    }
    if ((uint32_t)t.node >= sv.nbvh4 || t.fetched > sv.nbvh4) return kTravOverflow;
    const Ray3& r = t.r;
    const float4* np = reinterpret_cast<const float4*>(sv.bvh4 + t.node);
    const float4 nx = ldg(np + 0 + t.sx), fx = ldg(np + 1 - t.sx), ny = ldg(np + 2 + t.sy), fy = ldg(np + 3 - t.sy);
    const float4 nz = ldg(np + 4 + t.sz), fz = ldg(np + 5 - t.sz);
    const int4 ch = ldg(reinterpret_cast<const int4*>(np + 6));
    const float nxa[4] = {nx.x, nx.y, nx.z, nx.w}, fxa[4] = {fx.x, fx.y, fx.z, fx.w};
    const float nya[4] = {ny.x, ny.y, ny.z, ny.w}, fya[4] = {fy.x, fy.y, fy.z, fy.w};
    const float nza[4] = {nz.x, nz.y, nz.z, nz.w}, fza[4] = {fz.x, fz.y, fz.z, fz.w};
    const int32_t cha[4] = {ch.x, ch.y, ch.z, ch.w};
    float tn[4];
    bool hitc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float x0 = (nxa[c] - r.o.x) * t.inv.x, x1 = (fxa[c] - r.o.x) * t.inv.x;
        const float y0 = (nya[c] - r.o.y) * t.inv.y, y1 = (fya[c] - r.o.y) * t.inv.y;
        const float z0 = (nza[c] - r.o.z) * t.inv.z, z1 = (fza[c] - r.o.z) * t.inv.z;
        const float tnear = fmaxf(fmaxf(x0, y0), fmaxf(z0, r.tmin));
        const float tfar = fminf(fminf(x1, y1), fminf(z1, t.tmax));
        tn[c] = tnear;
        hitc[c] = tnear <= tfar * kSlabCons && cha[c] != kEmptyChild;
    }
    uint32_t leaves = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        leaves |= (hitc[c] && cha[c] < 0) ? 1u << c : 0u;
        t.pleaf[c] = cha[c];
    }
    t.pmask = leaves;
    if constexpr (ANY) {
        // as in trav_step: no distance order for any hit
        uint32_t m = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) m |= (hitc[c] && cha[c] >= 0) ? 1u << c : 0u;
        ++t.fetched;
        if (m != 0) {
            const int c0 = __builtin_ctz(m);
            m &= m - 1;
            while (m != 0) {
                const int c = 31 - __builtin_clz(m);
                m &= ~(1u << c);
                const int32_t ref = c == 1 ? cha[1] : c == 2 ? cha[2] : cha[3];
                if (!st.push(make_int2(ref, 0))) return kTravOverflow;
            }
            t.node = c0 == 0 ? cha[0] : c0 == 1 ? cha[1] : c0 == 2 ? cha[2] : cha[3];
        } else {
            t.node = -1;
        }
        return kTravMore;
    }
    uint32_t key[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const bool go = hitc[c] && cha[c] >= 0 && tn[c] <= t.tmax * kSlabCons;
        key[c] = go ? ((__float_as_uint(fmaxf(tn[c], 0.f)) & ~3u) | (uint32_t)c) : 0xffffffffu;
    }
#define GPR_CSWAP(a, b)                           \
    {                                             \
        const uint32_t lo_ = min(key[a], key[b]); \
        const uint32_t hi_ = max(key[a], key[b]); \
        key[a] = lo_;                             \
        key[b] = hi_;                             \
    }
    GPR_CSWAP(0, 1) GPR_CSWAP(2, 3) GPR_CSWAP(0, 2) GPR_CSWAP(1, 3) GPR_CSWAP(1, 2)
#undef GPR_CSWAP
    const int cnt = (key[0] != 0xffffffffu) + (key[1] != 0xffffffffu) + (key[2] != 0xffffffffu) +
                    (key[3] != 0xffffffffu);
    ++t.fetched;
    if (cnt > 0) {
        for (int k = cnt - 1; k >= 1; --k) {
            const int c = key[k] & 3;
            if (!st.push(make_int2(cha[c], __float_as_int(tn[c])))) return kTravOverflow;
        }
        t.node = cha[key[0] & 3];
    } else {
        t.node = -1;
    }
    return kTravMore;
}

// Tests the parked leaves in slot order; true when an any-hit query is done.
template <bool ANY>
GPR_D bool travd_leaves(const SceneView& sv, TravD& t) {
    uint32_t m = t.pmask;
    t.pmask = 0;
    while (m != 0) {
        const int c = __builtin_ctz(m);
        m &= m - 1;
        const int32_t leaf = c == 0 ? t.pleaf[0] : c == 1 ? t.pleaf[1] : c == 2 ? t.pleaf[2] : t.pleaf[3];
        tri_rec_hit(sv.tris, sv.ntri, ~leaf, t.r, t.tmax, t.hit, t.found);
        if (ANY && t.found) return true;
    }
    return false;
}

// Closest hit (ANY=false) or any hit (ANY=true) against the BVH4.  Returns
// false only if the stack capacity was exceeded (the ray is then reported
// with the hits found so far and the caller raises an error flag).
template <bool ANY, class Stack>
GPR_D bool fast_trace(const SceneView& sv, const Ray3& r, Stack& st, gpr_hit& hit) {
    Trav t;
    int rc = trav_begin<ANY>(sv, r, st, t) ? kTravDone : kTravMore;
    while (rc == kTravMore) rc = trav_step<ANY>(sv, st, t);
    hit = t.hit;
    return rc != kTravOverflow;
}

}  // namespace gpr
