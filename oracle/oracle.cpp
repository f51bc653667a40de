// oracle.cpp -- CPU restatement of the reference wavefront path-tracing core
// (slongle/GPU-Renderer).  TEST INFRASTRUCTURE ONLY: the product library never
// links, loads or calls this code; tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it as the checker / CPU baseline.
//
// Written from the reference's semantics (cited file:line per function), not
// copied.  Float op order follows the reference source left to right and the
// file is compiled with -ffp-contract=off, so the GPU kernels (which use the
// same op order) can be compared bit for bit.
//
// Pinning status of each part (what checks it besides the GPU):
//   morton / left_shift3      closed-form KATs + brute-force bit loop   (tests/test_oracle_build.py)
//   sort                      sortedness + permutation invariants        (tests/test_oracle_build.py)
//   lbvh topology             brute-force split search vs orc_lbvh_ranges (tests/test_oracle_build.py)
//   refit boxes               exact min/max vs numpy union               (tests/test_oracle_build.py)
//   intersect / trace         FP64 brute force over all triangles        (tests/test_oracle_trace.py)
//   tea / pcg32               published TEA and pcg32-demo vectors       (tests/test_oracle_rng.py)
//   sin / cos                 libm double within 1 ulp                   (tests/test_oracle_rng.py)
//   camera / transforms       numpy inverse + pinhole geometry           (tests/test_oracle_camera.py)
//   bsdfs                     closed forms, reciprocity, chi-square      (tests/test_oracle_bsdf.py)
//   integrator                furnace closed form, energy bounds         (tests/test_oracle_render.py)

#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr float kPi = 3.14159265358979323846f;     // fwd.h:6 PI
constexpr float kInvPi = 0.31830988618379067154f;  // fwd.h:7 INV_PI

// ---------------------------------------------------------------- vectors --
// float3 semantics of the reference's vector helpers: component-wise ops,
// dot = x*x' + y*y' + z*z' (left to right), host min/max = a<b?a:b.
struct f3 {
    float x, y, z;
};
inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
inline f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
inline f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline f3 operator*(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
inline f3 operator/(f3 a, f3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
inline f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float hmin(float a, float b) { return a < b ? a : b; }
inline float hmax(float a, float b) { return a > b ? a : b; }
inline f3 vmin(f3 a, f3 b) { return mk(hmin(a.x, b.x), hmin(a.y, b.y), hmin(a.z, b.z)); }
inline f3 vmax(f3 a, f3 b) { return mk(hmax(a.x, b.x), hmax(a.y, b.y), hmax(a.z, b.z)); }
// normalize = v * (1/sqrt(dot(v,v))) (host rsqrtf of the reference's helpers)
inline f3 normalize(f3 v) {
    float inv = 1.0f / std::sqrt(dot(v, v));
    return v * inv;
}
inline float length(f3 v) { return std::sqrt(dot(v, v)); }
inline f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

// RGBSpectrum (spectrum.h:4-54)
struct rgb {
    float r, g, b;
};
inline rgb rs(float c) { return rgb{c, c, c}; }
inline rgb operator+(rgb a, rgb b) { return rgb{a.r + b.r, a.g + b.g, a.b + b.b}; }
inline rgb operator-(rgb a, rgb b) { return rgb{a.r - b.r, a.g - b.g, a.b - b.b}; }
inline rgb operator*(rgb a, rgb b) { return rgb{a.r * b.r, a.g * b.g, a.b * b.b}; }
inline rgb operator*(rgb a, float s) { return rgb{a.r * s, a.g * s, a.b * s}; }
inline rgb operator*(float s, rgb a) { return rgb{s * a.r, s * a.g, s * a.b}; }
inline rgb operator/(rgb a, rgb b) { return rgb{a.r / b.r, a.g / b.g, a.b / b.b}; }
inline rgb operator/(rgb a, float s) { return rgb{a.r / s, a.g / s, a.b / s}; }
inline rgb vsqrt(rgb a) { return rgb{std::sqrt(a.r), std::sqrt(a.g), std::sqrt(a.b)}; }
inline float maxc(rgb a) { return hmax(a.r, hmax(a.g, a.b)); }
inline bool is_black(rgb a) { return a.r == 0 && a.g == 0 && a.b == 0; }
inline rgb ldc(const float* p) { return rgb{p[0], p[1], p[2]}; }

// ------------------------------------------------------------- sin / cos --
// Own sin/cos (textbook Taylor series on a quadrant-reduced argument, all in
// double, rounded once to float).  Used wherever the reference calls sin/cos
// on a float inside device code, so CPU and GPU agree bit for bit.
constexpr double kPio2Hi = 1.5707963267948966;
constexpr double kPio2Lo = 6.123233995736766e-17;
constexpr double kTwoOverPi = 0.6366197723675814;

inline void sincos_core(float x, double* s, double* c) {
    double xd = (double)x;
    double q = std::floor(xd * kTwoOverPi + 0.5);
    double r = xd - q * kPio2Hi;
    r = r - q * kPio2Lo;
    double r2 = r * r;
    // sin r = r + r^3 P(r^2), cos r = 1 + r^2 Q(r^2): Taylor coefficients to r^15 / r^14
    double ps = -1.0 / 1307674368000.0;
    ps = 1.0 / 6227020800.0 + r2 * ps;
    ps = -1.0 / 39916800.0 + r2 * ps;
    ps = 1.0 / 362880.0 + r2 * ps;
    ps = -1.0 / 5040.0 + r2 * ps;
    ps = 1.0 / 120.0 + r2 * ps;
    ps = -1.0 / 6.0 + r2 * ps;
    double sr = r + r * r2 * ps;
    double pc = -1.0 / 87178291200.0;
    pc = 1.0 / 479001600.0 + r2 * pc;
    pc = -1.0 / 3628800.0 + r2 * pc;
    pc = 1.0 / 40320.0 + r2 * pc;
    pc = -1.0 / 720.0 + r2 * pc;
    pc = 1.0 / 24.0 + r2 * pc;
    pc = -0.5 + r2 * pc;
    double cr = 1.0 + r2 * pc;
    int iq = (int)((long long)q & 3);
    switch (iq) {
        case 0: *s = sr; *c = cr; break;
        case 1: *s = cr; *c = -sr; break;
        case 2: *s = -sr; *c = -cr; break;
        default: *s = -cr; *c = sr; break;
    }
}
inline float osin(float x) { double s, c; sincos_core(x, &s, &c); return (float)s; }
inline float ocos(float x) { double s, c; sincos_core(x, &s, &c); return (float)c; }

// ================================================================ Morton ==
// LeftShift3 (bvh.h:12-25): spread the 10 low bits of x to every third bit;
// 1024 is clamped to 1023.  Restated as an explicit bit loop.
uint32_t left_shift3(uint32_t x) {
    if (x == 1024u) x = 1023u;
    uint32_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((x >> b) & 1u) << (3 * b);
    return r;
}

// float -> uint32 truncation of the key quantizer; NaN (0/0 on a flat axis)
// is defined as 0 here and on the GPU.
inline uint32_t q_to_u32(float v) {
    if (!(v == v)) return 0u;
    if (v <= 0.0f) return 0u;
    if (v >= 4294967040.0f) return 0xffffffffu;
    return (uint32_t)v;
}

// EncodeMorton3 (bvh.h:27-32): z<<2 | y<<1 | x on the truncated coordinates
uint32_t encode_morton3(f3 v) {
    return (left_shift3(q_to_u32(v.z)) << 2) | (left_shift3(q_to_u32(v.y)) << 1) |
           left_shift3(q_to_u32(v.x));
}

// Triangle::worldBound (triangle.h:96-101) = Union(AABB(p0,p1), p2)
inline void world_bound(const float* p, f3* lo, f3* hi) {
    f3 p0 = ld3(p), p1 = ld3(p + 3), p2 = ld3(p + 6);
    f3 a = vmin(p0, p1), b = vmax(p0, p1);
    *lo = vmin(a, p2);
    *hi = vmax(b, p2);
}
// AABB::centroid (aabb.h:23) = (min + max) * 0.5f
inline f3 centroid_of(const float* p) {
    f3 lo, hi;
    world_bound(p, &lo, &hi);
    return (lo + hi) * 0.5f;
}

// ================================================================== LBVH ==
struct BuildNode {
    f3 lo, hi;
    uint32_t tri;
    int32_t left, right;  // -1 for leaves
};

// recursive_build (bvh.h:143-184): equal end keys -> midpoint split
// ((l+r)>>1)-1; else first position of the maximal adjacent XOR.
int32_t recursive_build(const uint32_t* keys, const uint32_t* idx, const float* pos, uint32_t l,
                        uint32_t r, std::vector<BuildNode>& nodes) {
    if (l + 1 == r) {
        BuildNode n;
        n.tri = idx[l];
        world_bound(pos + 9ull * idx[l], &n.lo, &n.hi);
        n.left = n.right = -1;
        nodes.push_back(n);
        return (int32_t)(nodes.size() - 1);
    }
    uint32_t split = l;
    if (keys[l] == keys[r - 1]) {
        split = ((l + r) >> 1) - 1;
    } else {
        for (uint32_t i = l; i < r - 1; ++i) {
            uint32_t x_now = keys[i] ^ keys[i + 1];
            uint32_t x_split = keys[split] ^ keys[split + 1];
            if (x_now > x_split) split = i;
        }
    }
    int32_t lc = recursive_build(keys, idx, pos, l, split + 1, nodes);
    int32_t rc = recursive_build(keys, idx, pos, split + 1, r, nodes);
    BuildNode n;
    n.left = lc;
    n.right = rc;
    n.tri = 0;
    n.lo = vmin(nodes[lc].lo, nodes[rc].lo);  // Union (aabb.h:37-46)
    n.hi = vmax(nodes[lc].hi, nodes[rc].hi);
    nodes.push_back(n);
    return (int32_t)(nodes.size() - 1);
}

// flatten_nodes (bvh.h:121-141): DFS pre-order, left child = idx+1.
uint32_t flatten(const std::vector<BuildNode>& bn, int32_t cur, orc_node* out, uint32_t* count) {
    uint32_t me = (*count)++;
    const BuildNode& n = bn[cur];
    out[me].bmin[0] = n.lo.x; out[me].bmin[1] = n.lo.y; out[me].bmin[2] = n.lo.z;
    out[me].bmax[0] = n.hi.x; out[me].bmax[1] = n.hi.y; out[me].bmax[2] = n.hi.z;
    if (n.left < 0) {
        out[me].idx = n.tri;
        out[me].leaf = 1;
    } else {
        flatten(bn, n.left, out, count);
        out[me].idx = flatten(bn, n.right, out, count);
        out[me].leaf = 0;
    }
    return me;
}

void ranges_rec(const uint32_t* keys, uint32_t l, uint32_t r, uint32_t* lo, uint32_t* hi,
                uint32_t* count) {
    uint32_t me = (*count)++;
    lo[me] = l;
    hi[me] = r;
    if (l + 1 == r) return;
    uint32_t split = l;
    if (keys[l] == keys[r - 1]) {
        split = ((l + r) >> 1) - 1;
    } else {
        for (uint32_t i = l; i < r - 1; ++i)
            if ((keys[i] ^ keys[i + 1]) > (keys[split] ^ keys[split + 1])) split = i;
    }
    ranges_rec(keys, l, split + 1, lo, hi, count);
    ranges_rec(keys, split + 1, r, lo, hi, count);
}

// ============================================================= traversal ==
// triangle_intersect (pathtracer_kernel.h:119-152): Moller-Trumbore with float
// dots widened to double, double invDet; reject det==0, u<0|u>1, v<0|u+v>1,
// t<tMin|t>tMax (t==tMax accepted).
inline bool intersect_ref(f3 p0, f3 p1, f3 p2, f3 o, f3 d, float tmin, float tmax, float* th,
                          float* uh, float* vh) {
    f3 e1 = p1 - p0;
    f3 e2 = p2 - p0;
    f3 P = cross(d, e2);
    double det = (double)dot(P, e1);
    if (det == 0) return false;
    double inv_det = 1.0 / det;
    f3 T = o - p0;
    double u = (double)dot(P, T) * inv_det;
    if (u < 0.0 || u > 1.0) return false;
    f3 Q = cross(T, e1);
    double v = (double)dot(Q, d) * inv_det;
    if (v < 0.0 || u + v > 1.0) return false;
    double t = (double)dot(Q, e2) * inv_det;
    if (t < (double)tmin || t > (double)tmax) return false;
    *th = (float)t;
    *uh = (float)u;
    *vh = (float)v;
    return true;
}

// aabb_intersect (pathtracer_kernel.h:154-173)
inline bool aabb_ref(const orc_node& nd, f3 o, float tmin_r, float tmax_r, f3 inv, const int neg[3]) {
    const float* b[2] = {nd.bmin, nd.bmax};
    float tmin = (b[neg[0]][0] - o.x) * inv.x;
    float tmax = (b[1 - neg[0]][0] - o.x) * inv.x;
    float tymin = (b[neg[1]][1] - o.y) * inv.y;
    float tymax = (b[1 - neg[1]][1] - o.y) * inv.y;
    if (tmin > tymax || tymin > tmax) return false;
    tmin = hmax(tmin, tymin);
    tmax = hmin(tmax, tymax);
    float tzmin = (b[neg[2]][2] - o.z) * inv.z;
    float tzmax = (b[1 - neg[2]][2] - o.z) * inv.z;
    if (tmin > tzmax || tzmin > tmax) return false;
    tmin = hmax(tmin, tzmin);
    tmax = hmin(tmax, tzmax);
    return (tmin < tmax_r) && (tmax > tmin_r);
}

// trace_kernel / trace_shadow_kernel (pathtracer_kernel.h:175-308): right
// child first; a leaf hit shrinks tMax; any-hit returns on the first hit.
// Misses are reported as {t=inf, id=-1, u=0, v=0}.
inline orc_hit trace_one(const orc_node* nodes, const float* pos, const orc_ray& r, bool any,
                         uint64_t* nv, uint64_t* nt) {
    orc_hit hit{INFINITY, -1, 0.f, 0.f};
    f3 o = ld3(r.o), d = ld3(r.d);
    float tmin = r.tmin, tmax = r.tmax;
    f3 inv = mk(1 / d.x, 1 / d.y, 1 / d.z);
    int neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
    int cur = 0, sp = 0;
    int stack[64];
    uint64_t visits = 0, tests = 0;
    while (true) {
        const orc_node& nd = nodes[cur];
        ++visits;
        if (aabb_ref(nd, o, tmin, tmax, inv, neg)) {
            if (nd.leaf & 0xff) {
                const float* p = pos + 9ull * nd.idx;
                float t, u, v;
                ++tests;
                if (intersect_ref(ld3(p), ld3(p + 3), ld3(p + 6), o, d, tmin, tmax, &t, &u, &v)) {
                    tmax = t;
                    hit.t = t;
                    hit.u = u;
                    hit.v = v;
                    hit.id = (int32_t)nd.idx;
                    if (any) break;
                }
                if (sp == 0) break;
                cur = stack[--sp];
            } else {
                stack[sp++] = cur + 1;
                cur = (int)nd.idx;
            }
        } else {
            if (sp == 0) break;
            cur = stack[--sp];
        }
    }
    if (nv) *nv += visits;
    if (nt) *nt += tests;
    return hit;
}

// ================================================================== RNG ===
// InitRandom (sampling.h:31-46): TEA with the fixed key, `backoff` rounds.
const uint32_t kTeaKey[4] = {0xa341316cu, 0xc8013ea4u, 0xad90777du, 0x7e95761eu};
void tea(uint32_t* pv0, uint32_t* pv1, const uint32_t k[4], uint32_t rounds) {
    uint32_t v0 = *pv0, v1 = *pv1, s0 = 0;
    for (uint32_t n = 0; n < rounds; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + k[0]) ^ (v1 + s0) ^ ((v1 >> 5) + k[1]);
        v1 += ((v0 << 4) + k[2]) ^ (v0 + s0) ^ ((v0 >> 5) + k[3]);
    }
    *pv0 = v0;
    *pv1 = v1;
}
inline uint32_t init_random(uint32_t a, uint32_t b, uint32_t backoff = 16) {
    tea(&a, &b, kTeaKey, backoff);
    return a;
}
// NextRandom (sampling.h:48-54): Numerical-Recipes LCG, 24-bit float in [0,1)
inline float next_random(uint32_t& seed) {
    seed = 1664525u * seed + 1013904223u;
    return float(seed & 0x00FFFFFFu) / float(0x01000000);
}
// PCG32 (O'Neill 2014, XSH-RR over a 64-bit LCG).
constexpr uint64_t kPcgMult = 6364136223846793005ull;
inline uint32_t pcg_next(uint64_t& state, uint64_t inc) {
    uint64_t old = state;
    state = old * kPcgMult + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
inline void pcg_seed(uint64_t& state, uint64_t& inc, uint64_t initstate, uint64_t initseq) {
    state = 0u;
    inc = (initseq << 1u) | 1u;
    pcg_next(state, inc);
    state += initstate;
    pcg_next(state, inc);
}
inline float pcg_float(uint64_t& state, uint64_t inc) {
    return float(pcg_next(state, inc) >> 8) * (1.0f / 16777216.0f);
}

// ============================================================ transforms ==
struct M4 {
    float m[4][4];
};
M4 ident() {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = (i == j) ? 1.f : 0.f;
    return r;
}
// Matrix4x4::operator* (transform.h:41-50)
M4 mul(const M4& a, const M4& b) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] +
                        a.m[i][3] * b.m[3][j];
    return r;
}
// Inverse (transform.cpp:30-90): Gauss-Jordan with full pivoting.
M4 inverse(const M4& a) {
    int indxc[4], indxr[4];
    int ipiv[4] = {0, 0, 0, 0};
    float minv[4][4];
    std::memcpy(minv, a.m, sizeof(minv));
    for (int i = 0; i < 4; ++i) {
        int irow = 0, icol = 0;
        float big = 0.f;
        for (int j = 0; j < 4; ++j) {
            if (ipiv[j] != 1) {
                for (int k = 0; k < 4; ++k) {
                    if (ipiv[k] == 0) {
                        if (std::fabs(minv[j][k]) >= big) {
                            big = std::fabs(minv[j][k]);
                            irow = j;
                            icol = k;
                        }
                    }
                }
            }
        }
        ++ipiv[icol];
        if (irow != icol)
            for (int k = 0; k < 4; ++k) std::swap(minv[irow][k], minv[icol][k]);
        indxr[i] = irow;
        indxc[i] = icol;
        float pivinv = 1.f / minv[icol][icol];
        minv[icol][icol] = 1.f;
        for (int j = 0; j < 4; ++j) minv[icol][j] *= pivinv;
        for (int j = 0; j < 4; ++j) {
            if (j != icol) {
                float save = minv[j][icol];
                minv[j][icol] = 0;
                for (int k = 0; k < 4; ++k) minv[j][k] -= minv[icol][k] * save;
            }
        }
    }
    for (int j = 3; j >= 0; --j) {
        if (indxr[j] != indxc[j])
            for (int k = 0; k < 4; ++k) std::swap(minv[k][indxr[j]], minv[k][indxc[j]]);
    }
    M4 r;
    std::memcpy(r.m, minv, sizeof(minv));
    return r;
}
struct Xf {
    M4 m, inv;
};
// Transform::operator* (transform.h:106-108)
Xf xmul(const Xf& a, const Xf& b) { return Xf{mul(a.m, b.m), mul(b.inv, a.inv)}; }
// Scale / Translate / Perspective (transform.cpp:105-259)
Xf scale(float x, float y, float z) {
    Xf t{ident(), ident()};
    t.m.m[0][0] = x; t.m.m[1][1] = y; t.m.m[2][2] = z;
    t.inv.m[0][0] = (float)(1. / x); t.inv.m[1][1] = (float)(1. / y); t.inv.m[2][2] = (float)(1. / z);
    return t;
}
Xf translate(float x, float y, float z) {
    Xf t{ident(), ident()};
    t.m.m[0][3] = x; t.m.m[1][3] = y; t.m.m[2][3] = z;
    t.inv.m[0][3] = -x; t.inv.m[1][3] = -y; t.inv.m[2][3] = -z;
    return t;
}
Xf perspective(float fov, float n, float f) {
    float inv_tan = (float)(1. / (double)std::tan(fov * 0.5f));
    M4 p = ident();
    p.m[0][0] = inv_tan;
    p.m[1][1] = inv_tan;
    p.m[2][2] = f / (f - n);
    p.m[2][3] = -f * n / (f - n);
    p.m[3][2] = 1;
    p.m[3][3] = 0;
    return Xf{p, inverse(p)};
}
// Transform::transformPoint / transformVector (transform.h:162-183)
inline f3 xf_point(const M4& m, f3 p) {
    float x = m.m[0][0] * p.x + m.m[0][1] * p.y + m.m[0][2] * p.z + m.m[0][3];
    float y = m.m[1][0] * p.x + m.m[1][1] * p.y + m.m[1][2] * p.z + m.m[1][3];
    float z = m.m[2][0] * p.x + m.m[2][1] * p.y + m.m[2][2] * p.z + m.m[2][3];
    float w = m.m[3][0] * p.x + m.m[3][1] * p.y + m.m[3][2] * p.z + m.m[3][3];
    if (w == 1) return mk(x, y, z);
    return mk(x, y, z) / w;
}
inline f3 xf_vector(const M4& m, f3 v) {
    float x = m.m[0][0] * v.x + m.m[0][1] * v.y + m.m[0][2] * v.z;
    float y = m.m[1][0] * v.x + m.m[1][1] * v.y + m.m[1][2] * v.z;
    float z = m.m[2][0] * v.x + m.m[2][1] * v.y + m.m[2][2] * v.z;
    return mk(x, y, z);
}
// Camera::setup (camera.h:21-37), non-free (Mitsuba raster) mode.
struct Cam {
    M4 r2c, c2w;
    bool free_mode;
    f3 eye, dir, up;
    float fovx, fovy;
    float w, h;
};
Cam make_camera(const orc_scene* s) {
    Cam c;
    c.free_mode = s->cam_free != 0;
    c.eye = mk(s->eye[0], s->eye[1], s->eye[2]);
    c.dir = mk(s->dir[0], s->dir[1], s->dir[2]);
    c.up = mk(s->up[0], s->up[1], s->up[2]);
    c.fovx = s->fov_x;
    c.fovy = s->fov_y;
    c.w = (float)s->width;
    c.h = (float)s->height;
    std::memcpy(c.c2w.m, s->c2w, sizeof(float) * 16);
    float fov = s->fov_deg / 180 * kPi;  // degToRad (fwd.h:25-28)
    float aspect = float(s->width) / s->height;
    Xf c2r = xmul(xmul(xmul(scale((float)s->width, (float)s->height, 1),
                             scale(-0.5f, (float)(0.5 * aspect), 1)),
                        translate(-1, (float)(1. / aspect), 0)),
                   perspective(fov, 1e-4f, 10000.f));
    c.r2c = c2r.inv;  // Inverse(Transform) swaps m and mInv (transform.h:137-139)
    return c;
}
// Camera::generateRay (camera.h:86-119).
inline void gen_ray(const Cam& c, float px, float py, f3* o, f3* d) {
    if (c.free_mode) {
        // free mode (camera.h:89-107)
        f3 dir = normalize(c.dir);
        f3 horizontal_axis = normalize(cross(dir, c.up));
        f3 vertical_axis = normalize(cross(horizontal_axis, dir));
        f3 middle = c.eye + dir;
        f3 horizontal = horizontal_axis * std::tan(c.fovx * 0.5f);
        f3 vertical = vertical_axis * std::tan(c.fovy * 0.5f);
        f3 film = middle + horizontal * (2.f * (px / c.w) - 1.f) + vertical * (2.f * (py / c.h) - 1.f);
        *o = c.eye;
        *d = normalize(film - c.eye);
        return;
    }
    f3 pc = xf_point(c.r2c, mk(px, py, 0));
    *o = xf_point(c.c2w, mk(0.f, 0.f, 0.f));
    *d = normalize(xf_vector(c.c2w, normalize(pc)));
}

// ================================================================ BSDFs ===
// Shading helpers (sampling.h:56-181)
inline float cos_theta(f3 w) { return w.z; }
inline float abs_cos_theta(f3 w) { return std::fabs(w.z); }
inline float cos2_theta(f3 w) { return w.z * w.z; }
inline float sin2_theta(f3 w) { return hmax(0.f, 1.f - cos2_theta(w)); }
inline float sin_theta(f3 w) { return std::sqrt(sin2_theta(w)); }
inline float tan_theta(f3 w) { return sin_theta(w) / cos_theta(w); }
inline float tan2_theta(f3 w) { return sin2_theta(w) / cos2_theta(w); }
inline float clampf(float a, float l, float r) { return a > r ? r : (a < l ? l : a); }
inline float cos_phi(f3 w) {
    float st = sin_theta(w);
    return st == 0 ? 1.f : clampf(w.x / st, -1, 1);
}
inline float sin_phi(f3 w) {
    float st = sin_theta(w);
    return st == 0 ? 0.f : clampf(w.y / st, -1, 1);
}
inline float cos2_phi(f3 w) { return cos_phi(w) * cos_phi(w); }
inline float sin2_phi(f3 w) { return sin_phi(w) * sin_phi(w); }
inline bool same_hemi(f3 a, f3 b) { return a.z * b.z > 0; }

// UniformSampleTriangle / CosineSampleHemisphere (sampling.h:56-70)
inline void uniform_sample_triangle(float s0, float s1, float* u, float* v) {
    float a = std::sqrt(s0);
    *u = 1 - a;
    *v = s1 * a;
}
inline f3 cosine_sample_hemisphere(float u0, float u1) {
    float a = std::sqrt(u0);
    float b = kPi * 2 * u1;
    float z = std::sqrt(hmax(0.f, 1 - u0));
    return mk(a * ocos(b), a * osin(b), z);
}

// Fresnel (bsdfs/optical.h:43-141)
float fr_dielectric(float cos_i, float eta_i, float eta_t) {
    cos_i = hmax(-1.f, hmin(cos_i, 1.f));
    bool entering = cos_i > 0.f;
    if (!entering) {
        std::swap(eta_i, eta_t);
        cos_i = std::fabs(cos_i);
    }
    float sin_i = std::sqrt(hmax(0.f, 1 - cos_i * cos_i));
    float sin_t = eta_i / eta_t * sin_i;
    if (sin_t >= 1) return 1.f;
    float cos_t = std::sqrt(hmax(0.f, 1 - sin_t * sin_t));
    float rparl = ((eta_t * cos_i) - (eta_i * cos_t)) / ((eta_t * cos_i) + (eta_i * cos_t));
    float rperp = ((eta_i * cos_i) - (eta_t * cos_t)) / ((eta_i * cos_i) + (eta_t * cos_t));
    return (rparl * rparl + rperp * rperp) * 0.5f;
}
rgb fr_conductor(float cos_i, rgb etai, rgb etat, rgb k) {
    cos_i = hmax(-1.f, hmin(cos_i, 1.f));
    rgb eta = etat / etai;
    rgb etak = k / etai;
    float c2 = cos_i * cos_i;
    float s2 = 1.f - c2;
    rgb eta2 = eta * eta;
    rgb etak2 = etak * etak;
    rgb t0 = eta2 - etak2 - rs(s2);
    rgb a2pb2 = vsqrt(t0 * t0 + eta2 * etak2 * 4.f);
    rgb t1 = a2pb2 + rs(c2);
    rgb a = vsqrt((a2pb2 + t0) * 0.5f);
    rgb t2 = a * 2.f * cos_i;
    rgb Rs = (t1 - t2) / (t1 + t2);
    rgb t3 = a2pb2 * c2 + rs(s2 * s2);
    rgb t4 = t2 * s2;
    rgb Rp = Rs * (t3 - t4) / (t3 + t4);
    return (Rp + Rs) * 0.5f;
}

// Trowbridge-Reitz / GGX (bsdfs/microfacet.h:5-156)
inline float tr_lambda(f3 w, float ax, float ay) {
    float abs_tan = std::fabs(tan_theta(w));
    if (std::isinf(abs_tan)) return 0.;
    float alpha = std::sqrt(cos2_phi(w) * ax * ax + sin2_phi(w) * ay * ay);
    float a2t2 = (alpha * abs_tan) * (alpha * abs_tan);
    return (-1 + std::sqrt(1.f + a2t2)) / 2;
}
inline float tr_g1(f3 w, float ax, float ay) { return 1 / (1 + tr_lambda(w, ax, ay)); }
inline float tr_g(f3 wo, f3 wi, float ax, float ay) {
    return 1 / (1 + tr_lambda(wo, ax, ay) + tr_lambda(wi, ax, ay));
}
inline float tr_d(f3 wh, float ax, float ay) {
    float t2 = tan2_theta(wh);
    if (std::isinf(t2)) return 0.;
    const float cos4 = cos2_theta(wh) * cos2_theta(wh);
    float e = (cos2_phi(wh) / (ax * ax) + sin2_phi(wh) / (ay * ay)) * t2;
    return 1 / (kPi * ax * ay * cos4 * (1 + e) * (1 + e));
}
void tr_sample11(float cos_t, float U1, float U2, float* sx, float* sy) {
    if (cos_t > .9999) {
        float r = std::sqrt(U1 / (1 - U1));
        float phi = (float)(6.28318530718 * (double)U2);
        *sx = r * ocos(phi);
        *sy = r * osin(phi);
        return;
    }
    float sin_t = std::sqrt(hmax(0.f, 1.f - cos_t * cos_t));
    float tan_t = sin_t / cos_t;
    float a = 1 / tan_t;
    float G1 = 2 / (1 + std::sqrt(1.f + 1.f / (a * a)));
    float A = 2 * U1 / G1 - 1;
    float tmp = 1.f / (A * A - 1.f);
    if (tmp > 1e10) tmp = 1e10f;
    float B = tan_t;
    float D = std::sqrt(hmax(float(B * B * tmp * tmp - (A * A - B * B) * tmp), 0.f));
    float x1 = B * tmp - D;
    float x2 = B * tmp + D;
    *sx = (A < 0 || x2 > 1.f / tan_t) ? x1 : x2;
    float S;
    if (U2 > 0.5f) {
        S = 1.f;
        U2 = 2.f * (U2 - .5f);
    } else {
        S = -1.f;
        U2 = 2.f * (.5f - U2);
    }
    float z = (U2 * (U2 * (U2 * 0.27385f - 0.73369f) + 0.46341f)) /
              (U2 * (U2 * (U2 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
    *sy = S * z * std::sqrt(1.f + *sx * *sx);
}
f3 tr_sample(f3 wi, float ax, float ay, float U1, float U2) {
    f3 ws = normalize(mk(ax * wi.x, ay * wi.y, wi.z));
    float sx, sy;
    tr_sample11(cos_theta(ws), U1, U2, &sx, &sy);
    float tmp = cos_phi(ws) * sx - sin_phi(ws) * sy;
    sy = sin_phi(ws) * sx + cos_phi(ws) * sy;
    sx = tmp;
    sx = ax * sx;
    sy = ay * sy;
    return normalize(mk(-sx, -sy, 1.f));
}
f3 sample_wh(f3 wo, float u0, float u1, float ax, float ay) {
    bool flip = wo.z < 0;
    f3 wh = tr_sample(flip ? -wo : wo, ax, ay, u0, u1);
    if (flip) wh = -wh;
    return wh;
}
// Reflect / Refract (bsdfs/optical.h:11-33)
inline f3 reflect(f3 wo, f3 n) { return -wo + n * 2.f * dot(wo, n); }
inline bool refract(f3 wi, f3 n, float eta, f3* wt) {
    float cos_i = dot(n, wi);
    float sin2_i = hmax(0.f, 1 - cos_i * cos_i);
    float sin2_t = eta * eta * sin2_i;
    if (sin2_t >= 1) return false;
    float cos_t = std::sqrt(1 - sin2_t);
    *wt = -wi * eta + n * (eta * cos_i - cos_t);
    return true;
}

enum { BX_LAMBERT = 0, BX_FSPEC = 1, BX_MICRO = 3, BX_ROUGH_DIEL = 5 };
enum { P_R = 1, P_T = 2, P_DIFF = 4, P_GLOSSY = 8, P_SPEC = 16 };
struct Fres {
    int conductor;
    rgb eta_i, eta_t, k;
    rgb eval(float c) const {
        if (conductor) return fr_conductor(std::fabs(c), eta_i, eta_t, k);
        return rs(fr_dielectric(c, eta_i.r, eta_t.r));
    }
};
struct BxDF {
    int type;
    int prop;
    rgb color;
    float ior, ax, ay;
    Fres fr;
    bool delta() const { return (prop & P_SPEC) != 0; }
};

// Rough dielectric: GGX microfacet reflection + transmission (Walter et al.
// 2007), a build addition (the reference maps it to the smooth case).  The
// microfacet normal is sampled from D(wh)|cos wh| so sample() and pdf() agree;
// transmitted radiance carries the (eta_i/eta_t)^2 factor like the smooth
// case (bsdfs/specular.h:57-58).
inline float rd_eta(f3 wo, float ior) { return wo.z > 0 ? ior : 1.f / ior; }
void rough_diel_eval(const BxDF& b, f3 wo, f3 wi, rgb* f, float* pdf) {
    *f = rs(0.f);
    *pdf = 0.f;
    float co = wo.z, ci = wi.z;
    if (co == 0 || ci == 0) return;
    bool refl = co * ci > 0;
    float eta = rd_eta(wo, b.ior);
    f3 wh;
    if (refl) {
        wh = wi + wo;
        if (wh.x == 0 && wh.y == 0 && wh.z == 0) return;
        wh = normalize(wh);
    } else {
        wh = normalize(wo + wi * eta);
    }
    if (wh.z < 0) wh = -wh;
    float dwo = dot(wo, wh), dwi = dot(wi, wh);
    float F = fr_dielectric(dwo, 1.f, b.ior);
    float D = tr_d(wh, b.ax, b.ax);
    float G = tr_g(wo, wi, b.ax, b.ax);
    float pwh = D * abs_cos_theta(wh);
    if (refl) {
        if (dwo * co <= 0 || dwi * ci <= 0) return;
        *f = b.color * (D * G * F / (4 * std::fabs(co)));
        *pdf = F * pwh / (4 * std::fabs(dwo));
    } else {
        if (dwo * dwi > 0) return;
        if (dwo * co <= 0 || dwi * ci <= 0) return;
        float sd = dwo + eta * dwi;
        float jac = std::fabs(dwi) / (sd * sd);  // |dwh/dwi| without eta^2
        // f*|cos i| = (1-F) D G |wi.wh||wo.wh| / (|cos o| sd^2)   (eta^2 * (1/eta)^2 = 1)
        *f = b.color * ((1 - F) * D * G * std::fabs(dwo) * jac / std::fabs(co));
        *pdf = (1 - F) * pwh * eta * eta * jac;
    }
}
void rough_diel_sample(const BxDF& b, f3 wo, float u0, float u1, float u2, f3* wi, rgb* f,
                       float* pdf) {
    *f = rs(0.f);
    *pdf = 0.f;
    if (wo.z == 0) return;
    // wh ~ D(wh)|cos wh|: tan^2 = a^2 u0/(1-u0), phi = 2 pi u1
    float a = b.ax;
    float t2 = a * a * u0 / (1 - u0);
    float ct = 1 / std::sqrt(1 + t2);
    float st = std::sqrt(hmax(0.f, 1 - ct * ct));
    float phi = kPi * 2 * u1;
    f3 wh = mk(st * ocos(phi), st * osin(phi), ct);
    float dwo = dot(wo, wh);
    // only microfacets facing wo scatter it; a sample whose direction lands on
    // the wrong side for its lobe has no counterpart in eval/pdf and is dropped
    if (dwo * wo.z <= 0) return;
    float F = fr_dielectric(dwo, 1.f, b.ior);
    f3 w;
    if (u2 < F) {
        w = reflect(wo, wh);
        if (w.z * wo.z <= 0) return;
    } else {
        float eta = rd_eta(wo, b.ior);
        f3 n = wh;
        if (dwo < 0) n = -n;
        if (!refract(wo, n, 1.f / eta, &w)) return;
        if (w.z * wo.z >= 0) return;
    }
    *wi = w;
    rough_diel_eval(b, wo, w, f, pdf);
}

// BxDF dispatch (bsdfs/bxdf.h:73-157) with lambert.h, specular.h, microfacet.h
void bx_eval(const BxDF& b, f3 wo, f3 wi, rgb* f) {
    *f = rs(0.f);
    if (b.type == BX_LAMBERT) {
        if (!same_hemi(wo, wi)) return;
        *f = b.color * kInvPi * std::fabs(wi.z);
    } else if (b.type == BX_FSPEC) {
        *f = rs(0.f);
    } else if (b.type == BX_MICRO) {
        float co = abs_cos_theta(wo), ci = abs_cos_theta(wi);
        f3 wh = wi + wo;
        if (ci == 0 || co == 0) return;
        if (wh.x == 0 && wh.y == 0 && wh.z == 0) return;
        wh = normalize(wh);
        rgb F = b.fr.eval(dot(wi, wh));
        *f = b.color * tr_d(wh, b.ax, b.ay) * tr_g(wo, wi, b.ax, b.ay) * F / (4 * co);
    } else if (b.type == BX_ROUGH_DIEL) {
        float pdf;
        rough_diel_eval(b, wo, wi, f, &pdf);
    }
}
void bx_pdf(const BxDF& b, f3 wo, f3 wi, float* pdf) {
    *pdf = 0.f;
    if (b.type == BX_LAMBERT) {
        if (!same_hemi(wo, wi)) return;
        *pdf = std::fabs(wi.z) * kInvPi;
    } else if (b.type == BX_FSPEC) {
        *pdf = 0;
    } else if (b.type == BX_MICRO) {
        if (!same_hemi(wo, wi)) return;
        f3 wh = normalize(wo + wi);
        *pdf = tr_d(wh, b.ax, b.ay) * abs_cos_theta(wh) / (4 * dot(wo, wh));
    } else if (b.type == BX_ROUGH_DIEL) {
        rgb f;
        rough_diel_eval(b, wo, wi, &f, pdf);
    }
}
void bx_sample(const BxDF& b, float u0, float u1, float u2, f3 wo, f3* wi, rgb* f, float* pdf) {
    if (b.type == BX_LAMBERT) {
        *f = rs(0.f);
        *pdf = 0.f;
        *wi = cosine_sample_hemisphere(u0, u1);
        if (wo.z < 0) wi->z *= -1;
        *f = b.color * kInvPi * std::fabs(wi->z);
        *pdf = std::fabs(wi->z) * kInvPi;
    } else if (b.type == BX_FSPEC) {
        // FresnelSpecularSample (bsdfs/specular.h:29-71)
        float F = fr_dielectric(wo.z, 1, b.ior);
        if (u0 < F) {
            *wi = mk(-wo.x, -wo.y, wo.z);
            *pdf = F;
            *f = b.color * F;
        } else {
            bool entering = wo.z > 0;
            float ei = entering ? 1 : b.ior;
            float et = entering ? b.ior : 1;
            f3 n = mk(0, 0, 1);
            if (wo.z < 0) n = -n;
            if (!refract(wo, n, ei / et, wi)) {
                *f = rs(0.f);
                return;
            }
            rgb ft = b.color * (1 - F);
            ft = ft * ((ei * ei) / (et * et));
            *pdf = 1 - F;
            *f = ft;
        }
    } else if (b.type == BX_MICRO) {
        // MicrofacetReflectSample (bsdfs/microfacet.h:203-226); the pdf is
        // D|cos h|/(4 wo.wh) although wh is drawn from the visible normals.
        *f = rs(0.f);
        *pdf = 0.f;
        if (wo.z == 0) return;
        f3 wh = sample_wh(wo, u0, u1, b.ax, b.ay);
        *wi = reflect(wo, wh);
        if (!same_hemi(wo, *wi)) return;
        *pdf = tr_d(wh, b.ax, b.ay) * abs_cos_theta(wh) / (4 * dot(wo, wh));
        bx_eval(b, wo, *wi, f);
    } else if (b.type == BX_ROUGH_DIEL) {
        rough_diel_sample(b, wo, u0, u1, u2, wi, f, pdf);
    }
}

// Frame (sampling.h:124-163)
struct Frame {
    f3 s, t, n;
};
inline Frame make_frame(f3 n) {
    Frame fr;
    fr.n = n;
    if (std::fabs(n.x) > std::fabs(n.y)) {
        float inv = 1.f / std::sqrt(n.x * n.x + n.z * n.z);
        fr.s = mk(n.z * inv, 0, -n.x * inv);
    } else {
        float inv = 1.f / std::sqrt(n.y * n.y + n.z * n.z);
        fr.s = mk(0, n.z * inv, -n.y * inv);
    }
    fr.t = cross(fr.s, n);
    return fr;
}
inline f3 to_world(const Frame& fr, f3 v) { return normalize(fr.s * v.x + fr.t * v.y + fr.n * v.z); }
inline f3 to_local(const Frame& fr, f3 v) {
    return normalize(mk(dot(v, fr.s), dot(v, fr.t), dot(v, fr.n)));
}

// BSDF (bsdf.h:9-190)
struct BSDF {
    Frame fr;
    BxDF bx[2];
    uint32_t n;
};
BSDF make_bsdf(const orc_material& m, f3 ns) {
    BSDF b;
    b.fr = make_frame(ns);
    b.n = 0;
    BxDF x{};
    switch (m.type) {
        case 0:  // MATERIAL_DIFFUSE
            x.type = BX_LAMBERT; x.prop = P_R | P_DIFF; x.color = ldc(m.color);
            b.bx[b.n++] = x;
            break;
        case 1:  // MATERIAL_SPECULAR
            x.type = BX_FSPEC; x.prop = P_R | P_T | P_SPEC; x.color = ldc(m.color); x.ior = m.ior;
            b.bx[b.n++] = x;
            break;
        case 2:  // MATERIAL_ROUGH_CONDUCTOR
            x.type = BX_MICRO; x.prop = P_R | P_GLOSSY; x.color = ldc(m.color);
            x.ax = m.alpha_x; x.ay = m.alpha_y;
            x.fr = Fres{1, ldc(m.eta_i), ldc(m.eta_t), ldc(m.k)};
            b.bx[b.n++] = x;
            break;
        case 3:  // MATERIAL_ROUGH_DIELECTRIC (build addition)
            x.type = BX_ROUGH_DIEL; x.prop = P_R | P_T | P_GLOSSY; x.color = ldc(m.color);
            x.ior = m.ior; x.ax = m.alpha_x; x.ay = m.alpha_x;
            b.bx[b.n++] = x;
            break;
        default: {  // MATERIAL_ROUGH_PLASTIC: Lambert + GGX with dielectric Fresnel
            x.type = BX_LAMBERT; x.prop = P_R | P_DIFF; x.color = ldc(m.color);
            b.bx[b.n++] = x;
            BxDF y{};
            y.type = BX_MICRO; y.prop = P_R | P_GLOSSY; y.color = ldc(m.color1);
            y.ax = m.alpha_x; y.ay = m.alpha_y;
            y.fr = Fres{0, ldc(m.eta_i), ldc(m.eta_t), rs(0.f)};
            b.bx[b.n++] = y;
            break;
        }
    }
    return b;
}
inline bool bsdf_delta(const BSDF& b) {
    for (uint32_t i = 0; i < b.n; ++i)
        if (!b.bx[i].delta()) return false;
    return true;
}
inline bool lobe_match(const BxDF& x, bool refl) {
    return (refl && (x.prop & P_R)) || (!refl && (x.prop & P_T));
}
rgb bsdf_eval(const BSDF& b, f3 wo, f3 wi) {
    rgb f = rs(0.f);
    f3 lo = to_local(b.fr, wo), li = to_local(b.fr, wi);
    if (lo.z == 0) return f;
    bool refl = lo.z * li.z > 0;
    for (uint32_t i = 0; i < b.n; ++i) {
        rgb fi = rs(0.f);
        if (lobe_match(b.bx[i], refl)) bx_eval(b.bx[i], lo, li, &fi);
        f = f + fi;
    }
    return f;
}
float bsdf_pdf(const BSDF& b, f3 wo, f3 wi) {
    float pdf = 0.f;
    f3 lo = to_local(b.fr, wo), li = to_local(b.fr, wi);
    if (lo.z == 0) return pdf;
    for (uint32_t i = 0; i < b.n; ++i) {
        float p;
        bx_pdf(b.bx[i], lo, li, &p);
        pdf += p;
    }
    return pdf;  // the reference's `/ m_bxdf_num` is a discarded expression (bsdf.h:102)
}
struct BSample {
    f3 wi;
    rgb f;
    float pdf;
    bool specular;
};
BSample bsdf_sample(const BSDF& b, float u0, float u1, float u2, f3 wo) {
    BSample s;
    s.f = rs(0.f);
    s.pdf = 0.f;
    s.wi = mk(0, 0, 0);
    s.specular = false;
    f3 lo = to_local(b.fr, wo);
    f3 li = mk(0, 0, 0);
    uint32_t idx = (uint32_t)std::floor(u0 * (float)b.n);
    if (idx > b.n - 1) idx = b.n - 1;
    float v0 = u0 * (float)b.n - (float)idx;
    if (lo.z == 0) return s;
    bx_sample(b.bx[idx], v0, u1, u2, lo, &li, &s.f, &s.pdf);
    s.specular = b.bx[idx].delta();
    if (s.pdf == 0.f) return s;
    s.wi = to_world(b.fr, li);
    if (!b.bx[idx].delta()) {
        for (uint32_t i = 0; i < b.n; ++i) {
            if (i != idx) {
                float p;
                bx_pdf(b.bx[i], lo, li, &p);
                s.pdf += p;
            }
        }
    }
    if (!b.bx[idx].delta()) {
        bool refl = lo.z * li.z > 0;
        s.f = rs(0.f);
        for (uint32_t i = 0; i < b.n; ++i) {
            rgb fi = rs(0.f);
            if (lobe_match(b.bx[i], refl)) bx_eval(b.bx[i], lo, li, &fi);
            s.f = s.f + fi;
        }
    }
    return s;
}

// ============================================================ integrator ==
struct Geom {
    f3 p, ns, ng;
};
// Triangle::setupDifferential (triangle.h:104-138)
inline Geom setup_geom(const orc_scene* sc, uint32_t tri, float u, float v) {
    const float* p = sc->pos + 9ull * tri;
    f3 p0 = ld3(p), p1 = ld3(p + 3), p2 = ld3(p + 6);
    Geom g;
    g.p = p0 * (1 - u - v) + p1 * u + p2 * v;
    f3 dpdu = p0 - p2, dpdv = p1 - p2;
    g.ng = normalize(cross(dpdu, dpdv));
    if (sc->flags && (sc->flags[tri] & 1)) {
        const float* n = sc->nrm + 9ull * tri;
        g.ns = normalize(ld3(n) * (1 - u - v) + ld3(n + 3) * u + ld3(n + 6) * v);
    } else {
        g.ns = g.ng;
    }
    return g;
}
// Triangle::area (triangle.h:168-175)
inline float tri_area(const orc_scene* sc, uint32_t tri) {
    const float* p = sc->pos + 9ull * tri;
    f3 p0 = ld3(p), p1 = ld3(p + 3), p2 = ld3(p + 6);
    return length(cross(p1 - p0, p2 - p0)) * 0.5f;
}
inline rgb emission(const orc_scene* sc, uint32_t tri) { return ldc(sc->mats[sc->mat[tri]].emission); }

// float -> int of the GPU conversion instruction the reference's kernels run
// on: truncate, saturate, NaN -> 0 (a C++ cast leaves these undefined)
inline int cvt_int(float x) {
    if (std::isnan(x)) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int)x;
}
// TextureView::evalSpectrum (texture.h:37-44): uv' = uv * scale + offset,
// u2 = int(uv'.x * width) % width where width is uint32, so the int operand is
// converted to uint32 before the remainder; idx = v2 * width + u2.
inline rgb tex_eval(const orc_scene* sc, uint32_t handle, float u, float v) {
    const orc_texture& t = sc->tex[handle - 1];
    float nu = u * t.uv_scale[0] + t.uv_offset[0];
    float nv = v * t.uv_scale[1] + t.uv_offset[1];
    uint32_t u2 = (uint32_t)cvt_int(nu * (float)t.width) % t.width;
    uint32_t v2 = (uint32_t)cvt_int(nv * (float)t.height) % t.height;
    const float* px = sc->texels + 4ull * (t.offset + v2 * t.width + u2);
    return rgb{px[0], px[1], px[2]};
}
// The BxDF factories (bsdfs/bxdf.h:160-218) evaluate every material
// parameter at the hit's uv: Triangle::setupDifferential (triangle.h:126-134)
// interpolates the corner uvs when the mesh has them, else uses the barycentrics.
inline orc_material textured(const orc_scene* sc, const orc_material& m0, uint32_t tri, float bu, float bv) {
    orc_material m = m0;
    bool any = false;
    for (int k = 0; k < 8; ++k) any = any || m.tex[k] != 0;
    if (!any) return m;
    float u = bu, v = bv;
    if (sc->flags && (sc->flags[tri] & 2)) {
        const float* q = sc->uv + 6ull * tri;
        u = q[0] * (1 - bu - bv) + q[2] * bu + q[4] * bv;
        v = q[1] * (1 - bu - bv) + q[3] * bu + q[5] * bv;
    }
    float* spectra[8] = {m.color, m.color1, nullptr, nullptr, nullptr, m.eta_i, m.eta_t, m.k};
    float* scalars[8] = {nullptr, nullptr, &m.ior, &m.alpha_x, &m.alpha_y, nullptr, nullptr, nullptr};
    for (int k = 0; k < 8; ++k) {
        if (!m.tex[k]) continue;
        rgb c = tex_eval(sc, m.tex[k], u, v);
        if (spectra[k]) {
            spectra[k][0] = c.r; spectra[k][1] = c.g; spectra[k][2] = c.b;
        } else {
            *scalars[k] = c.r;  // evalFloat reads the red channel (texture.h:46-52)
        }
    }
    return m;
}
// EnvironmentLightView::Le (environment.h:16-28); spherical_theta/phi of
// helper_math (acos of the clamped z, atan2 shifted into [0, 2 pi)).  acos and
// atan2 are taken in double and rounded once.
inline rgb env_le(const orc_scene* sc, f3 o, f3 d) {
    f3 w = normalize(o + d * 100000.f - ld3(sc->env_center));
    const float* m = sc->env_w2o;
    w = normalize(f3{m[0] * w.x + m[1] * w.y + m[2] * w.z, m[3] * w.x + m[4] * w.y + m[5] * w.z,
                     m[6] * w.x + m[7] * w.y + m[8] * w.z});
    float z = std::fmax(-1.f, std::fmin(w.z, 1.f));
    float theta = (float)std::acos((double)z);
    float phi = (float)std::atan2((double)w.y, (double)w.x);
    if (phi < 0) phi = phi + 2 * 3.14159265358979323846f;
    float u = phi * 0.15915494309189533577f;  // INV_2_PI (fwd.h:25)
    float v = theta * 0.31830988618379067154f;  // INV_PI (fwd.h:24)
    const orc_texture& t = sc->tex[sc->env_tex - 1];
    int W = (int)t.width, H = (int)t.height;
    int u2 = std::min(W - 1, cvt_int(u * (float)W)), v2 = std::min(H - 1, cvt_int(v * (float)H));
    const float* px = sc->texels + 4ull * (t.offset + (uint32_t)(v2 * W + u2));
    return rgb{px[0], px[1], px[2]};
}
inline bool is_light(const orc_scene* sc, uint32_t tri) { return !is_black(emission(sc, tri)); }

// PowerHeuristic (pathtracer_kernel.h:480-486)
inline float power_heuristic(float a, float b) {
    a *= a;
    b *= b;
    return a / (a + b);
}

struct Ctx {
    const orc_scene* sc;
    const orc_node* nodes;
    const uint32_t* lights;
    uint32_t nlights;
    Cam cam;
    uint8_t* ties;  // per pixel: a closest hit of its paths had a near tie (may be null)
};

// Whether a triangle other than hit.id meets the ray at a t within 1e-6
// relative of hit.t.  Such a closest hit is decided by the visit order (the
// reference keeps the last triangle accepted, pathtracer_kernel.h:175-233), so
// a traversal in another order may pick the other triangle.
bool near_tie(const orc_node* nodes, const float* pos, const orc_ray& r, const orc_hit& hit) {
    if (hit.id < 0) return false;
    const double lo = (double)hit.t - 1e-6 * (double)hit.t;
    const float hi = (float)((double)hit.t + 1e-6 * (double)hit.t);
    f3 o = ld3(r.o), d = ld3(r.d);
    f3 inv = mk(1 / d.x, 1 / d.y, 1 / d.z);
    int neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
    int cur = 0, sp = 0;
    int stack[64];
    while (true) {
        const orc_node& nd = nodes[cur];
        if (aabb_ref(nd, o, r.tmin, hi, inv, neg)) {
            if (nd.leaf & 0xff) {
                const float* p = pos + 9ull * nd.idx;
                float t, u, v;
                if ((int32_t)nd.idx != hit.id &&
                    intersect_ref(ld3(p), ld3(p + 3), ld3(p + 6), o, d, r.tmin, hi, &t, &u, &v) && (double)t >= lo)
                    return true;
                if (sp == 0) break;
                cur = stack[--sp];
            } else {
                stack[sp++] = cur + 1;
                cur = (int)nd.idx;
            }
        } else {
            if (sp == 0) break;
            cur = stack[--sp];
        }
    }
    return false;
}

inline orc_hit trace_ray(const Ctx& c, f3 o, f3 d, float tmin, float tmax, bool any, uint32_t pix = ~0u) {
    orc_ray r;
    r.o[0] = o.x; r.o[1] = o.y; r.o[2] = o.z;
    r.d[0] = d.x; r.d[1] = d.y; r.d[2] = d.z;
    r.tmin = tmin;
    r.tmax = tmax;
    orc_hit h = trace_one(c.nodes, c.sc->pos, r, any, nullptr, nullptr);
    if (c.ties && !any && pix != ~0u && near_tie(c.nodes, c.sc->pos, r, h)) c.ties[pix] = 1;
    return h;
}

// Queue entry of the reference's RayQueue (rayqueue.h:6-47).
struct QE {
    f3 o, d;
    float tmin, tmax;
    orc_hit hit;
    rgb w;
    float pdf;
    bool spec;
    uint32_t pix;
    uint64_t rng;  // pcg mode: path-carried PCG state
};

// One shade step (shade_hit_with_MIS_kernel, pathtracer_kernel.h:488-628).
// Appends the shadow entry / scatter entry through the callbacks and adds the
// emission term to the frame buffer.  `samples` are the 8 draws.
template <class ShadowFn, class ScatterFn>
void shade(const Ctx& c, const QE& e, uint32_t bounce, const float* samples, float* fb,
           bool pcg_rr, uint32_t rr_start, ShadowFn&& push_shadow, ScatterFn&& push_scatter) {
    const orc_scene* sc = c.sc;
    float* px = fb + 3ull * e.pix;
    if (e.hit.id != -1) {
        uint32_t tri = (uint32_t)e.hit.id;
        Geom g = setup_geom(sc, tri, e.hit.u, e.hit.v);
        f3 wo = -normalize(e.d);
        const orc_material mat = textured(sc, sc->mats[sc->mat[tri]], tri, e.hit.u, e.hit.v);
        BSDF b = make_bsdf(mat, g.ns);
        if (bounce == 0 || e.spec) {
            if (is_light(sc, tri)) {
                rgb fl = dot(wo, g.ns) > 0 ? emission(sc, tri) : rs(0.f);
                rgb add = e.w * fl;
                px[0] = px[0] + add.r; px[1] = px[1] + add.g; px[2] = px[2] + add.b;
            }
        }
        if (!bsdf_delta(b) && c.nlights > 0) {
            // sample_lights (pathtracer_kernel.h:323-336)
            uint32_t ln = c.nlights;
            uint32_t lid = (uint32_t)(samples[0] * ln);
            if (lid > ln - 1) lid = ln - 1;
            uint32_t ltri = c.lights[lid];
            float lu, lv;
            uniform_sample_triangle(samples[1], samples[2], &lu, &lv);
            Geom lg = setup_geom(sc, ltri, lu, lv);
            float lpdf = 1.f / tri_area(sc, ltri);
            lpdf /= ln;
            f3 wi = lg.p - g.p;
            float d2 = dot(wi, wi);
            float dd = std::sqrt(d2);
            wi = wi / dd;
            lpdf *= d2 / std::fabs(dot(wi, lg.ns));
            if (std::isinf(lpdf)) lpdf = 0.f;
            rgb fl = dot(-wi, lg.ns) > 0 ? emission(sc, ltri) : rs(0.f);
            rgb fb_ = bsdf_eval(b, wo, wi);
            float bpdf = bsdf_pdf(b, wo, wi);
            float wgt = power_heuristic(lpdf, bpdf);
            if (lpdf > 0.f) {
                rgb ow = e.w * fl * fb_ * wgt / lpdf;
                if (!is_black(ow)) {
                    QE s;
                    s.o = g.p - e.d * 1.0e-4f;
                    s.d = lg.p - s.o;
                    s.tmin = 0.f;
                    s.tmax = 0.9999f;
                    s.w = ow;
                    s.spec = true;
                    s.pix = e.pix;
                    s.pdf = 0.f;
                    s.rng = 0;
                    push_shadow(s);
                }
            }
        }
        {
            BSample bs = bsdf_sample(b, samples[5], samples[6], samples[4], wo);
            rgb ow = bs.pdf == 0.f ? rs(0.f) : e.w * bs.f / bs.pdf;
            if (pcg_rr && bounce >= rr_start && !is_black(ow)) {
                // unbiased roulette (build addition): survive with 1-q, reweight
                float q = hmax(0.05f, 1 - maxc(ow));
                if (samples[7] < q)
                    ow = rs(0.f);
                else
                    ow = ow / (1 - q);
            }
            if (!is_black(ow)) {
                QE s;
                s.o = g.p;
                s.d = bs.wi;
                s.tmin = 1e-4f;
                s.tmax = 1e38f;
                s.w = ow;
                s.pdf = bs.pdf;
                s.spec = bs.specular;
                s.pix = e.pix;
                s.rng = e.rng;
                push_scatter(s);
            }
        }
    } else if (sc->has_env) {
        rgb le = sc->env_tex ? env_le(sc, e.o, e.d) : ldc(sc->env);
        rgb add = e.w * le;
        px[0] = px[0] + add.r; px[1] = px[1] + add.g; px[2] = px[2] + add.b;
    }
}

// accumulate_BSDF_sampling_contribution (pathtracer_kernel.h:684-749), light part
void bsdf_light_contrib(const Ctx& c, const QE& e, float* fb) {
    const orc_scene* sc = c.sc;
    if (e.hit.id != -1 && !e.spec) {
        uint32_t tri = (uint32_t)e.hit.id;
        if (is_light(sc, tri)) {
            Geom lg = setup_geom(sc, tri, e.hit.u, e.hit.v);
            float lpdf = 1.f / tri_area(sc, tri);
            lpdf /= c.nlights;
            f3 wi = e.d;
            float d = e.hit.t;
            lpdf *= d * d / std::fabs(dot(wi, lg.ns));
            rgb fl = dot(-e.d, lg.ns) > 0 ? emission(sc, tri) : rs(0.f);
            float wgt = power_heuristic(e.pdf, lpdf);
            rgb L = e.w * fl * wgt;
            float* px = fb + 3ull * e.pix;
            px[0] = px[0] + L.r; px[1] = px[1] + L.g; px[2] = px[2] + L.b;
        }
    }
}

inline void add_shadow(const Ctx& c, const QE& s, float* fb) {
    // accumulate_light_sampling_contribution (pathtracer_kernel.h:642-664)
    orc_hit h = trace_ray(c, s.o, s.d, s.tmin, s.tmax, true);
    if (h.id == -1) {
        float* px = fb + 3ull * s.pix;
        px[0] = px[0] + s.w.r; px[1] = px[1] + s.w.g; px[2] = px[2] + s.w.b;
    }
}

// pcg mode: one camera path, depth first.  Produces exactly the same
// per-pixel additions, in the same order, as the wavefront kernels.
void render_path_pcg(const Ctx& c, const orc_render_opts* op, uint32_t pix, uint32_t sample,
                     float* fb) {
    const orc_scene* sc = c.sc;
    uint64_t st, inc;
    pcg_seed(st, inc, op->seed, ((uint64_t)sample << 32) | (uint64_t)pix);
    uint32_t x = pix % sc->width, y = pix / sc->width;
    float jx = pcg_float(st, inc);
    float jy = pcg_float(st, inc);
    QE e;
    gen_ray(c.cam, (float)x + jx, (float)y + jy, &e.o, &e.d);
    e.tmin = 0.f;
    e.tmax = 1e34f;
    e.w = rs(1.f);
    e.spec = false;
    e.pdf = 0.f;
    e.pix = pix;
    e.hit = trace_ray(c, e.o, e.d, e.tmin, e.tmax, false, pix);
    for (uint32_t bounce = 0; bounce < op->max_depth; ++bounce) {
        float smp[8];
        for (int i = 0; i < 8; ++i) smp[i] = pcg_float(st, inc);
        e.rng = st;
        bool have_shadow = false, have_scatter = false;
        QE sh, scat;
        shade(c, e, bounce, smp, fb, true, op->rr_start,
              [&](const QE& s) { sh = s; have_shadow = true; },
              [&](const QE& s) { scat = s; have_scatter = true; });
        if (have_shadow) add_shadow(c, sh, fb);
        if (!have_scatter) break;
        scat.hit = trace_ray(c, scat.o, scat.d, scat.tmin, scat.tmax, false, pix);
        bsdf_light_contrib(c, scat, fb);
        st = scat.rng;
        e = scat;
    }
}

// reference mode: the wavefront loop of PathTracer::render (pathtracer.cu:41-150)
// run serially; queue slots are filled in serial order (what a stable
// compaction produces), seeds are slot-indexed and persist across samples
// (rayqueue.h:20-47 never copies m_seed), roulette terminates through
// hit.id = -1 without reweighting (pathtracer_kernel.h:738-748).
struct RefQueues {
    std::vector<QE> q[2];
    std::vector<uint32_t> seed[2];
};
void render_sample_ref(const Ctx& c, const orc_render_opts* op, uint32_t sample, float* fb,
                       RefQueues& Q) {
    const orc_scene* sc = c.sc;
    uint32_t npix = sc->width * sc->height;
    int in = 0, sc_ = 1;
    std::vector<QE>& q0 = Q.q[in];
    q0.resize(npix);
    for (uint32_t p = 0; p < npix; ++p) {
        uint32_t x = p % sc->width, y = p / sc->width;
        uint32_t& seed = Q.seed[in][p];
        seed = init_random(p, sample);
        float jx = next_random(seed);
        float jy = next_random(seed);
        QE& e = q0[p];
        gen_ray(c.cam, (float)x + jx, (float)y + jy, &e.o, &e.d);
        e.tmin = 0.f;
        e.tmax = 1e34f;
        e.w = rs(1.f);
        e.spec = false;
        e.pix = p;
        e.pdf = 0.f;
        e.rng = 0;
    }
    for (uint32_t i = 0; i < npix; ++i)
        q0[i].hit = trace_ray(c, q0[i].o, q0[i].d, q0[i].tmin, q0[i].tmax, false);
    std::vector<QE> shadow;
    for (uint32_t bounce = 0; bounce < op->max_depth; ++bounce) {
        std::vector<QE>& qi = Q.q[in];
        if (qi.empty()) break;
        std::vector<QE>& qs = Q.q[sc_];
        qs.clear();
        shadow.clear();
        for (size_t i = 0; i < qi.size(); ++i) {
            uint32_t& seed = Q.seed[in][i];
            float smp[8];
            for (int k = 0; k < 8; ++k) smp[k] = next_random(seed);
            shade(c, qi[i], bounce, smp, fb, false, 0,
                  [&](const QE& s) { shadow.push_back(s); },
                  [&](const QE& s) { qs.push_back(s); });
        }
        for (size_t i = 0; i < shadow.size(); ++i) add_shadow(c, shadow[i], fb);
        for (size_t i = 0; i < qs.size(); ++i)
            qs[i].hit = trace_ray(c, qs[i].o, qs[i].d, qs[i].tmin, qs[i].tmax, false);
        for (size_t i = 0; i < qs.size(); ++i) {
            QE& e = qs[i];
            float u = next_random(Q.seed[sc_][i]);
            bsdf_light_contrib(c, e, fb);
            if (maxc(e.w) < 1 && bounce > 3) {
                float q = hmax(0.05f, 1 - maxc(e.w));
                if (u < q) e.hit.id = -1;
            }
        }
        std::swap(in, sc_);
    }
    Q.q[0].clear();
    Q.q[1].clear();
}

}  // namespace

// ============================================================== C ABI =====
extern "C" {

float orc_sinf(float x) { return osin(x); }
float orc_cosf(float x) { return ocos(x); }

uint32_t orc_left_shift3(uint32_t x) { return left_shift3(x); }
uint32_t orc_encode_morton3(float x, float y, float z) { return encode_morton3(mk(x, y, z)); }

// Centroid-bounds loop of LBVH_build (bvh.h:192-196)
void orc_centroid_bounds(const float* pos, uint32_t ntri, float bmin[3], float bmax[3]) {
    f3 lo = mk(INFINITY, INFINITY, INFINITY), hi = mk(-INFINITY, -INFINITY, -INFINITY);
    for (uint32_t i = 0; i < ntri; ++i) {
        f3 c = centroid_of(pos + 9ull * i);
        lo = vmin(lo, c);
        hi = vmax(hi, c);
    }
    bmin[0] = lo.x; bmin[1] = lo.y; bmin[2] = lo.z;
    bmax[0] = hi.x; bmax[1] = hi.y; bmax[2] = hi.z;
}

// Morton loop of LBVH_build (bvh.h:199-204): (c - min) / (max - min) * 1024
void orc_morton_keys(const float* pos, uint32_t ntri, uint32_t* keys) {
    float a[3], b[3];
    orc_centroid_bounds(pos, ntri, a, b);
    f3 lo = ld3(a), hi = ld3(b);
    for (uint32_t i = 0; i < ntri; ++i) {
        f3 c = centroid_of(pos + 9ull * i);
        keys[i] = encode_morton3((c - lo) / (hi - lo) * 1024.f);
    }
}

// Sort of LBVH_build (bvh.h:207-208) defined stable: ties keep index order.
void orc_sort_pairs(const uint32_t* keys, uint32_t n, uint32_t* skeys, uint32_t* sidx) {
    std::vector<uint64_t> v(n);
    for (uint32_t i = 0; i < n; ++i) v[i] = ((uint64_t)keys[i] << 32) | i;
    std::sort(v.begin(), v.end());
    for (uint32_t i = 0; i < n; ++i) {
        skeys[i] = (uint32_t)(v[i] >> 32);
        sidx[i] = (uint32_t)v[i];
    }
}

uint32_t orc_lbvh_build(const float* pos, uint32_t ntri, orc_node* nodes, uint32_t* skeys,
                        uint32_t* sidx) {
    if (ntri == 0) return 0;
    std::vector<uint32_t> keys(ntri), sk(ntri), si(ntri);
    orc_morton_keys(pos, ntri, keys.data());
    orc_sort_pairs(keys.data(), ntri, sk.data(), si.data());
    std::vector<BuildNode> bn;
    bn.reserve(2ull * ntri);
    int32_t root = recursive_build(sk.data(), si.data(), pos, 0, ntri, bn);
    uint32_t count = 0;
    std::memset(nodes, 0, sizeof(orc_node) * (2ull * ntri - 1));
    flatten(bn, root, nodes, &count);
    if (skeys) std::memcpy(skeys, sk.data(), 4ull * ntri);
    if (sidx) std::memcpy(sidx, si.data(), 4ull * ntri);
    return count;
}

void orc_lbvh_ranges(const uint32_t* skeys, uint32_t n, uint32_t* lo, uint32_t* hi) {
    if (n == 0) return;
    uint32_t count = 0;
    ranges_rec(skeys, 0, n, lo, hi, &count);
}

int orc_intersect_ref(const float* p0, const float* p1, const float* p2, const orc_ray* r, float* t,
                      float* u, float* v) {
    return intersect_ref(ld3(p0), ld3(p1), ld3(p2), ld3(r->o), ld3(r->d), r->tmin, r->tmax, t, u, v)
               ? 1
               : 0;
}

void orc_trace(const orc_node* nodes, const float* pos, uint32_t ntri, const orc_ray* rays,
               uint64_t n, int any, orc_hit* hits, uint64_t* node_visits, uint64_t* tri_tests,
               int nthreads) {
    uint64_t nv = 0, nt = 0;
    if (ntri == 0) {
        for (uint64_t i = 0; i < n; ++i) hits[i] = orc_hit{INFINITY, -1, 0.f, 0.f};
    } else {
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : nv, nt) num_threads(nthreads > 0 ? nthreads : 1)
        for (long long i = 0; i < (long long)n; ++i)
            hits[i] = trace_one(nodes, pos, rays[i], any != 0, &nv, &nt);
    }
    if (node_visits) *node_visits = nv;
    if (tri_tests) *tri_tests = nt;
}

// Brute force over all triangles.  fp64=0: the reference's intersection
// arithmetic, closest = smallest t (first index on exact ties);
// fp64=1: every operation in double (independent of the float arithmetic).
void orc_brute(const float* pos, uint32_t ntri, const orc_ray* rays, uint64_t n, int any, int fp64,
               orc_hit* hits, double* t64, int nthreads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
    for (long long i = 0; i < (long long)n; ++i) {
        const orc_ray& r = rays[i];
        orc_hit best{INFINITY, -1, 0.f, 0.f};
        double bt = INFINITY;
        for (uint32_t k = 0; k < ntri; ++k) {
            const float* p = pos + 9ull * k;
            if (!fp64) {
                float t, u, v;
                if (intersect_ref(ld3(p), ld3(p + 3), ld3(p + 6), ld3(r.o), ld3(r.d), r.tmin, r.tmax, &t, &u, &v)) {
                    if ((double)t < bt) {
                        bt = t;
                        best = orc_hit{t, (int32_t)k, u, v};
                        if (any) break;
                    }
                }
            } else {
                double e1[3], e2[3], d[3], o[3], P[3], T[3], Q[3];
                for (int a = 0; a < 3; ++a) {
                    e1[a] = (double)p[3 + a] - p[a];
                    e2[a] = (double)p[6 + a] - p[a];
                    d[a] = r.d[a];
                    o[a] = r.o[a];
                    T[a] = o[a] - p[a];
                }
                P[0] = d[1] * e2[2] - d[2] * e2[1];
                P[1] = d[2] * e2[0] - d[0] * e2[2];
                P[2] = d[0] * e2[1] - d[1] * e2[0];
                double det = P[0] * e1[0] + P[1] * e1[1] + P[2] * e1[2];
                if (det == 0) continue;
                double u = (P[0] * T[0] + P[1] * T[1] + P[2] * T[2]) / det;
                if (u < 0 || u > 1) continue;
                Q[0] = T[1] * e1[2] - T[2] * e1[1];
                Q[1] = T[2] * e1[0] - T[0] * e1[2];
                Q[2] = T[0] * e1[1] - T[1] * e1[0];
                double v = (Q[0] * d[0] + Q[1] * d[1] + Q[2] * d[2]) / det;
                if (v < 0 || u + v > 1) continue;
                double t = (Q[0] * e2[0] + Q[1] * e2[1] + Q[2] * e2[2]) / det;
                if (t < r.tmin || t > r.tmax) continue;
                if (t < bt) {
                    bt = t;
                    best = orc_hit{(float)t, (int32_t)k, (float)u, (float)v};
                    if (any) break;
                }
            }
        }
        hits[i] = best;
        if (t64) t64[i] = bt;
    }
}

void orc_tea_keyed(uint32_t* v0, uint32_t* v1, const uint32_t key[4], uint32_t rounds) {
    tea(v0, v1, key, rounds);
}
uint32_t orc_init_random(uint32_t v0, uint32_t v1, uint32_t backoff) { return init_random(v0, v1, backoff); }
float orc_next_random(uint32_t* seed) { return next_random(*seed); }
void orc_pcg32_seed(uint64_t* state, uint64_t* inc, uint64_t initstate, uint64_t initseq) {
    pcg_seed(*state, *inc, initstate, initseq);
}
uint32_t orc_pcg32_next(uint64_t* state, uint64_t inc) { return pcg_next(*state, inc); }

// LookAt (transform.cpp:220-248); m and its Gauss-Jordan inverse.
void orc_lookat(const float target[3], const float origin[3], const float up[3], float m[16],
                float minv[16]) {
    M4 c = ident();
    f3 from = ld3(origin), to = ld3(target), u = ld3(up);
    c.m[0][3] = from.x;
    c.m[1][3] = from.y;
    c.m[2][3] = from.z;
    c.m[3][3] = 1;
    f3 fwd = normalize(to - from);
    f3 left = normalize(cross(u, fwd));
    f3 rup = cross(fwd, left);
    c.m[0][0] = left.x; c.m[1][0] = left.y; c.m[2][0] = left.z; c.m[3][0] = 0;
    c.m[0][1] = rup.x; c.m[1][1] = rup.y; c.m[2][1] = rup.z; c.m[3][1] = 0;
    c.m[0][2] = fwd.x; c.m[1][2] = fwd.y; c.m[2][2] = fwd.z; c.m[3][2] = 0;
    M4 ci = inverse(c);
    std::memcpy(m, c.m, sizeof(float) * 16);
    std::memcpy(minv, ci.m, sizeof(float) * 16);
}
void orc_matrix_inverse(const float m[16], float out[16]) {
    M4 a;
    std::memcpy(a.m, m, sizeof(float) * 16);
    M4 r = inverse(a);
    std::memcpy(out, r.m, sizeof(float) * 16);
}
void orc_camera_ray(const orc_scene* s, float px, float py, orc_ray* ray) {
    Cam c = make_camera(s);
    f3 o, d;
    gen_ray(c, px, py, &o, &d);
    ray->o[0] = o.x; ray->o[1] = o.y; ray->o[2] = o.z;
    ray->d[0] = d.x; ray->d[1] = d.y; ray->d[2] = d.z;
    ray->tmin = 0.f;
    ray->tmax = 1e34f;
}

void orc_tex_eval(const orc_scene* s, uint32_t handle, float u, float v, float out[3]) {
    rgb c = tex_eval(s, handle, u, v);
    out[0] = c.r; out[1] = c.g; out[2] = c.b;
}

void orc_env_le(const orc_scene* s, const float o[3], const float d[3], float out[3]) {
    rgb c = env_le(s, ld3(o), ld3(d));
    out[0] = c.r; out[1] = c.g; out[2] = c.b;
}

void orc_bsdf_eval(const orc_material* m, const float n[3], const float wo[3], const float wi[3],
                   float f[3], float* pdf) {
    BSDF b = make_bsdf(*m, ld3(n));
    rgb r = bsdf_eval(b, ld3(wo), ld3(wi));
    f[0] = r.r; f[1] = r.g; f[2] = r.b;
    *pdf = bsdf_pdf(b, ld3(wo), ld3(wi));
}
int orc_bsdf_sample(const orc_material* m, const float n[3], const float wo[3], float u0, float u1,
                    float u2, float wi[3], float f[3], float* pdf, int* specular) {
    BSDF b = make_bsdf(*m, ld3(n));
    BSample s = bsdf_sample(b, u0, u1, u2, ld3(wo));
    wi[0] = s.wi.x; wi[1] = s.wi.y; wi[2] = s.wi.z;
    f[0] = s.f.r; f[1] = s.f.g; f[2] = s.f.b;
    *pdf = s.pdf;
    *specular = s.specular ? 1 : 0;
    return s.pdf > 0 ? 1 : 0;
}
float orc_fr_dielectric(float c, float ei, float et) { return fr_dielectric(c, ei, et); }
void orc_fr_conductor(float c, const float ei[3], const float et[3], const float k[3], float out[3]) {
    rgb r = fr_conductor(c, ldc(ei), ldc(et), ldc(k));
    out[0] = r.r; out[1] = r.g; out[2] = r.b;
}
float orc_ggx_d(const float wh[3], float ax, float ay) { return tr_d(ld3(wh), ax, ay); }
float orc_ggx_g(const float wo[3], const float wi[3], float ax, float ay) {
    return tr_g(ld3(wo), ld3(wi), ax, ay);
}
float orc_ggx_lambda(const float w[3], float ax, float ay) { return tr_lambda(ld3(w), ax, ay); }
void orc_ggx_d_n(const float* wh, uint32_t n, float ax, float ay, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = tr_d(ld3(wh + 3 * i), ax, ay);
}
void orc_ggx_sample_wh(const float wo[3], const float* u, uint32_t n, float ax, float ay, float* wh) {
    const f3 w = ld3(wo);
    for (uint32_t i = 0; i < n; ++i) {
        const f3 h = sample_wh(w, u[2 * i], u[2 * i + 1], ax, ay);
        wh[3 * i] = h.x;
        wh[3 * i + 1] = h.y;
        wh[3 * i + 2] = h.z;
    }
}

int orc_render(const orc_scene* scene, const orc_node* nodes, const orc_render_opts* opts,
               float* rgb_sum, uint32_t* sample_num, int nthreads) {
    return orc_render_ties(scene, nodes, opts, rgb_sum, sample_num, nullptr, nthreads);
}

int orc_render_ties(const orc_scene* scene, const orc_node* nodes, const orc_render_opts* opts,
                    float* rgb_sum, uint32_t* sample_num, uint8_t* tie_pixels, int nthreads) {
    if (!scene || !opts || scene->width == 0 || scene->height == 0) return -1;
    if (tie_pixels && opts->rng_mode == 0) return -1;  // pcg mode only
    std::vector<uint32_t> lights;
    for (uint32_t i = 0; i < scene->ntri; ++i)
        if (is_light(scene, i)) lights.push_back(i);
    Ctx c;
    c.sc = scene;
    c.nodes = nodes;
    c.lights = lights.data();
    c.nlights = (uint32_t)lights.size();
    c.cam = make_camera(scene);
    c.ties = tie_pixels;
    uint32_t npix = scene->width * scene->height;
    if (scene->ntri == 0) {
        // nothing to hit: only the environment term (a constant one)
        if (scene->env_tex) return -1;
        for (uint32_t s = 0; s < opts->spp; ++s)
            for (uint32_t p = 0; p < npix; ++p) {
                if (scene->has_env)
                    for (int k = 0; k < 3; ++k) rgb_sum[3ull * p + k] += scene->env[k];
                sample_num[p]++;
            }
        return 0;
    }
    if (opts->rng_mode == 0) {
        RefQueues Q;
        Q.seed[0].assign(npix, 0u);
        Q.seed[1].assign(npix, 0u);
        for (uint32_t s = 0; s < opts->spp; ++s) {
            render_sample_ref(c, opts, opts->first_sample + s, rgb_sum, Q);
            for (uint32_t p = 0; p < npix; ++p) sample_num[p]++;
        }
    } else {
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
        for (long long p = 0; p < (long long)npix; ++p) {
            for (uint32_t s = 0; s < opts->spp; ++s) {
                render_path_pcg(c, opts, (uint32_t)p, opts->first_sample + s, rgb_sum);
                sample_num[p]++;
            }
        }
    }
    return 0;
}

}  // extern "C"
