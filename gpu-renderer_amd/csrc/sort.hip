// sort.hip -- LDS-staged radix sort of 30-bit Morton keys into (key, index)
// pairs for gfx950 (the std::sort of LBVH_build, src/renderer/bvh.h:207-208,
// defined stable: equal keys keep index order).
//
// The top digit has between 1 and 11 bits so that a bucket holds ~1K keys
// (10 bits = 1024 buckets at 1M keys).  Three launches, each reading its
// input once:
//   hist     8192-key tiles (1024 threads): LDS histograms of the top digit
//            per 4096-key half; every tile reserves its range inside each
//            bucket with one returning atomic per non-empty digit and hands
//            the halves consecutive sub-ranges.  The atomics make the tile
//            order inside a bucket arbitrary, which is harmless because the
//            last pass orders the full (key, index).
//   scatter  4096-key halves: bucket bases from an LDS scan of the digit
//            totals, LDS-atomic ranks inside the half, 8-byte
//            (key<<32 | index) composites staged in LDS in digit order and
//            written with consecutive lanes on consecutive slots.  The two
//            halves of a histogram tile run on one XCD.
//   local    one block per bucket staged in LDS: MSD split on the next 10 key
//            bits into 1024 sub-buckets (~1 element each), then every
//            composite is placed by its rank among its sub-bucket.  Buckets
//            with a crowded sub-bucket, or too large for LDS, use LDS bitonic
//            chunk sorts and pairwise merges.
// Every bucket block zeroes its digit total once it has read it, so each call
// leaves the totals cleared for the next call in stream order: eager calls and
// replays of captured graphs (of any number of sorts) may interleave freely.
#include "dev_math.h"
#include "gpr_internal.h"

#include <algorithm>
#include <mutex>
#include <string>

namespace gpr {
namespace {

constexpr int kMaxDigits = 2048;    // top-digit buckets (11 bits at most)
constexpr int kSubs = 2;            // scatter sub-tiles per histogram tile
constexpr int kScatterThreads = 1024;
// Reservation groups: tile t reserves its range of each bucket from the
// counter of group t % kGroups, so each counter takes ntiles / kGroups
// returning atomics instead of ntiles (device-scope atomics on one address
// serialise); a bucket holds the groups' ranges one after another.
constexpr int kGroups = 4;
constexpr uint32_t kAccWords = kMaxDigits * kGroups;
// counter of digit d in group g: one group's counters side by side, so the
// reservations of one wave touch two cache lines (64 digits)
GPR_D constexpr uint32_t acc_at(uint32_t d, uint32_t g) { return g * kMaxDigits + d; }
constexpr int kTile = 8192;         // keys per histogram tile
constexpr int kSubTile = kTile / kSubs;  // keys per scatter tile
constexpr int kTileThreads = 1024;
constexpr int kLocalThreads = 512;
constexpr int kFastCap = 1536;      // bucket size of the LDS fast path (A and B)
constexpr int kSub = 1024;          // sub-buckets of the LDS MSD split
constexpr int kSubMax = 48;         // largest sub-bucket sorted by one thread
constexpr uint32_t kWin = 4;        // sub-bucket entries a rank reads at once
static_assert(kSub % kLocalThreads == 0, "sub-bucket scan: whole counts per thread");

// Top-digit width: ~1K keys per bucket, between 1 and 11 bits.
uint32_t digit_bits(uint32_t n) {
    uint32_t b = 1;
    while (b < 11u && ((uint64_t)n >> b) > 1024u) ++b;
    return b;
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Inclusive scan of a wave64: Hillis-Steele inside each row of 16 lanes with
// row_shr DPP moves (1, 2, 4, 8), then the row totals carried across rows with
// row_bcast:15 (rows 1 and 3) and row_bcast:31 (rows 2 and 3).  Lanes a move
// does not reach receive 0.
template <int CTRL, int ROWS>
GPR_D uint32_t dpp_or_zero(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xf, false);
}
GPR_D uint32_t wave_scan_incl(uint32_t v) {
    v += dpp_or_zero<0x111, 0xf>(v);  // row_shr:1
    v += dpp_or_zero<0x112, 0xf>(v);  // row_shr:2
    v += dpp_or_zero<0x114, 0xf>(v);  // row_shr:4
    v += dpp_or_zero<0x118, 0xf>(v);  // row_shr:8
    v += dpp_or_zero<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
    v += dpp_or_zero<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
    return v;
}

// inclusive scan of one value per thread over a block of NT threads (wave64)
template <int NT>
GPR_D uint32_t scan_incl(uint32_t v, uint32_t* sh, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    v = wave_scan_incl(v);
    if (lane == 63) sh[wave] = v;
    __syncthreads();
    uint32_t add = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) {
        const uint32_t s = sh[w];
        add += w < wave ? s : 0u;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return v + add;
}

// inclusive scans of two values per thread over a block of NT threads,
// sharing their barriers
template <int NT>
GPR_D void scan_incl2(uint32_t& a, uint32_t& b, uint32_t* sh /* 2 * NT / 64 */) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    a = wave_scan_incl(a);
    b = wave_scan_incl(b);
    if (lane == 63) {
        sh[2 * wave] = a;
        sh[2 * wave + 1] = b;
    }
    __syncthreads();
    uint32_t adda = 0, addb = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) {
        adda += w < wave ? sh[2 * w] : 0u;
        addb += w < wave ? sh[2 * w + 1] : 0u;
    }
    __syncthreads();
    a += adda;
    b += addb;
}

// ---- hist: 8192-key tiles, per-half histograms, one reservation per digit
__global__ __launch_bounds__(kTileThreads) void k_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                       uint32_t shift, uint32_t nd,
                                                       uint32_t* __restrict__ acc,
                                                       uint32_t* __restrict__ off) {
    __shared__ uint32_t h[kSubs * kMaxDigits];
    // the keys are requested first (16-byte loads where the array allows,
    // thread t taking keys 4t..4t+3 of each 4096-key round), the LDS histograms
    // cleared while they are in flight
    constexpr int kRound = 4 * kTileThreads;  // keys of one round of 16-byte loads
    static_assert(kRound % kSubTile == 0 || kSubTile % kRound == 0, "load rounds tile the sub-tiles");
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    const bool vec = (reinterpret_cast<uintptr_t>(keys) & 15u) == 0;
    uint32_t k[kTile / kTileThreads];
#pragma unroll
    for (int x = 0; x < kTile / kRound; ++x) {
        const uint64_t i0 = base + (uint64_t)x * kRound + 4u * threadIdx.x;
        if (vec && i0 + 3 < n) {
            const uint4 v = *reinterpret_cast<const uint4*>(keys + i0);
            k[4 * x] = v.x;
            k[4 * x + 1] = v.y;
            k[4 * x + 2] = v.z;
            k[4 * x + 3] = v.w;
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) k[4 * x + c] = i0 + c < n ? keys[i0 + c] : 0xffffffffu;
        }
    }
    for (uint32_t i = threadIdx.x; i < kSubs * nd; i += kTileThreads) h[i] = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kTile / kTileThreads; ++j) {
        // k[4x + c] is key x * kRound + 4 * threadIdx.x + c of the tile
        const uint32_t q = ((uint32_t)(j / 4) * kRound + 4u * threadIdx.x) / kSubTile;
        if (k[j] != 0xffffffffu) atomicAdd(&h[q * nd + ((k[j] >> shift) & (nd - 1))], 1u);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < nd; d += kTileThreads) {
        uint32_t q[kSubs], tot = 0;
#pragma unroll
        for (int x = 0; x < kSubs; ++x) {
            q[x] = h[x * nd + d];
            tot += q[x];
        }
        uint32_t r = tot ? atomicAdd(&acc[acc_at(d, blockIdx.x % kGroups)], tot) : 0u;
#pragma unroll
        for (int x = 0; x < kSubs; ++x) {
            __hip_atomic_store(&off[((uint64_t)blockIdx.x * kSubs + x) * nd + d], r, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            r += q[x];
        }
    }
}

// ---- scatter: 4096-key sub-tiles, composites into their bucket.
// Workgroups are dispatched round-robin over the 8 XCDs (block b on XCD b%8)
// and the L2 of one XCD is not coherent with the others', so a cache line
// written from two XCDs reaches HBM as two partial writes.  The two halves
// of one histogram tile own consecutive sub-ranges of every bucket; mapping
// them onto one XCD lets its L2 merge their writes into whole lines.
constexpr int kXcds = 8;
GPR_D uint32_t sub_of_block(uint32_t b) {
    const uint32_t g = b / kXcds, x = b % kXcds;
    const uint32_t t = (g / kSubs) * kXcds + x;
    return t * kSubs + g % kSubs;
}

__global__ __launch_bounds__(kScatterThreads) void k_scatter(const uint32_t* __restrict__ keys, uint32_t n,
                                                             uint32_t shift, uint32_t nd,
                                                             const uint32_t* __restrict__ acc,
                                                             const uint32_t* __restrict__ off,
                                                             uint2* __restrict__ bases,
                                                             unsigned long long* __restrict__ tmp) {
    constexpr int NT = kScatterThreads, KPT = kSubTile / NT;
    constexpr int kPer = kMaxDigits / NT;  // digits per thread at most
    static_assert(kSubTile % NT == 0 && kMaxDigits % NT == 0, "whole keys and digits per thread");
    __shared__ uint32_t gpos[kMaxDigits];  // global slot of digit d's first element, minus its local start
    __shared__ uint32_t lcnt[kMaxDigits];  // digit counts of the sub-tile, then local starts
    __shared__ unsigned long long stage[kSubTile];
    __shared__ uint32_t sh[2 * NT / 64];
    const uint32_t sub = sub_of_block(blockIdx.x);
    const uint64_t b0 = (uint64_t)sub * kSubTile;
    if (b0 >= n) return;  // padding blocks of the XCD mapping
    // every global load is issued before the first dependent use
    uint32_t k[KPT];
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
        const uint64_t i = b0 + j * NT + threadIdx.x;
        k[j] = i < n ? keys[i] : 0u;
    }
    // thread t owns digits per*t .. per*t + per - 1 (nd is a power of two:
    // either a multiple of NT or below it, then one digit for threads < nd)
    const uint32_t per = nd >= (uint32_t)NT ? nd / NT : 1u;
    const bool owns = threadIdx.x * per < nd;
    const uint32_t d0 = threadIdx.x * per;
    // digit d's total over all groups, and the part of it in groups before
    // this tile's (the tile's range of bucket d follows those groups' ranges)
    const uint32_t grp = (sub / kSubs) % kGroups;
    const uint32_t* offs = off + (uint64_t)sub * nd;
    uint32_t a[kPer], p[kPer], o[kPer];
#pragma unroll
    for (int x = 0; x < kPer; ++x) {
        a[x] = p[x] = o[x] = 0u;
        if (owns && (uint32_t)x < per) {
            uint32_t g[kGroups];
#pragma unroll
            for (int y = 0; y < kGroups; ++y) g[y] = acc[acc_at(d0 + x, y)];
#pragma unroll
            for (int y = 0; y < kGroups; ++y) {
                p[x] += (uint32_t)y < grp ? g[y] : 0u;
                a[x] += g[y];
            }
            o[x] = offs[d0 + x];
        }
    }
    for (uint32_t d = threadIdx.x; d < nd; d += NT) lcnt[d] = 0;
    __syncthreads();
    // rank of every key among the sub-tile's keys of its digit
    uint32_t r[KPT];
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
        const uint64_t i = b0 + j * NT + threadIdx.x;
        r[j] = i < n ? atomicAdd(&lcnt[(k[j] >> shift) & (nd - 1)], 1u) : 0u;
    }
    __syncthreads();
    uint32_t l[kPer], ex = 0, lex = 0;
#pragma unroll
    for (int x = 0; x < kPer; ++x) {
        l[x] = owns && (uint32_t)x < per ? lcnt[d0 + x] : 0u;
        ex += a[x];
        lex += l[x];
    }
    // bucket bases (exclusive scan of the digit totals) and the sub-tile's
    // digit starts (exclusive scan of its digit counts) in one pass
    const uint32_t ta = ex, tl = lex;
    scan_incl2<NT>(ex, lex, sh);
    ex -= ta;
    lex -= tl;
    if (owns) {
#pragma unroll
        for (int x = 0; x < kPer; ++x) {
            if ((uint32_t)x < per) {
                lcnt[d0 + x] = lex;
                gpos[d0 + x] = ex + p[x] + o[x] - lex;
                if (sub == 0) bases[d0 + x] = make_uint2(ex, a[x]);
                ex += a[x];
                lex += l[x];
            }
        }
    }
    __syncthreads();
    // stage the composites in digit order, then write them with consecutive
    // lanes on consecutive slots of each digit's run
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
        const uint64_t i = b0 + j * NT + threadIdx.x;
        if (i < n)
            stage[lcnt[(k[j] >> shift) & (nd - 1)] + r[j]] = ((unsigned long long)k[j] << 32) | (unsigned long long)i;
    }
    __syncthreads();
    const uint32_t m = (uint32_t)min<uint64_t>(kSubTile, n - b0);
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
        const uint32_t q = j * NT + threadIdx.x;
        if (q < m) {
            const unsigned long long w = stage[q];
            __hip_atomic_store(&tmp[gpos[(uint32_t)(w >> (32 + shift)) & (nd - 1)] + q], w, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

GPR_D void bitonic_lds(unsigned long long* s, uint32_t P) {
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < (P >> 1); i += kLocalThreads) {
                const uint32_t a = ((i & ~(j - 1)) << 1) | (i & (j - 1));
                const uint32_t b = a + j;
                const unsigned long long x = s[a], y = s[b];
                if ((x > y) == ((a & k) == 0)) {
                    s[a] = y;
                    s[b] = x;
                }
            }
            __syncthreads();
        }
    }
}

GPR_D uint32_t lower_bound_u64(const unsigned long long* a, uint32_t n, unsigned long long v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Sorts x[0, n) (composites) into skeys / sidx[0, n): LDS bitonic sorts of
// 2048-entry chunks, then pairwise merges through y[0, n) (global scratch).
// The general fallback: any size, any key distribution.
GPR_D void sort_range_merge(unsigned long long* x, unsigned long long* y, uint32_t n, unsigned long long* buf,
                            uint32_t* __restrict__ skeys, uint32_t* __restrict__ sidx) {
    constexpr uint32_t kChunk = 2048;
    for (uint32_t c0 = 0; c0 < n; c0 += kChunk) {
        const uint32_t m = min(kChunk, n - c0);
        for (uint32_t i = threadIdx.x; i < kChunk; i += kLocalThreads) buf[i] = i < m ? x[c0 + i] : ~0ull;
        __syncthreads();
        bitonic_lds(buf, kChunk);
        for (uint32_t i = threadIdx.x; i < m; i += kLocalThreads) x[c0 + i] = buf[i];
        __syncthreads();
    }
    for (uint32_t w = kChunk; w < n; w <<= 1) {
        for (uint32_t e = threadIdx.x; e < n; e += kLocalThreads) {
            const uint32_t r0 = (e / (2 * w)) * (2 * w);
            const uint32_t mid = min(r0 + w, n), end = min(r0 + 2 * w, n);
            const unsigned long long v = x[e];
            const uint32_t o = e < mid ? (e - r0) + lower_bound_u64(x + mid, end - mid, v)
                                       : (e - mid) + lower_bound_u64(x + r0, mid - r0, v);
            y[r0 + o] = v;
        }
        __syncthreads();
        unsigned long long* t = x;
        x = y;
        y = t;
    }
    for (uint32_t i = threadIdx.x; i < n; i += kLocalThreads) {
        const unsigned long long v = x[i];
        skeys[i] = (uint32_t)(v >> 32);
        sidx[i] = (uint32_t)v;
    }
    __syncthreads();
}

// First sub-bucket start >= pos (starts are non-decreasing; an empty
// sub-bucket starts where the next one does), or n.
GPR_D uint32_t first_start_at(const uint32_t* start, uint32_t nsub, uint32_t n, uint32_t pos) {
    uint32_t lo = 0, hi = nsub;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (start[mid] < pos) lo = mid + 1;
        else hi = mid;
    }
    return lo < nsub ? start[lo] : n;
}

constexpr uint32_t kWinStep = 1536;   // window starts fall on sub-bucket starts past multiples of this
constexpr uint32_t kRankMax = 1024;   // largest sub-bucket placed by rank (O(size) LDS reads per entry)
static_assert(kWinStep + kRankMax <= 2 * kFastCap, "a window must fit the LDS buffer");

// A bucket of n composites at src (any order, any size): split by the
// sub-bucket digit (key bits below the top digit) into dst = the bucket's
// range of the scratch, then cut the split into windows of whole
// consecutive sub-buckets (each starts at the first sub-bucket start at or
// past a multiple of kWinStep, so it holds at most kWinStep + kRankMax - 1
// entries when no sub-bucket exceeds kRankMax) and place every entry of a
// window by its rank among its sub-bucket, all in LDS.  A window with a
// sub-bucket above kRankMax is sorted by sort_range_merge, with src (free
// after the split) as its scratch.
GPR_D void large_bucket(unsigned long long* src, unsigned long long* dst, uint32_t n, uint32_t sub_shift,
                        uint32_t nsub, uint32_t* __restrict__ skeys, uint32_t* __restrict__ sidx,
                        unsigned long long* buf, uint32_t* cnt, uint32_t* start, uint32_t* sh, uint32_t* flag) {
    // global reads in batches of kBatch per thread, all in flight before their first use
    constexpr uint32_t kBatch = 8;
    for (uint32_t s = threadIdx.x; s < kSub; s += kLocalThreads) cnt[s] = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += kBatch * kLocalThreads) {
        unsigned long long v[kBatch];
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j) {
            const uint32_t i = base + j * kLocalThreads + threadIdx.x;
            v[j] = i < n ? src[i] : 0ull;
        }
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j)
            if (base + j * kLocalThreads + threadIdx.x < n) atomicAdd(&cnt[(uint32_t)(v[j] >> sub_shift) & (nsub - 1)], 1u);
    }
    __syncthreads();
    constexpr int kQ = kSub / kLocalThreads;
    uint32_t cq[kQ], sq = 0;
#pragma unroll
    for (int k = 0; k < kQ; ++k) {
        cq[k] = cnt[kQ * threadIdx.x + k];
        sq += cq[k];
    }
    uint32_t tot;
    uint32_t run = scan_incl<kLocalThreads>(sq, sh, &tot) - sq;
    uint32_t* fill = reinterpret_cast<uint32_t*>(buf);  // buf is free during the split
#pragma unroll
    for (int k = 0; k < kQ; ++k) {
        start[kQ * threadIdx.x + k] = run;
        fill[kQ * threadIdx.x + k] = 0;
        run += cq[k];
    }
    __syncthreads();
    for (uint32_t base = 0; base < n; base += kBatch * kLocalThreads) {
        unsigned long long v[kBatch];
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j) {
            const uint32_t i = base + j * kLocalThreads + threadIdx.x;
            v[j] = i < n ? src[i] : 0ull;
        }
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j) {
            if (base + j * kLocalThreads + threadIdx.x < n) {
                const uint32_t sb = (uint32_t)(v[j] >> sub_shift) & (nsub - 1);
                dst[start[sb] + atomicAdd(&fill[sb], 1u)] = v[j];
            }
        }
    }
    __syncthreads();  // the split is in dst, seen by every thread of the block
    const uint32_t nw = (n + kWinStep - 1) / kWinStep;
    for (uint32_t w = 0; w < nw; ++w) {
        const uint32_t lo = first_start_at(start, nsub, n, w * kWinStep);
        const uint32_t hi = first_start_at(start, nsub, n, (w + 1) * kWinStep);
        const uint32_t m = hi - lo;
        if (m == 0) continue;
        if (threadIdx.x == 0) *flag = m > 2u * kFastCap ? 1u : 0u;
        __syncthreads();
        if (*flag == 0) {
            constexpr uint32_t kW = (2 * kFastCap + kLocalThreads - 1) / kLocalThreads;
            unsigned long long v[kW];
#pragma unroll
            for (uint32_t j = 0; j < kW; ++j) {
                const uint32_t i = j * kLocalThreads + threadIdx.x;
                v[j] = i < m ? dst[lo + i] : 0ull;
            }
#pragma unroll
            for (uint32_t j = 0; j < kW; ++j) {
                const uint32_t i = j * kLocalThreads + threadIdx.x;
                if (i < m) {
                    buf[i] = v[j];
                    if (cnt[(uint32_t)(v[j] >> sub_shift) & (nsub - 1)] > kRankMax) *flag = 1u;
                }
            }
        }
        __syncthreads();
        if (*flag != 0) {
            sort_range_merge(dst + lo, src + lo, m, buf, skeys + lo, sidx + lo);
            continue;
        }
        for (uint32_t j = threadIdx.x; j < m; j += kLocalThreads) {
            const unsigned long long v = buf[j];
            const uint32_t sb = (uint32_t)(v >> sub_shift) & (nsub - 1);
            const uint32_t b = start[sb] - lo, c = cnt[sb];
            // 8 independent LDS reads per step: the loop is bound by LDS
            // latency, not by the compares
            uint32_t rank = 0, x = 0;
            for (; x + 8 <= c; x += 8) {
                unsigned long long a[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) a[u] = buf[b + x + u];
#pragma unroll
                for (int u = 0; u < 8; ++u) rank += a[u] < v ? 1u : 0u;
            }
            for (; x < c; ++x) rank += buf[b + x] < v ? 1u : 0u;
            skeys[start[sb] + rank] = (uint32_t)(v >> 32);
            sidx[start[sb] + rank] = (uint32_t)v;
        }
        __syncthreads();  // buf is reloaded by the next window
    }
}

// ---- local: one bucket per block, LDS MSD split + rank inside each sub-bucket
__global__ __launch_bounds__(kLocalThreads) void k_local(uint32_t shift, uint32_t* __restrict__ acc,
                                                         const uint2* __restrict__ bases,
                                                         unsigned long long* __restrict__ tmp,
                                                         unsigned long long* __restrict__ alt,
                                                         uint32_t* __restrict__ skeys,
                                                         uint32_t* __restrict__ sidx) {
    __shared__ unsigned long long buf[2 * kFastCap];  // A | B
    __shared__ uint32_t cnt[kSub];
    __shared__ uint32_t start[kSub];
    __shared__ uint32_t sh[kLocalThreads / 64];
    __shared__ uint32_t crowded;
    const uint32_t d = blockIdx.x;
    // the bucket's base and size, as the scatter pass computed them
    const uint2 bn = bases[d];
    const uint32_t b0 = bn.x, n = bn.y;
    // the reservation counters start from zero in the next call
    if (threadIdx.x < kGroups) acc[acc_at(d, threadIdx.x)] = 0u;
    if (n == 0) return;
    unsigned long long* src = tmp + b0;
    const uint32_t sub_bits = min(10u, shift);
    const uint32_t nsub = 1u << sub_bits;
    const uint32_t sub_shift = 32 + shift - sub_bits;
    if (n <= (uint32_t)kFastCap) {
        unsigned long long* A = buf;
        unsigned long long* B = buf + kFastCap;
        // all loads in flight before the first LDS use
        constexpr int kPer = (kFastCap + kLocalThreads - 1) / kLocalThreads;
        unsigned long long v[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const uint32_t i = j * kLocalThreads + threadIdx.x;
            v[j] = i < n ? src[i] : 0ull;
        }
        for (uint32_t s = threadIdx.x; s < kSub; s += kLocalThreads) cnt[s] = 0;
        if (threadIdx.x == 0) crowded = 0;
        __syncthreads();
        uint32_t r[kPer], sb[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const uint32_t i = j * kLocalThreads + threadIdx.x;
            sb[j] = (uint32_t)(v[j] >> sub_shift) & (nsub - 1);
            r[j] = i < n ? atomicAdd(&cnt[sb[j]], 1u) : 0u;
        }
        __syncthreads();
        constexpr int kQ = kSub / kLocalThreads;  // sub-bucket counts scanned per thread
        uint32_t cq[kQ], sq = 0;
#pragma unroll
        for (int k = 0; k < kQ; ++k) {
            cq[k] = cnt[kQ * threadIdx.x + k];
            sq += cq[k];
            if (cq[k] > kSubMax) crowded = 1;
        }
        uint32_t tot;
        uint32_t run = scan_incl<kLocalThreads>(sq, sh, &tot) - sq;
#pragma unroll
        for (int k = 0; k < kQ; ++k) {
            start[kQ * threadIdx.x + k] = run;
            run += cq[k];
        }
        __syncthreads();
        if (!crowded) {
            // sub-bucket order (arrival order inside each sub-bucket) ...
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const uint32_t i = j * kLocalThreads + threadIdx.x;
                if (i < n) A[start[sb[j]] + r[j]] = v[j];
            }
            __syncthreads();
            // ... then every composite's rank among its sub-bucket (all distinct)
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const uint32_t i = j * kLocalThreads + threadIdx.x;
                if (i < n) {
                    // rank among the sub-bucket: its first kWin entries are read
                    // together (independent LDS loads); longer sub-buckets continue
                    const uint32_t b = start[sb[j]], c = cnt[sb[j]];
                    uint32_t rank = 0;
#pragma unroll
                    for (uint32_t x = 0; x < kWin; ++x) {
                        const unsigned long long a = A[min(b + x, (uint32_t)kFastCap - 1)];
                        rank += (x < c && a < v[j]) ? 1u : 0u;
                    }
                    for (uint32_t x = kWin; x < c; ++x) rank += A[b + x] < v[j] ? 1u : 0u;
                    B[b + rank] = v[j];
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const uint32_t i = j * kLocalThreads + threadIdx.x;
                if (i < n) {
                    const unsigned long long w = B[i];
                    __hip_atomic_store(&skeys[b0 + i], (uint32_t)(w >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&sidx[b0 + i], (uint32_t)w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            return;
        }
    }
    // Oversized bucket, or a crowded sub-bucket: a second MSD split on the
    // same next key bits, through the scratch `alt`, then windows of
    // consecutive sub-buckets staged in LDS and placed by rank.
    large_bucket(src, alt + b0, n, sub_shift, nsub, skeys + b0, sidx + b0, buf, cnt, start, sh, &crowded);
}

struct Layout {
    uint32_t* acc;
    uint2* bases;     // per bucket: first slot, size
    uint32_t* off;
    unsigned long long* tmp;
    unsigned long long* alt;
};
// Byte offsets of the scratch arrays for n keys, each on a 256-B boundary
// (constexpr: the no-allocation entry point sizes its fixed array with it).
struct Offsets {
    size_t acc, bases, off, tmp, alt, total;
};
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr Offsets offsets(uint32_t n) {
    const uint64_t ntiles = n > (uint32_t)kTile ? ((uint64_t)n + kTile - 1) / kTile : 1;
    const uint64_t nk = n ? n : 1;
    Offsets o{};
    o.acc = 0;
    o.bases = align256(o.acc + 4ull * kAccWords);
    o.off = align256(o.bases + 8ull * kMaxDigits);
    o.tmp = align256(o.off + 4ull * kMaxDigits * kSubs * ntiles);
    o.alt = align256(o.tmp + 8ull * nk);
    o.total = o.alt + 8ull * nk + 256;
    return o;
}
size_t layout(uint32_t n, char* base, Layout* L) {
    const Offsets o = offsets(n);
    if (L) {
        L->acc = reinterpret_cast<uint32_t*>(base + o.acc);
        L->bases = reinterpret_cast<uint2*>(base + o.bases);
        L->off = reinterpret_cast<uint32_t*>(base + o.off);
        L->tmp = reinterpret_cast<unsigned long long*>(base + o.tmp);
        L->alt = reinterpret_cast<unsigned long long*>(base + o.alt);
    }
    return o.total;
}

// zero fill of n words (graph-capturable, unlike a memset node)
__global__ __launch_bounds__(1024) void k_clear_words(uint32_t* __restrict__ p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

}  // namespace

size_t sort_scratch_bytes(uint32_t n) { return layout(n, nullptr, nullptr); }

int sort_prepare(void* scratch, size_t bytes, hipStream_t st) {
    // the digit totals (the first array of the layout) start cleared; afterwards
    // every call clears them after use
    GPR_HIP(hipMemsetAsync(scratch, 0, std::min<size_t>(bytes, 4ull * kAccWords), st));
    return GPR_OK;
}

int sort_pairs(const uint32_t* keys, uint32_t* skeys, uint32_t* sidx, uint32_t n, void* scratch,
               size_t scratch_bytes, hipStream_t st) {
    if (n == 0) return GPR_OK;
    if (scratch_bytes < sort_scratch_bytes(n)) return fail(GPR_ERR_INVALID, "sort scratch too small");
    Layout L;
    layout(n, reinterpret_cast<char*>(scratch), &L);
    const uint32_t dbits = digit_bits(n);
    const uint32_t nd = 1u << dbits, shift = 30 - dbits;
    const uint32_t ntiles = div_up(n, kTile);
    hipLaunchKernelGGL(k_hist, dim3(ntiles), dim3(kTileThreads), 0, st, keys, n, shift, nd, L.acc, L.off);
    // grid padded to whole groups of kXcds tiles for the XCD mapping
    const uint32_t nsub_blocks = div_up(ntiles, kXcds) * kXcds * kSubs;
    hipLaunchKernelGGL(k_scatter, dim3(nsub_blocks), dim3(kScatterThreads), 0, st, keys, n, shift, nd,
                       L.acc, L.off, L.bases, L.tmp);
    hipLaunchKernelGGL(k_local, dim3(nd), dim3(kLocalThreads), 0, st, shift, L.acc, L.bases, L.tmp,
                       L.alt, skeys, sidx);
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

}  // namespace gpr

// ---- the sort without allocation ------------------------------------------
// A fixed device array holds the scratch of sorts of up to kNoAllocMaxKeys
// keys, so gpr_sort_pairs_device_noalloc calls neither hipMalloc nor hipFree.
// Each GPU holds its own copy of the array until the process ends (18.9 MB);
// every call on a GPU shares it.
namespace {
constexpr uint32_t kNoAllocMaxKeys = 1u << 20;
constexpr size_t kNoAllocBytes = gpr::offsets(kNoAllocMaxKeys).total;
constexpr int kMaxDevices = 16;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
// The array's device address on the current device, looked up once per device
// with that device current.  In host code `scratch` names a host-side stand-in,
// never the device memory: only this address is passed to kernels.
int noalloc_scratch(void** out) {
    int dev = 0;
    GPR_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return gpr::fail(GPR_ERR_INVALID, "gpr_sort_pairs_device_noalloc: device index outside 0..15");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        GPR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return gpr::fail(GPR_ERR_HIP, "gpr_sort_pairs_device_noalloc: no device address for the scratch array");
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return GPR_OK;
}
}  // namespace

int gpr_sort_pairs_device_noalloc(const uint32_t* keys, uint32_t* sk, uint32_t* si, uint32_t n) {
    using namespace gpr;
    if (n == 0) return GPR_OK;
    if (!keys || !sk || !si) return fail(GPR_ERR_INVALID, "gpr_sort_pairs_device_noalloc: null pointer");
    if (n > kNoAllocMaxKeys)
        return fail(GPR_ERR_INVALID, "gpr_sort_pairs_device_noalloc: n = " + std::to_string(n) +
                                         " exceeds the fixed scratch's maximum of " +
                                         std::to_string(kNoAllocMaxKeys) + " keys");
    if (sort_scratch_bytes(n) > kNoAllocBytes)
        return fail(GPR_ERR_INVALID, "gpr_sort_pairs_device_noalloc: scratch layout exceeds the fixed array");
    void* p = nullptr;
    GPR_TRY(noalloc_scratch(&p));
    // the digit totals start at zero whatever an earlier call left in the array
    hipLaunchKernelGGL(k_clear_words, dim3(2), dim3(1024), 0, stream(), reinterpret_cast<uint32_t*>(p),
                       (uint32_t)kAccWords);
    GPR_HIP(hipGetLastError());
    return sort_pairs(keys, sk, si, n, p, kNoAllocBytes, stream());
}
