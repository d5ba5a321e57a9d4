// Stage (c): boundary points, their clusters and the cluster order.
//
//   k_point_count  one wave per row pair (y, y+1), y in 1..qh-2: the 4-bit
//                  neighbour mask of each pixel (fast/clusters.c:150-174),
//                  the points of each row counted (and the frame's cluster
//                  table cleared).
//   k_point_rows   the same masks again: each point written at its
//                  emission index with its cluster's table slot; each run
//                  of points of one cluster (id = max_root << 32 |
//                  min_root, fast/clusters.c:229) finds or inserts the id.
//   k_point_scan   the row offsets of the points.
//   k_cluster_rank one block per frame: the distinct clusters sorted by
//                  (hash bucket, id), the order of the reference's cluster
//                  table (fast/clusters.c:338-356), giving each its rank
//                  and its id.
//   k_rank_words   each point's word: its cluster's rank | its index above
//                  the rank's bits (frames of 2048 clusters or more also
//                  count each cluster's points here, k_cluster_starts scans
//                  the counts into the clusters' first points).
//   radix sort     stable LSD sort of the words, kRadixBits a pass (one
//                  pass below 2048 clusters, whose digit offsets are then
//                  the clusters' first points): stability keeps each
//                  cluster's points in emission order; the last pass
//                  writes each point itself in place.
#include <hip/hip_runtime.h>

#include "atg_internal.h"

namespace atg {

__device__ __forceinline__ uint32_t wave_incl_scan_u(uint32_t v) { return wave_scan_incl(v); }

// Exclusive scan over a 256-thread block; returns the prefix, *total the sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t wsum[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan_u(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int i = 0; i < 4; i++) {
        if (i < w) pre += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

__device__ __forceinline__ uint32_t opp(uint32_t a, uint32_t b) { return (a ^ b) == 255u ? 1u : 0u; }

// Whether a frame's radix sort moves packed words: its key bits
// (counters[kSortBits]: the cluster ranks' bits) and the bits of a point
// index (below pcap) fit one word, as they do for the cluster sort.
__device__ __forceinline__ bool sort_packed(const Frame &F) {
    return F.counters[kSortBits] + (64u - (uint32_t)__clzll((unsigned long long)F.pcap)) <= 64u;
}

constexpr uint64_t kSlotEmpty = ~0ull;  // (no cluster id: its two roots differ)
// Clusters below which the cluster sort's one pass gives their first points
// (digit d = rank d, and the end at digit ncl); from it on they are counted.
constexpr uint32_t kOnePass = 1u << kRadixBits;

// The slot of cluster `id` in the frame's table, inserted if new (its slot
// then listed in clist, counted in counters[kNClusters]); linear probing.
// A table fuller than half sets error flag 128 (the call is re-run with a
// table sized for the worst case).
__device__ uint32_t cluster_slot(const Frame &F, uint64_t id) {
    uint32_t s = (uint32_t)((id * 0x9E3779B97F4A7C15ull) >> 32) & F.hmask;
    for (uint32_t probe = 0; probe <= F.hmask; probe++, s = (s + 1) & F.hmask) {
        // a long probe: past an overflow the call is re-run with a larger
        // table, so an overfull table is not searched to its end
        if ((probe & 63u) == 63u && (__atomic_load_n(F.counters + kError, __ATOMIC_RELAXED) & 128u)) return 0;
        const uint64_t k = __atomic_load_n(&F.hkey[s], __ATOMIC_RELAXED);
        if (k == id) return s;
        if (k != kSlotEmpty) continue;
        const uint64_t old = atomicCAS((unsigned long long *)&F.hkey[s], (unsigned long long)kSlotEmpty,
                                       (unsigned long long)id);
        if (old == kSlotEmpty) {
            const uint32_t at = atomicAdd(F.counters + kNClusters, 1u);
            if (at < F.ccap) F.clist[at] = s;
            else atomicOr(F.counters + kError, 128u);
            return s;
        }
        if (old == id) return s;
    }
    atomicOr(F.counters + kError, 128u);
    return 0;
}
__device__ __forceinline__ uint64_t low_bits(uint32_t b) { return b >= 64 ? ~0ull : (1ull << b) - 1ull; }

// One pixel column of the row pair: t of rows a and b, their plab (root
// | ok << 31).
struct PixCol {
    uint32_t ta, tb, la, lb;
};

// The point mask of pixel x of row a against row b (fast/clusters.c:
// 150-174) from its column c, the columns left (l) and right (r) of it.
__device__ __forceinline__ uint32_t point_mask(const PixCol &c, const PixCol &l, const PixCol &r, int x) {
    if (x == 0) return (opp(c.ta, r.ta) && (c.la >> 31) && (r.la >> 31)) ? 1u : 0u;
    if (!(c.la >> 31)) return 0u;
    const uint32_t v = c.ta;
    const uint32_t c10 = opp(v, r.ta) & (r.la >> 31);
    const uint32_t c01 = opp(v, c.tb) & (c.lb >> 31);
    const uint32_t c11 = opp(v, r.tb) & (r.lb >> 31);
    const uint32_t cm1 = opp(v, l.tb) & (l.lb >> 31);
    // the (1,1) point of the pixel to the left is the same pair; x = 1 has
    // no pixel to the left in the C library's loop (fast/clusters.c:170-171)
    const uint32_t p11 = x > 1 ? (opp(l.ta, c.tb) & (l.la >> 31) & (c.lb >> 31)) : 0u;
    return c10 | (c01 << 1) | ((cm1 & (p11 ^ 1u)) << 2) | (c11 << 3);
}

// The columns of the neighbour lanes (the step's edge lanes take `fill`),
// and lane 63's column in every lane.
__device__ __forceinline__ PixCol col_from_left(const PixCol &c, const PixCol &fill) {
    return {wave_from_left(c.ta, fill.ta), wave_from_left(c.tb, fill.tb), wave_from_left(c.la, fill.la),
            wave_from_left(c.lb, fill.lb)};
}
__device__ __forceinline__ PixCol col_from_right(const PixCol &c, const PixCol &fill) {
    return {wave_from_right(c.ta, fill.ta), wave_from_right(c.tb, fill.tb), wave_from_right(c.la, fill.la),
            wave_from_right(c.lb, fill.lb)};
}
__device__ __forceinline__ PixCol col_of_lane63(const PixCol &c) {
    return {(uint32_t)__builtin_amdgcn_readlane((int)c.ta, 63), (uint32_t)__builtin_amdgcn_readlane((int)c.tb, 63),
            (uint32_t)__builtin_amdgcn_readlane((int)c.la, 63), (uint32_t)__builtin_amdgcn_readlane((int)c.lb, 63)};
}

// A wave per row pair (y, y + 1), 64 pixels a step, kG steps of columns
// loaded at once (each lane's column, and the column right of the group
// for lane 63 of the last step); neighbours by shuffles, the left one of
// lane 0 kept from the step before. Each step's points (at most 4 a pixel)
// are staged with their cluster ids in LDS and written out contiguously
// with their clusters' slots: the first point of each run of one cluster
// (in emission order, across steps) looks the cluster up.
__global__ __launch_bounds__(256) void k_point_rows(const Frame *frames) {
    __shared__ uint64_t st_pt[4][256], st_key[4][256];
    const Frame F = frames[blockIdx.y];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int y = 1 + blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (y > F.qh - 2) return;
    const int QW = F.qw;
    const uint8_t *a = F.t + (int64_t)y * QW, *b = a + QW;
    const uint32_t *la = F.plab + (int64_t)y * QW, *lb = la + QW;
    auto load = [&](int x) -> PixCol {
        PixCol c{0, 0, 0, 0};
        if (x < QW) {
            c.ta = a[x];
            c.tb = b[x];
            c.la = la[x];
            c.lb = lb[x];
        }
        return c;
    };
    uint32_t count = 0;
    const uint32_t base = F.poff[y];
    const uint64_t lt = (1ull << lane) - 1ull;
    uint64_t prev_id = kSlotEmpty;  // the cluster of the row pair's last written point (none yet)
    uint32_t prev_slot = 0;
    constexpr int kG = 4;
    PixCol left{0, 0, 0, 0};
    for (int g0 = 0; g0 < QW - 1; g0 += 64 * kG) {
        PixCol col[kG];
#pragma unroll
        for (int jg = 0; jg < kG; jg++) col[jg] = load(g0 + 64 * jg + lane);
        const PixCol after = load(g0 + 64 * kG);  // the column right of the group
#pragma unroll
        for (int jg = 0; jg < kG; jg++) {
            const int x0 = g0 + 64 * jg;
            if (x0 >= QW - 1) break;
            const PixCol cur = col[jg];
            const PixCol cr = jg + 1 < kG ? PixCol{(uint32_t)__builtin_amdgcn_readlane((int)col[jg + 1].ta, 0),
                                                   (uint32_t)__builtin_amdgcn_readlane((int)col[jg + 1].tb, 0),
                                                   (uint32_t)__builtin_amdgcn_readlane((int)col[jg + 1].la, 0),
                                                   (uint32_t)__builtin_amdgcn_readlane((int)col[jg + 1].lb, 0)}
                                          : after;
            const int x = x0 + lane;
            const PixCol r = col_from_right(cur, cr), l = col_from_left(cur, left);
            const uint32_t m = x < QW - 1 ? point_mask(cur, l, r, x) : 0u;
            const uint32_t c = __popc(m);
                const uint32_t inc = wave_incl_scan_u(c), tot = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
                uint32_t k = inc - c;  // this lane's first point in the step
                if (m) {
                    const int vd = 255 - 2 * (int)cur.ta;
                    const uint32_t r0 = cur.la & 0x7fffffffu;
                    for (int j = 0; j < 4; j++) {
                        if (!((m >> j) & 1u)) continue;
                        // partner (dx, dy): (1, 0), (0, 1), (-1, 1), (1, 1)
                        const int dx = j == 0 ? 1 : j == 1 ? 0 : j == 2 ? -1 : 1;
                        const int dy = j == 0 ? 0 : 1;
                        const uint32_t r1 = (j == 0 ? r.la : j == 1 ? cur.lb : j == 2 ? l.lb : r.lb) & 0x7fffffffu;
                        const uint32_t hi = r0 > r1 ? r0 : r1, lo = r0 > r1 ? r1 : r0;
                        const uint32_t px = (uint32_t)(2 * x + dx), py = (uint32_t)(2 * y + dy);
                        const uint32_t gx = (uint32_t)(dx * vd) & 0xffffu, gy = (uint32_t)(dy * vd) & 0xffffu;
                        st_pt[wv][k] = (uint64_t)px | ((uint64_t)py << 16) | ((uint64_t)gx << 32) | ((uint64_t)gy << 48);
                        st_key[wv][k] = ((uint64_t)hi << 32) | lo;
                        k++;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
                const uint32_t at0 = base + count;
                for (uint32_t q0 = 0; q0 < tot; q0 += 64) {
                    const uint32_t q = q0 + (uint32_t)lane;
                    const bool valid = q < tot;
                    const uint64_t id = valid ? st_key[wv][q] : 0ull;
                    const uint64_t before = !valid ? 0ull : q == 0 ? prev_id : st_key[wv][q - 1];
                    const bool head = valid && id != before;  // the first point of a run of its cluster
                    uint32_t slot = head ? cluster_slot(F, id) : 0u;
                    // the slot of the last head at or before each lane; lanes
                    // before the chunk's first head continue the previous run
                    const uint64_t upto = __ballot(head) & (lt | (1ull << lane));
                    const int src = upto ? 63 - __clzll((long long)upto) : 0;
                    const uint32_t from = (uint32_t)__shfl((int)slot, src, 64);
                    slot = upto ? from : prev_slot;
                    const int64_t at = (int64_t)at0 + q;
                    if (valid && at < F.pcap) {  // (beyond pcap: flagged by k_point_scan)
                        F.pts[at] = st_pt[wv][q];
                        F.sval[at] = slot;
                    }
                    const int lastl = (int)(min(tot - q0, 64u) - 1u);
                    prev_id = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(id >> 32), lastl, 64) << 32) |
                              (uint32_t)__shfl((int)(uint32_t)id, lastl, 64);
                    prev_slot = (uint32_t)__shfl((int)slot, lastl, 64);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
                count += tot;
            left = col_of_lane63(cur);
        }
    }
}

// Pass 0 of the emission on its own: a wave per row pair counts its
// points (the masks of k_point_rows) and clears its share of the frame's
// cluster table; kG steps of 64 columns loaded at once (the count's work a
// step is a few instructions: one round of reads a step was its time).
__global__ __launch_bounds__(256) void k_point_count(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int y = 1 + (int)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (y > F.qh - 2) return;
    const int QW = F.qw;
    {
        const uint32_t nrows = (uint32_t)(F.qh - 2), slots = F.hmask + 1;
        for (uint32_t i = (uint32_t)(y - 1) * 64u + (uint32_t)lane; i < slots; i += nrows * 64u) F.hkey[i] = kSlotEmpty;
    }
    const uint8_t *a = F.t + (int64_t)y * QW, *b = a + QW;
    const uint32_t *la = F.plab + (int64_t)y * QW, *lb = la + QW;
    auto load = [&](int x) -> PixCol {
        PixCol c{0, 0, 0, 0};
        if (x < QW) {
            c.ta = a[x];
            c.tb = b[x];
            c.la = la[x];
            c.lb = lb[x];
        }
        return c;
    };
    constexpr int kG = 4;
    uint32_t count = 0;
    PixCol left{0, 0, 0, 0};
    for (int g0 = 0; g0 < QW - 1; g0 += 64 * kG) {
        PixCol col[kG];
#pragma unroll
        for (int j = 0; j < kG; j++) col[j] = load(g0 + 64 * j + lane);
        const PixCol after = load(g0 + 64 * kG);  // the column right of the group
#pragma unroll
        for (int j = 0; j < kG; j++) {
            const int x0 = g0 + 64 * j;
            if (x0 >= QW - 1) break;
            const int x = x0 + lane;
            const PixCol fill = j + 1 < kG ? PixCol{(uint32_t)__builtin_amdgcn_readlane((int)col[j + 1].ta, 0),
                                                    (uint32_t)__builtin_amdgcn_readlane((int)col[j + 1].tb, 0),
                                                    (uint32_t)__builtin_amdgcn_readlane((int)col[j + 1].la, 0),
                                                    (uint32_t)__builtin_amdgcn_readlane((int)col[j + 1].lb, 0)}
                                           : after;
            const PixCol r = col_from_right(col[j], fill), l = col_from_left(col[j], left);
            const uint32_t m = x < QW - 1 ? point_mask(col[j], l, r, x) : 0u;
            count += (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan_u(__popc(m)), 63);
            left = col_of_lane63(col[j]);
        }
    }
    if (lane == 0) F.pcnt[y] = count;
}

// Row offsets of the points (one block per frame).
__global__ __launch_bounds__(256) void k_point_scan(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    uint32_t carry = 0;
    for (int y0 = 0; y0 < F.qh; y0 += 256) {
        int y = y0 + threadIdx.x;
        uint32_t c = (y >= 1 && y <= F.qh - 2) ? F.pcnt[y] : 0u;
        uint32_t tot;
        uint32_t pre = block_excl_scan(c, &tot);
        if (y < F.qh) F.poff[y] = carry + pre;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        uint32_t n = carry;
        if ((int64_t)n > F.pcap) {
            atomicOr(F.counters + kError, 2u);
            n = (uint32_t)F.pcap;
        }
        F.counters[kNPts] = n;
        F.poff[F.qh] = carry;
    }
}

// ---------------------------------------------------------------- radix sort

struct SortBufs {
    const uint64_t *kin;
    const uint32_t *vin;
    uint64_t *kout;
    uint32_t *vout;
};

__device__ __forceinline__ SortBufs sort_bufs(const Frame &F, int parity) {
    if (parity == 0) return {F.skey, F.sval, F.skey2, F.sval2};
    return {F.skey2, F.sval2, F.skey, F.sval};
}

// Block b's segment of n items, gridDim.x blocks a frame (256-item steps).
__device__ __forceinline__ void seg_of(uint32_t n, int b, uint32_t *s0, uint32_t *s1) {
    const uint32_t nb = gridDim.x;
    uint32_t seg = ((n + nb - 1) / nb + 255u) & ~255u;
    *s0 = min(n, (uint32_t)b * seg);
    *s1 = min(n, (uint32_t)(b + 1) * seg);
}

// The lanes of the wave holding the same digit as this lane (of those
// with `valid`).
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid) {
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < kRadixBits; bit++) {
        const uint64_t bb = __ballot((d >> bit) & 1u);
        peers &= ((d >> bit) & 1u) ? bb : ~bb;
    }
    return peers;
}

// The cluster sort's word of point i: its cluster's rank | i above the
// rank bits (formed where the first pass reads it, from the point's table
// slot, sval[i], and the slots' ranks).
__device__ __forceinline__ uint64_t rank_word(const Frame &F, uint32_t slot, uint32_t i, uint32_t sb) {
    return (uint64_t)F.hrank[slot] | ((uint64_t)i << sb);
}

// dyn: the pass is skipped where the frame's key (counters[kSortBits]) has
// no bits at shift; the skipped passes are the last ones. pk: the frame's
// words are packed where sort_packed says so (the key is their low
// kSortBits bits).
// Digit counts of each block's segment, rhist[block][digit]; each run of
// equal digits on adjacent lanes adds its length once (a digit that comes
// back later in the wave adds again: counts only, no order).
__global__ __launch_bounds__(256) void k_radix_hist(const Frame *frames, int which, int shift, int parity, int dyn,
                                                    int pk, int ranks) {
    __shared__ uint32_t h[kRadixBins];
    const Frame F = frames[blockIdx.y];
    const uint32_t sb = F.counters[kSortBits];
    if (dyn && shift >= (int)sb) return;
    if (F.counters[kError] & (2u | 128u)) return;  // a call to re-run: nothing to sort
    const uint64_t km = pk && sort_packed(F) ? low_bits(sb) : ~0ull;
    const uint32_t n = F.counters[which];
    const SortBufs B = sort_bufs(F, parity);
    const int t = threadIdx.x, lane = t & 63;
    for (int i = t; i < kRadixBins; i += 256) h[i] = 0;
    __syncthreads();
    uint32_t s0, s1;
    seg_of(n, blockIdx.x, &s0, &s1);
    const uint64_t lt = (1ull << lane) - 1ull;
    // four chunks a step: their slots' reads issued together, then their
    // words' (the counts do not depend on the order they are added in)
    constexpr int kG = 4;
    for (uint32_t c0 = s0; c0 < s1; c0 += 256 * kG) {
        uint64_t keys[kG];
        if (ranks) {
            uint32_t slot[kG];
#pragma unroll
            for (int g = 0; g < kG; g++) {
                const uint32_t i = c0 + (uint32_t)(g * 256 + t);
                slot[g] = i < s1 ? F.sval[i] : 0u;
            }
#pragma unroll
            for (int g = 0; g < kG; g++) {
                const uint32_t i = c0 + (uint32_t)(g * 256 + t);
                keys[g] = i < s1 ? rank_word(F, slot[g], i, sb) : 0ull;
            }
        } else {
#pragma unroll
            for (int g = 0; g < kG; g++) {
                const uint32_t i = c0 + (uint32_t)(g * 256 + t);
                keys[g] = i < s1 ? B.kin[i] : 0ull;
            }
        }
#pragma unroll
        for (int g = 0; g < kG; g++) {
            const uint32_t i = c0 + (uint32_t)(g * 256 + t);
            const bool valid = i < s1;
            const uint32_t d = valid ? (uint32_t)((keys[g] & km) >> shift) & (kRadixBins - 1) : 0xffffffffu;
            const uint32_t dl = wave_from_left(d, 0xfffffffeu);  // lane 0: a head
            const uint64_t heads = __ballot(valid && d != dl) | ~__ballot(valid);  // invalid lanes end runs
            if (valid && d != dl) {
                const uint64_t after = heads & ~(lt | (1ull << lane));
                const uint32_t end = after ? (uint32_t)(__ffsll((unsigned long long)after) - 1) : 64u;
                atomicAdd(&h[d], end - (uint32_t)lane);
            }
        }
    }
    __syncthreads();
    for (int i = t; i < kRadixBins; i += 256) F.rhist[blockIdx.x * kRadixBins + i] = h[i];
}

// Exclusive scan over a block of NT threads; *total the sum.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan_t(uint32_t v, uint32_t *total) {
    __shared__ uint32_t ws[NT / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan_u(v);
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int i = 0; i < NT / 64; i++) {
        pre += i < w ? ws[i] : 0u;
        tot += ws[i];
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

// Per frame (a block of kOffThreads): the digit counts turned into each
// block's first output position per digit, in place (digits in order,
// blocks in order within a digit). Thread t holds the digits t,
// t + kOffThreads, ...: a digit's total, its base among the digits (a
// block scan per group), then its blocks' positions.
constexpr int kOffThreads = 1024;
__global__ __launch_bounds__(kOffThreads) void k_radix_offsets(const Frame *frames, int shift, int dyn, int starts,
                                                                int nb) {
    constexpr int kPer = kRadixBins / kOffThreads;
    static_assert(kPer * kOffThreads == kRadixBins, "whole digits a thread");
    const Frame F = frames[blockIdx.y];
    if (dyn && shift >= (int)F.counters[kSortBits]) return;
    if (F.counters[kError] & (2u | 128u)) return;
    const int t = threadIdx.x;
    uint32_t carry = 0;
    for (int q = 0; q < kPer; q++) {
        const int d = q * kOffThreads + t;
        uint32_t v[kSortBlocks], tot = 0;
#pragma unroll
        for (int b = 0; b < kSortBlocks; b++) {
            v[b] = b < nb ? F.rhist[b * kRadixBins + d] : 0u;
            tot += v[b];
        }
        uint32_t all;
        uint32_t run = carry + block_excl_scan_t<kOffThreads>(tot, &all);
        carry += all;
        // a one-pass cluster sort: digit d is cluster rank d, its start
        // the cluster's first point
        if (starts && F.counters[kNClusters] < kOnePass && d <= (int)F.counters[kNClusters]) F.cstart[d] = run;
#pragma unroll
        for (int b = 0; b < kSortBlocks; b++) {
            if (b < nb) F.rhist[b * kRadixBins + d] = run;
            run += v[b];
        }
    }
}

// Stable scatter: chunks of 256 keys in order; inside a chunk, the rank of
// a key among the keys of its digit before it: within a wave by ballots
// (peers = lanes of the same digit), across waves by per-wave counts (a
// byte each: at most 64; the lane that set one clears it once the chunk's
// positions are taken). iota: the values are the keys' positions (the
// first pass of an unpacked cluster sort; its points were emitted in
// order), not read.
__global__ __launch_bounds__(256) void k_radix_scatter(const Frame *frames, int which, int shift, int parity,
                                                       int dyn, int iota, int pk, int gather, int ranks) {
    __shared__ uint32_t off[kRadixBins];
    __shared__ uint8_t wcnt[4][kRadixBins];
    const Frame F = frames[blockIdx.y];
    const uint32_t sb = F.counters[kSortBits];
    if (dyn && shift >= (int)sb) return;
    if (F.counters[kError] & (2u | 128u)) return;
    const bool packed = pk && sort_packed(F);
    const uint64_t km = packed ? low_bits(sb) : ~0ull;
    const uint32_t n = F.counters[which];
    const SortBufs B = sort_bufs(F, parity);
    // gather (the cluster sort's last pass): the point of each word to its place
    const bool to_points = gather && packed && shift + kRadixBits >= (int)sb;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int i = t; i < kRadixBins; i += 256) {
        off[i] = F.rhist[blockIdx.x * kRadixBins + i];
        for (int k = 0; k < 4; k++) wcnt[k][i] = 0;
    }
    __syncthreads();
    uint32_t s0, s1;
    seg_of(n, blockIdx.x, &s0, &s1);
    const uint64_t lt = (1ull << lane) - 1ull;
    // the chunks' loads ahead of their use: a ranked word's slot two chunks
    // ahead, each word (and value) one chunk ahead, the point a gather moves
    // at the top of its chunk (read during the ranking)
    auto slot_at = [&](uint32_t i) -> uint32_t { return ranks && i < s1 ? F.sval[i] : 0u; };
    auto word_at = [&](uint32_t i, uint32_t slot) -> uint64_t {
        if (i >= s1) return 0ull;
        return ranks ? rank_word(F, slot, i, sb) : B.kin[i];
    };
    auto val_at = [&](uint32_t i) -> uint32_t { return i < s1 && !packed ? (iota ? i : B.vin[i]) : 0u; };
    uint64_t key_n = word_at(s0 + t, slot_at(s0 + t));
    uint32_t val_n = val_at(s0 + t);
    uint32_t slot_n2 = slot_at(s0 + 256 + t);
    for (uint32_t c0 = s0; c0 < s1; c0 += 256) {
        const uint32_t i = c0 + t;
        const bool valid = i < s1;
        const uint64_t key = key_n;
        const uint32_t val = val_n;
        const uint64_t pt = valid && to_points ? F.pts[(uint32_t)(key >> sb)] : 0ull;
        key_n = word_at(i + 256, slot_n2);
        val_n = val_at(i + 256);
        slot_n2 = slot_at(i + 512);
        const uint32_t d = valid ? (uint32_t)((key & km) >> shift) & (kRadixBins - 1) : 0xffffffffu;
        // the wave's runs of equal digits on adjacent lanes; where every
        // digit of the wave is one run, a lane's rank is its place in the
        // run (each run head marks its digit with its lane: a head that reads
        // back another lane's mark shares the digit with another run, and
        // the wave then ranks by ballots over the digit bits)
        const uint32_t dl = wave_from_left(d, 0xfffffffeu);
        const bool head = valid && d != dl;
        const uint64_t hm = __ballot(head) | ~__ballot(valid);
        const uint64_t at_or_before = lt | (1ull << lane);
        uint32_t rank, cnt;
        bool lead;
        if (head) wcnt[w][d] = (uint8_t)(lane + 1);
        // (LDS fences: the read must see every head's store, not this lane's own)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        const bool clash = head && wcnt[w][d] != (uint8_t)(lane + 1);
        if (!__ballot(clash)) {
            const uint64_t upto = hm & at_or_before, after = hm & ~at_or_before;
            rank = (uint32_t)lane - (uint32_t)(63 - __clzll((long long)upto));
            cnt = (after ? (uint32_t)(__ffsll((unsigned long long)after) - 1) : 64u) - (uint32_t)lane;
            lead = head;
        } else {
            const uint64_t peers = digit_peers(d, valid);
            rank = (uint32_t)__popcll(peers & lt);
            cnt = (uint32_t)__popcll(peers);
            lead = valid && rank == 0;
        }
        if (head) wcnt[w][d] = 0;
        if (lead) wcnt[w][d] = (uint8_t)cnt;
        __syncthreads();
        if (valid) {
            uint32_t pos = off[d] + rank;
            for (int k = 0; k < w; k++) pos += wcnt[k][d];
            if (to_points) {
                F.cpts[pos] = pt;
            } else {
                B.kout[pos] = key;
                if (!packed) B.vout[pos] = val;
            }
        }
        __syncthreads();
        if (lead) {
            atomicAdd(&off[d], cnt);
            wcnt[w][d] = 0;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- clusters

// The reference's cluster order: (hash bucket, id) ascending, bucket =
// ((id * 2654435761) >> 32) & (buckets - 1) (fast/clusters.c:44,338-356).
// The padding entries (kSlotEmpty) sort last.
__device__ __forceinline__ bool cluster_before(uint64_t a, uint64_t b, uint64_t bm) {
    if (a == kSlotEmpty || b == kSlotEmpty) return b == kSlotEmpty && a != kSlotEmpty;
    const uint64_t ba = ((a * 2654435761ull) >> 32) & bm, bb = ((b * 2654435761ull) >> 32) & bm;
    return ba < bb || (ba == bb && a < b);
}

constexpr int kRankThreads = 1024;
constexpr uint32_t kRankLds = 16384;  // clusters sorted in LDS at most (else in the frame's csort)


// One block per frame: the clusters of the table in the reference's order
// (bitonic network; keys distinct), then each cluster's rank (a binary
// search in the sorted ids) and its id at its rank; the rank bits are the
// sort key's. (LDS: kRankLds ids, dynamic.)
__global__ __launch_bounds__(kRankThreads) void k_cluster_rank(const Frame *frames) {
    extern __shared__ uint64_t lk[];
    const Frame F = frames[blockIdx.y];
    const uint32_t t = threadIdx.x;
    if (F.counters[kError] & (2u | 128u)) {
        // the call is re-run with more room: the later stages see no cluster
        if (t == 0) {
            F.counters[kNClusters] = 0;
            F.cstart[0] = 0;
        }
        return;
    }
    const uint32_t ncl = F.counters[kNClusters];
    uint32_t m = 1;
    while (m < ncl) m <<= 1;
    const uint64_t bm = F.bucket_mask;
    if (m <= kRankThreads) {
        // a cluster per thread in registers: (bucket, id, table slot),
        // bitonic, partners in the wave by lane exchanges, the others
        // through LDS; after it thread t holds the cluster of rank t
        uint32_t slot = t < ncl ? F.clist[t] : 0u;
        uint64_t id = t < ncl ? F.hkey[slot] : kSlotEmpty;
        uint32_t bk = id == kSlotEmpty ? 0xffffffffu : (uint32_t)(((id * 2654435761ull) >> 32) & bm);
        uint64_t *xid = lk;
        uint32_t *xbk = reinterpret_cast<uint32_t *>(lk + kRankThreads), *xsl = xbk + kRankThreads;
        for (uint32_t k = 2; k <= m; k <<= 1)
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                const uint32_t p = t ^ j;
                uint32_t obk, osl;
                uint64_t oid;
                if (j >= 64) {
                    xid[t] = id;
                    xbk[t] = bk;
                    xsl[t] = slot;
                    __syncthreads();
                    oid = xid[p];
                    obk = xbk[p];
                    osl = xsl[p];
                    __syncthreads();
                } else {
                    obk = (uint32_t)__shfl_xor((int)bk, (int)j, 64);
                    osl = (uint32_t)__shfl_xor((int)slot, (int)j, 64);
                    oid = ((uint64_t)(uint32_t)__shfl_xor((int)(uint32_t)(id >> 32), (int)j, 64) << 32) |
                          (uint32_t)__shfl_xor((int)(uint32_t)id, (int)j, 64);
                }
                const bool ob = obk < bk || (obk == bk && oid < id), mb = bk < obk || (bk == obk && id < oid);
                // ascending blocks keep the smaller at the lower index
                if ((t < p) == ((t & k) == 0) ? ob : mb) {
                    bk = obk;
                    id = oid;
                    slot = osl;
                }
            }
        if (t < ncl) {
            F.hrank[slot] = t;
            F.cid[t] = id;
        }
        if (t == 0) F.counters[kSortBits] = ncl <= 1 ? 1u : 32u - (uint32_t)__clz(ncl - 1);
        return;
    }
    uint64_t *keys = m <= kRankLds ? lk : F.csort;
    for (uint32_t i = t; i < m; i += kRankThreads) keys[i] = i < ncl ? F.hkey[F.clist[i]] : kSlotEmpty;
    __syncthreads();
    for (uint32_t k = 2; k <= m; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = t; i < m; i += kRankThreads) {
                const uint32_t p = i ^ j;
                if (p <= i) continue;
                const uint64_t a = keys[i], b = keys[p];
                const bool up = (i & k) == 0;
                if (up ? cluster_before(b, a, bm) : cluster_before(a, b, bm)) {
                    keys[i] = b;
                    keys[p] = a;
                }
            }
            __syncthreads();
        }
    for (uint32_t i = t; i < ncl; i += kRankThreads) {
        const uint32_t slot = F.clist[i];
        const uint64_t id = F.hkey[slot];
        uint32_t lo = 0, hi = ncl;  // the first sorted key not before id
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (cluster_before(keys[mid], id, bm)) lo = mid + 1; else hi = mid;
        }
        F.hrank[slot] = lo;
        F.cid[lo] = id;
    }
    if (t == 0) F.counters[kSortBits] = ncl <= 1 ? 1u : 32u - (uint32_t)__clz(ncl - 1);
}

// Frames of kOnePass clusters or more: the counts (k_rank_words) scanned
// into the clusters' first points, in rank order (one block a frame).
__global__ __launch_bounds__(kRankThreads) void k_cluster_starts(const Frame *frames) {
    __shared__ uint32_t wsum[kRankThreads / 64];
    const Frame F = frames[blockIdx.y];
    const uint32_t ncl = F.counters[kNClusters];
    if (ncl < kOnePass || (F.counters[kError] & (2u | 128u))) return;
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < ncl; c0 += kRankThreads) {
        const uint32_t i = c0 + t;
        const uint32_t v = i < ncl ? F.hcnt[i] : 0u;  // counts by rank
        const uint32_t inc = wave_incl_scan_u(v);
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        for (uint32_t q = 0; q < kRankThreads / 64; q++) {
            pre += q < w ? wsum[q] : 0u;
            tot += wsum[q];
        }
        if (i < ncl) F.cstart[i] = carry + pre + inc - v;
        carry += tot;
        __syncthreads();
    }
    if (t == 0) F.cstart[ncl] = carry;
}

// Frames of kOnePass clusters or more: each cluster's points counted (by
// rank, in hcnt, free since the ranks were read: cleared by k_rank_clear),
// each run of one rank on adjacent lanes adding its length. (Smaller frames
// take their cluster starts from the sort's one pass.)
__global__ __launch_bounds__(256) void k_rank_counts(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    if (F.counters[kError] & (2u | 128u)) return;
    const uint32_t n = F.counters[kNPts], ncl = F.counters[kNClusters];
    if (ncl < kOnePass) return;
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t s0, s1;
    seg_of(n, blockIdx.x, &s0, &s1);
    for (uint32_t c0 = s0; c0 < s1; c0 += 256) {
        const uint32_t i = c0 + threadIdx.x;
        const bool valid = i < s1;
        const uint32_t r = valid ? F.hrank[F.sval[i]] : 0xffffffffu;
        const uint32_t rl = wave_from_left(r, 0xfffffffeu);
        const bool head = valid && r != rl;
        const uint64_t hm = __ballot(head) | ~__ballot(valid);
        if (head) {
            const uint64_t after = hm & ~(lt | (1ull << lane));
            const uint32_t end = after ? (uint32_t)(__ffsll((unsigned long long)after) - 1) : 64u;
            atomicAdd(F.hcnt + r, end - (uint32_t)lane);
        }
    }
}

// Frames of kOnePass clusters or more: the counts by rank cleared before
// k_rank_counts adds to them (hcnt holds the slots' nothing by then).
__global__ __launch_bounds__(256) void k_rank_clear(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    const uint32_t ncl = F.counters[kNClusters];
    if (ncl < kOnePass) return;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < ncl; i += gridDim.x * 256) F.hcnt[i] = 0u;
}

// ---------------------------------------------------------------- launchers

// pk: the cluster sort (packed frames move one word); else (key, value)
// pairs throughout.
// Radix-sort blocks a frame: all kSortBlocks for calls of a few frames
// (each frame's sort spread over more of the device), half for batches.
int sort_blocks(int nframes) { return nframes <= 16 ? kSortBlocks : kSortBlocks / 2; }

void launch_radix_sort(hipStream_t st, const Frame *dframes, int nframes, int which, int nbits, int *parity,
                       int dyn, int iota, int pk, int gather, int ranks) {
    const int nb = sort_blocks(nframes);
    for (int shift = 0; shift < nbits; shift += kRadixBits) {
        const int rk = ranks && shift == 0 ? 1 : 0;
        hipLaunchKernelGGL(k_radix_hist, dim3(nb, nframes), dim3(256), 0, st, dframes, which, shift, *parity, dyn,
                           pk, rk);
        hipLaunchKernelGGL(k_radix_offsets, dim3(1, nframes), dim3(kOffThreads), 0, st, dframes, shift, dyn,
                           gather && shift == 0 ? 1 : 0, nb);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb, nframes), dim3(256), 0, st, dframes, which, shift, *parity,
                           dyn, shift == 0 ? iota : 0, pk, gather, rk);
        *parity ^= 1;
    }
}

// Host set-up of the stage's kernels (before any stream capture): the rank
// kernel's dynamic LDS.
void prepare_points_stage() {
    static bool done = false;
    if (done) return;
    done = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_cluster_rank),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kRankLds * 8)) == hipSuccess;
    (void)hipGetLastError();
}

void launch_points_stage(hipStream_t st, const Frame *dframes, int nframes, int max_qh, int key_bits, const Tune &tu,
                         Marks *mk) {
    const int rw = tu.row_waves < 1 ? 1 : tu.row_waves > 4 ? 4 : tu.row_waves;  // launch bounds: 256 threads
    dim3 g((unsigned)((max_qh + rw - 1) / rw), nframes), b(64 * rw);
    hipLaunchKernelGGL(k_point_count, g, b, 0, st, dframes);
    hipLaunchKernelGGL(k_point_scan, dim3(1, nframes), dim3(256), 0, st, dframes);
    mk->mark(st, "points.count");
    hipLaunchKernelGGL(k_point_rows, g, b, 0, st, dframes);
    mk->mark(st, "points.write");
    hipLaunchKernelGGL(k_cluster_rank, dim3(1, nframes), dim3(kRankThreads), kRankLds * 8, st, dframes);
    mk->mark(st, "clusters.rank");
    // the point counts of frames of kOnePass clusters or more (none where
    // every frame's table holds fewer: one pass gives the starts)
    if (key_bits > kRadixBits) {
        hipLaunchKernelGGL(k_rank_clear, dim3(8, nframes), dim3(256), 0, st, dframes);
        hipLaunchKernelGGL(k_rank_counts, dim3(sort_blocks(nframes), nframes), dim3(256), 0, st, dframes);
        hipLaunchKernelGGL(k_cluster_starts, dim3(1, nframes), dim3(kRankThreads), 0, st, dframes);
    }
    // key_bits: the most rank bits the call's frames can have; the passes
    // past a frame's own return at once, the last one it runs writes the
    // points in place
    int parity = 0;
    launch_radix_sort(st, dframes, nframes, kNPts, key_bits, &parity, 1, 0, 1, 1, 1);
    mk->mark(st, "cluster_sort");
}

}  // namespace atg
