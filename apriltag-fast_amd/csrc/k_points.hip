// Stage (c): boundary points, their clusters and the cluster order.
//
//   k_root_*       the roots of clusters that can hold points (the ok
//                  roots) numbered densely in node order (launched before
//                  k_labels, which writes each pixel's number to dlab).
//   k_point_rows   one wave per row pair (y, y+1), y in 1..qh-2: the 4-bit
//                  neighbour mask of each pixel (fast/clusters.c:150-174);
//                  pass 0 counts points per row, pass 1 writes each point
//                  at its emission index (row offset + wave prefix) with
//                  its cluster sort key.
//   k_point_scan   the row offsets of the points.
//   radix sort     stable LSD sort of (key, index), kRadixBits a pass: the
//                  keys order the clusters by (hash bucket, id) as the
//                  reference's cluster table does (fast/clusters.c:338-356);
//                  stability keeps each cluster's points in emission order.
//                  Where the key and the point index fit one word (the
//                  index above the key's bits: a packed frame) the sort
//                  moves that word alone.
//   k_cluster_*    the points regrouped and the cluster starts compacted.
//
// Cluster sort key of a point: bucket << 2D | dense(max_root) << D |
// dense(min_root), D the bits of the ok-root count. The dense numbers keep
// the roots' order, so comparing keys equals comparing (bucket, id) for
// id = max_root << 32 | min_root; the passes beyond the key's 2D + bucket
// bits (known on the device only) are skipped.
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

// Whether the frame's cluster sort is packed: its key bits (counters
// [kSortBits], known once the ok roots are numbered) and the bits of a
// point index (below pcap) fit one word.
__device__ __forceinline__ bool sort_packed(const Frame &F) {
    return F.counters[kSortBits] + (64u - (uint32_t)__clzll((unsigned long long)F.pcap)) <= 64u;
}
__device__ __forceinline__ uint64_t low_bits(uint32_t b) { return b >= 64 ? ~0ull : (1ull << b) - 1ull; }

// One pixel column of the row pair: t of rows a and b, their plab (root
// | ok << 31) and, for the write pass, their dlab (dense root numbers).
struct PixCol {
    uint32_t ta, tb, la, lb, da, db;
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
            wave_from_left(c.lb, fill.lb), wave_from_left(c.da, fill.da), wave_from_left(c.db, fill.db)};
}
__device__ __forceinline__ PixCol col_from_right(const PixCol &c, const PixCol &fill) {
    return {wave_from_right(c.ta, fill.ta), wave_from_right(c.tb, fill.tb), wave_from_right(c.la, fill.la),
            wave_from_right(c.lb, fill.lb), wave_from_right(c.da, fill.da), wave_from_right(c.db, fill.db)};
}
__device__ __forceinline__ PixCol col_of_lane63(const PixCol &c) {
    return {(uint32_t)__builtin_amdgcn_readlane((int)c.ta, 63), (uint32_t)__builtin_amdgcn_readlane((int)c.tb, 63),
            (uint32_t)__builtin_amdgcn_readlane((int)c.la, 63), (uint32_t)__builtin_amdgcn_readlane((int)c.lb, 63),
            (uint32_t)__builtin_amdgcn_readlane((int)c.da, 63), (uint32_t)__builtin_amdgcn_readlane((int)c.db, 63)};
}

// A wave per row pair (y, y + 1), 64 pixels a step: each lane's column of
// the step, the column right of the step (for lane 63) and both of the
// next step loaded during this one; neighbours by shuffles, the left one
// of lane 0 kept from the step before. Pass 0 counts the points of each
// row; pass 1 stages each step's points (at most 4 a pixel) in LDS and
// writes them out contiguously with their cluster sort keys (bucket << 2D
// | dense(max root) << D | dense(min root), and the point's index above
// the key's bits when the frame's sort is packed).
__global__ __launch_bounds__(256) void k_point_rows(const Frame *frames, int pass) {
    __shared__ uint64_t st_pt[4][256], st_key[4][256];
    const Frame F = frames[blockIdx.y];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int y = 1 + blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (y > F.qh - 2) return;
    const int QW = F.qw;
    const uint8_t *a = F.t + (int64_t)y * QW, *b = a + QW;
    const uint32_t *la = F.plab + (int64_t)y * QW, *lb = la + QW;
    const uint32_t *da = F.dlab + (int64_t)y * QW, *db = da + QW;
    auto load = [&](int x) -> PixCol {
        PixCol c{0, 0, 0, 0, 0, 0};
        if (x < QW) {
            c.ta = a[x];
            c.tb = b[x];
            c.la = la[x];
            c.lb = lb[x];
            if (pass) {
                c.da = da[x];
                c.db = db[x];
            }
        }
        return c;
    };
    uint32_t count = 0;
    uint32_t base = pass ? F.poff[y] : 0u;
    const uint32_t D = pass ? F.counters[kRootBits] : 0u;
    const uint32_t sbits = pass ? F.counters[kSortBits] : 0u;
    const bool packed = pass && sort_packed(F);
    const uint64_t mask = F.bucket_mask;
    PixCol cur = load(lane), cr = load(64), left{0, 0, 0, 0, 0, 0};
    for (int x0 = 0; x0 < QW - 1; x0 += 64) {
        const PixCol nxt = load(x0 + 64 + lane), nr = load(x0 + 128);
        const int x = x0 + lane;
        const PixCol r = col_from_right(cur, cr), l = col_from_left(cur, left);
        const uint32_t m = x < QW - 1 ? point_mask(cur, l, r, x) : 0u;
        const uint32_t c = __popc(m);
        if (pass) {
            const uint32_t inc = wave_incl_scan_u(c), tot = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
            uint32_t k = inc - c;  // this lane's first point in the step
            if (m) {
                const int vd = 255 - 2 * (int)cur.ta;
                const uint32_t r0 = cur.la & 0x7fffffffu, d0 = cur.da;
                for (int j = 0; j < 4; j++) {
                    if (!((m >> j) & 1u)) continue;
                    // partner (dx, dy): (1, 0), (0, 1), (-1, 1), (1, 1)
                    const int dx = j == 0 ? 1 : j == 1 ? 0 : j == 2 ? -1 : 1;
                    const int dy = j == 0 ? 0 : 1;
                    const uint32_t r1 = (j == 0 ? r.la : j == 1 ? cur.lb : j == 2 ? l.lb : r.lb) & 0x7fffffffu;
                    const uint32_t d1 = j == 0 ? r.da : j == 1 ? cur.db : j == 2 ? l.db : r.db;
                    // dense numbers keep the roots' order
                    const uint32_t hi = r0 > r1 ? r0 : r1, lo = r0 > r1 ? r1 : r0;
                    const uint32_t dh = r0 > r1 ? d0 : d1, dl = r0 > r1 ? d1 : d0;
                    const uint64_t id = ((uint64_t)hi << 32) + lo;
                    const uint64_t bucket = ((id * 2654435761ull) >> 32) & mask;
                    const uint32_t px = (uint32_t)(2 * x + dx), py = (uint32_t)(2 * y + dy);
                    const uint32_t gx = (uint32_t)(dx * vd) & 0xffffu, gy = (uint32_t)(dy * vd) & 0xffffu;
                    st_pt[wv][k] = (uint64_t)px | ((uint64_t)py << 16) | ((uint64_t)gx << 32) | ((uint64_t)gy << 48);
                    st_key[wv][k] = (bucket << (2 * D)) | ((uint64_t)dh << D) | dl;
                    k++;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
            const uint32_t at0 = base + count;
            for (uint32_t q = (uint32_t)lane; q < tot; q += 64) {
                const int64_t at = (int64_t)at0 + q;
                if (at >= F.pcap) break;  // flagged by k_point_scan
                F.pts[at] = st_pt[wv][q];
                F.skey[at] = st_key[wv][q] | (packed ? (uint64_t)at << sbits : 0ull);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
            count += tot;
        } else {
            count += (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan_u(c), 63);
        }
        left = col_of_lane63(cur);
        cur = nxt;
        cr = nr;
    }
    if (!pass && lane == 0) F.pcnt[y] = count;
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

__device__ __forceinline__ void seg_of(uint32_t n, int b, uint32_t *s0, uint32_t *s1) {
    uint32_t seg = ((n + kSortBlocks - 1) / kSortBlocks + 255u) & ~255u;
    *s0 = min(n, (uint32_t)b * seg);
    *s1 = min(n, (uint32_t)(b + 1) * seg);
}

// count = counters[which]
// The ok roots (nlab = root | ok << 31 names itself) of each segment of
// nodes: counts, then dense numbers in node order (rdense) and back
// (rinv); D and the sort key's bits into the counters.
__device__ __forceinline__ bool ok_root(const Frame &F, uint32_t i) {
    const uint32_t v = F.nlab[i];
    return (v >> 31) && (v & 0x7fffffffu) == i;
}

__global__ __launch_bounds__(256) void k_root_count(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    const uint32_t nn = F.counters[kNRuns] + (uint32_t)F.qh;
    uint32_t s0, s1;
    seg_of(nn, blockIdx.x, &s0, &s1);
    uint32_t c = 0;
    for (uint32_t i = s0 + threadIdx.x; i < s1; i += 256) c += ok_root(F, i) ? 1u : 0u;
    uint32_t tot;
    block_excl_scan(c, &tot);
    if (threadIdx.x == 0) F.rhist[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_root_dense(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    const uint32_t nn = F.counters[kNRuns] + (uint32_t)F.qh;
    uint32_t base = 0, all = 0;
    for (int b = 0; b < kSortBlocks; b++) {
        const uint32_t v = F.rhist[b];
        if (b < (int)blockIdx.x) base += v;
        all += v;
    }
    uint32_t s0, s1;
    seg_of(nn, blockIdx.x, &s0, &s1);
    for (uint32_t c0 = s0; c0 < s1; c0 += 256) {
        const uint32_t i = c0 + threadIdx.x;
        const uint32_t f = (i < s1 && ok_root(F, i)) ? 1u : 0u;
        uint32_t tot;
        const uint32_t pre = block_excl_scan(f, &tot);
        if (f) {
            F.rdense[i] = base + pre;
            F.rinv[base + pre] = i;
        }
        base += tot;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t d = all <= 1 ? 1u : 32u - (uint32_t)__clz(all - 1);
        F.counters[kRootBits] = d;
        F.counters[kSortBits] = (uint32_t)__popcll(F.bucket_mask) + 2 * d;
    }
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

// dyn: the pass is skipped where the frame's key (counters[kSortBits]) has
// no bits at shift; the skipped passes are the last ones. pk: the frame's
// words are packed where sort_packed says so (the key is their low
// kSortBits bits).
// Digit counts of each block's segment, rhist[block][digit]; each run of
// equal digits on adjacent lanes adds its length once (a digit that comes
// back later in the wave adds again: counts only, no order).
__global__ __launch_bounds__(256) void k_radix_hist(const Frame *frames, int which, int shift, int parity, int dyn,
                                                    int pk) {
    __shared__ uint32_t h[kRadixBins];
    const Frame F = frames[blockIdx.y];
    const uint32_t sb = F.counters[kSortBits];
    if (dyn && shift >= (int)sb) return;
    const uint64_t km = pk && sort_packed(F) ? low_bits(sb) : ~0ull;
    const uint32_t n = F.counters[which];
    const SortBufs B = sort_bufs(F, parity);
    const int t = threadIdx.x, lane = t & 63;
    for (int i = t; i < kRadixBins; i += 256) h[i] = 0;
    __syncthreads();
    uint32_t s0, s1;
    seg_of(n, blockIdx.x, &s0, &s1);
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t c0 = s0; c0 < s1; c0 += 256) {
        const uint32_t i = c0 + t;
        const bool valid = i < s1;
        const uint32_t d = valid ? (uint32_t)((B.kin[i] & km) >> shift) & (kRadixBins - 1) : 0xffffffffu;
        const uint32_t dl = wave_from_left(d, 0xfffffffeu);  // lane 0: a head
        const uint64_t heads = __ballot(valid && d != dl) | ~__ballot(valid);  // invalid lanes end runs
        if (valid && d != dl) {
            const uint64_t after = heads & ~(lt | (1ull << lane));
            const uint32_t end = after ? (uint32_t)(__ffsll((unsigned long long)after) - 1) : 64u;
            atomicAdd(&h[d], end - (uint32_t)lane);
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
__global__ __launch_bounds__(kOffThreads) void k_radix_offsets(const Frame *frames, int shift, int dyn) {
    constexpr int kPer = kRadixBins / kOffThreads;
    static_assert(kPer * kOffThreads == kRadixBins, "whole digits a thread");
    const Frame F = frames[blockIdx.y];
    if (dyn && shift >= (int)F.counters[kSortBits]) return;
    const int t = threadIdx.x;
    uint32_t carry = 0;
    for (int q = 0; q < kPer; q++) {
        const int d = q * kOffThreads + t;
        uint32_t v[kSortBlocks], tot = 0;
#pragma unroll
        for (int b = 0; b < kSortBlocks; b++) {
            v[b] = F.rhist[b * kRadixBins + d];
            tot += v[b];
        }
        uint32_t all;
        uint32_t run = carry + block_excl_scan_t<kOffThreads>(tot, &all);
        carry += all;
#pragma unroll
        for (int b = 0; b < kSortBlocks; b++) {
            F.rhist[b * kRadixBins + d] = run;
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
                                                       int dyn, int iota, int pk) {
    __shared__ uint32_t off[kRadixBins];
    __shared__ uint8_t wcnt[4][kRadixBins];
    const Frame F = frames[blockIdx.y];
    const uint32_t sb = F.counters[kSortBits];
    if (dyn && shift >= (int)sb) return;
    const bool packed = pk && sort_packed(F);
    const uint64_t km = packed ? low_bits(sb) : ~0ull;
    const uint32_t n = F.counters[which];
    const SortBufs B = sort_bufs(F, parity);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int i = t; i < kRadixBins; i += 256) {
        off[i] = F.rhist[blockIdx.x * kRadixBins + i];
        for (int k = 0; k < 4; k++) wcnt[k][i] = 0;
    }
    __syncthreads();
    uint32_t s0, s1;
    seg_of(n, blockIdx.x, &s0, &s1);
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t c0 = s0; c0 < s1; c0 += 256) {
        const uint32_t i = c0 + t;
        const bool valid = i < s1;
        const uint64_t key = valid ? B.kin[i] : 0ull;
        const uint32_t val = (valid && !packed) ? (iota ? i : B.vin[i]) : 0u;
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
            B.kout[pos] = key;
            if (!packed) B.vout[pos] = val;
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

// The buffer of the cluster sort's result: one swap per executed pass.
__device__ __forceinline__ int sorted_parity(const Frame &F) {
    return (int)(((F.counters[kSortBits] + kRadixBits - 1) / kRadixBits) & 1);
}

// After the sort: the points in cluster order, and the number of cluster
// starts in each block's segment.
__global__ __launch_bounds__(256) void k_cluster_count(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    const int parity = sorted_parity(F);
    uint32_t n = F.counters[kNPts];
    const uint64_t *key = parity ? F.skey2 : F.skey;
    const uint32_t *val = parity ? F.sval2 : F.sval;
    const uint32_t sb = F.counters[kSortBits];
    const bool packed = sort_packed(F);
    const uint64_t km = packed ? low_bits(sb) : ~0ull;
    uint32_t s0, s1;
    seg_of(n, blockIdx.x, &s0, &s1);
    uint32_t c = 0;
    for (uint32_t i = s0 + threadIdx.x; i < s1; i += 256) {
        const uint64_t k = key[i];
        F.cpts[i] = F.pts[packed ? (uint32_t)(k >> sb) : val[i]];
        c += (i == 0 || (k & km) != (key[i - 1] & km)) ? 1u : 0u;
    }
    uint32_t tot;
    block_excl_scan(c, &tot);
    if (threadIdx.x == 0) F.rhist[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_cluster_write(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    const int parity = sorted_parity(F);
    uint32_t n = F.counters[kNPts];
    const uint64_t *key = parity ? F.skey2 : F.skey;
    const uint64_t km = sort_packed(F) ? low_bits(F.counters[kSortBits]) : ~0ull;
    uint32_t base = 0, all = 0;
    for (int b = 0; b < kSortBlocks; b++) {
        uint32_t v = F.rhist[b];
        if (b < (int)blockIdx.x) base += v;
        all += v;
    }
    uint32_t s0, s1;
    seg_of(n, blockIdx.x, &s0, &s1);
    for (uint32_t c0 = s0; c0 < s1; c0 += 256) {
        uint32_t i = c0 + threadIdx.x;
        const uint64_t k = i < s1 ? key[i] & km : 0ull;
        uint32_t f = (i < s1 && (i == 0 || k != (key[i - 1] & km))) ? 1u : 0u;
        uint32_t tot;
        uint32_t pre = block_excl_scan(f, &tot);
        if (f) {
            F.cstart[base + pre] = i;
            // the cluster id back from the sort key's dense roots
            const uint32_t D = F.counters[kRootBits];
            const uint64_t dm = ((uint64_t)1 << D) - 1;
            F.cid[base + pre] = ((uint64_t)F.rinv[(k >> D) & dm] << 32) | F.rinv[k & dm];
        }
        base += tot;
    }
    if (blockIdx.x == kSortBlocks - 1 && threadIdx.x == 0) {
        F.counters[kNClusters] = all;
        F.cstart[all] = n;
    }
}

// ---------------------------------------------------------------- launchers

// The ok roots numbered densely in node order (before k_labels, which
// writes each pixel's number).
void launch_root_numbers(hipStream_t st, const Frame *dframes, int nframes) {
    hipLaunchKernelGGL(k_root_count, dim3(kSortBlocks, nframes), dim3(256), 0, st, dframes);
    hipLaunchKernelGGL(k_root_dense, dim3(kSortBlocks, nframes), dim3(256), 0, st, dframes);
}

// pk: the cluster sort (packed frames move one word); else (key, value)
// pairs throughout.
void launch_radix_sort(hipStream_t st, const Frame *dframes, int nframes, int which, int nbits, int *parity,
                       int dyn, int iota, int pk) {
    for (int shift = 0; shift < nbits; shift += kRadixBits) {
        hipLaunchKernelGGL(k_radix_hist, dim3(kSortBlocks, nframes), dim3(256), 0, st, dframes, which, shift,
                           *parity, dyn, pk);
        hipLaunchKernelGGL(k_radix_offsets, dim3(1, nframes), dim3(kOffThreads), 0, st, dframes, shift, dyn);
        hipLaunchKernelGGL(k_radix_scatter, dim3(kSortBlocks, nframes), dim3(256), 0, st, dframes, which, shift,
                           *parity, dyn, shift == 0 ? iota : 0, pk);
        *parity ^= 1;
    }
}

void launch_points_stage(hipStream_t st, const Frame *dframes, int nframes, int max_qh, int key_bits, const Tune &tu,
                         Marks *mk) {
    const int rw = tu.row_waves < 1 ? 1 : tu.row_waves > 4 ? 4 : tu.row_waves;  // launch bounds: 256 threads
    dim3 g((unsigned)((max_qh + rw - 1) / rw), nframes), b(64 * rw);
    hipLaunchKernelGGL(k_point_rows, g, b, 0, st, dframes, 0);
    hipLaunchKernelGGL(k_point_scan, dim3(1, nframes), dim3(256), 0, st, dframes);
    mk->mark(st, "points.count");
    hipLaunchKernelGGL(k_point_rows, g, b, 0, st, dframes, 1);
    mk->mark(st, "points.write");
    // key_bits: the widest key the call's frames can have
    int parity = 0;
    launch_radix_sort(st, dframes, nframes, kNPts, key_bits, &parity, 1, 1, 1);
    mk->mark(st, "cluster_sort");
    hipLaunchKernelGGL(k_cluster_count, dim3(kSortBlocks, nframes), dim3(256), 0, st, dframes);
    mk->mark(st, "clusters.gather");
    hipLaunchKernelGGL(k_cluster_write, dim3(kSortBlocks, nframes), dim3(256), 0, st, dframes);
    mk->mark(st, "clusters.write");
}

}  // namespace atg
