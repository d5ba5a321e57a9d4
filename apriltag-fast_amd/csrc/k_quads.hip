// Stage (c, sort keys) and (d): fit a quad to each cluster.
//
//   k_fit_prep     one block per cluster: size and bounding-box rejects,
//                  the angle key of every point (fast/quads.c:438-453), the
//                  gradient dot product in the reference's 8-lane order
//                  (:467-512), the border polarity; writes (cluster, slope)
//                  sort items in reverse point order.
//   radix sort     all clusters' items at once: stable LSD by (cluster,
//                  slope), so equal slopes come out in descending point
//                  index = ascending complemented index: the full-key order.
//   k_fit_moments  the full keys; the leaf-block fix of runs of equal slope
//                  (the order of the reference's merge sort, :42-46,
//                  :241-327), the point weights and the six cumulative
//                  moment sums, sequential in angle order as the reference
//                  adds them (:1006-1035), staged through LDS.
//   k_fit_corners  window errors, the 7-tap filter, maxima, the maxima
//                  threshold, pair line fits and the best 4 corners, the
//                  four lines, corners, area and convexity (:539-959,
//                  :1044-1126).
//   k_quad_order   the fitted quads back into cluster order.
//
// Floating point follows the reference's operation order exactly; the
// file is compiled without FP contraction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "atg_internal.h"

#pragma clang fp contract(off)

namespace atg {

// Correctly rounded float sqrt and division. The device's single-precision
// sqrt is not correctly rounded; double has more than 2*24+2 bits, so one
// double operation rounded once to float is the correctly rounded result.
#ifndef ATG_SCAN_RING
#define ATG_SCAN_RING 16
#endif
__device__ __forceinline__ float fsqrt_cr(float x) { return (float)__dsqrt_rn((double)x); }
__device__ __forceinline__ float fdiv_cr(float a, float b) { return (float)__ddiv_rn((double)a, (double)b); }

__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }

__device__ __forceinline__ uint32_t wscan(uint32_t v) { return wave_scan_incl(v); }

__device__ __forceinline__ uint32_t bscan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t ws[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wscan(v);
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int i = 0; i < 4; i++) {
        if (i < w) pre += ws[i];
        tot += ws[i];
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

// slope_sort_key (fast/quads.c:429-433)
__device__ __forceinline__ uint32_t slope_key(float s) {
    uint32_t b = f2u(s), sg = (uint32_t)((int32_t)b >> 31);
    return b ^ (sg | 0x80000000u);
}

// The high word of point_key (fast/quads.c:438-453).
__device__ __forceinline__ uint32_t point_slope(float dx, float dy) {
    float quadrant = dy > 0.0f ? (dx > 0.0f ? 65536.0f : 131072.0f) : (dx > 0.0f ? 0.0f : -65536.0f);
    uint32_t neg = dy < 0.0f ? 0x80000000u : 0u;
    float fx = u2f(f2u(dx) ^ neg), fy = u2f(f2u(dy) ^ neg);
    float ndx = fx < 0.0f ? fy : fx;
    float ndy = fx < 0.0f ? -fx : fy;
    return slope_key(__fadd_rn(quadrant, fdiv_cr(ndy, ndx)));
}

__device__ __forceinline__ void pt_unpack(uint64_t p, int *x, int *y, int *gx, int *gy) {
    *x = (int)(p & 0xffffu);
    *y = (int)((p >> 16) & 0xffffu);
    *gx = (int)(int16_t)(uint16_t)((p >> 32) & 0xffffu);
    *gy = (int)(int16_t)(uint16_t)(p >> 48);
}

// ------------------------------------------------------------ prep

// Min and max over a wave by lane exchanges (xor 32 .. 1).
template <int J = 32>
__device__ __forceinline__ void wave_minmax(uint32_t &mn, uint32_t &mx) {
    mn = min(mn, lane_xor32<J>(mn));
    mx = max(mx, lane_xor32<J>(mx));
    if constexpr (J > 1) wave_minmax<J / 2>(mn, mx);
}

// Every cluster of every frame, NT threads a cluster (a wave: most clusters
// are a few hundred points), blocks taking clusters in strides. The first
// kR NT points stay in registers through the passes (bounding box, keys,
// dot product, weights); the others are read again at each.
template <int NT>
__global__ __launch_bounds__(NT) void k_fit_prep(const Frame *frames, const DevParams *P, int items) {
    constexpr int kDotChunk = NT * 16;  // dot-product terms staged at a time
    constexpr int kW = NT / 64;
    constexpr int kR = 4;
    __shared__ float terms[kDotChunk];
    __shared__ float tail[8];
    __shared__ uint32_t red[kW][4];
    __shared__ float lanes8[8];
    __shared__ int s_keep;
    const Frame F = frames[blockIdx.y];
    const uint32_t ncl = F.counters[kNClusters];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (uint32_t c = blockIdx.x; c < ncl; c += gridDim.x) {
        const uint32_t s0 = F.cstart[c], n = F.cstart[c + 1] - s0;
        const auto pts = gmem(F.cpts) + s0;
        bool in_range = (int64_t)n >= P->mcp && (int64_t)n <= F.max_px && n >= kMinFitPoints;
        uint64_t pv[kR];
#pragma unroll
        for (int r = 0; r < kR; r++) {
            const uint32_t i = t + r * NT;
            pv[r] = i < n ? pts[i] : 0ull;
        }
        // point i, for this thread's indices only (i = t + r NT)
        auto pt_at = [&](uint32_t i) -> uint64_t {
            const uint32_t r = (i - t) / NT;
            return r == 0 ? pv[0] : r == 1 ? pv[1] : r == 2 ? pv[2] : r == 3 ? pv[3] : pts[i];
        };
        // bounding box (fast/quads.c:515-528)
        uint32_t xmin = 0xffff, xmax = 0, ymin = 0xffff, ymax = 0;
        if (in_range)
            for (uint32_t i = t; i < n; i += NT) {
                int x, y, gx, gy;
                pt_unpack(pt_at(i), &x, &y, &gx, &gy);
                xmin = min(xmin, (uint32_t)x); xmax = max(xmax, (uint32_t)x);
                ymin = min(ymin, (uint32_t)y); ymax = max(ymax, (uint32_t)y);
            }
        wave_minmax(xmin, xmax);
        wave_minmax(ymin, ymax);
        if constexpr (kW > 1) {
            if (lane == 0) { red[w][0] = xmin; red[w][1] = xmax; red[w][2] = ymin; red[w][3] = ymax; }
            __syncthreads();
            for (int k2 = 0; k2 < kW; k2++) {
                xmin = min(xmin, red[k2][0]); xmax = max(xmax, red[k2][1]);
                ymin = min(ymin, red[k2][2]); ymax = max(ymax, red[k2][3]);
            }
            __syncthreads();
        }
        bool ok = in_range && ((int64_t)xmax - (int64_t)xmin) * ((int64_t)ymax - (int64_t)ymin) >= P->tag_width;
        float cx = (float)((double)(xmin + xmax) * 0.5 + 0.05118);
        float cy = (float)((double)(ymin + ymax) * 0.5 + -0.028581);
        // sort items (reverse point order) and the full keys of the points
        // (a rejected cluster's keys are never read: a third of the points
        // lie in the few clusters over max_px)
        for (uint32_t i = t; i < n && (ok || items); i += NT) {
            uint32_t slope = 0;
            if (ok) {
                int x, y, gx, gy;
                pt_unpack(pt_at(i), &x, &y, &gx, &gy);
                slope = point_slope(__fsub_rn((float)x, cx), __fsub_rn((float)y, cy));
                gmem(F.okey)[s0 + i] = ((uint64_t)slope << 32) | (uint32_t)~i;
            }
            if (items) {  // for the sort of all points at once
                uint32_t pos = s0 + (n - 1 - i);
                gmem(F.skey)[pos] = ((uint64_t)c << 32) | slope;
                gmem(F.sval)[pos] = i;
            }
        }
        if (!ok) {
            if (t == 0) F.cflag[c] = 0;
            continue;
        }
        // dot product: lane l of 8 sums the terms of points l, l+8, ...
        // below n - n % 8 in order, then the last n % 8 terms in order
        // (fast/quads.c:490-504)
        auto term = [&](uint64_t p) {
            int x, y, gx, gy;
            pt_unpack(p, &x, &y, &gx, &gy);
            float dx = __fsub_rn((float)x, cx), dy = __fsub_rn((float)y, cy);
            return __fadd_rn(__fmul_rn(dx, (float)gx), __fmul_rn(dy, (float)gy));
        };
        const uint32_t nv = n - n % 8;
        float acc = 0.0f;
        for (uint32_t c0 = 0; c0 < nv; c0 += kDotChunk) {
            uint32_t len = min((uint32_t)kDotChunk, nv - c0);
            for (uint32_t j = t; j < len; j += NT) terms[j] = term(pt_at(c0 + j));
            __syncthreads();
            if (t < 8) {
                // in order, 8 terms a step read at once (one LDS wait a step)
                uint32_t j = t;
                for (; j + 56 < len; j += 64) {
                    float v[8];
#pragma unroll
                    for (int q = 0; q < 8; q++) v[q] = terms[j + 8 * q];
#pragma unroll
                    for (int q = 0; q < 8; q++) acc = __fadd_rn(acc, v[q]);
                }
                for (; j < len; j += 8) acc = __fadd_rn(acc, terms[j]);
            }
            __syncthreads();
        }
        if (t < 8) lanes8[t] = acc;
        if (t < (int)(n - nv)) tail[t] = term(pts[nv + t]);
        __syncthreads();
        if (t == 0) {
            float dot = __fadd_rn(__fadd_rn(__fadd_rn(lanes8[0], lanes8[4]), __fadd_rn(lanes8[2], lanes8[6])),
                                  __fadd_rn(__fadd_rn(lanes8[1], lanes8[5]), __fadd_rn(lanes8[3], lanes8[7])));
            for (uint32_t i = 0; i < n - nv; i++) dot = __fadd_rn(dot, tail[i]);
            bool rev = dot < 0.0f;
            bool keep = !((!P->reversed_border && rev) || (!P->normal_border && !rev));
            F.cflag[c] = keep ? (1u | (rev ? 2u : 0u)) : 0u;
            if (keep) atomicAdd(F.counters + kNKeptPts, n);
            s_keep = keep ? 1 : 0;
        }
        __syncthreads();
        if (s_keep) {
            // for the moments: x, y and the squared image gradient at the
            // point's pixel (fast/quads.c:968-985), in cluster order
            const int iw = F.qw, ih = F.qh;
            for (uint32_t i = t; i < n; i += NT) {
                int px, py, gx, gy;
                pt_unpack(pt_at(i), &px, &py, &gx, &gy);
                int ix = (px >> 1) + (px & 1), iy = (py >> 1) + (py & 1);
                uint32_t g2 = 0;
                if (ix > 0 && ix + 1 < iw && iy > 0 && iy + 1 < ih) {
                    const auto im = gmem(F.qim) + (int64_t)iy * iw + ix;
                    int ggx = (int)im[1] - (int)im[-1], ggy = (int)im[iw] - (int)im[-iw];
                    g2 = (uint32_t)(ggx * ggx + ggy * ggy);
                }
                gmem(F.pw)[s0 + i] = (uint64_t)px | ((uint64_t)py << 16) | ((uint64_t)g2 << 32);
            }
        }
    }
}

// One compare-exchange of distinct keys (or equal padding): the lane takes
// its partner's key when the partner is the smaller one and the lane keeps
// the minimum, or the larger one and it keeps the maximum.
__device__ __forceinline__ uint64_t cx_take(uint64_t mine, uint64_t other, bool keep_min) {
    return (other < mine) == keep_min ? other : mine;
}

// Step (K, J) of a wave's bitonic sort of 64 R keys in registers (key
// i = r * 64 + lane in v[r]), then the steps after it.
template <int R, int K, int J>
__device__ __forceinline__ void wave_bitonic_step(uint64_t (&v)[R]) {
    const int lane = threadIdx.x & 63;
    if constexpr (J >= 64) {
        constexpr int jr = J >> 6;
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (r & jr) continue;
            const int r2 = r | jr;
            const bool up = ((r << 6) & K) == 0;
            const uint64_t a = v[r], b = v[r2];
            const bool sw = (a > b) == up;
            v[r] = sw ? b : a;
            v[r2] = sw ? a : b;
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint64_t o = lane_xor64<J>(v[r]);
            const bool up = (((r << 6) | lane) & K) == 0;
            v[r] = cx_take(v[r], o, ((lane & J) == 0) == up);
        }
    }
    if constexpr (J > 1) wave_bitonic_step<R, K, J / 2>(v);
}
template <int R, int K>
__device__ __forceinline__ void wave_bitonic_merge(uint64_t (&v)[R]) {
    wave_bitonic_step<R, K, K / 2>(v);
    if constexpr (K < 64 * R) wave_bitonic_merge<R, 2 * K>(v);
}

// A wave's bitonic sort of 64 R keys in registers: key i = r * 64 + lane
// sits in v[r]; partners within a register row by lane exchanges, across
// rows in the lane's own registers. Ascending.
template <int R>
__device__ __forceinline__ void wave_bitonic(uint64_t (&v)[R]) {
    wave_bitonic_merge<R, 2>(v);
}

// ------------------------------------------------------------ sort fix

__device__ __forceinline__ void kswap(uint64_t *k, int a, int b) {
    if ((k[a] >> 32) > (k[b] >> 32)) {
        uint64_t x = k[a];
        k[a] = k[b];
        k[b] = x;
    }
}

// The compare-exchange networks of the merge sort's leaves, high words
// only (fast/quads.c:56-85).
__device__ __forceinline__ void leaf_network(uint64_t *k, int n) {
    if (n <= 1) return;
    if (n == 2) { kswap(k, 0, 1); return; }
    if (n == 3) { kswap(k, 0, 1); kswap(k, 1, 2); kswap(k, 0, 1); return; }
    if (n == 4) { kswap(k, 0, 1); kswap(k, 2, 3); kswap(k, 0, 2); kswap(k, 1, 3); kswap(k, 1, 2); return; }
    kswap(k, 0, 1); kswap(k, 3, 4); kswap(k, 1, 2); kswap(k, 0, 1); kswap(k, 0, 3);
    kswap(k, 2, 4); kswap(k, 1, 2); kswap(k, 2, 3); kswap(k, 1, 2);
}

// The merge sort's leaf holding position p of n (fast/quads.c:215-237):
// split in four while half the range exceeds 5, then in halves.
__device__ void leaf_of(int64_t n, int64_t p, int64_t *lo_out, int64_t *len_out) {
    int64_t lo = 0, sz = n;
    while (sz / 2 > 5) {
        int64_t asz = sz / 2, a1 = asz / 2, b1 = asz + (sz - asz) / 2, r = p - lo;
        if (r < a1) sz = a1;
        else if (r < asz) { lo += a1; sz = asz - a1; }
        else if (r < b1) { lo += asz; sz = b1 - asz; }
        else { lo += b1; sz = sz - b1; }
    }
    while (sz > 5) {
        int64_t asz = sz / 2;
        if (p - lo >= asz) { lo += asz; sz -= asz; }
        else sz = asz;
    }
    *lo_out = lo;
    *len_out = sz;
}

__device__ __forceinline__ int64_t kidx(uint64_t k) { return (int64_t)(uint32_t)~(uint32_t)k; }

// The entry of key i of a run (keys[s, s + m) of slope hi) in the merge
// sort's order: the run's keys fall into groups of one merge-sort leaf
// each (consecutive, as the indices descend along the run, at most 5 long);
// the keys of a group are emitted in its leaf network's order, each with
// the running maximum as its block head (fast/quads.c:241-302).
__device__ __forceinline__ void run_entry(const uint64_t *orig, int64_t n, const uint64_t *keys, uint32_t s,
                                          uint32_t m, uint64_t hi, uint32_t i, uint64_t key, int64_t lo, int64_t len,
                                          int g0, int g1, uint64_t *ent, uint64_t *bh) {
    *ent = key;
    *bh = key;
    if (g1 - g0 <= 1) return;
    uint64_t lf[5];
#pragma unroll
    for (int q = 0; q < 5; q++) lf[q] = q < len ? gmem(orig)[lo + q] : 0ull;
    leaf_network(lf, (int)len);
    const int j = (int)i - g0;
    uint64_t top = 0;
    int c = 0;
#pragma unroll
    for (int q = 0; q < 5; q++) {
        if (q < len && (lf[q] & 0xffffffff00000000ull) == hi) {
            top = (c == 0 || lf[q] > top) ? lf[q] : top;
            if (c == j) {
                *ent = lf[q];
                *bh = top;
            }
            c++;
        }
    }
}

// order_run for a batch of runs of at most 64 keys together, a lane per
// key: lane l holds key j of the run starting at keys[rs] with m keys,
// whose first key is on lane base (lanes past the batch: m = 0). Each run's
// entries are ranked by block head, stably, among its own lanes.
__device__ __noinline__ void order_runs_batch(const uint64_t *orig, int64_t n, uint64_t *keys, uint32_t rs,
                                                 uint32_t m, int base) {
    const int lane = threadIdx.x & 63;
    const int j = lane - base;
    const bool act = m > 0;
    const uint64_t key = act ? keys[rs + (uint32_t)j] : 0ull;
    const uint64_t hi = key & 0xffffffff00000000ull;
    int64_t lo = -1, len = 0;
    if (act) leaf_of(n, kidx(key), &lo, &len);
    const int64_t lo_prev = __shfl_up(lo, 1, 64);
    const uint64_t starts = __ballot(act && (j == 0 || lo_prev != lo));
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int g0 = 63 - __clzll((long long)(starts & upto));
    const uint64_t after = starts & ~upto;
    int g1 = after ? __ffsll((unsigned long long)after) - 1 : 64;
    g1 = min(g1, base + (int)m);
    uint64_t ent = key, bh = key;
    if (act) run_entry(orig, n, keys, rs, m, hi, (uint32_t)j, key, lo, len, g0 - base, g1 - base, &ent, &bh);
    // every run's entries in (head, position) order at once: a register
    // sort of (run's first lane, head's low word, lane) (a run's heads share
    // its slope, the high word); after it, lane p of a run holds the entry
    // of rank p - base, from its source lane
    uint64_t cv[1] = {((uint64_t)(act ? base : 127) << 38) | ((bh & 0xffffffffull) << 6) | (uint64_t)lane};
    wave_bitonic<1>(cv);
    const int src = (int)(cv[0] & 63u);
    const uint64_t e = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(ent >> 32), src, 64) << 32) |
                       (uint32_t)__shfl((int)(uint32_t)ent, src, 64);
    if (act) keys[rs + (uint32_t)j] = e;
}

// order_run for a longer run, 64 keys at a time, the entries and heads in
// HBM scratch (ent, bh at the run's positions).
__device__ __noinline__ void order_run_long(const uint64_t *orig, int64_t n, uint64_t *keys, uint32_t s, uint32_t m,
                               uint64_t *ent_s, uint64_t *bh_s) {
    const int lane = threadIdx.x & 63;
    const uint64_t hi = keys[s] & 0xffffffff00000000ull;
    auto leaf_lo = [&](uint32_t i) {
        int64_t lo, len;
        leaf_of(n, kidx(keys[s + i]), &lo, &len);
        return lo;
    };
    for (uint32_t r0 = 0; r0 < m; r0 += 64) {
        const uint32_t i = r0 + (uint32_t)lane;
        if (i < m) {
            const uint64_t key = keys[s + i];
            int64_t lo, len;
            leaf_of(n, kidx(key), &lo, &len);
            int g0 = (int)i, g1 = (int)i + 1;
            while (g0 > 0 && leaf_lo((uint32_t)g0 - 1) == lo) g0--;
            while (g1 < (int)m && leaf_lo((uint32_t)g1) == lo) g1++;
            uint64_t e, b;
            run_entry(orig, n, keys, s, m, hi, i, key, lo, len, g0, g1, &e, &b);
            gmem(ent_s)[i] = e;
            gmem(bh_s)[i] = b;
        }
    }
    __threadfence_block();
    for (uint32_t r0 = 0; r0 < m; r0 += 64) {
        const uint32_t i = r0 + (uint32_t)lane;
        const uint64_t b = i < m ? gmem(bh_s)[i] : 0ull;
        uint32_t rank = 0;
        for (uint32_t j = 0; j < m; j++) {
            const uint64_t b2 = gmem(bh_s)[j];
            rank += (b2 < b || (b2 == b && j < i)) ? 1u : 0u;
        }
        if (i < m) keys[s + rank] = gmem(ent_s)[i];
    }
}

// fix_runs (fast/quads.c:304-327) by a block of NT threads: each run of
// equal slope holding two neighbours less than 5 apart is put in the merge
// sort's order. The keys go in segments of 8 NT positions: a lane at each
// run start in the segment reads its run (it may reach past the segment)
// and lists it if it needs the fix; the waves then split the list, pack
// runs of at most 64 keys onto their lanes (order_runs_batch) and take the
// longer ones one at a time (order_run_long). Runs are disjoint and a
// run's rewrite keeps every slope, so the runs go in any order. ls/lm hold
// 4 NT + 1 entries; ends with a barrier.
template <int NT>
__device__ void fix_runs_block(uint64_t *keys, uint32_t n, const uint64_t *orig, uint64_t *ent, uint64_t *bh,
                               uint32_t *ls, uint32_t *lm, uint32_t *lcnt) {
    constexpr uint32_t kSeg = 8 * NT;
    constexpr int kW = NT / 64;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (uint32_t g0 = 0; g0 < n; g0 += kSeg) {
        const uint32_t g1 = min(n, g0 + kSeg);
        if (t == 0) *lcnt = 0;
        __syncthreads();
        // thread t lists the runs starting at positions [p0, p0 + 8) of
        // the segment, p0 = g0 + 8 t: from its keys p0 - 1 .. p0 + 15 in
        // registers, bit r of `same` says position p0 + r continues the run
        // of p0 + r - 1 and bit r of `near` that it is also less than 5 from
        // it (fast/quads.c:313-318); a run is followed through the registers
        // and, past them (runs of more than 8 keys), through `keys`
        const uint32_t p0 = g0 + 8u * (uint32_t)t;
        uint64_t kr[17];
#pragma unroll
        for (int r = 0; r < 17; r++) {
            const uint32_t k = p0 + (uint32_t)r - 1u;
            kr[r] = (r > 0 || p0 > 0) && k < n ? keys[k] : 0ull;
        }
        uint32_t same = 0, near = 0;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const uint32_t k = p0 + (uint32_t)r;
            const bool sm = k > 0 && k < n && (kr[r + 1] >> 32) == (kr[r] >> 32);
            same |= (sm ? 1u : 0u) << r;
            near |= (sm && kr[r + 1] - kr[r] < 5 ? 1u : 0u) << r;
        }
        // starts among the first 8 positions (inside the segment)
        const uint32_t lim = g1 > p0 ? min(8u, g1 - p0) : 0u;
        uint32_t starts = ~same & ((1u << lim) - 1u);
        while (starts) {
            const int r0 = __ffs((int)starts) - 1;
            starts &= starts - 1;
            // the run's positions r0 + 1 .. e - 1 continue it
            const uint32_t rest = ~(same >> (r0 + 1)) & 0xffffu;
            const int len_in = __ffs((int)rest) - 1;  // keys after r0 in the registers, at most 15 - r0
            uint32_t e = p0 + (uint32_t)r0 + 1u + (uint32_t)len_in;
            bool need = ((near >> (r0 + 1)) & ((1u << len_in) - 1u)) != 0;
            if (r0 + 1 + len_in == 16 && e < n) {
                // the run goes on past the registers
                const uint64_t hk = kr[16] >> 32;
                uint64_t prev = kr[16];
                bool more = true;
                while (more) {
                    uint64_t v[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) v[q] = e + q < n ? keys[e + q] : 0ull;
                    const uint32_t e0 = e;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        if (more && e0 + q < n && (v[q] >> 32) == hk) {
                            need = need || v[q] - prev < 5;
                            prev = v[q];
                            e++;
                        } else {
                            more = false;
                        }
                    }
                }
            }
            if (need) {
                const uint32_t slot = atomicAdd(lcnt, 1u);
                ls[slot] = p0 + (uint32_t)r0;
                lm[slot] = e - (p0 + (uint32_t)r0);
            }
        }
        __syncthreads();
        const uint32_t L = *lcnt;
        const uint32_t a = L * (uint32_t)w / kW, b = L * (uint32_t)(w + 1) / kW;
        int used = 0, my_base = 0;
        uint32_t my_rs = 0, my_m = 0;
        for (uint32_t i0 = a; i0 < b; i0 += 64) {
            const uint32_t i = i0 + (uint32_t)lane;
            const uint32_t rs_l = i < b ? ls[i] : 0u, m_l = i < b ? lm[i] : 0u;
            const uint32_t cnt = min(64u, b - i0);
            for (uint32_t q = 0; q < cnt; q++) {
                const uint32_t rs = (uint32_t)__builtin_amdgcn_readlane((int)rs_l, (int)q),
                               m = (uint32_t)__builtin_amdgcn_readlane((int)m_l, (int)q);
                if (m <= 64) {
                    if (used + (int)m > 64) {
                        order_runs_batch(orig, n, keys, my_rs, my_m, my_base);
                        used = 0;
                        my_m = 0;
                    }
                    if (lane >= used && lane < used + (int)m) {
                        my_rs = rs;
                        my_m = m;
                        my_base = used;
                    }
                    used += (int)m;
                } else {
                    if (used) order_runs_batch(orig, n, keys, my_rs, my_m, my_base);
                    used = 0;
                    my_m = 0;
                    order_run_long(orig, n, keys, rs, m, ent + rs, bh + rs);
                }
            }
        }
        if (used) order_runs_batch(orig, n, keys, my_rs, my_m, my_base);
        __syncthreads();
    }
}

// Per frame (a block each): its kept clusters in cluster order, the huge
// ones (above kHugeCluster points) first, then the other large ones (above
// kLargeCluster), and the small ones at kcap (block scans over the flags),
// and their counts.
__global__ __launch_bounds__(256) void k_keep_lists(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    const uint32_t ncl = F.counters[kNClusters];
    const int t = threadIdx.x;
    uint32_t nh = 0, nl = 0, ns = 0;
    // the huge ones, then the large ones after them
    for (int pass = 0; pass < 2; pass++) {
        for (uint32_t c0 = 0; c0 < ncl; c0 += 256) {
            const uint32_t c = c0 + t;
            const bool kept = c < ncl && (F.cflag[c] & 1u);
            const uint32_t sz = kept ? F.cstart[c + 1] - F.cstart[c] : 0u;
            const bool huge = sz > kHugeCluster, large = sz > kLargeCluster && !huge;
            const bool mine = pass == 0 ? huge : large;
            uint32_t tot;
            const uint32_t pre = bscan(mine ? 1u : 0u, &tot);
            const uint32_t at = (pass == 0 ? 0u : nh) + (pass == 0 ? nh : nl) + pre;
            if (mine && at < F.kcap) F.klist[at] = c;
            if (pass == 0) {
                nh += tot;
            } else {
                nl += tot;
                uint32_t ts;
                const uint32_t ps = bscan(kept && sz <= kLargeCluster ? 1u : 0u, &ts);
                if (kept && sz <= kLargeCluster && ns + ps < F.kcap) F.klist[F.kcap + ns + ps] = c;
                ns += ts;
            }
        }
    }
    if (t == 0) {
        if (nh + nl > F.kcap || ns > F.kcap) atomicOr(F.counters + kError, 64u);
        nh = min(nh, F.kcap);
        F.counters[kNHuge] = nh;
        F.counters[kNLarge] = min(nl, F.kcap - nh);
        F.counters[kNSmall] = min(ns, F.kcap);
    }
}

// The cluster at global index g of the kernels of one cluster a block
// (l = 0: the huge clusters of every frame, list 4, then those of list 0;
// l = 4: the huge ones alone; l = 5: list 0 alone; l = 1: the small ones):
// frame fi and cluster c, false past the end. The cursors follow g, which
// only grows.
struct TakeCursor {
    uint32_t fh = 0, fl = 0;
};
// Wave: g is the same in every lane of the wave, which all take part (the
// frame found by the wave's 64-way search); otherwise by each lane alone.
template <bool Wave = true>
__device__ __forceinline__ bool cluster_at(const Frame *frames, const WorkQueue *Q, int l, uint32_t g,
                                           TakeCursor &cur, uint32_t *fi, uint32_t *c) {
    if (l == 0 || l == 4) {
        const uint32_t th = Q->total[4];
        if (g < th) {
            const uint32_t *off = Q->off[4];
            cur.fh = Wave ? frame_of_wave(off, Q->nframes, g, cur.fh) : frame_of(off, Q->nframes, g, cur.fh);
            *fi = cur.fh;
            *c = frames[cur.fh].klist[g - off[cur.fh]];
            return true;
        }
        if (l == 4) return false;
        g -= th;
    }
    if (l == 5) l = 0;
    if (g >= Q->total[l]) return false;
    const uint32_t *off = Q->off[l];
    cur.fl = Wave ? frame_of_wave(off, Q->nframes, g, cur.fl) : frame_of(off, Q->nframes, g, cur.fl);
    const Frame &F = frames[cur.fl];
    *fi = cur.fl;
    *c = F.klist[(l ? F.kcap : F.counters[kNHuge]) + (g - off[cur.fl])];
    return true;
}

// One block: where each frame's items start in list l (counter `which` of
// each frame), and the list's size.
// Block b builds list (lw >> 16 b) & 0xff from counter (lw >> 16 b + 8) & 0xff
// of every frame (one launch for up to four lists).
__global__ __launch_bounds__(256) void k_work_offsets(const Frame *frames, int nframes, WorkQueue *Q, uint64_t lw) {
    const uint32_t pair = (uint32_t)(lw >> (16 * blockIdx.x)) & 0xffffu;
    const int l = (int)(pair & 0xffu), which = (int)(pair >> 8);
    uint32_t carry = 0;
    for (int f0 = 0; f0 < nframes; f0 += 256) {
        const int f = f0 + (int)threadIdx.x;
        const uint32_t v = f < nframes ? frames[f].counters[which] : 0u;
        uint32_t tot;
        const uint32_t pre = bscan(v, &tot);
        if (f < nframes) Q->off[l][f] = carry + pre;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        Q->off[l][nframes] = carry;
        Q->total[l] = carry;
    }
}

void launch_work_offsets(hipStream_t st, const Frame *dframes, int nframes, WorkQueue *Q, int l, int which) {
    hipLaunchKernelGGL(k_work_offsets, dim3(1), dim3(256), 0, st, dframes, nframes, Q,
                       (uint64_t)(uint32_t)(l | which << 8));
}

// ------------------------------------------------------------ angle sort

// The pairs of one step of a bitonic network in which every comparison
// puts the smaller key first (the first step of each merge compares
// mirrored positions): keys past n act as +infinity and never move, so
// the network sorts any n in place. Pair q of the step (k, j) is (i, p):
// j = 0 is the mirror step of merge size k, j > 0 the step of distance j.
__device__ __forceinline__ void bitonic_pair(uint32_t q, uint32_t k, uint32_t j, uint32_t *i, uint32_t *p) {
    if (j == 0) {
        const uint32_t h = k >> 1, blk = q / h, r = q - blk * h;
        *i = blk * k + r;
        *p = blk * k + k - 1 - r;
    } else {
        *i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
        *p = *i + j;
    }
}

constexpr uint32_t kSortTile = 2048;  // keys of a cluster sorted together in LDS
constexpr uint32_t kBigSort = 8192;   // keys of a cluster sorted by one 1024-thread block


template <int R>
__device__ __forceinline__ void wave_sort_keys(const uint64_t *in, uint64_t *out, uint32_t n) {
    const int lane = threadIdx.x & 63;
    uint64_t v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t i = (uint32_t)(r * 64 + lane);
        v[r] = i < n ? gmem(in)[i] : ~0ull;
    }
    wave_bitonic<R>(v);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t i = (uint32_t)(r * 64 + lane);
        if (i < n) gmem(out)[i] = v[r];
    }
}

// The half-cleaner steps of distance J down to 1 of the all-ascending
// network on a wave's 512 keys in registers (key r * 64 + lane in v[r]):
// the smaller key to the lower position.
template <int J = 256>
__device__ __forceinline__ void reg_half_clean(uint64_t (&v)[8]) {
    const int lane = threadIdx.x & 63;
    if constexpr (J >= 64) {
        constexpr int jr = J >> 6;
#pragma unroll
        for (int r = 0; r < 8; r++) {
            if (r & jr) continue;
            const int r2 = r | jr;
            const uint64_t a = v[r], b = v[r2];
            v[r] = a < b ? a : b;
            v[r2] = a < b ? b : a;
        }
    } else {
#pragma unroll
        for (int r = 0; r < 8; r++) v[r] = cx_take(v[r], lane_xor64<J>(v[r]), (lane & J) == 0);
    }
    if constexpr (J > 1) reg_half_clean<J / 2>(v);
}

// One step of the network whose partners sit in other waves: through LDS
// (sk, kSortTile keys), partner p = position ^ x.
// act: the wave takes part (a wave whose keys lie beyond every pair of
// the step only meets the barriers)
__device__ __forceinline__ void lds_xor_step(uint64_t (&v)[8], uint64_t *sk, uint32_t x, bool act = true) {
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (act) {
#pragma unroll
        for (int r = 0; r < 8; r++) sk[w * 512 + r * 64 + lane] = v[r];
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const uint32_t i = w * 512 + r * 64 + lane, p = i ^ x;
            const uint64_t o = sk[p];
            v[r] = i < p ? (o < v[r] ? o : v[r]) : (o > v[r] ? o : v[r]);
        }
    }
    __syncthreads();
}

// The keys of each tile of kSortTile = 4 x 512 (a block's waves, 8 a lane in
// registers) of src[0, n) into dst: sorted (kk = 0), or the steps of
// distance below kSortTile of the merge of size kk. The merges across
// waves go through LDS, everything inside a wave in registers.
__device__ __forceinline__ void sort_tiles(const uint64_t *src, uint64_t *dst, uint32_t n, uint64_t *sk, uint32_t kk) {
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint32_t t0 = 0; t0 < n; t0 += kSortTile) {
        const uint32_t tn = min(kSortTile, n - t0);
        uint64_t v[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const uint32_t i = w * 512 + r * 64 + lane;
            v[r] = i < tn ? gmem(src)[t0 + i] : ~0ull;
        }
        if (kk == 0) {
            wave_bitonic<8>(v);        // each wave's 512 keys, ascending
            lds_xor_step(v, sk, 1023);  // merge of 1024: the mirror step
            reg_half_clean(v);
            lds_xor_step(v, sk, 2047);  // merge of 2048: the mirror step
            lds_xor_step(v, sk, 512);
            reg_half_clean(v);
        } else {
            lds_xor_step(v, sk, 1024);
            lds_xor_step(v, sk, 512);
            reg_half_clean(v);
        }
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const uint32_t i = w * 512 + r * 64 + lane;
            if (i < tn) gmem(dst)[t0 + i] = v[r];
        }
        __syncthreads();
    }
}

// Persistent over list 0 of kept clusters (above kLargeCluster points):
// the full keys of a cluster (slope << 32 | ~index, written by k_fit_prep
// in point order) sorted ascending, tiles of 2048 in registers and LDS, the
// wider steps over the cluster's keys in HBM. The keys are distinct, so
// any sort gives the full-key order the reference's sort starts from
// (fast/quads.c:379-427); its tie fix follows in k_fit_moments.
template <bool FIX>
__global__ __launch_bounds__(256) void k_angle_sort(const Frame *frames, const WorkQueue *Q, int l) {
    static_assert(kSortTile == 2048, "4 waves x 512 keys");
    constexpr bool big = FIX;  // calls of a few frames: the big sort takes the huge clusters, their keys fixed here
    __shared__ uint64_t sk[kSortTile];
    __shared__ uint32_t fls[FIX ? 4 * 256 + 1 : 1], flm[FIX ? 4 * 256 + 1 : 1], fcnt;
    const uint32_t t = threadIdx.x;
    TakeCursor cur;
    uint32_t fi, c;
    for (uint32_t g = blockIdx.x; cluster_at(frames, Q, l, g, cur, &fi, &c); g += gridDim.x) {
        const Frame &F = frames[fi];
        const uint32_t s0 = F.cstart[c], n = F.cstart[c + 1] - s0;
        if (big && n > kHugeCluster && n <= kBigSort) continue;  // k_angle_sort_big's
        uint64_t *keys = F.skey + s0;
        uint32_t m = kSortTile;
        while (m < n) m <<= 1;
        sort_tiles(F.okey + s0, keys, n, sk, 0);  // the unsorted keys stay for the tie fix
        for (uint32_t k = 2 * kSortTile; k <= m; k <<= 1) {
            for (uint32_t j = 0; j == 0 || j >= kSortTile; j = j ? j >> 1 : k >> 2) {
                for (uint32_t q = t; q < m / 2; q += 256) {
                    uint32_t i, p;
                    bitonic_pair(q, k, j, &i, &p);
                    if (p >= n) continue;
                    const uint64_t a = keys[i], b = keys[p];
                    if (a > b) {
                        keys[i] = b;
                        keys[p] = a;
                    }
                }
                __syncthreads();
            }
            sort_tiles(keys, keys, n, sk, k);
        }
        // the tie fix on the sorted keys (in LDS when they fit a tile);
        // for batches k_fit_moments fixes the large clusters' keys
        if constexpr (!FIX) {
        } else if (n <= kSortTile) {
            for (uint32_t i = t; i < n; i += 256) sk[i] = gmem(keys)[i];
            __syncthreads();
            fix_runs_block<256>(sk, n, F.okey + s0, reinterpret_cast<uint64_t *>(F.errs + s0),
                                reinterpret_cast<uint64_t *>(F.filt + s0), fls, flm, &fcnt);
            for (uint32_t i = t; i < n; i += 256) gmem(keys)[i] = sk[i];
            __syncthreads();
        } else {
            fix_runs_block<256>(keys, n, F.okey + s0, reinterpret_cast<uint64_t *>(F.errs + s0),
                                reinterpret_cast<uint64_t *>(F.filt + s0), fls, flm, &fcnt);
        }
    }
}

// The angle sort of the clusters of kHugeCluster + 1 to kBigSort points
// (of the huge list), for calls of a few frames, a cluster per block of
// 1024 threads: the keys in the
// registers of 16 waves (512 a wave), each wave's sorted in registers, the
// merges across waves through LDS (skb, dynamic), the steps inside a wave
// in registers; padding keys (all ones) sort last.
__global__ __launch_bounds__(1024) void k_angle_sort_big(const Frame *frames, const WorkQueue *Q) {
    static_assert(kBigSort == 16 * 512, "16 waves x 512 keys");
    extern __shared__ uint64_t skb[];
    __shared__ uint32_t fls[4 * 1024 + 1], flm[4 * 1024 + 1], fcnt;
    const uint32_t t = threadIdx.x, w = t >> 6, lane = t & 63;
    const uint32_t total = Q->total[4];
    TakeCursor cur;
    uint32_t fi, c;
    for (uint32_t g = blockIdx.x; g < total && cluster_at(frames, Q, 0, g, cur, &fi, &c); g += gridDim.x) {
        const Frame &F = frames[fi];
        const uint32_t s0 = F.cstart[c], n = F.cstart[c + 1] - s0;
        if (n <= kHugeCluster || n > kBigSort) continue;
        uint32_t m = 1024;
        while (m < n) m <<= 1;
        // the waves of keys [0, m): every step pairs keys below m, so the
        // waves beyond (padding only, sorted) just meet the barriers
        const bool act = w * 512 < m;
        uint64_t v[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const uint32_t i = w * 512 + (uint32_t)r * 64 + lane;
            v[r] = i < n ? gmem(F.okey)[s0 + i] : ~0ull;
        }
        if (act) wave_bitonic<8>(v);
        for (uint32_t k = 1024; k <= m; k <<= 1) {
            lds_xor_step(v, skb, k - 1, act);  // the merge's mirror step
            for (uint32_t j = k >> 2; j >= 512; j >>= 1) lds_xor_step(v, skb, j, act);
            if (act) reg_half_clean(v);
        }
        // the tie fix on the sorted keys, in LDS (one segment of 8192)
        if (act) {
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const uint32_t i = w * 512 + (uint32_t)r * 64 + lane;
                skb[i] = v[r];
            }
        }
        __syncthreads();
        fix_runs_block<1024>(skb, n, F.okey + s0, reinterpret_cast<uint64_t *>(F.errs + s0),
                             reinterpret_cast<uint64_t *>(F.filt + s0), fls, flm, &fcnt);
        for (uint32_t i = t; i < n; i += 1024) gmem(F.skey)[s0 + i] = skb[i];
        __syncthreads();
    }
}

// The angle sort of list 1 (at most kLargeCluster points), a wave per
// cluster, persistent.
template <int R>
__device__ __forceinline__ void wave_sort_keys_lds(const uint64_t *in, uint64_t *out, uint32_t n) {
    const int lane = threadIdx.x & 63;
    uint64_t v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t i = (uint32_t)(r * 64 + lane);
        v[r] = i < n ? gmem(in)[i] : ~0ull;
    }
    wave_bitonic<R>(v);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t i = (uint32_t)(r * 64 + lane);
        if (i < n) out[i] = v[r];
    }
}

// (96 registers: five waves a SIMD instead of four)
__global__ __launch_bounds__(64, 5) void k_angle_sort_wave(const Frame *frames, const WorkQueue *Q) {
    static_assert(kLargeCluster <= 512, "eight keys a lane at most");
    __shared__ uint64_t kl[kLargeCluster];
    __shared__ uint32_t fls[4 * 64 + 1], flm[4 * 64 + 1], fcnt;
    const uint32_t total = Q->total[1];
    const uint32_t *off = Q->off[1];
    const int lane = threadIdx.x;
    uint32_t fi = 0;
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
        fi = frame_of_wave(off, Q->nframes, g, fi);
        const Frame &F = frames[fi];
        const uint32_t c = F.klist[F.kcap + (g - off[fi])];
        const uint32_t s0 = F.cstart[c], n = F.cstart[c + 1] - s0;
        if (n <= 64)
            wave_sort_keys_lds<1>(F.okey + s0, kl, n);
        else if (n <= 128)
            wave_sort_keys_lds<2>(F.okey + s0, kl, n);
        else if (n <= 256)
            wave_sort_keys_lds<4>(F.okey + s0, kl, n);
        else
            wave_sort_keys_lds<8>(F.okey + s0, kl, n);
        __syncthreads();
        // the tie fix on the sorted keys (fast/quads.c:304-327)
        fix_runs_block<64>(kl, n, F.okey + s0, reinterpret_cast<uint64_t *>(F.errs + s0),
                           reinterpret_cast<uint64_t *>(F.filt + s0), fls, flm, &fcnt);
        for (uint32_t i = (uint32_t)lane; i < n; i += 64) gmem(F.skey)[s0 + i] = kl[i];
        __syncthreads();
    }
}

// ------------------------------------------------------------ moments

// Lane q < 6 of a wave: running sum q continued from s over the terms
// a[0, len) (LDS), each partial sum stored to out[0, len): in place over
// its term (InPlace: a position is rewritten only after its term was read
// into the ring; the block then copies the sums out with whole-wave
// stores), or the cluster's moment array q in HBM straight from the chain
// (no LDS writes beside the ring's reads: faster where the chain itself is
// the critical path, the long chunks of single frames). Each of the kRing
// registers takes the term kRing ahead as soon as its add has used it.
// Returns the sum.
template <bool InPlace>
__device__ __forceinline__ double scan_chunk(double *a, uint32_t len, double s, double *out_p) {
    auto out = [&]() {
        if constexpr (InPlace) return a;
        else return gmem(out_p);
    }();
    constexpr uint32_t kRing = ATG_SCAN_RING;
    static_assert(kRing % 2 == 0, "pairs of sums a store");
    uint32_t j = 0;
    if (len >= kRing) {
        double v[kRing];
#pragma unroll
        for (uint32_t q = 0; q < kRing; q++) v[q] = a[q];
        for (; j + 2 * kRing <= len; j += kRing) {
#pragma unroll
            for (uint32_t q = 0; q < kRing; q += 2) {
                const double s0 = __dadd_rn(s, v[q]);
                const double s1 = __dadd_rn(s0, v[q + 1]);
                s = s1;
                v[q] = a[j + kRing + q];
                v[q + 1] = a[j + kRing + q + 1];
                out[j + q] = s0;
                out[j + q + 1] = s1;
            }
        }
#pragma unroll
        for (uint32_t q = 0; q < kRing; q++) {
            s = __dadd_rn(s, v[q]);
            out[j + q] = s;
        }
        j += kRing;
    }
    // the rest (fewer than kRing), its reads in flight together
    if (j < len) {
        double v[kRing];
#pragma unroll
        for (uint32_t q = 0; q < kRing; q++) v[q] = j + q < len ? a[j + q] : 0.0;
#pragma unroll
        for (uint32_t q = 0; q < kRing; q++) {
            if (j + q < len) {
                s = __dadd_rn(s, v[q]);
                out[j + q] = s;
            }
        }
    }
    return s;
}

// The moment terms of points [c0, c0 + len) of a cluster in key order, in
// three steps (so that a caller can keep the loads a chunk or two ahead):
// the points' indices from the keys, their packed (x, y, g2) words, and
// the six terms (fast/quads.c:968-1035: weight sqrt(g2) + 1, centres
// x / 2 + 1 / 2) into b[q][j], for j = t0, t0 + stride, ... (KP of them).
template <int KP>
__device__ __forceinline__ void term_points(const uint64_t *keys, uint32_t c0, uint32_t len, int t0, int stride,
                                            uint32_t (&ix)[KP]) {
#pragma unroll
    for (int r = 0; r < KP; r++) {
        const uint32_t j = (uint32_t)(t0 + r * stride);
        ix[r] = j < len ? (uint32_t)~(uint32_t)gmem(keys)[c0 + j] : 0u;
    }
}
template <int KP>
__device__ __forceinline__ void term_words(const Frame &F, uint32_t s0, uint32_t len, int t0, int stride,
                                           const uint32_t (&ix)[KP], uint64_t (&pwv)[KP]) {
#pragma unroll
    for (int r = 0; r < KP; r++) {
        const uint32_t j = (uint32_t)(t0 + r * stride);
        pwv[r] = j < len ? gmem(F.pw)[s0 + ix[r]] : 0ull;
    }
}
template <int KP, int CH>
__device__ __forceinline__ void term_values(const uint64_t (&pwv)[KP], uint32_t len, double (*b)[CH + 2], int t0,
                                            int stride) {
#pragma unroll
    for (int r = 0; r < KP; r++) {
        const uint32_t j = (uint32_t)(t0 + r * stride);
        if (j >= len) break;
        const int px = (int)(pwv[r] & 0xffffu), py = (int)((pwv[r] >> 16) & 0xffffu);
        const uint32_t g2 = (uint32_t)(pwv[r] >> 32);
        const double wgt = __dadd_rn(__dsqrt_rn((double)g2), 1.0);
        const double fx = __dadd_rn(__dmul_rn((double)px, 0.5), 0.5);
        const double fy = __dadd_rn(__dmul_rn((double)py, 0.5), 0.5);
        const double wfx = __dmul_rn(wgt, fx), wfy = __dmul_rn(wgt, fy);
        b[0][j] = wfx;
        b[1][j] = wfy;
        b[2][j] = __dmul_rn(wfx, fx);
        b[3][j] = __dmul_rn(wfx, fy);
        b[4][j] = __dmul_rn(wfy, fy);
        b[5][j] = wgt;
    }
}
template <int KP, int CH>
__device__ __forceinline__ void moment_terms(const Frame &F, uint32_t s0, const uint64_t *keys, uint32_t c0,
                                             uint32_t len, double (*b)[CH + 2], int t0, int stride) {
    uint32_t ix[KP];
    uint64_t pwv[KP];
    term_points<KP>(keys, c0, len, t0, stride, ix);
    term_words<KP>(F, s0, len, t0, stride, ix, pwv);
    term_values<KP, CH>(pwv, len, b, t0, stride);
}

// The running sums of positions j = t0, t0 + stride, ... (KP of them, below
// len) of a chunk at c0 from LDS to the cluster's six moment arrays
// (array q at mom + q pc), consecutive positions on consecutive lanes.
template <int KP, int CH>
__device__ __forceinline__ void sums_out(const double (*b)[CH + 2], uint32_t len, double *mom, int64_t pc, uint32_t c0,
                                         int t0, int stride) {
    auto m = gmem(mom);
#pragma unroll
    for (int r = 0; r < KP; r++) {
        const uint32_t j = (uint32_t)(t0 + r * stride);
        if (j >= len) break;
#pragma unroll
        for (int q = 0; q < 6; q++) m[(int64_t)q * pc + c0 + j] = b[q][j];
    }
}

// A barrier of the block that orders LDS only (loads from HBM issued before
// it stay in flight).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Persistent over list l of kept clusters, block b taking global indices
// b, b + G, ... (a frame cursor): NT threads a cluster, the running sums
// in chunks of CH points. With more than one wave the chunks are double
// buffered: lanes 0-5 of wave 0 run the six sums over one chunk while the
// other waves form the next chunk's terms.
template <int NT, int CH>
__global__ __launch_bounds__(NT, NT == 128 ? 8 : 1) void k_fit_moments(const Frame *frames, int parity, const WorkQueue *Q,
                                                   int keys_ready, int l, int fix) {
    constexpr int kBufs = NT > 64 ? 2 : 1;
    // the sums in place and copied out by the block (short chunks of many
    // clusters at once), or stored by the chain lanes
    constexpr bool kInPlace = kBufs == 2 && CH <= 128;
    // rows 2 doubles apart in the banks: the scan's lanes conflict-free
    __shared__ double sm[kBufs][6][CH + 2];
    __shared__ uint32_t fls[4 * NT + 1], flm[4 * NT + 1], fcnt;
    const int t = threadIdx.x, w = t >> 6;
    {
    TakeCursor cur;
    uint32_t fi, c;
    for (uint32_t g = blockIdx.x; cluster_at(frames, Q, l, g, cur, &fi, &c); g += gridDim.x) {
        const Frame &F = frames[fi];
        uint64_t *skey = parity && !keys_ready ? F.skey2 : F.skey;
        const uint32_t *sval = parity ? F.sval2 : F.sval;
        const int64_t pc = F.pcap;
        const uint32_t s0 = F.cstart[c], n = F.cstart[c + 1] - s0;
        uint64_t *keys = skey + s0;
        if (!keys_ready) {
            // the full keys in order: slope << 32 | ~index, from the sorted
            // (cluster << 32 | slope, index) items
            for (uint32_t k = t; k < n; k += NT)
                keys[k] = ((keys[k] & 0xffffffffull) << 32) | (uint32_t)~sval[s0 + k];
            __syncthreads();
        }
        // fix_runs (fast/quads.c:304-327), the keys in LDS (the term
        // buffers, free until then) when they fit, unless the angle sort
        // fixed them (fix 0)
        if (fix) {
            constexpr uint32_t kFixLds = kBufs * 6 * CH;
            uint64_t *kl = reinterpret_cast<uint64_t *>(&sm[0][0][0]);
            const bool in_lds = n <= kFixLds;
            // (the copies four keys a thread at a time, their loads together)
            if (in_lds) {
                for (uint32_t i0 = 0; i0 < n; i0 += 4 * NT) {
                    uint64_t cv[4];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const uint32_t i = i0 + t + (uint32_t)(r * NT);
                        cv[r] = i < n ? gmem(keys)[i] : 0ull;
                    }
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const uint32_t i = i0 + t + (uint32_t)(r * NT);
                        if (i < n) kl[i] = cv[r];
                    }
                }
                __syncthreads();
            }
            fix_runs_block<NT>(in_lds ? kl : keys, n, F.okey + s0, reinterpret_cast<uint64_t *>(F.errs + s0),
                               reinterpret_cast<uint64_t *>(F.filt + s0), fls, flm, &fcnt);
            if (in_lds) {
                for (uint32_t i0 = 0; i0 < n; i0 += 4 * NT) {
                    uint64_t cv[4];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const uint32_t i = i0 + t + (uint32_t)(r * NT);
                        cv[r] = i < n ? kl[i] : 0ull;
                    }
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const uint32_t i = i0 + t + (uint32_t)(r * NT);
                        if (i < n) gmem(keys)[i] = cv[r];
                    }
                }
                __syncthreads();
            }
        }
        // the six running sums (fast/quads.c:1006-1035), stored by the
        // chain lanes into the moment arrays
        double run = 0.0;  // lane q < 6: running sum q
        double *mom = F.mom + s0;
        double *out = mom + (t < 6 ? t : 0) * pc;
        const uint32_t nch = (n + CH - 1) / CH;
        if constexpr (kBufs == 2) {
            // the other waves keep their loads ahead of the chain: the
            // points of the chunk after next and the words of the next one
            // in registers (the barriers order LDS only)
            constexpr int KP = (CH + NT - 65) / (NT - 64);
            const int th = t - 64, sh = NT - 64;
            auto clen = [&](uint32_t ci) { return ci < nch ? min((uint32_t)CH, n - ci * CH) : 0u; };
            uint32_t ix[KP];
            uint64_t pwv[KP];
            moment_terms<(CH + NT - 1) / NT, CH>(F, s0, keys, 0, clen(0), sm[0], t, NT);
            if (w > 0) {
                term_points<KP>(keys, CH, clen(1), th, sh, ix);
                term_words<KP>(F, s0, clen(1), th, sh, ix, pwv);
                term_points<KP>(keys, 2 * CH, clen(2), th, sh, ix);
            }
            __syncthreads();
            for (uint32_t ci = 0; ci < nch; ci++) {
                const uint32_t len = clen(ci);
                if (w == 0) {
                    if (t < 6) run = scan_chunk<kInPlace>(sm[ci & 1][t], len, run, out + ci * CH);
                } else {
                    // the sums of the chunk before out of the other buffer,
                    // each thread its own positions, which it then refills
                    // with the next chunk's terms
                    if (kInPlace && ci > 0)
                        sums_out<KP, CH>(sm[(ci + 1) & 1], clen(ci - 1), mom, pc, (ci - 1) * CH, th, sh);
                    if (ci + 1 < nch) {
                        term_values<KP, CH>(pwv, clen(ci + 1), sm[(ci + 1) & 1], th, sh);
                        term_words<KP>(F, s0, clen(ci + 2), th, sh, ix, pwv);
                        term_points<KP>(keys, (ci + 3) * CH, clen(ci + 3), th, sh, ix);
                    }
                }
                lds_barrier();
            }
            if (kInPlace && nch > 0) {
                sums_out<(CH + NT - 1) / NT, CH>(sm[(nch - 1) & 1], clen(nch - 1), mom, pc, (nch - 1) * CH, t, NT);
                __syncthreads();  // the buffers are the next cluster's
            }
        } else {
            for (uint32_t ci = 0; ci < nch; ci++) {
                const uint32_t c0 = ci * CH, len = min((uint32_t)CH, n - c0);
                moment_terms<(CH + NT - 1) / NT, CH>(F, s0, keys, c0, len, sm[0], t, NT);
                __syncthreads();
                if (t < 6) run = scan_chunk<false>(sm[0][t], len, run, out + c0);
                __syncthreads();
            }
        }
    }
    }
}

// The moments of batches, G clusters a wave (persistent over groups of G
// consecutive entries of list l; their keys sorted and fixed by the angle
// sort): the six running sums of all G clusters advance together, lane 6 k + q holding sum q of cluster
// k: for each chunk of C positions the wave forms the terms of every
// cluster's points there into LDS, the chain lanes run their sums over
// them in registers (each sum in its cluster's order, as the reference
// adds them), and the wave copies the partial sums out to the moment
// arrays. A step of the chains costs one wave instruction for G clusters.
template <int G, int C>
__global__ __launch_bounds__(64) void k_fit_moments_packed(const Frame *frames, const WorkQueue *Q, int l) {
    static_assert(6 * G <= 64 && (G * C) % 64 == 0, "a lane a chain; whole items a lane");
    __shared__ double tb[G][6][C + 2];  // rows 2 doubles apart: the chain lanes' reads spread over the banks
    // each cluster of the group: its keys, point words and moment arrays
    // (at its first point), the arrays' stride, its size
    __shared__ const uint64_t *gkey[G];
    __shared__ const uint64_t *gpw[G];
    __shared__ double *gmom[G];
    __shared__ int64_t gpc[G];
    __shared__ uint32_t gn[G];
    const int lane = threadIdx.x;
    const uint32_t total = Q->total[l] + (l == 0 ? Q->total[4] : 0u);
    TakeCursor cur;  // this lane's (entries g0 + lane, lane < G)
    for (uint32_t g0 = blockIdx.x * G; g0 < total; g0 += gridDim.x * G) {
        if (lane < G) {
            uint32_t fi = 0, c = 0;
            const bool ok = g0 + lane < total && cluster_at<false>(frames, Q, l, g0 + lane, cur, &fi, &c);
            const Frame &F = frames[ok ? fi : 0];
            const uint32_t s0 = ok ? F.cstart[c] : 0u;
            gn[lane] = ok ? F.cstart[c + 1] - s0 : 0u;
            gkey[lane] = F.skey + s0;
            gpw[lane] = F.pw + s0;
            gmom[lane] = F.mom + s0;
            gpc[lane] = F.pcap;
        }
        __syncthreads();
        uint32_t maxlen = 0;
        for (int k = 0; k < G; k++) maxlen = max(maxlen, gn[k]);
        // the packed running sums
        const int ck = lane / 6, cq = lane % 6;
        const bool chain = lane < 6 * G;
        double run = 0.0;
        // terms of (cluster k, position c0 + j), item k C + j: the points'
        // indices two chunks ahead and their words one chunk ahead, so that
        // a chunk's loads are in flight through the chunk before it
        constexpr int kItems = G * C / 64;
        auto load_ix = [&](uint32_t c0, uint32_t (&ix)[kItems]) {
#pragma unroll
            for (int r = 0; r < kItems; r++) {
                const int it = lane + 64 * r, k = it / C, j = it % C;
                ix[r] = c0 + (uint32_t)j < gn[k] ? (uint32_t)~(uint32_t)gmem(gkey[k])[c0 + j] : 0xffffffffu;
            }
        };
        auto load_pw = [&](const uint32_t (&ix)[kItems], uint64_t (&pwv)[kItems]) {
#pragma unroll
            for (int r = 0; r < kItems; r++) {
                const int it = lane + 64 * r, k = it / C;
                pwv[r] = ix[r] != 0xffffffffu ? gmem(gpw[k])[ix[r]] : 0ull;
            }
        };
        uint32_t ix[kItems], ixn[kItems];
        uint64_t pwv[kItems];
        load_ix(0, ix);
        load_pw(ix, pwv);
        load_ix(C, ixn);
        for (uint32_t c0 = 0; c0 < maxlen; c0 += C) {
            uint32_t ixc[kItems];
            uint64_t pwc[kItems];
#pragma unroll
            for (int r = 0; r < kItems; r++) {
                ixc[r] = ix[r];
                pwc[r] = pwv[r];
            }
            if (c0 + C < maxlen) {
#pragma unroll
                for (int r = 0; r < kItems; r++) ix[r] = ixn[r];
                load_pw(ix, pwv);
                load_ix(c0 + 2 * C, ixn);
            }
#pragma unroll
            for (int r = 0; r < kItems; r++) {
                const int it = lane + 64 * r, k = it / C, j = it % C;
                if (ixc[r] == 0xffffffffu) continue;
                const int px = (int)(pwc[r] & 0xffffu), py = (int)((pwc[r] >> 16) & 0xffffu);
                const uint32_t g2 = (uint32_t)(pwc[r] >> 32);
                const double wgt = __dadd_rn(__dsqrt_rn((double)g2), 1.0);
                const double fx = __dadd_rn(__dmul_rn((double)px, 0.5), 0.5);
                const double fy = __dadd_rn(__dmul_rn((double)py, 0.5), 0.5);
                const double wfx = __dmul_rn(wgt, fx), wfy = __dmul_rn(wgt, fy);
                tb[k][0][j] = wfx;
                tb[k][1][j] = wfy;
                tb[k][2][j] = __dmul_rn(wfx, fx);
                tb[k][3][j] = __dmul_rn(wfx, fy);
                tb[k][4][j] = __dmul_rn(wfy, fy);
                tb[k][5][j] = wgt;
            }
            __syncthreads();
            // the chains: each lane its sum over its cluster's positions here
            if (chain) {
                const uint32_t n = gn[ck];
                const uint32_t len = c0 < n ? min((uint32_t)C, n - c0) : 0u;
                double *a = tb[ck][cq];
                double v[C];
#pragma unroll
                for (int j = 0; j < C; j++) v[j] = a[j];
#pragma unroll
                for (int j = 0; j < C; j++)
                    if ((uint32_t)j < len) {
                        run = __dadd_rn(run, v[j]);
                        v[j] = run;
                    }
#pragma unroll
                for (int j = 0; j < C; j++) a[j] = v[j];
            }
            __syncthreads();
            // out to the moment arrays, consecutive positions on consecutive lanes
            for (int it = lane; it < G * 6 * C; it += 64) {
                const int k = it / (6 * C), q = (it / C) % 6, j = it % C;
                if (c0 + (uint32_t)j < gn[k]) gmem(gmom[k])[(int64_t)q * gpc[k] + c0 + j] = tb[k][q][j];
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------ corners

// The six cumulative moment arrays of a cluster (mx, my, mxx, mxy, myy, w),
// cluster-relative: array q starts at b + q * stride.
struct Mom {
    const __attribute__((address_space(1))) double *b;
    int64_t stride;
    __device__ __forceinline__ const __attribute__((address_space(1))) double *a(int q) const { return b + q * stride; }
};

// The window sums of [i0, i1], wrapping, as fit_line forms them
// (fast/quads.c:539-572); a wrapping range has i0 > i1 >= 0. Every operand
// is loaded whichever case applies (the loads of all six sums in flight at
// once, without divergent branches), the case picks the result.
__device__ __forceinline__ void range_sums(const Mom &M, int64_t sz, int64_t i0, int64_t i1, double s[6]) {
    const bool wrap = i0 >= i1, head = i0 > 0;
    const int64_t ib = head ? i0 - 1 : 0;
    double a[6], b[6], c[6];
    #pragma unroll
    for (int q = 0; q < 6; q++) {
        const auto v = M.a(q);
        a[q] = v[sz - 1];
        b[q] = v[ib];
        c[q] = v[i1];
    }
    #pragma unroll
    for (int q = 0; q < 6; q++) {
        const double r1 = __dadd_rn(__dsub_rn(a[q], b[q]), c[q]), r2 = __dsub_rn(c[q], b[q]);
        s[q] = wrap ? r1 : (head ? r2 : c[q]);
    }
}

// fit_line (fast/quads.c:539-600) from the sums: line (ex, ey, nx, ny), mse.
__device__ __forceinline__ void line_from_sums(const double s[6], double lp[4], double *mse) {
    double w = s[5];
    double ex = __ddiv_rn(s[0], w), ey = __ddiv_rn(s[1], w);
    double cxx = __dsub_rn(__ddiv_rn(s[2], w), __dmul_rn(ex, ex));
    double cxy = __dsub_rn(__ddiv_rn(s[3], w), __dmul_rn(ex, ey));
    double cyy = __dsub_rn(__ddiv_rn(s[4], w), __dmul_rn(ey, ey));
    double dd = __dsub_rn(cxx, cyy);
    double arg = __dadd_rn(__dmul_rn(dd, dd), __dmul_rn(__dmul_rn(4.0, cxy), cxy));
    double root = (double)fsqrt_cr((float)arg);
    double es = __dmul_rn(0.5, __dsub_rn(__dadd_rn(cxx, cyy), root));
    double eig = __dmul_rn(0.5, __dadd_rn(__dadd_rn(cxx, cyy), root));
    double nx1 = __dsub_rn(cxx, eig), ny1 = cxy, m1 = __dadd_rn(__dmul_rn(nx1, nx1), __dmul_rn(ny1, ny1));
    double nx2 = cxy, ny2 = __dsub_rn(cyy, eig), m2 = __dadd_rn(__dmul_rn(nx2, nx2), __dmul_rn(ny2, ny2));
    double nx = nx2, ny = ny2, m = m2;
    if (m1 > m2) { nx = nx1; ny = ny1; m = m1; }
    double len = (double)fsqrt_cr((float)m);
    lp[0] = ex;
    lp[1] = ey;
    lp[2] = 0.0;
    lp[3] = 0.0;
    if (!(fabs(len) < 1e-12)) {
        lp[2] = __ddiv_rn(nx, len);
        lp[3] = __ddiv_rn(ny, len);
    }
    *mse = es;
}

// The window error of window_errors / moment_errors (fast/quads.c:604-652).
__device__ __forceinline__ double win_err(const double s[6], double nn) {
    double w = s[5];
    double ex = __ddiv_rn(s[0], w), ey = __ddiv_rn(s[1], w);
    double cxx = __dsub_rn(__ddiv_rn(s[2], w), __dmul_rn(ex, ex));
    double cxy = __dsub_rn(__ddiv_rn(s[3], w), __dmul_rn(ex, ey));
    double cyy = __dsub_rn(__ddiv_rn(s[4], w), __dmul_rn(ey, ey));
    double d = __dsub_rn(cxx, cyy);
    double root = (double)fsqrt_cr((float)__dadd_rn(__dmul_rn(d, d), __dmul_rn(4.0, __dmul_rn(cxy, cxy))));
    return __dmul_rn(nn, __dmul_rn(0.5, __dsub_rn(__dadd_rn(cxx, cyy), root)));
}

// order-preserving integer of a double
__device__ __forceinline__ uint64_t dkey(double d) {
    uint64_t b = (uint64_t)__double_as_longlong(d);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
// the double of an order-preserving integer
__device__ __forceinline__ double dkey_value(uint64_t k) {
    return __longlong_as_double((long long)((k >> 63) ? (k ^ 0x8000000000000000ull) : ~k));
}

// The first R rows of a wave's keys (R = 1, 2 or 4; key i = r * 64 + lane
// in kv[r]) sorted ascending in place.
__device__ __forceinline__ void wave_sort_rows(uint64_t (&kv)[4], int R) {
    if (R == 1) {
        uint64_t a[1] = {kv[0]};
        wave_bitonic<1>(a);
        kv[0] = a[0];
    } else if (R == 2) {
        uint64_t a[2] = {kv[0], kv[1]};
        wave_bitonic<2>(a);
        kv[0] = a[0];
        kv[1] = a[1];
    } else {
        wave_bitonic<4>(kv);
    }
}

// Key pos (wave-uniform) of a wave's rows, in every lane.
__device__ __forceinline__ uint64_t wave_key_at(const uint64_t (&kv)[4], uint32_t pos) {
    const uint64_t row = pos < 64 ? kv[0] : pos < 128 ? kv[1] : pos < 192 ? kv[2] : kv[3];
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(row >> 32), (int)(pos & 63)) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)row, (int)(pos & 63));
}

// The triples m0 < m1 < m2 in colex order (by m2, then m1, then m0): the
// first C(k, 3) of them are those with m2 < k. Packed m0 | m1 << 5 | m2 << 10.
struct TriTable {
    uint16_t v[4495];  // C(31, 3): m2 up to 30
    constexpr TriTable() : v() {
        int n = 0;
        for (int m2 = 2; m2 < 31; m2++)
            for (int m1 = 1; m1 < m2; m1++)
                for (int m0 = 0; m0 < m1; m0++) v[n++] = (uint16_t)(m0 | m1 << 5 | m2 << 10);
    }
};
__constant__ const TriTable kTri{};

// Block shapes of the corner search: 256 threads for the large clusters,
// a wave for the rest (at most kLargeCluster points, all in LDS).
template <int NT>
struct CornerShape;
template <>
struct CornerShape<256> {
    static constexpr int kMaxNm = kMaxMaxima;  // maxima kept at most
    static constexpr int kLdsPts = 1024;  // clusters up to this size keep their errors in LDS
    static constexpr int kMaxima = 1024;  // maxima values in LDS at most
};
template <>
struct CornerShape<64> {
    static constexpr int kMaxNm = 16;
    static constexpr int kLdsPts = (int)kLargeCluster;
    static constexpr int kMaxima = (int)kLargeCluster / 2;  // strict cyclic maxima of n points: at most n / 2
};

template <int NT>
__device__ __forceinline__ uint32_t bscan_t(uint32_t v, uint32_t *total) {
    if constexpr (NT == 64) {
        const uint32_t inc = wscan(v);
        *total = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
        return inc - v;
    } else {
        return bscan(v, total);
    }
}

// Persistent over list l of kept clusters, a cluster per block of NT threads.
template <int NT>
// (the 256-thread shape held to 128 registers: four blocks a CU, as its LDS allows)
__global__ __launch_bounds__(NT, 4) void k_fit_corners(const Frame *frames, const DevParams *P, const WorkQueue *Q,
                                                    int l) {
    constexpr int kMaxNm = CornerShape<NT>::kMaxNm, kLdsPts = CornerShape<NT>::kLdsPts;
    constexpr int kMaxima = CornerShape<NT>::kMaxima;
    // errors and filtered errors, and (256 threads) the waves' top maxima
    // keys of the threshold's merge
    constexpr int kPart1 = 2 * kLdsPts + (NT == 256 ? 256 : 0);
    constexpr int kShared = kPart1 > 4 * kMaxNm * kMaxNm ? kPart1 : 4 * kMaxNm * kMaxNm;
    constexpr int kWaves = NT / 64;
    // The errors, filtered errors and maxima values of a cluster (first
    // part) share LDS with the pair-fit tables (second part).
    __shared__ double sh[kShared];
    __shared__ uint32_t lidx[kMaxima];
    __shared__ uint32_t mx[kMaxNm];
    __shared__ uint32_t hist[NT == 256 ? 256 : 1];
    constexpr uint32_t kSegPts = NT == 256 ? 16384 : (uint32_t)kLdsPts;  // maxima flags a pass
    __shared__ uint64_t fw[kSegPts / 64];
    __shared__ uint64_t s_prefix;
    __shared__ uint32_t s_rank;
    __shared__ double s_wbest[kWaves];
    __shared__ uint32_t s_word[kWaves];
    __shared__ int64_t s_idx[4];
    __shared__ double s_lines[4][4];
    __shared__ float s_qp[4][2];
    __shared__ int s_ok, s_lok[4], s_dok[4], s_cvx[4];
    __shared__ double s_side[6];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const double max_mse = (double)P->max_mse, max_dot = (double)P->cos_crit;
    const double INF = __longlong_as_double(0x7ff0000000000000ll);
    double *pe = sh, *pm = sh + kMaxNm * kMaxNm, *pnx = sh + 2 * kMaxNm * kMaxNm, *pny = sh + 3 * kMaxNm * kMaxNm;
    {
    TakeCursor cur;
    uint32_t fi, c;
    for (uint32_t g = blockIdx.x; cluster_at(frames, Q, l, g, cur, &fi, &c); g += gridDim.x) {
        const Frame &F = frames[fi];
        const int64_t pc = F.pcap;
        uint32_t flag = F.cflag[c];
        if (!(flag & 1u)) continue;
        const uint32_t s0 = F.cstart[c];
        const int64_t sz = F.cstart[c + 1] - s0;
        const Mom M = {gmem(F.mom) + s0, pc};
        int64_t ksz = min((int64_t)20, sz / 12);
        if (ksz < 2) continue;
        double nn = (double)(2 * ksz + 1);
        const bool inl = sz <= kLdsPts;
        if (NT == 64 && !inl) {  // the list holds no such cluster
            if (t == 0) atomicOr(F.counters + kError, 4u);
            continue;
        }
        double *errs = inl ? sh : F.errs + s0;
        double *y = inl ? sh + kLdsPts : F.filt + s0;
        // window errors (fast/quads.c:654-686 edges, :604-635 middle); the
        // operands of all four cases loaded at once, the case picks the sum
        double tot[6];
        #pragma unroll
        for (int q = 0; q < 6; q++) tot[q] = M.a(q)[sz - 1];
        for (int64_t i = t; i < sz; i += NT) {
            const int64_t i0 = i - ksz, i1 = i + ksz;
            const bool lo = i0 < 0, hi = i1 >= sz, head = i0 > 0;  // ksz <= sz / 12: not both
            const int64_t ib = lo ? i0 + sz - 1 : (head ? i0 - 1 : 0), ic = hi ? i1 - sz : i1;
            double b[6], c[6], s[6];
            #pragma unroll
            for (int q = 0; q < 6; q++) {
                const auto v = M.a(q);
                b[q] = v[ib];
                c[q] = v[ic];
            }
            #pragma unroll
            for (int q = 0; q < 6; q++) {
                const double r1 = __dadd_rn(__dsub_rn(tot[q], b[q]), c[q]), r2 = __dsub_rn(c[q], b[q]);
                s[q] = (lo || hi) ? r1 : (head ? r2 : c[q]);
            }
            errs[i] = win_err(s, nn);
        }
        __syncthreads();
        // 7-tap filter, wrapping (fast/quads.c:706-735)
        for (int64_t i = t; i < sz; i += NT) {
            double acc;
            if (i >= 3 && i < sz - 3) {
                acc = __dmul_rn(errs[i - 3], P->qsm_kern[0]);
                for (int k = 1; k < 7; k++) acc = __dadd_rn(acc, __dmul_rn(errs[i - 3 + k], P->qsm_kern[k]));
            } else {
                acc = 0.0;
                int64_t j = i - 3;
                if (j < 0) j += sz;
                for (int k = 0; k < 7; k++) {
                    acc = __dadd_rn(acc, __dmul_rn(errs[j], P->qsm_kern[k]));
                    j++;
                    if (j == sz) j = 0;
                }
            }
            y[i] = acc;
        }
        __syncthreads();
        // cyclic local maxima, in order (fast/quads.c:805-823): by segments
        // of kSegPts points, the flags as ballot words in LDS (coalesced
        // reads of the filtered errors), then each thread lists the maxima
        // of a contiguous run of words after one block scan of the counts
        uint32_t *midx = inl ? lidx : F.midx + s0;
        uint32_t base = 0;
        for (int64_t g0 = 0; g0 < sz; g0 += kSegPts) {
            const uint32_t gn = (uint32_t)min((int64_t)kSegPts, sz - g0), nw = (gn + 63) / 64;
            for (uint32_t c0 = 0; c0 < gn; c0 += NT) {
                const int64_t i = g0 + c0 + t;
                bool f = false;
                if (i < sz) {
                    const double e = y[i], yl = y[i == 0 ? sz - 1 : i - 1], yr = y[i == sz - 1 ? 0 : i + 1];
                    f = e > yl && e > yr;
                }
                const uint64_t m = __ballot(f);
                if (lane == 0) fw[(c0 >> 6) + (uint32_t)w] = m;
            }
            __syncthreads();
            const uint32_t per = (nw + NT - 1) / NT, w0 = min(nw, (uint32_t)t * per), w1 = min(nw, w0 + per);
            uint32_t cnt = 0;
            for (uint32_t k = w0; k < w1; k++) cnt += (uint32_t)__popcll(fw[k]);
            uint32_t tot;
            uint32_t o = base + bscan_t<NT>(cnt, &tot);
            for (uint32_t k = w0; k < w1; k++)
                for (uint64_t m = fw[k]; m; m &= m - 1) midx[o++] = (uint32_t)(g0 + 64 * k + (uint32_t)__ffsll((long long)m) - 1);
            base += tot;
            __syncthreads();
        }
        uint32_t nm = base;
        if (nm < 4) continue;
        if ((int64_t)nm > P->max_nmaxima) {
            // the (max_nmaxima+1)-th largest maximum error (fast/quads.c:
            // 826-862 finds the same value by quickselect); the maxima
            // above it are kept, in order
            const uint32_t rank = (uint32_t)P->max_nmaxima;  // 0-based, descending
            double thr;
            if (NT == 64 && nm <= (uint32_t)kMaxima) {
                // a register sort of the maxima's order-preserving keys
                // (absent entries 0, first): the value at descending rank
                // `rank` (which of equal values is taken does not matter)
                uint64_t kv[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const uint32_t i = (uint32_t)(r * 64 + lane);
                    kv[r] = i < nm ? dkey(y[midx[i]]) : 0ull;
                }
                const int R = nm <= 64 ? 1 : nm <= 128 ? 2 : 4;
                wave_sort_rows(kv, R);
                thr = dkey_value(wave_key_at(kv, (uint32_t)(64 * R) - 1u - rank));
            } else if (NT == 256 && nm <= (uint32_t)kMaxima && rank < 64) {
                // the maxima's keys sorted in registers a wave at a time,
                // 256 a wave (absent entries 0, first); past one wave's
                // worth, the waves' top 64 keys (which hold the overall top
                // 64) sorted again by wave 0. The value goes to every wave
                // through LDS.
                const uint32_t nwv = (nm + 255) / 256;
                uint64_t *top = (uint64_t *)(sh + 2 * kLdsPts);
                if ((uint32_t)w < nwv) {
                    const uint32_t cnt = min(256u, nm - (uint32_t)w * 256u);
                    const int R = cnt <= 64 ? 1 : cnt <= 128 ? 2 : 4;
                    uint64_t kv[4];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const uint32_t i = (uint32_t)(w * 256 + r * 64 + lane);
                        kv[r] = (r < R && i < nm) ? dkey(y[midx[i]]) : 0ull;
                    }
                    wave_sort_rows(kv, R);
                    if (nwv == 1) {
                        const double v = dkey_value(wave_key_at(kv, (uint32_t)(64 * R) - 1u - rank));
                        if (lane == 0) s_wbest[0] = v;
                    } else {
                        top[w * 64 + lane] = R == 1 ? kv[0] : R == 2 ? kv[1] : kv[3];
                    }
                }
                if (nwv > 1) {
                    __syncthreads();
                    if (w == 0) {
                        uint64_t kv[4];
#pragma unroll
                        for (int r = 0; r < 4; r++) kv[r] = (uint32_t)r < nwv ? top[r * 64 + lane] : 0ull;
                        const int R = nwv <= 2 ? 2 : 4;
                        wave_sort_rows(kv, R);
                        const double v = dkey_value(wave_key_at(kv, (uint32_t)(64 * R) - 1u - rank));
                        if (lane == 0) s_wbest[0] = v;
                    }
                }
                __syncthreads();
                thr = s_wbest[0];
            } else if constexpr (NT == 256) {
                // radix select over the order-preserving bits
                uint32_t rk = rank;
                uint64_t prefix = 0;
                for (int shift = 56; shift >= 0; shift -= 8) {
                    hist[t] = 0;
                    __syncthreads();
                    uint64_t hmask = shift == 56 ? 0ull : (~0ull << (shift + 8));
                    for (uint32_t i = t; i < nm; i += 256) {
                        uint64_t k = dkey(y[midx[i]]);
                        if ((k & hmask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
                    }
                    __syncthreads();
                    {
                        // bins in descending order, one per thread: the bin
                        // that holds the rk-th largest
                        uint32_t d = 255u - (uint32_t)t, cnt = hist[d], tot;
                        uint32_t before = bscan(cnt, &tot);
                        if (rk >= before && rk < before + cnt) {
                            s_prefix = prefix | ((uint64_t)d << shift);
                            s_rank = rk - before;
                        }
                    }
                    __syncthreads();
                    prefix = s_prefix;
                    rk = s_rank;
                    __syncthreads();
                }
                thr = dkey_value(prefix);
            } else {
                thr = 0.0;  // the wave shape holds at most kMaxima maxima
            }
            // keep the maxima above the threshold (a comparison of doubles,
            // as the reference's), in order
            uint32_t o = 0;
            for (uint32_t c0 = 0; c0 < nm; c0 += NT) {
                uint32_t i = c0 + t;
                uint32_t f = 0, v = 0;
                if (i < nm) {
                    v = midx[i];
                    f = y[v] > thr ? 1u : 0u;
                }
                uint32_t tot;
                uint32_t pre = bscan_t<NT>(f, &tot);
                if (f && o + pre < (uint32_t)kMaxNm) mx[o + pre] = v;
                o += tot;
            }
            nm = o;
        } else {
            for (uint32_t i = t; i < nm && i < (uint32_t)kMaxNm; i += NT) mx[i] = midx[i];
        }
        __syncthreads();
        if (nm > (uint32_t)kMaxNm) {
            if (t == 0) atomicOr(F.counters + kError, 4u);
            continue;
        }
        // the selection keeps fewer than four maxima when ties fill the top
        // max_nmaxima + 1 values (or max_nmaxima < 4): no 4-tuple exists, the
        // reference's search finds nothing and rejects the cluster
        // (fast/quads.c:864-956); nm is the same in every thread
        if (nm < 4) continue;
        // pair fits (fast/quads.c:869-928): (a, b) for a < b, closing (b, a) for b >= a + 3
        for (uint32_t pq = t; pq < nm * nm; pq += NT) {
            uint32_t p = pq / nm, q = pq % nm;
            bool fwd = p < q, back = q + 3 <= p;
            if (!fwd && !back) continue;
            int64_t i0 = mx[p], i1 = mx[q];
            double s[6], lp[4], mse;
            range_sums(M, sz, i0, i1, s);
            line_from_sums(s, lp, &mse);
            double cnt = fwd ? (double)(i1 - i0 + 1) : (double)(sz - i0 + i1 + 1);
            pe[pq] = __dmul_rn(cnt, mse);
            pm[pq] = mse;
            pnx[pq] = lp[2];
            pny[pq] = lp[3];
        }
        __syncthreads();
        // the first minimum over (m0, m1, m2, m3) in loop order (:930-956)
        double best = INF;
        uint32_t bord = 0xffffffffu;
        // the triples with m2 + 1 < nm (an m3 after m2), in any order: the
        // loop order is in the key
        const uint32_t ntri = (nm - 1) * (nm - 2) * (nm - 3) / 6;
        for (uint32_t k3 = t; k3 < ntri; k3 += NT) {
            const uint32_t pk = kTri.v[k3];
            const uint32_t m0 = pk & 31u, m1 = (pk >> 5) & 31u, m2 = pk >> 10;
            const uint32_t tri = (m0 * nm + m1) * nm + m2;
            uint32_t p01 = m0 * nm + m1, p12 = m1 * nm + m2;
            if (pm[p01] > max_mse || pm[p12] > max_mse) continue;
            double dot = __dadd_rn(__dmul_rn(pnx[p01], pnx[p12]), __dmul_rn(pny[p01], pny[p12]));
            if (fabs(dot) > max_dot) continue;
            double e012 = __dadd_rn(pe[p01], pe[p12]);
            for (uint32_t m3 = m2 + 1; m3 < nm; m3++) {
                uint32_t p23 = m2 * nm + m3, p30 = m3 * nm + m0;
                double f23 = pm[p23] > max_mse ? INF : pe[p23];
                double f30 = pm[p30] > max_mse ? INF : pe[p30];
                double err = __dadd_rn(__dadd_rn(e012, f23), f30);
                uint32_t ord = tri * nm + m3;
                if (err < best) {
                    best = err;
                    bord = ord;
                }
            }
        }
        // the least (error, loop order) over the block (equal infinite
        // errors reject the cluster whatever their order)
        for (int d = 32; d >= 1; d >>= 1) {
            double e2 = __shfl_xor(best, d, 64);
            uint32_t o2 = (uint32_t)__shfl_xor((int)bord, d, 64);
            if (e2 < best || (e2 == best && o2 < bord)) {
                best = e2;
                bord = o2;
            }
        }
        if (lane == 0) {
            s_wbest[w] = best;
            s_word[w] = bord;
        }
        __syncthreads();
        if (t == 0) {
            double b = INF;
            uint32_t o = 0xffffffffu;
            for (int k = 0; k < kWaves; k++) {
                double e = s_wbest[k];
                if (e < b || (e == b && s_word[k] < o)) {
                    b = e;
                    o = s_word[k];
                }
            }
            const bool ok0 = b < INF && __ddiv_rn(b, (double)sz) < max_mse;
            s_ok = ok0 ? 1 : 0;
            if (ok0) {
                uint32_t m3 = o % nm, tri = o / nm;
                uint32_t m0 = tri / (nm * nm), m1 = (tri / nm) % nm, m2 = tri % nm;
                s_idx[0] = mx[m0];
                s_idx[1] = mx[m1];
                s_idx[2] = mx[m2];
                s_idx[3] = mx[m3];
            }
        }
        __syncthreads();
        if (!s_ok) continue;
        // the four lines (fast/quads.c:1058-1075), one thread each
        if (t < 4) {
            double s[6], lp[4], mse;
            range_sums(M, sz, s_idx[t], s_idx[(t + 1) & 3], s);
            line_from_sums(s, lp, &mse);
            for (int k = 0; k < 4; k++) s_lines[t][k] = lp[k];
            s_lok[t] = mse > max_mse ? 0 : 1;
        }
        __syncthreads();
        // the corners of lines i and i + 1 (fast/quads.c:1077-1092), one
        // thread each; then the six side lengths of the area's two Heron
        // triangles (:1094-1110) and the four convexity checks (:1112-1124)
        // on threads of their own, and their reduction (in the reference's
        // order) on thread 0. A corner, side or angle of a quad rejected
        // earlier in the reference's order is computed and not used.
        auto sq = [](double v) { return __dmul_rn(v, v); };
        auto fd = [](float a, float b) { return (double)__fsub_rn(a, b); };
        if (t < 4) {
            const int i = t, j = (i + 1) & 3;
            double a00 = s_lines[i][3], a01 = -s_lines[j][3], a10 = -s_lines[i][2], a11 = s_lines[j][2];
            double b0 = __dadd_rn(-s_lines[i][0], s_lines[j][0]), b1 = __dadd_rn(-s_lines[i][1], s_lines[j][1]);
            double det = __dsub_rn(__dmul_rn(a00, a11), __dmul_rn(a10, a01));
            const bool dok = !(fabs(det) < 0.001);
            s_dok[i] = dok ? 1 : 0;
            if (dok) {
                double w00 = __ddiv_rn(a11, det), w01 = __ddiv_rn(-a01, det);
                double l0 = __dadd_rn(__dmul_rn(w00, b0), __dmul_rn(w01, b1));
                s_qp[i][0] = (float)__dadd_rn(s_lines[i][0], __dmul_rn(l0, a00));
                s_qp[i][1] = (float)__dadd_rn(s_lines[i][1], __dmul_rn(l0, a10));
            }
        }
        __syncthreads();
        if (!(s_lok[0] && s_lok[1] && s_lok[2] && s_lok[3] && s_dok[0] && s_dok[1] && s_dok[2] && s_dok[3])) continue;
        if (t < 6) {
            // triangle (0, 1, 2) then (2, 3, 0): sides (0,1) (1,2) (2,0) (2,3) (3,0) (0,2)
            const int ia = t < 3 ? t : (t == 3 ? 2 : (t == 4 ? 3 : 0));
            const int ib = t < 3 ? (t + 1) % 3 : (t == 3 ? 3 : (t == 4 ? 0 : 2));
            s_side[t] = __dsqrt_rn(__dadd_rn(sq(fd(s_qp[ib][0], s_qp[ia][0])), sq(fd(s_qp[ib][1], s_qp[ia][1]))));
        } else if (t >= 8 && t < 12) {
            const int i0 = t - 8, i1 = (i0 + 1) & 3, i2 = (i0 + 2) & 3;
            const double ccr = (double)P->cos_crit;
            double dx1 = fd(s_qp[i1][0], s_qp[i0][0]), dy1 = fd(s_qp[i1][1], s_qp[i0][1]);
            double dx2 = fd(s_qp[i2][0], s_qp[i1][0]), dy2 = fd(s_qp[i2][1], s_qp[i1][1]);
            double den = __dsqrt_rn(__dmul_rn(__dadd_rn(__dmul_rn(dx1, dx1), __dmul_rn(dy1, dy1)),
                                              __dadd_rn(__dmul_rn(dx2, dx2), __dmul_rn(dy2, dy2))));
            bool cok = den != 0.0;
            if (cok) {
                double cd = __ddiv_rn(__dadd_rn(__dmul_rn(dx1, dx2), __dmul_rn(dy1, dy2)), den);
                if ((cd > ccr || cd < -ccr) || __dmul_rn(dx1, dy2) < __dmul_rn(dy1, dx2)) cok = false;
            }
            s_cvx[i0] = cok ? 1 : 0;
        }
        __syncthreads();
        if (t == 0) {
            bool ok = true;
            QuadRec q;
            for (int i = 0; i < 4; i++) {
                q.p[i][0] = s_qp[i][0];
                q.p[i][1] = s_qp[i][1];
            }
            double area = 0.0;
            for (int k = 0; k < 2; k++) {
                const double l0 = s_side[3 * k], l1 = s_side[3 * k + 1], l2 = s_side[3 * k + 2];
                double pp = __ddiv_rn(__dadd_rn(__dadd_rn(l0, l1), l2), 2.0);
                area = __dadd_rn(area, __dsqrt_rn(__dmul_rn(__dmul_rn(__dmul_rn(pp, __dsub_rn(pp, l0)),
                                                                      __dsub_rn(pp, l1)),
                                                            __dsub_rn(pp, l2))));
            }
            double tw = (double)P->tag_width;
            if (area < __dmul_rn(__dmul_rn(0.95, tw), tw)) ok = false;
            ok = ok && s_cvx[0] && s_cvx[1] && s_cvx[2] && s_cvx[3];
            if (ok) {
                q.reversed = (flag & 2u) ? 1 : 0;
                q.hok = 0;
                uint32_t slot = atomicAdd(F.counters + kNQuads, 1u);
                if (slot < F.qcap) {
                    F.qtmp[slot] = q;
                    F.qcl[slot] = c;
                } else {
                    atomicOr(F.counters + kError, 8u);
                }
            }
        }
        __syncthreads();
    }
    }
}

// The fitted quads in cluster order (one block per frame): rank of each
// by its cluster index.
__global__ __launch_bounds__(256) void k_quad_order(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    const uint32_t nq = min(F.counters[kNQuads], F.qcap), ncl = F.counters[kNClusters];
    // a quad's rank = the clusters before its own that have a quad (one a
    // cluster at most): marks over the clusters (midx, free after the
    // corner search), scanned
    uint32_t *mark = F.midx;
    for (uint32_t c = threadIdx.x; c < ncl; c += 256) mark[c] = 0u;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nq; i += 256) mark[F.qcl[i]] = 1u;
    __syncthreads();
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < ncl; c0 += 256) {
        const uint32_t c = c0 + threadIdx.x;
        const uint32_t v = c < ncl ? mark[c] : 0u;
        uint32_t tot;
        const uint32_t pre = bscan(v, &tot);
        if (c < ncl) mark[c] = carry + pre;
        carry += tot;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nq; i += 256) F.quads[mark[F.qcl[i]]] = F.qtmp[i];
    if (threadIdx.x == 0) F.counters[kNQuads] = nq;
}

void launch_radix_sort(hipStream_t st, const Frame *dframes, int nframes, int which, int nbits, int *parity,
                       int dyn, int iota, int pk, int gather, int ranks);

void warm_decode_grid(int fit_blocks);

// Resident blocks per kernel shape, looked up in a fixed table (a lookup
// never allocates, so calls that must not allocate can launch through it).
unsigned persistent_grid(const void *fn, int threads, size_t lds, int want) {
    struct Entry {
        const void *fn;
        int threads;
        size_t lds;
        int resident;
    };
    constexpr int kSlots = 256;
    static std::mutex mu;
    static Entry table[kSlots];
    int resident = 0;
    {
        std::lock_guard<std::mutex> g(mu);
        size_t h = (reinterpret_cast<size_t>(fn) >> 4) * 31 + (size_t)threads * 7 + lds;
        for (int k = 0; k < kSlots; k++) {
            Entry &e = table[(h + (size_t)k) % kSlots];
            if (e.fn == fn && e.threads == threads && e.lds == lds) {
                resident = e.resident;
                break;
            }
            if (e.fn == nullptr) {
                int dev = 0, cus = 0, per_cu = 0;
                if (hipGetDevice(&dev) != hipSuccess ||
                    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds) != hipSuccess)
                    cus = per_cu = 0;
                (void)hipGetLastError();
                e = Entry{fn, threads, lds, std::max(1, per_cu * cus)};
                resident = e.resident;
                break;
            }
        }
        if (resident == 0) resident = 1;  // a full table (never: a few dozen shapes exist)
    }
    return (unsigned)std::max(1, std::min(want, resident));
}

template <typename K>
static unsigned pgrid(K *fn, int threads, int want) {
    return persistent_grid(reinterpret_cast<const void *>(fn), threads, 0, want);
}

void warm_persistent_grids(int fit_blocks) {
    pgrid(k_angle_sort<true>, 256, fit_blocks);
    pgrid(k_angle_sort<false>, 256, fit_blocks);
    {
        static bool big_lds = false;  // the big sort's LDS: its keys across waves
        if (!big_lds)
            big_lds = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_angle_sort_big),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kBigSort * 8)) == hipSuccess;
        (void)hipGetLastError();
        persistent_grid(reinterpret_cast<const void *>(&k_angle_sort_big), 1024, kBigSort * 8, fit_blocks);
    }
    pgrid(k_angle_sort_wave, 64, fit_blocks);
    pgrid(k_fit_moments<128, 128>, 128, 2 * fit_blocks);
    pgrid(k_fit_moments<256, 512>, 256, fit_blocks);
    pgrid(k_fit_moments<64, 128>, 64, fit_blocks);
    pgrid(k_fit_moments_packed<10, 32>, 64, fit_blocks);
    pgrid(k_fit_corners<256>, 256, fit_blocks);
    pgrid(k_fit_corners<64>, 64, fit_blocks);
    warm_decode_grid(fit_blocks);
}

// Whether the angle sort runs per cluster (any size: its steps wider than
// a 2048-key tile go over the cluster's keys in HBM), or (forced, for
// tests) over all points at once.
// 0: the angle sort of all points at once in HBM (forced); otherwise per
// cluster, 2 when a cluster may exceed kBigSort points (the frame bounds
// a fitted cluster by 2 (2 qw + 2 qh), fast/quads.c:1148-1154), else 1.
int angle_sort_cap(int max_cluster, bool force_global) {
    if (force_global) return 0;
    return max_cluster > (int)kBigSort ? 2 : 1;
}

void launch_quads_stage(hipStream_t st, const Frame *dframes, const DevParams *dparams, int nframes, int cl_bits,
                        int fit_blocks, int sort_cap, bool wave_corners, WorkQueue *Q, Marks *mk, Fork *fk) {
    // k_fit_prep: every cluster of every frame, a grid per frame; the fit
    // kernels after it: persistent over the work lists, fit_blocks blocks of
    // 256 threads or four times as many waves, at most as many as are
    // resident at once
    const int wf = 4 * fit_blocks;
    // k_fit_prep: a wave a cluster for batches; for a few frames, whose
    // clusters all fit at once, 256 threads a cluster (the large ones sooner)
    if (nframes > 4)
        hipLaunchKernelGGL(k_fit_prep<64>, dim3(512, nframes), dim3(64), 0, st, dframes, dparams, sort_cap ? 0 : 1);
    else
        hipLaunchKernelGGL(k_fit_prep<256>, dim3(nframes <= 2 ? 1024 : 256, nframes), dim3(256), 0, st, dframes, dparams,
                           sort_cap ? 0 : 1);
    mk->mark(st, "fit_prep.prep");
    hipLaunchKernelGGL(k_keep_lists, dim3(1, nframes), dim3(256), 0, st, dframes);
    // the huge, large and small lists, a block each
    const uint64_t lw = (uint64_t)(4 | kNHuge << 8) | (uint64_t)(0 | kNLarge << 8) << 16 |
                        (uint64_t)(1 | kNSmall << 8) << 32;
    hipLaunchKernelGGL(k_work_offsets, dim3(3), dim3(256), 0, st, dframes, nframes, Q, lw);
    mk->mark(st, "fit_prep.lists");
    // The clusters' fits are independent from here to the quad order. With
    // the fork on, the small ones (list 1) go on one forked stream and, for
    // calls of a few frames, the large ones (list 0) on the other, beside
    // the huge ones (list 4) on the main stream: parallel branches of the
    // call's graph. Batches keep the huge and large ones together.
    int parity = 0;
    hipStream_t fs[Fork::kForks] = {st, st};
    const bool few = nframes <= 16;
    if (sort_cap) {
        fk->begin(st, fs);
        hipStream_t sl = few ? fs[1] : st;  // the large clusters' stream
        if (few) {
            // the huge clusters of up to kBigSort points each by a
            // 1024-thread block, larger ones (possible from 4096 quad-image
            // pixels of perimeter on) by the tiled sort
            hipLaunchKernelGGL(k_angle_sort_big,
                               dim3(persistent_grid(reinterpret_cast<const void *>(&k_angle_sort_big), 1024,
                                                    kBigSort * 8, std::min(fit_blocks, 64 * nframes))),
                               dim3(1024), kBigSort * 8, st, dframes, Q);
            if (sort_cap == 2)
                hipLaunchKernelGGL(k_angle_sort<true>, dim3(pgrid(k_angle_sort<true>, 256, fit_blocks)), dim3(256), 0,
                                   st, dframes, Q, 4);
            hipLaunchKernelGGL(k_angle_sort<true>, dim3(pgrid(k_angle_sort<true>, 256, fit_blocks)), dim3(256), 0, sl,
                               dframes, Q, 5);
        } else {
            hipLaunchKernelGGL(k_angle_sort<false>, dim3(pgrid(k_angle_sort<false>, 256, fit_blocks)), dim3(256), 0,
                               st, dframes, Q, 0);
        }
        mk->mark(st, "angle_sort.large");
        hipLaunchKernelGGL(k_angle_sort_wave, dim3(pgrid(k_angle_sort_wave, 64, wf)), dim3(64), 0, fs[0], dframes, Q);
        mk->mark(fs[0], "angle_sort.small");
    } else {
        launch_radix_sort(st, dframes, nframes, kNPts, 32 + cl_bits, &parity, 0, 0, 0, 0, 0);
        mk->mark(st, "angle_sort.all");
        fk->begin(st, fs);
    }
    // the huge and large clusters: 128-thread blocks of 128-point chunks for
    // batches (more blocks a CU: more running sums in flight; 1.48 -> 1.36
    // ms at 256 frames against 256-point chunks), 256 threads for up to 16
    // frames (each cluster sooner), one launch a list then
    if (!few) {
        hipLaunchKernelGGL((k_fit_moments<128, 128>), dim3(pgrid(k_fit_moments<128, 128>, 128, 2 * fit_blocks)),
                           dim3(128), 0, st, dframes, parity, Q, sort_cap ? 1 : 0, 0, 1);
    } else {
        const int fix = sort_cap ? 0 : 1;
        hipLaunchKernelGGL((k_fit_moments<256, 512>), dim3(pgrid(k_fit_moments<256, 512>, 256, fit_blocks)), dim3(256),
                           0, st, dframes, parity, Q, sort_cap ? 1 : 0, 4, fix);
        hipLaunchKernelGGL((k_fit_moments<256, 512>), dim3(pgrid(k_fit_moments<256, 512>, 256, fit_blocks)), dim3(256),
                           0, fs[1], dframes, parity, Q, sort_cap ? 1 : 0, 5, fix);
    }
    mk->mark(st, "moments.large");
    if (!few && sort_cap)  // batches: ten small clusters' sums a wave
        hipLaunchKernelGGL((k_fit_moments_packed<10, 32>), dim3(pgrid(k_fit_moments_packed<10, 32>, 64, wf)),
                           dim3(64), 0, fs[0], dframes, Q, 1);
    else
        hipLaunchKernelGGL((k_fit_moments<64, 128>), dim3(pgrid(k_fit_moments<64, 128>, 64, wf)), dim3(64), 0, fs[0],
                           dframes, parity, Q, sort_cap ? 1 : 0, 1, sort_cap ? 0 : 1);
    mk->mark(fs[0], "moments.small");
    if (!few) {
        hipLaunchKernelGGL(k_fit_corners<256>, dim3(pgrid(k_fit_corners<256>, 256, fit_blocks)), dim3(256), 0, st,
                           dframes, dparams, Q, 0);
    } else {
        hipLaunchKernelGGL(k_fit_corners<256>, dim3(pgrid(k_fit_corners<256>, 256, fit_blocks)), dim3(256), 0, st,
                           dframes, dparams, Q, 4);
        hipLaunchKernelGGL(k_fit_corners<256>, dim3(pgrid(k_fit_corners<256>, 256, fit_blocks)), dim3(256), 0, fs[1],
                           dframes, dparams, Q, 5);
    }
    mk->mark(st, "corners.large");
    if (wave_corners)
        hipLaunchKernelGGL(k_fit_corners<64>, dim3(pgrid(k_fit_corners<64>, 64, wf)), dim3(64), 0, fs[0], dframes,
                           dparams, Q, 1);
    else
        hipLaunchKernelGGL(k_fit_corners<256>, dim3(pgrid(k_fit_corners<256>, 256, fit_blocks)), dim3(256), 0, fs[0],
                           dframes, dparams, Q, 1);
    mk->mark(fs[0], "corners.small");
    fk->end(st);
    hipLaunchKernelGGL(k_quad_order, dim3(1, nframes), dim3(256), 0, st, dframes);
    launch_work_offsets(st, dframes, nframes, Q, 2, (int)kNQuads);
    mk->mark(st, "corners.order");
}

}  // namespace atg
