// Stage (b): connected components over row runs, with the reference's
// exact union-find representatives, and the per-pixel labels of (c).
//
// The reference links sets in a fixed row-major edge order with union by
// size (fast/unionfind.c:15-45, 112-146, 223-237): the root of a merged
// set is the root of the heavier part, ties to the previous-row side.
// Those roots become cluster ids, and cluster ids decide the cluster
// order, so a min-label GPU CCL (right partition, other roots) is not
// enough. Two facts make the replay parallel: the outcome of a link
// depends only on the two set weights at that moment, and weights are
// pixel counts, which do not depend on the order. So:
//
//   k_uf_edges   one wave per row, all rows at once: the row's links in the
//                reference's order (current run ascending, then previous
//                run ascending).
//   k_front      one block per frame, an unordered lock-free union-find row
//                by row: which runs of row y share a set over rows 0..y
//                (the row's records) and each record's weight.
//   k_replay     one wave per row, all rows at once: the row's links replayed
//                in order over the previous row's records (weights known,
//                roots named symbolically) and the row's runs.
//   k_root_jump  pointer jumping over the records, all rows at once: the
//   k_lost       symbolic roots become nodes; each record root that lost
//                gets its parent, each set that ends its final weight.
//   k_resolve    one thread per node follows parent[] to the final root;
//                nlab = root | (pixel count >= min_cluster_pixels) << 31
//                (fast/clusters.c:73-94).
//   k_labels     one wave per row: the nlab of the run under each pixel,
//                the run index counted with ballots over the row's changes,
//                and the dense number of its root (numbered just before).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "atg_internal.h"

namespace atg {

__device__ __forceinline__ uint32_t run_start(uint32_t r) { return r & 0x7fffu; }
__device__ __forceinline__ uint32_t run_end(uint32_t r) { return (r >> 15) & 0x7fffu; }
__device__ __forceinline__ uint32_t run_white(uint32_t r) { return (r >> 30) & 1u; }

// inclusive wave scan of v (64 lanes)
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) { return wave_scan_incl(v); }

// Whether current run a and previous run b of the same value are linked
// (fast/unionfind.c:130-141): overlap on columns >= 1; white also touches
// diagonally.
__device__ __forceinline__ bool runs_join(int a0, int a1, int b0, int b1, bool white) {
    if (max(max(a0, b0), 1) <= min(a1, b1)) return true;
    if (!white) return false;
    if (max(max(a0, b0 + 1), 1) <= min(a1, b1 + 1)) return true;
    return max(max(a0, b0 - 1), 1) <= min(a1, b1 - 1);
}

// ecnt[y]: row y's link count, and in this bit whether its last run is
// also linked to the row's virtual node (fast/unionfind.c:93-97: a white
// run ending at column qw - 2 under a white pixel at qw - 1 whose left
// neighbour is black, in row y - 1)
constexpr uint32_t kLastColLink = 0x80000000u;
__device__ __forceinline__ int edge_count(uint32_t ecnt) { return (int)(ecnt & ~kLastColLink); }

// The links of every row, in the reference's order (current run
// ascending, then previous run ascending): one wave per row, all rows at
// once. uedge[y][e] = i << 16 | k (current run i of row y, previous run k).
// L: the previous row's runs are first copied to LDS (lds_qw words a wave),
// where the search and scans read them.
template <bool L>
__global__ __launch_bounds__(256) void k_uf_edges(const Frame *frames, int lds_qw) {
    extern __shared__ uint32_t sprv[];
    const Frame F = frames[blockIdx.y];
    const int lane = threadIdx.x & 63;
    int y = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (y < 1 || y >= F.qh) return;
    const int QW = F.qw;
    const int ncur = (int)F.rcnt[y], nprev = (int)F.rcnt[y - 1];
    const uint32_t *cur = F.runs + (int64_t)y * QW;
    const uint32_t *prv = F.runs + (int64_t)(y - 1) * QW;
    // the virtual-node link of the row's last run (its loads in flight
    // through the row's links)
    uint32_t lastrun = 0u;
    uint8_t tq1 = 0, tq2 = 0;
    if (lane == 0 && ncur > 0) {
        const uint8_t *trow = F.t + (int64_t)(y - 1) * QW;
        lastrun = cur[ncur - 1];
        tq1 = trow[QW - 1];
        tq2 = trow[QW - 2];
    }
    // the first 64 current runs, read while the previous row is copied
    uint32_t rnext = lane < ncur ? cur[lane] : 0u;
    if constexpr (L) {
        uint32_t *sp = sprv + (threadIdx.x >> 6) * lds_qw;
        // 8 loads a lane in flight at once
        for (int k0 = 0; k0 < nprev; k0 += 512) {
            uint32_t v[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int k = k0 + q * 64 + lane;
                v[q] = k < nprev ? prv[k] : 0u;
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int k = k0 + q * 64 + lane;
                if (k < nprev) sp[k] = v[q];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        prv = sp;
    }
    uint32_t *out = F.uedge + (int64_t)y * 3 * QW;
    const uint32_t cap = 3u * (uint32_t)QW;
    uint32_t carry = 0;
    for (int i0 = 0; i0 < ncur; i0 += 64) {
        int i = i0 + lane;
        int a0 = 0, a1 = -1, klo = nprev;
        bool white = false;
        uint32_t cnt = 0;
        const uint32_t r = rnext;  // the next 64 read a step ahead
        if (i + 64 < ncur) rnext = cur[i + 64];
        if (i < ncur) {
            a0 = (int)run_start(r);
            a1 = (int)run_end(r);
            white = run_white(r) != 0;
            int lo = 0, hi = nprev;
            while (lo < hi) {  // the first previous run with end + 1 >= a0
                int mid = (lo + hi) >> 1;
                if ((int)run_end(prv[mid]) + 1 < a0) lo = mid + 1; else hi = mid;
            }
            klo = lo;
            for (int k = klo; k < nprev; k++) {
                uint32_t pk = prv[k];
                if ((int)run_start(pk) > a1 + 1) break;
                if ((run_white(pk) != 0) == white && runs_join(a0, a1, (int)run_start(pk), (int)run_end(pk), white))
                    cnt++;
            }
        }
        uint32_t inc = wave_incl_scan(cnt);
        uint32_t at = carry + inc - cnt;
        if (cnt && at + cnt <= cap) {
            for (int k = klo; k < nprev; k++) {
                uint32_t pk = prv[k];
                if ((int)run_start(pk) > a1 + 1) break;
                if ((run_white(pk) != 0) == white && runs_join(a0, a1, (int)run_start(pk), (int)run_end(pk), white))
                    out[at++] = ((uint32_t)i << 16) | (uint32_t)k;
            }
        }
        carry += (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
    }
    if (lane == 0) {
        if (carry > cap) atomicOr(F.counters + kError, 1u);
        const bool lastcol = ncur > 0 && (int)run_end(lastrun) == QW - 2 && run_white(lastrun) && tq1 == 255 && tq2 != 255;
        F.ecnt[y] = min(carry, cap) | (lastcol ? kLastColLink : 0u);
        atomicAdd(F.counters + kNEdges, min(carry, cap));
    }
}

// Memory of the three replay kernels: LDS, or (for very wide frames) HBM
// scratch through volatile pointers, so that each lane sees the others'
// writes after a barrier.
template <bool G>
struct Mem {
    using U32 = typename std::conditional<G, volatile uint32_t, uint32_t>::type;
    using U16 = typename std::conditional<G, volatile uint16_t, uint16_t>::type;
};

// A barrier of a one-wave block. LDS operations of a wave complete in
// order once lgkmcnt is 0, so the LDS variant only has to keep the
// compiler (and the LDS queue) in order; the scratch variant orders HBM.
template <bool G>
__device__ __forceinline__ void wsync() {
    if constexpr (G) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    }
}

constexpr uint32_t kSym = 0x80000000u;     // a record of the previous row, by slot
constexpr uint32_t kLostNone = 0xffffffffu; // the record's root goes on into the row
constexpr uint32_t kLostRetired = 0xfffffffeu;  // the record's set ended at the previous row

// ---------------------------------------------------------------------
// Step 1, k_front: the sets that touch each row, with their weights.
//
// Connectivity and pixel counts do not depend on the link order, so one
// block per frame sweeps the rows with an unordered lock-free union-find:
// the elements of row y are the records of row y-1 (distinct sets), the
// current runs and the row's virtual node; after the unions, the roots
// that hold a current run become row y's records, numbered in element
// order. Output per row: rlab[run] = its record, recw[slot] = the record's
// weight (slot = row_off[y] + record), nrec[y] = the record count.

template <bool G>
__device__ __forceinline__ uint32_t fr_find(typename Mem<G>::U32 *p, uint32_t x) {
    for (;;) {  // path halving: only non-roots are rewritten, to an ancestor
        uint32_t q = p[x];
        if (q == x) return x;
        uint32_t g = p[q];
        if (g == q) return q;
        p[x] = g;
        x = g;
    }
}

template <bool G>
__device__ __forceinline__ void fr_unite(typename Mem<G>::U32 *p, uint32_t a, uint32_t b) {
    for (;;) {  // hook the higher root under the lower; retry if it moved
        a = fr_find<G>(p, a);
        b = fr_find<G>(p, b);
        if (a == b) return;
        if (a < b) {
            uint32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS((uint32_t *)&p[a], a, b) == a) return;
    }
}

// The roots of a and b found together with path halving, each step's reads
// of the two paths issued at once: the two dependent chains of LDS reads
// overlap. qa, qb: the parents of a and b, already read.
template <bool G>
__device__ __forceinline__ void fr_find_pair(typename Mem<G>::U32 *p, uint32_t &a, uint32_t qa, uint32_t &b,
                                             uint32_t qb) {
    bool da = qa == a, db = qb == b;
    while (!(da && db)) {
        const uint32_t ga = da ? a : (uint32_t)p[qa];
        const uint32_t gb = db ? b : (uint32_t)p[qb];
        if (!da) {
            if (ga == qa) {
                a = qa;
                da = true;
            } else {
                p[a] = ga;
                a = ga;
            }
        }
        if (!db) {
            if (gb == qb) {
                b = qb;
                db = true;
            } else {
                p[b] = gb;
                b = gb;
            }
        }
        if (!da) qa = p[a];
        if (!db) qb = p[b];
    }
}

// The roots of N elements a[i] (q[i]: their parents, already read), found
// together by reads only (one read a path a step, the N issued at once): a
// settling element's own write of its root is then the only write to it.
template <bool G, int N>
__device__ __forceinline__ void fr_roots(typename Mem<G>::U32 *p, uint32_t (&a)[N], uint32_t (&q)[N]) {
    for (;;) {
        bool more = false;
        uint32_t g[N];
#pragma unroll
        for (int i = 0; i < N; i++) {
            const bool root = q[i] == a[i];
            g[i] = root ? q[i] : (uint32_t)p[q[i]];
            more |= !root;
        }
        if (!more) return;
#pragma unroll
        for (int i = 0; i < N; i++) {
            if (q[i] != a[i]) {
                a[i] = q[i];
                q[i] = g[i];
            }
        }
    }
}

// fr_unite with the two finds together (fr_find_pair)
template <bool G>
__device__ __forceinline__ void fr_unite_pair(typename Mem<G>::U32 *p, uint32_t a, uint32_t b) {
    for (;;) {
        fr_find_pair<G>(p, a, p[a], b, p[b]);
        if (a == b) return;
        if (a < b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS((uint32_t *)&p[a], a, b) == a) return;
    }
}

// The first kStage runs and links of a row come from registers loaded one
// row ahead; the rest (rare) from HBM at use.
constexpr int kStage = 1024;

size_t front_bytes(int qw) { return (size_t)48 * qw + 128; }  // two sets of three [2 qw + 2] arrays
size_t front_band_bytes(int qw) { return (size_t)68 * qw + 128; }  // and the ports' least ports and records

// A barrier of the whole block over LDS only (the scratch variant orders
// HBM too).
template <bool G>
__device__ __forceinline__ void bsync() {
    if constexpr (G) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    }
}

// How a sweep starts and what it writes. The whole frame (kSweepFull)
// starts from row 0, whose runs are each their own set. A band of rows
// y0..y1 of a split frame starts from row y0 - 1: as ports (kSweepLocal),
// or from row y0 - 1's final records and weights (kSweepFinal).
//
// kSweepLocal: each run of row y0 - 1 (a port) is a set of weight 0; the
// band's sets that hold ports are its port components, each named by its
// least port. The sweep keeps each set's least port (as its least run)
// and follows each port's record row by row; a component ends either in
// row y1 (live) or earlier (its set has no run in the next row). It
// writes, at the port row's slots: rlost[slot + x] = the component of port
// x; rroot[slot + g] for each component g = kPortLive | its record in row
// y1, or its weight (its pixels in the band) if it ended earlier.
enum { kSweepFull = 0, kSweepLocal = 1, kSweepFinal = 2 };
constexpr uint32_t kPortEnded = 0x80000000u;  // (port tracking) the port's component ended: | its component
constexpr uint32_t kPortLive = 0x80000000u;   // (component record) reaches row y1: | its record there

// The sweep of rows y0..y1 (kSweepFull: 0..QH-1). Writes rlab/recw of
// rows wlo..y1 (kSweepLocal: row y1 only, its sets without what lies above
// the band); row_off and the virtual nodes of rows olo..ohi (the last
// band also row_off[QH] and the run count).
template <bool G, int kFrontThreads, int mode>
__device__ __forceinline__ void front_sweep(const Frame &F, uint8_t *base, uint32_t *lds_small, int y0, int y1, int olo, int ohi) {
    constexpr int kPer = kStage / kFrontThreads;  // prefetched runs and links per thread
    // the frame's arrays in HBM, as global pointers (flat accesses would
    // hold up the waits for the LDS union-find)
    const auto g_runs = gmem(F.runs);
    const auto g_uedge = gmem(F.uedge);
    const auto g_rcnt = gmem(F.rcnt);
    const auto g_ecnt = gmem(F.ecnt);
    const auto g_rlab = gmem(F.rlab);
    const auto g_recw = gmem(F.recw);
    const auto g_rroot = gmem(F.rroot);
    const auto g_rlost = gmem(F.rlost);
    const auto g_parent = gmem(F.parent);
    const auto g_fw = gmem(F.fw);
    const auto g_row_off = gmem(F.row_off);
    const auto g_counters = gmem(F.counters);
    using U32 = typename Mem<G>::U32;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int QW = F.qw, QH = F.qh;
    // elements of row y: the records of row y-1 at [0, QW) (each named by
    // its first run in row y-1), the runs of row y at QW + i, the virtual
    // node at 2 QW. Two sets of (parent, weight, least run) arrays take
    // the rows in turn: row y unites in one while the other holds row y-1
    // settled (every element under its root), where a run i's record is
    // pmin[pp[QW + i]] and, at the record's first run, its weight
    // pacc[pp[QW + i]].
    const int EM = 2 * QW + 2;
    const uint32_t VN = 2u * (uint32_t)QW;
    U32 *p = reinterpret_cast<U32 *>(base);
    U32 *acc = p + EM;
    U32 *minrun = acc + EM;
    U32 *pp = minrun + EM;
    U32 *pacc = pp + EM;
    U32 *pmin = pacc + EM;
    // kSweepLocal: the least port of each set (two sets, as minrun), and
    // each port's record in the last settled row (or kPortEnded | component)
    U32 *mp = pmin + EM;
    U32 *pmp = mp + EM;
    U32 *port = pmp + EM;
    const int yp = mode == kSweepFull ? 0 : y0 - 1;  // the row the sweep starts from
    const int wlo = mode == kSweepFull ? 0 : mode == kSweepFinal ? y0 : QH;  // rows written in the loop
    // row offsets: exclusive scan of the per-row run counts (wave 0); the
    // base of row yp for the sweep
    if (wave == 0) {
        uint32_t carry = 0;
        for (int y00 = 0; y00 < QH; y00 += 64) {
            int y = y00 + lane;
            uint32_t c = y < QH ? g_rcnt[y] : 0u;
            uint32_t inc = wave_incl_scan(c);
            if (y < QH && y >= olo && y <= ohi) g_row_off[y] = carry + inc - c;
            if (y == yp) lds_small[0] = carry + inc - c;
            carry += (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
        }
        if (lane == 0 && ohi == QH - 1) {
            g_row_off[QH] = carry;
            g_counters[kNRuns] = carry;
        }
        for (int y = olo + lane; y <= ohi; y += 64) {
            g_parent[carry + y] = carry + y;
            g_fw[carry + y] = 1;
        }
    }
    __syncthreads();
    uint32_t pbase = lds_small[0];  // row_off of the previous row
    // the starting row, as settled
    int pcur = (int)g_rcnt[yp];
    for (int i = tid; i < pcur; i += kFrontThreads) {
        if (mode == kSweepFull) {
            const uint32_t r = g_runs[i];
            pp[QW + i] = (uint32_t)(QW + i);
            pmin[QW + i] = (uint32_t)i;
            pacc[QW + i] = run_end(r) - run_start(r) + 1;
            g_parent[i] = (uint32_t)i;
        } else if (mode == kSweepLocal) {
            pp[QW + i] = (uint32_t)(QW + i);
            pmin[QW + i] = (uint32_t)i;
            pacc[QW + i] = 0u;
            pmp[QW + i] = (uint32_t)i;
            port[i] = (uint32_t)i;
        } else {
            const uint32_t l = g_rlab[pbase + i];
            pp[QW + i] = (uint32_t)QW + l;
            if (l == (uint32_t)i) {
                pmin[QW + i] = (uint32_t)i;
                pacc[QW + i] = g_recw[pbase + i];
            }
        }
    }
    const int nport = mode == kSweepLocal ? pcur : 0;
    const uint32_t portslot = pbase;  // the port row's slots
    // the next row, loaded one row ahead: thread tid holds runs and links
    // tid + j kFrontThreads. Two buffers take the rows in turn (the loop
    // body twice), so that no copy of a buffer waits for its loads.
    // The row's run and link counts are loaded lane by lane (lane 0 the
    // run count, lane 1 the link count) and read out of the lanes where the
    // row begins: a uniform load would be moved to scalar registers as it
    // returns, a wait on every load in flight.
    struct Ahead {
        uint32_t cnt, r[kPer], e[kPer];
    };
    auto prefetch = [&](int y, Ahead &nx) {
        const auto cur = g_runs + (int64_t)y * QW;
        const auto edges = g_uedge + (int64_t)y * 3 * QW;
        nx.cnt = (lane & 1) ? g_ecnt[y] : g_rcnt[y];
#pragma unroll
        for (int j = 0; j < kPer; j++) {
            int i = tid + j * kFrontThreads;
            nx.r[j] = i < QW ? cur[i] : 0u;
            nx.e[j] = i < 3 * QW ? edges[i] : 0u;
        }
    };
    auto sweep_row = [&](int y, const Ahead &cu, Ahead &nx) {
        const uint32_t rc = (uint32_t)__builtin_amdgcn_readlane((int)cu.cnt, 0);
        const uint32_t ec = (uint32_t)__builtin_amdgcn_readlane((int)cu.cnt, 1);
        const int ncur = (int)rc, ne = edge_count(ec);
        const bool vlink = (ec & kLastColLink) != 0;
        const uint32_t *rr = cu.r, *ee = cu.e;
        const uint32_t cbase = pbase + (uint32_t)pcur;
        if (y + 1 <= y1) prefetch(y + 1, nx);
        const auto cur = g_runs + (int64_t)y * QW;
        const auto edges = g_uedge + (int64_t)y * 3 * QW;
        // the previous row's outputs (each run's record, its first run,
        // and at a record's first run its weight), and the elements: the
        // records of row y-1 (weight 0 at the other slots, which no link
        // names), the runs of row y, the virtual node
        // (pp[QW + x] then holds run x's record, which its links read)
        const bool out = y - 1 >= wlo;
        for (int x = tid; x < pcur; x += kFrontThreads) {
            const uint32_t r = pp[QW + x];
            const uint32_t l = pmin[r], pw = pacc[r];
            const uint32_t wgt = l == (uint32_t)x ? pw : 0u;
            pp[QW + x] = l;
            if (out) {
                g_rlab[pbase + x] = l;
                if (l == (uint32_t)x) g_recw[pbase + x] = wgt;
            }
            p[x] = (uint32_t)x;
            acc[x] = wgt;
            minrun[x] = 0xffffffffu;
            if (mode == kSweepLocal) mp[x] = l == (uint32_t)x ? pmp[r] : 0xffffffffu;
        }
        // each port's record in row y-1 (from its record in row y-2), or
        // the end of its component in row y-2
        if (mode == kSweepLocal && y > yp + 1)
            for (int x = tid; x < nport; x += kFrontThreads) {
                const uint32_t q = port[x];
                if (q & kPortEnded) continue;
                const uint32_t r = pp[q], m = pmin[r];
                if (m != 0xffffffffu) {
                    port[x] = m;
                } else {
                    const uint32_t g = pmp[r];
                    port[x] = kPortEnded | g;
                    g_rroot[portslot + g] = pacc[r];  // (every port of the component writes the same)
                }
            }
        bool lastcol = false;  // fast/unionfind.c:93-97, on the thread of the last run
#pragma unroll
        for (int j = 0; j < kPer; j++) {
            const int i = tid + j * kFrontThreads;
            if (i < ncur) {
                const uint32_t r = rr[j];
                p[QW + i] = (uint32_t)(QW + i);
                acc[QW + i] = run_end(r) - run_start(r) + 1;
                minrun[QW + i] = 0xffffffffu;
                if (mode == kSweepLocal) mp[QW + i] = 0xffffffffu;
                if (i == ncur - 1) lastcol = vlink;
            }
        }
        for (int i = kStage + tid; i < ncur; i += kFrontThreads) {
            const uint32_t r = cur[i];
            p[QW + i] = (uint32_t)(QW + i);
            acc[QW + i] = run_end(r) - run_start(r) + 1;
            minrun[QW + i] = 0xffffffffu;
            if (mode == kSweepLocal) mp[QW + i] = 0xffffffffu;
            if (i == ncur - 1) lastcol = vlink;
        }
        // the virtual node on the last thread (the one least likely to hold
        // a record or a run)
        const bool vthread = tid == kFrontThreads - 1;
        if (vthread) {
            p[VN] = VN;
            acc[VN] = 1;
            minrun[VN] = 0xffffffffu;
            if (mode == kSweepLocal) mp[VN] = 0xffffffffu;
        }
        bsync<G>();
        // a link joins a current run and the record of a previous-row run
#pragma unroll
        for (int j = 0; j < kPer; j++) {
            const int e = tid + j * kFrontThreads;
            if (e < ne) fr_unite_pair<G>(p, (uint32_t)QW + (ee[j] >> 16), pp[QW + (ee[j] & 0xffffu)]);
        }
        for (int e = kStage + tid; e < ne; e += kFrontThreads) {
            const uint32_t ed = edges[e];
            fr_unite_pair<G>(p, (uint32_t)QW + (ed >> 16), pp[QW + (ed & 0xffffu)]);
        }
        if (lastcol) fr_unite_pair<G>(p, (uint32_t)(QW + ncur - 1), VN);
        bsync<G>();
        // each element under its root: weights into the roots, and the
        // least run of each set that holds one. A thread's record x and run
        // QW + x go together, their reads issued at once (the weight of a
        // non-root never changes: it may be read before its root is known).
        const int npair = max(pcur, ncur);
        for (int x = tid; x < npair; x += kFrontThreads) {
            const bool hr = x < pcur, hn = x < ncur;
            // (a thread with one of the two takes it on both sides)
            const uint32_t er = hr ? (uint32_t)x : (uint32_t)QW + (uint32_t)x;
            const uint32_t en = hn ? (uint32_t)QW + (uint32_t)x : er;
            uint32_t rt[2] = {er, en}, qt[2] = {p[er], p[en]};
            const uint32_t wr = acc[er], wn = acc[en];
            const uint32_t mr = mode == kSweepLocal ? (uint32_t)mp[er] : 0u;
            fr_roots<G, 2>(p, rt, qt);
            const uint32_t rr0 = rt[0], rn0 = rt[1];
            if (hr && rr0 != er) {
                p[er] = rr0;
                atomicAdd((uint32_t *)&acc[rr0], wr);
                if (mode == kSweepLocal) atomicMin((uint32_t *)&mp[rr0], mr);
            }
            if (hn) {
                if (rn0 != en) {
                    p[en] = rn0;
                    atomicAdd((uint32_t *)&acc[rn0], wn);
                }
                atomicMin((uint32_t *)&minrun[rn0], (uint32_t)x);
            }
        }
        if (vthread) {
            uint32_t rt[1] = {VN}, qt[1] = {p[VN]};
            fr_roots<G, 1>(p, rt, qt);
            const uint32_t rv = rt[0];
            if (rv != VN) {
                p[VN] = rv;
                atomicAdd((uint32_t *)&acc[rv], (uint32_t)acc[VN]);
                if (mode == kSweepLocal) atomicMin((uint32_t *)&mp[rv], (uint32_t)mp[VN]);
            }
        }
        bsync<G>();
        // row y settled: the previous set for row y + 1 (the links of row y
        // read the other set, which row y + 1 rewrites, before the barrier
        // after the unions)
        pbase = cbase;
        pcur = ncur;
        U32 *t0 = p; p = pp; pp = t0;
        U32 *t1a = acc; acc = pacc; pacc = t1a;
        U32 *t2m = minrun; minrun = pmin; pmin = t2m;
        U32 *t3 = mp; mp = pmp; pmp = t3;
    };
    Ahead ah0{}, ah1{};
    if (yp + 1 <= y1) prefetch(yp + 1, ah0);
    bsync<G>();
    for (int y = yp + 1; y <= y1; y += 2) {
        sweep_row(y, ah0, ah1);
        if (y + 1 <= y1) sweep_row(y + 1, ah1, ah0);
    }
    // the last row's outputs, and (kSweepLocal) each port's record in it
    for (int i = tid; i < pcur; i += kFrontThreads) {
        const uint32_t r = pp[QW + i];
        const uint32_t l = pmin[r];
        g_rlab[pbase + i] = l;
        if (l == (uint32_t)i) g_recw[pbase + i] = pacc[r];
    }
    if (mode == kSweepLocal) {
        for (int x = tid; x < nport; x += kFrontThreads) {
            const uint32_t q = port[x];  // its record in row y1 - 1 (row yp when y1 = yp + 1)
            uint32_t g;
            if (q & kPortEnded) {
                g = q & ~kPortEnded;
            } else {
                const uint32_t r = pp[q], m = pmin[r];
                g = pmp[r];
                g_rroot[portslot + g] = m != 0xffffffffu ? kPortLive | m : pacc[r];
            }
            g_rlost[portslot + x] = g;
        }
    }
}

template <int T>
__global__ __launch_bounds__(T) void k_front(const Frame *frames) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ uint32_t small[T / 64];
    const Frame F = frames[blockIdx.y];
    if (F.fscr)
        front_sweep<true, T, kSweepFull>(F, F.fscr, small, 0, F.qh - 1, 0, F.qh - 1);
    else
        front_sweep<false, T, kSweepFull>(F, lds, small, 0, F.qh - 1, 0, F.qh - 1);
}

constexpr int kBandRows = 16;      // rows of a band of the split front at least
constexpr int kMaxFrontBands = 64;  // bands of a frame at most

// Row bands of a frame for the split front: band k of nb holds rows
// [k qh / nb, (k + 1) qh / nb); nb = qh / band_rows, at most max_bands
// (one band: the whole-frame sweep). A frame whose front state is in HBM
// has a region of it for each band (front_band_stride bytes apart).
__device__ __forceinline__ int front_bands(const Frame &F, int band_rows, int max_bands) {
    return max(1, min(max_bands, F.qh / max(band_rows, 2)));
}
__host__ __device__ __forceinline__ size_t front_band_stride(int qw) {
    return ((size_t)68 * qw + 128 + 255) & ~(size_t)255;
}
__device__ __forceinline__ int band_start(const Frame &F, int k, int nb) { return (int)((int64_t)k * F.qh / nb); }

// Step 1 of the split front: band 0 sweeps rows from 0 as the whole-frame
// sweep does (its outputs are final); every other band sweeps its rows
// from ports (kSweepLocal). A frame of one band is swept whole here.
template <int T>
__global__ __launch_bounds__(T) void k_front_local(const Frame *frames, int band_rows, int max_bands) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ uint32_t small[T / 64];
    const Frame F = frames[blockIdx.y];
    const int nb = front_bands(F, band_rows, max_bands), k = (int)blockIdx.x;
    if (k >= nb) return;
    const int y0 = band_start(F, k, nb), y1 = band_start(F, k + 1, nb) - 1;
    if (F.fscr) {
        uint8_t *mine = F.fscr + (size_t)k * front_band_stride(F.qw);
        if (k == 0)
            front_sweep<true, T, kSweepFull>(F, mine, small, 0, y1, 0, y1);
        else
            front_sweep<true, T, kSweepLocal>(F, mine, small, y0, y1, y0, y1);
    } else if (k == 0) {
        front_sweep<false, T, kSweepFull>(F, lds, small, 0, y1, 0, y1);
    } else {
        front_sweep<false, T, kSweepLocal>(F, lds, small, y0, y1, y0, y1);
    }
}

// Step 2: band by band (one block per frame), the final records of each
// band's last row: the sets of the row above the band (final records and
// weights), the band's port components and its sets in its last row
// (records and weights without what lies above) joined. Elements: the
// records above at [0, QW), the band's last-row records at QW + [0, QW),
// the port components (by least port) at 2 QW + [0, QW).
//
// The bands form a chain through the row above each band only: band k's
// final last-row records are band k + 1's records above. Element x of a
// row belongs to thread x mod T (slot x / T of its registers) in both
// roles, so each step hands its results to the next in registers, and the
// step's other inputs (the ports of the row above, the band's last-row
// records) are loaded one band ahead: a step waits on LDS only.
template <int T>
__device__ void front_carry_loop(const Frame &F, int nb, uint32_t *p, uint32_t *w, uint32_t *mn) {
    const uint32_t QW = (uint32_t)F.qw;
    const int tid = threadIdx.x;
    auto settle = [&](uint32_t x) {  // (reads only until the element's own root write)
        uint32_t rt[1] = {x}, qt[1] = {p[x]};
        fr_roots<false, 1>(p, rt, qt);
        const uint32_t r = rt[0];
        if (r != x) {
            p[x] = r;
            atomicAdd(&w[r], w[x]);  // (roots no longer change: a non-root's weight is read once)
        }
        return r;
    };
    for (int k = 1; k < nb; k++) {
        const int r0 = band_start(F, k, nb) - 1, e = band_start(F, k + 1, nb) - 1;
        const uint32_t b0 = F.row_off[r0], n0 = F.rcnt[r0], be = F.row_off[e], ne = F.rcnt[e];
        for (uint32_t x = tid; x < n0; x += T) {
            p[x] = x;
            w[x] = F.rlab[b0 + x] == x ? F.recw[b0 + x] : 0u;
            mn[x] = 0xffffffffu;
            // a component's element: its weight if it ended inside the band
            // (a live one's pixels are in its last-row record's weight)
            const bool comp = F.rlost[b0 + x] == x;
            const uint32_t info = comp ? F.rroot[b0 + x] : 0u;
            p[2 * QW + x] = 2 * QW + x;
            w[2 * QW + x] = comp && !(info & kPortLive) ? info : 0u;
            mn[2 * QW + x] = 0xffffffffu;
        }
        for (uint32_t i = tid; i < ne; i += T) {
            p[QW + i] = QW + i;
            w[QW + i] = F.rlab[be + i] == i ? F.recw[be + i] : 0u;
            mn[QW + i] = 0xffffffffu;
        }
        __syncthreads();
        for (uint32_t x = tid; x < n0; x += T) {
            const uint32_t g = F.rlost[b0 + x];
            fr_unite<false>(p, F.rlab[b0 + x], 2 * QW + g);
            if (g == x) {
                const uint32_t info = F.rroot[b0 + x];
                if (info & kPortLive) fr_unite<false>(p, 2 * QW + x, QW + (info & ~kPortLive));
            }
        }
        __syncthreads();
        for (uint32_t x = tid; x < n0; x += T) {
            settle(x);
            settle(2 * QW + x);
        }
        for (uint32_t i = tid; i < ne; i += T) {
            const uint32_t r = settle(QW + i);
            if (F.rlab[be + i] == i) atomicMin(&mn[r], i);  // a record of the band: its first run
        }
        __syncthreads();
        for (uint32_t i = tid; i < ne; i += T) {
            const uint32_t r = p[QW + F.rlab[be + i]], l = mn[r];
            F.rlab[be + i] = l;
            if (l == i) F.recw[be + i] = w[r];
        }
        __syncthreads();
    }
}

template <int T>
__global__ __launch_bounds__(T) void k_front_carry(const Frame *frames, int band_rows, int max_bands) {
    extern __shared__ __align__(16) uint8_t lds[];
    constexpr int kR = 4;  // elements of a row a thread at most (rows of up to kR T runs)
    const Frame F = frames[blockIdx.y];
    const int nb = front_bands(F, band_rows, max_bands);
    if (nb < 2) return;
    const uint32_t QW = (uint32_t)F.qw;
    uint32_t *p = reinterpret_cast<uint32_t *>(lds);
    uint32_t *w = p + 3 * QW;
    uint32_t *mn = w + 3 * QW;
    if (QW > (uint32_t)(kR * T)) {  // a row may hold more runs than the registers: every input from HBM
        front_carry_loop<T>(F, nb, p, w, mn);
        return;
    }
    const int tid = threadIdx.x;
    const auto g_rlost = gmem(F.rlost), g_rroot = gmem(F.rroot), g_rlab = gmem(F.rlab), g_recw = gmem(F.recw);
    const auto g_row_off = gmem(F.row_off), g_rcnt = gmem(F.rcnt);
    // each band's row above and last row: first slot and run count, all
    // read at once
    __shared__ uint32_t rows[kMaxFrontBands][4];
    for (int k = tid; k < nb; k += T) {
        const int r0 = band_start(F, k, nb) - 1, e = band_start(F, k + 1, nb) - 1;
        if (k > 0) {
            rows[k][0] = g_row_off[r0];
            rows[k][1] = g_rcnt[r0];
            rows[k][2] = g_row_off[e];
            rows[k][3] = g_rcnt[e];
        }
    }
    __syncthreads();
    // a band's inputs: the ports of its row above (component of each port,
    // and the component's weight or live record) and its last-row records
    struct In {
        uint32_t b0, n0, be, ne;
        uint32_t g[kR], info[kR], lab[kR], wt[kR];
    };
    auto load = [&](int k, In &in) {
        in.b0 = rows[k][0];
        in.n0 = rows[k][1];
        in.be = rows[k][2];
        in.ne = rows[k][3];
#pragma unroll
        for (int j = 0; j < kR; j++) {
            const uint32_t x = (uint32_t)(tid + j * T);
            in.g[j] = x < in.n0 ? g_rlost[in.b0 + x] : 0u;
            in.info[j] = x < in.n0 ? g_rroot[in.b0 + x] : 0u;
            in.lab[j] = x < in.ne ? g_rlab[in.be + x] : 0u;
            in.wt[j] = x < in.ne ? g_recw[in.be + x] : 0u;
        }
    };
    // the final records above band 1: band 0's last row (final after step 1)
    uint32_t alab[kR], aw[kR];
    {
        const uint32_t b0 = rows[1][0], n0 = rows[1][1];
#pragma unroll
        for (int j = 0; j < kR; j++) {
            const uint32_t x = (uint32_t)(tid + j * T);
            alab[j] = x < n0 ? g_rlab[b0 + x] : 0u;
            aw[j] = x < n0 && alab[j] == x ? g_recw[b0 + x] : 0u;
        }
    }
    In cur, nxt;
    load(1, cur);
    for (int k = 1; k < nb; k++) {
        if (k + 1 < nb) load(k + 1, nxt);  // in flight through this band's step
        const uint32_t n0 = cur.n0, ne = cur.ne;
#pragma unroll
        for (int j = 0; j < kR; j++) {
            const uint32_t x = (uint32_t)(tid + j * T);
            if (x < n0) {
                p[x] = x;
                w[x] = aw[j];
                mn[x] = 0xffffffffu;
                // a component's element: its weight if it ended inside the
                // band (a live one's pixels are in its last-row record's
                // weight); its mn slot holds its least record above
                const bool comp = cur.g[j] == x;
                p[2 * QW + x] = 2 * QW + x;
                w[2 * QW + x] = comp && !(cur.info[j] & kPortLive) ? cur.info[j] : 0u;
                mn[2 * QW + x] = 0xffffffffu;
            }
            if (x < ne) {
                p[QW + x] = QW + x;
                w[QW + x] = cur.lab[j] == x ? cur.wt[j] : 0u;
                mn[QW + x] = 0xffffffffu;
            }
        }
        bsync<false>();  // LDS only: the next band's loads stay in flight
        // each component's least record above (the records above are roots
        // here: singletons until the unions below)
#pragma unroll
        for (int j = 0; j < kR; j++) {
            const uint32_t x = (uint32_t)(tid + j * T);
            if (x < n0) atomicMin(&mn[2 * QW + cur.g[j]], alab[j]);
        }
        bsync<false>();
        // the component under its least record, every port's record united
        // with it, and a live component with its last-row record: the sets
        // of the port-by-port unions, with fewer hooks racing on one root
        // (a component's node is never a root after this, so its mn slot is
        // free; roots keep their mn)
#pragma unroll
        for (int j = 0; j < kR; j++) {
            const uint32_t x = (uint32_t)(tid + j * T);
            if (x < n0) {
                const uint32_t g = cur.g[j], m = mn[2 * QW + g];
                if (g == x) p[2 * QW + x] = m;
                if (alab[j] != m) fr_unite_pair<false>(p, alab[j], m);
                if (g == x && (cur.info[j] & kPortLive))
                    fr_unite_pair<false>(p, 2 * QW + x, QW + (cur.info[j] & ~kPortLive));
            }
        }
        bsync<false>();
        // each element under its root (a position's three elements found
        // together; a missing one takes another's place), the weights of
        // the non-roots into the roots
#pragma unroll
        for (int j = 0; j < kR; j++) {
            const uint32_t x = (uint32_t)(tid + j * T);
            const bool ha = x < n0, hl = x < ne;
            if (!ha && !hl) continue;
            const uint32_t ea = ha ? x : QW + x, ec = ha ? 2 * QW + x : QW + x, el = hl ? QW + x : x;
            uint32_t rt[3] = {ea, ec, el}, qt[3] = {p[ea], p[ec], p[el]};
            const uint32_t wa = w[ea], wc = w[ec], wl = w[el];
            fr_roots<false, 3>(p, rt, qt);
            if (ha) {
                if (rt[0] != ea) {
                    p[ea] = rt[0];
                    atomicAdd(&w[rt[0]], wa);
                }
                if (rt[1] != ec) {
                    p[ec] = rt[1];
                    atomicAdd(&w[rt[1]], wc);
                }
            }
            if (hl) {
                if (rt[2] != el) {
                    p[el] = rt[2];
                    atomicAdd(&w[rt[2]], wl);
                }
                if (cur.lab[j] == x) atomicMin(&mn[rt[2]], x);  // a record of the band: its first run
            }
        }
        bsync<false>();
        // the band's final last-row records: to HBM, and (the next band's
        // records above) to the registers of the same elements
#pragma unroll
        for (int j = 0; j < kR; j++) {
            const uint32_t i = (uint32_t)(tid + j * T);
            if (i < ne) {
                const uint32_t r = p[QW + cur.lab[j]], l = mn[r];
                const uint32_t wr = w[r];
                g_rlab[cur.be + i] = l;
                if (l == i) g_recw[cur.be + i] = wr;
                alab[j] = l;
                aw[j] = l == i ? wr : 0u;
            }
        }
        bsync<false>();
        cur = nxt;
    }
}

// Step 3: every band but the first sweeps its rows again from the final
// records above it (its last row is final already).
template <int T>
__global__ __launch_bounds__(T) void k_front_final(const Frame *frames, int band_rows, int max_bands) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ uint32_t small[T / 64];
    const Frame F = frames[blockIdx.y];
    const int nb = front_bands(F, band_rows, max_bands), k = (int)blockIdx.x + 1;
    if (k >= nb) return;
    const int y0 = band_start(F, k, nb), y1 = band_start(F, k + 1, nb) - 2;
    if (y1 < y0) return;
    if (F.fscr)
        front_sweep<true, T, kSweepFinal>(F, F.fscr + (size_t)k * front_band_stride(F.qw), small, y0, y1, 1, 0);
    else
        front_sweep<false, T, kSweepFinal>(F, lds, small, y0, y1, 1, 0);
}

// A wave per row: the front names each record of a row by its first run;
// here the records are numbered densely in that order (rlab per run, recw
// per record slot, nrec), the names mapped through rroot (free until the
// replay, and for row 0, whose records are its runs, the identity it
// keeps).
__global__ __launch_bounds__(256) void k_front_dense(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int y = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (y >= F.qh) return;
    const uint32_t b = F.row_off[y], n = F.rcnt[y];
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t carry = 0;
    constexpr int kG = 16;  // rows of up to 1024 runs: every chunk's reads at once
    if (n <= 64u * kG) {
        uint32_t lab[kG], wv[kG], dn[kG];
#pragma unroll
        for (int j = 0; j < kG; j++) {
            const uint32_t i = 64u * j + (uint32_t)lane;
            lab[j] = i < n ? F.rlab[b + i] : 0xffffffffu;
            wv[j] = i < n ? F.recw[b + i] : 0u;  // (every read before any write)
        }
#pragma unroll
        for (int j = 0; j < kG; j++) {
            const uint32_t i = 64u * j + (uint32_t)lane;
            const uint64_t fb = __ballot(lab[j] == i);
            dn[j] = carry + (uint32_t)__popcll(fb & lt);
            carry += (uint32_t)__popcll(fb);
        }
#pragma unroll
        for (int j = 0; j < kG; j++) {
            const uint32_t i = 64u * j + (uint32_t)lane;
            if (lab[j] == i) {
                F.rroot[b + i] = dn[j];
                F.recw[b + dn[j]] = wv[j];
            }
        }
        __threadfence_block();
        uint32_t nl[kG];
#pragma unroll
        for (int j = 0; j < kG; j++) {
            const uint32_t i = 64u * j + (uint32_t)lane;
            nl[j] = i < n ? F.rroot[b + lab[j]] : 0u;
        }
#pragma unroll
        for (int j = 0; j < kG; j++) {
            const uint32_t i = 64u * j + (uint32_t)lane;
            if (i < n) F.rlab[b + i] = nl[j];
        }
    } else {
        for (uint32_t c0 = 0; c0 < n; c0 += 64) {
            const uint32_t i = c0 + (uint32_t)lane;
            const bool first = i < n && F.rlab[b + i] == i;
            const uint64_t fb = __ballot(first);
            const uint32_t d = carry + (uint32_t)__popcll(fb & lt);
            const uint32_t w = first ? F.recw[b + i] : 0u;  // read before any lane of the chunk writes
            if (first) {
                F.rroot[b + i] = d;
                F.recw[b + d] = w;  // d <= i: slots already read
            }
            carry += (uint32_t)__popcll(fb);
        }
        __threadfence_block();
        for (uint32_t i = (uint32_t)lane; i < n; i += 64) F.rlab[b + i] = F.rroot[b + F.rlab[b + i]];
    }
    if (lane == 0) F.nrec[y] = carry;
}

// ---------------------------------------------------------------------
// Step 2, k_replay: every row's links in the reference's order, all rows
// at once.
//
// At the start of row y the sets touching row y-1 are distinct, and their
// weights are the ones the sequential replay has there (a weight is a
// pixel count: step 1 knows it). So row y can be replayed on its own: the
// elements are those records (weights from step 1, roots unknown: named
// by slot, kSym | slot), the current runs and the virtual node. A link
// only ever touches the sets of one row-y record (the group of its
// current run, rlab from step 1), so the groups are independent: each
// lane replays the links of one group in order (fast/unionfind.c:15-45,
// 112-146), and the last-column link comes last.
// Output: parent[] of the row's runs and virtual node (a node or a
// record slot), rlost[slot] of each previous record (the winner it lost
// to, or kLostNone / kLostRetired), rroot[slot] of each row-y record (the
// root of its set as replayed: a node or kSym | a previous record's slot).
//
// Rows of at most C = 448 runs (512 for calls of up to 32 frames; each of the
// two rows) and replay_links(C) links replay in LDS (k_replay, a wave per
// row, about 9 KiB at 448: 18 rows in flight a CU); the wider rows, which it
// lists, in k_replay_wide: first those of up to 704 runs, then the others with
// LDS for the widest row of the call (C = qw, 3C links); rows beyond that, and
// all rows when asked, in HBM scratch sized for the row width (k_replay_hbm).

// Byte offsets of a row's replay memory for C runs, L links and ix-byte
// link indices.
struct ReplayLayout {
    size_t le, nxt, head, labA, hasl, total;
};
__host__ __device__ inline ReplayLayout replay_layout(int C, int L, int ix, int eb = 8) {
    ReplayLayout r;
    size_t o = (size_t)eb * (2 * C + 2);  // ent: [2C + 2] x (parent, weight), 8 bytes or packed in 4
    r.le = o;
    o += (size_t)4 * L;                  // link: current run << 16 | record, by group
    r.nxt = o;
    o = (o + (size_t)ix * C + 3) & ~(size_t)3;  // each current run's record
    r.head = o;
    o = (o + (size_t)ix * C + 3) & ~(size_t)3;  // each group's end in le
    r.labA = o;
    o = (o + (size_t)2 * C + 3) & ~(size_t)3;   // record of each previous-row run
    r.hasl = o;
    o = (o + (size_t)C + 15) & ~(size_t)15;     // record has a link
    r.total = o;
    return r;
}

// LDS capacity (runs per row) of the first replay launch for frames up to
// max_qw wide, and of the second (the rows wider than the first's).
// Links of a row replayed in LDS with cap runs (rows with more go to the
// wide launch): a quarter more than the runs, in 64s (a 1080p row has at
// most about 1.05 links a run).
__host__ __device__ inline int replay_links(int cap) { return (cap + cap / 4 + 63) & ~63; }
// Runs a row of the first replay launch: 448 for calls of more than 32
// frames (more rows in flight a CU; the few wider rows spread over the wide
// launch), 512 for smaller calls, whose wider rows would be a serial tail
// there.
int replay_cap(int max_qw, int nframes) {
    return std::min(nframes > 32 ? 448 : 512, (std::max(max_qw, 1) + 63) & ~63);
}
// The listed wider rows replay in two LDS launches: rows of up to 704 runs
// (a 2448x2048 row has at most about 620) with replay_links(704) links, about
// 14 KiB packed (11 rows in flight a CU), then the rest of the rows up to the
// widest row of the call (3C links, 3 rows a CU at C = 1280).
constexpr int kReplayMidCap = 704;
int replay_wide_cap(int max_qw) { return (std::max(max_qw, 1) + 63) & ~63; }
int replay_mid_cap(int max_qw) { return std::min(replay_wide_cap(max_qw), kReplayMidCap); }
size_t replay_lds_bytes(int cap, bool pk) { return replay_layout(cap, replay_links(cap), 2, pk ? 4 : 8).total; }
size_t replay_wide_lds_bytes(int cap) { return replay_layout(cap, 3 * cap, 2).total; }
__host__ __device__ inline size_t replay_hbm_bytes_d(int qw) { return replay_layout(qw, 3 * qw, 4).total; }
size_t replay_hbm_bytes(int qw) { return replay_hbm_bytes_d(qw); }

// Waves of k_replay_hbm per frame: a few for the odd wide row, 64 when
// every row takes it.
constexpr int kReplayHbmWaves = 16, kReplayHbmWavesAll = 64;

__device__ __forceinline__ bool replay_fits(const Frame &F, int y, int cap, int lcap) {
    return (int)F.rcnt[y - 1] <= cap && (int)F.rcnt[y] <= cap && edge_count(F.ecnt[y]) <= lcap;
}

// A row's elements: (parent, weight) pairs, or (PK) packed in one word as
// parent << 21 | weight, for rows of at most 1023 runs in frames of under
// 2^21 pixels (every weight is a pixel count of the frame).
template <bool G, bool PK>
struct Ents {
    using U32 = typename Mem<G>::U32;
    static constexpr uint32_t kWB = 21, kWM = (1u << kWB) - 1u;
    U32 *e;
    __device__ __forceinline__ uint2 get(uint32_t x) const {
        if constexpr (PK) {
            const uint32_t v = e[x];
            return make_uint2(v >> kWB, v & kWM);
        } else if constexpr (G) {
            uint2 v;
            v.x = e[2 * x];
            v.y = e[2 * x + 1];
            return v;
        } else {
            return *reinterpret_cast<const uint2 *>(e + 2 * x);
        }
    }
    __device__ __forceinline__ uint32_t par(uint32_t x) const {
        if constexpr (PK) return e[x] >> kWB;
        else return e[2 * x];
    }
    __device__ __forceinline__ uint32_t wt(uint32_t x) const {
        if constexpr (PK) return e[x] & kWM;
        else return e[2 * x + 1];
    }
    __device__ __forceinline__ void set(uint32_t x, uint32_t p, uint32_t w) const {
        if constexpr (PK) {
            e[x] = (p << kWB) | w;
        } else {
            e[2 * x] = p;
            e[2 * x + 1] = w;
        }
    }
    // a root x hooked under p: its weight is never read again (a find
    // returns its root's entry), so the packed word is written, not merged
    __device__ __forceinline__ void set_par(uint32_t x, uint32_t p) const {
        if constexpr (PK) e[x] = p << kWB;
        else e[2 * x] = p;
    }
    // the weight of a root x (its own parent): the packed word written whole
    __device__ __forceinline__ void set_wt(uint32_t x, uint32_t w) const {
        if constexpr (PK) e[x] = (x << kWB) | w;
        else e[2 * x + 1] = w;
    }
    // find with path halving; returns the root's entry with .x = the root
    __device__ __forceinline__ uint2 find(uint32_t x) const {
        for (;;) {
            const uint2 v = get(x);
            if (v.x == x) return v;
            const uint2 pv = get(v.x);
            if (pv.x == v.x) return pv;
            if constexpr (PK) e[x] = (pv.x << kWB) | v.y;
            else e[2 * x] = pv.x;
            x = pv.x;
        }
    }
    __device__ __forceinline__ uint32_t find_ro(uint32_t x) const {
        for (uint32_t q; (q = par(x)) != x;) x = q;
        return x;
    }
};

// Replay of row y with capacity C runs and L links (the row fits).
template <bool G, bool PK>
__device__ __forceinline__ void replay_row(const Frame &F, int y, uint8_t *base, int C, int L) {
    using U32 = typename Mem<G>::U32;
    using U16 = typename Mem<G>::U16;
    using IX = typename std::conditional<G, volatile uint32_t, uint16_t>::type;
    using U8 = typename std::conditional<G, volatile uint8_t, uint8_t>::type;
    const int lane = threadIdx.x;
    // the frame's arrays in HBM as global pointers (no flat accesses among the LDS work)
    const auto g_row_off = gmem(F.row_off);
    const auto g_nrec = gmem(F.nrec);
    const auto g_rcnt = gmem(F.rcnt);
    const auto g_ecnt = gmem(F.ecnt);
    const auto g_runs = gmem(F.runs);
    const auto g_uedge = gmem(F.uedge);
    const auto g_rlab = gmem(F.rlab);
    const auto g_recw = gmem(F.recw);
    const auto g_rroot = gmem(F.rroot);
    const auto g_parent = gmem(F.parent);
    const auto g_rlost = gmem(F.rlost);
    const int QW = F.qw;
    const ReplayLayout ly = replay_layout(C, L, G ? 4 : 2, PK ? 4 : 8);
    const Ents<G, PK> ent{reinterpret_cast<U32 *>(base)};
    U32 *le = reinterpret_cast<U32 *>(base + ly.le);
    IX *nxt = reinterpret_cast<IX *>(base + ly.nxt);
    IX *head = reinterpret_cast<IX *>(base + ly.head);
    U16 *labA = reinterpret_cast<U16 *>(base + ly.labA);
    U8 *hasl = reinterpret_cast<U8 *>(base + ly.hasl);
    const uint32_t pb = g_row_off[y - 1], cb = g_row_off[y], vb = g_row_off[F.qh];
    const int m = (int)g_nrec[y - 1], nprev = (int)g_rcnt[y - 1];
    const uint32_t ecnt = g_ecnt[y];
    const int ncur = (int)g_rcnt[y], ne = edge_count(ecnt), ng = (int)g_nrec[y];
    const auto cur = g_runs + (int64_t)y * QW;
    const auto edges = g_uedge + (int64_t)y * 3 * QW;
    const auto rlab_cur = g_rlab + cb;
    const int ev = m + ncur;
    // the elements (records with their weights, current runs with their
    // lengths), each previous run's record (labA) and each current run's
    // (nxt, until the group ends take it over), 4 entries a lane of each
    // array at a time, their loads issued together
    for (int b0 = 0; b0 < m || b0 < ncur || b0 < nprev; b0 += 256) {
        uint32_t rw[4], rr[4], rp[4], rc[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = b0 + 64 * q + lane;
            rw[q] = i < m ? g_recw[pb + i] : 0u;
            rr[q] = i < ncur ? cur[i] : 0u;
            rc[q] = i < ncur ? rlab_cur[i] : 0u;
            rp[q] = i < nprev ? g_rlab[pb + i] : 0u;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = b0 + 64 * q + lane;
            if (i < m) {
                ent.set((uint32_t)i, (uint32_t)i, rw[q]);
                hasl[i] = 0;
            }
            if (i < ncur) {
                ent.set((uint32_t)(m + i), (uint32_t)(m + i), run_end(rr[q]) - run_start(rr[q]) + 1);
                nxt[i] = rc[q];
            }
            if (i < nprev) labA[i] = (uint16_t)rp[q];
        }
    }
    if (lane == 0) {
        ent.set((uint32_t)ev, (uint32_t)ev, 1u);
    }
    const bool lastcol = ncur > 0 && (ecnt & kLastColLink) != 0;
    wsync<G>();
    // the links by group, stably (each group's links in order, contiguous
    // in le; group g's links from head[g - 1] (0 for g = 0) to head[g]):
    // counts per group, their starts by a scan, then each link at its
    // group's running end plus its rank among the group's lanes of its
    // chunk of 64 (peers by ballot over the group bits; one lane of each
    // group adds). A link's group: its current run's record (nxt). Each
    // chunk's edges are loaded while the chunk before it runs.
    const int gbits = ng <= 1 ? 1 : 32 - __clz(ng - 1);
    const uint64_t lt = (1ull << lane) - 1ull;
    auto group_peers = [&](uint32_t g, bool valid) {
        uint64_t peers = __ballot(valid);
        for (int bit = 0; bit < gbits; bit++) {
            const uint64_t bb = __ballot((g >> bit) & 1u);
            peers &= ((g >> bit) & 1u) ? bb : ~bb;
        }
        return peers;
    };
    for (int g = lane; g < ng; g += 64) head[g] = 0;
    wsync<G>();
    auto scan_heads = [&]() {  // each group's count to its first link
        for (int g0 = 0, carry = 0; g0 < ng; g0 += 64) {
            const int g = g0 + lane;
            const uint32_t c = g < ng ? (uint32_t)head[g] : 0u;
            const uint32_t inc = wave_scan_incl(c);
            if (g < ng) head[g] = (uint32_t)carry + inc - c;
            carry += __builtin_amdgcn_readlane((int)inc, 63);
        }
        wsync<G>();
    };
    // Rows of up to kRegChunks chunks of links (every row of the first
    // launch): the links, their groups and records and each chunk's peer
    // masks stay in registers from the count to the placement, the loads
    // of all chunks (HBM, then LDS) issued at once. Other rows: each pass
    // streams the links, a chunk loaded ahead.
    constexpr int kRegChunks = 10;
    if (ne <= 64 * kRegChunks) {
        uint32_t eds[kRegChunks], gs[kRegChunks], las[kRegChunks];
        uint64_t prs[kRegChunks];
#pragma unroll
        for (int c = 0; c < kRegChunks; c++) {
            const int e = 64 * c + lane;
            eds[c] = e < ne ? (uint32_t)edges[e] : 0u;
        }
#pragma unroll
        for (int c = 0; c < kRegChunks; c++) {
            const bool valid = 64 * c + lane < ne;
            gs[c] = valid ? (uint32_t)nxt[eds[c] >> 16] : 0u;
            las[c] = valid ? (uint32_t)labA[eds[c] & 0xffffu] : 0u;
        }
#pragma unroll
        for (int c = 0; c < kRegChunks; c++) {
            if (64 * c >= ne) break;
            const bool valid = 64 * c + lane < ne;
            if (valid) hasl[las[c]] = 1;
            prs[c] = group_peers(gs[c], valid);
            if (valid && !(prs[c] & lt)) head[gs[c]] = (uint32_t)head[gs[c]] + (uint32_t)__popcll(prs[c]);
            wsync<G>();
        }
        scan_heads();
#pragma unroll
        for (int c = 0; c < kRegChunks; c++) {
            if (64 * c >= ne) break;
            const bool valid = 64 * c + lane < ne;
            if (valid) le[(uint32_t)head[gs[c]] + (uint32_t)__popcll(prs[c] & lt)] = (eds[c] & 0xffff0000u) | las[c];
            wsync<G>();
            if (valid && !(prs[c] & lt)) head[gs[c]] = (uint32_t)head[gs[c]] + (uint32_t)__popcll(prs[c]);
            wsync<G>();
        }
    } else {
        uint32_t edn = lane < ne ? edges[lane] : 0u;
        for (int b0 = 0; b0 < ne; b0 += 64) {
            const int e = b0 + lane;
            const bool valid = e < ne;
            const uint32_t ed = edn;
            edn = e + 64 < ne ? edges[e + 64] : 0u;
            const uint32_t g = valid ? (uint32_t)nxt[ed >> 16] : 0u;
            if (valid) hasl[labA[ed & 0xffffu]] = 1;
            const uint64_t peers = group_peers(g, valid);
            if (valid && !(peers & lt)) head[g] = (uint32_t)head[g] + (uint32_t)__popcll(peers);
            wsync<G>();
        }
        scan_heads();
        edn = lane < ne ? edges[lane] : 0u;
        for (int b0 = 0; b0 < ne; b0 += 64) {
            const int e = b0 + lane;
            const bool valid = e < ne;
            const uint32_t ed = edn;
            edn = e + 64 < ne ? edges[e + 64] : 0u;
            const uint32_t g = valid ? (uint32_t)nxt[ed >> 16] : 0u;
            const uint64_t peers = group_peers(g, valid);
            if (valid) le[(uint32_t)head[g] + (uint32_t)__popcll(peers & lt)] = (ed & 0xffff0000u) | labA[ed & 0xffffu];
            wsync<G>();
            if (valid && !(peers & lt)) head[g] = (uint32_t)head[g] + (uint32_t)__popcll(peers);
            wsync<G>();
        }
    }
    auto gstart = [&](uint32_t g) -> uint32_t { return g ? (uint32_t)head[g - 1] : 0u; };
    // the groups that have links, listed (in labA, free by now), so that
    // the lanes share out those and not the rows' many linkless records
    U16 *glist = labA;
    int ngl = 0;
    for (int g0 = 0; g0 < ng; g0 += 64) {
        const int g = g0 + lane;
        const bool has = g < ng && (uint32_t)head[g] != gstart((uint32_t)g);
        const uint64_t hb = __ballot(has);
        if (has) glist[ngl + __popcll(hb & lt)] = (uint16_t)g;
        ngl += __popcll(hb);
    }
    wsync<G>();
    // each lane replays its groups' links in order; the current run's set
    // is kept in registers across its links, the next link is read ahead.
    // cr: the group's last root, cw its weight (as stored, since the run
    // change that left it; while cr == ra the stored weight lags, unread):
    // a record whose parent is cr needs no read of cr (find's first step)
    for (int k = lane; k < ngl; k += 64) {
        const uint32_t g = glist[k];
        uint32_t e = gstart(g);
        const uint32_t e1 = head[g];
        int cur_i = -1;
        uint32_t ra = 0, wa = 0, cr = 0xffffffffu, cw = 0;
        uint32_t ed = le[e];
        for (;;) {
            const uint32_t edn = e + 1 < e1 ? (uint32_t)le[e + 1] : 0u;
            const uint32_t i = ed >> 16, l = ed & 0xffffu;
            if ((int)i != cur_i) {
                if (cur_i >= 0) {
                    ent.set_wt(ra, wa);
                    cr = ra;
                    cw = wa;
                }
                cur_i = (int)i;
                ra = (uint32_t)m + i;  // a current run's first link: a fresh root
                wa = ent.wt(ra);
            }
            uint2 b;
            {
                const uint2 v = ent.get(l);
                if (v.x == l) {
                    b = v;
                } else if (v.x == cr) {
                    b = make_uint2(cr, cw);
                } else {
                    const uint2 pv = ent.get(v.x);
                    if (pv.x == v.x) {
                        b = pv;
                    } else {
                        ent.set(l, pv.x, v.y);  // find's halving step on l
                        b = ent.find(pv.x);
                    }
                }
            }
            if (b.x != ra) {
                if (wa > b.y) {
                    ent.set_par(b.x, ra);
                } else {
                    ent.set_par(ra, b.x);
                    ra = b.x;
                }
                wa += b.y;
                cr = ra;
            }
            if (++e == e1) break;
            ed = edn;
        }
        ent.set_wt(ra, wa);
    }
    wsync<G>();
    if (lane == 0 && lastcol) {  // fast/unionfind.c:93-97, last in order
        uint2 a = ent.find((uint32_t)(ev - 1));
        if (a.y > 1u) {
            ent.set_par((uint32_t)ev, a.x);
            ent.set_wt(a.x, a.y + 1);
        } else {
            ent.set_par(a.x, (uint32_t)ev);
            ent.set_wt((uint32_t)ev, a.y + 1);
        }
    }
    wsync<G>();
    auto name = [&](uint32_t x) -> uint32_t {
        if ((int)x < m) return kSym | (pb + x);
        if ((int)x < ev) return cb + (x - (uint32_t)m);
        return vb + (uint32_t)y - 1;
    };
    for (int l = lane; l < m; l += 64) {
        uint32_t q = ent.par((uint32_t)l);
        g_rlost[pb + l] = q == (uint32_t)l ? (hasl[l] ? kLostNone : kLostRetired) : name(q);
    }
    for (int x = m + lane; x <= ev; x += 64) g_parent[name((uint32_t)x)] = name(ent.par((uint32_t)x));
    for (int i = lane; i < ncur; i += 64) g_rroot[cb + (uint32_t)nxt[i]] = name(ent.find_ro((uint32_t)(m + i)));
}

// A wave per row, the rows that fit C = cap runs and replay_links(C) links; the others
// are listed (rwide) for the wide launch. PK: packed elements.
template <bool PK>
__global__ __launch_bounds__(64) void k_replay(const Frame *frames, int cap) {
    extern __shared__ __align__(16) uint8_t lds[];
    const Frame &F = frames[blockIdx.y];
    const int y = 1 + (int)blockIdx.x;
    if (y >= F.qh) return;
    if (!replay_fits(F, y, cap, replay_links(cap))) {
        if (threadIdx.x == 0) F.rwide[atomicAdd(F.counters + kNWideRows, 1u)] = (uint32_t)y;
        return;
    }
    replay_row<false, PK>(F, y, lds, cap, replay_links(cap));
}

// The listed rows of every frame (list 3 of the work queue) that fit
// C = cap runs and lcap links but not pcap runs and plcap links (pcap 0:
// every listed row that fits): persistent, blocks taking list entries in
// strides (a frame cursor). PK: packed elements.
template <bool PK>
__global__ __launch_bounds__(64) void k_replay_wide(const Frame *frames, const WorkQueue *Q, int cap, int lcap,
                                                    int pcap, int plcap) {
    extern __shared__ __align__(16) uint8_t lds[];
    const uint32_t total = Q->total[3];
    const uint32_t *off = Q->off[3];
    uint32_t fi = 0;
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
        fi = frame_of_wave(off, Q->nframes, g, fi);
        const Frame &F = frames[fi];
        const int y = (int)F.rwide[g - off[fi]];
        if (replay_fits(F, y, cap, lcap) && !(pcap && replay_fits(F, y, pcap, plcap)))
            replay_row<false, PK>(F, y, lds, cap, lcap);
        __syncthreads();  // the next row reuses the memory
    }
}

// The other rows (all rows for cap 0): wave b takes rows 1 + b + k W,
// 64 at a time checked by the lanes, in its own part of the frame's rscr.
// cap, lcap: the largest runs and links the LDS launches took.
__global__ __launch_bounds__(64) void k_replay_hbm(const Frame *frames, int cap, int lcap) {
    const Frame &F = frames[blockIdx.y];
    const int lane = threadIdx.x, W = (int)gridDim.x, b = (int)blockIdx.x;
    const size_t wstride = (replay_hbm_bytes_d(F.qw) + 255) & ~(size_t)255;
    uint8_t *mem = F.rscr + (size_t)b * wstride;
    for (int k0 = 0; 1 + b + W * k0 < F.qh; k0 += 64) {
        const int y = 1 + b + W * (k0 + lane);
        uint64_t todo = __ballot(y < F.qh && (cap == 0 || !replay_fits(F, y, cap, lcap)));
        while (todo) {
            const int j = __ffsll((unsigned long long)todo) - 1;
            todo &= todo - 1;
            replay_row<true, false>(F, 1 + b + W * (k0 + j), mem, F.qw, 3 * F.qw);
            __syncthreads();  // the next row reuses the memory
        }
    }
}

// ---------------------------------------------------------------------
// Step 3: the actual roots. rroot[slot] names a node or a record of the
// previous row, so the roots resolve by pointer jumping over the slots
// (k_root_jump, one pass per fivefold growth of the resolved distance, every row
// at once); then k_lost gives each record root that lost in the next row
// its winner as parent, and each set that ended its final weight.

// Hops a k_root_jump pass follows. If every slot's value was at least D
// links up its chain when the pass began, it is at least (kRootHops + 1) D
// up when the pass ends (each value read spans D more), so P passes with
// (kRootHops + 1)^P >= qh reach every root (a chain climbs a row a link).
constexpr int kRootHops = 4;
__global__ __launch_bounds__(256) void k_root_jump(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int y = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (y < 1 || y >= F.qh) return;
    const uint32_t cb = F.row_off[y];
    const int n = (int)F.nrec[y];
    constexpr int kG = 16;  // rows of up to 1024 records: each hop's reads of all records at once
    if (n <= 64 * kG) {
        uint32_t v[kG];
        bool sym[kG];
#pragma unroll
        for (int j = 0; j < kG; j++) {
            const int l = 64 * j + lane;
            v[j] = l < n ? F.rroot[cb + l] : 0u;
            sym[j] = (v[j] & kSym) != 0;
        }
        // kRootHops ancestors' values in turn: each is never a wrong root
#pragma unroll
        for (int h = 0; h < kRootHops; h++) {
#pragma unroll
            for (int j = 0; j < kG; j++)
                if (v[j] & kSym) v[j] = F.rroot[v[j] & ~kSym];
        }
#pragma unroll
        for (int j = 0; j < kG; j++)
            if (sym[j]) F.rroot[cb + 64 * j + lane] = v[j];
        return;
    }
    for (int l = lane; l < n; l += 64) {
        uint32_t v = F.rroot[cb + l];
        if (!(v & kSym)) continue;
#pragma unroll
        for (int h = 0; h < kRootHops; h++)
            if (v & kSym) v = F.rroot[v & ~kSym];
        F.rroot[cb + l] = v;
    }
}

// Row y (1..qh) settles the records of row y-1; row qh ends them all.
__global__ __launch_bounds__(256) void k_lost(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int y = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (y < 1 || y > F.qh) return;
    const uint32_t pb = F.row_off[y - 1];
    const int m = (int)F.nrec[y - 1];
    constexpr int kG = 16;  // rows of up to 1024 records: the reads of all records at once
    if (m <= 64 * kG) {
        uint32_t root[kG], x[kG], w[kG];
#pragma unroll
        for (int j = 0; j < kG; j++) {
            const int l = 64 * j + lane;
            root[j] = l < m ? F.rroot[pb + l] : 0u;
            x[j] = l < m ? (y < F.qh ? F.rlost[pb + l] : kLostRetired) : kLostNone;
            w[j] = l < m ? F.recw[pb + l] : 0u;
        }
#pragma unroll
        for (int j = 0; j < kG; j++)
            if (x[j] != kLostNone && x[j] != kLostRetired && (x[j] & kSym)) x[j] = F.rroot[x[j] & ~kSym];
#pragma unroll
        for (int j = 0; j < kG; j++) {
            if (x[j] == kLostRetired) F.fw[root[j]] = w[j];
            else if (x[j] != kLostNone) F.parent[root[j]] = x[j];
        }
        return;
    }
    for (int l = lane; l < m; l += 64) {
        uint32_t root = F.rroot[pb + l];
        uint32_t x = y < F.qh ? F.rlost[pb + l] : kLostRetired;
        if (x == kLostRetired) {
            F.fw[root] = F.recw[pb + l];
        } else if (x != kLostNone) {
            F.parent[root] = (x & kSym) ? F.rroot[x & ~kSym] : x;
        }
    }
}

// One thread per node: the final root (a parent naming a record slot goes
// to that record's root) and the ok flag (fast/clusters.c:73-94).
__global__ void k_resolve(const Frame *frames, const DevParams *P) {
    const Frame F = frames[blockIdx.y];
    int64_t nn = (int64_t)F.counters[kNRuns] + F.qh;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    uint32_t r = (uint32_t)i;
    // every link points to a node or record that was a root when it was
    // made, so the walk ends; the bound only guards against a corrupted table
    int64_t steps = 0;
    for (; steps <= 2 * nn; steps++) {
        uint32_t q = F.parent[r];
        if (q & kSym) q = F.rroot[q & ~kSym];
        if (q == r) break;
        r = q;
    }
    if (steps > 2 * nn) atomicOr(F.counters + kError, 32u);
    uint32_t ok = (int64_t)F.fw[r] >= (int64_t)P->mcp ? 1u : 0u;
    F.nlab[i] = r | (ok << 31);
}

// One wave per row: plab[y][x] = nlab of the run under pixel x (0 where
// t is 127), the row's virtual node at x = qw-1 (fast/clusters.c:123-142).
__global__ __launch_bounds__(256) void k_labels(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    int y = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (y >= F.qh) return;
    const int lane = threadIdx.x & 63;
    const uint8_t *trow = F.t + (int64_t)y * F.qw;
    uint32_t *out = F.plab + (int64_t)y * F.qw;
    const uint32_t *nl = F.nlab + F.row_off[y];
    const int xmax = F.qw - 2;
    int count = 0;
    uint32_t carry_t = 0;
    const uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    // kG chunks of 64 pixels at a time: their bytes loaded at once, each
    // pixel's run index from the chunks' ballots in order, then the runs'
    // labels gathered at once (two rounds of reads for 1024 pixels)
    constexpr int kG = 16;
    for (int g0 = 0; g0 <= xmax; g0 += 64 * kG) {
        uint32_t tv[kG], lab[kG];
        int idx[kG];
#pragma unroll
        for (int j = 0; j < kG; j++) {
            const int x = g0 + 64 * j + lane;
            tv[j] = x <= xmax ? trow[x] : 0u;
        }
#pragma unroll
        for (int j = 0; j < kG; j++) {
            if (g0 + 64 * j > xmax) break;
            const int x = g0 + 64 * j + lane;
            const bool valid = x <= xmax;
            const uint32_t tp = wave_from_left(tv[j], carry_t);
            const bool st = valid && (x == 0 || tv[j] != tp) && tv[j] != 127u;
            const uint64_t sm = __ballot(st);
            idx[j] = count + __popcll(sm & le) - 1;
            count += __popcll(sm);
            carry_t = (uint32_t)__builtin_amdgcn_readlane((int)tv[j], 63);
        }
#pragma unroll
        for (int j = 0; j < kG; j++) {
            const int x = g0 + 64 * j + lane;
            lab[j] = x <= xmax && tv[j] != 127u ? nl[idx[j]] : 0u;
        }
#pragma unroll
        for (int j = 0; j < kG; j++) {
            const int x = g0 + 64 * j + lane;
            if (x <= xmax) out[x] = lab[j];
        }
    }
    if (lane == 0) {
        out[F.qw - 1] = F.nlab[F.row_off[F.qh] + y];
    }
}

// Lets the front's kernels of T threads take `bytes` of dynamic LDS.
template <int T>
static bool front_lds_allow(int bytes) {
    const void *fs[] = {reinterpret_cast<const void *>(&k_front<T>), reinterpret_cast<const void *>(&k_front_local<T>),
                        reinterpret_cast<const void *>(&k_front_carry<T>),
                        reinterpret_cast<const void *>(&k_front_final<T>)};
    for (const void *f : fs)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess) return false;
    return true;
}

// The front's dynamic LDS limit: 64 KiB, or most of the 160 KiB of a CU
// where the kernels accept that much (asked once; a refusal is not left
// behind as the runtime's last error).
static size_t front_lds_limit() {
    static size_t lim = 0;
    if (!lim) {
        const int big = 159 * 1024;
        bool ok = front_lds_allow<64>(big) && front_lds_allow<128>(big) && front_lds_allow<256>(big) &&
                  front_lds_allow<512>(big) && front_lds_allow<1024>(big);
        (void)hipGetLastError();
        lim = ok ? (size_t)big : (size_t)64 * 1024;
    }
    return lim;
}

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// A frame whose front state does not fit LDS (or forced) keeps it in HBM:
// a region for each band the split front can give it.
size_t front_scratch_bytes(int qw, bool force) {
    size_t f = front_bytes(qw);
    return (force || f > front_lds_limit()) ? front_band_stride(qw) * (size_t)kMaxFrontBands : 0;
}

size_t replay_scratch_bytes(int qw, bool force) {
    return al256(replay_hbm_bytes(qw)) * (size_t)(force ? kReplayHbmWavesAll : kReplayHbmWaves);
}

// The launch shapes of a call: max_lds_qw is the widest frame whose front
// runs in LDS (0: none), max_qw the widest frame, max_weight the most
// pixels a set of a frame can weigh (qw qh plus the virtual nodes).
CclPlan ccl_plan(int max_lds_qw, int max_qw, int64_t max_weight, bool force_scratch, int nframes) {
    CclPlan c{};
    c.front = max_lds_qw > 0 ? (int)front_bytes(max_lds_qw) : 0;
    // the split front: LDS for the state of its LDS frames (with the
    // ports) and for the band-by-band join of the widest frame (three
    // arrays of 3 qw), when both fit
    {
        const size_t need = std::max(max_lds_qw > 0 ? front_band_bytes(max_lds_qw) : (size_t)0,
                                     (size_t)36 * (size_t)max_qw + 64);
        c.front_band = need <= front_lds_limit() ? (int)need : 0;
    }
    // rows of up to replay_cap runs in LDS, the others (all when forced) in HBM
    c.replay_cap = force_scratch ? 0 : replay_cap(max_qw, nframes);
    // packed elements when every weight fits 21 bits (and a parent 11)
    c.replay_pack = c.replay_cap && c.replay_cap < 1023 && max_weight < (1 << 21);
    c.replay = c.replay_cap ? (int)replay_lds_bytes(c.replay_cap, c.replay_pack) : 0;
    // the wider rows: LDS launches up to 704 runs, then up to the widest
    // row when that fits LDS
    c.replay_mid_cap = 0;
    c.replay_wide_cap = 0;
    c.replay_wide = 0;
    if (c.replay_cap && replay_wide_cap(max_qw) > c.replay_cap) {
        c.replay_mid_cap = replay_mid_cap(max_qw);
        c.replay_mid_pack = c.replay_mid_cap < 1023 && max_weight < (1 << 21);
        c.replay_mid = (int)replay_lds_bytes(c.replay_mid_cap, c.replay_mid_pack);
        const void *fm = c.replay_mid_pack ? reinterpret_cast<const void *>(&k_replay_wide<true>)
                                           : reinterpret_cast<const void *>(&k_replay_wide<false>);
        c.replay_mid_grid = (int)persistent_grid(fm, 64, (size_t)c.replay_mid, 1 << 16);
    }
    if (c.replay_mid_cap && replay_wide_cap(max_qw) > c.replay_mid_cap) {
        const void *fw = reinterpret_cast<const void *>(&k_replay_wide<false>);
        const size_t wb = replay_wide_lds_bytes(replay_wide_cap(max_qw));
        bool ok = wb <= 64 * 1024;
        if (!ok && wb <= 159 * 1024) {
            ok = hipFuncSetAttribute(fw, hipFuncAttributeMaxDynamicSharedMemorySize, (int)wb) == hipSuccess;
            (void)hipGetLastError();
        }
        if (ok) {
            c.replay_wide_cap = replay_wide_cap(max_qw);
            c.replay_wide = (int)wb;
            c.replay_wide_grid = (int)persistent_grid(fw, 64, wb, 1 << 16);
        }
    }
    c.replay_waves = force_scratch ? kReplayHbmWavesAll : kReplayHbmWaves;
    return c;
}

void launch_work_offsets(hipStream_t st, const Frame *dframes, int nframes, WorkQueue *Q, int l, int which);

void launch_ccl_stage(hipStream_t st, const Frame *dframes, const DevParams *dparams, int nframes, int max_qw,
                      int max_qh, const CclPlan &c, WorkQueue *Q, const Tune &tu, Marks *mk) {
    // per-row kernels: tu.row_waves rows (one wave each) per block, at
    // most 4: their launch bounds are 256 threads
    const int rw = std::min(std::max(tu.row_waves, 1), 4);
    const dim3 rb(64 * rw), rg((unsigned)((max_qh + rw) / rw), nframes);
    // the previous rows in LDS when rw rows of the widest frame fit 64 KiB
    // (and the LDS paths are not switched off: replay_cap 0)
    const size_t elds = (size_t)rw * (size_t)max_qw * 4;
    if (elds <= 64 * 1024 && c.replay_cap > 0)
        hipLaunchKernelGGL(k_uf_edges<true>, rg, rb, elds, st, dframes, max_qw);
    else
        hipLaunchKernelGGL(k_uf_edges<false>, rg, rb, 0, st, dframes, 0);
    mk->mark(st, "uf_edges");
    // the split front: bands of at least kBandRows rows, at most mb a frame
    // (tu.front_bands, or by the call's size: more bands for fewer frames)
    const int mb = tu.front_bands > 0 ? std::min(tu.front_bands, kMaxFrontBands)
                                      : std::max(1, std::min(kMaxFrontBands, 256 / std::max(1, nframes)));
    if (mb > 1 && c.front_band > 0) {
#define ATG_FRONT_BANDS(T)                                                                                     \
    hipLaunchKernelGGL(k_front_local<T>, dim3(mb, nframes), dim3(T), (size_t)c.front_band, st, dframes, kBandRows, mb); \
    hipLaunchKernelGGL(k_front_carry<T>, dim3(1, nframes), dim3(T), (size_t)c.front_band, st, dframes, kBandRows, mb); \
    hipLaunchKernelGGL(k_front_final<T>, dim3(mb - 1, nframes), dim3(T), (size_t)c.front_band, st, dframes, kBandRows, mb)
        switch (tu.front_threads) {
        case 64: ATG_FRONT_BANDS(64); break;
        case 128: ATG_FRONT_BANDS(128); break;
        case 512: ATG_FRONT_BANDS(512); break;
        case 1024: ATG_FRONT_BANDS(1024); break;
        default: ATG_FRONT_BANDS(256);
        }
#undef ATG_FRONT_BANDS
    } else switch (tu.front_threads) {
    case 64:
        hipLaunchKernelGGL(k_front<64>, dim3(1, nframes), dim3(64), (size_t)c.front, st, dframes);
        break;
    case 128:
        hipLaunchKernelGGL(k_front<128>, dim3(1, nframes), dim3(128), (size_t)c.front, st, dframes);
        break;
    case 512:
        hipLaunchKernelGGL(k_front<512>, dim3(1, nframes), dim3(512), (size_t)c.front, st, dframes);
        break;
    case 1024:
        hipLaunchKernelGGL(k_front<1024>, dim3(1, nframes), dim3(1024), (size_t)c.front, st, dframes);
        break;
    default:
        hipLaunchKernelGGL(k_front<256>, dim3(1, nframes), dim3(256), (size_t)c.front, st, dframes);
    }
    mk->mark(st, "uf_front.sweep");
    hipLaunchKernelGGL(k_front_dense, rg, rb, 0, st, dframes);
    mk->mark(st, "uf_front.dense");
    const dim3 rows((unsigned)std::max(1, max_qh - 1), nframes);
    if (c.replay_cap && c.replay_pack)
        hipLaunchKernelGGL(k_replay<true>, rows, dim3(64), (size_t)c.replay, st, dframes, c.replay_cap);
    else if (c.replay_cap)
        hipLaunchKernelGGL(k_replay<false>, rows, dim3(64), (size_t)c.replay, st, dframes, c.replay_cap);
    mk->mark(st, "uf_replay.rows");
    int hcap = c.replay_cap, hlcap = replay_links(c.replay_cap);
    if (c.replay_mid_cap) {
        launch_work_offsets(st, dframes, nframes, Q, 3, (int)kNWideRows);
        const int mcap = c.replay_mid_cap, mlcap = replay_links(mcap);
        if (c.replay_mid_pack)
            hipLaunchKernelGGL(k_replay_wide<true>, dim3((unsigned)c.replay_mid_grid), dim3(64), (size_t)c.replay_mid,
                               st, dframes, Q, mcap, mlcap, 0, 0);
        else
            hipLaunchKernelGGL(k_replay_wide<false>, dim3((unsigned)c.replay_mid_grid), dim3(64), (size_t)c.replay_mid,
                               st, dframes, Q, mcap, mlcap, 0, 0);
        hcap = mcap;
        hlcap = mlcap;
    }
    if (c.replay_wide_cap) {
        hipLaunchKernelGGL(k_replay_wide<false>, dim3((unsigned)c.replay_wide_grid), dim3(64), (size_t)c.replay_wide,
                           st, dframes, Q, c.replay_wide_cap, 3 * c.replay_wide_cap, hcap, hlcap);
        hcap = c.replay_wide_cap;
        hlcap = 3 * c.replay_wide_cap;
    }
    hipLaunchKernelGGL(k_replay_hbm, dim3((unsigned)c.replay_waves, nframes), dim3(64), 0, st, dframes, hcap, hlcap);
    mk->mark(st, "uf_replay.wide");
    for (int64_t d = 1; d < max_qh; d *= kRootHops + 1) hipLaunchKernelGGL(k_root_jump, rg, rb, 0, st, dframes);
    mk->mark(st, "uf_stitch.jump");
    hipLaunchKernelGGL(k_lost, rg, rb, 0, st, dframes);
    mk->mark(st, "uf_stitch.lost");
    int64_t maxnodes = (int64_t)max_qw * max_qh + max_qh;
    hipLaunchKernelGGL(k_resolve, dim3((unsigned)((maxnodes + 255) / 256), nframes), dim3(256), 0, st, dframes,
                       dparams);
    mk->mark(st, "resolve");
    hipLaunchKernelGGL(k_labels, rg, rb, 0, st, dframes);
    mk->mark(st, "labels.rows");
}

}  // namespace atg
