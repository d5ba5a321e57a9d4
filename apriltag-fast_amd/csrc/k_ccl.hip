// Stage (b): connected components over row runs, with the reference's
// exact union-find representatives, and the per-pixel labels of (c).
//
// The reference links sets in a fixed row-major edge order with union by
// size (fast/unionfind.c:15-45, 112-146, 223-237): the root of a merged
// set is the root of the heavier part, ties to the previous-row side.
// Those roots become cluster ids, and cluster ids decide the cluster
// order, so a min-label GPU CCL (right partition, other roots) is not
// enough. The outcome of each link depends only on the two set weights at
// that moment, and a set that does not reach row y can never be linked
// again. So the replay keeps only the sets that touch the previous row:
//
//   k_uf_sweep   one wave per frame walks the rows in order. Per row, the
//                lanes build the row's edge list in the reference's order
//                (cur run ascending, then previous run ascending, then the
//                last-column edge) into LDS; lane 0 links them over a small
//                union-find of set records (weight, root node) held in LDS;
//                the lanes then relabel the current runs and retire the
//                sets that ended. Every link that makes a root lose writes
//                parent[loser] = winner to HBM.
//   k_resolve    one thread per node follows parent[] to the final root;
//                nlab = root | (pixel count >= min_cluster_pixels) << 31
//                (fast/clusters.c:73-94).
//   k_labels     one wave per row: the nlab of the run under each pixel,
//                the run index counted with ballots over the row's changes.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "atg_internal.h"

namespace atg {

__device__ __forceinline__ uint32_t run_start(uint32_t r) { return r & 0x7fffu; }
__device__ __forceinline__ uint32_t run_end(uint32_t r) { return (r >> 15) & 0x7fffu; }
__device__ __forceinline__ uint32_t run_white(uint32_t r) { return (r >> 30) & 1u; }

// inclusive wave scan of v (64 lanes)
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Whether current run a and previous run b of the same value are linked
// (fast/unionfind.c:130-141): overlap on columns >= 1; white also touches
// diagonally.
__device__ __forceinline__ bool runs_join(int a0, int a1, int b0, int b1, bool white) {
    if (max(max(a0, b0), 1) <= min(a1, b1)) return true;
    if (!white) return false;
    if (max(max(a0, b0 + 1), 1) <= min(a1, b1 + 1)) return true;
    return max(max(a0, b0 - 1), 1) <= min(a1, b1 - 1);
}

// The links of every row, in the reference's order (current run
// ascending, then previous run ascending): one wave per row, all rows at
// once. uedge[y][e] = i << 16 | k (current run i of row y, previous run k).
__global__ __launch_bounds__(256) void k_uf_edges(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    const int lane = threadIdx.x & 63;
    int y = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (y < 1 || y >= F.qh) return;
    const int QW = F.qw;
    const int ncur = (int)F.rcnt[y], nprev = (int)F.rcnt[y - 1];
    const uint32_t *cur = F.runs + (int64_t)y * QW, *prv = F.runs + (int64_t)(y - 1) * QW;
    uint32_t *out = F.uedge + (int64_t)y * 3 * QW;
    const uint32_t cap = 3u * (uint32_t)QW;
    uint32_t carry = 0;
    for (int i0 = 0; i0 < ncur; i0 += 64) {
        int i = i0 + lane;
        int a0 = 0, a1 = -1, klo = nprev;
        bool white = false;
        uint32_t cnt = 0;
        if (i < ncur) {
            uint32_t r = cur[i];
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
        carry += __shfl(inc, 63, 64);
    }
    if (lane == 0) {
        if (carry > cap) atomicOr(F.counters + kError, 1u);
        F.ecnt[y] = min(carry, cap);
    }
}

// The sweep's state, in LDS (or, for very wide frames, in global scratch
// accessed through volatile pointers, so that lanes see lane 0's writes).
template <bool G>
struct SweepMem {
    using U32 = typename std::conditional<G, volatile uint32_t, uint32_t>::type;
    using U16 = typename std::conditional<G, volatile uint16_t, uint16_t>::type;
    U32 *ent;   // [E][4]: union-find parent, weight (pixels), root node, unused
    U32 *deg;   // [QW] links of each current run
    U32 *lcnt;  // [QW] links into each set record
    U16 *flag;  // [E] (over deg/lcnt) compaction index + 1 of roots holding a current run
    U32 *le;    // [3 QW] links as (i << 16 | record), bit 31 once resolved; then [2][QW + 1] new records
    U16 *labA;  // [QW] record of each previous-row run
    U16 *labB;  // [QW] record of each current-row run
};

// An element's entry in one load: x = parent, y = weight, z = root node.
template <bool G>
__device__ __forceinline__ uint4 ld_ent(SweepMem<G> &M, uint32_t x) {
    if constexpr (G) {
        uint4 v;
        v.x = M.ent[4 * x];
        v.y = M.ent[4 * x + 1];
        v.z = M.ent[4 * x + 2];
        v.w = 0;
        return v;
    } else {
        return *reinterpret_cast<const uint4 *>(M.ent + 4 * x);
    }
}

// Find with path halving; returns the root's entry with .x = the root.
template <bool G>
__device__ __forceinline__ uint4 sw_find(SweepMem<G> &M, uint32_t x) {
    for (;;) {
        uint4 e = ld_ent(M, x);
        if (e.x == x) return e;
        uint4 pe = ld_ent(M, e.x);
        if (pe.x == e.x) return pe;
        M.ent[4 * x] = pe.x;
        x = pe.x;
    }
}

template <bool G>
__device__ __forceinline__ uint32_t sw_find_ro(SweepMem<G> &M, uint32_t x) {
    for (uint32_t p; (p = M.ent[4 * x]) != x;) x = p;
    return x;
}

// A barrier of the one-wave sweep: LDS operations of a wave execute in
// order, so only the compiler must not move memory operations across it;
// the global-scratch variant needs the full barrier.
template <bool G>
__device__ __forceinline__ void sw_sync() {
    if constexpr (G) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// One wave per frame walks the rows in order. Per row, the set records of
// the previous row and the current runs are union-find elements; a link
// where one current run meets one record and nothing else links either is
// independent of all others and all lanes resolve those at once; lane 0
// resolves the rest in the reference's order; then the roots that hold
// current runs become the next row's records and the others are final.
template <bool G>
__device__ void uf_sweep(const Frame &F, uint8_t *base) {
    const int lane = threadIdx.x;
    const int QW = F.qw, QH = F.qh;
    const int E = 2 * QW + 2;
    SweepMem<G> M;
    {
        using U32 = typename SweepMem<G>::U32;
        using U16 = typename SweepMem<G>::U16;
        uint8_t *p = base;
        M.ent = reinterpret_cast<U32 *>(p); p += 16 * (size_t)E;
        M.deg = reinterpret_cast<U32 *>(p);
        M.flag = reinterpret_cast<U16 *>(p); p += 4 * (size_t)QW;
        M.lcnt = reinterpret_cast<U32 *>(p); p += 4 * (size_t)QW + 8;
        M.le = reinterpret_cast<U32 *>(p); p += 12 * (size_t)QW;
        M.labA = reinterpret_cast<U16 *>(p); p += 2 * (size_t)QW;
        M.labB = reinterpret_cast<U16 *>(p);
    }
    // row offsets: exclusive scan of the per-row run counts
    uint32_t carry = 0;
    for (int y0 = 0; y0 < QH; y0 += 64) {
        int y = y0 + lane;
        uint32_t c = y < QH ? F.rcnt[y] : 0u;
        uint32_t inc = wave_incl_scan(c);
        if (y < QH) F.row_off[y] = carry + inc - c;
        carry += __shfl(inc, 63, 64);
    }
    const uint32_t vb = carry;  // the virtual last-column nodes start here
    if (lane == 0) {
        F.row_off[QH] = vb;
        F.counters[kNRuns] = vb;
    }
    for (int y = lane; y < QH; y += 64) {
        F.parent[vb + y] = vb + y;
        F.fw[vb + y] = 1;
    }
    __syncthreads();  // row_off is read back below
    int m;
    {
        int n0 = (int)F.rcnt[0];
        for (int i = lane; i < n0; i += 64) {
            uint32_t r = F.runs[i];
            M.labA[i] = (uint16_t)i;
            M.ent[4 * i + 1] = run_end(r) - run_start(r) + 1;
            M.ent[4 * i + 2] = (uint32_t)i;
            F.parent[i] = (uint32_t)i;
        }
        m = n0;
    }
    uint64_t cyc[6] = {0, 0, 0, 0, 0, 0};
    uint64_t tc = clock64();
    auto tick = [&](int k) {
        uint64_t n = clock64();
        cyc[k] += n - tc;
        tc = n;
    };
    for (int y = 1; y < QH; y++) {
        sw_sync<G>();
        const uint32_t cbase = F.row_off[y];
        const int ncur = (int)F.rcnt[y];
        const int ne = (int)F.ecnt[y];
        const uint32_t *cur = F.runs + (int64_t)y * QW;
        const uint32_t *edges = F.uedge + (int64_t)y * 3 * QW;
        const int ev = m + ncur;  // the virtual node of row y-1
        for (int l = lane; l < m; l += 64) {
            M.ent[4 * l] = (uint32_t)l;
            M.lcnt[l] = 0;
        }
        for (int i = lane; i < ncur; i += 64) {
            uint32_t r = cur[i];
            M.ent[4 * (m + i)] = (uint32_t)(m + i);
            M.ent[4 * (m + i) + 1] = run_end(r) - run_start(r) + 1;
            M.ent[4 * (m + i) + 2] = cbase + i;
            M.deg[i] = 0;
            F.parent[cbase + i] = cbase + i;
        }
        if (lane == 0) {
            M.ent[4 * ev] = (uint32_t)ev;
            M.ent[4 * ev + 1] = 1;
            M.ent[4 * ev + 2] = vb + y - 1;
        }
        sw_sync<G>();
        tick(0);
        // the links in terms of set records, and how many touch each side
        for (int e = lane; e < ne; e += 64) {
            uint32_t ed = edges[e];
            uint32_t i = ed >> 16, l = M.labA[ed & 0xffffu];
            M.le[e] = (i << 16) | l;
            atomicAdd((uint32_t *)&M.deg[i], 1u);
            atomicAdd((uint32_t *)&M.lcnt[l], 1u);
        }
        sw_sync<G>();
        tick(1);
        // one current run meeting one record, nothing else linking either:
        // an independent link (fast/unionfind.c:36-44)
        for (int e = lane; e < ne; e += 64) {
            uint32_t ed = M.le[e];
            uint32_t i = ed >> 16, l = ed & 0xffffu;
            if (M.deg[i] != 1u || M.lcnt[l] != 1u) continue;
            uint32_t a = (uint32_t)m + i;
            uint32_t wa = M.ent[4 * a + 1], wb = M.ent[4 * l + 1];
            uint32_t Ra = M.ent[4 * a + 2], Rb = M.ent[4 * l + 2];
            if (wa > wb) {
                M.ent[4 * l] = a;
                M.ent[4 * a + 1] = wa + wb;
                F.parent[Rb] = Ra;
            } else {
                M.ent[4 * a] = l;
                M.ent[4 * l + 1] = wa + wb;
                F.parent[Ra] = Rb;
            }
            M.le[e] = ed | 0x80000000u;
        }
        bool lastcol = false;
        if (ncur > 0) {
            uint32_t lr = cur[ncur - 1];
            const uint8_t *trow = F.t + (int64_t)(y - 1) * QW;
            lastcol = (int)run_end(lr) == QW - 2 && run_white(lr) && trow[QW - 1] == 255 && trow[QW - 2] != 255;
        }
        sw_sync<G>();
        tick(2);
        // the other links, in order, on lane 0; the current run's set is
        // kept in registers across its links
        if (lane == 0) {
            int cur_i = -1;
            uint32_t ra = 0, wa = 0, Ra = 0;
            for (int e0 = 0; e0 < ne; e0 += 8) {
                uint32_t eb[8];
#pragma unroll
                for (int q = 0; q < 8; q++) eb[q] = e0 + q < ne ? M.le[e0 + q] : 0x80000000u;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    uint32_t ed = eb[q];
                    if (ed & 0x80000000u) continue;
                    uint32_t i = ed >> 16, l = ed & 0xffffu;
                    if ((int)i != cur_i) {
                        if (cur_i >= 0) M.ent[4 * ra + 1] = wa;
                        cur_i = (int)i;
                        ra = (uint32_t)m + i;  // a current run's first link: a fresh root
                        uint4 ea = ld_ent(M, ra);
                        wa = ea.y;
                        Ra = ea.z;
                    }
                    uint4 eb4 = sw_find(M, l);
                    uint32_t rb = eb4.x;
                    if (rb == ra) continue;
                    uint32_t wb = eb4.y, Rb = eb4.z;
                    if (wa > wb) {
                        M.ent[4 * rb] = ra;
                        F.parent[Rb] = Ra;
                    } else {
                        M.ent[4 * ra] = rb;
                        F.parent[Ra] = Rb;
                        ra = rb;
                        Ra = Rb;
                    }
                    wa += wb;
                }
            }
            if (cur_i >= 0) M.ent[4 * ra + 1] = wa;
            // the last-column link (fast/unionfind.c:93-97), last in order
            if (lastcol) {
                uint4 ea = sw_find(M, (uint32_t)(m + ncur - 1));
                uint32_t a = ea.x, w1 = ea.y, R1 = ea.z;
                if (w1 > 1u) {
                    M.ent[4 * ev] = a;
                    M.ent[4 * a + 1] = w1 + 1;
                    F.parent[vb + y - 1] = R1;
                } else {
                    M.ent[4 * a] = (uint32_t)ev;
                    M.ent[4 * ev + 1] = w1 + 1;
                    F.parent[R1] = vb + y - 1;
                }
            }
        }
        sw_sync<G>();
        // roots that hold a current run go on; the others are final
        const int nel = ev + 1;
        for (int e = lane; e < nel; e += 64) M.flag[e] = 0;
        sw_sync<G>();
        for (int i = lane; i < ncur; i += 64) M.flag[sw_find_ro(M, (uint32_t)(m + i))] = 1;
        sw_sync<G>();
        uint32_t ncarry = 0;
        auto *tW = M.le;
        auto *tR = M.le + QW + 1;
        for (int e0 = 0; e0 < nel; e0 += 64) {
            int e = e0 + lane;
            uint32_t f = e < nel ? (uint32_t)M.flag[e] : 0u;
            uint32_t inc = wave_incl_scan(f);
            if (e < nel) {
                if (f) {
                    uint32_t j = ncarry + inc - 1;
                    M.flag[e] = (uint16_t)(j + 1);
                    tW[j] = M.ent[4 * e + 1];
                    tR[j] = M.ent[4 * e + 2];
                } else if (M.ent[4 * e] == (uint32_t)e) {
                    F.fw[M.ent[4 * e + 2]] = M.ent[4 * e + 1];
                }
            }
            ncarry += __shfl(inc, 63, 64);
        }
        sw_sync<G>();
        tick(4);
        for (int i = lane; i < ncur; i += 64)
            M.labB[i] = (uint16_t)(M.flag[sw_find_ro(M, (uint32_t)(m + i))] - 1);
        sw_sync<G>();
        m = (int)ncarry;
        for (int j = lane; j < m; j += 64) {
            M.ent[4 * j + 1] = tW[j];
            M.ent[4 * j + 2] = tR[j];
        }
        auto *sw = M.labA;
        M.labA = M.labB;
        M.labB = sw;
        tick(5);
    }
    sw_sync<G>();
    for (int l = lane; l < m; l += 64) F.fw[M.ent[4 * l + 2]] = M.ent[4 * l + 1];
    // cycles per phase (lane 0's clock), for tuning: counters[8..13]
    if (lane == 0)
        for (int k = 0; k < 6; k++) F.counters[8 + k] = (uint32_t)min(cyc[k] >> 10, (uint64_t)0xffffffffu);
}

__global__ __launch_bounds__(64) void k_uf_sweep(const Frame *frames, int lds_bytes, uint8_t *gscratch,
                                                 int64_t gstride) {
    extern __shared__ __align__(16) uint8_t lds[];
    const Frame F = frames[blockIdx.y];
    if (lds_bytes > 0)
        uf_sweep<false>(F, lds);
    else
        uf_sweep<true>(F, gscratch + (int64_t)blockIdx.y * gstride);
}

// One thread per node: the final root and the ok flag.
__global__ void k_resolve(const Frame *frames, const DevParams *P) {
    const Frame F = frames[blockIdx.y];
    int64_t nn = (int64_t)F.counters[kNRuns] + F.qh;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    uint32_t r = (uint32_t)i;
    // every link points to a node that was a root when it was made, so the
    // walk ends; the bound only guards against a corrupted table
    int64_t steps = 0;
    for (uint32_t p; (p = F.parent[r]) != r && steps <= nn; steps++) r = p;
    if (steps > nn) atomicOr(F.counters + kError, 32u);
    uint32_t ok = (int64_t)F.fw[r] >= (int64_t)P->mcp ? 1u : 0u;
    F.nlab[i] = r | (ok << 31);
}

// One wave per row: plab[y][x] = nlab of the run under pixel x (0 where
// t is 127), the row's virtual node at x = qw-1 (fast/clusters.c:123-142).
__global__ __launch_bounds__(256) void k_labels(const Frame *frames) {
    const Frame F = frames[blockIdx.y];
    int y = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (y >= F.qh) return;
    const int lane = threadIdx.x & 63;
    const uint8_t *trow = F.t + (int64_t)y * F.qw;
    uint32_t *out = F.plab + (int64_t)y * F.qw;
    const uint32_t *nl = F.nlab + F.row_off[y];
    const int xmax = F.qw - 2;
    int count = 0;
    uint32_t carry_t = 0;
    for (int x0 = 0; x0 <= xmax; x0 += 64) {
        int x = x0 + lane;
        bool valid = x <= xmax;
        uint32_t tv = valid ? trow[x] : 0u;
        uint32_t tp = __shfl_up(tv, 1, 64);
        if (lane == 0) tp = carry_t;
        bool st = valid && (x == 0 || tv != tp) && tv != 127u;
        uint64_t sm = __ballot(st);
        uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
        if (valid) {
            int idx = count + __popcll(sm & le) - 1;
            out[x] = tv == 127u ? 0u : nl[idx];
        }
        count += __popcll(sm);
        carry_t = __shfl(tv, 63, 64);
    }
    if (lane == 0) out[F.qw - 1] = F.nlab[F.row_off[F.qh] + y];
}

size_t sweep_lds_bytes(int max_qw) { return (size_t)56 * max_qw + 64; }

void launch_ccl_stage(hipStream_t st, const Frame *dframes, const DevParams *dparams, int nframes, int max_qw,
                      int max_qh, uint8_t *gscratch, int64_t gstride, Marks *mk) {
    hipLaunchKernelGGL(k_uf_edges, dim3((unsigned)((max_qh + 3) / 4), nframes), dim3(256), 0, st, dframes);
    mk->mark(st, "uf_edges");
    // LDS when it fits (above 64 KB by the per-kernel limit, up to gfx950's
    // 160 KB), else the per-frame scratch in HBM
    size_t need = sweep_lds_bytes(max_qw);
    int lds = (int)need;
    if (need > 64 * 1024) {
        static int big_ok = -1;
        if (big_ok < 0)
            big_ok = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_uf_sweep),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) == hipSuccess;
        if (!big_ok || need > 160 * 1024 || gstride == 0) lds = 0;
    }
    if (lds == 0 && gstride == 0) return;  // cannot happen: the caller sizes the scratch
    hipLaunchKernelGGL(k_uf_sweep, dim3(1, nframes), dim3(64), (size_t)lds, st, dframes, lds, gscratch, gstride);
    mk->mark(st, "uf_sweep");
    int64_t maxnodes = (int64_t)max_qw * max_qh + max_qh;
    hipLaunchKernelGGL(k_resolve, dim3((unsigned)((maxnodes + 255) / 256), nframes), dim3(256), 0, st, dframes,
                       dparams);
    mk->mark(st, "resolve");
    hipLaunchKernelGGL(k_labels, dim3((unsigned)((max_qh + 3) / 4), nframes), dim3(256), 0, st, dframes);
    mk->mark(st, "labels");
}

}  // namespace atg
