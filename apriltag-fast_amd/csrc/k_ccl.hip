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

struct SweepMem {
    uint16_t *par;   // [E] local union-find parent
    uint32_t *W;     // [E] set weight (pixels)
    uint32_t *R;     // [E] set root (global node id)
    uint16_t *flag;  // [E] compaction index + 1 of roots holding a current run
    uint16_t *labA;  // [QW] element of each previous-row run
    uint16_t *labB;  // [QW] element of each current-row run (next labA)
    uint32_t *edge;  // [3 QW] (a << 16 | b); reused as [2][QW] u32 temp
};

__device__ __forceinline__ uint32_t uf_find(uint16_t *par, uint32_t x) {
    while (par[x] != x) {
        uint32_t p = par[x], g = par[p];
        par[x] = (uint16_t)g;
        x = g;
    }
    return x;
}

__device__ __forceinline__ uint32_t uf_find_ro(const uint16_t *par, uint32_t x) {
    while (par[x] != x) x = par[x];
    return x;
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

__global__ __launch_bounds__(64) void k_uf_sweep(const Frame *frames, const DevParams *P, int lds_bytes,
                                                 uint8_t *gscratch, int64_t gstride) {
    extern __shared__ __align__(16) uint8_t lds[];
    const Frame F = frames[blockIdx.y];
    const int lane = threadIdx.x;
    const int QW = F.qw, QH = F.qh;
    const int E = 2 * QW + 2;
    uint8_t *base = lds_bytes > 0 ? lds : gscratch + (int64_t)blockIdx.y * gstride;
    SweepMem M;
    {
        uint8_t *p = base;
        M.W = reinterpret_cast<uint32_t *>(p); p += 4 * (size_t)E;
        M.R = reinterpret_cast<uint32_t *>(p); p += 4 * (size_t)E;
        M.edge = reinterpret_cast<uint32_t *>(p); p += 4 * (size_t)(3 * QW + 2);
        M.par = reinterpret_cast<uint16_t *>(p); p += 2 * (size_t)E;
        M.flag = reinterpret_cast<uint16_t *>(p); p += 2 * (size_t)E;
        M.labA = reinterpret_cast<uint16_t *>(p); p += 2 * (size_t)QW;
        M.labB = reinterpret_cast<uint16_t *>(p);
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
    // row 0: every run is its own set
    int m;
    {
        int n0 = (int)F.rcnt[0];
        const uint32_t *rr = F.runs;
        for (int i = lane; i < n0; i += 64) {
            uint32_t r = rr[i];
            M.labA[i] = (uint16_t)i;
            M.W[i] = run_end(r) - run_start(r) + 1;
            M.R[i] = (uint32_t)i;
            F.parent[i] = (uint32_t)i;
        }
        m = n0;
    }
    uint32_t *err = F.counters + kError;
    for (int y = 1; y < QH; y++) {
        __syncthreads();
        const uint32_t cbase = F.row_off[y];
        const int ncur = (int)F.rcnt[y], nprev = (int)F.rcnt[y - 1];
        const uint32_t *cur = F.runs + (int64_t)y * QW;
        const uint32_t *prv = F.runs + (int64_t)(y - 1) * QW;
        const int ev = m + ncur;  // the virtual node of row y-1
        // elements: set records 0..m-1, current runs m..m+ncur-1, vnode
        for (int l = lane; l < m; l += 64) M.par[l] = (uint16_t)l;
        for (int i = lane; i < ncur; i += 64) {
            uint32_t r = cur[i];
            M.par[m + i] = (uint16_t)(m + i);
            M.W[m + i] = run_end(r) - run_start(r) + 1;
            M.R[m + i] = cbase + i;
            F.parent[cbase + i] = cbase + i;
        }
        if (lane == 0) {
            M.par[ev] = (uint16_t)ev;
            M.W[ev] = 1;
            M.R[ev] = vb + y - 1;
        }
        // the edge list in the reference's order
        uint32_t ecarry = 0;
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
                // first previous run with end + 1 >= a0
                int lo = 0, hi = nprev;
                while (lo < hi) {
                    int mid = (lo + hi) >> 1;
                    if ((int)run_end(prv[mid]) + 1 < a0) lo = mid + 1; else hi = mid;
                }
                klo = lo;
                for (int k = klo; k < nprev; k++) {
                    uint32_t pk = prv[k];
                    if ((int)run_start(pk) > a1 + 1) break;
                    if ((run_white(pk) != 0) != white) continue;
                    if (runs_join(a0, a1, (int)run_start(pk), (int)run_end(pk), white)) cnt++;
                }
            }
            uint32_t inc = wave_incl_scan(cnt);
            uint32_t at = ecarry + inc - cnt;
            if (at + cnt > (uint32_t)(3 * QW)) {
                atomicOr(err, 1u);
                cnt = 0;
            }
            if (cnt) {
                for (int k = klo; k < nprev; k++) {
                    uint32_t pk = prv[k];
                    if ((int)run_start(pk) > a1 + 1) break;
                    if ((run_white(pk) != 0) != white) continue;
                    if (runs_join(a0, a1, (int)run_start(pk), (int)run_end(pk), white))
                        M.edge[at++] = ((uint32_t)(m + i) << 16) | M.labA[k];
                }
            }
            ecarry += __shfl(inc, 63, 64);
        }
        const uint32_t nedges = min(ecarry, (uint32_t)(3 * QW));
        // the last-column edge (fast/unionfind.c:93-97)
        bool lastcol = false;
        if (ncur > 0) {
            uint32_t lr = cur[ncur - 1];
            const uint8_t *trow = F.t + (int64_t)(y - 1) * QW;
            lastcol = (int)run_end(lr) == QW - 2 && run_white(lr) && trow[QW - 1] == 255 && trow[QW - 2] != 255;
        }
        __syncthreads();
        if (lane == 0) {
            uint16_t *par = M.par;
            uint32_t *W = M.W, *R = M.R;
            for (uint32_t e = 0; e <= nedges; e++) {
                uint32_t a, b;
                if (e < nedges) {
                    uint32_t ed = M.edge[e];
                    a = ed >> 16;
                    b = ed & 0xffffu;
                } else {
                    if (!lastcol) break;
                    a = (uint32_t)(m + ncur - 1);
                    b = (uint32_t)ev;
                }
                if (a > (uint32_t)ev || b > (uint32_t)ev) continue;
                uint32_t ra = uf_find(par, a), rb = uf_find(par, b);
                if (ra == rb) continue;
                uint32_t wa = W[ra], wb = W[rb];
                if (wa > wb) {
                    par[rb] = (uint16_t)ra;
                    W[ra] = wa + wb;
                    F.parent[R[rb]] = R[ra];
                } else {
                    par[ra] = (uint16_t)rb;
                    W[rb] = wa + wb;
                    F.parent[R[ra]] = R[rb];
                }
            }
        }
        __syncthreads();
        // roots that hold a current run go on; the others are final
        const int ne = ev + 1;
        for (int e = lane; e < ne; e += 64) M.flag[e] = 0;
        __syncthreads();
        for (int i = lane; i < ncur; i += 64) M.flag[uf_find_ro(M.par, m + i)] = 1;
        __syncthreads();
        uint32_t ncarry = 0;
        uint32_t *tW = M.edge, *tR = M.edge + QW + 1;
        for (int e0 = 0; e0 < ne; e0 += 64) {
            int e = e0 + lane;
            uint32_t f = e < ne ? M.flag[e] : 0u;
            uint32_t inc = wave_incl_scan(f);
            if (e < ne) {
                if (f) {
                    uint32_t j = ncarry + inc - 1;
                    M.flag[e] = (uint16_t)(j + 1);
                    tW[j] = M.W[e];
                    tR[j] = M.R[e];
                } else if (M.par[e] == e) {
                    F.fw[M.R[e]] = M.W[e];
                }
            }
            ncarry += __shfl(inc, 63, 64);
        }
        __syncthreads();
        for (int i = lane; i < ncur; i += 64) M.labB[i] = (uint16_t)(M.flag[uf_find_ro(M.par, m + i)] - 1);
        __syncthreads();
        m = (int)ncarry;
        for (int j = lane; j < m; j += 64) {
            M.W[j] = tW[j];
            M.R[j] = tR[j];
        }
        uint16_t *sw = M.labA;
        M.labA = M.labB;
        M.labB = sw;
    }
    __syncthreads();
    for (int l = lane; l < m; l += 64) F.fw[M.R[l]] = M.W[l];
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

void launch_ccl_stage(hipStream_t st, const Frame *dframes, const DevParams *dparams, int nframes, int max_qw,
                      int max_qh, uint8_t *gscratch, int64_t gstride, Marks *mk) {
    int64_t need = (int64_t)40 * max_qw + 64;
    int lds = need <= 160 * 1024 ? (int)need : 0;
    if (lds > 64 * 1024) {
        if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_uf_sweep),
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess)
            lds = 0;
    }
    hipLaunchKernelGGL(k_uf_sweep, dim3(1, nframes), dim3(64), (size_t)lds, st, dframes, dparams, lds, gscratch,
                       gstride);
    mk->mark(st, "uf_sweep");
    int64_t maxnodes = (int64_t)max_qw * max_qh + max_qh;
    hipLaunchKernelGGL(k_resolve, dim3((unsigned)((maxnodes + 255) / 256), nframes), dim3(256), 0, st, dframes,
                       dparams);
    mk->mark(st, "resolve");
    hipLaunchKernelGGL(k_labels, dim3((unsigned)((max_qh + 3) / 4), nframes), dim3(256), 0, st, dframes);
    mk->mark(st, "labels");
}

}  // namespace atg
