// Internal types of the MI355X detection path: the per-frame device
// descriptor every kernel reads, the shared parameters, and the launch
// wrappers of each kernel file.
//
// Data layout in HBM (one arena per context, one region per frame, all
// offsets 256-byte aligned; qw x qh is the quad-detection image):
//   qim     u8  [qh][qw]       decimated (blurred/sharpened) image
//   tmx/tmn u8  [th][tw]       4x4 tile max / min
//   t       u8  [qh][qw]       thresholded image (0, 127, 255)
//   runs    u32 [qh][qw]       row runs, row-strided: start | end << 15 | white << 30
//   rcnt    u32 [qh]           runs per row; row_off u32 [qh + 1]
//   parent  u32 [nodes]        union-find links: a node, or kSym | record slot
//   fw      u32 [nodes]        component pixel count, valid at final roots
//   nlab    u32 [nodes]        representative | ok << 31; before that, the
//                              record of each run within its row (rlab)
//   recw/rlost/rroot u32 [nodes]  per record slot (row_off[y] + record):
//                              weight, what its root lost to, its root
//   nrec    u32 [qh]           records (sets touching the row) per row
//   plab    u32 [qh][qw]       per-pixel nlab (ok 0 where t is 127)
//   pcnt/poff u32 [qh + 1]     points per row / their offsets
//   pts     u64 [pcap]         boundary points in emission order
//   skey/sval, skey2/sval2     radix-sort ping-pong (u64 key, u32 value)
//   cpts    u64 [pcap]         points grouped by cluster, clusters in order
//   cstart/ccount u32 [pcap]   cluster starts and sizes
//   okey    u64 [pcap]         unsorted angle keys; mom f64 [6][pcap] moments
//   quads   QuadRec [pcap/24]  fitted quads in cluster order; dets at_detection
// nodes = qw * qh + qh (runs plus one virtual last-column node per row).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "apriltag_fast.h"

namespace atg {

constexpr int kMaxFamilies = 16;
constexpr int kSortBlocks = 64;      // radix-sort blocks per frame at most
constexpr int kRadixBits = 11;       // bits of a radix-sort digit
constexpr int kRadixBins = 1 << kRadixBits;
constexpr int kInlineDets = 32;      // detections of a frame returned with its counters
// A cluster of fewer points never makes a quad (the corner search's window
// ksz = min(20, n / 12) must reach 2, fast/quads.c:789-790), so only
// clusters of at least this many points reach the fit: a frame's quads are
// at most its points / kMinFitPoints.
constexpr uint32_t kMinFitPoints = 24;
// max_nmaxima at most: the corner search holds the pair tables of this
// many maxima in LDS.
constexpr int kMaxMaxima = 32;

struct DevFamily {
    int32_t nbits, wab, tw, reversed;
    int32_t maxhamming, chunk_size, ncodes, pad;
    int64_t capacity;
    int32_t bit_x[64], bit_y[64];
    const uint32_t *offsets;  // [4][capacity + 1]
    const uint16_t *ids;      // [4][ncodes]
    const uint64_t *codes;    // [4][ncodes]
};

struct DevParams {
    float decimate, sigma;
    int32_t refine, deglitch;
    double sharpening;
    int32_t mcp, max_nmaxima, mwbd, nfam;
    float cos_crit, max_mse;
    int32_t tag_width, normal_border, reversed_border, blur_ksz;
    uint8_t blur_k[64];        // u8 blur kernel (host exp), blur_ksz taps
    double qsm_kern[7];        // f32-rounded exp(-j^2/2), j = -3..3 (host exp)
    double rot_cs[4][2];       // cos, sin of r * pi / 2 (host cos/sin)
    DevFamily fam[kMaxFamilies];
};

// A quad candidate (fast/common.h:278-282) plus its decode results.
struct QuadRec {
    float p[4][2];
    int32_t reversed;
    int32_t hok;
    double H[9];
};

// Wave-wide data movement by DPP (VALU moves, no LDS round trip; the
// controls of the GFX9 family).
// Inclusive sum over the 64 lanes: shifts of 1, 2, 4, 8 within rows of 16,
// then row 15's and lane 31's totals into the rows after them.
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15, rows 1, 3
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31, rows 2, 3
    return v;
}
// The global-memory view of a pointer into device memory (HBM). Through
// a plain pointer loaded from a frame descriptor the compiler uses flat
// instructions, which also count against the LDS wait counter: a wait for
// an LDS load then waits for the HBM accesses in flight as well.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem(T *p) {
    return (__attribute__((address_space(1))) T *)p;
}

// x of lane i - 1; lane 0 gets fill
__device__ __forceinline__ uint32_t wave_from_left(uint32_t x, uint32_t fill) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)x, 0x138, 0xF, 0xF, false);  // wave_shr:1
}
// x of lane i + 1; lane 63 gets fill
__device__ __forceinline__ uint32_t wave_from_right(uint32_t x, uint32_t fill) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)x, 0x130, 0xF, 0xF, false);  // wave_shl:1
}

// The value of lane ^ J of a wave (J a power of two below 64): a DPP quad
// permutation for 1 and 2, a swizzle over 32 lanes up to 16, a permute for
// 32.
template <int J>
__device__ __forceinline__ uint32_t lane_xor32(uint32_t x) {
    if constexpr (J == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);  // [1,0,3,2]
    else if constexpr (J == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);  // [2,3,0,1]
    else if constexpr (J < 32) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, (J << 10) | 0x1F);  // xor J
    else return (uint32_t)__shfl_xor((int)x, J, 64);
}
template <int J>
__device__ __forceinline__ uint64_t lane_xor64(uint64_t v) {
    return ((uint64_t)lane_xor32<J>((uint32_t)(v >> 32)) << 32) | lane_xor32<J>((uint32_t)v);
}

// counters[] of a frame
enum {
    kNRuns = 0, kNPts, kNClusters, kNQuads, kNDets, kError,
    kNEdges,    // union-find links (for the byte counts of the measurements)
    kNKeptPts,  // points of the clusters that go on to the quad fit
    kNLarge = 8,  // kept clusters above kLargeCluster points, at most kHugeCluster
    kNSmall = 9,  // kept clusters of at most kLargeCluster points
    kSortBits = 11,  // bits of the cluster sort key: the cluster ranks' bits
    kNWideRows = 12,  // rows the first replay launch left (over its run or link capacity), listed in rwide
    kNHuge = 13,  // kept clusters above kHugeCluster points
    kNCounters = 16
};

struct Frame {
    const uint8_t *img;
    int64_t stride;
    int32_t W, H;
    int32_t qw, qh, tw, th;
    uint64_t bucket_mask; // cluster hash buckets - 1 (fast/clusters.c:97-103)
    int32_t max_px;       // 2 * (2 qw + 2 qh) (fast/quads.c:1150)
    int32_t blurred;      // the quad image was blurred or sharpened
    int32_t skip;         // a frame the reference returns nothing for: an 8x8 zero image
    int64_t pcap;         // point capacity
    uint8_t *qim, *t, *tmx, *tmn, *tbuf;
    uint8_t *fscr;        // union-find front state in HBM (frames too wide for LDS), else null
    uint8_t *rscr;        // union-find replay state of rows too wide for LDS: [kReplayHbmWaves][..]
    uint32_t *runs, *rcnt, *row_off;
    uint32_t *parent, *fw, *nlab, *plab;
    uint32_t *rlab;           // = nlab until k_resolve: record of each run within its row
    uint32_t *recw, *rlost, *rroot, *nrec;
    uint32_t *uedge, *ecnt;   // [qh][3 qw] union-find links per row, [qh] their counts
    uint32_t *pcnt, *poff;
    uint64_t *pts, *cpts;
    uint64_t *skey, *skey2;
    uint32_t *sval, *sval2;
    uint32_t *rhist;      // [kSortBlocks][kRadixBins]
    uint64_t *pw;         // [pcap] kept clusters' points: x | y << 16 | squared gradient << 32,
                          // in cluster order (the memory of pts, free by then)
    uint32_t *rwide;      // [qh] rows for the replay's wide launch
    // the frame's cluster table (open addressing, hmask + 1 slots): each
    // cluster's id (kSlotEmpty where none), point count and rank; the
    // slots in insertion order ([hmask + 1) / 2]); the sort's scratch;
    // ccap clusters at most (more: error flag 128)
    uint32_t hmask, ccap;
    uint64_t *hkey;
    uint32_t *hcnt, *hrank, *clist;
    uint64_t *csort;
    uint32_t *cstart, *ccount, *cflag;
    uint32_t *klist;      // [2][kcap] kept clusters: the huge ones, then the large ones; at kcap the small ones
    uint32_t kcap;
    uint64_t *cid;        // [pcap] cluster ids (max_root << 32 | min_root), in order
    uint64_t *okey;       // unsorted angle keys, by cluster position
    double *mom;          // [6][pcap] cumulative moments in angle order
    double *errs, *filt;  // [pcap] window errors, filtered
    uint32_t *midx;       // [pcap] maxima positions
    uint32_t qcap;        // quad capacity: pcap / kMinFitPoints + 1, more than the frame can have
    QuadRec *qtmp;        // [qcap] fitted quads, any order
    uint32_t *qcl;        // [qcap] their cluster index
    QuadRec *quads;       // [qcap] in cluster order
    double *qline;        // [qcap][4][4] per quad edge: refined line (x, y, cos, sin);
                          // before k_lines: mean x, mean y and atan2's two arguments
    at_detection *dtmp;   // [qcap * nfam] per (quad, family)
    at_detection *dfin;   // [qcap * nfam] the final list, sorted by id
    uint32_t *dflag;      // [qcap * nfam]
    at_detection *inl;    // [kInlineDets] the final list's first entries, beside the counters
    uint32_t *counters;
};

// The work of the persistent kernels, across the frames of a call: list 0
// the kept clusters of kLargeCluster to kHugeCluster points, list 1 the
// smaller kept clusters, list 2 the quads, list 3 the rows of the
// union-find replay's wide launch, list 4 the kept clusters above
// kHugeCluster points. Each frame lists its own items (klist, rwide);
// off[l][f] is where frame f's items start in list l's global index,
// total[l] its size. Blocks take global indices in strides, without
// atomics; the kernels of one cluster a block go over the huge clusters of
// every frame first, then list 0, so that the longest items start early.
constexpr uint32_t kLargeCluster = 512;
constexpr uint32_t kHugeCluster = 1024;
constexpr int kLists = 5;
struct WorkQueue {
    uint32_t total[kLists];
    uint32_t nframes, pad[2];
    uint32_t *off[kLists];   // [nframes + 1] each
};

// The frame of entry g of a work list: the last f with off[f] <= g (off[]
// ascending, off[nframes] = the list's total > g), by binary search from a
// frame lo with off[lo] <= g (a block's entries ascend, so its last frame
// is a valid start). A persistent block's stride crosses many frames at
// once: a search costs log2(nframes) reads where a walk costs one a frame.
__device__ __forceinline__ uint32_t frame_of(const uint32_t *off, uint32_t nframes, uint32_t g, uint32_t lo) {
    uint32_t hi = nframes;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// frame_of for an entry g the same in every lane of the wave (every lane
// takes part): a 64-way search, each lane reading one candidate frame's
// offset and a ballot keeping the last at or below g, two rounds of reads
// for up to 4096 frames where the binary search takes log2(nframes).
__device__ __forceinline__ uint32_t frame_of_wave(const uint32_t *off, uint32_t nframes, uint32_t g, uint32_t lo) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t hi = nframes;  // off[lo] <= g < off[hi]
    while (hi - lo > 1) {
        const uint32_t step = (hi - lo + 63u) >> 6;
        const uint32_t idx = lo + lane * step;
        const uint64_t at_or_below = __ballot(idx < hi && off[idx] <= g);  // a prefix of the lanes, lane 0 set
        lo += (uint32_t)(63 - __clzll((long long)at_or_below)) * step;
        hi = min(hi, lo + step);
    }
    return lo;
}

// Work splits of the kernel chain (the autotuner's choices); any values
// give the same results.
struct Tune {
    int32_t thr_tile_rows;  // tile rows per k_threshold block: 1, 2, 4
    int32_t row_waves;      // rows (one wave each) per block of the per-row kernels: 1, 2, 4
                            // (their launch bounds are 256 threads)
    int32_t front_threads;  // k_front block: 128, 256, 512
    int32_t fit_blocks;     // blocks of the persistent fit and decode kernels: 32..4096
    int32_t front_bands;    // row bands of a frame's union-find front at most (1: one sweep; 0: by the
                            // call's size)
};

// The replay kernels' memory plan for one call (LDS where it fits, else
// HBM scratch).
struct CclPlan {
    int front;                  // LDS bytes of the front (frames without fscr)
    int front_band;             // LDS bytes of the split front's kernels, 0 = not used
    int replay;                 // LDS bytes of the replay
    int replay_cap;             // runs per row replayed in LDS, 0 = none
    bool replay_pack;           // its elements packed (parent, weight) in one word
    int replay_mid, replay_mid_cap;    // the second LDS launch (listed rows up to 704 runs): bytes, runs; 0 = none
    bool replay_mid_pack;       // its elements packed
    int replay_mid_grid;        // its persistent grid
    int replay_wide, replay_wide_cap;  // the third (the wider rows, up to the widest): bytes, runs; 0 = none
    int replay_wide_grid;       // its persistent grid
    int replay_waves;           // waves per frame of k_replay_hbm
};

// HBM scratch of one frame's union-find, by its own width: the front's
// when it does not fit LDS (or forced), the replay's for wide rows.
size_t front_scratch_bytes(int qw, bool force);
size_t replay_scratch_bytes(int qw, bool force);

// The grid of a persistent kernel: the blocks of `fn` resident at once on
// the current device (occupancy x CUs), at most `want`, so that no block
// starts only when another has done its whole share. Looked up once per
// kernel and block shape (warm_persistent_grids does so before any stream
// capture).
unsigned persistent_grid(const void *fn, int threads, size_t lds, int want);
void warm_persistent_grids(int fit_blocks);

// Optional timing marks between the kernels of one call (consecutive
// marks on one stream bracket the kernels launched between them). Level 1:
// a one-lane kernel writes the device wall clock into stamps[i]; being a
// kernel, the mark is captured in the call's HIP graph like any other
// launch. Level 2: HIP events (the call then runs without a graph).
constexpr int kMaxMarks = 48;
void launch_stamp(hipStream_t st, uint64_t *slot);
// Streams for chain work that is independent of the main stream's:
// begin() forks kForks streams off the main stream, end() joins them
// back. Under stream capture they become parallel branches of the graph.
struct Fork {
    static constexpr int kForks = 2;
    hipStream_t s[kForks] = {};
    hipEvent_t go = nullptr, done[kForks] = {};
    bool on = false;
    // the streams to use: the forked ones, or the main one when off
    void begin(hipStream_t st, hipStream_t (&out)[kForks]) {
        for (int k = 0; k < kForks; k++) out[k] = on ? s[k] : st;
        if (!on) return;
        hipEventRecord(go, st);
        for (int k = 0; k < kForks; k++) hipStreamWaitEvent(s[k], go, 0);
    }
    void end(hipStream_t st) {
        if (!on) return;
        for (int k = 0; k < kForks; k++) {
            hipEventRecord(done[k], s[k]);
            hipStreamWaitEvent(st, done[k], 0);
        }
    }
};

struct Marks {
    uint64_t *stamps = nullptr;  // device [kMaxMarks]
    hipEvent_t ev[kMaxMarks] = {};
    const char *name[kMaxMarks];
    int n = 0;
    int level = 0;
    void mark(hipStream_t st, const char *what) {
        if (!level || n >= kMaxMarks) return;
        name[n] = what;
        if (level == 2)
            hipEventRecord(ev[n++], st);
        else
            launch_stamp(st, stamps + n++);
    }
};

}  // namespace atg
