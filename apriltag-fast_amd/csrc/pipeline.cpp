// Host side of the atg C-ABI: the context, the HBM arena and its per-frame
// layout, the decode tables, the kernel chain of one (batched) call and
// the readback of the detections.
//
// One call = one pass of the chain over all its frames: every kernel is
// launched once with blockIdx.y = frame, so a batch of uneven frames costs
// the same number of launches as one frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "atg.h"
#include "atg_internal.h"
#include "chain.h"

namespace atg {
void launch_init(hipStream_t st, const Frame *dframes, int nframes, WorkQueue *Q);
void launch_copy(hipStream_t st, void *dst, const void *src, int64_t bytes);
CclPlan ccl_plan(int max_lds_qw, int max_qw, int64_t max_weight, bool force_scratch, int nframes);
void launch_image_stage(hipStream_t st, const Frame *dframes, const DevParams *dparams, int nframes, int max_qw,
                        int max_qh, bool blur, bool deglitch, const Tune &tu, Marks *mk);
void launch_ccl_stage(hipStream_t st, const Frame *dframes, const DevParams *dparams, int nframes, int max_qw,
                      int max_qh, const CclPlan &c, WorkQueue *Q, const Tune &tu, Marks *mk);
void launch_points_stage(hipStream_t st, const Frame *dframes, int nframes, int max_qh, int key_bits, const Tune &tu,
                         Marks *mk);
void launch_quads_stage(hipStream_t st, const Frame *dframes, const DevParams *dparams, int nframes, int cl_bits,
                        int fit_blocks, int sort_cap, bool wave_corners, WorkQueue *Q, Marks *mk, Fork *fk);
int angle_sort_cap(int max_cluster, bool force_global);
void launch_decode_stage(hipStream_t st, const Frame *dframes, const DevParams *dparams, int nframes, int blocks,
                         bool refine, WorkQueue *Q, Marks *mk);
void launch_debug_finish(hipStream_t st, uint8_t *buf, uint32_t n, int64_t *nout);
void launch_debug_trig(hipStream_t st, const float *y, const float *x, const float *th, int64_t n, float *at,
                       float *co, float *si, int slow);
hipError_t launch_pose(hipStream_t st, const atg_pose_in *in, int64_t n, int iters, int flags, atg_pose_out *out);
void prepare_points_stage();
}  // namespace atg

using namespace atg;

namespace {

constexpr size_t kCountersBytes = 64;
// a frame's result block: its counters and the first detections (the rest
// stay in its arena region)
constexpr size_t kResultBlock = kCountersBytes + (size_t)kInlineDets * sizeof(at_detection);
// fit_blocks 2048: the one-wave fit kernels that fit more than 16 waves a CU
// (the small-cluster angle sort, edge refinement) get them all resident
constexpr Tune kDefaultTune = {1, 4, 1024, 2048, 0};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int ceil_log2(uint64_t v) {
    int b = 0;
    while (((uint64_t)1 << b) < v) b++;
    return b;
}

// fast/image.c:6-20
void decimated_size(int64_t w, int64_t h, float f, int64_t *ow, int64_t *oh) {
    if (!(f > 1.0f)) { *ow = w; *oh = h; return; }
    if (f == 1.5f) { *ow = w / 3 * 2; *oh = h / 3 * 2; return; }
    int64_t k = (int64_t)f;
    if (k == 2) { *ow = w / 2; *oh = h / 2; return; }
    *ow = 1 + (w - 1) / k;
    *oh = 1 + (h - 1) / k;
}

// fast/clusters.c:97-103 with one thread
int64_t cluster_map_size(int64_t w, int64_t h) {
    int64_t chunk = 1 + (h - 1) / 6, n = 1024;
    while (n < chunk * w / 64 && n < 65536) n <<= 1;
    return n;
}

// The persistent kernels' work lists of a call (in the arena, after the
// frames' regions): the header and an offset array over the frames per list.
size_t queue_bytes(int64_t nframes) {
    return align256(sizeof(WorkQueue)) + kLists * align256((size_t)(nframes + 1) * 4);
}

// Kept clusters need at least a few points each, so a quarter of the point
// capacity bounds a frame's list of them.
int64_t cluster_cap(int64_t pcap) { return pcap / 4 + 64; }

// Slots of a frame's cluster table (a power of two, filled at most half):
// first, for quad images of up to kOnePassPixels, kOnePassSlots (a
// 960x540 image of the workload has about a thousand clusters: room for
// 2047, whose ranks take one radix pass), for larger ones a cluster per
// 32 points; after an overflow (error flag 128) for every point a cluster
// of its own.
constexpr int64_t kOnePassPixels = 960 * 540, kOnePassSlots = 4096;
bool one_pass_table(int64_t px, bool full) { return !full && px <= kOnePassPixels; }
int64_t hash_slots(int64_t pcap, int64_t px, bool full) {
    if (one_pass_table(px, full)) return kOnePassSlots;
    const int64_t want = full ? 2 * pcap + 2 : std::max<int64_t>(2048, pcap / 16);
    int64_t n = 1;
    while (n < want) n <<= 1;
    return n;
}

struct FramePlan {
    bool skip;            // the reference returns nothing for this frame
    bool hfull;           // its cluster table sized for the worst case
    int64_t W, H, stride;
    int64_t qw, qh;
    int64_t pcap;
    bool force;           // the HBM paths of wide frames for every frame
    size_t off, size;     // arena region
    size_t in_off;        // staged input (host frames), within the region
};

}  // namespace

struct atg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    // families and decode tables
    std::vector<at_family> fams;
    std::vector<int32_t> bits;
    uint8_t *dtables = nullptr;
    size_t tab_off[kMaxFamilies][3] = {};  // per family: offsets, ids, codes within dtables
    DevParams hp{};
    DevParams *dparams = nullptr;
    bool params_valid = false;
    atg_params last_params{};
    // arena
    uint8_t *arena = nullptr;
    size_t arena_size = 0;
    uint8_t *results = nullptr;       // result blocks, contiguous
    size_t results_size = 0;
    Frame *dframes = nullptr;
    size_t dframes_cap = 0;
    uint8_t *pinned = nullptr;        // host staging: frames in
    size_t pinned_size = 0;
    uint8_t *pinned_out = nullptr;    // host staging: result prefixes out
    size_t pinned_out_size = 0;
    // last call
    std::vector<Frame> hframes;
    std::vector<FramePlan> plans;
    std::vector<std::vector<at_detection>> dets;
    std::vector<int64_t> nquads;
    double chain_ms = 0.0;
    int64_t pcap_scale = 1;           // points per quad-image pixel
    bool hash_full = false;           // cluster tables sized for the worst case
    bool need_points = false, need_table = false;  // the last call ran out of point or table room
    Marks marks;                      // per-stage clock stamps, when profiling
    uint64_t *hstamps = nullptr;      // pinned host copy of the stamps
    double clock_khz = 0.0;           // the device wall clock's rate
    int gmarks_n = 0;                 // the marks of the captured graph's chain
    const char *gmarks_name[kMaxMarks];
    int profiling = 0;                // marks: 0 none, 1 clock stamps, 2 HIP events
    bool force_scratch = false;       // union-find replay state in HBM even where LDS fits
    bool arena_fixed = false;         // the arena never grows (a serving lane's budget)
    hipEvent_t ev2 = nullptr;         // the end of the last enqueued chain (with its readback)
    int32_t inflight = 0;             // frames of the enqueued chain
    WorkQueue qstage{};               // the work lists' header of the call
    Tune tune = kDefaultTune;         // work splits of the chain
    bool graphs = true;               // replay the chain as a HIP graph when the call shape repeats
    hipGraphExec_t gexec = nullptr;   // the graph of the last call shape
    std::vector<int64_t> gkey;        // that shape
    std::vector<int> fstatus;         // per frame of the last call: ATG_OK, 1 (point capacity), ATG_ERR_LIMIT
    std::vector<double> stage_ms;
    std::vector<const char *> stage_names;
    // atg_detect_noalloc: this device's static scratch, and pinned staging
    // for its descriptors and results (both set up at creation)
    uint8_t *scratch = nullptr;
    uint8_t *na_stage = nullptr;
    // the chain's forked streams (the fit of the small, large and huge
    // clusters side by side); used when no per-stage timing is taken
    Fork fork;
    // atg_estimate_poses: device copies of the tags in and the poses out
    uint8_t *pose_buf = nullptr;
    size_t pose_cap = 0;
};

namespace {

int fail(atg_ctx *c, int code, const std::string &m) {
    c->err = m;
    return code;
}

#define HIPCHK(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(ctx, ATG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// quick_decode_init (fast/decode.c:20-58): per chunk j, the codes sorted by
// the chunk's value (counting sort, id order kept).
void build_tables(const at_family &f, std::vector<uint32_t> &off, std::vector<uint16_t> &ids,
                  std::vector<uint64_t> &codes, int64_t *chunk, int64_t *cap) {
    int64_t cs = (f.nbits + 3) / 4, capacity = (int64_t)1 << cs;
    uint64_t mask = (uint64_t)(capacity - 1);
    off.assign((size_t)(4 * (capacity + 1)), 0);
    ids.assign((size_t)(4 * f.ncodes), 0);
    codes.assign((size_t)(4 * f.ncodes), 0);
    for (int64_t i = 0; i < f.ncodes; i++)
        for (int64_t j = 0; j < 4; j++) off[(size_t)(j * (capacity + 1) + (int64_t)((f.codes[i] >> (j * cs)) & mask) + 1)]++;
    for (int64_t j = 0; j < 4; j++)
        for (int64_t v = 0; v < capacity; v++) off[(size_t)(j * (capacity + 1) + v + 1)] += off[(size_t)(j * (capacity + 1) + v)];
    std::vector<uint32_t> cur(off);
    for (int64_t i = 0; i < f.ncodes; i++)
        for (int64_t j = 0; j < 4; j++) {
            int64_t c = j * (capacity + 1) + (int64_t)((f.codes[i] >> (j * cs)) & mask);
            size_t at = (size_t)(j * f.ncodes + cur[(size_t)c]++);
            ids[at] = (uint16_t)i;
            codes[at] = f.codes[i];
        }
    *chunk = cs;
    *cap = capacity;
}

// The device parameters of a call (host computation only, nothing
// allocated): the detector parameters, min_tag_width, the blur kernel and
// host-libm constants as the reference computes them, and each family's
// geometry with the device addresses of its decode tables (resident since
// atg_set_families).
int fill_params(atg_ctx *ctx, const atg_params *p, DevParams *out) {
    // (the reference selects the (max_nmaxima+1)-th largest maximum,
    // fast/quads.c:826-855: a negative value reads outside its array)
    if (p->max_nmaxima < 0 || p->max_nmaxima > kMaxMaxima)
        return fail(ctx, ATG_ERR_ARG, "max_nmaxima outside 0..32 (the corner search keeps at most 32 maxima)");
    DevParams &h = *out;
    memset(&h, 0, sizeof h);
    h.decimate = p->quad_decimate;
    h.sigma = p->quad_sigma;
    h.refine = p->refine_edges ? 1 : 0;
    h.deglitch = p->deglitch ? 1 : 0;
    h.sharpening = p->decode_sharpening;
    h.mcp = p->min_cluster_pixels;
    h.max_nmaxima = p->max_nmaxima;
    h.mwbd = p->min_white_black_diff;
    h.cos_crit = p->cos_critical_rad;
    h.max_mse = p->max_line_fit_mse;
    h.nfam = (int32_t)ctx->fams.size();
    // min_tag_width (fast/detector.c:67-81)
    int32_t mtw = 1000000;
    bool nb = false, rb = false;
    for (auto &f : ctx->fams) {
        mtw = std::min(mtw, f.width_at_border);
        nb = nb || !f.reversed_border;
        rb = rb || f.reversed_border;
    }
    if (p->quad_decimate > 1.0f) mtw = (int32_t)((float)mtw / p->quad_decimate);
    if (mtw < 3) mtw = 3;
    h.tag_width = mtw;
    h.normal_border = nb;
    h.reversed_border = rb;
    // blur kernel (fast/image.c:111-147): host exp, as the reference
    int64_t ksz = 0;
    if (p->quad_sigma != 0.0f) {
        ksz = (int64_t)(4.0f * fabsf(p->quad_sigma));
        if ((ksz & 1) == 0) ksz++;
    }
    if (ksz > 63) return fail(ctx, ATG_ERR_ARG, "quad_sigma too large (blur kernel over 63 taps)");
    h.blur_ksz = ksz > 1 ? (int32_t)ksz : 0;
    if (ksz > 1) {
        double sigma = fabsf(p->quad_sigma), dk[64], acc = 0.0;
        for (int64_t i = 0; i < ksz; i++) {
            double t = (double)(-(ksz / 2) + i) / sigma;
            dk[i] = exp(-0.5 * (t * t));
        }
        for (int64_t i = 0; i < ksz; i++) acc += dk[i];
        for (int64_t i = 0; i < ksz; i++) h.blur_k[i] = (uint8_t)(dk[i] / acc * 255.0);
    }
    // qsm_kernel_init (fast/quads.c:740-745), make_detection's rotations
    // (fast/detector.c:85-86): host libm, as the reference
    for (int i = 0; i < 7; i++) h.qsm_kern[i] = (float)exp(-(double)((i - 3) * (i - 3)) / 2.0);
    for (int r = 0; r < 4; r++) {
        double th = (double)r * M_PI / 2.0;
        h.rot_cs[r][0] = cos(th);
        h.rot_cs[r][1] = sin(th);
    }
    for (size_t i = 0; i < ctx->fams.size(); i++) {
        const at_family &f = ctx->fams[i];
        DevFamily &d = h.fam[i];
        d.nbits = f.nbits;
        d.wab = f.width_at_border;
        d.tw = f.total_width;
        d.reversed = f.reversed_border ? 1 : 0;
        d.maxhamming = ctx->bits[i];
        d.chunk_size = (int32_t)((f.nbits + 3) / 4);
        d.ncodes = (int32_t)f.ncodes;
        d.capacity = (int64_t)1 << d.chunk_size;
        for (int k = 0; k < 64; k++) {
            d.bit_x[k] = f.bit_x[k];
            d.bit_y[k] = f.bit_y[k];
        }
        d.offsets = reinterpret_cast<const uint32_t *>(ctx->dtables + ctx->tab_off[i][0]);
        d.ids = reinterpret_cast<const uint16_t *>(ctx->dtables + ctx->tab_off[i][1]);
        d.codes = reinterpret_cast<const uint64_t *>(ctx->dtables + ctx->tab_off[i][2]);
    }
    return ATG_OK;
}

// The decode tables of the context's families into HBM, where they stay
// until the families change (set-up time: this allocates).
int upload_tables(atg_ctx *ctx) {
    size_t at = 0;
    std::vector<uint8_t> blob;
    for (size_t i = 0; i < ctx->fams.size(); i++) {
        std::vector<uint32_t> off;
        std::vector<uint16_t> ids;
        std::vector<uint64_t> codes;
        int64_t cs, cap;
        build_tables(ctx->fams[i], off, ids, codes, &cs, &cap);
        size_t o_off = align256(at), o_ids = align256(o_off + off.size() * 4), o_codes = align256(o_ids + ids.size() * 2);
        at = o_codes + codes.size() * 8;
        blob.resize(align256(at));
        memcpy(blob.data() + o_off, off.data(), off.size() * 4);
        memcpy(blob.data() + o_ids, ids.data(), ids.size() * 2);
        memcpy(blob.data() + o_codes, codes.data(), codes.size() * 8);
        ctx->tab_off[i][0] = o_off;
        ctx->tab_off[i][1] = o_ids;
        ctx->tab_off[i][2] = o_codes;
    }
    HIPCHK(hipSetDevice(ctx->device));
    if (ctx->stream) HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->dtables) HIPCHK(hipFree(ctx->dtables));
    ctx->dtables = nullptr;
    if (!blob.empty()) {
        HIPCHK(hipMalloc(&ctx->dtables, blob.size()));
        HIPCHK(hipMemcpy(ctx->dtables, blob.data(), blob.size(), hipMemcpyHostToDevice));
    }
    return ATG_OK;
}

int upload_params(atg_ctx *ctx, const atg_params *p) {
    int rc = fill_params(ctx, p, &ctx->hp);
    if (rc) return rc;
    if (!ctx->dparams) HIPCHK(hipMalloc(&ctx->dparams, sizeof(DevParams)));
    HIPCHK(hipMemcpy(ctx->dparams, &ctx->hp, sizeof(DevParams), hipMemcpyHostToDevice));
    ctx->last_params = *p;
    ctx->params_valid = true;
    return ATG_OK;
}

bool same_params(const atg_params &a, const atg_params &b) {
    return a.quad_decimate == b.quad_decimate && a.quad_sigma == b.quad_sigma && a.refine_edges == b.refine_edges &&
           a.decode_sharpening == b.decode_sharpening && a.min_cluster_pixels == b.min_cluster_pixels &&
           a.max_nmaxima == b.max_nmaxima && a.cos_critical_rad == b.cos_critical_rad &&
           a.max_line_fit_mse == b.max_line_fit_mse && a.min_white_black_diff == b.min_white_black_diff &&
           a.deglitch == b.deglitch;
}

// The region of one frame: offsets of every array (also its size).
size_t layout(const FramePlan &fp, int nfam, bool host_input, Frame *F, uint8_t *base) {
    size_t at = 0;
    auto take = [&](size_t bytes) {
        size_t o = at;
        at = align256(at + bytes);
        return base ? base + o : nullptr;
    };
    int64_t qw = fp.qw, qh = fp.qh, px = qw * qh, tw = qw / 4, th = qh / 4;
    int64_t nodes = px + qh, pc = fp.pcap;
    uint8_t *in = (host_input || fp.skip) ? take((size_t)(fp.W * fp.H)) : nullptr;
    if (F) F->img = in;
    F->qim = take((size_t)px);
    F->t = take((size_t)px);
    F->tbuf = take((size_t)px);
    F->tmx = take((size_t)(tw * th + 1));
    F->tmn = take((size_t)(tw * th + 1));
    F->runs = (uint32_t *)take((size_t)px * 4);
    F->rcnt = (uint32_t *)take((size_t)(qh + 1) * 4);
    F->row_off = (uint32_t *)take((size_t)(qh + 2) * 4);
    F->parent = (uint32_t *)take((size_t)nodes * 4);
    F->fw = (uint32_t *)take((size_t)nodes * 4);
    F->nlab = (uint32_t *)take((size_t)nodes * 4);
    F->rlab = F->nlab;
    F->recw = (uint32_t *)take((size_t)nodes * 4);
    F->rlost = (uint32_t *)take((size_t)nodes * 4);
    F->rroot = (uint32_t *)take((size_t)nodes * 4);
    F->nrec = (uint32_t *)take((size_t)(qh + 1) * 4);
    F->plab = (uint32_t *)take((size_t)px * 4);
    F->uedge = (uint32_t *)take((size_t)px * 12);
    F->ecnt = (uint32_t *)take((size_t)(qh + 1) * 4);
    F->pcnt = (uint32_t *)take((size_t)(qh + 1) * 4);
    F->poff = (uint32_t *)take((size_t)(qh + 2) * 4);
    F->pts = (uint64_t *)take((size_t)pc * 8);
    F->cpts = (uint64_t *)take((size_t)pc * 8);
    F->pw = F->pts;
    F->skey = (uint64_t *)take((size_t)pc * 8);
    F->skey2 = (uint64_t *)take((size_t)pc * 8);
    F->sval = (uint32_t *)take((size_t)pc * 4);
    F->sval2 = (uint32_t *)take((size_t)pc * 4);
    F->rhist = (uint32_t *)take((size_t)kRadixBins * kSortBlocks * 4);
    F->rwide = (uint32_t *)take((size_t)(qh + 1) * 4);
    const int64_t hs = hash_slots(pc, px, fp.hfull);
    F->hmask = (uint32_t)(hs - 1);
    // clusters the table takes before the overflow flag: one radix pass's
    // worth (2047) for a one-pass table
    F->ccap = (uint32_t)(one_pass_table(px, fp.hfull) ? (1 << kRadixBits) - 1 : hs / 2);
    F->hkey = (uint64_t *)take((size_t)hs * 8);
    F->hcnt = (uint32_t *)take((size_t)hs * 4);
    F->hrank = (uint32_t *)take((size_t)hs * 4);
    F->clist = (uint32_t *)take((size_t)hs / 2 * 4);
    F->csort = (uint64_t *)take((size_t)hs / 2 * 8);
    const size_t fs = front_scratch_bytes((int)qw, fp.force);
    F->fscr = fs ? take(fs) : nullptr;
    F->rscr = take(replay_scratch_bytes((int)qw, fp.force));
    F->cstart = (uint32_t *)take((size_t)(pc + 2) * 4);
    F->ccount = nullptr;
    F->cflag = (uint32_t *)take((size_t)(pc + 1) * 4);
    F->kcap = (uint32_t)cluster_cap(pc);
    F->klist = (uint32_t *)take((size_t)F->kcap * 8);
    F->cid = (uint64_t *)take((size_t)(pc + 1) * 8);
    F->okey = (uint64_t *)take((size_t)pc * 8);
    F->mom = (double *)take((size_t)pc * 6 * 8);
    F->errs = (double *)take((size_t)pc * 8);
    F->filt = (double *)take((size_t)pc * 8);
    F->midx = (uint32_t *)take((size_t)pc * 4);
    const size_t qc = (size_t)(pc / kMinFitPoints + 1);
    F->qcap = (uint32_t)qc;
    F->qtmp = (QuadRec *)take(sizeof(QuadRec) * qc);
    F->qline = (double *)take(sizeof(double) * 16 * qc);
    F->qcl = (uint32_t *)take(4 * qc);
    F->quads = (QuadRec *)take(sizeof(QuadRec) * qc);
    F->dtmp = (at_detection *)take(sizeof(at_detection) * qc * std::max(nfam, 1));
    F->dfin = (at_detection *)take(sizeof(at_detection) * qc * std::max(nfam, 1));
    F->dflag = (uint32_t *)take(4 * qc * std::max(nfam, 1));
    return at;
}

}  // namespace

extern "C" void atg_params_default(atg_params *p) {
    // fast/detector.c:8-20
    memset(p, 0, sizeof *p);
    p->quad_decimate = 2.0f;
    p->quad_sigma = 0.0f;
    p->refine_edges = true;
    p->decode_sharpening = 0.25;
    p->min_cluster_pixels = 24;
    p->max_nmaxima = 10;
    p->cos_critical_rad = (float)cos(10.0 * M_PI / 180.0);
    p->max_line_fit_mse = 10.0f;
    p->min_white_black_diff = 5;
    p->deglitch = false;
}

// ------------------------------------------------------------------
// The scratch of atg_detect_noalloc: every buffer of its one-frame call
// (the frame descriptor, the work lists' header, the parameters, the
// counters and first detections, the frame's arena region) lies in this
// array. Each GPU holds its own copy for the life of the process. Host
// code never passes the array's name to a kernel or a copy: its device
// address on each device is looked up once (hipGetSymbolAddress, with that
// device current) and kept.
static __device__ __align__(256) uint8_t scratch[ATG_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch[kMaxDevices];

// The offsets of the scratch's parts (256-byte aligned): frame descriptor,
// work lists, parameters, result block, then the frame's arena region.
constexpr size_t kNaFrame = 0;
constexpr size_t kNaQueue = (sizeof(Frame) + 255) & ~(size_t)255;
constexpr size_t kNaParams = kNaQueue + ((sizeof(WorkQueue) + 255) & ~(size_t)255) + kLists * 256;
constexpr size_t kNaResult = kNaParams + ((sizeof(DevParams) + 255) & ~(size_t)255);
constexpr size_t kNaArena = kNaResult + ((kResultBlock + 255) & ~(size_t)255);
// pinned staging: the same head (descriptor, work lists, parameters, result)
constexpr size_t kNoAllocStage = kNaArena;
static_assert(kNaQueue + ((sizeof(WorkQueue) + 255) & ~(size_t)255) >= kNaQueue + 256, "queue header");

// The device address of this device's scratch (set-up: once per device,
// with the context's device current).
int noalloc_setup(atg_ctx *ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return fail(ctx, ATG_ERR_ARG, "device ordinal over 63");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[ctx->device]) {
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch[ctx->device] = static_cast<uint8_t *>(ptr);
    }
    ctx->scratch = g_scratch[ctx->device];
    return ATG_OK;
}

}  // namespace

extern "C" int atg_ctx_create(int device, atg_ctx **out) {
    *out = nullptr;
    atg_ctx *ctx = new atg_ctx();
    ctx->device = device;
    int rc = ATG_OK;
    [&]() -> int {
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&ctx->ev0));
        HIPCHK(hipEventCreate(&ctx->ev1));
        HIPCHK(hipEventCreateWithFlags(&ctx->ev2, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&ctx->fork.go, hipEventDisableTiming));
        for (int k = 0; k < Fork::kForks; k++) {
            HIPCHK(hipStreamCreateWithFlags(&ctx->fork.s[k], hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&ctx->fork.done[k], hipEventDisableTiming));
        }
        HIPCHK(hipMalloc(&ctx->marks.stamps, kMaxMarks * sizeof(uint64_t)));
        HIPCHK(hipHostMalloc(&ctx->hstamps, kMaxMarks * sizeof(uint64_t)));
        int khz = 0;
        HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
        ctx->clock_khz = khz > 0 ? khz : 100000.0;
        int rc = noalloc_setup(ctx);
        if (rc) return rc;
        HIPCHK(hipHostMalloc(&ctx->na_stage, kNoAllocStage));
        warm_persistent_grids(ctx->tune.fit_blocks);
        prepare_points_stage();
        return ATG_OK;
    }();
    if (!ctx->err.empty()) {
        fprintf(stderr, "atg: %s\n", ctx->err.c_str());
        rc = ATG_ERR_HIP;
        atg_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return ATG_OK;
}

extern "C" void atg_ctx_destroy(atg_ctx *ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    hipFree(ctx->arena);
    hipFree(ctx->results);
    hipFree(ctx->dframes);
    hipFree(ctx->dtables);
    hipFree(ctx->dparams);
    hipFree(ctx->marks.stamps);
    for (hipEvent_t e : ctx->marks.ev)
        if (e) hipEventDestroy(e);
    hipHostFree(ctx->hstamps);
    hipHostFree(ctx->na_stage);
    hipHostFree(ctx->pinned);
    hipHostFree(ctx->pinned_out);
    hipFree(ctx->pose_buf);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    if (ctx->ev2) hipEventDestroy(ctx->ev2);
    if (ctx->fork.go) hipEventDestroy(ctx->fork.go);
    for (int k = 0; k < Fork::kForks; k++) {
        if (ctx->fork.done[k]) hipEventDestroy(ctx->fork.done[k]);
        if (ctx->fork.s[k]) hipStreamDestroy(ctx->fork.s[k]);
    }
    if (ctx->gexec) hipGraphExecDestroy(ctx->gexec);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char *atg_last_error(const atg_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

extern "C" int atg_set_families(atg_ctx *ctx, const at_family *const *fams, const int32_t *bits, int32_t n) {
    if (n < 0 || n > kMaxFamilies) return fail(ctx, ATG_ERR_ARG, "at most 16 families");
    for (int32_t i = 0; i < n; i++) {
        // at_add_family's checks (fast/detector.c:45-54)
        if (bits[i] < 0 || bits[i] > 3) return fail(ctx, ATG_ERR_ARG, "bits_corrected outside 0..3");
        if (fams[i]->ncodes >= 65536 || fams[i]->nbits <= 0 || fams[i]->nbits > 64)
            return fail(ctx, ATG_ERR_ARG, "family outside the supported sizes");
    }
    ctx->fams.clear();
    ctx->bits.clear();
    for (int32_t i = 0; i < n; i++) {
        ctx->fams.push_back(*fams[i]);
        ctx->bits.push_back(bits[i]);
    }
    ctx->params_valid = false;
    return upload_tables(ctx);
}

namespace atg {

// The plan of one frame: its sizes, point capacity and the bytes of its
// arena region (a frame the reference returns nothing for becomes an 8x8
// zero frame).
int plan_frame(atg_ctx *ctx, const atg_params *p, int nfam, const atg_frame &f, int on_device, FramePlan &fp) {
    fp = FramePlan{};
    fp.W = f.width;
    fp.H = f.height;
    fp.stride = f.stride;
    fp.skip = nfam == 0 || fp.W < 8 || fp.H < 8 || !f.pixels;
    decimated_size(fp.W, fp.H, p->quad_decimate, &fp.qw, &fp.qh);
    if (!fp.skip && (fp.qw < 4 || fp.qh < 4)) fp.skip = true;
    if (fp.skip) { fp.qw = fp.qh = 4; fp.W = fp.H = 8; }
    if (fp.qw > 32767 || fp.qh > 32767) return fail(ctx, ATG_ERR_LIMIT, "quad image wider or taller than 32767");
    // node ids (runs and a virtual node a row) take 31 bits (the ok flag is
    // the 32nd of a label)
    const int64_t nodes = fp.qw * fp.qh + fp.qh;
    if (nodes >= ((int64_t)1 << 31)) return fail(ctx, ATG_ERR_LIMIT, "quad image over 2^31 pixels");
    fp.pcap = fp.qw * fp.qh * ctx->pcap_scale + 64;
    fp.force = ctx->force_scratch;
    fp.hfull = ctx->hash_full;
    Frame F{};
    fp.size = layout(fp, nfam, !on_device, &F, nullptr);
    return ATG_OK;
}

// The launch shapes of a call, from its frames' plans.
struct ChainShape {
    int max_qw = 4, max_qh = 4, key_bits = 0, cl_bits = 0, max_lds_qw = 0;
    int64_t max_weight = 0;
    int max_cluster = 0;
    void add(const FramePlan &fp) {
        // (the sizing layout decides the front's memory as layout() does)
        if (front_scratch_bytes((int)fp.qw, fp.force) == 0) max_lds_qw = std::max<int>(max_lds_qw, (int)fp.qw);
        const int64_t nodes = fp.qw * fp.qh + fp.qh;
        max_weight = std::max<int64_t>(max_weight, nodes);
        max_qw = std::max<int>(max_qw, (int)fp.qw);
        max_qh = std::max<int>(max_qh, (int)fp.qh);
        key_bits = std::max(key_bits, ceil_log2((uint64_t)hash_slots(fp.pcap, fp.qw * fp.qh, fp.hfull) / 2));  // cluster ranks
        cl_bits = std::max(cl_bits, ceil_log2((uint64_t)fp.pcap + 1));
        max_cluster = std::max(max_cluster, (int)(2 * (2 * fp.qw + 2 * fp.qh)));
    }
};

// The descriptor of a planned frame whose arena region starts at `region`
// and whose counters and first detections go to `result` (the image
// pointer is set by the caller).
void describe_frame(const DevParams &hp, const FramePlan &fp, int nfam, bool host_input, int key_bits,
                    uint8_t *region, uint8_t *result, Frame &F) {
    F = Frame{};
    layout(fp, nfam, host_input, &F, region);
    F.W = (int32_t)fp.W;
    F.H = (int32_t)fp.H;
    F.qw = (int32_t)fp.qw;
    F.qh = (int32_t)fp.qh;
    F.tw = (int32_t)(fp.qw / 4);
    F.th = (int32_t)(fp.qh / 4);
    (void)key_bits;
    F.bucket_mask = (uint64_t)cluster_map_size(fp.qw, fp.qh) - 1;
    F.max_px = (int32_t)(2 * (2 * fp.qw + 2 * fp.qh));
    F.blurred = hp.blur_ksz > 1 ? 1 : 0;
    F.pcap = fp.pcap;
    F.counters = reinterpret_cast<uint32_t *>(result);
    F.inl = reinterpret_cast<at_detection *>(result + kCountersBytes);
    if (fp.skip) {
        // an 8x8 zero frame (k_init clears it): no tag can come out of it
        F.skip = 1;
        F.stride = 8;
    }
}

// The kernel chain of a call over n described frames (dframes, dparams
// and the work lists' header dq in device memory), on stream st.
void launch_chain(hipStream_t st, const ChainShape &cs, const CclPlan &ccl, const DevParams &hp, bool deglitch,
                  const Frame *dframes, const DevParams *dparams, int n, WorkQueue *dq, const Tune &tu, bool force,
                  Marks *mk, Fork *fk) {
    // the angle sort in LDS per cluster, or (clusters too large) of all points in HBM
    const int sort_cap = angle_sort_cap(cs.max_cluster, force);
    // the corner search of the small clusters a wave each, when its pair
    // tables hold the maxima kept
    const bool wave_corners = hp.max_nmaxima <= 16;
    const bool blur = hp.blur_ksz > 1;
    launch_init(st, dframes, n, dq);
    mk->mark(st, "init");
    launch_image_stage(st, dframes, dparams, n, cs.max_qw, cs.max_qh, blur, deglitch, tu, mk);
    launch_ccl_stage(st, dframes, dparams, n, cs.max_qw, cs.max_qh, ccl, dq, tu, mk);
    launch_points_stage(st, dframes, n, cs.max_qh, cs.key_bits, tu, mk);
    launch_quads_stage(st, dframes, dparams, n, cs.cl_bits, tu.fit_blocks, sort_cap, wave_corners, dq, mk, fk);
    launch_decode_stage(st, dframes, dparams, n, tu.fit_blocks, hp.refine != 0, dq, mk);
}

// The first half of a call: plans, arena, frame descriptors, the kernel
// chain and the readback copy, all queued on the context's stream; the
// event ev2 marks the end. chain_collect reads the results once it has
// passed.
int chain_enqueue(atg_ctx *ctx, const atg_params *p, const atg_frame *frames, int32_t n, int on_device) {
    HIPCHK(hipSetDevice(ctx->device));
    if (!ctx->params_valid || !same_params(ctx->last_params, *p)) {
        int rc = upload_params(ctx, p);
        if (rc) return rc;
    }
    const int nfam = (int)ctx->fams.size();
    ctx->plans.assign((size_t)n, FramePlan{});
    ctx->hframes.assign((size_t)n, Frame{});
    size_t total = 0, pinned_need = 0;
    ChainShape cs;
    std::vector<uint8_t> pinned_src((size_t)n, 0);
    for (int32_t i = 0; i < n; i++) {
        FramePlan &fp = ctx->plans[(size_t)i];
        int rc = plan_frame(ctx, p, nfam, frames[i], on_device, fp);
        if (rc) return rc;
        cs.add(fp);
        fp.off = total;
        total += fp.size;
        if (!on_device && !fp.skip) {
            // host frames: dense rows in page-locked memory go by DMA, the
            // others through the context's pinned staging
            hipPointerAttribute_t attr{};
            const bool locked = fp.stride == fp.W &&
                                hipPointerGetAttributes(&attr, frames[i].pixels) == hipSuccess &&
                                attr.type == hipMemoryTypeHost;
            (void)hipGetLastError();  // (pageable memory is not known to HIP: not an error here)
            pinned_src[(size_t)i] = locked ? 1 : 0;
            if (!locked) pinned_need += align256((size_t)(fp.W * fp.H));
        }
    }
    const int max_qw = cs.max_qw, max_qh = cs.max_qh, key_bits = cs.key_bits, cl_bits = cs.cl_bits;
    const int max_lds_qw = cs.max_lds_qw;
    // the union-find's launch shapes (its HBM state is in each frame's region)
    const CclPlan ccl = ccl_plan(max_lds_qw, max_qw, cs.max_weight, ctx->force_scratch, (int)n);
    const size_t qoff = total;
    total += queue_bytes(n);
    if (total > ctx->arena_size && ctx->arena_fixed) return fail(ctx, ATG_ERR_LIMIT, "batch over the arena budget");
    if (total > ctx->arena_size) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipFree(ctx->arena));
        ctx->arena = nullptr;
        HIPCHK(hipMalloc(&ctx->arena, total));
        ctx->arena_size = total;
    }
    size_t rneed = kResultBlock * (size_t)n;
    if (rneed > ctx->results_size) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipFree(ctx->results));
        ctx->results = nullptr;
        HIPCHK(hipMalloc(&ctx->results, rneed));
        ctx->results_size = rneed;
    }
    if (sizeof(Frame) * (size_t)n > ctx->dframes_cap) {
        HIPCHK(hipFree(ctx->dframes));
        ctx->dframes = nullptr;
        HIPCHK(hipMalloc(&ctx->dframes, sizeof(Frame) * (size_t)n));
        ctx->dframes_cap = sizeof(Frame) * (size_t)n;
    }
    size_t out_need = (kCountersBytes + kInlineDets * sizeof(at_detection)) * (size_t)n + sizeof(Frame) * (size_t)n +
                      align256(sizeof(WorkQueue)) + 256;
    if (pinned_need > ctx->pinned_size) {
        HIPCHK(hipHostFree(ctx->pinned));
        ctx->pinned = nullptr;
        HIPCHK(hipHostMalloc(&ctx->pinned, pinned_need));
        ctx->pinned_size = pinned_need;
    }
    if (out_need > ctx->pinned_out_size) {
        HIPCHK(hipHostFree(ctx->pinned_out));
        ctx->pinned_out = nullptr;
        HIPCHK(hipHostMalloc(&ctx->pinned_out, out_need));
        ctx->pinned_out_size = out_need;
    }
    // frame descriptors, and the host frames into pinned staging; the
    // stream work (copies in, kernels, results out) is issued below
    struct Copy {
        void *dst;
        const void *src;
        size_t bytes;
    };
    std::vector<Copy> copies;
    size_t pin_at = 0;
    // the work lists: the header staged in pinned memory, copied in with the frames
    WorkQueue *dq = reinterpret_cast<WorkQueue *>(ctx->arena + qoff);
    {
        uint8_t *qb = ctx->arena + qoff + align256(sizeof(WorkQueue));
        WorkQueue hq{};
        hq.nframes = (uint32_t)n;
        for (int l = 0; l < kLists; l++)
            hq.off[l] = reinterpret_cast<uint32_t *>(qb + (size_t)l * align256((size_t)(n + 1) * 4));
        ctx->qstage = hq;
    }
    for (int32_t i = 0; i < n; i++) {
        FramePlan &fp = ctx->plans[(size_t)i];
        Frame &F = ctx->hframes[(size_t)i];
        describe_frame(ctx->hp, fp, nfam, !on_device, key_bits, ctx->arena + fp.off,
                       ctx->results + kResultBlock * (size_t)i, F);
        if (fp.skip) {
        } else if (on_device) {
            F.img = frames[i].pixels;
            F.stride = fp.stride;
        } else if (pinned_src[(size_t)i]) {
            // page-locked caller memory: copied in by DMA as it lies
            copies.push_back(Copy{(void *)F.img, frames[i].pixels, (size_t)(fp.W * fp.H)});
            F.stride = fp.W;
        } else {
            uint8_t *dst = ctx->pinned + pin_at;
            for (int64_t y = 0; y < fp.H; y++) memcpy(dst + y * fp.W, frames[i].pixels + y * fp.stride, (size_t)fp.W);
            copies.push_back(Copy{(void *)F.img, dst, (size_t)(fp.W * fp.H)});
            pin_at += align256((size_t)(fp.W * fp.H));
            F.stride = fp.W;
        }
    }
    Frame *fstage = reinterpret_cast<Frame *>(ctx->pinned_out);
    memcpy(fstage, ctx->hframes.data(), sizeof(Frame) * (size_t)n);
    copies.push_back(Copy{ctx->dframes, fstage, sizeof(Frame) * (size_t)n});
    const size_t pre = kCountersBytes + kInlineDets * sizeof(at_detection);
    uint8_t *hout = ctx->pinned_out + sizeof(Frame) * (size_t)n;
    uint8_t *qstage = ctx->pinned_out + align256(sizeof(Frame) * (size_t)n + pre * (size_t)n);
    memcpy(qstage, &ctx->qstage, sizeof(WorkQueue));
    copies.push_back(Copy{dq, qstage, sizeof(WorkQueue)});
    hipStream_t st = ctx->stream;
    Marks *mk = &ctx->marks;
    mk->level = ctx->profiling;
    ctx->fork.on = ctx->profiling == 0;
    const Tune tu = ctx->tune;
    const int sort_cap = angle_sort_cap(cs.max_cluster, ctx->force_scratch);
    const bool wave_corners = ctx->hp.max_nmaxima <= 16;
    const bool blur = ctx->hp.blur_ksz > 1, deglitch = p->deglitch;
    warm_persistent_grids(tu.fit_blocks);  // device queries stay out of stream capture
    prepare_points_stage();
    auto issue = [&]() -> hipError_t {
        mk->n = 0;
        for (const Copy &c : copies) {
            hipError_t e = hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyHostToDevice, st);
            if (e != hipSuccess) return e;
        }
        launch_chain(st, cs, ccl, ctx->hp, deglitch, ctx->dframes, ctx->dparams, n, dq, tu, ctx->force_scratch, mk,
                     &ctx->fork);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (mk->n > 0 && mk->level == 1) {
            e = hipMemcpyAsync(ctx->hstamps, mk->stamps, sizeof(uint64_t) * (size_t)mk->n, hipMemcpyDeviceToHost, st);
            if (e != hipSuccess) return e;
        }
        // results: counters and the first detections of every frame in one copy
        return hipMemcpy2DAsync(hout, pre, ctx->results, kResultBlock, pre, (size_t)n, hipMemcpyDeviceToHost, st);
    };
    HIPCHK(hipEventRecord(ctx->ev0, st));
    if (ctx->graphs && ctx->profiling != 2) {
        // one graph per call shape: everything the stream work depends on
        std::vector<int64_t> key = {n,       max_qw,         max_qh,         key_bits,      cl_bits,
                                    blur,    deglitch,       nfam,           ccl.front,     ccl.replay,
                                    ccl.replay_cap, ccl.replay_pack, ccl.replay_mid, ccl.replay_mid_cap, ccl.replay_mid_pack,
                                    ccl.replay_mid_grid, ccl.replay_wide, ccl.replay_wide_cap, ccl.replay_wide_grid,
                                    ccl.replay_waves, max_lds_qw,
                                    tu.thr_tile_rows, tu.row_waves, tu.front_threads, tu.fit_blocks, tu.front_bands,
                                    ccl.front_band, sort_cap, wave_corners, ctx->hp.refine,
                                    ctx->profiling,
                                    (int64_t)(intptr_t)ctx->arena, (int64_t)(intptr_t)ctx->results,
                                    (int64_t)(intptr_t)ctx->dframes, (int64_t)(intptr_t)ctx->pinned_out,
                                    (int64_t)(intptr_t)ctx->dparams};
        for (const Copy &c : copies) {
            key.push_back((int64_t)(intptr_t)c.dst);
            key.push_back((int64_t)(intptr_t)c.src);
            key.push_back((int64_t)c.bytes);
        }
        if (!ctx->gexec || key != ctx->gkey) {
            if (ctx->gexec) hipGraphExecDestroy(ctx->gexec);
            ctx->gexec = nullptr;
            ctx->gkey.clear();
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            hipError_t e = issue();
            hipError_t e2 = hipStreamEndCapture(st, &g);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (g) hipGraphDestroy(g);
                return fail(ctx, ATG_ERR_HIP, std::string("graph capture: ") +
                                                  hipGetErrorString(e != hipSuccess ? e : e2));
            }
            e = hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) {
                ctx->gexec = nullptr;
                return fail(ctx, ATG_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
            }
            ctx->gkey = key;
            ctx->gmarks_n = mk->n;
            for (int k = 0; k < mk->n; k++) ctx->gmarks_name[k] = mk->name[k];
        }
        // the marks of the captured chain (a replay issues nothing itself)
        mk->n = ctx->gmarks_n;
        for (int k = 0; k < mk->n; k++) mk->name[k] = ctx->gmarks_name[k];
        HIPCHK(hipGraphLaunch(ctx->gexec, st));
    } else {
        HIPCHK(issue());
    }
    HIPCHK(hipEventRecord(ctx->ev1, st));
    HIPCHK(hipEventRecord(ctx->ev2, st));
    ctx->inflight = n;
    return ATG_OK;
}

bool chain_done(atg_ctx *ctx) { return hipEventQuery(ctx->ev2) == hipSuccess; }

// The second half: after ev2, the counters and detections of every frame.
// Returns 1 if a frame ran out of point capacity (the caller re-runs the
// call with more).
int chain_collect(atg_ctx *ctx) {
    const int32_t n = ctx->inflight;
    ctx->inflight = 0;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipEventSynchronize(ctx->ev2));
    Marks *mk = &ctx->marks;
    size_t pre = kCountersBytes + kInlineDets * sizeof(at_detection);
    uint8_t *hout = ctx->pinned_out + sizeof(Frame) * (size_t)n;
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->chain_ms = ms;
    ctx->stage_ms.clear();
    ctx->stage_names.clear();
    for (int k = 1; k < mk->n; k++) {
        float ev_ms = 0.0f;
        if (mk->level == 2) HIPCHK(hipEventElapsedTime(&ev_ms, mk->ev[k - 1], mk->ev[k]));
        ctx->stage_ms.push_back(mk->level == 2 ? (double)ev_ms
                                               : (double)(ctx->hstamps[k] - ctx->hstamps[k - 1]) / ctx->clock_khz);
        ctx->stage_names.push_back(mk->name[k]);
    }
    ctx->dets.assign((size_t)n, {});
    ctx->nquads.assign((size_t)n, 0);
    ctx->fstatus.assign((size_t)n, ATG_OK);
    ctx->need_points = ctx->need_table = false;
    bool overflow = false;
    std::string first_err;
    for (int32_t i = 0; i < n; i++) {
        const uint32_t *cnt = reinterpret_cast<const uint32_t *>(hout + pre * (size_t)i);
        const at_detection *d = reinterpret_cast<const at_detection *>(hout + pre * (size_t)i + kCountersBytes);
        if (ctx->plans[(size_t)i].skip) continue;
        if (cnt[kError] & (2u | 128u)) {
            // out of point (2) or cluster-table (128) room: re-run with more
            overflow = true;
            ctx->fstatus[(size_t)i] = 1;
            ctx->need_points = ctx->need_points || (cnt[kError] & 2u);
            ctx->need_table = ctx->need_table || (cnt[kError] & 128u);
            continue;
        }
        if (cnt[kError]) {
            char m[128];
            snprintf(m, sizeof m, "frame %d: device capacity exceeded (flags 0x%x)", i, cnt[kError]);
            if (first_err.empty()) first_err = m;
            ctx->fstatus[(size_t)i] = ATG_ERR_LIMIT;
            continue;
        }
        uint32_t nd = cnt[kNDets];
        ctx->nquads[(size_t)i] = cnt[kNQuads];
        auto &v = ctx->dets[(size_t)i];
        v.resize(nd);
        uint32_t first = std::min<uint32_t>(nd, kInlineDets);
        if (first) memcpy(v.data(), d, first * sizeof(at_detection));
        if (nd > first)
            HIPCHK(hipMemcpy(v.data() + first, ctx->hframes[(size_t)i].dfin + first, (nd - first) * sizeof(at_detection),
                             hipMemcpyDeviceToHost));
    }
    if (!first_err.empty()) return fail(ctx, ATG_ERR_LIMIT, first_err);
    if (overflow) return 1;  // the caller retries with more room (retry_room)
    return ATG_OK;
}

int run_chain(atg_ctx *ctx, const atg_params *p, const atg_frame *frames, int32_t n, int on_device) {
    int rc = chain_enqueue(ctx, p, frames, n, on_device);
    if (rc) return rc;
    return chain_collect(ctx);
}

// After a call that ran out of room: more room for the next run (false
// when the worst-case sizes were used already).
bool retry_room(atg_ctx *ctx) {
    bool more = false;
    if (ctx->need_points && ctx->pcap_scale < 4) {
        ctx->pcap_scale = 4;  // more boundary points than quad-image pixels: the worst case is 4
        more = true;
    }
    if (ctx->need_table && !ctx->hash_full) {
        ctx->hash_full = true;
        more = true;
    }
    return more;
}

// A call with retries until the frames fit (at most two: the worst cases).
int run_chain_fit(atg_ctx *ctx, const atg_params *p, const atg_frame *frames, int32_t n, int on_device) {
    ctx->pcap_scale = 1;
    ctx->hash_full = false;
    int rc = run_chain(ctx, p, frames, n, on_device);
    while (rc == 1 && retry_room(ctx)) rc = run_chain(ctx, p, frames, n, on_device);
    if (rc == 1) rc = fail(ctx, ATG_ERR_LIMIT, "point or cluster capacity exceeded");
    return rc;
}

// Arena bytes of a call over these frames (what chain_enqueue allocates).
size_t batch_arena_bytes(const atg_ctx *ctx, const atg_params *p, const atg_frame *frames, int32_t n,
                         int on_device) {
    const int nfam = (int)ctx->fams.size();
    size_t total = 0;
    int max_qw = 4, max_qh = 4;
    for (int32_t i = 0; i < n; i++) {
        FramePlan fp{};
        fp.W = frames[i].width;
        fp.H = frames[i].height;
        fp.skip = nfam == 0 || fp.W < 8 || fp.H < 8 || !frames[i].pixels;
        decimated_size(fp.W, fp.H, p->quad_decimate, &fp.qw, &fp.qh);
        if (!fp.skip && (fp.qw < 4 || fp.qh < 4)) fp.skip = true;
        if (fp.skip) { fp.qw = fp.qh = 4; fp.W = fp.H = 8; }
        fp.pcap = fp.qw * fp.qh * ctx->pcap_scale + 64;
        fp.force = ctx->force_scratch;
        fp.hfull = ctx->hash_full;
        Frame F{};
        total += layout(fp, nfam, !on_device, &F, nullptr);
    }
    return total + queue_bytes(n);
}

// A fixed arena of `bytes` and result space for `max_frames`: later calls
// that need more fail instead of growing it.
int ctx_reserve(atg_ctx *ctx, size_t bytes, int32_t max_frames) {
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipFree(ctx->arena));
    ctx->arena = nullptr;
    HIPCHK(hipMalloc(&ctx->arena, bytes));
    ctx->arena_size = bytes;
    ctx->arena_fixed = true;
    HIPCHK(hipFree(ctx->results));
    ctx->results = nullptr;
    ctx->results_size = kResultBlock * (size_t)max_frames;
    HIPCHK(hipMalloc(&ctx->results, ctx->results_size));
    HIPCHK(hipFree(ctx->dframes));
    ctx->dframes = nullptr;
    ctx->dframes_cap = sizeof(Frame) * (size_t)max_frames;
    HIPCHK(hipMalloc(&ctx->dframes, ctx->dframes_cap));
    size_t out_need = (kCountersBytes + kInlineDets * sizeof(at_detection) + sizeof(Frame)) * (size_t)max_frames +
                      align256(sizeof(WorkQueue)) + 256;
    HIPCHK(hipHostFree(ctx->pinned_out));
    ctx->pinned_out = nullptr;
    HIPCHK(hipHostMalloc(&ctx->pinned_out, out_need));
    ctx->pinned_out_size = out_need;
    return ATG_OK;
}

hipStream_t ctx_stream(atg_ctx *ctx) { return ctx->stream; }
const std::vector<at_detection> &ctx_dets(const atg_ctx *ctx, int32_t i) { return ctx->dets[(size_t)i]; }
int ctx_frame_status(const atg_ctx *ctx, int32_t i) { return ctx->fstatus[(size_t)i]; }
int64_t ctx_nquads(const atg_ctx *ctx, int32_t i) { return ctx->nquads[(size_t)i]; }
size_t ctx_result_bytes(int32_t max_frames) { return kResultBlock * (size_t)max_frames; }

}  // namespace atg

extern "C" int atg_detect_batch(atg_ctx *ctx, const atg_params *p, const atg_frame *frames, int32_t n, int on_device,
                                atg_result *out) {
    if (!ctx || !p || (n > 0 && (!frames || !out)) || n < 0) return ATG_ERR_ARG;
    if (n == 0) return ATG_OK;
    ctx->err.clear();
    int rc = run_chain_fit(ctx, p, frames, n, on_device);
    if (rc) return rc;
    for (int32_t i = 0; i < n; i++) {
        out[i].dets = ctx->dets[(size_t)i].data();
        out[i].ndets = (int64_t)ctx->dets[(size_t)i].size();
        out[i].nquads = ctx->nquads[(size_t)i];
    }
    return ATG_OK;
}

extern "C" int atg_detect(atg_ctx *ctx, const atg_params *p, const uint8_t *pixels, int64_t width, int64_t height,
                          int64_t stride, int on_device, atg_result *out) {
    atg_frame f{pixels, width, height, stride};
    return atg_detect_batch(ctx, p, &f, 1, on_device, out);
}

// One frame with every buffer in the device's static scratch: no device,
// pinned or host allocation (the decode tables are resident since
// atg_set_families; descriptors go through the context's pinned staging).
extern "C" int atg_detect_noalloc(atg_ctx *ctx, const atg_params *p, const uint8_t *pixels, int64_t width,
                                  int64_t height, int64_t stride, int on_device, at_detection *dets, int64_t cap,
                                  int64_t *ndets, int64_t *nquads) {
    if (!ctx || !p || !ndets || !nquads || cap < 0 || (cap > 0 && !dets)) return ATG_ERR_ARG;
    ctx->err.clear();
    *ndets = 0;
    *nquads = 0;
    const int nfam = (int)ctx->fams.size();
    // no families, or under 8x8: nothing (fast/detector.c:254-255)
    if (nfam == 0 || width < 8 || height < 8 || !pixels) return ATG_OK;
    if (stride < width) return fail(ctx, ATG_ERR_ARG, "stride under the width");
    if (!ctx->scratch || !ctx->na_stage) return fail(ctx, ATG_ERR_HIP, "no device scratch on this context");
    HIPCHK(hipSetDevice(ctx->device));
    uint8_t *stage = ctx->na_stage;
    DevParams &hp = *reinterpret_cast<DevParams *>(stage + kNaParams);
    int rc = fill_params(ctx, p, &hp);
    if (rc) return rc;
    const atg_frame f{pixels, width, height, stride};
    hipStream_t st = ctx->stream;
    // the context's sizing, restored on every way out
    struct Restore {
        atg_ctx *c;
        int64_t scale;
        bool full;
        ~Restore() {
            c->pcap_scale = scale;
            c->hash_full = full;
        }
    } restore{ctx, ctx->pcap_scale, ctx->hash_full};
    ctx->pcap_scale = 1;
    ctx->hash_full = false;
    for (int attempt = 0; attempt < 3; attempt++) {
        FramePlan fp;
        rc = plan_frame(ctx, p, nfam, f, on_device, fp);
        if (rc) return rc;
        if (fp.skip) return ATG_OK;  // a quad image under 4x4: nothing
        if (kNaArena + fp.size > ATG_NOALLOC_SCRATCH_BYTES) {
            char m[200];
            snprintf(m, sizeof m, "the frame needs %zu bytes of device scratch; atg_detect_noalloc holds %zu",
                     kNaArena + fp.size, (size_t)ATG_NOALLOC_SCRATCH_BYTES);
            return fail(ctx, ATG_ERR_LIMIT, m);
        }
        ChainShape cs;
        cs.add(fp);
        const CclPlan ccl = ccl_plan(cs.max_lds_qw, cs.max_qw, cs.max_weight, ctx->force_scratch, 1);
        uint8_t *base = ctx->scratch;
        Frame &F = *reinterpret_cast<Frame *>(stage + kNaFrame);
        describe_frame(hp, fp, nfam, !on_device, cs.key_bits, base + kNaArena, base + kNaResult, F);
        F.stride = on_device ? stride : width;
        if (on_device) F.img = pixels;
        WorkQueue &q = *reinterpret_cast<WorkQueue *>(stage + kNaQueue);
        memset(&q, 0, sizeof q);
        q.nframes = 1;
        for (int l = 0; l < kLists; l++)
            q.off[l] = reinterpret_cast<uint32_t *>(base + kNaQueue + align256(sizeof(WorkQueue)) + (size_t)l * 256);
        if (!on_device)
            HIPCHK(hipMemcpy2DAsync(const_cast<uint8_t *>(F.img), (size_t)width, pixels, (size_t)stride, (size_t)width,
                                    (size_t)height, hipMemcpyHostToDevice, st));
        // descriptor, work lists' header and parameters in one copy
        HIPCHK(hipMemcpyAsync(base, stage, kNaResult, hipMemcpyHostToDevice, st));
        Marks off;
        Fork nofork;
        launch_chain(st, cs, ccl, hp, p->deglitch, reinterpret_cast<const Frame *>(base + kNaFrame),
                     reinterpret_cast<const DevParams *>(base + kNaParams), 1,
                     reinterpret_cast<WorkQueue *>(base + kNaQueue), ctx->tune, ctx->force_scratch, &off, &nofork);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(stage + kNaResult, base + kNaResult, kResultBlock, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        const uint32_t *cnt = reinterpret_cast<const uint32_t *>(stage + kNaResult);
        if (cnt[kError] & (2u | 128u)) {
            // out of point or cluster-table room: again with the worst case
            ctx->need_points = (cnt[kError] & 2u) != 0;
            ctx->need_table = (cnt[kError] & 128u) != 0;
            if (retry_room(ctx)) continue;
            break;
        }
        if (cnt[kError]) {
            char m[96];
            snprintf(m, sizeof m, "device capacity exceeded (flags 0x%x)", cnt[kError]);
            return fail(ctx, ATG_ERR_LIMIT, m);
        }
        const int64_t nd = cnt[kNDets], take = std::min(nd, cap), first = std::min<int64_t>(take, kInlineDets);
        *ndets = nd;
        *nquads = cnt[kNQuads];
        if (first) memcpy(dets, stage + kNaResult + kCountersBytes, (size_t)first * sizeof(at_detection));
        if (take > first)
            HIPCHK(hipMemcpy(dets + first, F.dfin + first, (size_t)(take - first) * sizeof(at_detection),
                             hipMemcpyDeviceToHost));
        return ATG_OK;
    }
    return fail(ctx, ATG_ERR_LIMIT, "point or cluster capacity exceeded");
}

extern "C" int atg_device_alloc(atg_ctx *ctx, size_t bytes, void **out) {
    if (!ctx || !out) return ATG_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMalloc(out, bytes));
    return ATG_OK;
}

extern "C" int atg_device_free(atg_ctx *ctx, void *p) {
    if (!ctx) return ATG_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipFree(p));
    return ATG_OK;
}

extern "C" int atg_copy_to_device(atg_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return ATG_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return ATG_OK;
}

extern "C" int atg_set_profiling(atg_ctx *ctx, int on) {
    if (!ctx || on < 0 || on > 2) return ATG_ERR_ARG;
    if (on == 2 && !ctx->marks.ev[0]) {
        HIPCHK(hipSetDevice(ctx->device));
        for (int k = 0; k < kMaxMarks; k++) HIPCHK(hipEventCreate(&ctx->marks.ev[k]));
    }
    ctx->profiling = on;
    return ATG_OK;
}

static bool valid_tune(const Tune &t) {
    auto in = [](int v, std::initializer_list<int> l) { return std::find(l.begin(), l.end(), v) != l.end(); };
    // row_waves: the per-row kernels are compiled for blocks of at most 256
    // threads (4 waves)
    return in(t.thr_tile_rows, {1, 2, 4}) && in(t.row_waves, {1, 2, 4}) && in(t.front_threads, {64, 128, 256, 512, 1024}) &&
           t.fit_blocks >= 32 && t.fit_blocks <= 4096 && t.front_bands >= 0 && t.front_bands <= 64;
}

extern "C" int atg_get_tune(const atg_ctx *ctx, atg_tune *out) {
    if (!ctx || !out) return ATG_ERR_ARG;
    static_assert(sizeof(atg_tune) == sizeof(Tune), "atg_tune mirrors Tune");
    memcpy(out, &ctx->tune, sizeof(Tune));
    return ATG_OK;
}

extern "C" int atg_set_tune(atg_ctx *ctx, const atg_tune *t) {
    if (!ctx || !t) return ATG_ERR_ARG;
    Tune v;
    memcpy(&v, t, sizeof v);
    if (!valid_tune(v)) return fail(ctx, ATG_ERR_ARG, "work split outside the supported values");
    ctx->tune = v;
    return ATG_OK;
}

extern "C" int atg_set_graphs(atg_ctx *ctx, int on) {
    if (!ctx) return ATG_ERR_ARG;
    ctx->graphs = on != 0;
    return ATG_OK;
}

// Coordinate descent over the work splits on the caller's batch: each
// value of one split with the others at their best so far, the device
// time of the chain (median of three runs) deciding; every candidate's
// detections must equal the first run's.
extern "C" int atg_autotune(atg_ctx *ctx, const atg_params *p, const atg_frame *frames, int32_t n, int on_device,
                            atg_tune *best_out, double *best_ms) {
    if (!ctx || !p || !frames || n < 1) return ATG_ERR_ARG;
    const Tune start = ctx->tune;
    const bool graphs = ctx->graphs;
    const int prof = ctx->profiling;
    ctx->graphs = false;
    ctx->profiling = 0;
    std::vector<std::vector<at_detection>> want;
    std::vector<int64_t> want_q;
    auto measure = [&](const Tune &t, double *ms) -> int {
        ctx->tune = t;
        std::vector<double> v;
        for (int rep = 0; rep < 4; rep++) {
            int rc = run_chain_fit(ctx, p, frames, n, on_device);
            if (rc) return rc;
            if (rep > 0) v.push_back(ctx->chain_ms);
        }
        std::sort(v.begin(), v.end());
        *ms = v[v.size() / 2];
        if (want.empty()) {
            want = ctx->dets;
            want_q = ctx->nquads;
            return ATG_OK;
        }
        for (int32_t i = 0; i < n; i++)
            if (ctx->nquads[(size_t)i] != want_q[(size_t)i] || ctx->dets[(size_t)i].size() != want[(size_t)i].size() ||
                memcmp(ctx->dets[(size_t)i].data(), want[(size_t)i].data(),
                       want[(size_t)i].size() * sizeof(at_detection)) != 0)
                return fail(ctx, ATG_ERR_LIMIT, "a work split changed the detections");
        return ATG_OK;
    };
    Tune best = start;
    double bms = 0.0;
    int rc = measure(best, &bms);
    const std::vector<int> cand[5] = {{1, 2, 4},
                                      {1, 2, 4},
                                      {128, 256, 512, 1024},
                                      {256, 512, 1024, 2048, 4096},
                                      {0, 1, 2, 4, 8, 16, 32, 64}};
    for (int k = 0; k < 5 && rc == ATG_OK; k++) {
        for (int v : cand[k]) {
            Tune t = best;
            int32_t *f[5] = {&t.thr_tile_rows, &t.row_waves, &t.front_threads, &t.fit_blocks, &t.front_bands};
            if (*f[k] == v) continue;
            *f[k] = v;
            double ms = 0.0;
            rc = measure(t, &ms);
            if (rc) break;
            if (ms < bms) {
                bms = ms;
                best = t;
            }
        }
    }
    ctx->graphs = graphs;
    ctx->profiling = prof;
    ctx->tune = rc == ATG_OK ? best : start;
    if (rc) return rc;
    if (best_out) memcpy(best_out, &best, sizeof best);
    if (best_ms) *best_ms = bms;
    return ATG_OK;
}

extern "C" int atg_debug_copy(atg_ctx *ctx, void *dst, const void *src, int64_t bytes, double *ms) {
    if (!ctx || !dst || !src || bytes < 16 || bytes % 16) return ATG_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    launch_copy(ctx->stream, dst, src, bytes);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float f = 0.0f;
    HIPCHK(hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
    if (ms) *ms = f;
    return ATG_OK;
}

extern "C" int atg_debug_finish(atg_ctx *ctx, at_detection *dets, int64_t *n) {
    if (!ctx || !n || *n < 0 || *n > (int64_t)1 << 24 || (*n > 0 && !dets)) return ATG_ERR_ARG;
    if (*n == 0) return ATG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    const size_t bytes = sizeof(at_detection) * (size_t)*n;
    // the list, the output, 5 n u32 of scratch, the count
    const size_t total = align256(2 * bytes + 20 * (size_t)*n);
    uint8_t *buf = nullptr;
    HIPCHK(hipMalloc(&buf, total + 256));
    int64_t *dn = reinterpret_cast<int64_t *>(buf + total);
    int rc = [&]() -> int {
        HIPCHK(hipMemcpyAsync(buf, dets, bytes, hipMemcpyHostToDevice, ctx->stream));
        launch_debug_finish(ctx->stream, buf, (uint32_t)*n, dn);
        HIPCHK(hipGetLastError());
        int64_t kept = 0;
        HIPCHK(hipMemcpyAsync(&kept, dn, sizeof kept, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipMemcpy(dets, buf + bytes, sizeof(at_detection) * (size_t)kept, hipMemcpyDeviceToHost));
        *n = kept;
        return ATG_OK;
    }();
    hipFree(buf);
    return rc;
}

extern "C" int atg_estimate_poses(atg_ctx *ctx, const atg_pose_in *in, int64_t n, int32_t n_iters, int32_t flags,
                                  atg_pose_out *out) {
    if (!ctx) return ATG_ERR_ARG;
    if (n < 0 || n > ((int64_t)1 << 24) || (n > 0 && (!in || !out)))
        return fail(ctx, ATG_ERR_ARG, "atg_estimate_poses: 0..2^24 tags, non-null arrays");
    if (flags & ~ATG_POSE_HOMOGRAPHY_ONLY) return fail(ctx, ATG_ERR_ARG, "atg_estimate_poses: unknown flags");
    if (n == 0) return ATG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    const size_t bin = sizeof(atg_pose_in) * (size_t)n, bout = sizeof(atg_pose_out) * (size_t)n;
    const size_t need = align256(bin) + bout;
    if (need > ctx->pose_cap) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (ctx->pose_buf) HIPCHK(hipFree(ctx->pose_buf));
        ctx->pose_buf = nullptr;
        ctx->pose_cap = 0;
        HIPCHK(hipMalloc(&ctx->pose_buf, need));
        ctx->pose_cap = need;
    }
    auto *din = reinterpret_cast<atg_pose_in *>(ctx->pose_buf);
    auto *dout = reinterpret_cast<atg_pose_out *>(ctx->pose_buf + align256(bin));
    HIPCHK(hipMemcpyAsync(din, in, bin, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(launch_pose(ctx->stream, din, n, n_iters, flags, dout));
    HIPCHK(hipMemcpyAsync(out, dout, bout, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ATG_OK;
}

extern "C" int atg_debug_trig(atg_ctx *ctx, const float *y, const float *x, const float *th, int64_t n,
                              float *atan2_out, float *cos_out, float *sin_out, int slow) {
    if (!ctx || n < 0 || n > ((int64_t)1 << 26) || (n > 0 && (!y || !x || !th || !atan2_out || !cos_out || !sin_out)))
        return ATG_ERR_ARG;
    if (n == 0) return ATG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    const size_t b = sizeof(float) * (size_t)n;
    float *buf = nullptr;
    HIPCHK(hipMalloc(&buf, 6 * b));
    int rc = [&]() -> int {
        HIPCHK(hipMemcpyAsync(buf, y, b, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(buf + n, x, b, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(buf + 2 * n, th, b, hipMemcpyHostToDevice, ctx->stream));
        launch_debug_trig(ctx->stream, buf, buf + n, buf + 2 * n, n, buf + 3 * n, buf + 4 * n, buf + 5 * n, slow);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(atan2_out, buf + 3 * n, b, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(cos_out, buf + 4 * n, b, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(sin_out, buf + 5 * n, b, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return ATG_OK;
    }();
    hipFree(buf);
    return rc;
}

extern "C" int atg_debug_force_scratch(atg_ctx *ctx, int on) {
    if (!ctx) return ATG_ERR_ARG;
    ctx->force_scratch = on != 0;
    return ATG_OK;
}

extern "C" int atg_stage_times(const atg_ctx *ctx, const char **names, double *ms, int32_t cap) {
    if (!ctx) return ATG_ERR_ARG;
    int32_t n = (int32_t)ctx->stage_ms.size();
    for (int32_t i = 0; i < n && i < cap; i++) {
        names[i] = ctx->stage_names[(size_t)i];
        ms[i] = ctx->stage_ms[(size_t)i];
    }
    return n;
}

extern "C" int atg_last_timing(const atg_ctx *ctx, double *chain_ms) {
    if (!ctx) return ATG_ERR_ARG;
    *chain_ms = ctx->chain_ms;
    return ATG_OK;
}

extern "C" void atg_stages_free(atg_stages *s) {
    free(s->qim); free(s->t); free(s->runs); free(s->row_off); free(s->rep); free(s->csize); free(s->ok);
    free(s->cluster_ids); free(s->ccount); free(s->pts); free(s->quads);
    memset(s, 0, sizeof *s);
}

extern "C" int atg_debug_counters(atg_ctx *ctx, int32_t frame, uint32_t *out16) {
    if (!ctx || frame < 0 || (size_t)frame >= ctx->hframes.size()) return ATG_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(out16, ctx->hframes[(size_t)frame].counters, 16 * 4, hipMemcpyDeviceToHost));
    return ATG_OK;
}

extern "C" int atg_debug_stages(atg_ctx *ctx, int32_t frame, atg_stages *out) {
    memset(out, 0, sizeof *out);
    if (!ctx || frame < 0 || (size_t)frame >= ctx->hframes.size()) return ATG_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const Frame &F = ctx->hframes[(size_t)frame];
    int64_t qw = F.qw, qh = F.qh, px = qw * qh;
    uint32_t cnt[kNCounters];
    HIPCHK(hipMemcpy(cnt, F.counters, sizeof cnt, hipMemcpyDeviceToHost));
    out->qw = qw;
    out->qh = qh;
    out->qim = (uint8_t *)malloc((size_t)px);
    out->t = (uint8_t *)malloc((size_t)px);
    HIPCHK(hipMemcpy(out->qim, F.qim, (size_t)px, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(out->t, F.t, (size_t)px, hipMemcpyDeviceToHost));
    // runs, packed in row order
    std::vector<uint32_t> rcnt((size_t)qh), runs((size_t)px);
    HIPCHK(hipMemcpy(rcnt.data(), F.rcnt, 4 * (size_t)qh, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(runs.data(), F.runs, 4 * (size_t)px, hipMemcpyDeviceToHost));
    int64_t nr = 0;
    for (int64_t y = 0; y < qh; y++) nr += rcnt[(size_t)y];
    out->nruns = nr;
    out->runs = (uint32_t *)malloc(12 * (size_t)nr + 4);
    out->row_off = (uint32_t *)malloc(4 * (size_t)(qh + 1));
    int64_t k = 0;
    for (int64_t y = 0; y < qh; y++) {
        out->row_off[y] = (uint32_t)k;
        for (uint32_t i = 0; i < rcnt[(size_t)y]; i++, k++) {
            uint32_t r = runs[(size_t)(y * qw + i)];
            out->runs[3 * k] = r & 0x7fff;
            out->runs[3 * k + 1] = (r >> 15) & 0x7fff;
            out->runs[3 * k + 2] = (r >> 30) & 1 ? 255 : 0;
        }
    }
    out->row_off[qh] = (uint32_t)k;
    int64_t nn = nr + qh;
    out->nnodes = nn;
    std::vector<uint32_t> nlab((size_t)nn), fw((size_t)nn);
    HIPCHK(hipMemcpy(nlab.data(), F.nlab, 4 * (size_t)nn, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fw.data(), F.fw, 4 * (size_t)nn, hipMemcpyDeviceToHost));
    out->rep = (uint32_t *)malloc(4 * (size_t)nn + 4);
    out->csize = (uint32_t *)malloc(4 * (size_t)nn + 4);
    out->ok = (uint8_t *)malloc((size_t)nn + 1);
    for (int64_t i = 0; i < nn; i++) {
        uint32_t r = nlab[(size_t)i] & 0x7fffffffu;
        out->rep[i] = r;
        out->ok[i] = (uint8_t)(nlab[(size_t)i] >> 31);
        out->csize[i] = r < (uint32_t)nn ? fw[r] : 0;
    }
    // clusters and their points in output order
    uint32_t np = cnt[kNPts], nc = cnt[kNClusters];
    out->npts = np;
    out->nclusters = nc;
    out->pts = (uint64_t *)malloc(8 * (size_t)np + 8);
    HIPCHK(hipMemcpy(out->pts, F.cpts, 8 * (size_t)np, hipMemcpyDeviceToHost));
    std::vector<uint32_t> cs((size_t)nc + 1);
    HIPCHK(hipMemcpy(cs.data(), F.cstart, 4 * ((size_t)nc + 1), hipMemcpyDeviceToHost));
    out->ccount = (uint32_t *)malloc(4 * (size_t)nc + 4);
    out->cluster_ids = (uint64_t *)malloc(8 * (size_t)nc + 8);
    HIPCHK(hipMemcpy(out->cluster_ids, F.cid, 8 * (size_t)nc, hipMemcpyDeviceToHost));
    for (uint32_t c = 0; c < nc; c++) out->ccount[c] = cs[c + 1] - cs[c];
    uint32_t nq = std::min<uint32_t>(cnt[kNQuads], F.qcap);
    out->nquads = nq;
    std::vector<QuadRec> q(nq);
    HIPCHK(hipMemcpy(q.data(), F.quads, sizeof(QuadRec) * nq, hipMemcpyDeviceToHost));
    out->quads = (float *)malloc(36 * (size_t)nq + 4);
    for (uint32_t i = 0; i < nq; i++) {
        for (int j = 0; j < 4; j++) {
            out->quads[9 * i + 2 * j] = q[i].p[j][0];
            out->quads[9 * i + 2 * j + 1] = q[i].p[j][1];
        }
        out->quads[9 * i + 8] = q[i].reversed ? 1.0f : 0.0f;
    }
    return ATG_OK;
}
