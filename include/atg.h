/* atg: the thin C-ABI of the MI355X AprilTag detection path.
 *
 * Plain pointers and sizes only. The native API (apriltag_fast.h) is a
 * layer over this one; the batched and serving forms are entry points
 * beside the single-frame one, not changes to it.
 *
 * Reference interfaces each entry point replaces or extends:
 *   atg_ctx_create/destroy  - the device half of at_detector_create /
 *                             at_detector_destroy (fast/detector.c:4-43)
 *   atg_set_families        - at_add_family / at_clear_families
 *                             (fast/detector.c:45-60): uploads the
 *                             quick-decode tables (fast/decode.c:20-58)
 *   atg_detect              - at_detect (fast/detector.c:246-335) for one
 *                             frame, in host or device memory
 *   atg_detect_batch        - no counterpart: many frames of any sizes in
 *                             one pass of the kernel chain
 *   atg_debug_stages        - no counterpart: the stage arrays of the
 *                             last call, in the layout of the reference's
 *                             quad_thresh_single intermediates
 *                             (fast/detector.c:198-244)
 *
 * Errors are negative return codes; atg_last_error() describes the last
 * one. All calls on one context must come from one thread at a time. */
#ifndef ATG_H
#define ATG_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#include "apriltag_fast.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ATG_OK 0
#define ATG_ERR_ARG -1     /* bad argument */
#define ATG_ERR_HIP -2     /* a HIP call failed */
#define ATG_ERR_LIMIT -3   /* an image beyond the supported size */
#define ATG_ERR_NOFAM -4   /* no families (the reference returns nothing) */

typedef struct atg_ctx atg_ctx;

/* The detection parameters (the parameter head of at_detector,
 * fast/apriltag_fast.h:73-87). max_nmaxima must be 0..32 (ATG_ERR_ARG
 * otherwise): the corner search keeps at most 32 maxima on the device. */
typedef struct {
    float quad_decimate;
    float quad_sigma;
    bool refine_edges;
    double decode_sharpening;
    int32_t min_cluster_pixels;
    int32_t max_nmaxima;
    float cos_critical_rad;
    float max_line_fit_mse;
    int32_t min_white_black_diff;
    bool deglitch;
} atg_params;

typedef struct {
    const uint8_t *pixels;
    int64_t width, height, stride;
} atg_frame;

/* The detections of one frame: dets points into the context (valid until
 * the next detect call on it), sorted by id as at_detect's. */
typedef struct {
    const at_detection *dets;
    int64_t ndets;
    int64_t nquads;
} atg_result;

/* Stage arrays of one frame, malloc'd (free with atg_stages_free). The
 * layout matches the stage dump of the reference harness. */
typedef struct {
    int64_t qw, qh;
    uint8_t *qim;
    uint8_t *t;
    int64_t nruns;
    uint32_t *runs;     /* nruns x 3: start, end, v */
    uint32_t *row_off;  /* qh + 1 */
    int64_t nnodes;
    uint32_t *rep;      /* representative (union-find root) of each node */
    uint32_t *csize;    /* pixel count of each node's component */
    uint8_t *ok;
    int64_t nclusters;
    uint64_t *cluster_ids;
    uint32_t *ccount;
    int64_t npts;
    uint64_t *pts;      /* grouped by cluster, clusters in output order */
    int64_t nquads;
    float *quads;       /* nquads x 9: p[4][2] on the quad image, reversed flag */
} atg_stages;

void atg_params_default(atg_params *p);

/* device: HIP device ordinal. Returns ATG_OK and *out, or an error. */
int atg_ctx_create(int device, atg_ctx **out);
void atg_ctx_destroy(atg_ctx *ctx);
const char *atg_last_error(const atg_ctx *ctx);

/* Replace the context's families: fams[i] with bits[i] bits corrected.
 * The same validation as at_add_family (fast/detector.c:45-54). */
int atg_set_families(atg_ctx *ctx, const at_family *const *fams, const int32_t *bits, int32_t n);

/* One frame. on_device != 0: pixels is a device pointer. */
int atg_detect(atg_ctx *ctx, const atg_params *p, const uint8_t *pixels, int64_t width, int64_t height,
               int64_t stride, int on_device, atg_result *out);

/* n frames of any sizes through one pass of the kernel chain. out[i] is
 * what atg_detect gives for frames[i] alone. */
int atg_detect_batch(atg_ctx *ctx, const atg_params *p, const atg_frame *frames, int32_t n, int on_device,
                     atg_result *out);

/* Stage arrays of frame `frame` of the last detect call. */
int atg_debug_stages(atg_ctx *ctx, int32_t frame, atg_stages *out);
void atg_stages_free(atg_stages *s);
/* The 16 device counters of frame `frame` of the last call (runs, points,
 * clusters, quads, detections, error flags, then stage-internal
 * diagnostics). */
int atg_debug_counters(atg_ctx *ctx, int32_t frame, uint32_t *out16);

/* Work splits of the kernel chain (any values give the same detections):
 * threshold tile rows per block (1, 2, 4), rows per block of the per-row
 * kernels (1, 2, 4), threads of the union-find front sweep (64, 128, 256,
 * 512, 1024), blocks of the persistent fit and decode kernels (32..4096; at most
 * as many as are resident at once are launched), row bands of a frame's
 * union-find front at most (1..64, each of 16 rows or more, swept at once;
 * 0: chosen by the number of frames of the call). */
typedef struct {
    int32_t thr_tile_rows, row_waves, front_threads, fit_blocks, front_bands;
} atg_tune;
int atg_get_tune(const atg_ctx *ctx, atg_tune *out);
int atg_set_tune(atg_ctx *ctx, const atg_tune *t);
/* Chooses the work splits for this batch by timing candidates on it
 * (coordinate descent, device time), checks that each gives the same
 * detections, and keeps the fastest in the context. */
int atg_autotune(atg_ctx *ctx, const atg_params *p, const atg_frame *frames, int32_t n, int on_device,
                 atg_tune *best, double *best_ms);
/* on != 0 (the default): calls whose shape (frame sizes, parameters, work
 * splits, buffers) repeats replay the captured HIP graph of the chain
 * instead of launching each kernel. */
int atg_set_graphs(atg_ctx *ctx, int on);

/* A plain device-to-device copy kernel of `bytes` (a multiple of 16) on
 * the context's stream, its device time in *ms: the known-size reference
 * that HBM traffic counters are scaled by. */
int atg_debug_copy(atg_ctx *ctx, void *dst, const void *src, int64_t bytes, double *ms);

/* The device's last step of a frame on a caller's detection list, in
 * place: reconcile (fast/decode.c:860-891), then the stable sort by id
 * (fast/detector.c:164-179); *n becomes the count kept. */
int atg_debug_finish(atg_ctx *ctx, at_detection *dets, int64_t *n);

/* The refinement's correctly rounded float functions (the libm calls of
 * refine_edges, fast/decode.c:740-744) on host arrays of n: atan2_out[i]
 * = atan2f(y[i], x[i]), cos_out[i] = cosf(th[i]), sin_out[i] = sinf(th[i])
 * as the device computes them (|th| <= 4). slow != 0: by the
 * double-double path wherever it applies (which the device otherwise takes
 * only next to a float rounding boundary). */
int atg_debug_trig(atg_ctx *ctx, const float *y, const float *x, const float *th, int64_t n, float *atan2_out,
                   float *cos_out, float *sin_out, int slow);

/* on != 0: the following calls take the paths of frames too large for
 * LDS (the union-find replay's state in HBM scratch, the angle sort of all
 * points at once in HBM), so that tests can check them on any frame. */
int atg_debug_force_scratch(atg_ctx *ctx, int on);

/* Device time (ms, HIP events on the context's stream) of the last
 * detect call's kernel chain. */
int atg_last_timing(const atg_ctx *ctx, double *chain_ms);

/* Per-kernel device times of the following calls (on: 1) a one-lane
 * kernel between the kernel groups stamps the device wall clock (a few us
 * each; graph replay stays on); 2) HIP events between them on the
 * context's stream (the calls then run without graphs); 0) off. */
int atg_set_profiling(atg_ctx *ctx, int on);
/* The stages of the last call: names[i] and ms[i] for i < the return
 * value (at most cap are written). */
int atg_stage_times(const atg_ctx *ctx, const char **names, double *ms, int32_t cap);

/* Bytes of the static device scratch of atg_detect_noalloc. Every GPU the
 * library runs on holds one such array until the process ends. */
#define ATG_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)

/* No counterpart in the reference: atg_detect for callers that cannot
 * allocate (tight serving loops, harnesses that forbid hipMalloc). One
 * frame (on_device != 0: pixels is a device pointer); every buffer of the
 * call is the library's static device scratch of ATG_NOALLOC_SCRATCH_BYTES,
 * so the call allocates nothing: no device, pinned or host memory (the
 * decode tables are put on the device by atg_set_families, the context's
 * pinned staging by atg_ctx_create). A frame whose buffers do not fit
 * (a 1920x1080 frame at decimate 2 takes about 110 MB) is refused with
 * ATG_ERR_LIMIT. The detections, sorted by id as atg_detect's, go to
 * dets: *ndets is the frame's count, of which at most cap are written;
 * *nquads the quad count. All calls on one GPU share its scratch, which
 * keeps the last call's contents (the chain clears what it counts):
 * calls that overlap on one GPU, from any contexts, race. */
int atg_detect_noalloc(atg_ctx *ctx, const atg_params *p, const uint8_t *pixels, int64_t width, int64_t height,
                       int64_t stride, int on_device, at_detection *dets, int64_t cap, int64_t *ndets,
                       int64_t *nquads);

/* ------------------------------------------------------------------
 * Tag pose (apriltag_pose.h:1-77 in the reference, the consumer of the
 * detections): many tags in one launch, a device thread each.
 *
 * One tag: its detection's homography and corners (apriltag_detection_t
 * H and p, apriltag.h:215-250) and the camera (apriltag_detection_info_t,
 * apriltag_pose.h:11-18). */
typedef struct {
    double H[9];          /* row-major */
    double p[4][2];       /* corners, as the detection's */
    double tagsize;       /* metres */
    double fx, fy, cx, cy;
} atg_pose_in;

/* The poses of one tag (R row-major 3x3, t 3x1):
 *   hpose   homography_to_pose(H, -fx, fy, cx, cy) (common/homography.c:
 *           275-351), the top three rows of its 4x4 [R | T]
 *   R0, t0  estimate_pose_for_tag_homography (apriltag_pose.c:455-488)
 *   R1..err2  estimate_tag_pose_orthogonal_iteration (apriltag_pose.c:
 *           490-526): the refined homography pose and its object-space
 *           error, and the second local minimum (has2 = 0 and err2 =
 *           HUGE_VAL when there is none)
 *   best    estimate_tag_pose's choice (apriltag_pose.c:528-547): 1 when
 *           err1 <= err2, else 2 */
typedef struct {
    double hpose[12];
    double R0[9], t0[3];
    double R1[9], t1[3], err1;
    double R2[9], t2[3], err2;
    int32_t has2, best;
} atg_pose_out;

/* flags of atg_estimate_poses */
#define ATG_POSE_HOMOGRAPHY_ONLY 1  /* hpose, R0 and t0 only */

/* No counterpart as one call in the reference (which estimates one tag a
 * call, apriltag_pose.c:490-547): the poses of n tags, host arrays in and
 * out, n_iters orthogonal-iteration steps (the reference's estimate_tag_pose
 * uses 50). Floating point follows the reference's formulas; its SVD
 * polar factors and 3x3 inverses are formed by other, equally accurate
 * methods, so results agree to rounding (tests bound them by 1e-9
 * relative), not bit for bit. */
int atg_estimate_poses(atg_ctx *ctx, const atg_pose_in *in, int64_t n, int32_t n_iters, int32_t flags,
                       atg_pose_out *out);

/* Device buffers on the context's device, for callers without their own
 * allocator (frames kept resident in HBM between calls). */
int atg_device_alloc(atg_ctx *ctx, size_t bytes, void **out);
int atg_device_free(atg_ctx *ctx, void *p);
int atg_copy_to_device(atg_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ------------------------------------------------------------------
 * Serving (no counterpart in the reference): frames from many camera
 * streams, admitted as they arrive into batches that overlap on the
 * device. `lanes` batches run at once, each a kernel chain on its own
 * stream with a fixed arena; a lane takes the oldest waiting frames while
 * its arena holds them, and the chain copies each frame from host memory
 * into the batch's arena (by DMA from page-locked memory). The budget
 * covers the lanes' arenas (frames included) and result buffers. A
 * frame's detections are the ones atg_detect gives for it alone. */
typedef struct atg_server atg_server;

typedef struct {
    int32_t device;
    size_t budget_bytes;       /* HBM for frames in flight */
    int32_t lanes;             /* batches in flight at once (1..8) */
    int32_t max_batch;         /* frames per batch at most */
    int32_t slots_per_stream;  /* a stream's frames in flight at most (0: no cap) */
} atg_server_config;

/* One completed frame. Times are on the server clock (atg_server_now_ns). */
typedef struct {
    int32_t stream;
    int32_t status;            /* ATG_OK or an error code */
    uint64_t tag;
    int64_t arrival_ns, admit_ns, done_ns;
    int64_t nquads;
    int64_t ndets;
    const at_detection *dets;  /* valid until the next atg_server_poll */
    int32_t batch_frames;      /* frames of the batch it ran in */
} atg_served;

/* A frame to arrive at_ns after atg_server_feed is called. */
typedef struct {
    int32_t stream;
    int64_t at_ns;
    const uint8_t *pixels;     /* host memory, borrowed until the result is polled */
    int64_t stride;
    uint64_t tag;
} atg_arrival;

typedef struct {
    int64_t submitted, dropped, completed, batches, max_batch_seen;
    int32_t lanes;
    size_t lane_arena_bytes, slot_bytes;  /* slot_bytes: 0 (frames live in the lanes' arenas) */
} atg_server_counts;

void atg_server_config_default(atg_server_config *c);  /* device 0, 16 GiB, 3 lanes, 128, no cap */
int atg_server_create(const atg_server_config *c, const atg_params *p, const at_family *const *fams,
                      const int32_t *bits, int32_t nfam, atg_server **out);
/* Stops the worker; frames not yet arrived are cancelled, admitted ones finish. */
void atg_server_destroy(atg_server *s);
const char *atg_server_error(const atg_server *s);
int64_t atg_server_now_ns(const atg_server *s);
/* Streams are added before the first frame (the lanes are sized then). */
int atg_server_add_stream(atg_server *s, int32_t width, int32_t height, int32_t *stream_id);
/* A frame of `stream` arrives now (pixels borrowed until its result is
 * polled). ATG_ERR_LIMIT: the stream's queue is full and the frame dropped. */
int atg_server_submit(atg_server *s, int32_t stream, const uint8_t *pixels, int64_t stride, uint64_t tag);
/* Frames that arrive at their times, released by the server's own feeder
 * (precise arrival processes without a caller thread). A fed frame that
 * finds its stream's queue full is dropped: its record comes back from
 * atg_server_poll with status ATG_ERR_LIMIT and no detections. */
int atg_server_feed(atg_server *s, const atg_arrival *arrivals, int64_t n);
/* Completed frames, in completion order: up to cap into out; returns the count. */
int atg_server_poll(atg_server *s, atg_served *out, int32_t cap);
/* Waits until every submitted and fed frame is done, at most timeout_ms. */
int atg_server_drain(atg_server *s, int64_t timeout_ms);
int atg_server_stats(atg_server *s, atg_server_counts *out);
/* Pinned host memory (frames that the server copies in by DMA). */
int atg_host_alloc(size_t bytes, void **out);
int atg_host_free(void *p);

/* ------------------------------------------------------------------
 * Seeded synthetic frames (host code), for tests and benchmarks. */
typedef struct {
    int32_t width, height;
    int32_t ntags_min, ntags_max;
    int32_t side_min, side_max;
    double max_tilt;        /* tan(tilt/2) bound of each of two tilts */
    int32_t family_a;       /* index into the built-in families (3 = tag36h11) */
    int32_t family_b;       /* second family or -1 (8 = tagStandard52h13) */
    int32_t background;     /* 0: uniform + linear gradient, 1: value noise */
    int32_t bg_amp;
    double blur_min, blur_max;
    double noise_min, noise_max;
    int32_t nbars, nclutter, ncheckers;
} atg_synth_spec;

typedef struct {
    int32_t family, id;
    double corners[8];
} atg_synth_tag;

void atg_synth_default(atg_synth_spec *s, int config);
/* Writes a frame into out (stride bytes a row); returns the number of
 * tags placed (truth gets up to max_truth of them), or -1. */
int atg_synth_frame(const atg_synth_spec *spec, uint64_t seed, uint8_t *out, int64_t stride, atg_synth_tag *truth,
                    int32_t max_truth);

/* Frame `step` of a camera stream: the frame of `seed` with each tag moved
 * step frames along its own slow drift (step 0 = atg_synth_frame). */
int atg_synth_frame_drift(const atg_synth_spec *spec, uint64_t seed, int64_t step, uint8_t *out, int64_t stride,
                          atg_synth_tag *truth, int32_t max_truth);

/* The built-in families in the order tag16h5, tag25h9, tag36h10,
 * tag36h11, tagCircle21h7, tagCircle49h12, tagCustom48h12,
 * tagStandard41h12, tagStandard52h13 (index 0..8), or NULL. */
const at_family *atg_builtin_family(int32_t index);

#ifdef __cplusplus
}
#endif

#endif
