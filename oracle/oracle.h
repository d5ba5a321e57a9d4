/* CPU restatement of the reference AprilTag detector (single thread), used
 * ONLY as the checker in tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg. The product never links or calls it.
 *
 * Every function in oracle.c names the reference file:line it follows.
 * The stage struct has the same layout as the reference harness's
 * (oracle/ref_harness.c RefStages), so one ctypes mirror reads both. */
#ifndef ATG_ORACLE_H
#define ATG_ORACLE_H
#include <stdbool.h>
#include <stdint.h>

typedef struct {
    char name[32];
    const uint64_t *codes;
    int64_t ncodes;
    int32_t nbits, h, width_at_border, total_width;
    bool reversed_border;
    int32_t bit_x[64], bit_y[64];
} orc_family; /* the layout of at_family (fast/apriltag_fast.h:24-36) */

typedef struct {
    int32_t nthreads;
    float quad_decimate, quad_sigma;
    bool refine_edges;
    double decode_sharpening;
    bool debug;
    int32_t min_cluster_pixels, max_nmaxima;
    float critical_rad, cos_critical_rad, max_line_fit_mse;
    int32_t min_white_black_diff;
    bool deglitch;
} orc_params; /* the parameter head of at_detector (fast/apriltag_fast.h:73-87) */

typedef struct {
    int32_t family, id, hamming;
    float decision_margin;
    double H[9], c[2], p[4][2];
} orc_detection; /* the layout of at_detection (fast/apriltag_fast.h:52-64) */

typedef struct {
    int64_t qw, qh;
    uint8_t *qim;
    uint8_t *t;
    int64_t nruns;
    uint32_t *runs;
    uint32_t *row_off;
    int64_t nnodes;
    uint32_t *rep;
    uint32_t *csize;
    uint8_t *ok;
    int64_t nclusters;
    uint64_t *cluster_ids;
    uint32_t *ccount;
    int64_t npts;
    uint64_t *pts;
    int64_t nquads;
    float *quads;
} orc_stages;

typedef struct {
    orc_detection *dets;
    int64_t ndets;
    int64_t nquads;
    int64_t nraw; /* detections before reconcile */
} orc_result;

/* Detect tags. `bits` holds bits_corrected per family. stages may be NULL.
 * Returns 0, or -1 when the reference returns early (no families, or the
 * image under 8x8). Free with orc_result_free / orc_stages_free. */
int orc_detect(const orc_params *p, const orc_family *fams, const int32_t *bits, int32_t nfam,
               const uint8_t *pixels, int64_t w, int64_t h, int64_t stride, orc_result *res, orc_stages *stages);
void orc_result_free(orc_result *r);
void orc_stages_free(orc_stages *s);
void orc_default_params(orc_params *p);
/* Removes overlapping detections of the same tag in place, keeping the
 * better one (fast/decode.c:860-891). */
void orc_reconcile(orc_detection *d, int64_t *n);

/* Correctly rounded float atan2, cos and sin (cr_libm.c): the libm calls of
 * refine_edges (fast/decode.c:740-744). */
float cr_atan2f(float y, float x);
float cr_cosf(float x);
float cr_sinf(float x);

#endif
