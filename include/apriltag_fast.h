/* The native AprilTag detection API, served by the MI355X detector.
 *
 * Drop-in for the reference's native API (fast/apriltag_fast.h:24-141):
 * the same functions, the same argument meaning and the same struct
 * layouts, so code written against the CPU detector links against
 * libapriltag_amd.so unchanged. The detection itself runs as HIP kernels
 * on the GPU (see atg.h); at_detect copies the frame to the device,
 * detects, and copies the detections back into td->detections.
 *
 *     at_detector *td = at_detector_create();
 *     at_add_family(td, at_tag36h11(), 2);
 *     at_detect(td, pixels, width, height, stride);
 *     for (int64_t i = 0; i < td->ndetections; i++) { ... td->detections[i] ... }
 *     at_detector_destroy(td);
 */
#ifndef APRILTAG_FAST_H
#define APRILTAG_FAST_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A tag family: code words and bit layout (fast/apriltag_fast.h:24-36).
 * The codes must outlive every detector the family is added to. */
typedef struct at_family {
    char name[32];
    const uint64_t *codes;
    int64_t ncodes;
    int32_t nbits;
    int32_t h;
    int32_t width_at_border;
    int32_t total_width;
    bool reversed_border;
    int32_t bit_x[64];
    int32_t bit_y[64];
} at_family;

/* Quad detection parameters (fast/apriltag_fast.h:40-48). */
typedef struct at_quad_thresh_params {
    int32_t min_cluster_pixels;   /* 24 */
    int32_t max_nmaxima;          /* 10 */
    float critical_rad;           /* not used */
    float cos_critical_rad;       /* cos(10 degrees) */
    float max_line_fit_mse;       /* 10.0 */
    int32_t min_white_black_diff; /* 5 */
    bool deglitch;                /* false */
} at_quad_thresh_params;

/* One detected tag (fast/apriltag_fast.h:52-64). */
typedef struct at_detection {
    int32_t family;
    int32_t id;
    int32_t hamming;
    float decision_margin;
    double H[9];
    double c[2];
    double p[4][2];
} at_detection;

/* Time spent in each stage of the last detection, in seconds
 * (fast/apriltag_fast.h:67-69). On the GPU the stages overlap in one
 * stream; `total` is the wall time of at_detect and `decode` holds the
 * device time of the whole pipeline. */
typedef struct at_timing {
    double decimate, blur, threshold, unionfind, clusters, fit_quads, decode, reconcile, total;
} at_timing;

struct at_internal;

/* fast/apriltag_fast.h:73-96. nthreads is accepted and ignored: the
 * detection runs on the GPU. */
typedef struct at_detector {
    int32_t nthreads;
    float quad_decimate;
    float quad_sigma;
    bool refine_edges;
    double decode_sharpening;
    bool debug;
    at_quad_thresh_params qtp;
    at_detection *detections;
    int64_t ndetections;
    int64_t nquads;
    at_timing timing;
    struct at_internal *in;
} at_detector;

/* fast/apriltag_fast.h:99-101 */
at_detector *at_detector_create(void);
void at_detector_destroy(at_detector *td);

/* fast/apriltag_fast.h:103-108: false for bits_corrected outside 0..3,
 * a 17th family, ncodes >= 65536 or nbits outside 1..64. */
bool at_add_family(at_detector *td, const at_family *fam, int bits_corrected);
void at_clear_families(at_detector *td);
/* fast/apriltag_fast.h:110 */
const char *at_family_name(const at_detector *td, const at_detection *d);

/* fast/apriltag_fast.h:112-114. Aborts with a message on a GPU error
 * (the reference aborts on arena failure, fast/util.c:26-29). */
void at_detect(at_detector *td, const uint8_t *pixels, int64_t width, int64_t height, int64_t stride);

/* fast/apriltag_fast.h:116-117: releases the detector's device buffers;
 * the next at_detect allocates them again. */
void at_shutdown(at_detector *td);

/* fast/apriltag_fast.h:121-131 */
const at_family *at_tag16h5(void);
const at_family *at_tag25h9(void);
const at_family *at_tag36h10(void);
const at_family *at_tag36h11(void);
const at_family *at_tagCircle21h7(void);
const at_family *at_tagCircle49h12(void);
const at_family *at_tagCustom48h12(void);
const at_family *at_tagStandard41h12(void);
const at_family *at_tagStandard52h13(void);
const at_family *at_family_by_name(const char *name);

/* fast/apriltag_fast.h:136-141 */
bool at_read_pnm(const char *path, int64_t *width, int64_t *height, uint8_t **pixels);
bool at_read_pnm_into(const char *path, int64_t *width, int64_t *height, uint8_t **pixels, int64_t *cap);
bool at_write_pgm(const char *path, const uint8_t *pixels, int64_t width, int64_t height);

#ifdef __cplusplus
}
#endif

#endif
