/* The upstream AprilTag C API (apriltag.h of the reference, :73-300), over
 * the MI355X detector: the same types, struct layouts and functions, so
 * that code written against the upstream library builds and links against
 * this one unchanged. apriltag_detector_detect runs at_detect
 * (apriltag_fast.h) on the GPU and returns the same detections as the
 * reference's apriltag_detector_detect (apriltag.c:140-230).
 *
 * As in the reference: nedges and nsegments stay 0 and wp stays NULL;
 * families with bits_corrected above 3, or a 17th family, are removed
 * with a message on stderr at the next detect. */
#ifndef ATG_APRILTAG_H
#define ATG_APRILTAG_H

#include <pthread.h>
#include <stdbool.h>
#include <stdint.h>

#include "common/image_u8.h"
#include "common/matd.h"
#include "common/timeprofile.h"
#include "common/workerpool.h"
#include "common/zarray.h"

#ifdef __cplusplus
extern "C" {
#endif

/* apriltag.h:73-100 */
typedef struct apriltag_family apriltag_family_t;
struct apriltag_family {
    uint32_t ncodes;
    uint64_t *codes;
    int width_at_border;
    int total_width;
    bool reversed_border;
    uint32_t nbits;
    uint32_t *bit_x;
    uint32_t *bit_y;
    uint32_t h;
    char *name;
    void *impl;
};

/* apriltag.h:103-133 */
struct apriltag_quad_thresh_params {
    int min_cluster_pixels;
    int max_nmaxima;
    float critical_rad;
    float cos_critical_rad;
    float max_line_fit_mse;
    int min_white_black_diff;
    int deglitch;
};

struct at_detector;

/* apriltag.h:139-210 (160 bytes on x86-64) */
typedef struct apriltag_detector apriltag_detector_t;
struct apriltag_detector {
    int nthreads;
    float quad_decimate;
    float quad_sigma;
    bool refine_edges;
    double decode_sharpening;
    bool debug;
    struct apriltag_quad_thresh_params qtp;
    timeprofile_t *tp;
    uint32_t nedges;
    uint32_t nsegments;
    uint32_t nquads;
    zarray_t *tag_families;
    workerpool_t *wp;
    pthread_mutex_t mutex;
    struct at_detector *fast;
    zarray_t *family_bits;
    bool families_changed;
};

/* apriltag.h:215-250 (112 bytes) */
typedef struct apriltag_detection apriltag_detection_t;
struct apriltag_detection {
    apriltag_family_t *family;
    int id;
    int hamming;
    float decision_margin;
    matd_t *H;
    double c[2];
    double p[4][2];
};

apriltag_detector_t *apriltag_detector_create(void);
void apriltag_detector_add_family_bits(apriltag_detector_t *td, apriltag_family_t *fam, int bits_corrected);
static inline void apriltag_detector_add_family(apriltag_detector_t *td, apriltag_family_t *fam) {
    apriltag_detector_add_family_bits(td, fam, 2);
}
void apriltag_detector_remove_family(apriltag_detector_t *td, apriltag_family_t *fam);
void apriltag_detector_clear_families(apriltag_detector_t *td);
void apriltag_detector_destroy(apriltag_detector_t *td);
/* A new zarray of apriltag_detection_t*, sorted by id; free it with
 * apriltag_detections_destroy. The image is read in place. */
zarray_t *apriltag_detector_detect(apriltag_detector_t *td, image_u8_t *im_orig);
void apriltag_detection_destroy(apriltag_detection_t *det);
void apriltag_detections_destroy(zarray_t *detections);
void apriltag_detection_copy(apriltag_detection_t *src, apriltag_detection_t *dst);
zarray_t *apriltag_detections_copy(zarray_t *detections);
/* The parameters of src, no families, a detector of its own. */
apriltag_detector_t *apriltag_detector_copy(apriltag_detector_t *src);
/* The tag image of code idx: total_width squared, 1-pixel cells. */
image_u8_t *apriltag_to_image(apriltag_family_t *fam, uint32_t idx);

#ifdef __cplusplus
}
#endif

#endif
