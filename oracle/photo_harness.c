/* Test infrastructure (the checker), not the product: the reference's own
 * JPEG decoder (common/pjpeg.c, pjpeg-idct.c), compiled unchanged into
 * oracle/_ref/libjpegref.so, to turn the photos of the reference's tests
 * (test/data/*.jpg) into the exact gray pixels its tests detect in
 * (pjpeg_to_u8_baseline, as test/test_detection.c:74-76 does).
 *
 * ref_jpeg_gray: decodes path; *w, *h its size; writes w*h pixels (rows
 * packed) into out when cap is enough. Returns 0, or -1. */
#include <stdint.h>
#include <string.h>

#include "common/pjpeg.h"

int ref_jpeg_gray(const char *path, uint8_t *out, int64_t cap, int *w, int *h) {
    int err = 0;
    pjpeg_t *pj = pjpeg_create_from_file(path, 0, &err);
    if (!pj || err) return -1;
    image_u8_t *im = pjpeg_to_u8_baseline(pj);
    pjpeg_destroy(pj);
    if (!im) return -1;
    *w = im->width;
    *h = im->height;
    if (out && cap >= (int64_t)im->width * im->height)
        for (int y = 0; y < im->height; y++) memcpy(out + (int64_t)y * im->width, im->buf + (int64_t)y * im->stride, im->width);
    image_u8_destroy(im);
    return 0;
}
