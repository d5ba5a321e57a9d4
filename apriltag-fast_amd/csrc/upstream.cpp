// The upstream AprilTag API (include/apriltag.h) over the native at_* API:
// the detector keeps the upstream families and bits in zarrays, hands them
// to its at_detector when they changed, runs at_detect on the image in
// place and turns the at_detection records into apriltag_detection_t
// (apriltag.c:18-230 of the reference); plus the nine tag*_create
// families, matd_t and image_u8_t as those functions need them.
#include <pthread.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

#include "apriltag.h"
#include "apriltag_fast.h"

namespace {

// apriltag_family_t -> at_family (fast/families.c:19-35): the codes are
// borrowed, the bit layout copied.
bool to_fast_family(const apriltag_family_t *src, at_family *dst) {
    if (src->nbits > 64) return false;
    memset(dst, 0, sizeof *dst);
    strncpy(dst->name, src->name ? src->name : "", sizeof dst->name - 1);
    dst->codes = src->codes;
    dst->ncodes = src->ncodes;
    dst->nbits = (int32_t)src->nbits;
    dst->h = (int32_t)src->h;
    dst->width_at_border = src->width_at_border;
    dst->total_width = src->total_width;
    dst->reversed_border = src->reversed_border;
    for (uint32_t i = 0; i < src->nbits; i++) {
        dst->bit_x[i] = (int32_t)src->bit_x[i];
        dst->bit_y[i] = (int32_t)src->bit_y[i];
    }
    return true;
}

// The families to the native detector again; one it refuses (bits above
// 3, a 17th family) is removed so that family indices stay aligned
// (apriltag.c:97-120).
void sync_families(apriltag_detector_t *td) {
    at_clear_families(td->fast);
    for (int i = 0; i < zarray_size(td->tag_families);) {
        apriltag_family_t *fam;
        int bits;
        zarray_get(td->tag_families, i, &fam);
        zarray_get(td->family_bits, i, &bits);
        at_family f;
        if (to_fast_family(fam, &f) && at_add_family(td->fast, &f, bits)) {
            i++;
            continue;
        }
        fprintf(stderr, "apriltag: cannot add family %s (bits_corrected %d), removed\n", fam->name, bits);
        zarray_remove_index(td->tag_families, i, 0);
        zarray_remove_index(td->family_bits, i, 0);
    }
    td->families_changed = false;
}

// td->tp as the reference fills it (apriltag.c:123-138): one stamp per
// stage time of the native detector.
void fill_timeprofile(timeprofile_t *tp, const at_timing *t) {
    int64_t u = tp->utime;
    zarray_clear(tp->stamps);
    const char *names[] = {"init", "decimate", "blur/sharp", "threshold", "unionfind", "make clusters",
                           "fit quads to clusters", "decode+refinement", "reconcile"};
    const double secs[] = {0.0, t->decimate, t->blur, t->threshold, t->unionfind, t->clusters, t->fit_quads,
                           t->decode, t->reconcile};
    for (size_t i = 0; i < sizeof secs / sizeof secs[0]; i++) {
        u += (int64_t)(secs[i] * 1e6);
        struct timeprofile_entry e;
        memset(&e, 0, sizeof e);
        strncpy(e.name, names[i], sizeof e.name - 1);
        e.utime = u;
        zarray_add(tp->stamps, &e);
    }
}

apriltag_family_t *family_from_builtin(const at_family *f) {
    if (!f) return nullptr;
    apriltag_family_t *tf = (apriltag_family_t *)calloc(1, sizeof(apriltag_family_t));
    tf->name = strdup(f->name);
    tf->h = (uint32_t)f->h;
    tf->ncodes = (uint32_t)f->ncodes;
    tf->codes = const_cast<uint64_t *>(f->codes);  // the embedded table, shared
    tf->nbits = (uint32_t)f->nbits;
    tf->width_at_border = f->width_at_border;
    tf->total_width = f->total_width;
    tf->reversed_border = f->reversed_border;
    tf->bit_x = (uint32_t *)calloc((size_t)f->nbits, sizeof(uint32_t));
    tf->bit_y = (uint32_t *)calloc((size_t)f->nbits, sizeof(uint32_t));
    for (int i = 0; i < f->nbits; i++) {
        tf->bit_x[i] = (uint32_t)f->bit_x[i];
        tf->bit_y[i] = (uint32_t)f->bit_y[i];
    }
    return tf;
}

void family_destroy(apriltag_family_t *tf) {
    if (!tf) return;
    free(tf->bit_x);
    free(tf->bit_y);
    free(tf->name);
    free(tf);
}

image_u8_t *new_image(unsigned int w, unsigned int h, unsigned int stride) {
    image_u8_t tmp = {(int32_t)w, (int32_t)h, (int32_t)stride, (uint8_t *)calloc((size_t)h * stride, 1)};
    image_u8_t *im = (image_u8_t *)malloc(sizeof(image_u8_t));
    memcpy(im, &tmp, sizeof tmp);
    return im;
}

}  // namespace

extern "C" {

apriltag_detector_t *apriltag_detector_create(void) {
    // apriltag.c:18-40: the upstream defaults, the native detector's own
    apriltag_detector_t *td = (apriltag_detector_t *)calloc(1, sizeof(apriltag_detector_t));
    td->nthreads = 1;
    td->quad_decimate = 2.0f;
    td->quad_sigma = 0.0f;
    td->qtp.max_nmaxima = 10;
    td->qtp.min_cluster_pixels = 24;
    td->qtp.max_line_fit_mse = 10.0f;
    td->qtp.cos_critical_rad = (float)cos(10 * M_PI / 180);
    td->qtp.deglitch = false;
    td->qtp.min_white_black_diff = 5;
    td->tag_families = zarray_create(sizeof(apriltag_family_t *));
    td->family_bits = zarray_create(sizeof(int));
    pthread_mutex_init(&td->mutex, nullptr);
    td->tp = timeprofile_create();
    td->refine_edges = true;
    td->decode_sharpening = 0.25;
    td->debug = false;
    td->fast = at_detector_create();
    return td;
}

void apriltag_detector_destroy(apriltag_detector_t *td) {
    if (!td) return;
    at_detector_destroy(td->fast);
    timeprofile_destroy(td->tp);
    zarray_destroy(td->tag_families);
    zarray_destroy(td->family_bits);
    pthread_mutex_destroy(&td->mutex);
    free(td);
}

apriltag_detector_t *apriltag_detector_copy(apriltag_detector_t *src) {
    apriltag_detector_t *dst = (apriltag_detector_t *)malloc(sizeof(apriltag_detector_t));
    memcpy(dst, src, sizeof *dst);
    dst->tag_families = zarray_create(sizeof(apriltag_family_t *));
    dst->family_bits = zarray_create(sizeof(int));
    dst->tp = timeprofile_create();
    dst->wp = nullptr;
    pthread_mutex_init(&dst->mutex, nullptr);
    dst->fast = at_detector_create();
    dst->families_changed = false;
    return dst;
}

void apriltag_detector_add_family_bits(apriltag_detector_t *td, apriltag_family_t *fam, int bits_corrected) {
    zarray_add(td->tag_families, &fam);
    zarray_add(td->family_bits, &bits_corrected);
    td->families_changed = true;
}

void apriltag_detector_remove_family(apriltag_detector_t *td, apriltag_family_t *fam) {
    for (int i = 0; i < zarray_size(td->tag_families); i++) {
        apriltag_family_t *f;
        zarray_get(td->tag_families, i, &f);
        if (f != fam) continue;
        zarray_remove_index(td->tag_families, i, 0);
        zarray_remove_index(td->family_bits, i, 0);
        td->families_changed = true;
        return;
    }
}

void apriltag_detector_clear_families(apriltag_detector_t *td) {
    zarray_clear(td->tag_families);
    zarray_clear(td->family_bits);
    td->families_changed = true;
}

zarray_t *apriltag_detector_detect(apriltag_detector_t *td, image_u8_t *im_orig) {
    zarray_t *detections = zarray_create(sizeof(apriltag_detection_t *));
    if (zarray_size(td->tag_families) == 0) return detections;      // apriltag.c:156-159
    if (im_orig->width < 8 || im_orig->height < 8) return detections;  // apriltag.c:160-164
    if (td->families_changed) sync_families(td);
    at_detector *fd = td->fast;
    fd->nthreads = td->nthreads > 0 ? td->nthreads : 1;
    fd->quad_decimate = td->quad_decimate;
    fd->quad_sigma = td->quad_sigma;
    fd->refine_edges = td->refine_edges;
    fd->decode_sharpening = td->decode_sharpening;
    fd->debug = td->debug;
    fd->qtp.min_cluster_pixels = td->qtp.min_cluster_pixels;
    fd->qtp.max_nmaxima = td->qtp.max_nmaxima;
    fd->qtp.critical_rad = td->qtp.critical_rad;
    fd->qtp.cos_critical_rad = td->qtp.cos_critical_rad;
    fd->qtp.max_line_fit_mse = td->qtp.max_line_fit_mse;
    fd->qtp.min_white_black_diff = td->qtp.min_white_black_diff;
    fd->qtp.deglitch = td->qtp.deglitch != 0;
    td->tp->utime = utime_now();
    at_detect(fd, im_orig->buf, im_orig->width, im_orig->height, im_orig->stride);
    fill_timeprofile(td->tp, &fd->timing);
    td->nquads = (uint32_t)fd->nquads;
    td->nedges = 0;
    td->nsegments = 0;
    zarray_ensure_capacity(detections, (int)fd->ndetections);
    for (int64_t i = 0; i < fd->ndetections; i++) {
        const at_detection *d = &fd->detections[i];
        apriltag_detection_t *det = (apriltag_detection_t *)calloc(1, sizeof(apriltag_detection_t));
        zarray_get(td->tag_families, d->family, &det->family);
        det->id = d->id;
        det->hamming = d->hamming;
        det->decision_margin = d->decision_margin;
        det->H = matd_create_data(3, 3, d->H);
        det->c[0] = d->c[0];
        det->c[1] = d->c[1];
        memcpy(det->p, d->p, sizeof det->p);
        zarray_add(detections, &det);
    }
    return detections;
}

void apriltag_detection_destroy(apriltag_detection_t *det) {
    if (!det) return;
    matd_destroy(det->H);
    free(det);
}

void apriltag_detections_destroy(zarray_t *detections) {
    for (int i = 0; i < zarray_size(detections); i++) {
        apriltag_detection_t *det;
        zarray_get(detections, i, &det);
        apriltag_detection_destroy(det);
    }
    zarray_destroy(detections);
}

void apriltag_detection_copy(apriltag_detection_t *src, apriltag_detection_t *dst) {
    if (dst->H) matd_destroy(dst->H);
    dst->H = matd_copy(src->H);
    dst->c[0] = src->c[0];
    dst->c[1] = src->c[1];
    memcpy(dst->p, src->p, sizeof dst->p);
    dst->id = src->id;
    dst->family = src->family;
    dst->hamming = src->hamming;
    dst->decision_margin = src->decision_margin;
}

zarray_t *apriltag_detections_copy(zarray_t *detections) {
    zarray_t *out = zarray_create(sizeof(apriltag_detection_t *));
    for (int i = 0; i < zarray_size(detections); i++) {
        apriltag_detection_t *det;
        zarray_get(detections, i, &det);
        apriltag_detection_t *c = (apriltag_detection_t *)calloc(1, sizeof(apriltag_detection_t));
        apriltag_detection_copy(det, c);
        zarray_add(out, &c);
    }
    return out;
}

// apriltag.c:263-297: the white ring inside the border, then the code
// bits, 1 pixel per cell; the rest black.
image_u8_t *apriltag_to_image(apriltag_family_t *fam, uint32_t idx) {
    if (!fam || idx >= fam->ncodes) return nullptr;
    const int tw = fam->total_width;
    image_u8_t *im = image_u8_create((unsigned)tw, (unsigned)tw);
    const uint64_t code = fam->codes[idx];
    const int ww = fam->width_at_border + (fam->reversed_border ? 0 : 2);
    const int w0 = (tw - ww) / 2, w1 = tw - 1 - w0;
    for (int i = 0; i < ww - 1; i++) {
        im->buf[w0 * im->stride + w0 + i] = 255;             // top, from the left
        im->buf[(w0 + i) * im->stride + w1] = 255;           // right, from the top
        im->buf[w1 * im->stride + w0 + i + 1] = 255;         // bottom, from the second
        im->buf[(w0 + 1 + i) * im->stride + w0] = 255;       // left, from the second
    }
    const int b0 = (tw - fam->width_at_border) / 2;
    for (uint32_t i = 0; i < fam->nbits; i++)
        if (code & ((uint64_t)1 << (fam->nbits - i - 1)))
            im->buf[((int)fam->bit_y[i] + b0) * im->stride + (int)fam->bit_x[i] + b0] = 255;
    return im;
}

// ---------------------------------------------------------------- families

#define ATG_FAMILY(name, fn)                                                      \
    apriltag_family_t *name##_create(void) { return family_from_builtin(fn()); } \
    void name##_destroy(apriltag_family_t *tf) { family_destroy(tf); }
ATG_FAMILY(tag16h5, at_tag16h5)
ATG_FAMILY(tag25h9, at_tag25h9)
ATG_FAMILY(tag36h10, at_tag36h10)
ATG_FAMILY(tag36h11, at_tag36h11)
ATG_FAMILY(tagCircle21h7, at_tagCircle21h7)
ATG_FAMILY(tagCircle49h12, at_tagCircle49h12)
ATG_FAMILY(tagCustom48h12, at_tagCustom48h12)
ATG_FAMILY(tagStandard41h12, at_tagStandard41h12)
ATG_FAMILY(tagStandard52h13, at_tagStandard52h13)
#undef ATG_FAMILY

// ---------------------------------------------------------------- matd_t

matd_t *matd_create(int rows, int cols) {
    matd_t *m = (matd_t *)calloc(1, sizeof(matd_t));
    m->nrows = (unsigned)rows;
    m->ncols = (unsigned)cols;
    m->data = (double *)calloc((size_t)(rows * cols > 0 ? rows * cols : 1), sizeof(double));
    return m;
}

matd_t *matd_create_data(int rows, int cols, const double *data) {
    matd_t *m = matd_create(rows, cols);
    memcpy(m->data, data, sizeof(double) * (size_t)rows * (size_t)cols);
    return m;
}

matd_t *matd_copy(const matd_t *m) {
    if (!m) return nullptr;
    return matd_create_data((int)m->nrows, (int)m->ncols, m->data);
}

matd_t *matd_identity(int dim) {
    matd_t *m = matd_create(dim, dim);
    for (int i = 0; i < dim; i++) MATD_EL(m, i, i) = 1.0;
    return m;
}

void matd_destroy(matd_t *m) {
    if (!m) return;
    free(m->data);
    free(m);
}

// ---------------------------------------------------------------- image_u8_t

image_u8_t *image_u8_create_stride(unsigned int width, unsigned int height, unsigned int stride) {
    return new_image(width, height, stride);
}

image_u8_t *image_u8_create_alignment(unsigned int width, unsigned int height, unsigned int alignment) {
    unsigned int stride = width;
    if (alignment && stride % alignment) stride += alignment - stride % alignment;
    return new_image(width, height, stride);
}

image_u8_t *image_u8_create(unsigned int width, unsigned int height) {
    return image_u8_create_alignment(width, height, 96);  // the reference's default row alignment
}

image_u8_t *image_u8_copy(const image_u8_t *in) {
    image_u8_t *im = new_image((unsigned)in->width, (unsigned)in->height, (unsigned)in->stride);
    memcpy(im->buf, in->buf, (size_t)in->height * (size_t)in->stride);
    return im;
}

image_u8_t *image_u8_create_from_pnm(const char *path) {
    int64_t w = 0, h = 0;
    uint8_t *px = nullptr;
    if (!at_read_pnm(path, &w, &h, &px)) {
        free(px);
        return nullptr;
    }
    image_u8_t *im = image_u8_create((unsigned)w, (unsigned)h);
    for (int64_t y = 0; y < h; y++) memcpy(im->buf + y * im->stride, px + y * w, (size_t)w);
    free(px);
    return im;
}

int image_u8_write_pnm(const image_u8_t *im, const char *path) {
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    fprintf(f, "P5\n%d %d\n255\n", im->width, im->height);
    bool ok = true;
    for (int y = 0; y < im->height && ok; y++)
        ok = fwrite(im->buf + (size_t)y * im->stride, 1, (size_t)im->width, f) == (size_t)im->width;
    return fclose(f) == 0 && ok ? 0 : -1;
}

void image_u8_destroy(image_u8_t *im) {
    if (!im) return;
    free(im->buf);
    free(im);
}

}  // extern "C"
