// The upstream AprilTag API (include/apriltag.h) over the native at_* API:
// the detector keeps the upstream families and bits in zarrays, hands them
// to its at_detector when they changed, runs at_detect on the image in
// place and turns the at_detection records into apriltag_detection_t
// (apriltag.c:18-230 of the reference); plus the nine tag*_create
// families, matd_t and image_u8_t as those functions need them.
#include <pthread.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <utility>
#include <vector>

#include "apriltag.h"
#include "apriltag_fast.h"

namespace {

// The native family record of an upstream family: the code list is shared
// (the caller keeps it alive, as for at_add_family), the bit positions are
// copied into the record's fixed arrays. False for a layout the native
// record cannot hold (more than 64 bits).
bool native_family(const apriltag_family_t &up, at_family *out) {
    if (up.nbits > 64) return false;
    at_family f{};
    snprintf(f.name, sizeof f.name, "%s", up.name ? up.name : "");
    f.codes = up.codes;
    f.ncodes = up.ncodes;
    f.nbits = (int32_t)up.nbits;
    f.h = (int32_t)up.h;
    f.width_at_border = up.width_at_border;
    f.total_width = up.total_width;
    f.reversed_border = up.reversed_border;
    std::copy(up.bit_x, up.bit_x + up.nbits, f.bit_x);
    std::copy(up.bit_y, up.bit_y + up.nbits, f.bit_y);
    *out = f;
    return true;
}

// Hands the upstream family list to the native detector. A family it
// refuses (bits_corrected outside 0..3, a 17th family, a layout it cannot
// hold) leaves both upstream lists, with a line on stderr, so that a
// detection's family index names the same family on both sides.
void push_families(apriltag_detector_t *td) {
    std::vector<apriltag_family_t *> fams;
    std::vector<int> bits;
    for (int k = 0; k < zarray_size(td->tag_families); k++) {
        apriltag_family_t *up;
        int b;
        zarray_get(td->tag_families, k, &up);
        zarray_get(td->family_bits, k, &b);
        fams.push_back(up);
        bits.push_back(b);
    }
    at_clear_families(td->fast);
    zarray_clear(td->tag_families);
    zarray_clear(td->family_bits);
    for (size_t k = 0; k < fams.size(); k++) {
        at_family f;
        if (!native_family(*fams[k], &f) || !at_add_family(td->fast, &f, bits[k])) {
            fprintf(stderr, "apriltag: cannot add family %s (bits_corrected %d), removed\n", fams[k]->name, bits[k]);
            continue;
        }
        zarray_add(td->tag_families, &fams[k]);
        zarray_add(td->family_bits, &bits[k]);
    }
    td->families_changed = false;
}

// td->tp after a detection: one stamp per stage of the upstream detector's
// profile, in its order and with its names, each at the start time plus
// the stage times so far.
void stamp_profile(timeprofile_t *tp, const at_timing &t) {
    const std::pair<const char *, double> stages[] = {
        {"init", 0.0},
        {"decimate", t.decimate},
        {"blur/sharp", t.blur},
        {"threshold", t.threshold},
        {"unionfind", t.unionfind},
        {"make clusters", t.clusters},
        {"fit quads to clusters", t.fit_quads},
        {"decode+refinement", t.decode},
        {"reconcile", t.reconcile},
    };
    zarray_clear(tp->stamps);
    int64_t at = tp->utime;
    for (const auto &st : stages) {
        at += (int64_t)(st.second * 1e6);
        struct timeprofile_entry e{};
        snprintf(e.name, sizeof e.name, "%s", st.first);
        e.utime = at;
        zarray_add(tp->stamps, &e);
    }
}

// The native detector's parameters from the upstream detector's.
void copy_params(const apriltag_detector_t &up, at_detector *fd) {
    fd->nthreads = std::max(up.nthreads, 1);
    fd->quad_decimate = up.quad_decimate;
    fd->quad_sigma = up.quad_sigma;
    fd->refine_edges = up.refine_edges;
    fd->decode_sharpening = up.decode_sharpening;
    fd->debug = up.debug;
    fd->qtp.min_cluster_pixels = up.qtp.min_cluster_pixels;
    fd->qtp.max_nmaxima = up.qtp.max_nmaxima;
    fd->qtp.critical_rad = up.qtp.critical_rad;
    fd->qtp.cos_critical_rad = up.qtp.cos_critical_rad;
    fd->qtp.max_line_fit_mse = up.qtp.max_line_fit_mse;
    fd->qtp.min_white_black_diff = up.qtp.min_white_black_diff;
    fd->qtp.deglitch = up.qtp.deglitch != 0;
}

// A new upstream detection record (calloc'd, H a new 3x3 matd_t) for a
// native one; the family pointer is borrowed from the detector's list.
apriltag_detection_t *upstream_detection(const at_detection &d, zarray_t *families) {
    apriltag_detection_t *det = (apriltag_detection_t *)calloc(1, sizeof(apriltag_detection_t));
    zarray_get(families, d.family, &det->family);
    det->id = d.id;
    det->hamming = d.hamming;
    det->decision_margin = d.decision_margin;
    det->H = matd_create_data(3, 3, d.H);
    std::copy(d.c, d.c + 2, det->c);
    std::copy(&d.p[0][0], &d.p[0][0] + 8, &det->p[0][0]);
    return det;
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
    zarray_t *out = zarray_create(sizeof(apriltag_detection_t *));
    // no families, or an image under 8x8: no detections
    const bool runs = zarray_size(td->tag_families) > 0 && im_orig->width >= 8 && im_orig->height >= 8;
    if (!runs) return out;
    if (td->families_changed) push_families(td);
    at_detector *fd = td->fast;
    copy_params(*td, fd);
    td->tp->utime = utime_now();
    at_detect(fd, im_orig->buf, im_orig->width, im_orig->height, im_orig->stride);
    stamp_profile(td->tp, fd->timing);
    td->nquads = (uint32_t)fd->nquads;
    td->nedges = 0;
    td->nsegments = 0;
    zarray_ensure_capacity(out, (int)fd->ndetections);
    for (const at_detection *d = fd->detections; d != fd->detections + fd->ndetections; ++d) {
        apriltag_detection_t *det = upstream_detection(*d, td->tag_families);
        zarray_add(out, &det);
    }
    return out;
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
    matd_t *h = matd_copy(src->H);  // dst's own H is replaced
    matd_destroy(dst->H);
    *dst = *src;
    dst->H = h;
}

zarray_t *apriltag_detections_copy(zarray_t *detections) {
    const int n = zarray_size(detections);
    zarray_t *out = zarray_create(sizeof(apriltag_detection_t *));
    zarray_ensure_capacity(out, n);
    for (int k = 0; k < n; k++) {
        apriltag_detection_t *src, *dst = (apriltag_detection_t *)calloc(1, sizeof(apriltag_detection_t));
        zarray_get(detections, k, &src);
        apriltag_detection_copy(src, dst);
        zarray_add(out, &dst);
    }
    return out;
}

// The tag image of code idx, one pixel per cell, total_width square, black
// but for: the one-cell white ring just inside where the black border
// ends (for a normal-border family, two cells wider than the data square:
// the white quiet zone; for a reversed-border family, the data square's own
// outline), drawn as four strokes of width-1 cells, one from each corner
// of the ring going clockwise; and the white data bits, bit i (most
// significant first) at (bit_x[i], bit_y[i]) from the data square's corner.
image_u8_t *apriltag_to_image(apriltag_family_t *fam, uint32_t idx) {
    if (!fam || idx >= fam->ncodes) return nullptr;
    const int tw = fam->total_width;
    image_u8_t *im = image_u8_create((unsigned)tw, (unsigned)tw);
    auto white = [&](int x, int y) { im->buf[y * im->stride + x] = 255; };
    const int ring = fam->width_at_border + (fam->reversed_border ? 0 : 2);
    const int lo = (tw - ring) / 2, hi = tw - 1 - lo;
    // strokes: start x, start y, step x, step y
    const int strokes[4][4] = {{lo, lo, 1, 0}, {hi, lo, 0, 1}, {lo + 1, hi, 1, 0}, {lo, lo + 1, 0, 1}};
    for (const auto &st : strokes)
        for (int k = 0; k < ring - 1; k++) white(st[0] + k * st[2], st[1] + k * st[3]);
    const int data = (tw - fam->width_at_border) / 2;
    const uint64_t code = fam->codes[idx];
    for (uint32_t i = 0; i < fam->nbits; i++)
        if ((code >> (fam->nbits - 1 - i)) & 1u) white(data + (int)fam->bit_x[i], data + (int)fam->bit_y[i]);
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
    if (rows == 0 || cols == 0) return matd_create_scalar(0.0);
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
