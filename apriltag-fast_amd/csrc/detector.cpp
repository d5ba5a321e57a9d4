// The native at_* API (include/apriltag_fast.h) over the atg C-ABI: the
// detector keeps its families and parameters as the reference's does
// (fast/detector.c:4-64), and at_detect runs the HIP chain on the
// detector's device context, created on first use.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <chrono>

#include "atg.h"

struct at_internal {
    std::vector<at_family> fams;
    std::vector<int32_t> bits;
    bool fams_dirty = true;
    atg_ctx *ctx = nullptr;
    std::vector<at_detection> dets;
};

extern "C" at_detector *at_detector_create(void) {
    at_detector *td = (at_detector *)calloc(1, sizeof *td);
    if (!td) abort();
    // fast/detector.c:8-20
    td->nthreads = 1;
    td->quad_decimate = 2.0f;
    td->quad_sigma = 0.0f;
    td->refine_edges = true;
    td->decode_sharpening = 0.25;
    td->debug = false;
    td->qtp.min_cluster_pixels = 24;
    td->qtp.max_nmaxima = 10;
    td->qtp.critical_rad = 0.0f;
    td->qtp.cos_critical_rad = (float)cos(10.0 * M_PI / 180.0);
    td->qtp.max_line_fit_mse = 10.0f;
    td->qtp.min_white_black_diff = 5;
    td->qtp.deglitch = false;
    td->in = new at_internal();
    return td;
}

extern "C" void at_shutdown(at_detector *td) {
    if (!td) return;
    atg_ctx_destroy(td->in->ctx);
    td->in->ctx = nullptr;
    td->in->fams_dirty = true;
}

extern "C" void at_detector_destroy(at_detector *td) {
    if (!td) return;
    at_shutdown(td);
    delete td->in;
    free(td);
}

extern "C" bool at_add_family(at_detector *td, const at_family *fam, int bits_corrected) {
    at_internal *in = td->in;
    // fast/detector.c:45-54
    if (bits_corrected < 0 || bits_corrected > 3) return false;
    if (in->fams.size() == 16) return false;
    if (fam->ncodes >= 65536 || fam->nbits <= 0 || fam->nbits > 64) return false;
    in->fams.push_back(*fam);
    in->bits.push_back(bits_corrected);
    in->fams_dirty = true;
    return true;
}

extern "C" void at_clear_families(at_detector *td) {
    td->in->fams.clear();
    td->in->bits.clear();
    td->in->fams_dirty = true;
}

extern "C" const char *at_family_name(const at_detector *td, const at_detection *d) {
    return td->in->fams[(size_t)d->family].name;
}

// td->timing (fast/apriltag_fast.h:67-69) from the device times of the
// chain's kernels (named group.kernel): each reference stage is the sum of
// the groups that do its work.
static void fill_timing(atg_ctx *ctx, at_timing *t) {
    static const struct {
        const char *group;
        double at_timing::*field;
    } map[] = {{"decimate", &at_timing::decimate},   {"blur", &at_timing::blur},
               {"threshold_rle", &at_timing::threshold}, {"uf_edges", &at_timing::unionfind},
               {"uf_front", &at_timing::unionfind},   {"uf_replay", &at_timing::unionfind},
               {"uf_stitch", &at_timing::unionfind},  {"resolve", &at_timing::unionfind},
               {"labels", &at_timing::clusters},      {"points", &at_timing::clusters},
               {"cluster_sort", &at_timing::clusters}, {"clusters", &at_timing::clusters},
               {"fit_prep", &at_timing::fit_quads},   {"angle_sort", &at_timing::fit_quads},
               {"moments", &at_timing::fit_quads},    {"corners", &at_timing::fit_quads},
               {"decode", &at_timing::decode},        {"finish", &at_timing::reconcile}};
    const char *names[64];
    double ms[64];
    int n = atg_stage_times(ctx, names, ms, 64);
    for (int i = 0; i < n && i < 64; i++) {
        const char *dot = strchr(names[i], '.');
        const size_t len = dot ? (size_t)(dot - names[i]) : strlen(names[i]);
        for (const auto &m : map)
            if (strlen(m.group) == len && strncmp(names[i], m.group, len) == 0) t->*m.field += ms[i] * 1e-3;
    }
}

extern "C" void at_detect(at_detector *td, const uint8_t *pixels, int64_t width, int64_t height, int64_t stride) {
    at_internal *in = td->in;
    auto t0 = std::chrono::steady_clock::now();
    in->dets.clear();
    td->detections = in->dets.data();
    td->ndetections = 0;
    td->nquads = 0;
    memset(&td->timing, 0, sizeof td->timing);
    // fast/detector.c:254-255
    if (in->fams.empty()) return;
    if (width < 8 || height < 8) return;
    if (!in->ctx) {
        if (atg_ctx_create(0, &in->ctx) != ATG_OK) {
            fprintf(stderr, "apriltag: no GPU context\n");
            abort();
        }
        atg_set_profiling(in->ctx, 1);  // td->timing per stage
    }
    if (in->fams_dirty) {
        std::vector<const at_family *> fp;
        for (auto &f : in->fams) fp.push_back(&f);
        if (atg_set_families(in->ctx, fp.data(), in->bits.data(), (int32_t)fp.size()) != ATG_OK) {
            fprintf(stderr, "apriltag: %s\n", atg_last_error(in->ctx));
            abort();
        }
        in->fams_dirty = false;
    }
    atg_params p;
    p.quad_decimate = td->quad_decimate;
    p.quad_sigma = td->quad_sigma;
    p.refine_edges = td->refine_edges;
    p.decode_sharpening = td->decode_sharpening;
    p.min_cluster_pixels = td->qtp.min_cluster_pixels;
    p.max_nmaxima = td->qtp.max_nmaxima;
    p.cos_critical_rad = td->qtp.cos_critical_rad;
    p.max_line_fit_mse = td->qtp.max_line_fit_mse;
    p.min_white_black_diff = td->qtp.min_white_black_diff;
    p.deglitch = td->qtp.deglitch;
    atg_result r;
    if (atg_detect(in->ctx, &p, pixels, width, height, stride, 0, &r) != ATG_OK) {
        fprintf(stderr, "apriltag: %s\n", atg_last_error(in->ctx));
        abort();
    }
    in->dets.assign(r.dets, r.dets + r.ndets);
    td->detections = in->dets.data();
    td->ndetections = r.ndets;
    td->nquads = r.nquads;
    fill_timing(in->ctx, &td->timing);
    td->timing.total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ------------------------------------------------------------------ PNM
// (fast/image.c:178-266): binary P5/P6, color to gray as (r + 2g + b) / 4

static bool decode_pnm(const uint8_t *data, int64_t len, int64_t *width, int64_t *height, uint8_t **pixels,
                       int64_t *cap) {
    if (len < 2 || data[0] != 'P' || (data[1] != '5' && data[1] != '6')) return false;
    int64_t at = 2;
    auto num = [&]() -> int64_t {
        while (at < len) {
            uint8_t c = data[at];
            if (c == '#') {
                while (at < len && data[at] != '\n') at++;
            } else if (c == ' ' || c == '\n' || c == '\r' || c == '\t') {
                at++;
            } else {
                break;
            }
        }
        int64_t v = 0;
        while (at < len && data[at] >= '0' && data[at] <= '9') v = v * 10 + (data[at++] - '0');
        return v;
    };
    int64_t w = num(), h = num(), maxval = num();
    at++;
    if (w <= 0 || h <= 0 || maxval != 255) return false;
    int64_t ch = data[1] == '5' ? 1 : 3;
    if (len - at < w * h * ch) return false;
    if (!*pixels || *cap < w * h) {
        uint8_t *px = (uint8_t *)aligned_alloc(64, (size_t)((w * h + 63) / 64 * 64));
        if (!px) return false;
        free(*pixels);
        *pixels = px;
        *cap = w * h;
    }
    if (ch == 1) {
        memcpy(*pixels, data + at, (size_t)(w * h));
    } else {
        for (int64_t i = 0; i < w * h; i++) {
            int r = data[at + 3 * i], g = data[at + 3 * i + 1], b = data[at + 3 * i + 2];
            (*pixels)[i] = (uint8_t)((r + g + g + b) / 4);
        }
    }
    *width = w;
    *height = h;
    return true;
}

extern "C" bool at_read_pnm_into(const char *path, int64_t *width, int64_t *height, uint8_t **pixels, int64_t *cap) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    bool ok = false;
    std::vector<uint8_t> data;
    if (fseek(f, 0, SEEK_END) == 0) {
        long len = ftell(f);
        if (len > 0 && fseek(f, 0, SEEK_SET) == 0) {
            data.resize((size_t)len);
            if (fread(data.data(), 1, (size_t)len, f) == (size_t)len)
                ok = decode_pnm(data.data(), len, width, height, pixels, cap);
        }
    }
    fclose(f);
    return ok;
}

extern "C" bool at_read_pnm(const char *path, int64_t *width, int64_t *height, uint8_t **pixels) {
    int64_t cap = 0;
    *pixels = nullptr;
    return at_read_pnm_into(path, width, height, pixels, &cap);
}

extern "C" bool at_write_pgm(const char *path, const uint8_t *pixels, int64_t width, int64_t height) {
    FILE *f = fopen(path, "wb");
    if (!f) return false;
    fprintf(f, "P5\n%lld %lld\n255\n", (long long)width, (long long)height);
    bool ok = fwrite(pixels, 1, (size_t)(width * height), f) == (size_t)(width * height);
    return fclose(f) == 0 && ok;
}
