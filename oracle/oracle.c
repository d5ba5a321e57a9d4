/* CPU restatement of the reference AprilTag detector, single thread.
 *
 * TEST INFRASTRUCTURE (the oracle): only tests/, __graft_entry__.smoke()
 * and bench.py's CPU-baseline leg load this file's library, and only as
 * the checker. The product (the HIP path) never links or calls it.
 *
 * It restates the algorithm of the reference's single-thread path
 * (fast/detector.c:246-335 with quad_thresh_single :198-244 and
 * decode_quads :113-139) as plain scalar C, without the reference's SIMD
 * emulations: where the reference emulates an order (keysort_blocks for
 * keysort4, refine_batch for lane_weight, the bucketed quick decode), this
 * file computes the defining form directly, so agreement with the
 * reference build pins both. Floating point follows the reference's
 * operation order literally; compile with -ffp-contract=off. */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct { uint16_t start, end; uint8_t v; } ORun;
typedef struct { uint32_t parent, size; } ONode;
typedef struct { uint16_t x, y; int16_t gx, gy; } OPt;
typedef struct { float p[4][2]; bool reversed_border; double H[9]; } OQuad;

static int64_t imin(int64_t a, int64_t b) { return a < b ? a : b; }
static int64_t imax(int64_t a, int64_t b) { return a > b ? a : b; }
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

void orc_default_params(orc_params *p) {
    /* fast/detector.c:8-20 */
    memset(p, 0, sizeof *p);
    p->nthreads = 1;
    p->quad_decimate = 2.0f;
    p->quad_sigma = 0.0f;
    p->refine_edges = true;
    p->decode_sharpening = 0.25;
    p->min_cluster_pixels = 24;
    p->max_nmaxima = 10;
    p->cos_critical_rad = (float)cos(10.0 * M_PI / 180.0);
    p->max_line_fit_mse = 10.0f;
    p->min_white_black_diff = 5;
}

/* ------------------------------------------------------------------ (a) */

/* fast/image.c:6-20 */
static void dec_size(int64_t w, int64_t h, float f, int64_t *ow, int64_t *oh) {
    if (f == 1.5f) { *ow = w / 3 * 2; *oh = h / 3 * 2; return; }
    int64_t k = (int64_t)f;
    if (k == 2) { *ow = w / 2; *oh = h / 2; return; }
    *ow = 1 + (w - 1) / k;
    *oh = 1 + (h - 1) / k;
}

/* fast/image.c:26-46 (x2, double rounding for the first ow - ow%32
 * columns), :50-62 (x1.5), :83-87 (integer point sampling) */
static void decimate(const uint8_t *src, int64_t s, float f, uint8_t *out, int64_t ow, int64_t oh) {
    if (f == 1.5f) {
        for (int64_t sy = 0; sy < oh; sy += 2) {
            const uint8_t *r0 = src + sy / 2 * 3 * s, *r1 = r0 + s, *r2 = r1 + s;
            uint8_t *o0 = out + sy * ow, *o1 = o0 + ow;
            for (int64_t k = 0; k < ow / 2; k++) {
                int64_t x = 3 * k;
                int a = r0[x], b = r0[x + 1], c = r0[x + 2];
                int d = r1[x], e = r1[x + 1], ff = r1[x + 2];
                int g = r2[x], hh = r2[x + 1], i = r2[x + 2];
                o0[2 * k] = (uint8_t)((4 * a + 2 * b + 2 * d + e) / 9);
                o0[2 * k + 1] = (uint8_t)((4 * c + 2 * b + 2 * ff + e) / 9);
                o1[2 * k] = (uint8_t)((4 * g + 2 * d + 2 * hh + e) / 9);
                o1[2 * k + 1] = (uint8_t)((4 * i + 2 * ff + 2 * hh + e) / 9);
            }
        }
        return;
    }
    int64_t k = (int64_t)f;
    if (k == 2) {
        int64_t nv = ow - ow % 32;
        for (int64_t y = 0; y < oh; y++) {
            const uint8_t *a = src + 2 * y * s, *b = a + s;
            for (int64_t x = 0; x < ow; x++) {
                int p = a[2 * x], q = a[2 * x + 1], r = b[2 * x], t = b[2 * x + 1];
                if (x < nv)
                    out[y * ow + x] = (uint8_t)((((p + r + 1) >> 1) + ((q + t + 1) >> 1) + 1) >> 1);
                else
                    out[y * ow + x] = (uint8_t)((p + q + r + t + 2) >> 2);
            }
        }
        return;
    }
    for (int64_t y = 0; y < oh; y++)
        for (int64_t x = 0; x < ow; x++) out[y * ow + x] = src[y * k * s + x * k];
}

/* fast/image.c:93-108 */
static void convolve(const uint8_t *x, uint8_t *y, int64_t sz, const uint8_t *k, int64_t ksz, bool par) {
    int64_t i = 0;
    while (i < ksz / 2 && i < sz) { y[i] = x[i]; i++; }
    int64_t n = par ? sz - ksz + 1 : sz - ksz;
    for (i = 0; i < n; i++) {
        uint32_t acc = 0;
        for (int64_t j = 0; j < ksz; j++) acc += (uint32_t)k[j] * x[i + j];
        y[ksz / 2 + i] = (uint8_t)(acc >> 8);
    }
    int64_t tail = par ? sz - ksz / 2 : sz - ksz + ksz / 2;
    for (i = imax(tail, 0); i < sz; i++) y[i] = x[i];
}

/* fast/image.c:142-147 */
static int64_t blur_ksz(float sigma) {
    if (sigma == 0.0f) return 0;
    int64_t k = (int64_t)(4.0f * fabsf(sigma));
    if ((k & 1) == 0) k++;
    return k;
}

/* fast/image.c:111-139 */
static void gblur(uint8_t *im, int64_t w, int64_t h, double sigma, int64_t ksz) {
    double *dk = malloc(sizeof(double) * ksz);
    uint8_t *k = malloc(ksz);
    double acc = 0.0;
    for (int64_t i = 0; i < ksz; i++) {
        double t = (double)(-(ksz / 2) + i) / sigma;
        dk[i] = exp(-0.5 * (t * t));
    }
    for (int64_t i = 0; i < ksz; i++) acc += dk[i];
    for (int64_t i = 0; i < ksz; i++) k[i] = (uint8_t)(dk[i] / acc * 255.0);
    bool par = w * h >= 65536;
    uint8_t *tmp = malloc(imax(w, h)), *tmp2 = malloc(h);
    for (int64_t y = 0; y < h; y++) {
        memcpy(tmp, im + y * w, w);
        convolve(tmp, im + y * w, w, k, ksz, par);
    }
    for (int64_t x = 0; x < w; x++) {
        for (int64_t y = 0; y < h; y++) tmp[y] = im[y * w + x];
        convolve(tmp, tmp2, h, k, ksz, par);
        for (int64_t y = 0; y < h; y++) im[y * w + x] = tmp2[y];
    }
    free(dk); free(k); free(tmp); free(tmp2);
}

/* fast/image.c:158-173 (+ sharpen :149-154) */
static bool blur_or_sharpen(uint8_t *im, int64_t w, int64_t h, float qs) {
    int64_t ksz = blur_ksz(qs);
    if (ksz <= 1) return false;
    double sigma = fabsf(qs);
    if (qs > 0.0f) { gblur(im, w, h, sigma, ksz); return true; }
    uint8_t *orig = malloc(w * h);
    memcpy(orig, im, w * h);
    gblur(im, w, h, sigma, ksz);
    for (int64_t i = 0; i < w * h; i++) {
        int v = 2 * (int)orig[i] - im[i];
        im[i] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
    }
    free(orig);
    return true;
}

/* fast/threshold.c:109-161 with :15-102, written per pixel: 4x4 tile
 * min/max, then the 3x3 max/min over the tile grid clipped at its edge,
 * then t = lo + (hi - lo)/2. Columns right of the last whole tile use the
 * last tile and are never flat; so are rows below the last whole tile row. */
static void threshold(const uint8_t *im, int64_t w, int64_t h, int diff, uint8_t *out) {
    int64_t tw = w / 4, th = h / 4;
    uint8_t *mx = malloc(tw * th), *mn = malloc(tw * th), *hx = malloc(tw * th), *hn = malloc(tw * th);
    for (int64_t ty = 0; ty < th; ty++)
        for (int64_t tx = 0; tx < tw; tx++) {
            uint8_t hi = 0, lo = 255;
            for (int dy = 0; dy < 4; dy++)
                for (int dx = 0; dx < 4; dx++) {
                    uint8_t v = im[(4 * ty + dy) * w + 4 * tx + dx];
                    if (v > hi) hi = v;
                    if (v < lo) lo = v;
                }
            mx[ty * tw + tx] = hi;
            mn[ty * tw + tx] = lo;
        }
    for (int64_t ty = 0; ty < th; ty++)
        for (int64_t tx = 0; tx < tw; tx++) {
            uint8_t hi = 0, lo = 255;
            for (int64_t yy = imax(ty - 1, 0); yy <= imin(ty + 1, th - 1); yy++)
                for (int64_t xx = imax(tx - 1, 0); xx <= imin(tx + 1, tw - 1); xx++) {
                    if (mx[yy * tw + xx] > hi) hi = mx[yy * tw + xx];
                    if (mn[yy * tw + xx] < lo) lo = mn[yy * tw + xx];
                }
            hx[ty * tw + tx] = hi;
            hn[ty * tw + tx] = lo;
        }
    for (int64_t y = 0; y < h; y++) {
        bool yflat = y < 4 * th;
        int64_t ty = yflat ? y / 4 : th - 1;
        for (int64_t x = 0; x < w; x++) {
            bool xflat = x < 4 * tw;
            int64_t tx = xflat ? x / 4 : tw - 1;
            int hi = hx[ty * tw + tx], lo = hn[ty * tw + tx];
            uint8_t t = (uint8_t)(lo + (hi - lo) / 2);
            bool flat = yflat && xflat && hi - lo < diff;
            out[y * w + x] = flat ? 127 : (im[y * w + x] > t ? 255 : 0);
        }
    }
    free(mx); free(mn); free(hx); free(hn);
}

/* fast/threshold.c:173-193 */
static void deglitch(uint8_t *t, int64_t w, int64_t h) {
    uint8_t *tmp = calloc(w * h, 1);
    for (int64_t y = 1; y < h - 1; y++)
        for (int64_t x = 1; x < w - 1; x++) {
            uint8_t m = 0;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) if (t[(y + dy) * w + x + dx] > m) m = t[(y + dy) * w + x + dx];
            tmp[y * w + x] = m;
        }
    for (int64_t y = 1; y < h - 1; y++)
        for (int64_t x = 1; x < w - 1; x++) {
            uint8_t m = 255;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) if (tmp[(y + dy) * w + x + dx] < m) m = tmp[(y + dy) * w + x + dx];
            t[y * w + x] = m;
        }
    free(tmp);
}

/* ------------------------------------------------------------------ (b) */

/* fast/threshold.c:213-266: runs of equal values over columns [0, w-2];
 * runs of 127 are not stored. */
static int64_t rle_row(const uint8_t *row, int64_t w, ORun *out) {
    int64_t xmax = w - 2, n = 0, start = 0;
    uint8_t v = row[0];
    for (int64_t x = 1; x <= xmax; x++) {
        if (row[x] == row[x - 1]) continue;
        if (v != 127) out[n++] = (ORun){(uint16_t)start, (uint16_t)(x - 1), v};
        start = x;
        v = row[x];
    }
    if (v != 127) out[n++] = (ORun){(uint16_t)start, (uint16_t)xmax, v};
    return n;
}

static uint32_t uf_find(ONode *uf, uint32_t a) {
    uint32_t r = a;
    while (uf[r].parent != r) r = uf[r].parent;
    while (uf[a].parent != r) { uint32_t t = uf[a].parent; uf[a].parent = r; a = t; }
    return r;
}

/* fast/unionfind.c:15-45: the root of the heavier set wins (weight =
 * size + 1), ties to b's root. Path compression never changes a root. */
static void uf_union(ONode *uf, uint32_t a, uint32_t b) {
    uint32_t ra = uf_find(uf, a), rb = uf_find(uf, b);
    if (ra == rb) return;
    uint32_t as = uf[ra].size + 1, bs = uf[rb].size + 1;
    if (as > bs) { uf[rb].parent = ra; uf[ra].size += bs; }
    else { uf[ra].parent = rb; uf[rb].size += as; }
}

/* fast/unionfind.c:284-298 (1 thread: one task over rows [1, h)) with
 * connect_rows_any :112-146 and last_column :93-97. */
static void components(ONode *uf, const uint8_t *t, int64_t w, int64_t h, const ORun *runs, const uint32_t *ro) {
    int64_t vb = ro[h];
    for (int64_t y = 0; y < h; y++) uf[vb + y] = (ONode){(uint32_t)(vb + y), 0};
    for (int64_t i = ro[0]; i < ro[1]; i++) uf[i] = (ONode){(uint32_t)i, (uint32_t)(runs[i].end - runs[i].start)};
    for (int64_t y = 1; y < h; y++) {
        const ORun *prev = runs + ro[y - 1], *cur = runs + ro[y];
        int64_t np = ro[y] - ro[y - 1], nc = ro[y + 1] - ro[y];
        for (int64_t i = 0; i < nc; i++)
            uf[ro[y] + i] = (ONode){(uint32_t)(ro[y] + i), (uint32_t)(cur[i].end - cur[i].start)};
        int64_t j = 0;
        for (int64_t i = 0; i < nc; i++) {
            int64_t a0 = cur[i].start, a1 = cur[i].end;
            uint8_t v = cur[i].v;
            while (j < np && (int64_t)prev[j].end + 1 < a0) j++;
            for (int64_t k = j; k < np && prev[k].start <= a1 + 1; k++) {
                if (prev[k].v != v) continue;
                int64_t b0 = prev[k].start, b1 = prev[k].end;
                bool join = imax(imax(a0, b0), 1) <= imin(a1, b1);
                if (!join && v == 255)
                    join = imax(imax(a0, b0 + 1), 1) <= imin(a1, b1 + 1) ||
                           imax(imax(a0, b0 - 1), 1) <= imin(a1, b1 - 1);
                if (join) uf_union(uf, (uint32_t)(ro[y] + i), (uint32_t)(ro[y - 1] + k));
            }
        }
        uint8_t above_last = t[(y - 1) * w + w - 1], above_prev = t[(y - 1) * w + w - 2];
        if (nc > 0 && cur[nc - 1].end == w - 2 && cur[nc - 1].v == 255 && above_last == 255 && above_prev != 255)
            uf_union(uf, (uint32_t)(ro[y] + nc - 1), (uint32_t)(vb + y - 1));
    }
}

/* ------------------------------------------------------------------ (c) */

/* fast/clusters.c:97-103 */
static int64_t cluster_map_size(int64_t w, int64_t h) {
    int64_t chunk = 1 + (h - 1) / 6, n = 1024;
    while (n < chunk * w / 64 && n < 65536) n <<= 1;
    return n;
}

typedef struct { uint64_t key_bucket, id; int64_t seq; OPt p; } OEmit;

static int cmp_emit(const void *a, const void *b) {
    const OEmit *x = a, *y = b;
    if (x->key_bucket != y->key_bucket) return x->key_bucket < y->key_bucket ? -1 : 1;
    if (x->id != y->id) return x->id < y->id ? -1 : 1;
    return x->seq < y->seq ? -1 : x->seq > y->seq;
}

/* The labels of one row: the representative and ok flag of the run over
 * each pixel; pixels without a run get ok 0; the last column is the
 * row's virtual node (fast/clusters.c:123-142). */
static void label_row(const ORun *runs, int64_t n, const uint32_t *rep, const uint8_t *ok, uint32_t vr, uint8_t vo,
                      int64_t w, uint32_t *lab, uint8_t *okb) {
    memset(okb, 0, w);
    memset(lab, 0, w * 4);
    for (int64_t i = 0; i < n; i++)
        for (int64_t x = runs[i].start; x <= runs[i].end; x++) { lab[x] = rep[i]; okb[x] = ok[i]; }
    lab[w - 1] = vr;
    okb[w - 1] = vo;
}

static int opp(uint8_t a, uint8_t b) { return (a ^ b) == 255; }

/* fast/clusters.c:278-308 with point_masks :150-174 and emit_row :194-252:
 * each pixel x of rows 1..h-2 against its neighbors (1,0), (0,1), (-1,1),
 * (1,1), in that order; then the clusters ordered by (hash bucket, id)
 * (:338-356) with each cluster's points in emission order. */
static OEmit *gradient_points(const uint8_t *t, int64_t w, int64_t h, const ORun *runs, const uint32_t *ro,
                              const uint32_t *rep, const uint8_t *ok, int64_t *npts) {
    int64_t vb = ro[h];
    int64_t cap = 1024, n = 0;
    OEmit *em = malloc(sizeof(OEmit) * cap);
    uint64_t mask = (uint64_t)cluster_map_size(w, h) - 1;
    uint32_t *la = malloc(w * 4), *lb = malloc(w * 4);
    uint8_t *oa = malloc(w), *ob = malloc(w);
    for (int64_t y = 1; y + 1 < h; y++) {
        label_row(runs + ro[y], ro[y + 1] - ro[y], rep + ro[y], ok + ro[y], rep[vb + y], ok[vb + y], w, la, oa);
        label_row(runs + ro[y + 1], ro[y + 2] - ro[y + 1], rep + ro[y + 1], ok + ro[y + 1], rep[vb + y + 1],
                  ok[vb + y + 1], w, lb, ob);
        const uint8_t *a = t + y * w, *b = a + w;
        for (int64_t x = 0; x < w - 1; x++) {
            int m = 0;
            if (x == 0) {
                m = opp(a[0], a[1]) && oa[0] && oa[1] ? 1 : 0;
            } else if (oa[x]) {
                uint8_t v = a[x];
                int c10 = opp(v, a[x + 1]) & oa[x + 1];
                int c01 = opp(v, b[x]) & ob[x];
                int c11 = opp(v, b[x + 1]) & ob[x + 1];
                int cm1 = opp(v, b[x - 1]) & ob[x - 1];
                /* the pixel to the left gave this pair as its (1,1) point;
                   x = 1 has no pixel to the left in the C loop */
                int p11 = x > 1 ? (opp(a[x - 1], b[x]) & oa[x - 1] & ob[x]) : 0;
                m = c10 | (c01 << 1) | ((cm1 & (p11 ^ 1)) << 2) | (c11 << 3);
            }
            static const int DX[4] = {1, 0, -1, 1}, DY[4] = {0, 1, 1, 1};
            for (int j = 0; j < 4; j++) {
                if (!(m >> j & 1)) continue;
                int dx = DX[j], dy = DY[j];
                int vd = 255 - 2 * (int)a[x];
                uint32_t r0 = la[x], r1 = dy ? lb[x + dx] : la[x + dx];
                uint64_t id = r0 < r1 ? ((uint64_t)r1 << 32) + r0 : ((uint64_t)r0 << 32) + r1;
                if (n == cap) { cap *= 2; em = realloc(em, sizeof(OEmit) * cap); }
                em[n].id = id;
                em[n].key_bucket = ((id * 2654435761u) >> 32) & mask;
                em[n].seq = n;
                em[n].p = (OPt){(uint16_t)(2 * x + dx), (uint16_t)(2 * y + dy), (int16_t)(dx * vd), (int16_t)(dy * vd)};
                n++;
            }
        }
    }
    free(la); free(lb); free(oa); free(ob);
    qsort(em, n, sizeof(OEmit), cmp_emit);
    *npts = n;
    return em;
}

/* ------------------------------------------------------------------ (d) */

/* fast/quads.c:429-433 */
static uint64_t slope_key(float s) {
    uint32_t b = f2u(s), sg = (uint32_t)((int32_t)b >> 31);
    return b ^ (sg | 0x80000000u);
}

/* fast/quads.c:438-453 */
static uint64_t point_key(float dx, float dy, int64_t idx) {
    float quadrant = dy > 0.0f ? (dx > 0.0f ? 65536.0f : 131072.0f) : (dx > 0.0f ? 0.0f : -65536.0f);
    uint32_t neg = dy < 0.0f ? 0x80000000u : 0u;
    float fx = u2f(f2u(dx) ^ neg), fy = u2f(f2u(dy) ^ neg);
    float ndx = fx < 0.0f ? fy : fx;
    float ndy = fx < 0.0f ? -fx : fy;
    return (slope_key(quadrant + ndy / ndx) << 32) | (uint32_t)~(uint32_t)idx;
}

/* The angle sort: the merge sort whose order the reference reproduces
 * (fast/quads.c:42-46, keysort :161-177): leaves of at most 5 keys by the
 * compare-exchange networks of key_network_sort :56-85, which compare the
 * high words only; merges by the full key, taking the right run on ties. */
static void kswap(uint64_t *k, int a, int b) {
    if ((k[a] >> 32) > (k[b] >> 32)) { uint64_t t = k[a]; k[a] = k[b]; k[b] = t; }
}

static void knet(uint64_t *k, int64_t n) {
    if (n <= 1) return;
    if (n == 2) { kswap(k, 0, 1); return; }
    if (n == 3) { kswap(k, 0, 1); kswap(k, 1, 2); kswap(k, 0, 1); return; }
    if (n == 4) { kswap(k, 0, 1); kswap(k, 2, 3); kswap(k, 0, 2); kswap(k, 1, 3); kswap(k, 1, 2); return; }
    kswap(k, 0, 1); kswap(k, 3, 4); kswap(k, 1, 2); kswap(k, 0, 1); kswap(k, 0, 3);
    kswap(k, 2, 4); kswap(k, 1, 2); kswap(k, 2, 3); kswap(k, 1, 2);
}

static void kmerge(const uint64_t *a, int64_t na, const uint64_t *b, int64_t nb, uint64_t *out) {
    int64_t i = 0, j = 0, o = 0;
    while (i < na && j < nb) out[o++] = a[i] < b[j] ? a[i++] : b[j++];
    while (i < na) out[o++] = a[i++];
    while (j < nb) out[o++] = b[j++];
}

static void ksort(uint64_t *a, uint64_t *tmp, int64_t n) {
    if (n <= 5) { knet(a, n); return; }
    int64_t h = n / 2;
    ksort(a, tmp, h);
    ksort(a + h, tmp, n - h);
    kmerge(a, h, a + h, n - h, tmp);
    memcpy(a, tmp, n * 8);
}

typedef struct { double *mx, *my, *mxx, *mxy, *myy, *w; } OMom;

/* fast/quads.c:539-600 */
static void fit_line(OMom a, int64_t sz, int64_t i0, int64_t i1, double lp[4], double *mse) {
    double mx, my, mxx, mxy, myy, w;
    if (i0 < i1) {
        mx = a.mx[i1]; my = a.my[i1]; mxx = a.mxx[i1]; mxy = a.mxy[i1]; myy = a.myy[i1]; w = a.w[i1];
        if (i0 > 0) {
            mx -= a.mx[i0 - 1]; my -= a.my[i0 - 1]; mxx -= a.mxx[i0 - 1];
            mxy -= a.mxy[i0 - 1]; myy -= a.myy[i0 - 1]; w -= a.w[i0 - 1];
        }
    } else {
        mx = a.mx[sz - 1] - a.mx[i0 - 1]; my = a.my[sz - 1] - a.my[i0 - 1];
        mxx = a.mxx[sz - 1] - a.mxx[i0 - 1]; mxy = a.mxy[sz - 1] - a.mxy[i0 - 1];
        myy = a.myy[sz - 1] - a.myy[i0 - 1]; w = a.w[sz - 1] - a.w[i0 - 1];
        mx += a.mx[i1]; my += a.my[i1]; mxx += a.mxx[i1]; mxy += a.mxy[i1]; myy += a.myy[i1]; w += a.w[i1];
    }
    double ex = mx / w, ey = my / w;
    double cxx = mxx / w - ex * ex, cxy = mxy / w - ex * ey, cyy = myy / w - ey * ey;
    double root = (double)sqrtf((float)((cxx - cyy) * (cxx - cyy) + 4.0 * cxy * cxy));
    double es = 0.5 * (cxx + cyy - root), eig = 0.5 * (cxx + cyy + root);
    double nx1 = cxx - eig, ny1 = cxy, m1 = nx1 * nx1 + ny1 * ny1;
    double nx2 = cxy, ny2 = cyy - eig, m2 = nx2 * nx2 + ny2 * ny2;
    double nx = nx2, ny = ny2, m = m2;
    if (m1 > m2) { nx = nx1; ny = ny1; m = m1; }
    double len = (double)sqrtf((float)m);
    lp[0] = ex; lp[1] = ey; lp[2] = 0.0; lp[3] = 0.0;
    if (!(fabs(len) < 1e-12)) { lp[2] = nx / len; lp[3] = ny / len; }
    *mse = es;
}

/* The line-fit error of a window from its moment sums, as
 * window_errors/moment_errors (fast/quads.c:604-652) write it. */
static double win_err(double mx, double my, double mxx, double mxy, double myy, double w, double nn) {
    double ex = mx / w, ey = my / w;
    double cxx = mxx / w - ex * ex, cxy = mxy / w - ex * ey, cyy = myy / w - ey * ey;
    double d = cxx - cyy;
    double root = (double)sqrtf((float)(d * d + 4.0 * (cxy * cxy)));
    return nn * (0.5 * (cxx + cyy - root));
}

static double qsm_kern[7];

/* fast/quads.c:788-959 */
static bool segment_maxima(const orc_params *qp, OMom a, int64_t sz, int64_t idx[4]) {
    int64_t ksz = imin(20, sz / 12);
    if (ksz < 2) return false;
    double nn = (double)(2 * ksz + 1);
    double *errs = malloc(8 * sz), *y = malloc(8 * sz);
    for (int64_t i = 0; i < sz; i++) {
        int64_t i0 = i - ksz, i1 = i + ksz;
        double m[6];
        double *arr[6] = {a.mx, a.my, a.mxx, a.mxy, a.myy, a.w};
        for (int q = 0; q < 6; q++) {
            const double *v = arr[q];
            if (i0 < 0) m[q] = (v[sz - 1] - v[i0 + sz - 1]) + v[i1];   /* wraps at the start (:654-660) */
            else if (i1 >= sz) m[q] = (v[sz - 1] - v[i0 - 1]) + v[i1 - sz]; /* wraps at the end (:677-686) */
            else if (i0 > 0) m[q] = v[i1] - v[i0 - 1];
            else m[q] = v[i1];
        }
        errs[i] = win_err(m[0], m[1], m[2], m[3], m[4], m[5], nn);
    }
    /* 7-tap filter, wrapping (:706-735) */
    for (int64_t k = 0; k + 6 < sz; k++) {
        double acc = errs[k] * qsm_kern[0];
        for (int t = 1; t < 7; t++) acc += errs[k + t] * qsm_kern[t];
        y[k + 3] = acc;
    }
    for (int e = 0; e < 6; e++) {
        int64_t iy = e < 3 ? e : sz - 6 + e;
        double acc = 0.0;
        int64_t j = iy - 3;
        if (j < 0) j += sz;
        for (int i = 0; i < 7; i++) { acc += errs[j] * qsm_kern[i]; j++; if (j == sz) j = 0; }
        y[iy] = acc;
    }
    memcpy(errs, y, 8 * sz);
    /* cyclic local maxima */
    int64_t *mxi = malloc(8 * sz);
    double *mxe = malloc(8 * sz);
    int64_t nm = 0;
    for (int64_t i = 0; i < sz; i++) {
        double e = errs[i], l = errs[i == 0 ? sz - 1 : i - 1], r = errs[i == sz - 1 ? 0 : i + 1];
        if (e > l && e > r) { mxi[nm] = i; mxe[nm] = e; nm++; }
    }
    bool ret = false;
    if (nm < 4) goto done;
    if (nm > qp->max_nmaxima) {
        /* keep the maxima above the (max_nmaxima+1)-th largest error, in
           order (the value quickselect finds, :826-862) */
        double *s = malloc(8 * nm);
        memcpy(s, mxe, 8 * nm);
        for (int64_t i = 1; i < nm; i++) { /* sort descending */
            double v = s[i]; int64_t j = i;
            while (j > 0 && s[j - 1] < v) { s[j] = s[j - 1]; j--; }
            s[j] = v;
        }
        double th = s[qp->max_nmaxima];
        free(s);
        int64_t o = 0;
        for (int64_t i = 0; i < nm; i++) if (mxe[i] > th) { mxi[o] = mxi[i]; mxe[o] = mxe[i]; o++; }
        nm = o;
    }
    {
        double max_mse = qp->max_line_fit_mse, max_dot = qp->cos_critical_rad;
        double *perr = malloc(8 * nm * nm), *pmse = malloc(8 * nm * nm), *pnx = malloc(8 * nm * nm),
               *pny = malloc(8 * nm * nm);
        for (int64_t p = 0; p < nm; p++)
            for (int64_t q = 0; q < nm; q++) {
                bool fwd = p < q, back = q + 3 <= p;
                if (!fwd && !back) continue;
                int64_t i0 = mxi[p], i1 = mxi[q];
                double lp[4], mse;
                /* pair_lines (:750-786) is fit_line on the same sums; its error is
                   the point count times the smaller eigenvalue */
                fit_line(a, sz, i0, i1, lp, &mse);
                double cnt = fwd ? (double)(i1 - i0 + 1) : (double)(sz - i0 + i1 + 1);
                perr[p * nm + q] = cnt * mse;
                pmse[p * nm + q] = mse;
                pnx[p * nm + q] = lp[2];
                pny[p * nm + q] = lp[3];
            }
        double best = (double)INFINITY;
        int64_t bi[4] = {0, 0, 0, 0};
        for (int64_t m0 = 0; m0 < nm - 3; m0++)
            for (int64_t m1 = m0 + 1; m1 < nm - 2; m1++) {
                if (pmse[m0 * nm + m1] > max_mse) continue;
                for (int64_t m2 = m1 + 1; m2 < nm - 1; m2++) {
                    if (pmse[m1 * nm + m2] > max_mse) continue;
                    double dot = pnx[m0 * nm + m1] * pnx[m1 * nm + m2] + pny[m0 * nm + m1] * pny[m1 * nm + m2];
                    if (fabs(dot) > max_dot) continue;
                    double e012 = perr[m0 * nm + m1] + perr[m1 * nm + m2];
                    for (int64_t m3 = m2 + 1; m3 < nm; m3++) {
                        double f23 = pmse[m2 * nm + m3] > max_mse ? (double)INFINITY : perr[m2 * nm + m3];
                        double f30 = pmse[m3 * nm + m0] > max_mse ? (double)INFINITY : perr[m3 * nm + m0];
                        double err = e012 + f23 + f30;
                        if (err < best) { best = err; bi[0] = mxi[m0]; bi[1] = mxi[m1]; bi[2] = mxi[m2]; bi[3] = mxi[m3]; }
                    }
                }
            }
        free(perr); free(pmse); free(pnx); free(pny);
        if (best == (double)INFINITY) goto done;
        memcpy(idx, bi, sizeof bi);
        ret = best / (double)sz < max_mse;
    }
done:
    free(errs); free(y); free(mxi); free(mxe);
    return ret;
}

static double sq(double x) { return x * x; }

/* fast/quads.c:1128-1143 and fit_quad_in :1044-1126 */
static bool fit_quad(const orc_params *qp, const OPt *pts, int64_t sz, const uint8_t *im, int64_t iw, int64_t ih,
                     int64_t tag_width, bool normal_border, bool reversed_border, OQuad *quad) {
    uint16_t xmin = 0xffff, xmax = 0, ymin = 0xffff, ymax = 0;
    for (int64_t i = 0; i < sz; i++) {
        if (pts[i].x < xmin) xmin = pts[i].x;
        if (pts[i].x > xmax) xmax = pts[i].x;
        if (pts[i].y < ymin) ymin = pts[i].y;
        if (pts[i].y > ymax) ymax = pts[i].y;
    }
    if (((int64_t)xmax - xmin) * ((int64_t)ymax - ymin) < tag_width) return false;
    float cx = (float)((double)(xmin + xmax) * 0.5 + 0.05118);
    float cy = (float)((double)(ymin + ymax) * 0.5 + -0.028581);
    uint64_t *keys = malloc(8 * sz), *tmp = malloc(8 * sz);
    float lanes[8] = {0};
    int64_t nv = sz - sz % 8;
    for (int64_t i = 0; i < sz; i++) {
        float dx = pts[i].x - cx, dy = pts[i].y - cy;
        keys[i] = point_key(dx, dy, i);
        if (i < nv) lanes[i % 8] = lanes[i % 8] + (dx * (float)pts[i].gx + dy * (float)pts[i].gy);
    }
    float dot = ((lanes[0] + lanes[4]) + (lanes[2] + lanes[6])) + ((lanes[1] + lanes[5]) + (lanes[3] + lanes[7]));
    for (int64_t i = nv; i < sz; i++) {
        float dx = pts[i].x - cx, dy = pts[i].y - cy;
        dot += dx * pts[i].gx + dy * pts[i].gy;
    }
    bool ok = false;
    double *buf = NULL;
    quad->reversed_border = dot < 0.0f;
    if (!reversed_border && quad->reversed_border) goto out;
    if (!normal_border && !quad->reversed_border) goto out;
    ksort(keys, tmp, sz);
    /* compute_lfps (:968-1035) */
    buf = malloc(8 * 6 * sz);
    OMom a = {buf, buf + sz, buf + 2 * sz, buf + 3 * sz, buf + 4 * sz, buf + 5 * sz};
    double smx = 0, smy = 0, smxx = 0, smyy = 0, smxy = 0, sw = 0;
    for (int64_t k = 0; k < sz; k++) {
        const OPt *p = &pts[(uint32_t)~(uint32_t)keys[k]];
        int32_t ix = (p->x >> 1) + (p->x & 1), iy = (p->y >> 1) + (p->y & 1);
        int32_t g2 = 0;
        if (ix > 0 && ix + 1 < iw && iy > 0 && iy + 1 < ih) {
            int32_t c = iy * (int32_t)iw + ix;
            int32_t gx = (int32_t)im[c + 1] - im[c - 1], gy = (int32_t)im[c + iw] - im[c - iw];
            g2 = gx * gx + gy * gy;
        }
        double w = sqrt((double)g2) + 1.0;
        double fx = (double)p->x * 0.5 + 0.5, fy = (double)p->y * 0.5 + 0.5;
        smx += w * fx; smy += w * fy; smxx += w * fx * fx; smxy += w * fx * fy; smyy += w * fy * fy; sw += w;
        a.mx[k] = smx; a.my[k] = smy; a.mxx[k] = smxx; a.mxy[k] = smxy; a.myy[k] = smyy; a.w[k] = sw;
    }
    int64_t idx[4];
    if (!segment_maxima(qp, a, sz, idx)) goto out;
    double lines[4][4];
    for (int i = 0; i < 4; i++) {
        double mse;
        fit_line(a, sz, idx[i], idx[(i + 1) & 3], lines[i], &mse);
        if (mse > qp->max_line_fit_mse) goto out;
    }
    for (int i = 0; i < 4; i++) {
        int j = (i + 1) & 3;
        double a00 = lines[i][3], a01 = -lines[j][3], a10 = -lines[i][2], a11 = lines[j][2];
        double b0 = -lines[i][0] + lines[j][0], b1 = -lines[i][1] + lines[j][1];
        double det = a00 * a11 - a10 * a01;
        if (fabs(det) < 0.001) goto out;
        double w00 = a11 / det, w01 = -a01 / det;
        double l0 = w00 * b0 + w01 * b1;
        quad->p[i][0] = (float)(lines[i][0] + l0 * a00);
        quad->p[i][1] = (float)(lines[i][1] + l0 * a10);
    }
    {
        /* area (two Heron triangles) and convexity (:1094-1124); the
           coordinate differences are float */
        double area = 0.0, len[3];
        for (int i = 0; i < 3; i++) {
            int ia = i, ib = (i + 1) % 3;
            len[i] = sqrt(sq(quad->p[ib][0] - quad->p[ia][0]) + sq(quad->p[ib][1] - quad->p[ia][1]));
        }
        double p = (len[0] + len[1] + len[2]) / 2.0;
        area += sqrt(p * (p - len[0]) * (p - len[1]) * (p - len[2]));
        static const int idxs[4] = {2, 3, 0, 2};
        for (int i = 0; i < 3; i++) {
            int ia = idxs[i], ib = idxs[i + 1];
            len[i] = sqrt(sq(quad->p[ib][0] - quad->p[ia][0]) + sq(quad->p[ib][1] - quad->p[ia][1]));
        }
        p = (len[0] + len[1] + len[2]) / 2.0;
        area += sqrt(p * (p - len[0]) * (p - len[1]) * (p - len[2]));
        if (area < 0.95 * (double)tag_width * (double)tag_width) goto out;
        double ccr = qp->cos_critical_rad;
        for (int i = 0; i < 4; i++) {
            int i0 = i, i1 = (i + 1) & 3, i2 = (i + 2) & 3;
            double dx1 = quad->p[i1][0] - quad->p[i0][0], dy1 = quad->p[i1][1] - quad->p[i0][1];
            double dx2 = quad->p[i2][0] - quad->p[i1][0], dy2 = quad->p[i2][1] - quad->p[i1][1];
            double den = sqrt((dx1 * dx1 + dy1 * dy1) * (dx2 * dx2 + dy2 * dy2));
            if (den == 0.0) goto out;
            double cd = (dx1 * dx2 + dy1 * dy2) / den;
            if ((cd > ccr || cd < -ccr) || dx1 * dy2 < dy1 * dx2) goto out;
        }
    }
    ok = true;
out:
    free(keys); free(tmp); free(buf);
    return ok;
}

/* ------------------------------------------------------------------ (e) */

typedef struct { const uint8_t *im; int64_t s, w, h; } OView;

/* fast/decode.c:8-11 */
static int64_t ctrunc(double x) {
    if (x > -2147483649.0 && x < 2147483648.0) return (int64_t)x;
    return -2147483648LL;
}

/* One step of the edge search: the weight (g2 - g1)^2 of the samples at
 * n+1 and n-1 along the normal, or 0 where either is outside the image or
 * g1 < g2 (fast/decode.c:440-456, the bilinear sum in its lane order). */
static double lane_weight(const OView *v, double x0, double y0, double nx, double ny, double n) {
    double x1 = x0 + (n + 1.0) * nx - 0.5, y1 = y0 + (n + 1.0) * ny - 0.5;
    double x2 = x0 + (n - 1.0) * nx - 0.5, y2 = y0 + (n - 1.0) * ny - 0.5;
    double wl = (double)(v->w - 1), hl = (double)(v->h - 1);
    if (!(x1 > -1.0 && x1 < wl && y1 > -1.0 && y1 < hl && x2 > -1.0 && x2 < wl && y2 > -1.0 && y2 < hl)) return 0.0;
    double tx1 = trunc(x1), ty1 = trunc(y1), tx2 = trunc(x2), ty2 = trunc(y2);
    const uint8_t *p = v->im + (int64_t)ty1 * v->s + (int64_t)tx1, *q = v->im + (int64_t)ty2 * v->s + (int64_t)tx2;
    int64_t s = v->s;
    double a1 = x1 - tx1, b1 = y1 - ty1, a2 = x2 - tx2, b2 = y2 - ty2;
    double na1 = 1.0 - a1, nb1 = 1.0 - b1, na2 = 1.0 - a2, nb2 = 1.0 - b2;
    double g1 = ((na1 * nb1) * p[0] + (a1 * nb1) * p[1]) + ((na1 * b1) * p[s] + (a1 * b1) * p[s + 1]);
    double g2 = ((na2 * nb2) * q[0] + (a2 * nb2) * q[1]) + ((na2 * b2) * q[s] + (a2 * b2) * q[s + 1]);
    double d = g2 - g1;
    return !(g1 < g2) ? d * d : 0.0;
}

/* fast/decode.c:408-411 */
static double bilinear1(const OView *v, int64_t xi, int64_t yi, double a, double b) {
    const uint8_t *p = v->im + yi * v->s + xi;
    return (1.0 - a) * (1.0 - b) * p[0] + a * (1.0 - b) * p[1] + (1.0 - a) * b * p[v->s] + a * b * p[v->s + 1];
}

/* refine_edges (fast/decode.c:657-761): steps below vsteps (a multiple of
 * 4) are summed in 4 lanes, lane l taking the steps = l mod 4, combined
 * (0+2)+(1+3) (refine_lanes :549-626); the rest follow in order. */
static void refine_edges(float qd, const OView *v, OQuad *q) {
    int64_t w = v->w, h = v->h;
    double lines[4][4];
    int64_t range = (int64_t)(qd + 1.0f);
    int64_t max_steps = 2 * 4 * range + 1;
    int64_t vsteps = imax(max_steps - max_steps % 4, 0);
    for (int edge = 0; edge < 4; edge++) {
        int a = edge, b = (edge + 1) & 3;
        double nx = (double)(q->p[b][1] - q->p[a][1]);
        double ny = (double)(-q->p[b][0] + q->p[a][0]);
        double mag = sqrt(nx * nx + ny * ny);
        nx /= mag;
        ny /= mag;
        if (q->reversed_border) { nx = -nx; ny = -ny; }
        int64_t ns = imax(16, ctrunc(mag / 8.0));
        double mx = 0.0, my = 0.0, mxx = 0.0, mxy = 0.0, myy = 0.0, nn = 0.0;
        for (int64_t si = 0; si < ns; si++) {
            double alpha = (1.0 + (double)si) / (double)(ns + 1);
            double x0 = alpha * q->p[a][0] + (1.0 - alpha) * q->p[b][0];
            double y0 = alpha * q->p[a][1] + (1.0 - alpha) * q->p[b][1];
            double rf = (double)-range;
            double ln[4] = {0, 0, 0, 0}, lc[4] = {0, 0, 0, 0};
            for (int64_t st = 0; st < vsteps; st++) {
                double n = rf + 0.25 * (double)st;
                double wt = lane_weight(v, x0, y0, nx, ny, n);
                ln[st & 3] += wt * n;
                lc[st & 3] += wt;
            }
            double mn = 0.0 + ((ln[0] + ln[2]) + (ln[1] + ln[3]));
            double mc = 0.0 + ((lc[0] + lc[2]) + (lc[1] + lc[3]));
            for (int64_t st = vsteps; st < max_steps; st++) {
                double n = (double)-range + 0.25 * (double)st;
                double x1 = x0 + (n + 1.0) * nx - 0.5, y1 = y0 + (n + 1.0) * ny - 0.5;
                int64_t x1i = ctrunc(x1), y1i = ctrunc(y1);
                if (x1i < 0 || x1i + 1 >= w || y1i < 0 || y1i + 1 >= h) continue;
                double x2 = x0 + (n - 1.0) * nx - 0.5, y2 = y0 + (n - 1.0) * ny - 0.5;
                int64_t x2i = ctrunc(x2), y2i = ctrunc(y2);
                if (x2i < 0 || x2i + 1 >= w || y2i < 0 || y2i + 1 >= h) continue;
                double g1 = bilinear1(v, x1i, y1i, x1 - (double)x1i, y1 - (double)y1i);
                double g2 = bilinear1(v, x2i, y2i, x2 - (double)x2i, y2 - (double)y2i);
                if (g1 < g2) continue;
                double wt = (g2 - g1) * (g2 - g1);
                mn += wt * n;
                mc += wt;
            }
            if (mc == 0.0) continue;
            double n0 = mn / mc;
            double bx = x0 + n0 * nx, by = y0 + n0 * ny;
            mx += bx; my += by; mxx += bx * bx; mxy += bx * by; myy += by * by; nn += 1.0;
        }
        double ex = mx / nn, ey = my / nn;
        double cxx = mxx / nn - ex * ex, cxy = mxy / nn - ex * ey, cyy = myy / nn - ey * ey;
        /* correctly rounded float atan2/cos/sin (cr_libm.c) */
        float th2 = cr_atan2f((float)(-2.0 * cxy), (float)(cyy - cxx));
        double theta = 0.5 * (double)th2;
        lines[edge][0] = ex;
        lines[edge][1] = ey;
        lines[edge][2] = cr_cosf((float)theta);
        lines[edge][3] = cr_sinf((float)theta);
    }
    for (int i = 0; i < 4; i++) {
        int j = (i + 1) & 3;
        double a00 = lines[i][3], a01 = -lines[j][3], a10 = -lines[i][2], a11 = lines[j][2];
        double b0 = -lines[i][0] + lines[j][0], b1 = -lines[i][1] + lines[j][1];
        double det = a00 * a11 - a10 * a01;
        if (fabs(det) > 0.001) {
            double w00 = a11 / det, w01 = -a01 / det;
            double l0 = w00 * b0 + w01 * b1;
            q->p[j][0] = (float)(lines[i][0] + l0 * a00);
            q->p[j][1] = (float)(lines[i][1] + l0 * a10);
        }
    }
}

/* fast/decode.c:114-165 */
static bool homography(const double c[4][4], double hr[9]) {
    double a[72];
    memset(a, 0, sizeof a);
    for (int i = 0; i < 4; i++) {
        int r0 = 18 * i, r1 = r0 + 9;
        a[r0 + 0] = c[i][0]; a[r0 + 1] = c[i][1]; a[r0 + 2] = 1.0;
        a[r0 + 6] = -c[i][0] * c[i][2]; a[r0 + 7] = -c[i][1] * c[i][2]; a[r0 + 8] = c[i][2];
        a[r1 + 3] = c[i][0]; a[r1 + 4] = c[i][1]; a[r1 + 5] = 1.0;
        a[r1 + 6] = -c[i][0] * c[i][3]; a[r1 + 7] = -c[i][1] * c[i][3]; a[r1 + 8] = c[i][3];
    }
    for (int col = 0; col < 8; col++) {
        double mv = 0.0;
        int mi = -1;
        for (int row = col; row < 8; row++) {
            double v = fabs(a[row * 9 + col]);
            if (v > mv) { mv = v; mi = row; }
        }
        if (mi < 0 || mv < 1e-10) return false;
        if (mi != col)
            for (int i = col; i < 9; i++) { double t = a[col * 9 + i]; a[col * 9 + i] = a[mi * 9 + i]; a[mi * 9 + i] = t; }
        for (int i = col + 1; i < 8; i++) {
            double f = a[i * 9 + col] / a[col * 9 + col];
            a[i * 9 + col] = 0.0;
            for (int j = col + 1; j < 9; j++) a[i * 9 + j] -= f * a[col * 9 + j];
        }
    }
    for (int col = 7; col >= 0; col--) {
        double sum = 0.0;
        for (int i = col + 1; i < 8; i++) sum += a[col * 9 + i] * a[i * 9 + 8];
        a[col * 9 + 8] = (a[col * 9 + 8] - sum) / a[col * 9 + col];
    }
    const double hh[9] = {a[8], a[17], a[26], a[35], a[44], a[53], a[62], a[71], 1.0};
    memcpy(hr, hh, sizeof hh);
    return true;
}

/* fast/decode.c:168-196 */
static bool invertible(const double m[9]) {
    double lu[9];
    memcpy(lu, m, sizeof lu);
    for (int j = 0; j < 3; j++) {
        for (int i = 0; i < 3; i++) {
            int km = i < j ? i : j;
            double acc = 0.0;
            for (int k = 0; k < km; k++) acc += lu[i * 3 + k] * lu[k * 3 + j];
            lu[i * 3 + j] -= acc;
        }
        int p = j;
        for (int i = j + 1; i < 3; i++) if (fabs(lu[i * 3 + j]) > fabs(lu[p * 3 + j])) p = i;
        if (p != j) for (int k = 0; k < 3; k++) { double t = lu[p * 3 + k]; lu[p * 3 + k] = lu[j * 3 + k]; lu[j * 3 + k] = t; }
        double d = lu[j * 3 + j];
        if (fabs(d) < 1e-8) return false;
        if (d != 0.0) { double inv = 1.0 / d; for (int i = j + 1; i < 3; i++) lu[i * 3 + j] *= inv; }
    }
    return true;
}

static void project(const double h[9], double x, double y, double *ox, double *oy) {
    double xx = h[0] * x + h[1] * y + h[2], yy = h[3] * x + h[4] * y + h[5], zz = h[6] * x + h[7] * y + h[8];
    *ox = xx / zz;
    *oy = yy / zz;
}

typedef struct { double a[9], b[3], c[3]; } OGray;

static void gm_add(OGray *g, double x, double y, double v) {
    g->a[0] += x * x; g->a[1] += x * y; g->a[2] += x; g->a[4] += y * y; g->a[5] += y; g->a[8] += 1.0;
    g->b[0] += x * v; g->b[1] += y * v; g->b[2] += v;
}

/* fast/decode.c:242-265 */
static void gm_solve(OGray *g) {
    const double *a = g->a;
    double l[9] = {0}, m[9] = {0};
    l[0] = sqrt(a[0]);
    l[3] = a[1] / l[0];
    l[6] = a[2] / l[0];
    l[4] = sqrt(a[4] - l[3] * l[3]);
    l[7] = (a[5] - l[3] * l[6]) / l[4];
    l[8] = sqrt(a[8] - l[6] * l[6] - l[7] * l[7]);
    m[0] = 1.0 / l[0];
    m[3] = -l[3] * m[0] / l[4];
    m[4] = 1.0 / l[4];
    m[6] = (-l[6] * m[0] - l[7] * m[3]) / l[8];
    m[7] = -l[7] * m[4] / l[8];
    m[8] = 1.0 / l[8];
    const double *b = g->b;
    double t0 = m[0] * b[0], t1 = m[3] * b[0] + m[4] * b[1], t2 = m[6] * b[0] + m[7] * b[1] + m[8] * b[2];
    g->c[0] = m[0] * t0 + m[3] * t1 + m[6] * t2;
    g->c[1] = m[4] * t1 + m[7] * t2;
    g->c[2] = m[8] * t2;
}

/* fast/decode.c:268-278 */
static double pixel_value(const OView *v, double px, double py) {
    int64_t x1 = ctrunc(floor(px - 0.5)), x2 = ctrunc(ceil(px - 0.5));
    double x = px - 0.5 - (double)x1;
    int64_t y1 = ctrunc(floor(py - 0.5)), y2 = ctrunc(ceil(py - 0.5));
    double y = py - 0.5 - (double)y1;
    if (x1 < 0 || x2 >= v->w || y1 < 0 || y2 >= v->h) return -1.0;
    const uint8_t *r1 = v->im + y1 * v->s, *r2 = v->im + y2 * v->s;
    return r1[x1] * (1.0 - x) * (1.0 - y) + r1[x2] * x * (1.0 - y) + r2[x1] * (1.0 - x) * y + r2[x2] * x * y;
}

/* fast/decode.c:291-402. Returns false when the border polarity is wrong. */
static bool sample(const orc_family *f, double sharpening, const OView *v, const OQuad *q, float *margin,
                   uint64_t *rcode_out) {
    int32_t wab = f->width_at_border;
    float wf = (float)wab;
    const float pat[40] = {
        -0.5f, 0.5f, 0.0f, 1.0f, 1.0f,  0.5f, 0.5f, 0.0f, 1.0f, 0.0f,
        (float)(wab + 0.5), 0.5f, 0.0f, 1.0f, 1.0f,  (float)(wab - 0.5), 0.5f, 0.0f, 1.0f, 0.0f,
        0.5f, -0.5f, 1.0f, 0.0f, 1.0f,  0.5f, 0.5f, 1.0f, 0.0f, 0.0f,
        0.5f, (float)(wab + 0.5), 1.0f, 0.0f, 1.0f,  0.5f, (float)(wab - 0.5), 1.0f, 0.0f, 0.0f,
    };
    OGray wh, bl;
    memset(&wh, 0, sizeof wh);
    memset(&bl, 0, sizeof bl);
    const double *H = q->H;
    for (int pi = 0; pi < 8; pi++) {
        bool white = pat[pi * 5 + 4] != 0.0f;
        for (int32_t i = 0; i < wab; i++) {
            double tx01 = (double)((pat[pi * 5] + (float)i * pat[pi * 5 + 2]) / wf);
            double ty01 = (double)((pat[pi * 5 + 1] + (float)i * pat[pi * 5 + 3]) / wf);
            double tx = 2.0 * (tx01 - 0.5), ty = 2.0 * (ty01 - 0.5);
            double zz = H[6] * tx + H[7] * ty + H[8];
            double px = (H[0] * tx + H[1] * ty + H[2]) / zz, py = (H[3] * tx + H[4] * ty + H[5]) / zz;
            int64_t ix = ctrunc(px), iy = ctrunc(py);
            if (ix < 0 || iy < 0 || ix >= v->w || iy >= v->h) continue;
            gm_add(white ? &wh : &bl, tx, ty, (double)v->im[iy * v->s + ix]);
        }
    }
    gm_solve(&wh);
    if (wab > 1) gm_solve(&bl);
    else { bl.c[0] = 0.0; bl.c[1] = 0.0; bl.c[2] = bl.b[2] / 4.0; }
    double wi = wh.c[0] * 0.0 + wh.c[1] * 0.0 + wh.c[2], bi = bl.c[0] * 0.0 + bl.c[1] * 0.0 + bl.c[2];
    if ((wi - bi < 0.0) != f->reversed_border) return false;
    int64_t tw = f->total_width;
    double *vals = calloc(tw * tw, 8), *sharp = calloc(tw * tw, 8);
    int64_t mc = (wab - tw) / 2;
    for (int64_t i = 0; i < f->nbits; i++) {
        int64_t by = f->bit_y[i], bx = f->bit_x[i];
        double tx01 = ((double)bx + 0.5) / wab, ty01 = ((double)by + 0.5) / wab;
        double tx = 2.0 * (tx01 - 0.5), ty = 2.0 * (ty01 - 0.5);
        double zz = H[6] * tx + H[7] * ty + H[8];
        double px = (H[0] * tx + H[1] * ty + H[2]) / zz, py = (H[3] * tx + H[4] * ty + H[5]) / zz;
        double val = pixel_value(v, px, py);
        if (val == -1.0) continue;
        double th = ((bl.c[0] * tx + bl.c[1] * ty + bl.c[2]) + (wh.c[0] * tx + wh.c[1] * ty + wh.c[2])) / 2.0;
        vals[tw * (by - mc) + (bx - mc)] = val - th;
    }
    for (int64_t y = 0; y < tw; y++)
        for (int64_t x = 0; x < tw; x++) {
            double acc = 0.0;
            if (y > 0) acc += -vals[(y - 1) * tw + x];
            if (x > 0) acc += -vals[y * tw + x - 1];
            acc += vals[y * tw + x] * 4.0;
            if (x + 1 < tw) acc += -vals[y * tw + x + 1];
            if (y + 1 < tw) acc += -vals[(y + 1) * tw + x];
            sharp[y * tw + x] = acc;
        }
    for (int64_t i = 0; i < tw * tw; i++) vals[i] = vals[i] + sharpening * sharp[i];
    float bs = 0.0f, ws = 0.0f, bc = 1.0f, wc = 1.0f;
    uint64_t rc = 0;
    for (int64_t i = 0; i < f->nbits; i++) {
        rc <<= 1;
        double vv = vals[(f->bit_y[i] - mc) * tw + (f->bit_x[i] - mc)];
        if (vv > 0.0) { ws = (float)(ws + vv); wc += 1.0f; rc |= 1; }
        else { bs = (float)(bs - vv); bc += 1.0f; }
    }
    double x = (double)(ws / wc), y = (double)(bs / bc);
    double m = x != x ? y : y != y ? x : x < y ? x : y;
    *margin = (float)m;
    *rcode_out = rc;
    free(vals); free(sharp);
    return true;
}

/* fast/decode.c:61-71 */
static uint64_t rotate90(uint64_t w0, int64_t nbits) {
    int64_t p = nbits;
    uint64_t l = 0;
    if (nbits % 4 == 1) { p = nbits - 1; l = 1; }
    uint64_t w = ((w0 >> l) << (p / 4 + (int64_t)l)) | (w0 >> (3 * p / 4 + (int64_t)l) << l) | (w0 & l);
    return w & (((uint64_t)1 << nbits) - 1);
}

/* The code lookup of fast/decode.c:87-107, as the search it indexes: for
 * each rotation (0..3), each chunk (0..3), the codes whose chunk equals
 * the received code's, in id order, the first within maxhamming. */
static void lookup(const orc_family *f, int maxh, uint64_t rcode, int64_t *id, int64_t *ham, int64_t *rot) {
    int64_t cs = (f->nbits + 3) / 4;
    uint64_t mask = ((uint64_t)1 << cs) - 1;
    for (int64_t r = 0; r < 4; r++) {
        for (int64_t i = 0; i < 4; i++) {
            uint64_t val = (rcode >> (i * cs)) & mask;
            for (int64_t c = 0; c < f->ncodes; c++) {
                if (((f->codes[c] >> (i * cs)) & mask) != val) continue;
                int64_t hm = __builtin_popcountll(f->codes[c] ^ rcode);
                if (hm <= maxh) { *id = c; *ham = hm; *rot = r; return; }
            }
        }
        rcode = rotate90(rcode, f->nbits);
    }
    *id = 65535; *ham = 255; *rot = 0;
}

/* fast/detector.c:83-107 */
static orc_detection make_det(int64_t fi, int64_t id, int64_t ham, int64_t rot, float margin, const double qh[9]) {
    double th = (double)rot * M_PI / 2.0;
    double c = cos(th), s = sin(th);
    const double r[9] = {c, -s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0};
    orc_detection d;
    memset(&d, 0, sizeof d);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
            for (int k = 0; k < 3; k++) acc += qh[i * 3 + k] * r[k * 3 + j];
            d.H[i * 3 + j] = acc;
        }
    d.family = (int32_t)fi;
    d.id = (int32_t)id;
    d.hamming = (int32_t)ham;
    d.decision_margin = margin;
    project(d.H, 0.0, 0.0, &d.c[0], &d.c[1]);
    for (int i = 0; i < 4; i++) project(d.H, i == 1 || i == 2 ? 1.0 : -1.0, i < 2 ? 1.0 : -1.0, &d.p[i][0], &d.p[i][1]);
    return d;
}

/* g2d overlap test, fast/decode.c:766-843 */
typedef struct { double p[2], u[2]; } OLine;
static OLine g2d_line(const double a[2], const double b[2]) {
    double u0 = b[0] - a[0], u1 = b[1] - a[1];
    double mag = (double)sqrtf((float)(u0 * u0 + u1 * u1));
    return (OLine){{a[0], a[1]}, {u0 / mag, u1 / mag}};
}
static double g2d_coord(const OLine *l, const double q[2]) { return (q[0] - l->p[0]) * l->u[0] + (q[1] - l->p[1]) * l->u[1]; }
static bool g2d_isect(const OLine *la, const OLine *lb, double out[2]) {
    double m00 = la->u[0], m01 = -lb->u[0], m10 = la->u[1], m11 = -lb->u[1];
    double det = m00 * m11 - m01 * m10;
    if (fabs(det) < 0.00000001) return false;
    double i00 = m11 / det, i01 = -m01 / det;
    double b00 = lb->p[0] - la->p[0], b10 = lb->p[1] - la->p[1];
    double x00 = i00 * b00 + i01 * b10;
    out[0] = la->u[0] * x00 + la->p[0];
    out[1] = la->u[1] * x00 + la->p[1];
    return true;
}
static bool g2d_seg(const double a0[2], const double a1[2], const double b0[2], const double b1[2]) {
    OLine la = g2d_line(a0, a1), lb = g2d_line(b0, b1);
    double t[2];
    if (!g2d_isect(&la, &lb, t)) return false;
    double a = g2d_coord(&la, la.p), b = g2d_coord(&la, a1), c = g2d_coord(&la, t);
    if ((c < a && c < b) || (c > a && c > b)) return false;
    a = g2d_coord(&lb, lb.p); b = g2d_coord(&lb, b1); c = g2d_coord(&lb, t);
    if ((c < a && c < b) || (c > a && c > b)) return false;
    return true;
}
static bool g2d_contains(const double poly[4][2], const double q[2]) {
    int last = 0, acc = 0;
    for (int i = 0; i < 5; i++) {
        const double *p = poly[i % 4];
        int quadrant = p[0] < q[0] ? (p[1] < q[1] ? 2 : 1) : (p[1] < q[1] ? 3 : 0);
        if (i > 0) {
            int dq = quadrant - last;
            if (dq == -3 || dq == 1) acc++;
            else if (dq == -1 || dq == 3) acc--;
            else if (dq == -2 || dq == 2) {
                const double *p0 = poly[i - 1];
                double nx = p[1] - q[1], ny = -p[0] + q[0];
                double dot = nx * (p0[0] - q[0]) + ny * (p0[1] - q[1]);
                acc += dot < 0.0 ? -2 : 2;
            }
        }
        last = quadrant;
    }
    return acc >= 2 || acc <= -2;
}
static bool g2d_overlaps(const double pa[4][2], const double pb[4][2]) {
    for (int ia = 0; ia < 4; ia++)
        for (int ib = 0; ib < 4; ib++)
            if (g2d_seg(pa[ia], pa[(ia + 1) % 4], pb[ib], pb[(ib + 1) % 4])) return true;
    const double qb[2] = {(pb[0][0] + pb[1][0] + pb[2][0]) / 3.0, (pb[0][1] + pb[1][1] + pb[2][1]) / 3.0};
    if (g2d_contains(pa, qb)) return true;
    const double qa[2] = {(pa[0][0] + pa[1][0] + pa[2][0]) / 3.0, (pa[0][1] + pa[1][1] + pa[2][1]) / 3.0};
    return g2d_contains(pb, qa);
}
static int pref(int p, double a, double b) { return p != 0 ? p : a < b ? -1 : b < a ? 1 : 0; }

/* fast/decode.c:860-891 */
void orc_reconcile(orc_detection *d, int64_t *n) {
    int64_t i0 = 0;
    while (i0 < *n) {
        bool removed0 = false;
        int64_t i1 = i0 + 1;
        while (i1 < *n) {
            if (d[i0].id != d[i1].id || d[i0].family != d[i1].family) { i1++; continue; }
            if (g2d_overlaps(d[i0].p, d[i1].p)) {
                int p = 0;
                p = pref(p, d[i0].hamming, d[i1].hamming);
                p = pref(p, -(double)d[i0].decision_margin, -(double)d[i1].decision_margin);
                for (int i = 0; i < 4; i++) { p = pref(p, d[i0].p[i][0], d[i1].p[i][0]); p = pref(p, d[i0].p[i][1], d[i1].p[i][1]); }
                if (p < 0) { memmove(d + i1, d + i1 + 1, (*n - i1 - 1) * sizeof *d); (*n)--; continue; }
                memmove(d + i0, d + i0 + 1, (*n - i0 - 1) * sizeof *d);
                (*n)--;
                removed0 = true;
                break;
            }
            i1++;
        }
        if (!removed0) i0++;
    }
}

/* ------------------------------------------------------------------ */

void orc_result_free(orc_result *r) { free(r->dets); memset(r, 0, sizeof *r); }

void orc_stages_free(orc_stages *s) {
    free(s->qim); free(s->t); free(s->runs); free(s->row_off); free(s->rep); free(s->csize); free(s->ok);
    free(s->cluster_ids); free(s->ccount); free(s->pts); free(s->quads);
    memset(s, 0, sizeof *s);
}

/* fast/detector.c:246-335 (nthreads 1) */
int orc_detect(const orc_params *P, const orc_family *fams, const int32_t *bits, int32_t nfam, const uint8_t *px,
               int64_t W, int64_t Hh, int64_t stride, orc_result *res, orc_stages *st) {
    memset(res, 0, sizeof *res);
    if (st) memset(st, 0, sizeof *st);
    if (nfam == 0 || W < 8 || Hh < 8) return -1;
    for (int i = 0; i < 7; i++) qsm_kern[i] = (float)exp(-(double)((i - 3) * (i - 3)) / 2.0);
    int64_t w = W, h = Hh;
    if (P->quad_decimate > 1.0f) dec_size(W, Hh, P->quad_decimate, &w, &h);
    uint8_t *qim = malloc(w * h);
    if (P->quad_decimate > 1.0f) decimate(px, stride, P->quad_decimate, qim, w, h);
    else for (int64_t y = 0; y < h; y++) memcpy(qim + y * w, px + y * stride, w);
    bool blurred = blur_or_sharpen(qim, w, h, P->quad_sigma);
    uint8_t *t = malloc(w * h);
    threshold(qim, w, h, P->min_white_black_diff, t);
    if (P->deglitch) deglitch(t, w, h);
    ORun *runs = malloc(sizeof(ORun) * (w * h + 1));
    uint32_t *ro = malloc(4 * (h + 2));
    ro[0] = 0;
    for (int64_t y = 0; y < h; y++) ro[y + 1] = ro[y] + (uint32_t)rle_row(t + y * w, w, runs + ro[y]);
    ro[h + 1] = ro[h];
    int64_t nruns = ro[h], nnodes = nruns + h;
    ONode *uf = malloc(sizeof(ONode) * nnodes);
    components(uf, t, w, h, runs, ro);
    uint32_t *rep = malloc(4 * nnodes);
    uint8_t *ok = malloc(nnodes);
    for (int64_t i = 0; i < nnodes; i++) {
        rep[i] = uf_find(uf, (uint32_t)i);
        ok[i] = (int64_t)uf[rep[i]].size + 1 >= P->min_cluster_pixels;
    }
    /* min_tag_width, fast/detector.c:67-81 */
    bool nb = false, rb = false;
    int32_t mtw = 1000000;
    for (int i = 0; i < nfam; i++) {
        if (fams[i].width_at_border < mtw) mtw = fams[i].width_at_border;
        nb = nb || !fams[i].reversed_border;
        rb = rb || fams[i].reversed_border;
    }
    if (P->quad_decimate > 1.0f) mtw = (int32_t)((float)mtw / P->quad_decimate);
    if (mtw < 3) mtw = 3;
    int64_t npts = 0;
    OEmit *em = h >= 3 ? gradient_points(t, w, h, runs, ro, rep, ok, &npts) : NULL;
    /* clusters: contiguous runs of one id in em */
    int64_t ncl = 0;
    int64_t *cs = malloc(8 * (npts + 1));
    for (int64_t i = 0; i < npts; i++) if (i == 0 || em[i].id != em[i - 1].id) cs[ncl++] = i;
    cs[ncl] = npts;
    OPt *cp = malloc(sizeof(OPt) * (npts + 1));
    for (int64_t i = 0; i < npts; i++) cp[i] = em[i].p;
    OQuad *quads = malloc(sizeof(OQuad) * (ncl + 1));
    int64_t nq = 0;
    int64_t max_px = 2 * (2 * w + 2 * h);
    for (int64_t c = 0; c < ncl; c++) {
        int64_t n = cs[c + 1] - cs[c];
        if (n < P->min_cluster_pixels || n > max_px) continue;
        OQuad q;
        memset(&q, 0, sizeof q);
        if (fit_quad(P, cp + cs[c], n, qim, w, h, mtw, nb, rb, &q)) quads[nq++] = q;
    }
    if (st) {
        st->qw = w; st->qh = h;
        st->qim = malloc(w * h); memcpy(st->qim, qim, w * h);
        st->t = malloc(w * h); memcpy(st->t, t, w * h);
        st->nruns = nruns;
        st->runs = malloc(4 * (3 * nruns + 1));
        for (int64_t i = 0; i < nruns; i++) {
            st->runs[3 * i] = runs[i].start; st->runs[3 * i + 1] = runs[i].end; st->runs[3 * i + 2] = runs[i].v;
        }
        st->row_off = malloc(4 * (h + 1)); memcpy(st->row_off, ro, 4 * (h + 1));
        st->nnodes = nnodes;
        st->rep = malloc(4 * nnodes + 4); memcpy(st->rep, rep, 4 * nnodes);
        st->ok = malloc(nnodes + 1); memcpy(st->ok, ok, nnodes);
        st->csize = malloc(4 * nnodes + 4);
        for (int64_t i = 0; i < nnodes; i++) st->csize[i] = uf[rep[i]].size + 1;
        st->nclusters = ncl;
        st->cluster_ids = malloc(8 * (ncl + 1));
        st->ccount = malloc(4 * (ncl + 1));
        for (int64_t c = 0; c < ncl; c++) { st->cluster_ids[c] = em[cs[c]].id; st->ccount[c] = (uint32_t)(cs[c + 1] - cs[c]); }
        st->npts = npts;
        st->pts = malloc(8 * (npts + 1));
        memcpy(st->pts, cp, 8 * npts);
        st->nquads = nq;
        st->quads = malloc(4 * (9 * nq + 1));
        for (int64_t i = 0; i < nq; i++) {
            for (int j = 0; j < 4; j++) { st->quads[9 * i + 2 * j] = quads[i].p[j][0]; st->quads[9 * i + 2 * j + 1] = quads[i].p[j][1]; }
            st->quads[9 * i + 8] = quads[i].reversed_border ? 1.0f : 0.0f;
        }
    }
    /* decode (fast/detector.c:313-327) */
    if (P->quad_decimate > 1.0f)
        for (int64_t i = 0; i < nq; i++)
            for (int j = 0; j < 4; j++) { quads[i].p[j][0] *= P->quad_decimate; quads[i].p[j][1] *= P->quad_decimate; }
    OView view = {px, stride, W, Hh};
    if (blurred && !(P->quad_decimate > 1.0f)) view = (OView){qim, W, W, Hh};
    int64_t dcap = 16, nd = 0;
    orc_detection *dets = malloc(sizeof(orc_detection) * dcap);
    for (int64_t qi = 0; qi < nq; qi++) {
        OQuad *q = &quads[qi];
        if (P->refine_edges) refine_edges(P->quad_decimate, &view, q);
        double corr[4][4];
        for (int i = 0; i < 4; i++) {
            corr[i][0] = i == 0 || i == 3 ? -1.0 : 1.0;
            corr[i][1] = i == 0 || i == 1 ? -1.0 : 1.0;
            corr[i][2] = q->p[i][0];
            corr[i][3] = q->p[i][1];
        }
        double hm[9];
        if (!homography(corr, hm)) continue;
        memcpy(q->H, hm, sizeof hm);
        if (!invertible(hm)) continue;
        for (int fi = 0; fi < nfam; fi++) {
            if (fams[fi].reversed_border != q->reversed_border) continue;
            float margin;
            uint64_t rcode;
            if (!sample(&fams[fi], P->decode_sharpening, &view, q, &margin, &rcode)) continue;
            int64_t id, ham, rot;
            lookup(&fams[fi], bits[fi], rcode, &id, &ham, &rot);
            if (margin >= 0.0f && ham < 255) {
                if (nd == dcap) { dcap *= 2; dets = realloc(dets, sizeof(orc_detection) * dcap); }
                dets[nd++] = make_det(fi, id, ham, rot, margin, q->H);
            }
        }
    }
    res->nraw = nd;
    orc_reconcile(dets, &nd);
    /* stable sort by id (fast/detector.c:164-179) */
    for (int64_t i = 1; i < nd; i++) {
        orc_detection v = dets[i];
        int64_t j = i;
        while (j > 0 && dets[j - 1].id > v.id) { dets[j] = dets[j - 1]; j--; }
        dets[j] = v;
    }
    res->dets = dets;
    res->ndets = nd;
    res->nquads = nq;
    free(qim); free(t); free(runs); free(ro); free(uf); free(rep); free(ok); free(em); free(cs); free(cp); free(quads);
    return 0;
}
