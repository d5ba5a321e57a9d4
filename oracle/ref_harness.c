/* Stage-dump harness over the reference detector, compiled from the
 * reference's own sources where they lie (the unity build
 * fast/apriltag_all.c is #included, so its static stage functions are
 * callable here). Built by oracle/Makefile into oracle/_ref/libatref.so.
 *
 * ref_stages() runs the single-thread flow of at_detect
 * (fast/detector.c:246-335) and quad_thresh_single (fast/detector.c:198-244)
 * step by step and copies out what each stage produced: the quad image,
 * the thresholded image, the runs, the union-find representatives and ok
 * flags, the clusters in output order with their points, and the quads
 * before scaling. The public at_* API of the same library gives the
 * detections.
 *
 * Test infrastructure only: used by tests/ and by bench.py's CPU-baseline
 * leg, never by the product. */
#include "fast/apriltag_all.c"

typedef struct {
    int64_t qw, qh;
    uint8_t *qim;       /* qw*qh: decimated, blurred or sharpened */
    uint8_t *t;         /* qw*qh: thresholded (after deglitch) */
    int64_t nruns;
    uint32_t *runs;     /* nruns x 3: start, end, v */
    uint32_t *row_off;  /* qh + 1 */
    int64_t nnodes;     /* nruns + qh */
    uint32_t *rep;      /* nnodes */
    uint32_t *csize;    /* nnodes: size + 1 of each node's component */
    uint8_t *ok;        /* nnodes */
    int64_t nclusters;
    uint64_t *cluster_ids; /* nclusters, output order */
    uint32_t *ccount;   /* nclusters: points of each */
    int64_t npts;       /* total points */
    uint64_t *pts;      /* npts, grouped by cluster in output order (no padding) */
    int64_t nquads;
    float *quads;       /* nquads x 9: p[4][2] (quad image scale), reversed_border */
} RefStages;

void ref_stages_free(RefStages *s) {
    free(s->qim);
    free(s->t);
    free(s->runs);
    free(s->row_off);
    free(s->rep);
    free(s->csize);
    free(s->ok);
    free(s->cluster_ids);
    free(s->ccount);
    free(s->pts);
    free(s->quads);
    memset(s, 0, sizeof *s);
}

/* Returns 0 on success, -1 when at_detect would return early. */
int ref_stages(at_detector *td, const uint8_t *pixels, int64_t width, int64_t height, int64_t stride,
               RefStages *out) {
    memset(out, 0, sizeof *out);
    struct at_internal *in = td->in;
    if (in->nfamilies == 0 || width < 8 || height < 8) return -1;
    int32_t saved_threads = td->nthreads;
    bool saved_debug = td->debug;
    td->nthreads = 1;
    td->debug = false;
    /* size the arenas exactly as at_detect does */
    at_detect(td, pixels, width, height, stride);
    at_ctx *cx = &in->mcx;
    int64_t qw = width, qh = height;
    if (td->quad_decimate > 1.0f) decimated_size(width, height, td->quad_decimate, &qw, &qh);
    in->icx.sp = 1;
    in->buf_qim = STK_TAKE(&in->icx, uint8_t);
    in->buf_t = STK_TAKE(&in->icx, uint8_t);
    in->buf_uf = STK_TAKE(&in->icx, UfNode);
    at_mark mk = stk_enter(cx);
    uint8_t *qim = in->buf_qim;
    if (td->quad_decimate > 1.0f)
        decimate_rows(pixels, stride, td->quad_decimate, qim, qw, 0, qh);
    else
        for (int64_t y = 0; y < qh; y++) memcpy(qim + y * qw, pixels + y * stride, (size_t)qw);
    blur_or_sharpen(cx, qim, qw, qh, td->quad_sigma);
    out->qw = qw;
    out->qh = qh;
    out->qim = malloc((size_t)(qw * qh));
    memcpy(out->qim, qim, (size_t)(qw * qh));

    /* quad_thresh_single, stage by stage */
    int64_t w = qw, h = qh;
    uint8_t *t = in->buf_t;
    vec_run runs;
    svec_take(cx, runs);
    uint32_t *row_off = STK_ALLOC(cx, uint32_t, h + 1);
    threshold_image(cx, qim, w, h, w, td->qtp.min_white_black_diff, t);
    if (td->qtp.deglitch) deglitch(cx, t, w, h);
    rle_image(cx, t, w, h, &runs, row_off);
    out->t = malloc((size_t)(w * h));
    memcpy(out->t, t, (size_t)(w * h));
    int64_t nruns = row_off[h];
    int64_t nnodes = nruns + h;
    out->nruns = nruns;
    out->runs = malloc((size_t)(3 * nruns + 1) * 4);
    for (int64_t i = 0; i < nruns; i++) {
        out->runs[3 * i] = runs.data[i].start;
        out->runs[3 * i + 1] = runs.data[i].end;
        out->runs[3 * i + 2] = runs.data[i].v;
    }
    out->row_off = malloc((size_t)(h + 1) * 4);
    memcpy(out->row_off, row_off, (size_t)(h + 1) * 4);

    UfNode *uf = in->buf_uf;
    uint8_t *edge = STK_ALLOC(cx, uint8_t, 2 * h);
    row_edges(t, w, h, edge);
    connected_components(cx, uf, edge, w, h, runs.data, row_off, 1);

    ClusterTable ct;
    ct_init(cx, &ct, cluster_map_size(w, h, 1));
    bool normal_border, reversed_border;
    int64_t tag_width = min_tag_width(td, &normal_border, &reversed_border);
    uint32_t *rep = STK_ALLOC(cx, uint32_t, nnodes);
    uint8_t *ok = STK_ALLOC(cx, uint8_t, nnodes);
    resolve_range(uf, 0, nnodes, td->qtp.min_cluster_pixels, rep, ok);
    out->nnodes = nnodes;
    out->rep = malloc((size_t)nnodes * 4 + 4);
    out->csize = malloc((size_t)nnodes * 4 + 4);
    out->ok = malloc((size_t)nnodes + 1);
    memcpy(out->rep, rep, (size_t)nnodes * 4);
    memcpy(out->ok, ok, (size_t)nnodes);
    for (int64_t i = 0; i < nnodes; i++) out->csize[i] = uf[rep[i]].size + 1;
    if (h >= 3) {
        PackedRows pr = {runs.data, row_off + 1, rep, ok};
        gradient_points(cx, &ct, t + w, w, 1, h - 2, packed_row, &pr, rep + nruns + 1, ok + nruns + 1);
    }
    vec_pt cpts;
    vec_u32 cstart;
    svec_take(cx, cpts);
    svec_take(cx, cstart);
    order_clusters(cx, &ct, &cpts, &cstart);
    /* the cluster ids in output order (order_clusters drops none: every
       slot has a point) */
    {
        int64_t nslots = ct.ids.len;
        uint32_t *order = STK_ALLOC(cx, uint32_t, nslots);
        cluster_order(&ct, order);
        int64_t nc = cstart.len - 1;
        out->nclusters = nc;
        out->cluster_ids = malloc((size_t)(nc + 1) * 8);
        out->ccount = malloc((size_t)(nc + 1) * 4);
        int64_t c = 0, np = 0;
        for (int64_t i = 0; i < nslots; i++) {
            uint32_t s = order[i];
            if (ct.counts.data[s] == 0) continue;
            out->cluster_ids[c] = ct.ids.data[s];
            out->ccount[c] = ct.counts.data[s];
            np += ct.counts.data[s];
            c++;
        }
        out->npts = np;
        out->pts = malloc((size_t)(np + 1) * 8);
        int64_t at = 0;
        for (int64_t k = 0; k < nc; k++) {
            int64_t s0 = cstart.data[k], n = (int64_t)cstart.data[k + 1] - PT_PAD - s0;
            memcpy(out->pts + at, cpts.data + s0, (size_t)n * 8);
            at += n;
        }
    }
    vec_quad quads;
    svec_take(cx, quads);
    fit_quads(cx, &td->qtp, cpts.data, cstart.data, cstart.len - 1, PT_PAD, qim, w, h, tag_width, normal_border,
              reversed_border, &quads);
    out->nquads = quads.len;
    out->quads = malloc((size_t)(quads.len * 9 + 1) * 4);
    for (int64_t i = 0; i < quads.len; i++) {
        for (int j = 0; j < 4; j++) {
            out->quads[9 * i + 2 * j] = quads.data[i].p[j][0];
            out->quads[9 * i + 2 * j + 1] = quads.data[i].p[j][1];
        }
        out->quads[9 * i + 8] = quads.data[i].reversed_border ? 1.0f : 0.0f;
    }
    stk_leave(cx, mk);
    td->nthreads = saved_threads;
    td->debug = saved_debug;
    /* leave the detector's result as at_detect made it */
    at_detect(td, pixels, width, height, stride);
    return 0;
}

/* Offsets of the public detector fields, for a ctypes mirror. */
int64_t ref_sizeof_detector(void) { return (int64_t)sizeof(at_detector); }
int64_t ref_sizeof_detection(void) { return (int64_t)sizeof(at_detection); }
int64_t ref_sizeof_family(void) { return (int64_t)sizeof(at_family); }

/* The reference's reconcile (fast/decode.c:860-891) on a caller's list. */
void ref_reconcile(at_detection *dets, int64_t *ndets) { reconcile(dets, ndets); }

/* The detections of at_detect's single-thread path before reconcile and
 * the sort (fast/detector.c:283-327): decimate, blur, quad_thresh_single,
 * scale, decode_quads. Copies up to cap of them into out; returns their
 * count, or -1 when at_detect would return early. */
int64_t ref_raw_detections(at_detector *td, const uint8_t *pixels, int64_t width, int64_t height, int64_t stride,
                           at_detection *out, int64_t cap) {
    struct at_internal *in = td->in;
    if (in->nfamilies == 0 || width < 8 || height < 8) return -1;
    int32_t saved_threads = td->nthreads;
    bool saved_debug = td->debug;
    td->nthreads = 1;
    td->debug = false;
    at_detect(td, pixels, width, height, stride); /* sizes the arenas */
    at_ctx *cx = &in->mcx;
    int64_t qw = width, qh = height;
    if (td->quad_decimate > 1.0f) decimated_size(width, height, td->quad_decimate, &qw, &qh);
    in->icx.sp = 1;
    in->buf_qim = STK_TAKE(&in->icx, uint8_t);
    in->buf_t = STK_TAKE(&in->icx, uint8_t);
    in->buf_uf = STK_TAKE(&in->icx, UfNode);
    at_mark mk = stk_enter(cx);
    uint8_t *qim = in->buf_qim;
    if (td->quad_decimate > 1.0f)
        decimate_rows(pixels, stride, td->quad_decimate, qim, qw, 0, qh);
    else
        for (int64_t y = 0; y < qh; y++) memcpy(qim + y * qw, pixels + y * stride, (size_t)qw);
    bool blurred = blur_or_sharpen(cx, qim, qw, qh, td->quad_sigma);
    ImView view = {pixels, stride, width, height};
    if (blurred && !(td->quad_decimate > 1.0f)) view = (ImView){qim, width, width, height};
    vec_quad quads;
    svec_take(cx, quads);
    quad_thresh_single(td, qim, qw, qh, &quads);
    if (td->quad_decimate > 1.0f)
        for (int64_t i = 0; i < quads.len; i++)
            for (int j = 0; j < 4; j++) {
                quads.data[i].p[j][0] *= td->quad_decimate;
                quads.data[i].p[j][1] *= td->quad_decimate;
            }
    in->dets.len = 0;
    decode_quads(td, quads.data, quads.len, &view);
    int64_t n = in->dets.len;
    memcpy(out, in->dets.data, (size_t)(n < cap ? n : cap) * sizeof(at_detection));
    stk_leave(cx, mk);
    td->nthreads = saved_threads;
    td->debug = saved_debug;
    at_detect(td, pixels, width, height, stride);
    return n;
}
