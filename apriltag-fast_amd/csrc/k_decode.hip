// Stage (d, refine) and (e): per quad, edge refinement on the full image,
// homography, bit sampling and code lookup; per frame, the overlap
// reconcile and the stable sort by id.
//
//   k_decode   persistent, a wave a quad (or, for calls of a few frames, a
//              256-thread block with a wave per edge), over the quads of all
//              frames. refine_edges (fast/decode.c:657-761): the edges in turn, four
//              lanes a sample, each lane one of the reference's four step
//              sums (:549-626), the edge's sums added in sample order by
//              lane 0. The homography's elimination runs in LDS, its row
//              updates in parallel. The bit sampler (:291-402) gathers its
//              pixel samples in parallel; the gray models and the scores are
//              added in the reference's order by one lane; each cell's
//              threshold and sharpening are a lane's. Lookup (:87-107): the
//              quick-decode buckets of the 16 (rotation, chunk) pairs at
//              once, the first match in the reference's order kept.
//   k_finish   one block per frame: the detections in (quad, family)
//              order, reconcile (:860-891), stable sort by id
//              (fast/detector.c:164-179).
//
// atan2f/cosf/sinf must be correctly rounded (the reference's goldens are):
// cr_trig.h.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "atg_internal.h"
#include "cr_trig.h"

#pragma clang fp contract(off)

namespace atg {

// Correctly rounded float sqrt and division. The device's single-precision
// sqrt is not correctly rounded; double has more than 2*24+2 bits, so one
// double operation rounded once to float is the correctly rounded result.
__device__ __forceinline__ float fsqrt_cr(float x) { return (float)__dsqrt_rn((double)x); }
__device__ __forceinline__ float fdiv_cr(float a, float b) { return (float)__ddiv_rn((double)a, (double)b); }

#define DADD __dadd_rn
#define DSUB __dsub_rn
#define DMUL __dmul_rn
#define DDIV __ddiv_rn

// C's (int) of a double on x86-64 (fast/decode.c:8-11)
__device__ __forceinline__ int64_t ctrunc(double x) {
    if (x > -2147483649.0 && x < 2147483648.0) return (int64_t)x;
    return -2147483648LL;
}

struct View {
    const uint8_t *im;
    int64_t s, w, h;
};

// lane_weight (fast/decode.c:440-456): (g2 - g1)^2, or 0 outside the
// image or where g1 < g2; bilinear sums in the lane order.
__device__ __forceinline__ double lane_weight(const View &v, double x0, double y0, double nx, double ny, double n) {
    double x1 = DSUB(DADD(x0, DMUL(DADD(n, 1.0), nx)), 0.5), y1 = DSUB(DADD(y0, DMUL(DADD(n, 1.0), ny)), 0.5);
    double x2 = DSUB(DADD(x0, DMUL(DSUB(n, 1.0), nx)), 0.5), y2 = DSUB(DADD(y0, DMUL(DSUB(n, 1.0), ny)), 0.5);
    double wl = (double)(v.w - 1), hl = (double)(v.h - 1);
    if (!(x1 > -1.0 && x1 < wl && y1 > -1.0 && y1 < hl && x2 > -1.0 && x2 < wl && y2 > -1.0 && y2 < hl)) return 0.0;
    double tx1 = trunc(x1), ty1 = trunc(y1), tx2 = trunc(x2), ty2 = trunc(y2);
    const uint8_t *p = v.im + (int64_t)ty1 * v.s + (int64_t)tx1, *q = v.im + (int64_t)ty2 * v.s + (int64_t)tx2;
    int64_t s = v.s;
    double a1 = DSUB(x1, tx1), b1 = DSUB(y1, ty1), a2 = DSUB(x2, tx2), b2 = DSUB(y2, ty2);
    double na1 = DSUB(1.0, a1), nb1 = DSUB(1.0, b1), na2 = DSUB(1.0, a2), nb2 = DSUB(1.0, b2);
    double g1 = DADD(DADD(DMUL(DMUL(na1, nb1), (double)p[0]), DMUL(DMUL(a1, nb1), (double)p[1])),
                     DADD(DMUL(DMUL(na1, b1), (double)p[s]), DMUL(DMUL(a1, b1), (double)p[s + 1])));
    double g2 = DADD(DADD(DMUL(DMUL(na2, nb2), (double)q[0]), DMUL(DMUL(a2, nb2), (double)q[1])),
                     DADD(DMUL(DMUL(na2, b2), (double)q[s]), DMUL(DMUL(a2, b2), (double)q[s + 1])));
    double d = DSUB(g2, g1);
    return !(g1 < g2) ? DMUL(d, d) : 0.0;
}

// bilinear1 (fast/decode.c:408-411), left to right
__device__ __forceinline__ double bilinear1(const View &v, int64_t xi, int64_t yi, double a, double b) {
    const uint8_t *p = v.im + yi * v.s + xi;
    double na = DSUB(1.0, a), nb = DSUB(1.0, b);
    double r = DMUL(DMUL(na, nb), (double)p[0]);
    r = DADD(r, DMUL(DMUL(a, nb), (double)p[1]));
    r = DADD(r, DMUL(DMUL(na, b), (double)p[v.s]));
    r = DADD(r, DMUL(DMUL(a, b), (double)p[v.s + 1]));
    return r;
}

// One sample point of an edge: the weighted mean offset along the normal
// (fast/decode.c:690-730). The steps below vsteps (a multiple of 4) add to
// four partial sums by step & 3, in order; refine_partial makes partial q
// (on its own lane), refine_finish combines the four as the reference
// does, adds the remaining steps in order and gives the sample's point.
__device__ __forceinline__ void refine_partial(const View &v, int64_t range, int64_t vsteps, double x0, double y0,
                                               double nx, double ny, int q, double *ln, double *lc) {
    const double rf = (double)-range;
    double a = 0.0, c = 0.0;
    for (int64_t s0 = 0; s0 < vsteps; s0 += 4) {
        const double n = DADD(rf, DMUL(0.25, (double)(s0 + q)));
        const double w = lane_weight(v, x0, y0, nx, ny, n);
        a = DADD(a, DMUL(w, n));
        c = DADD(c, w);
    }
    *ln = a;
    *lc = c;
}

__device__ bool refine_finish(const View &v, int64_t range, int64_t max_steps, int64_t vsteps, double x0, double y0,
                              double nx, double ny, const double ln[4], const double lc[4], double *bx, double *by) {
    const double rf = (double)-range;
    double mn = DADD(0.0, DADD(DADD(ln[0], ln[2]), DADD(ln[1], ln[3])));
    double mc = DADD(0.0, DADD(DADD(lc[0], lc[2]), DADD(lc[1], lc[3])));
    for (int64_t st = vsteps; st < max_steps; st++) {
        double n = DADD(rf, DMUL(0.25, (double)st));
        double x1 = DSUB(DADD(x0, DMUL(DADD(n, 1.0), nx)), 0.5), y1 = DSUB(DADD(y0, DMUL(DADD(n, 1.0), ny)), 0.5);
        int64_t x1i = ctrunc(x1), y1i = ctrunc(y1);
        if (x1i < 0 || x1i + 1 >= v.w || y1i < 0 || y1i + 1 >= v.h) continue;
        double x2 = DSUB(DADD(x0, DMUL(DSUB(n, 1.0), nx)), 0.5), y2 = DSUB(DADD(y0, DMUL(DSUB(n, 1.0), ny)), 0.5);
        int64_t x2i = ctrunc(x2), y2i = ctrunc(y2);
        if (x2i < 0 || x2i + 1 >= v.w || y2i < 0 || y2i + 1 >= v.h) continue;
        double g1 = bilinear1(v, x1i, y1i, DSUB(x1, (double)x1i), DSUB(y1, (double)y1i));
        double g2 = bilinear1(v, x2i, y2i, DSUB(x2, (double)x2i), DSUB(y2, (double)y2i));
        if (g1 < g2) continue;
        double w = DMUL(DSUB(g2, g1), DSUB(g2, g1));
        mn = DADD(mn, DMUL(w, n));
        mc = DADD(mc, w);
    }
    if (mc == 0.0) return false;
    double n0 = DDIV(mn, mc);
    *bx = DADD(x0, DMUL(n0, nx));
    *by = DADD(y0, DMUL(n0, ny));
    return true;
}

// A barrier of a block of NT threads over LDS; for one wave, only the
// ordering of its LDS operations.
template <int NT>
__device__ __forceinline__ void bsync_t() {
    if constexpr (NT == 64) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// homography_compute2 (fast/decode.c:114-165): the 8x9 system of the four
// corner correspondences, Gaussian elimination with partial pivoting, in
// LDS. Each pivot step's row updates are independent element by element,
// so the block's threads make them (the same operations per element as
// the reference's loops); the pivot search and back substitution are one
// thread's. c[i] = (x, y) of the tag corner, qp[i] its image point. Every
// thread returns the same result; hr is written by thread 0.
template <int NT>
__device__ bool homography_block(const float qp[4][2], double *A, int *s_piv, double *hr) {
    const int t = threadIdx.x;
    for (int e = t; e < 72; e += NT) {
        const int row = e / 9, col = e % 9, i = row >> 1;
        const double cx = (i == 0 || i == 3) ? -1.0 : 1.0, cy = (i == 0 || i == 1) ? -1.0 : 1.0;
        const double px = (double)qp[i][0], py = (double)qp[i][1];
        const double pv = (row & 1) ? py : px;  // the image coordinate of this row
        const int o = (row & 1) ? 3 : 0;        // where (x, y, 1) sits
        double v = 0.0;
        if (col == o) v = cx;
        else if (col == o + 1) v = cy;
        else if (col == o + 2) v = 1.0;
        else if (col == 6) v = DMUL(-cx, pv);
        else if (col == 7) v = DMUL(-cy, pv);
        else if (col == 8) v = pv;
        A[e] = v;
    }
    bsync_t<NT>();
    bool ok = true;
    for (int col = 0; col < 8; col++) {
        if (t == 0) {
            double mv = 0.0;
            int mi = -1;
            for (int row = col; row < 8; row++) {
                double v = fabs(A[row * 9 + col]);
                if (v > mv) { mv = v; mi = row; }
            }
            *s_piv = (mi < 0 || mv < 1e-10) ? -1 : mi;
        }
        bsync_t<NT>();
        const int mi = *s_piv;
        if (mi < 0) { ok = false; break; }
        if (mi != col && t >= col && t < 9) {
            double x = A[col * 9 + t];
            A[col * 9 + t] = A[mi * 9 + t];
            A[mi * 9 + t] = x;
        }
        bsync_t<NT>();
        constexpr int kR = (72 + NT - 1) / NT;  // elements a thread
        double nv[kR];
        bool wr[kR];
#pragma unroll
        for (int r = 0; r < kR; r++) {
            const int e = t + r * NT;
            nv[r] = 0.0;
            wr[r] = false;
            if (e < 72) {
                const int i = e / 9, j = e % 9;
                if (i > col && j >= col) {
                    double f = DDIV(A[i * 9 + col], A[col * 9 + col]);
                    nv[r] = j == col ? 0.0 : DSUB(A[i * 9 + j], DMUL(f, A[col * 9 + j]));
                    wr[r] = true;
                }
            }
        }
        bsync_t<NT>();
#pragma unroll
        for (int r = 0; r < kR; r++)
            if (wr[r]) A[t + r * NT] = nv[r];
        bsync_t<NT>();
    }
    if (ok && t == 0) {
        for (int col = 7; col >= 0; col--) {
            double sum = 0.0;
            for (int i = col + 1; i < 8; i++) sum = DADD(sum, DMUL(A[col * 9 + i], A[i * 9 + 8]));
            A[col * 9 + 8] = DDIV(DSUB(A[col * 9 + 8], sum), A[col * 9 + col]);
        }
        for (int r = 0; r < 8; r++) hr[r] = A[r * 9 + 8];
        hr[8] = 1.0;
    }
    return ok;
}

// matd_invertible (fast/decode.c:168-196)
__device__ bool invertible(const double m[9], double *lu) {
    for (int i = 0; i < 9; i++) lu[i] = m[i];
    for (int j = 0; j < 3; j++) {
        for (int i = 0; i < 3; i++) {
            int km = i < j ? i : j;
            double acc = 0.0;
            for (int k = 0; k < km; k++) acc = DADD(acc, DMUL(lu[i * 3 + k], lu[k * 3 + j]));
            lu[i * 3 + j] = DSUB(lu[i * 3 + j], acc);
        }
        int p = j;
        for (int i = j + 1; i < 3; i++) if (fabs(lu[i * 3 + j]) > fabs(lu[p * 3 + j])) p = i;
        if (p != j)
            for (int k = 0; k < 3; k++) { double t = lu[p * 3 + k]; lu[p * 3 + k] = lu[j * 3 + k]; lu[j * 3 + k] = t; }
        double d = lu[j * 3 + j];
        if (fabs(d) < 1e-8) return false;
        if (d != 0.0) {
            double inv = DDIV(1.0, d);
            for (int i = j + 1; i < 3; i++) lu[i * 3 + j] = DMUL(lu[i * 3 + j], inv);
        }
    }
    return true;
}

__device__ __forceinline__ void project(const double h[9], double x, double y, double *ox, double *oy) {
    double xx = DADD(DADD(DMUL(h[0], x), DMUL(h[1], y)), h[2]);
    double yy = DADD(DADD(DMUL(h[3], x), DMUL(h[4], y)), h[5]);
    double zz = DADD(DADD(DMUL(h[6], x), DMUL(h[7], y)), h[8]);
    *ox = DDIV(xx, zz);
    *oy = DDIV(yy, zz);
}

struct Gray {
    double a[9], b[3], c[3];
};

__device__ __forceinline__ void gray_add(Gray &g, double x, double y, double v) {
    g.a[0] = DADD(g.a[0], DMUL(x, x)); g.a[1] = DADD(g.a[1], DMUL(x, y)); g.a[2] = DADD(g.a[2], x);
    g.a[4] = DADD(g.a[4], DMUL(y, y)); g.a[5] = DADD(g.a[5], y); g.a[8] = DADD(g.a[8], 1.0);
    g.b[0] = DADD(g.b[0], DMUL(x, v)); g.b[1] = DADD(g.b[1], DMUL(y, v)); g.b[2] = DADD(g.b[2], v);
}

// graymodel_solve (fast/decode.c:242-265)
__device__ void gray_solve(Gray &g) {
    const double *a = g.a;
    double l[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    l[0] = __dsqrt_rn(a[0]);
    l[3] = DDIV(a[1], l[0]);
    l[6] = DDIV(a[2], l[0]);
    l[4] = __dsqrt_rn(DSUB(a[4], DMUL(l[3], l[3])));
    l[7] = DDIV(DSUB(a[5], DMUL(l[3], l[6])), l[4]);
    l[8] = __dsqrt_rn(DSUB(DSUB(a[8], DMUL(l[6], l[6])), DMUL(l[7], l[7])));
    m[0] = DDIV(1.0, l[0]);
    m[3] = DDIV(DMUL(-l[3], m[0]), l[4]);
    m[4] = DDIV(1.0, l[4]);
    m[6] = DDIV(DSUB(DMUL(-l[6], m[0]), DMUL(l[7], m[3])), l[8]);
    m[7] = DDIV(DMUL(-l[7], m[4]), l[8]);
    m[8] = DDIV(1.0, l[8]);
    const double *b = g.b;
    double t0 = DMUL(m[0], b[0]);
    double t1 = DADD(DMUL(m[3], b[0]), DMUL(m[4], b[1]));
    double t2 = DADD(DADD(DMUL(m[6], b[0]), DMUL(m[7], b[1])), DMUL(m[8], b[2]));
    g.c[0] = DADD(DADD(DMUL(m[0], t0), DMUL(m[3], t1)), DMUL(m[6], t2));
    g.c[1] = DADD(DMUL(m[4], t1), DMUL(m[7], t2));
    g.c[2] = DMUL(m[8], t2);
}

// value_for_pixel (fast/decode.c:268-278)
__device__ double pixel_value(const View &v, double px, double py) {
    double xm = DSUB(px, 0.5), ym = DSUB(py, 0.5);
    int64_t x1 = ctrunc(floor(xm)), x2 = ctrunc(ceil(xm));
    double x = DSUB(xm, (double)x1);
    int64_t y1 = ctrunc(floor(ym)), y2 = ctrunc(ceil(ym));
    double y = DSUB(ym, (double)y1);
    if (x1 < 0 || x2 >= v.w || y1 < 0 || y2 >= v.h) return -1.0;
    const uint8_t *r1 = v.im + y1 * v.s, *r2 = v.im + y2 * v.s;
    double nx = DSUB(1.0, x), ny = DSUB(1.0, y);
    double r = DMUL(DMUL((double)r1[x1], nx), ny);
    r = DADD(r, DMUL(DMUL((double)r1[x2], x), ny));
    r = DADD(r, DMUL(DMUL((double)r2[x1], nx), y));
    r = DADD(r, DMUL(DMUL((double)r2[x2], x), y));
    return r;
}

// rotate90 (fast/decode.c:61-71)
__device__ __forceinline__ uint64_t rotate90(uint64_t w0, int64_t nbits) {
    int64_t p = nbits;
    uint64_t l = 0;
    if (nbits % 4 == 1) { p = nbits - 1; l = 1; }
    uint64_t w = ((w0 >> l) << (p / 4 + (int64_t)l)) | (w0 >> (3 * p / 4 + (int64_t)l) << l) | (w0 & l);
    return w & (((uint64_t)1 << nbits) - 1);
}

// Bit-sampler points, 8 wab + nbits: at most 8 * 12 + 64 for the families
// decoded here (total width up to 12, width at border no more).
constexpr int kMaxSamples = 160;

// The border sampling patterns of quad_sample (fast/decode.c:296-305):
// start x, start y, step x, step y, and whether the side is white.
__device__ __forceinline__ void border_pattern(int pi, int wab, float *x0, float *y0, float *dx, float *dy,
                                               bool *white) {
    const float hi = (float)((double)wab + 0.5), lo = (float)((double)wab - 0.5);
    const bool vert = pi < 4;  // the first four walk down a column
    const int k = pi & 3;
    const float a = k == 0 ? -0.5f : k == 1 ? 0.5f : k == 2 ? hi : lo;
    *x0 = vert ? a : 0.5f;
    *y0 = vert ? 0.5f : a;
    *dx = vert ? 0.0f : 1.0f;
    *dy = vert ? 1.0f : 0.0f;
    *white = (k & 1) == 0;
}


// Persistent: the quads of every frame, block b taking global indices b,
// b + G, ... (a frame cursor). NT threads a quad: 256 (a wave per edge) or
// one wave (the edges in turn). The edges' sample sums of refine_edges
// (fast/decode.c:657-761) up to the line's angle: per edge, the mean point
// and atan2's two arguments into qline (k_lines takes the angle; its
// library calls need far more registers than the rest of the decode).
template <int NT>
__global__ __launch_bounds__(NT) void k_refine(const Frame *frames, const DevParams *P, const WorkQueue *Q) {
    __shared__ double wsx[4][16], wsy[4][16];  // an edge's refined sample points, 16 at a time
    __shared__ int wsv[4][16];
    __shared__ float qp[4][2];
    const int t = threadIdx.x, lane = t & 63;
    const uint32_t total = Q->total[2];
    const uint32_t *off = Q->off[2];
    uint32_t fi = 0;
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
        fi = frame_of_wave(off, Q->nframes, g, fi);
        const Frame &F = frames[fi];
        const uint32_t qi = g - off[fi];
        View v = {F.img, F.stride, F.W, F.H};
        if (F.blurred && !(P->decimate > 1.0f)) v = View{F.qim, F.W, F.W, F.H};
        QuadRec q = F.quads[qi];
        if (t == 0) {
            for (int i = 0; i < 4; i++) {
                float f = P->decimate;
                qp[i][0] = f > 1.0f ? __fmul_rn(q.p[i][0], f) : q.p[i][0];
                qp[i][1] = f > 1.0f ? __fmul_rn(q.p[i][1], f) : q.p[i][1];
            }
        }
        bsync_t<NT>();
        if (P->refine) {
            // refine_edges (fast/decode.c:657-761): edge e on wave e of a
            // 256-thread block, or the edges in turn on a one-wave block;
            // four lanes a sample (one partial sum each), 16 samples at a
            // time; lane 0 adds the samples up in order
            auto refine_edge = [&](const int e) {
                const int64_t range = (int64_t)(P->decimate + 1.0f);
                const int64_t max_steps = 2 * 4 * range + 1;
                int64_t vsteps = max_steps - max_steps % 4;
                if (vsteps < 0) vsteps = 0;
                const int ea = e, eb = (e + 1) & 3;
                double nx = (double)__fsub_rn(qp[eb][1], qp[ea][1]);
                double ny = (double)__fadd_rn(-qp[eb][0], qp[ea][0]);
                const double mag = __dsqrt_rn(DADD(DMUL(nx, nx), DMUL(ny, ny)));
                nx = DDIV(nx, mag);
                ny = DDIV(ny, mag);
                if (q.reversed) { nx = -nx; ny = -ny; }
                int64_t ns = ctrunc(DDIV(mag, 8.0));
                if (ns < 16) ns = 16;
                double mx = 0.0, my = 0.0, mxx = 0.0, mxy = 0.0, myy = 0.0, nn = 0.0;
                const int part = lane & 3, sl = lane >> 2;
                for (int64_t c0 = 0; c0 < ns; c0 += 16) {
                    const int64_t si = c0 + sl;
                    double x0 = 0.0, y0 = 0.0, ln = 0.0, lc = 0.0;
                    if (si < ns) {
                        double alpha = DDIV(DADD(1.0, (double)si), (double)(ns + 1));
                        x0 = DADD(DMUL(alpha, (double)qp[ea][0]), DMUL(DSUB(1.0, alpha), (double)qp[eb][0]));
                        y0 = DADD(DMUL(alpha, (double)qp[ea][1]), DMUL(DSUB(1.0, alpha), (double)qp[eb][1]));
                        refine_partial(v, range, vsteps, x0, y0, nx, ny, part, &ln, &lc);
                    }
                    // the quad's four partial sums in every lane of the quad
                    // (DPP quad broadcasts)
                    auto qb = [](double v, auto kc) {
                        constexpr int ctl = decltype(kc)::value * 0x55;
                        const uint64_t b = (uint64_t)__double_as_longlong(v);
                        const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)b, ctl, 0xF, 0xF, false);
                        const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), ctl, 0xF, 0xF, false);
                        return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
                    };
                    const double lns[4] = {qb(ln, std::integral_constant<int, 0>{}), qb(ln, std::integral_constant<int, 1>{}),
                                           qb(ln, std::integral_constant<int, 2>{}), qb(ln, std::integral_constant<int, 3>{})};
                    const double lcs[4] = {qb(lc, std::integral_constant<int, 0>{}), qb(lc, std::integral_constant<int, 1>{}),
                                           qb(lc, std::integral_constant<int, 2>{}), qb(lc, std::integral_constant<int, 3>{})};
                    if (part == 0) {
                        int ok = 0;
                        if (si < ns) {
                            double bx = 0, by = 0;
                            ok = refine_finish(v, range, max_steps, vsteps, x0, y0, nx, ny, lns, lcs, &bx, &by) ? 1 : 0;
                            wsx[e][sl] = bx;
                            wsy[e][sl] = by;
                        }
                        wsv[e][sl] = ok;
                    }
                    bsync_t<64>();
                    if (lane == 0) {
                        const int64_t len = ns - c0 < 16 ? ns - c0 : 16;
                        for (int64_t k = 0; k < len; k++) {
                            if (!wsv[e][k]) continue;
                            const double bx = wsx[e][k], by = wsy[e][k];
                            mx = DADD(mx, bx); my = DADD(my, by);
                            mxx = DADD(mxx, DMUL(bx, bx)); mxy = DADD(mxy, DMUL(bx, by)); myy = DADD(myy, DMUL(by, by));
                            nn = DADD(nn, 1.0);
                        }
                    }
                    bsync_t<64>();
                }
                if (lane == 0) {
                    double ex = DDIV(mx, nn), ey = DDIV(my, nn);
                    double cxx = DSUB(DDIV(mxx, nn), DMUL(ex, ex));
                    double cxy = DSUB(DDIV(mxy, nn), DMUL(ex, ey));
                    double cyy = DSUB(DDIV(myy, nn), DMUL(ey, ey));
                    double *ln = F.qline + (size_t)qi * 16 + e * 4;
                    ln[0] = ex;
                    ln[1] = ey;
                    ln[2] = (double)(float)DMUL(-2.0, cxy);
                    ln[3] = (double)(float)DSUB(cyy, cxx);
                }
            };
            if constexpr (NT == 64) {
                for (int e = 0; e < 4; e++) refine_edge(e);
            } else {
                refine_edge(t >> 6);
            }
            bsync_t<NT>();
        }
        bsync_t<NT>();
    }
}

// atan2, cos and sin of refine_edges' lines (fast/decode.c:750-760), a
// thread per quad edge of every frame (the quads' list 2, four edges each):
// qline's atan2 arguments become the line's direction.
__global__ __launch_bounds__(256) void k_lines(const Frame *frames, const WorkQueue *Q) {
    const uint32_t total = 4 * Q->total[2];
    const uint32_t *off = Q->off[2];
    uint32_t fi = 0;
    for (uint32_t g = blockIdx.x * 256 + threadIdx.x; g < total; g += gridDim.x * 256) {
        fi = frame_of(off, Q->nframes, g >> 2, fi);  // (4 off[f] <= g iff off[f] <= g / 4)
        double *ln = frames[fi].qline + (size_t)(g - 4 * off[fi]) * 4;
        const float th2 = cr::atan2f_cr((float)ln[2], (float)ln[3]);
        const double theta = DMUL(0.5, (double)th2);
        ln[2] = (double)cr::cosf_cr((float)theta);
        ln[3] = (double)cr::sinf_cr((float)theta);
    }
}

// The refinement's correctly rounded atan2f, cosf and sinf on caller
// arrays (a check of cr_trig.h): a thread an element; slow: the
// double-double path for every element.
__global__ __launch_bounds__(256) void k_debug_trig(const float *y, const float *x, const float *th, int64_t n,
                                                    float *at, float *co, float *si, int slow) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    at[i] = cr::atan2f_cr(y[i], x[i], slow != 0);
    co[i] = cr::cosf_cr(th[i], slow != 0);
    si[i] = cr::sinf_cr(th[i], slow != 0);
}

void launch_debug_trig(hipStream_t st, const float *y, const float *x, const float *th, int64_t n, float *at,
                       float *co, float *si, int slow) {
    if (n > 0)
        hipLaunchKernelGGL(k_debug_trig, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, y, x, th, n, at, co, si,
                           slow);
}

// Persistent: the quads of every frame, block b taking global indices b,
// b + G, ... (a frame cursor).
// NT threads a quad: 256 or one wave. The corners from the refined lines
// (k_refine, k_lines), the homography, the samples and the decoding.
template <int NT>
__global__ __launch_bounds__(NT, NT == 256 ? 4 : 1) void k_decode(const Frame *frames, const DevParams *P, const WorkQueue *Q) {
    __shared__ double smp[3][kMaxSamples];  // tag x, tag y, value
    __shared__ int sflag[kMaxSamples];
    __shared__ float qp[4][2];
    __shared__ double lines[4][4];
    __shared__ double sH[9], sA[72], sLU[9];
    __shared__ int s_hok, s_piv, s_polar;
    __shared__ Gray g_white, g_black;
    __shared__ double vals[12 * 12], sharp[12 * 12];
    __shared__ uint64_t s_rc;
    __shared__ float s_margin;
    __shared__ int s_mj[16], s_mh[16];
    const int t = threadIdx.x, lane = t & 63;
    const int nfam = P->nfam;
    const uint32_t total = Q->total[2];
    const uint32_t *off = Q->off[2];
    uint32_t fi = 0;
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
        fi = frame_of_wave(off, Q->nframes, g, fi);
        const Frame &F = frames[fi];
        const uint32_t qi = g - off[fi];
        View v = {F.img, F.stride, F.W, F.H};
        if (F.blurred && !(P->decimate > 1.0f)) v = View{F.qim, F.W, F.W, F.H};
        QuadRec q = F.quads[qi];
        if (t == 0) {
            for (int i = 0; i < 4; i++) {
                float f = P->decimate;
                qp[i][0] = f > 1.0f ? __fmul_rn(q.p[i][0], f) : q.p[i][0];
                qp[i][1] = f > 1.0f ? __fmul_rn(q.p[i][1], f) : q.p[i][1];
            }
        }
        bsync_t<NT>();
        if (P->refine) {
            // the refined lines (k_refine, k_lines) and their intersections
            if (t < 16) lines[t >> 2][t & 3] = F.qline[(size_t)qi * 16 + t];
            bsync_t<NT>();
            if (t == 0) {
                for (int i = 0; i < 4; i++) {
                    int j = (i + 1) & 3;
                    double a00 = lines[i][3], a01 = -lines[j][3], a10 = -lines[i][2], a11 = lines[j][2];
                    double b0 = DADD(-lines[i][0], lines[j][0]), b1 = DADD(-lines[i][1], lines[j][1]);
                    double det = DSUB(DMUL(a00, a11), DMUL(a10, a01));
                    if (fabs(det) > 0.001) {
                        double w00 = DDIV(a11, det), w01 = DDIV(-a01, det);
                        double l0 = DADD(DMUL(w00, b0), DMUL(w01, b1));
                        qp[j][0] = (float)DADD(lines[i][0], DMUL(l0, a00));
                        qp[j][1] = (float)DADD(lines[i][1], DMUL(l0, a10));
                    }
                }
            }
            bsync_t<NT>();
        }
        // homography (fast/decode.c:198-210) and its invertibility
        bool hok = homography_block<NT>(qp, sA, &s_piv, sH);
        if (t == 0) s_hok = hok && invertible(sH, sLU) ? 1 : 0;
        bsync_t<NT>();
        for (int fam = 0; fam < nfam; fam++) {
            if (t == 0) F.dflag[qi * nfam + fam] = 0;
            const DevFamily &fm = P->fam[fam];
            if (!s_hok || fm.reversed != q.reversed) continue;
            const int wab = fm.wab;
            const float wf = (float)wab;
            const int nb = 8 * wab;
            if (nb + fm.nbits > kMaxSamples || fm.tw > 12) continue;
            // border samples (fast/decode.c:310-326), in parallel
            for (int k = t; k < nb; k += NT) {
                const int pi = k / wab, i = k % wab;
                float bx0, by0, bdx, bdy;
                bool white;
                border_pattern(pi, wab, &bx0, &by0, &bdx, &bdy, &white);
                double tx01 = (double)fdiv_cr(__fadd_rn(bx0, __fmul_rn((float)i, bdx)), wf);
                double ty01 = (double)fdiv_cr(__fadd_rn(by0, __fmul_rn((float)i, bdy)), wf);
                double tx = DMUL(2.0, DSUB(tx01, 0.5)), ty = DMUL(2.0, DSUB(ty01, 0.5));
                double zz = DADD(DADD(DMUL(sH[6], tx), DMUL(sH[7], ty)), sH[8]);
                double px = DDIV(DADD(DADD(DMUL(sH[0], tx), DMUL(sH[1], ty)), sH[2]), zz);
                double py = DDIV(DADD(DADD(DMUL(sH[3], tx), DMUL(sH[4], ty)), sH[5]), zz);
                int64_t ix = ctrunc(px), iy = ctrunc(py);
                int f = 0;
                double val = 0.0;
                if (!(ix < 0 || iy < 0 || ix >= v.w || iy >= v.h)) {
                    f = 1 | (white ? 2 : 0);
                    val = (double)v.im[iy * v.s + ix];
                }
                smp[0][k] = tx;
                smp[1][k] = ty;
                smp[2][k] = val;
                sflag[k] = f;
            }
            // bit samples (fast/decode.c:346-362)
            const int mc = (wab - fm.tw) / 2;
            for (int i = t; i < fm.nbits; i += NT) {
                int k = nb + i;
                double tx01 = DDIV(DADD((double)fm.bit_x[i], 0.5), (double)wab);
                double ty01 = DDIV(DADD((double)fm.bit_y[i], 0.5), (double)wab);
                double tx = DMUL(2.0, DSUB(tx01, 0.5)), ty = DMUL(2.0, DSUB(ty01, 0.5));
                double zz = DADD(DADD(DMUL(sH[6], tx), DMUL(sH[7], ty)), sH[8]);
                double px = DDIV(DADD(DADD(DMUL(sH[0], tx), DMUL(sH[1], ty)), sH[2]), zz);
                double py = DDIV(DADD(DADD(DMUL(sH[3], tx), DMUL(sH[4], ty)), sH[5]), zz);
                smp[0][k] = tx;
                smp[1][k] = ty;
                smp[2][k] = pixel_value(v, px, py);
            }
            bsync_t<NT>();
            // the gray models, in sample order (fast/decode.c:229-265)
            if (t == 0) {
                Gray wh, bl;
                for (int i = 0; i < 9; i++) { wh.a[i] = 0; bl.a[i] = 0; }
                for (int i = 0; i < 3; i++) { wh.b[i] = 0; bl.b[i] = 0; wh.c[i] = 0; bl.c[i] = 0; }
                for (int k = 0; k < nb; k++) {
                    const int f = sflag[k];
                    if (!(f & 1)) continue;
                    if (f & 2) gray_add(wh, smp[0][k], smp[1][k], smp[2][k]);
                    else gray_add(bl, smp[0][k], smp[1][k], smp[2][k]);
                }
                gray_solve(wh);
                if (wab > 1) gray_solve(bl);
                else { bl.c[0] = 0.0; bl.c[1] = 0.0; bl.c[2] = DDIV(bl.b[2], 4.0); }
                double wi = DADD(DADD(DMUL(wh.c[0], 0.0), DMUL(wh.c[1], 0.0)), wh.c[2]);
                double bi = DADD(DADD(DMUL(bl.c[0], 0.0), DMUL(bl.c[1], 0.0)), bl.c[2]);
                s_polar = (DSUB(wi, bi) < 0.0) == (fm.reversed != 0) ? 1 : 0;
                g_white = wh;
                g_black = bl;
            }
            bsync_t<NT>();
            if (s_polar) {
                const int tw = fm.tw;
                for (int i = t; i < tw * tw; i += NT) vals[i] = 0.0;
                bsync_t<NT>();
                // each bit's value against the models' threshold, in its cell
                for (int i = t; i < fm.nbits; i += NT) {
                    double val = smp[2][nb + i];
                    if (val == -1.0) continue;
                    double tx = smp[0][nb + i], ty = smp[1][nb + i];
                    const double *bc = g_black.c, *wc = g_white.c;
                    double th = DDIV(DADD(DADD(DADD(DMUL(bc[0], tx), DMUL(bc[1], ty)), bc[2]),
                                          DADD(DADD(DMUL(wc[0], tx), DMUL(wc[1], ty)), wc[2])),
                                     2.0);
                    vals[tw * (fm.bit_y[i] - mc) + (fm.bit_x[i] - mc)] = DSUB(val, th);
                }
                bsync_t<NT>();
                // sharpening (fast/decode.c:370-385), each cell in tap order
                for (int i = t; i < tw * tw; i += NT) {
                    const int y = i / tw, x = i % tw;
                    double acc = 0.0;
                    if (y > 0) acc = DADD(acc, -vals[(y - 1) * tw + x]);
                    if (x > 0) acc = DADD(acc, -vals[y * tw + x - 1]);
                    acc = DADD(acc, DMUL(vals[y * tw + x], 4.0));
                    if (x + 1 < tw) acc = DADD(acc, -vals[y * tw + x + 1]);
                    if (y + 1 < tw) acc = DADD(acc, -vals[(y + 1) * tw + x]);
                    sharp[i] = acc;
                }
                bsync_t<NT>();
                for (int i = t; i < tw * tw; i += NT) vals[i] = DADD(vals[i], DMUL(P->sharpening, sharp[i]));
                bsync_t<NT>();
                if (t == 0) {
                    float bs = 0.0f, ws = 0.0f, bc = 1.0f, wc = 1.0f;
                    uint64_t rc = 0;
                    for (int i = 0; i < fm.nbits; i++) {
                        rc <<= 1;
                        double vv = vals[(fm.bit_y[i] - mc) * tw + (fm.bit_x[i] - mc)];
                        if (vv > 0.0) { ws = (float)DADD((double)ws, vv); wc = __fadd_rn(wc, 1.0f); rc |= 1; }
                        else { bs = (float)DSUB((double)bs, vv); bc = __fadd_rn(bc, 1.0f); }
                    }
                    double xm = (double)fdiv_cr(ws, wc), ym = (double)fdiv_cr(bs, bc);
                    double mg = xm != xm ? ym : ym != ym ? xm : xm < ym ? xm : ym;
                    s_rc = rc;
                    s_margin = (float)mg;
                }
                bsync_t<NT>();
                // quick decode (fast/decode.c:87-107): the 16 (rotation,
                // chunk) lookups at once, thread 4 r + i each; the first
                // match in the reference's order (rotation, chunk, entry)
                // is the lowest thread's
                if (t < 16) {
                    const int r = t >> 2, i = t & 3;
                    uint64_t rcode = s_rc;
                    for (int k = 0; k < r; k++) rcode = rotate90(rcode, fm.nbits);
                    const uint64_t cmask = (uint64_t)(fm.capacity - 1);
                    const int64_t val = (int64_t)((rcode >> (i * fm.chunk_size)) & cmask);
                    const int64_t ob = i * (fm.capacity + 1) + val;
                    const int64_t basei = (int64_t)i * fm.ncodes;
                    int mj = -1, mh = 255;
                    for (int64_t j = basei + fm.offsets[ob]; j < basei + fm.offsets[ob + 1]; j++) {
                        const int64_t hm = __popcll(fm.codes[j] ^ rcode);
                        if (hm <= fm.maxhamming) {
                            mj = (int)j;
                            mh = (int)hm;
                            break;
                        }
                    }
                    s_mj[t] = mj;
                    s_mh[t] = mh;
                }
                bsync_t<NT>();
                if (t == 0) {
                    const float margin = s_margin;
                    int64_t id = 65535, ham = 255, rot = 0;
                    for (int k = 0; k < 16; k++) {
                        if (s_mj[k] >= 0) {
                            id = fm.ids[s_mj[k]];
                            ham = s_mh[k];
                            rot = k >> 2;
                            break;
                        }
                    }
                    if (margin >= 0.0f && ham < 255) {
                        // make_detection (fast/detector.c:83-107)
                        at_detection d;
                        double c = P->rot_cs[rot][0], sn = P->rot_cs[rot][1];
                        const double rr[9] = {c, -sn, 0.0, sn, c, 0.0, 0.0, 0.0, 1.0};
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++) {
                                double acc = 0.0;
                                for (int k = 0; k < 3; k++) acc = DADD(acc, DMUL(sH[i * 3 + k], rr[k * 3 + j]));
                                d.H[i * 3 + j] = acc;
                            }
                        d.family = fam;
                        d.id = (int32_t)id;
                        d.hamming = (int32_t)ham;
                        d.decision_margin = margin;
                        project(d.H, 0.0, 0.0, &d.c[0], &d.c[1]);
                        for (int i = 0; i < 4; i++)
                            project(d.H, i == 1 || i == 2 ? 1.0 : -1.0, i < 2 ? 1.0 : -1.0, &d.p[i][0], &d.p[i][1]);
                        F.dtmp[qi * nfam + fam] = d;
                        F.dflag[qi * nfam + fam] = 1;
                    }
                }
            }
            bsync_t<NT>();
        }
        if (t == 0) {
            QuadRec &o = F.quads[qi];
            o.hok = s_hok;
            for (int i = 0; i < 9; i++) o.H[i] = sH[i];
        }
        bsync_t<NT>();
    }
}

// ------------------------------------------------------------ reconcile

struct G2Line {
    double p[2], u[2];
};

__device__ G2Line g2_line(const double *a, const double *b) {
    double u0 = DSUB(b[0], a[0]), u1 = DSUB(b[1], a[1]);
    double mag = (double)fsqrt_cr((float)DADD(DMUL(u0, u0), DMUL(u1, u1)));
    G2Line l;
    l.p[0] = a[0]; l.p[1] = a[1]; l.u[0] = DDIV(u0, mag); l.u[1] = DDIV(u1, mag);
    return l;
}
__device__ double g2_coord(const G2Line &l, const double *q) {
    return DADD(DMUL(DSUB(q[0], l.p[0]), l.u[0]), DMUL(DSUB(q[1], l.p[1]), l.u[1]));
}
__device__ bool g2_isect(const G2Line &la, const G2Line &lb, double out[2]) {
    double m00 = la.u[0], m01 = -lb.u[0], m10 = la.u[1], m11 = -lb.u[1];
    double det = DSUB(DMUL(m00, m11), DMUL(m01, m10));
    if (fabs(det) < 0.00000001) return false;
    double i00 = DDIV(m11, det), i01 = DDIV(-m01, det);
    double b00 = DSUB(lb.p[0], la.p[0]), b10 = DSUB(lb.p[1], la.p[1]);
    double x00 = DADD(DMUL(i00, b00), DMUL(i01, b10));
    out[0] = DADD(DMUL(la.u[0], x00), la.p[0]);
    out[1] = DADD(DMUL(la.u[1], x00), la.p[1]);
    return true;
}
__device__ bool g2_seg(const double *a0, const double *a1, const double *b0, const double *b1) {
    G2Line la = g2_line(a0, a1), lb = g2_line(b0, b1);
    double t[2];
    if (!g2_isect(la, lb, t)) return false;
    double a = g2_coord(la, la.p), b = g2_coord(la, a1), c = g2_coord(la, t);
    if ((c < a && c < b) || (c > a && c > b)) return false;
    a = g2_coord(lb, lb.p); b = g2_coord(lb, b1); c = g2_coord(lb, t);
    if ((c < a && c < b) || (c > a && c > b)) return false;
    return true;
}
__device__ bool g2_contains(const double poly[4][2], const double q[2]) {
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
                double nx = DSUB(p[1], q[1]), ny = DADD(-p[0], q[0]);
                double dot = DADD(DMUL(nx, DSUB(p0[0], q[0])), DMUL(ny, DSUB(p0[1], q[1])));
                acc += dot < 0.0 ? -2 : 2;
            }
        }
        last = quadrant;
    }
    return acc >= 2 || acc <= -2;
}
__device__ bool g2_overlaps(const double pa[4][2], const double pb[4][2]) {
    for (int ia = 0; ia < 4; ia++)
        for (int ib = 0; ib < 4; ib++)
            if (g2_seg(pa[ia], pa[(ia + 1) % 4], pb[ib], pb[(ib + 1) % 4])) return true;
    double qb[2] = {DDIV(DADD(DADD(pb[0][0], pb[1][0]), pb[2][0]), 3.0), DDIV(DADD(DADD(pb[0][1], pb[1][1]), pb[2][1]), 3.0)};
    if (g2_contains(pa, qb)) return true;
    double qa[2] = {DDIV(DADD(DADD(pa[0][0], pa[1][0]), pa[2][0]), 3.0), DDIV(DADD(DADD(pa[0][1], pa[1][1]), pa[2][1]), 3.0)};
    return g2_contains(pb, qa);
}
__device__ __forceinline__ int prefer(int p, double a, double b) { return p != 0 ? p : a < b ? -1 : b < a ? 1 : 0; }

// The last step of a frame, by one wave: the detections of the decoded
// slots (src[k] where flag[k], all nslots when flag is null), in slot
// order, reconciled (fast/decode.c:860-891) and stably sorted by id
// (fast/detector.c:164-179) into out; returns their count. Scratch: A the
// kept slots in order, K their (id, family) keys, B the list ordered by
// key (stable), R removal flags, S the survivors (each nslots u32).
//
// Reconcile only compares detections of the same id and family, and a
// removal never changes the order of the others, so the reference's one
// pass over the list gives, within each (id, family) group, what the same
// pass over that group alone gives: each lane runs the pass over a group.
__device__ __forceinline__ uint64_t wave_lt_mask() { return (1ull << (threadIdx.x & 63)) - 1ull; }

__device__ uint32_t finish_list(const at_detection *src, const uint32_t *flag, uint32_t nslots, at_detection *out,
                                uint32_t *A, uint32_t *K, uint32_t *B, uint32_t *R, uint32_t *S) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt = wave_lt_mask();
    // the decoded slots in order (ballot compaction)
    uint32_t n = 0;
    for (uint32_t c0 = 0; c0 < nslots; c0 += 64) {
        const uint32_t k = c0 + lane;
        const bool f = k < nslots && (flag == nullptr || flag[k] != 0u);
        const uint64_t b = __ballot(f);
        if (f) {
            const uint32_t at = n + (uint32_t)__popcll(b & lt);
            A[at] = k;
            K[at] = (uint32_t)src[k].id | ((uint32_t)src[k].family << 16);
            R[at] = 0u;
        }
        n += (uint32_t)__popcll(b);
    }
    __threadfence_block();
    // B: positions ordered by key, stable (rank of each)
    for (uint32_t i = lane; i < n; i += 64) {
        const uint32_t ki = K[i];
        uint32_t r = 0;
        for (uint32_t j = 0; j < n; j++) {
            const uint32_t kj = K[j];
            r += (kj < ki || (kj == ki && j < i)) ? 1u : 0u;
        }
        B[r] = i;
    }
    __threadfence_block();
    // one lane a group: the reference's pass over the group's detections
    for (uint32_t g0 = 0; g0 < n; g0 += 64) {
        const uint32_t g = g0 + lane;
        if (g >= n || (g > 0 && K[B[g]] == K[B[g - 1]])) continue;  // not a group's first
        uint32_t e = g + 1;
        while (e < n && K[B[e]] == K[B[g]]) e++;
        for (uint32_t i0 = g; i0 < e; i0++) {
            const uint32_t p0 = B[i0];
            if (R[p0]) continue;
            const at_detection &d0 = src[A[p0]];
            for (uint32_t i1 = i0 + 1; i1 < e; i1++) {
                const uint32_t p1 = B[i1];
                if (R[p1]) continue;
                const at_detection &d1 = src[A[p1]];
                if (!g2_overlaps(d0.p, d1.p)) continue;
                int pr = 0;
                pr = prefer(pr, (double)d0.hamming, (double)d1.hamming);
                pr = prefer(pr, -(double)d0.decision_margin, -(double)d1.decision_margin);
                for (int i = 0; i < 4; i++) {
                    pr = prefer(pr, d0.p[i][0], d1.p[i][0]);
                    pr = prefer(pr, d0.p[i][1], d1.p[i][1]);
                }
                if (pr < 0) {
                    R[p1] = 1u;  // the later one goes; the pass goes on from the next
                    continue;
                }
                R[p0] = 1u;  // the earlier one goes: its pass ends
                break;
            }
        }
    }
    __threadfence_block();
    // the survivors in order
    uint32_t m = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += 64) {
        const uint32_t i = c0 + lane;
        const bool f = i < n && R[i] == 0u;
        const uint64_t b = __ballot(f);
        if (f) S[m + (uint32_t)__popcll(b & lt)] = A[i];
        m += (uint32_t)__popcll(b);
    }
    __threadfence_block();
    // stable by id: the rank of each survivor, its record copied there
    for (uint32_t i = lane; i < m; i += 64) {
        const int32_t id = src[S[i]].id;
        uint32_t r = 0;
        for (uint32_t j = 0; j < m; j++) {
            const int32_t o = src[S[j]].id;
            r += (o < id || (o == id && j < i)) ? 1u : 0u;
        }
        out[r] = src[S[i]];
    }
    return m;
}

__global__ __launch_bounds__(64) void k_finish(const Frame *frames, const DevParams *P) {
    const Frame F = frames[blockIdx.y];
    const uint32_t nslots = F.counters[kNQuads] * (uint32_t)P->nfam;
    // scratch: the corner search's arrays, free by now
    uint32_t *A = F.midx, *K = reinterpret_cast<uint32_t *>(F.errs), *R = reinterpret_cast<uint32_t *>(F.filt);
    const uint32_t n = finish_list(F.dtmp, F.dflag, nslots, F.dfin, A, K, K + nslots, R, R + nslots);
    if (threadIdx.x == 0) F.counters[kNDets] = n;
    __threadfence_block();
    const uint32_t words = min(n, (uint32_t)kInlineDets) * (uint32_t)(sizeof(at_detection) / 8);
    const uint64_t *s = reinterpret_cast<const uint64_t *>(F.dfin);
    uint64_t *d = reinterpret_cast<uint64_t *>(F.inl);
    for (uint32_t i = threadIdx.x; i < words; i += 64) d[i] = s[i];
}

// The last step on a caller's list (debug entry point): buf holds the n
// records, then n records of output, then 5 n u32 of scratch.
__global__ __launch_bounds__(64) void k_debug_finish(uint8_t *buf, uint32_t n, int64_t *nout) {
    at_detection *in = reinterpret_cast<at_detection *>(buf), *out = in + n;
    uint32_t *s = reinterpret_cast<uint32_t *>(out + n);
    const uint32_t m = finish_list(in, nullptr, n, out, s, s + n, s + 2 * n, s + 3 * n, s + 4 * n);
    if (threadIdx.x == 0) *nout = m;
}

void launch_debug_finish(hipStream_t st, uint8_t *buf, uint32_t n, int64_t *nout) {
    hipLaunchKernelGGL(k_debug_finish, dim3(1), dim3(64), 0, st, buf, n, nout);
}

__global__ void k_init(const Frame *frames, WorkQueue *Q) {
    const Frame F = frames[blockIdx.y];
    if (threadIdx.x < kNCounters) F.counters[threadIdx.x] = 0;
    if (F.skip) const_cast<uint8_t *>(F.img)[threadIdx.x] = 0;  // 64 bytes: the 8x8 zero frame
    if (blockIdx.y == 0 && threadIdx.x < kLists) Q->total[threadIdx.x] = 0;
}

void launch_init(hipStream_t st, const Frame *dframes, int nframes, WorkQueue *Q) {
    hipLaunchKernelGGL(k_init, dim3(1, nframes), dim3(64), 0, st, dframes, Q);
}

void warm_decode_grid(int fit_blocks) {
    persistent_grid(reinterpret_cast<const void *>(&k_decode<64>), 64, 0, 4 * fit_blocks);
    persistent_grid(reinterpret_cast<const void *>(&k_decode<256>), 256, 0, fit_blocks);
    persistent_grid(reinterpret_cast<const void *>(&k_refine<64>), 64, 0, 4 * fit_blocks);
    persistent_grid(reinterpret_cast<const void *>(&k_refine<256>), 256, 0, fit_blocks);
    persistent_grid(reinterpret_cast<const void *>(&k_lines), 256, 0, fit_blocks);
}

void launch_decode_stage(hipStream_t st, const Frame *dframes, const DevParams *dparams, int nframes, int blocks,
                         bool P_refine, WorkQueue *Q, Marks *mk) {
    // a wave a quad (four times as many quads in flight as 256-thread
    // blocks); for a few frames, whose quads all fit at once, a 256-thread
    // block a quad, its edges refined at the same time
    if (P_refine) {
        if (nframes > 4)
            hipLaunchKernelGGL(k_refine<64>,
                               dim3(persistent_grid(reinterpret_cast<const void *>(&k_refine<64>), 64, 0, 4 * blocks)),
                               dim3(64), 0, st, dframes, dparams, Q);
        else
            hipLaunchKernelGGL(k_refine<256>,
                               dim3(persistent_grid(reinterpret_cast<const void *>(&k_refine<256>), 256, 0, blocks)),
                               dim3(256), 0, st, dframes, dparams, Q);
        mk->mark(st, "decode.refine");
        hipLaunchKernelGGL(k_lines, dim3(persistent_grid(reinterpret_cast<const void *>(&k_lines), 256, 0, blocks)),
                           dim3(256), 0, st, dframes, Q);
        mk->mark(st, "decode.lines");
    }
    if (nframes > 4)
        hipLaunchKernelGGL(k_decode<64>,
                           dim3(persistent_grid(reinterpret_cast<const void *>(&k_decode<64>), 64, 0, 4 * blocks)),
                           dim3(64), 0, st, dframes, dparams, Q);
    else
        hipLaunchKernelGGL(k_decode<256>,
                           dim3(persistent_grid(reinterpret_cast<const void *>(&k_decode<256>), 256, 0, blocks)),
                           dim3(256), 0, st, dframes, dparams, Q);
    mk->mark(st, "decode.sample");
    hipLaunchKernelGGL(k_finish, dim3(1, nframes), dim3(64), 0, st, dframes, dparams);
    mk->mark(st, "finish");
}

}  // namespace atg
