// Seeded, bit-reproducible synthetic AprilTag frames (host code).
//
// Everything is integer arithmetic or IEEE +,-,*,/ in double (no libm
// transcendental), compiled without FP contraction, so a seed gives the
// same bytes on any x86-64 host:
//   - PRNG: splitmix64;
//   - rotations: the rational parametrisation cos = (1-t^2)/(1+t^2),
//     sin = 2t/(1+t^2) of a uniform t in [-1, 1], times a quarter turn;
//   - perspective: a 3D rotation of the tag plane by two tilts (rational
//     as above), projected with a pinhole camera;
//   - tags: cells of the family layout (the image of apriltag_to_image,
//     apriltag.c:271-297: tw x tw cells, quiet zone white), rendered by
//     4x4 supersampling through the inverse homography;
//   - blur: a separable integer kernel (weights from a rational
//     approximation of the Gaussian), noise: Irwin-Hall sums of 12
//     uniform 16-bit integers.
#include <cstdint>
#include <cstring>
#include <vector>

#include "atg.h"
#include "atg_internal.h"

#pragma clang fp contract(off)

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    // uniform integer in [lo, hi]
    int64_t range(int64_t lo, int64_t hi) { return lo + (int64_t)(next() % (uint64_t)(hi - lo + 1)); }
    // uniform double in [0, 1)
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double lo, double hi) { return lo + (hi - lo) * unit(); }
    // approximately N(0, 1): Irwin-Hall of 12 uniforms, in 1/65536 units
    int64_t gauss16() {
        int64_t s = 0;
        for (int i = 0; i < 12; i++) s += (int64_t)(next() & 0xffff);
        return s - 6 * 65536;
    }
};

struct Rot {
    double c, s;
};

// a rotation by a uniform-ish angle in [-max, max], as a rational point on
// the circle: t = tan(angle/2) ~ angle/2 for the small tilts used here
Rot rational_rot(double t) {
    double d = 1.0 + t * t;
    return {(1.0 - t * t) / d, 2.0 * t / d};
}

Rot random_rotation(Rng &r) {
    Rot a = rational_rot(r.uniform(-1.0, 1.0));  // [-90, 90] degrees
    switch (r.range(0, 3)) {
    case 1: return {-a.s, a.c};
    case 2: return {-a.c, -a.s};
    case 3: return {a.s, -a.c};
    default: return a;
    }
}

// 3x3 homography solve from 4 correspondences (tag (u,v) -> image (x,y))
bool solve_h(const double src[4][2], const double dst[4][2], double H[9]) {
    double A[8][9];
    for (int i = 0; i < 4; i++) {
        double u = src[i][0], v = src[i][1], x = dst[i][0], y = dst[i][1];
        double r0[9] = {u, v, 1, 0, 0, 0, -u * x, -v * x, x};
        double r1[9] = {0, 0, 0, u, v, 1, -u * y, -v * y, y};
        memcpy(A[2 * i], r0, sizeof r0);
        memcpy(A[2 * i + 1], r1, sizeof r1);
    }
    for (int c = 0; c < 8; c++) {
        int p = c;
        for (int r = c + 1; r < 8; r++)
            if ((A[r][c] < 0 ? -A[r][c] : A[r][c]) > (A[p][c] < 0 ? -A[p][c] : A[p][c])) p = r;
        if (A[p][c] == 0.0) return false;
        if (p != c)
            for (int k = 0; k < 9; k++) { double t = A[c][k]; A[c][k] = A[p][k]; A[p][k] = t; }
        for (int r = 0; r < 8; r++) {
            if (r == c) continue;
            double f = A[r][c] / A[c][c];
            for (int k = c; k < 9; k++) A[r][k] -= f * A[c][k];
        }
    }
    for (int i = 0; i < 8; i++) H[i] = A[i][8] / A[i][i];
    H[8] = 1.0;
    return true;
}

bool invert3(const double m[9], double o[9]) {
    double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
    double A = e * i - f * h, B = -(d * i - f * g), Cc = d * h - e * g;
    double det = a * A + b * B + c * Cc;
    if (det == 0.0) return false;
    o[0] = A / det; o[1] = -(b * i - c * h) / det; o[2] = (b * f - c * e) / det;
    o[3] = B / det; o[4] = (a * i - c * g) / det; o[5] = -(a * f - c * d) / det;
    o[6] = Cc / det; o[7] = -(a * h - b * g) / det; o[8] = (a * e - b * d) / det;
    return true;
}

// The cells of a tag: tw x tw, 255 white, 0 black (the layout of
// apriltag_to_image, apriltag.c:271-297).
void tag_cells(const at_family *f, int64_t id, std::vector<uint8_t> &cells) {
    int tw = f->total_width;
    cells.assign((size_t)tw * tw, 0);
    int wbw = f->width_at_border + (f->reversed_border ? 0 : 2);
    int wbs = (tw - wbw) / 2;
    for (int i = 0; i < wbw - 1; i++) {
        cells[(size_t)wbs * tw + wbs + i] = 255;
        cells[(size_t)(wbs + i) * tw + tw - 1 - wbs] = 255;
        cells[(size_t)(tw - 1 - wbs) * tw + wbs + i + 1] = 255;
        cells[(size_t)(wbs + 1 + i) * tw + wbs] = 255;
    }
    int bs = (tw - f->width_at_border) / 2;
    uint64_t code = f->codes[id];
    for (int i = 0; i < f->nbits; i++)
        if (code & ((uint64_t)1 << (f->nbits - i - 1)))
            cells[(size_t)(f->bit_y[i] + bs) * tw + f->bit_x[i] + bs] = 255;
}

struct Canvas {
    int64_t w, h;
    std::vector<int32_t> px;  // gray, 0..255 (int while composing)
};

// Fill the quad (x,y image points of the unit square's image under H)
// with the pattern f(u, v) in [0,1)^2 -> gray, 4x4 supersampled.
template <class F>
void paint(Canvas &cv, const double H[9], double x0, double y0, double x1, double y1, F pattern) {
    double Hi[9];
    if (!invert3(H, Hi)) return;
    int64_t ix0 = (int64_t)(x0 < 0 ? 0 : x0), iy0 = (int64_t)(y0 < 0 ? 0 : y0);
    int64_t ix1 = (int64_t)(x1 + 1 > cv.w ? cv.w : x1 + 1), iy1 = (int64_t)(y1 + 1 > cv.h ? cv.h : y1 + 1);
    for (int64_t y = iy0; y < iy1; y++)
        for (int64_t x = ix0; x < ix1; x++) {
            int inside = 0, acc = 0;
            for (int sy = 0; sy < 4; sy++)
                for (int sx = 0; sx < 4; sx++) {
                    double px = (double)x + (sx + 0.5) / 4.0, py = (double)y + (sy + 0.5) / 4.0;
                    double z = Hi[6] * px + Hi[7] * py + Hi[8];
                    double u = (Hi[0] * px + Hi[1] * py + Hi[2]) / z;
                    double v = (Hi[3] * px + Hi[4] * py + Hi[5]) / z;
                    if (!(u >= 0.0 && u < 1.0 && v >= 0.0 && v < 1.0)) continue;
                    inside++;
                    acc += pattern(u, v);
                }
            if (inside == 0) continue;
            int32_t &p = cv.px[(size_t)(y * cv.w + x)];
            // coverage-weighted blend with what is below
            p = (int32_t)((acc + p * (16 - inside) + 8) / 16);
        }
}

// the image corners of the unit square under a pinhole view of a square
// of side `side` centred at (cx, cy), in-plane rotation rot, tilts t1, t2
void place_square(double cx, double cy, double side, Rot rot, Rot tx, Rot ty, double focal, double out[4][2]) {
    const double uv[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    for (int i = 0; i < 4; i++) {
        double X = (uv[i][0] - 0.5) * side, Y = (uv[i][1] - 0.5) * side, Z = 0.0;
        // tilt about x, then about y
        double Y1 = Y * tx.c - Z * tx.s, Z1 = Y * tx.s + Z * tx.c;
        double X2 = X * ty.c + Z1 * ty.s, Z2 = -X * ty.s + Z1 * ty.c;
        // in-plane rotation
        double X3 = X2 * rot.c - Y1 * rot.s, Y3 = X2 * rot.s + Y1 * rot.c;
        double zc = focal + Z2;
        out[i][0] = cx + focal * X3 / zc;
        out[i][1] = cy + focal * Y3 / zc;
    }
}

void box_blur_gauss(Canvas &cv, double sigma) {
    if (sigma <= 0.0) return;
    // weights 4096 / (1 + x^2/(2 sigma^2))^4: a rational bell close to the
    // Gaussian for |x| <= 3 sigma
    int r = (int)(3.0 * sigma + 0.5);
    if (r < 1) r = 1;
    std::vector<int64_t> k(2 * r + 1);
    int64_t ks = 0;
    for (int i = -r; i <= r; i++) {
        double q = 1.0 + (double)(i * i) / (2.0 * sigma * sigma) / 4.0;
        double v = 4096.0 / (q * q * q * q);
        k[i + r] = (int64_t)(v + 0.5);
        ks += k[i + r];
    }
    std::vector<int32_t> tmp(cv.px.size());
    for (int64_t y = 0; y < cv.h; y++)
        for (int64_t x = 0; x < cv.w; x++) {
            int64_t acc = 0;
            for (int i = -r; i <= r; i++) {
                int64_t xx = x + i;
                xx = xx < 0 ? 0 : xx >= cv.w ? cv.w - 1 : xx;
                acc += k[i + r] * cv.px[(size_t)(y * cv.w + xx)];
            }
            tmp[(size_t)(y * cv.w + x)] = (int32_t)((acc + ks / 2) / ks);
        }
    for (int64_t y = 0; y < cv.h; y++)
        for (int64_t x = 0; x < cv.w; x++) {
            int64_t acc = 0;
            for (int i = -r; i <= r; i++) {
                int64_t yy = y + i;
                yy = yy < 0 ? 0 : yy >= cv.h ? cv.h - 1 : yy;
                acc += k[i + r] * tmp[(size_t)(yy * cv.w + x)];
            }
            cv.px[(size_t)(y * cv.w + x)] = (int32_t)((acc + ks / 2) / ks);
        }
}

// value noise: a coarse random grid, bilinear, a few octaves (integer)
void value_noise(Canvas &cv, Rng &r, int amp) {
    int cell = 128;
    for (int oct = 0; oct < 3; oct++, cell /= 2) {
        int gw = (int)(cv.w / cell) + 2, gh = (int)(cv.h / cell) + 2;
        std::vector<int32_t> g((size_t)gw * gh);
        int a = amp >> oct;
        for (auto &v : g) v = (int32_t)r.range(-a, a);
        for (int64_t y = 0; y < cv.h; y++) {
            int64_t gy = y / cell, fy = y % cell;
            for (int64_t x = 0; x < cv.w; x++) {
                int64_t gx = x / cell, fx = x % cell;
                int64_t v00 = g[(size_t)(gy * gw + gx)], v01 = g[(size_t)(gy * gw + gx + 1)];
                int64_t v10 = g[(size_t)((gy + 1) * gw + gx)], v11 = g[(size_t)((gy + 1) * gw + gx + 1)];
                int64_t top = v00 * (cell - fx) + v01 * fx, bot = v10 * (cell - fx) + v11 * fx;
                cv.px[(size_t)(y * cv.w + x)] += (int32_t)((top * (cell - fy) + bot * fy) / ((int64_t)cell * cell) / 2);
            }
        }
    }
}

}  // namespace

extern "C" void atg_synth_default(atg_synth_spec *s, int config) {
    memset(s, 0, sizeof *s);
    s->family_a = 3;  // tag36h11
    s->family_b = -1;
    switch (config) {
    default:
    case 1:
        s->width = 640; s->height = 480; s->ntags_min = s->ntags_max = 4;
        s->side_min = 60; s->side_max = 120; s->max_tilt = 0.27;  // tan(15 deg): tilt <= 30 deg
        s->background = 0; s->bg_amp = 30;
        s->blur_min = s->blur_max = 0.8; s->noise_min = s->noise_max = 4.0;
        break;
    case 2:
        s->width = 1920; s->height = 1080; s->ntags_min = s->ntags_max = 20;
        s->side_min = 40; s->side_max = 300; s->max_tilt = 0.41;  // tan(22.5 deg): tilt <= 45 deg
        s->background = 1; s->bg_amp = 60; s->nbars = 6;
        s->blur_min = 0.0; s->blur_max = 1.5; s->noise_min = 2.0; s->noise_max = 8.0;
        break;
    case 3:
        s->width = 3840; s->height = 2160; s->ntags_min = s->ntags_max = 60;
        s->side_min = 24; s->side_max = 400; s->max_tilt = 0.41;
        s->family_b = 8;  // tagStandard52h13
        s->background = 1; s->bg_amp = 60; s->nclutter = 2000; s->ncheckers = 8;
        s->blur_min = s->blur_max = 0.0; s->noise_min = s->noise_max = 6.0;
        break;
    case 4:  // a config-1-like frame; the caller sets the size and tag count
        s->width = 1280; s->height = 720; s->ntags_min = 5; s->ntags_max = 30;
        s->side_min = 40; s->side_max = 160; s->max_tilt = 0.27;
        s->background = 0; s->bg_amp = 30;
        s->blur_min = s->blur_max = 0.8; s->noise_min = s->noise_max = 4.0;
        break;
    }
}

// The frame of `seed` with its tags moved `step` frames along their drift:
// each tag has its own velocity (up to 1.5 px a frame) and spin (up to
// about 0.6 degrees a frame) from a generator of its own, so step 0 is the
// plain frame; the noise of a step > 0 comes from its own generator.
static int synth_frame(const atg_synth_spec *spec, uint64_t seed, int64_t step, uint8_t *out, int64_t stride,
                       atg_synth_tag *truth, int32_t max_truth) {
    Rng r{seed * 0x2545f4914f6cdd1dull + 0x1234567ull};
    Canvas cv{spec->width, spec->height, std::vector<int32_t>((size_t)spec->width * spec->height)};
    int64_t W = spec->width, H = spec->height;
    // background
    int32_t base = (int32_t)r.range(90, 170);
    if (spec->background == 0) {
        // a linear gradient of +-bg_amp across the frame, random direction
        int64_t gx = r.range(-1000, 1000), gy = r.range(-1000, 1000);
        int64_t norm = (gx < 0 ? -gx : gx) * W / 2 + (gy < 0 ? -gy : gy) * H / 2 + 1;
        for (int64_t y = 0; y < H; y++)
            for (int64_t x = 0; x < W; x++)
                cv.px[(size_t)(y * W + x)] =
                    base + (int32_t)((gx * (2 * x - W) / 2 + gy * (2 * y - H) / 2) * spec->bg_amp / norm);
    } else {
        for (auto &p : cv.px) p = base;
        value_noise(cv, r, spec->bg_amp);
    }
    double focal = 2.0 * (double)(W > H ? W : H);
    // clutter: filled rectangles and quadrilaterals
    for (int i = 0; i < spec->nclutter; i++) {
        double side = (double)r.range(5, 80);
        double cx = r.uniform(0, (double)W), cy = r.uniform(0, (double)H);
        Rot rot = random_rotation(r);
        Rot t0 = rational_rot(r.uniform(-0.3, 0.3)), t1 = rational_rot(r.uniform(-0.3, 0.3));
        double dst[4][2];
        place_square(cx, cy, side * (0.5 + r.unit()), rot, t0, t1, focal, dst);
        const double src[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        double Hm[9];
        if (!solve_h(src, dst, Hm)) continue;
        int32_t g = (int32_t)r.range(0, 255);
        double x0 = 1e30, y0 = 1e30, x1 = -1e30, y1 = -1e30;
        for (auto &p : dst) { x0 = p[0] < x0 ? p[0] : x0; x1 = p[0] > x1 ? p[0] : x1; y0 = p[1] < y0 ? p[1] : y0; y1 = p[1] > y1 ? p[1] : y1; }
        paint(cv, Hm, x0, y0, x1, y1, [g](double, double) { return g; });
    }
    // checkerboards
    for (int i = 0; i < spec->ncheckers; i++) {
        int n = (int)r.range(4, 9);
        double side = (double)r.range(80, 300);
        double cx = r.uniform(side, (double)W - side), cy = r.uniform(side, (double)H - side);
        Rot rot = random_rotation(r);
        Rot t0 = rational_rot(r.uniform(-0.3, 0.3)), t1 = rational_rot(r.uniform(-0.3, 0.3));
        double dst[4][2];
        place_square(cx, cy, side, rot, t0, t1, focal, dst);
        const double src[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        double Hm[9];
        if (!solve_h(src, dst, Hm)) continue;
        double x0 = 1e30, y0 = 1e30, x1 = -1e30, y1 = -1e30;
        for (auto &p : dst) { x0 = p[0] < x0 ? p[0] : x0; x1 = p[0] > x1 ? p[0] : x1; y0 = p[1] < y0 ? p[1] : y0; y1 = p[1] > y1 ? p[1] : y1; }
        paint(cv, Hm, x0, y0, x1, y1, [n](double u, double v) {
            return (((int)(u * n) + (int)(v * n)) & 1) ? 230 : 25;
        });
    }
    // tags
    int ntags = (int)r.range(spec->ntags_min, spec->ntags_max);
    std::vector<double> placed;  // cx, cy, radius
    int nt = 0;
    std::vector<uint8_t> cells;
    for (int i = 0; i < ntags; i++) {
        int fidx = spec->family_a;
        if (spec->family_b >= 0 && (r.next() & 1)) fidx = spec->family_b;
        const at_family *f = atg_builtin_family(fidx);
        if (!f) return -1;
        double side = (double)r.range(spec->side_min, spec->side_max);
        double cx = 0, cy = 0;
        bool ok = false;
        for (int tries = 0; tries < 200 && !ok; tries++) {
            double m = side * 0.75;
            if ((double)W <= 2 * m || (double)H <= 2 * m) break;
            cx = r.uniform(m, (double)W - m);
            cy = r.uniform(m, (double)H - m);
            ok = true;
            for (size_t k = 0; k < placed.size(); k += 3) {
                double dx = cx - placed[k], dy = cy - placed[k + 1], rr = placed[k + 2] + side * 0.75;
                if (dx * dx + dy * dy < rr * rr) { ok = false; break; }
            }
        }
        if (!ok) continue;
        placed.push_back(cx); placed.push_back(cy); placed.push_back(side * 0.75);
        int64_t id = r.range(0, f->ncodes - 1);
        tag_cells(f, id, cells);
        Rot rot = random_rotation(r);
        Rot t0 = rational_rot(r.uniform(-spec->max_tilt, spec->max_tilt));
        Rot t1 = rational_rot(r.uniform(-spec->max_tilt, spec->max_tilt));
        if (step != 0) {
            Rng d{(seed + 0x9e3779b97f4a7c15ull * (uint64_t)(i + 1)) ^ 0xd1b54a32d192ed03ull};
            double vx = d.uniform(-1.5, 1.5), vy = d.uniform(-1.5, 1.5), w = d.uniform(-0.005, 0.005);
            cx += vx * (double)step;
            cy += vy * (double)step;
            Rot dr = rational_rot(w * (double)step);
            rot = Rot{rot.c * dr.c - rot.s * dr.s, rot.s * dr.c + rot.c * dr.s};
        }
        double dst[4][2];
        place_square(cx, cy, side, rot, t0, t1, focal, dst);
        const double src[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        double Hm[9];
        if (!solve_h(src, dst, Hm)) continue;
        double x0 = 1e30, y0 = 1e30, x1 = -1e30, y1 = -1e30;
        for (auto &p : dst) { x0 = p[0] < x0 ? p[0] : x0; x1 = p[0] > x1 ? p[0] : x1; y0 = p[1] < y0 ? p[1] : y0; y1 = p[1] > y1 ? p[1] : y1; }
        int tw = f->total_width;
        int32_t lo = (int32_t)r.range(10, 50), hi = (int32_t)r.range(200, 250);
        paint(cv, Hm, x0, y0, x1, y1, [&](double u, double v) {
            int cxi = (int)(u * tw), cyi = (int)(v * tw);
            return cells[(size_t)cyi * tw + cxi] ? hi : lo;
        });
        if (nt < max_truth && truth) {
            truth[nt].family = fidx;
            truth[nt].id = (int32_t)id;
            for (int k = 0; k < 4; k++) { truth[nt].corners[2 * k] = dst[k][0]; truth[nt].corners[2 * k + 1] = dst[k][1]; }
        }
        nt++;
    }
    // occluding dark bars over some tags
    for (int i = 0; i < spec->nbars && !placed.empty(); i++) {
        size_t k = (size_t)r.range(0, (int64_t)placed.size() / 3 - 1) * 3;
        double len = placed[k + 2] * 1.2, thick = placed[k + 2] * 0.15 + 2.0;
        Rot rot = random_rotation(r);
        double cx = placed[k] + r.uniform(-0.5, 0.5) * placed[k + 2], cy = placed[k + 1] + r.uniform(-0.5, 0.5) * placed[k + 2];
        double dst[4][2];
        const double uv[4][2] = {{-0.5, -0.5}, {0.5, -0.5}, {0.5, 0.5}, {-0.5, 0.5}};
        for (int q = 0; q < 4; q++) {
            double X = uv[q][0] * len, Y = uv[q][1] * thick;
            dst[q][0] = cx + X * rot.c - Y * rot.s;
            dst[q][1] = cy + X * rot.s + Y * rot.c;
        }
        const double src[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        double Hm[9];
        if (!solve_h(src, dst, Hm)) continue;
        double x0 = 1e30, y0 = 1e30, x1 = -1e30, y1 = -1e30;
        for (auto &p : dst) { x0 = p[0] < x0 ? p[0] : x0; x1 = p[0] > x1 ? p[0] : x1; y0 = p[1] < y0 ? p[1] : y0; y1 = p[1] > y1 ? p[1] : y1; }
        int32_t g = (int32_t)r.range(5, 40);
        paint(cv, Hm, x0, y0, x1, y1, [g](double, double) { return g; });
    }
    double blur = r.uniform(spec->blur_min, spec->blur_max);
    if (spec->blur_max > 0.0) box_blur_gauss(cv, blur);
    double noise = r.uniform(spec->noise_min, spec->noise_max);
    if (step != 0) r = Rng{(seed ^ 0x6a09e667f3bcc909ull) * 0x2545f4914f6cdd1dull + (uint64_t)step * 0x9e3779b97f4a7c15ull};
    int64_t nq = (int64_t)(noise * 256.0 + 0.5);  // sigma in 1/256 gray levels
    for (int64_t y = 0; y < H; y++)
        for (int64_t x = 0; x < W; x++) {
            int64_t v = (int64_t)cv.px[(size_t)(y * W + x)] * 65536 * 256 + r.gauss16() * nq;
            v = v >= 0 ? (v + 65536 * 128) / (65536 * 256) : -((-v + 65536 * 128) / (65536 * 256));
            out[y * stride + x] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
        }
    return nt;
}

extern "C" int atg_synth_frame(const atg_synth_spec *spec, uint64_t seed, uint8_t *out, int64_t stride,
                               atg_synth_tag *truth, int32_t max_truth) {
    return synth_frame(spec, seed, 0, out, stride, truth, max_truth);
}

extern "C" int atg_synth_frame_drift(const atg_synth_spec *spec, uint64_t seed, int64_t step, uint8_t *out,
                                     int64_t stride, atg_synth_tag *truth, int32_t max_truth) {
    return synth_frame(spec, seed, step, out, stride, truth, max_truth);
}
