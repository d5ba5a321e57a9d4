// Tag pose from detections (the reference's apriltag_pose.c and
// homography_to_pose, common/homography.c:275-351): one device thread a
// tag, all in f64 registers.
//
//   hpose     camera-frame [R | T] from the homography: the columns from
//             H and the intrinsics, scaled to unit length (their f32
//             square roots as the reference takes them), the third
//             column their cross product, then the polar factor.
//   oi        orthogonal iteration (Lu, Hager, Mjolsness 2000;
//             apriltag_pose.c:43-137): translation by the closed form,
//             rotation by the polar factor of the cross-covariance.
//   ambiguity the second local minimum (Schweighofer, Pinz 2006;
//             apriltag_pose.c:259-452): the error as a quartic in
//             tan(beta/2), its stationary points by the reference's
//             Newton-bisection over the derivative's roots
//             (apriltag_pose.c:143-257).
//
// The reference's polar factors come from a general SVD (U V') and its
// 3x3 inverses from an LU solve; here the polar factor of the full-rank
// homography rotation is Newton's polar iteration, that of the rank-2
// cross-covariance (its third column is zero: the tag is planar) the
// closed form A (A'A)^-1/2, and inverses the adjugate over the
// determinant. These are the same matrices to rounding.
#include <hip/hip_runtime.h>

#include <cmath>

#include "atg.h"

#pragma clang fp contract(off)

namespace atg {

namespace {

struct V3 {
    double x[3];
};
struct M3 {
    double a[9];  // row-major
    __device__ double &operator()(int r, int c) { return a[3 * r + c]; }
    __device__ double operator()(int r, int c) const { return a[3 * r + c]; }
};

__device__ __forceinline__ M3 mzero() {
    M3 m;
    for (int i = 0; i < 9; i++) m.a[i] = 0.0;
    return m;
}
__device__ __forceinline__ M3 meye() {
    M3 m = mzero();
    m.a[0] = m.a[4] = m.a[8] = 1.0;
    return m;
}
__device__ __forceinline__ M3 mmul(const M3 &A, const M3 &B) {
    M3 C;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += A(r, k) * B(k, c);
            C(r, c) = s;
        }
    return C;
}
// A' B
__device__ __forceinline__ M3 mtmul(const M3 &A, const M3 &B) {
    M3 C;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += A(k, r) * B(k, c);
            C(r, c) = s;
        }
    return C;
}
// A B'
__device__ __forceinline__ M3 mmult(const M3 &A, const M3 &B) {
    M3 C;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += A(r, k) * B(c, k);
            C(r, c) = s;
        }
    return C;
}
__device__ __forceinline__ V3 mv(const M3 &A, const V3 &v) {
    V3 o;
    for (int r = 0; r < 3; r++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++) s += A(r, k) * v.x[k];
        o.x[r] = s;
    }
    return o;
}
// A' v
__device__ __forceinline__ V3 mtv(const M3 &A, const V3 &v) {
    V3 o;
    for (int r = 0; r < 3; r++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++) s += A(k, r) * v.x[k];
        o.x[r] = s;
    }
    return o;
}
__device__ __forceinline__ M3 madd(const M3 &A, const M3 &B) {
    M3 C;
    for (int i = 0; i < 9; i++) C.a[i] = A.a[i] + B.a[i];
    return C;
}
__device__ __forceinline__ M3 msub(const M3 &A, const M3 &B) {
    M3 C;
    for (int i = 0; i < 9; i++) C.a[i] = A.a[i] - B.a[i];
    return C;
}
__device__ __forceinline__ M3 mscale(const M3 &A, double s) {
    M3 C;
    for (int i = 0; i < 9; i++) C.a[i] = A.a[i] * s;
    return C;
}
__device__ __forceinline__ V3 vadd(const V3 &a, const V3 &b) {
    return V3{{a.x[0] + b.x[0], a.x[1] + b.x[1], a.x[2] + b.x[2]}};
}
__device__ __forceinline__ V3 vsub(const V3 &a, const V3 &b) {
    return V3{{a.x[0] - b.x[0], a.x[1] - b.x[1], a.x[2] - b.x[2]}};
}
__device__ __forceinline__ V3 vscale(const V3 &a, double s) { return V3{{a.x[0] * s, a.x[1] * s, a.x[2] * s}}; }
__device__ __forceinline__ double vdot(const V3 &a, const V3 &b) {
    double s = 0.0;
    for (int k = 0; k < 3; k++) s += a.x[k] * b.x[k];
    return s;
}
__device__ __forceinline__ V3 vcross(const V3 &a, const V3 &b) {
    return V3{{a.x[1] * b.x[2] - a.x[2] * b.x[1], a.x[2] * b.x[0] - a.x[0] * b.x[2], a.x[0] * b.x[1] - a.x[1] * b.x[0]}};
}
__device__ __forceinline__ V3 vunit(const V3 &a) {
    const double m = sqrt(vdot(a, a));
    return V3{{a.x[0] / m, a.x[1] / m, a.x[2] / m}};
}
__device__ __forceinline__ double mdet(const M3 &A) {
    return A(0, 0) * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) - A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) +
           A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
}
// The inverse by the adjugate (a singular matrix gives infinities, as the
// reference's NULL inverse would end its computation).
__device__ __forceinline__ M3 minv(const M3 &A) {
    M3 C;
    C(0, 0) = A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1);
    C(0, 1) = A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2);
    C(0, 2) = A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1);
    C(1, 0) = A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2);
    C(1, 1) = A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0);
    C(1, 2) = A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2);
    C(2, 0) = A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0);
    C(2, 1) = A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1);
    C(2, 2) = A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0);
    const double det = A(0, 0) * C(0, 0) + A(0, 1) * C(1, 0) + A(0, 2) * C(2, 0);
    return mscale(C, 1.0 / det);
}
__device__ __forceinline__ M3 mtrans(const M3 &A) {
    M3 C;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) C(r, c) = A(c, r);
    return C;
}

// The projection onto the ray v: v v' / (v' v) (apriltag_pose.c:12-19).
__device__ __forceinline__ M3 ray_projector(const V3 &v) {
    const double inv = 1.0 / vdot(v, v);
    M3 F;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) F(r, c) = (v.x[r] * v.x[c]) * inv;
    return F;
}

// The polar factor of a nonsingular A (the U V' of its SVD): Newton's
// iteration X <- (X + X^-T) / 2, quadratically convergent from any
// nonsingular start; stopped when a step changes no entry by more than
// 1e-15 (or after 40 steps).
__device__ M3 polar_newton(M3 X) {
    for (int it = 0; it < 40; it++) {
        const M3 Y = mscale(madd(X, mtrans(minv(X))), 0.5);
        double d = 0.0;
        for (int i = 0; i < 9; i++) d = fmax(d, fabs(Y.a[i] - X.a[i]));
        X = Y;
        if (d <= 1e-15) break;
    }
    return X;
}

// The proper rotation of a cross-covariance whose third column is zero
// (apriltag_pose.c:98-112: U V' of its SVD, the third column negated when
// the determinant is negative): the first two columns are the polar
// factor Q = A (A'A)^-1/2 of the 3x2 block A, the third q0 x q1.
// (A'A)^-1/2 of the 2x2 B = [a b; b c] is [c+s, -b; -b, a+s] / (s t),
// s = sqrt(det B), t = sqrt(a + c + 2 s).
__device__ M3 rotation_of_rank2(const M3 &M) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int k = 0; k < 3; k++) {
        a += M(k, 0) * M(k, 0);
        b += M(k, 0) * M(k, 1);
        c += M(k, 1) * M(k, 1);
    }
    const double s = sqrt(a * c - b * b), t = sqrt(a + c + 2.0 * s), f = 1.0 / (s * t);
    const double i00 = (c + s) * f, i01 = -b * f, i11 = (a + s) * f;
    M3 R;
    V3 q0, q1;
    for (int k = 0; k < 3; k++) {
        q0.x[k] = M(k, 0) * i00 + M(k, 1) * i01;
        q1.x[k] = M(k, 0) * i01 + M(k, 1) * i11;
    }
    const V3 q2 = vcross(q0, q1);
    for (int k = 0; k < 3; k++) {
        R(k, 0) = q0.x[k];
        R(k, 1) = q1.x[k];
        R(k, 2) = q2.x[k];
    }
    return R;
}

// homography_to_pose (common/homography.c:275-351): the camera looks down
// -z; the column lengths are the f32 square roots of double sums.
__device__ void homography_pose(const double *H, double fx, double fy, double cx, double cy, double out[12]) {
    double r20 = H[6], r21 = H[7], tz = H[8];
    double r00 = (H[0] - cx * r20) / fx, r01 = (H[1] - cx * r21) / fx, tx = (H[2] - cx * tz) / fx;
    double r10 = (H[3] - cy * r20) / fy, r11 = (H[4] - cy * r21) / fy, ty = (H[5] - cy * tz) / fy;
    const double len1 = (double)(float)__dsqrt_rn((double)(float)(r00 * r00 + r10 * r10 + r20 * r20));
    const double len2 = (double)(float)__dsqrt_rn((double)(float)(r01 * r01 + r11 * r11 + r21 * r21));
    double s = 1.0 / (double)(float)__dsqrt_rn((double)(float)(len1 * len2));
    if (tz > 0) s = -s;
    r20 *= s;
    r21 *= s;
    tz *= s;
    r00 *= s;
    r01 *= s;
    tx *= s;
    r10 *= s;
    r11 *= s;
    ty *= s;
    M3 R;
    R(0, 0) = r00;
    R(0, 1) = r01;
    R(0, 2) = r10 * r21 - r20 * r11;
    R(1, 0) = r10;
    R(1, 1) = r11;
    R(1, 2) = r20 * r01 - r00 * r21;
    R(2, 0) = r20;
    R(2, 1) = r21;
    R(2, 2) = r00 * r11 - r10 * r01;
    R = polar_newton(R);
    const double T[3] = {tx, ty, tz};
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++) out[4 * r + c] = R(r, c);
        out[4 * r + 3] = T[r];
    }
}

// Orthogonal iteration (apriltag_pose.c:43-137) over the four corners:
// v the image rays, p the object points; R in and out, t out. Returns the
// object-space error of the last step (HUGE_VAL after no step).
__device__ double orthogonal_iteration(const V3 (&v)[4], const V3 (&p)[4], V3 &t, M3 &R, int steps) {
    constexpr int n = 4;
    V3 pm{{0.0, 0.0, 0.0}};
    for (int i = 0; i < n; i++) pm = vadd(pm, p[i]);
    pm = vscale(pm, 1.0 / n);
    V3 pr[n];
    M3 F[n];
    M3 Fm = mzero();
    for (int i = 0; i < n; i++) {
        pr[i] = vsub(p[i], pm);
        F[i] = ray_projector(v[i]);
        Fm = madd(Fm, F[i]);
    }
    Fm = mscale(Fm, 1.0 / n);
    const M3 I = meye();
    const M3 M1inv = minv(msub(I, Fm));
    double err = HUGE_VAL;
    for (int step = 0; step < steps; step++) {
        // translation: M1^-1 mean((F_j - I) R p_j)
        V3 m2{{0.0, 0.0, 0.0}};
        for (int j = 0; j < n; j++) m2 = vadd(m2, mv(mmul(msub(F[j], I), R), p[j]));
        m2 = vscale(m2, 1.0 / n);
        t = mv(M1inv, m2);
        // rotation: the polar factor of sum (q_j - mean q) (p_j - mean p)'
        V3 q[n];
        V3 qm{{0.0, 0.0, 0.0}};
        for (int j = 0; j < n; j++) {
            q[j] = mv(F[j], vadd(mv(R, p[j]), t));
            qm = vadd(qm, q[j]);
        }
        qm = vscale(qm, 1.0 / n);
        M3 M3c = mzero();
        for (int j = 0; j < n; j++) {
            const V3 d = vsub(q[j], qm);
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) M3c(r, c) += d.x[r] * pr[j].x[c];
        }
        R = rotation_of_rank2(M3c);
        err = 0.0;
        for (int j = 0; j < n; j++) {
            const V3 e = mv(msub(I, F[j]), vadd(mv(R, p[j]), t));
            err += vdot(e, e);
        }
    }
    return err;
}

// p[0] + p[1] x + ... + p[deg] x^deg, term by term as the reference's
// polyval (apriltag_pose.c:143-149).
__device__ __forceinline__ double poly_at(const double *p, int deg, double x) {
    double r = 0.0;
    for (int i = 0; i <= deg; i++) r += p[i] * pow(x, (double)i);
    return r;
}

// The reference's solve_poly_approx (apriltag_pose.c:160-257) for degree
// 4, its recursion unrolled bottom-up: the roots of each derivative bound
// the intervals searched for the roots of the one above; roots beyond
// +-1000 are ignored; each sign change is refined by Newton steps kept
// inside the bracket (bisection otherwise), 100 steps at most.
__device__ int quartic_roots(const double (&p4)[5], double roots[4]) {
    constexpr double kMax = 1000.0;
    double P[5][5];  // P[d]: the derivative of degree d, P[4] = p4
    for (int i = 0; i < 5; i++) P[4][i] = p4[i];
    for (int d = 4; d > 1; d--)
        for (int i = 0; i < d; i++) P[d - 1][i] = (i + 1) * P[d][i + 1];
    double cur[4];
    int ncur = 0;
    // degree 1
    if (!(fabs(P[1][0]) > kMax * fabs(P[1][1]))) {
        cur[0] = -P[1][0] / P[1][1];
        ncur = 1;
    }
    for (int d = 2; d <= 4; d++) {
        const double *f = P[d], *df = P[d - 1];
        double next[4];
        int nn = 0;
        for (int i = 0; i <= ncur; i++) {
            const double lo = i == 0 ? -kMax : cur[i - 1];
            const double hi = i == ncur ? kMax : cur[i];
            const double flo = poly_at(f, d, lo), fhi = poly_at(f, d, hi);
            if (flo * fhi < 0) {
                double lower, upper;
                if (poly_at(f, d, lo) < poly_at(f, d, hi)) {
                    lower = lo;
                    upper = hi;
                } else {
                    lower = hi;
                    upper = lo;
                }
                double x = 0.5 * (lower + upper);
                double step_old = upper - lower, step = step_old;
                double fx = poly_at(f, d, x), dfx = poly_at(df, d - 1, x);
                for (int k = 0; k < 100; k++) {
                    if (((fx + dfx * (upper - x)) * (fx + dfx * (lower - x)) > 0) || (fabs(2 * fx) > fabs(step_old * dfx))) {
                        step_old = step;
                        step = 0.5 * (upper - lower);
                        x = lower + step;
                    } else {
                        step_old = step;
                        step = -fx / dfx;
                        x += step;
                    }
                    if (x == upper || x == lower) break;
                    fx = poly_at(f, d, x);
                    dfx = poly_at(df, d - 1, x);
                    if (fx > 0)
                        upper = x;
                    else
                        lower = x;
                }
                next[nn++] = x;
            } else if (poly_at(f, d, hi) == 0) {
                next[nn++] = hi;  // a double or triple root
            }
        }
        for (int i = 0; i < nn; i++) cur[i] = next[i];
        ncur = nn;
    }
    for (int i = 0; i < ncur; i++) roots[i] = cur[i];
    return ncur;
}

// The second local minimum of the object-space error (apriltag_pose.c:
// 259-452), given the first (t, R). Returns false when there is none (or
// more than one new minimum, which the reference also discards).
__device__ bool second_minimum(const V3 (&v)[4], const V3 (&p)[4], const V3 &t, const M3 &R, M3 &out) {
    constexpr int n = 4;
    const M3 I = meye();
    // 1. the frame whose third axis is the direction of t
    const V3 rt3 = vunit(t);
    const V3 ex{{1.0, 0.0, 0.0}};
    const V3 rt1 = vunit(vsub(ex, vscale(rt3, vdot(ex, rt3))));
    const V3 rt2 = vcross(rt3, rt1);
    M3 Rt;
    for (int k = 0; k < 3; k++) {
        Rt(0, k) = rt1.x[k];
        Rt(1, k) = rt2.x[k];
        Rt(2, k) = rt3.x[k];
    }
    // 2. the rotation about z taking R's third row onto the x-z plane
    const M3 R1p = mmul(Rt, R);
    double r31 = R1p(2, 0), r32 = R1p(2, 1);
    double hyp = sqrt(r31 * r31 + r32 * r32);
    if (hyp < 1e-100) {
        r31 = 1.0;
        r32 = 0.0;
        hyp = 1.0;
    }
    M3 Rz = meye();
    Rz(0, 0) = r31 / hyp;
    Rz(0, 1) = -r32 / hyp;
    Rz(1, 0) = r32 / hyp;
    Rz(1, 1) = r31 / hyp;
    // 3. the error as a function of the remaining angle beta
    const M3 Rtr = mmul(R1p, Rz);
    const double sg = -Rtr(0, 1), cg = Rtr(1, 1);
    M3 Rg = meye();
    Rg(0, 0) = cg;
    Rg(0, 1) = -sg;
    Rg(1, 0) = sg;
    Rg(1, 1) = cg;
    const double beta0 = atan2(-Rtr(2, 0), Rtr(2, 2));
    V3 pt[n];
    M3 Ft[n];
    M3 Fm = mzero();
    for (int i = 0; i < n; i++) {
        pt[i] = mtv(Rz, p[i]);
        Ft[i] = ray_projector(mv(Rt, v[i]));
        Fm = madd(Fm, Ft[i]);
    }
    Fm = mscale(Fm, 1.0 / n);
    const M3 G = mscale(minv(msub(I, Fm)), 1.0 / n);
    M3 B1 = mzero(), B2 = mzero();  // the beta terms: tan(beta/2) and its square
    B1(0, 2) = 2.0;
    B1(2, 0) = -2.0;
    B2(0, 0) = -1.0;
    B2(1, 1) = 1.0;
    B2(2, 2) = -1.0;
    V3 b0{{0, 0, 0}}, b1{{0, 0, 0}}, b2{{0, 0, 0}};
    for (int i = 0; i < n; i++) {
        const M3 FI = mmul(msub(Ft[i], I), Rg);
        b0 = vadd(b0, mv(FI, pt[i]));
        b1 = vadd(b1, mv(mmul(FI, B1), pt[i]));
        b2 = vadd(b2, mv(mmul(FI, B2), pt[i]));
    }
    b0 = mv(G, b0);
    b1 = mv(G, b1);
    b2 = mv(G, b2);
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
    for (int i = 0; i < n; i++) {
        const M3 IF = msub(I, Ft[i]);
        const V3 c0 = mv(IF, vadd(mv(Rg, pt[i]), b0));
        const V3 c1 = mv(IF, vadd(mv(mmul(Rg, B1), pt[i]), b1));
        const V3 c2 = mv(IF, vadd(mv(mmul(Rg, B2), pt[i]), b2));
        a0 += vdot(c0, c0);
        a1 += 2 * vdot(c0, c1);
        a2 += vdot(c1, c1) + 2 * vdot(c0, c2);
        a3 += 2 * vdot(c1, c2);
        a4 += vdot(c2, c2);
    }
    // 4. the stationary points of the quartic error in tan(beta/2)
    const double poly[5] = {a1, 2 * a2 - 4 * a0, 3 * a3 - 3 * a1, 4 * a4 - 2 * a2, -a3};
    double roots[4];
    const int nr = quartic_roots(poly, roots);
    double minimum = 0.0;
    int nmin = 0;
    for (int i = 0; i < nr; i++) {
        const double t1 = roots[i], t2 = t1 * t1, t3 = t1 * t2, t4 = t1 * t3, t5 = t1 * t4;
        const double curv = a2 - 2 * a0 + (3 * a3 - 6 * a1) * t1 + (6 * a4 - 8 * a2 + 10 * a0) * t2 +
                            (-8 * a3 + 6 * a1) * t3 + (-6 * a4 + 3 * a2) * t4 + a3 * t5;
        if (curv >= 0 && fabs(2 * atan(t1) - beta0) > 0.1) {
            minimum = t1;
            nmin++;
        }
    }
    if (nmin != 1) return false;
    // 5. the rotation of that minimum
    const double tb = minimum;
    const M3 Rb = mscale(madd(mscale(madd(mscale(B2, tb), B1), tb), I), 1 / (1 + tb * tb));
    out = mmult(mmul(mtmul(Rt, Rg), Rb), Rz);
    return true;
}

__device__ __forceinline__ void put_rt(const M3 &R, const V3 &t, double *Ro, double *to) {
    for (int i = 0; i < 9; i++) Ro[i] = R.a[i];
    for (int i = 0; i < 3; i++) to[i] = t.x[i];
}

__global__ __launch_bounds__(64) void k_pose(const atg_pose_in *in, int64_t n, int iters, int flags,
                                             atg_pose_out *out) {
    const int64_t g = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (g >= n) return;
    const atg_pose_in &P = in[g];
    atg_pose_out o;
    // estimate_pose_for_tag_homography (apriltag_pose.c:455-488): the
    // translation scaled by half the tag size, then y and z negated
    homography_pose(P.H, -P.fx, P.fy, P.cx, P.cy, o.hpose);
    const double half = P.tagsize / 2.0;
    M3 R0;
    V3 t0;
    for (int r = 0; r < 3; r++) {
        const double sg = r == 0 ? 1.0 : -1.0;
        for (int c = 0; c < 3; c++) R0(r, c) = sg * o.hpose[4 * r + c];
        t0.x[r] = sg * (o.hpose[4 * r + 3] * half);
    }
    put_rt(R0, t0, o.R0, o.t0);
    M3 R1 = R0, R2 = mzero();
    V3 t1 = t0, t2{{0.0, 0.0, 0.0}};
    o.err1 = o.err2 = HUGE_VAL;
    o.has2 = 0;
    if (!(flags & ATG_POSE_HOMOGRAPHY_ONLY)) {
        // estimate_tag_pose_orthogonal_iteration (apriltag_pose.c:490-526)
        const V3 p[4] = {{{-half, half, 0.0}}, {{half, half, 0.0}}, {{half, -half, 0.0}}, {{-half, -half, 0.0}}};
        V3 v[4];
        for (int i = 0; i < 4; i++) v[i] = V3{{(P.p[i][0] - P.cx) / P.fx, (P.p[i][1] - P.cy) / P.fy, 1.0}};
        o.err1 = orthogonal_iteration(v, p, t1, R1, iters);
        if (second_minimum(v, p, t1, R1, R2)) {
            o.has2 = 1;
            o.err2 = orthogonal_iteration(v, p, t2, R2, iters);
        }
    }
    put_rt(R1, t1, o.R1, o.t1);
    put_rt(R2, t2, o.R2, o.t2);
    o.best = o.err1 <= o.err2 ? 1 : 2;
    out[g] = o;
}

}  // namespace

hipError_t launch_pose(hipStream_t st, const atg_pose_in *in, int64_t n, int iters, int flags, atg_pose_out *out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pose, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, in, n, iters, flags, out);
    return hipGetLastError();
}

}  // namespace atg
