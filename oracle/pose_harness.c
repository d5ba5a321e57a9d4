/* Test infrastructure (the checker), not the product: a flat-array entry
 * into the reference's own pose estimation, compiled from its sources
 * (apriltag_pose.c, common/homography.c, common/matd.c, common/svd22.c)
 * into oracle/_ref/libposeref.so by `make -C oracle ref`.
 *
 * ref_pose() fills the layout of atg_pose_out (include/atg.h) from
 *   homography_to_pose(H, -fx, fy, cx, cy)   (common/homography.c:275)
 *   estimate_pose_for_tag_homography        (apriltag_pose.c:455)
 *   estimate_tag_pose_orthogonal_iteration  (apriltag_pose.c:490)
 *   estimate_tag_pose's choice              (apriltag_pose.c:528-547)
 * and ref_homography_to_pose() is homography_to_pose with the caller's
 * arguments. */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "apriltag_pose.h"
#include "common/homography.h"

typedef struct {
    double hpose[12];
    double R0[9], t0[3];
    double R1[9], t1[3], err1;
    double R2[9], t2[3], err2;
    int32_t has2, best;
} ref_pose_out;

static void put(const matd_t *m, double *dst, int n) {
    for (int i = 0; i < n; i++) dst[i] = m->data[i];
}

void ref_homography_to_pose(const double *H, double fx, double fy, double cx, double cy, double *out16) {
    matd_t *h = matd_create_data(3, 3, H);
    matd_t *m = homography_to_pose(h, fx, fy, cx, cy);
    put(m, out16, 16);
    matd_destroy(m);
    matd_destroy(h);
}

void ref_pose(const double *H, const double *corners, double tagsize, double fx, double fy, double cx, double cy,
              int iters, ref_pose_out *o) {
    memset(o, 0, sizeof *o);
    apriltag_detection_t det;
    memset(&det, 0, sizeof det);
    det.H = matd_create_data(3, 3, H);
    for (int i = 0; i < 4; i++) {
        det.p[i][0] = corners[2 * i];
        det.p[i][1] = corners[2 * i + 1];
    }
    apriltag_detection_info_t info = {&det, tagsize, fx, fy, cx, cy};

    double m16[16];
    ref_homography_to_pose(H, -fx, fy, cx, cy, m16);
    memcpy(o->hpose, m16, sizeof o->hpose);

    apriltag_pose_t p0;
    estimate_pose_for_tag_homography(&info, &p0);
    put(p0.R, o->R0, 9);
    put(p0.t, o->t0, 3);
    matd_destroy(p0.R);
    matd_destroy(p0.t);

    apriltag_pose_t p1, p2;
    double e1, e2;
    estimate_tag_pose_orthogonal_iteration(&info, &e1, &p1, &e2, &p2, iters);
    put(p1.R, o->R1, 9);
    put(p1.t, o->t1, 3);
    o->err1 = e1;
    o->err2 = e2;
    o->has2 = p2.R != NULL;
    if (p2.R) {
        put(p2.R, o->R2, 9);
        put(p2.t, o->t2, 3);
        matd_destroy(p2.R);
        matd_destroy(p2.t);
    }
    matd_destroy(p1.R);
    matd_destroy(p1.t);

    o->best = e1 <= e2 ? 1 : 2;  /* estimate_tag_pose's rule on this run */
    matd_destroy(det.H);
}

/* estimate_tag_pose itself (50 steps): its error, R and t. */
double ref_estimate_tag_pose(const double *H, const double *corners, double tagsize, double fx, double fy, double cx,
                             double cy, double *R, double *t) {
    apriltag_detection_t det;
    memset(&det, 0, sizeof det);
    det.H = matd_create_data(3, 3, H);
    for (int i = 0; i < 4; i++) {
        det.p[i][0] = corners[2 * i];
        det.p[i][1] = corners[2 * i + 1];
    }
    apriltag_detection_info_t info = {&det, tagsize, fx, fy, cx, cy};
    apriltag_pose_t pb;
    double e = estimate_tag_pose(&info, &pb);
    put(pb.R, R, 9);
    put(pb.t, t, 3);
    matd_destroy(pb.R);
    matd_destroy(pb.t);
    matd_destroy(det.H);
    return e;
}

/* estimate_tag_pose over n tags in the atg_pose_in layout (H[9], p[8],
 * tagsize, fx, fy, cx, cy: 22 doubles a tag), each error into err[i]: the
 * reference CPU side of the pose timing. */
void ref_estimate_many(const double *tags, int64_t n, double *err) {
    for (int64_t i = 0; i < n; i++) {
        const double *t = tags + 22 * i;
        double R[9], tt[3];
        err[i] = ref_estimate_tag_pose(t, t + 9, t[17], t[18], t[19], t[20], t[21], R, tt);
    }
}
