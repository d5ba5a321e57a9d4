// The upstream pose API (apriltag_pose.h and homography_to_pose of the
// reference) over atg_estimate_poses: one tag a call, on device 0, through
// a context created at the first call and kept for the process; calls take
// turns on it.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "apriltag_pose.h"
#include "atg.h"
#include "common/homography.h"

namespace {

std::mutex g_pose_mu;
atg_ctx *g_pose_ctx = nullptr;

// The reference cannot fail here (its pose is host arithmetic); a device
// that cannot run the kernel is reported and ends the process, as the
// reference ends it on an allocation failure.
atg_pose_out pose_one(const atg_pose_in &in, int iters, int flags) {
    std::lock_guard<std::mutex> lk(g_pose_mu);
    if (!g_pose_ctx && atg_ctx_create(0, &g_pose_ctx) != ATG_OK) {
        std::fprintf(stderr, "apriltag pose: no HIP device 0\n");
        std::abort();
    }
    atg_pose_out o;
    if (atg_estimate_poses(g_pose_ctx, &in, 1, iters, flags, &o) != ATG_OK) {
        std::fprintf(stderr, "apriltag pose: %s\n", atg_last_error(g_pose_ctx));
        std::abort();
    }
    return o;
}

atg_pose_in tag_of(const apriltag_detection_info_t *info) {
    atg_pose_in in;
    std::memcpy(in.H, info->det->H->data, sizeof in.H);
    std::memcpy(in.p, info->det->p, sizeof in.p);
    in.tagsize = info->tagsize;
    in.fx = info->fx;
    in.fy = info->fy;
    in.cx = info->cx;
    in.cy = info->cy;
    return in;
}

void set_pose(apriltag_pose_t *pose, const double *R, const double *t) {
    pose->R = matd_create_data(3, 3, R);
    pose->t = matd_create_data(3, 1, t);
}

}  // namespace

extern "C" {

matd_t *homography_to_pose(const matd_t *H, double fx, double fy, double cx, double cy) {
    atg_pose_in in{};
    std::memcpy(in.H, H->data, sizeof in.H);
    // the kernel forms homography_to_pose(H, -fx, ...) for the tag pose
    in.fx = -fx;
    in.fy = fy;
    in.cx = cx;
    in.cy = cy;
    in.tagsize = 2.0;
    const atg_pose_out o = pose_one(in, 0, ATG_POSE_HOMOGRAPHY_ONLY);
    double m[16];
    std::memcpy(m, o.hpose, sizeof o.hpose);
    m[12] = m[13] = m[14] = 0.0;
    m[15] = 1.0;
    return matd_create_data(4, 4, m);
}

void estimate_pose_for_tag_homography(apriltag_detection_info_t *info, apriltag_pose_t *pose) {
    const atg_pose_out o = pose_one(tag_of(info), 0, ATG_POSE_HOMOGRAPHY_ONLY);
    set_pose(pose, o.R0, o.t0);
}

void estimate_tag_pose_orthogonal_iteration(apriltag_detection_info_t *info, double *err1, apriltag_pose_t *pose1,
                                            double *err2, apriltag_pose_t *pose2, int nIters) {
    const atg_pose_out o = pose_one(tag_of(info), nIters, 0);
    set_pose(pose1, o.R1, o.t1);
    *err1 = o.err1;
    *err2 = o.err2;
    if (o.has2) {
        set_pose(pose2, o.R2, o.t2);
    } else {
        pose2->R = nullptr;
        pose2->t = nullptr;
    }
}

double estimate_tag_pose(apriltag_detection_info_t *info, apriltag_pose_t *pose) {
    const atg_pose_out o = pose_one(tag_of(info), 50, 0);
    if (o.best == 1) {
        set_pose(pose, o.R1, o.t1);
        return o.err1;
    }
    set_pose(pose, o.R2, o.t2);
    return o.err2;
}

}  // extern "C"
