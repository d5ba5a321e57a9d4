/* Tag pose from a detection: the reference's apriltag_pose.h API
 * (apriltag_pose.h:1-77), computed on the GPU (atg_estimate_poses in
 * atg.h; a batch of tags is one launch there).
 *
 * The upstream functions below run on device 0 through a context of
 * their own, created at the first call; calls from several threads take
 * turns. R (3x3) and t (3x1) are new matd_t the caller frees with
 * matd_destroy, as in the reference. */
#ifndef ATG_APRILTAG_POSE_H
#define ATG_APRILTAG_POSE_H

#include "apriltag.h"
#include "common/matd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* apriltag_pose.h:11-18 */
typedef struct {
    apriltag_detection_t *det;
    double tagsize;  /* metres */
    double fx, fy;   /* pixels */
    double cx, cy;   /* pixels */
} apriltag_detection_info_t;

/* apriltag_pose.h:25-28: the tag in the camera frame */
typedef struct {
    matd_t *R;  /* 3x3 */
    matd_t *t;  /* 3x1 */
} apriltag_pose_t;

/* apriltag_pose.h:34-36 (apriltag_pose.c:455-488): the pose from the
 * homography alone. */
void estimate_pose_for_tag_homography(apriltag_detection_info_t *info, apriltag_pose_t *pose);

/* apriltag_pose.h:60-66 (apriltag_pose.c:490-526): the homography pose
 * refined by nIters steps of orthogonal iteration, its error in *err1;
 * the second local minimum refined the same way in pose2 (pose2->R and
 * pose2->t NULL and *err2 = HUGE_VAL when there is none). */
void estimate_tag_pose_orthogonal_iteration(apriltag_detection_info_t *info, double *err1, apriltag_pose_t *pose1,
                                            double *err2, apriltag_pose_t *pose2, int nIters);

/* apriltag_pose.h:75 (apriltag_pose.c:528-547): of the two, the one of
 * smaller object-space error (the first on a tie), 50 steps; returns its
 * error. */
double estimate_tag_pose(apriltag_detection_info_t *info, apriltag_pose_t *pose);

#ifdef __cplusplus
}
#endif

#endif
