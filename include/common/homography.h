/* The homography helpers of the upstream API the pose needs
 * (common/homography.h:169 of the reference). */
#ifndef ATG_COMMON_HOMOGRAPHY_H
#define ATG_COMMON_HOMOGRAPHY_H

#include "matd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* common/homography.c:275-351: the 4x4 [R | T; 0 0 0 1] of the camera
 * (looking down -z) whose projection with intrinsics (fx, fy, cx, cy) the
 * 3x3 homography H is; R is the nearest rotation (polar factor). Runs on
 * the GPU as estimate_pose_for_tag_homography does. */
matd_t *homography_to_pose(const matd_t *H, double fx, double fy, double cx, double cy);

#ifdef __cplusplus
}
#endif

#endif
