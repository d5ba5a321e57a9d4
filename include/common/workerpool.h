/* workerpool_t: only the type, for apriltag_detector_t.wp, which stays
 * NULL here (the GPU takes the worker pool's place). */
#ifndef ATG_COMMON_WORKERPOOL_H
#define ATG_COMMON_WORKERPOOL_H
typedef struct workerpool workerpool_t;
#endif
