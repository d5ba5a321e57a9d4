/* The tag16h5 family for the upstream API (tag16h5.h of the reference). */
#ifndef ATG_TAG16H5_H
#define ATG_TAG16H5_H

#include "apriltag.h"

#ifdef __cplusplus
extern "C" {
#endif

apriltag_family_t *tag16h5_create(void);
void tag16h5_destroy(apriltag_family_t *tf);

#ifdef __cplusplus
}
#endif

#endif
