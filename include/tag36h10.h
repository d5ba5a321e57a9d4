/* The tag36h10 family for the upstream API (tag36h10.h of the reference). */
#ifndef ATG_TAG36H10_H
#define ATG_TAG36H10_H

#include "apriltag.h"

#ifdef __cplusplus
extern "C" {
#endif

apriltag_family_t *tag36h10_create(void);
void tag36h10_destroy(apriltag_family_t *tf);

#ifdef __cplusplus
}
#endif

#endif
