/* The tag36h11 family for the upstream API (tag36h11.h of the reference). */
#ifndef ATG_TAG36H11_H
#define ATG_TAG36H11_H

#include "apriltag.h"

#ifdef __cplusplus
extern "C" {
#endif

apriltag_family_t *tag36h11_create(void);
void tag36h11_destroy(apriltag_family_t *tf);

#ifdef __cplusplus
}
#endif

#endif
