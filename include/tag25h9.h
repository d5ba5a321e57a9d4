/* The tag25h9 family for the upstream API (tag25h9.h of the reference). */
#ifndef ATG_TAG25H9_H
#define ATG_TAG25H9_H

#include "apriltag.h"

#ifdef __cplusplus
extern "C" {
#endif

apriltag_family_t *tag25h9_create(void);
void tag25h9_destroy(apriltag_family_t *tf);

#ifdef __cplusplus
}
#endif

#endif
