/* The tagStandard52h13 family for the upstream API (tagStandard52h13.h of the reference). */
#ifndef ATG_TAGSTANDARD52H13_H
#define ATG_TAGSTANDARD52H13_H

#include "apriltag.h"

#ifdef __cplusplus
extern "C" {
#endif

apriltag_family_t *tagStandard52h13_create(void);
void tagStandard52h13_destroy(apriltag_family_t *tf);

#ifdef __cplusplus
}
#endif

#endif
