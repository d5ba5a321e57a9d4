/* The tagCircle21h7 family for the upstream API (tagCircle21h7.h of the reference). */
#ifndef ATG_TAGCIRCLE21H7_H
#define ATG_TAGCIRCLE21H7_H

#include "apriltag.h"

#ifdef __cplusplus
extern "C" {
#endif

apriltag_family_t *tagCircle21h7_create(void);
void tagCircle21h7_destroy(apriltag_family_t *tf);

#ifdef __cplusplus
}
#endif

#endif
