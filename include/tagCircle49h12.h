/* The tagCircle49h12 family for the upstream API (tagCircle49h12.h of the reference). */
#ifndef ATG_TAGCIRCLE49H12_H
#define ATG_TAGCIRCLE49H12_H

#include "apriltag.h"

#ifdef __cplusplus
extern "C" {
#endif

apriltag_family_t *tagCircle49h12_create(void);
void tagCircle49h12_destroy(apriltag_family_t *tf);

#ifdef __cplusplus
}
#endif

#endif
