/* The tagStandard41h12 family for the upstream API (tagStandard41h12.h of the reference). */
#ifndef ATG_TAGSTANDARD41H12_H
#define ATG_TAGSTANDARD41H12_H

#include "apriltag.h"

#ifdef __cplusplus
extern "C" {
#endif

apriltag_family_t *tagStandard41h12_create(void);
void tagStandard41h12_destroy(apriltag_family_t *tf);

#ifdef __cplusplus
}
#endif

#endif
