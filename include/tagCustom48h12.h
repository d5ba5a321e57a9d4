/* The tagCustom48h12 family for the upstream API (tagCustom48h12.h of the reference). */
#ifndef ATG_TAGCUSTOM48H12_H
#define ATG_TAGCUSTOM48H12_H

#include "apriltag.h"

#ifdef __cplusplus
extern "C" {
#endif

apriltag_family_t *tagCustom48h12_create(void);
void tagCustom48h12_destroy(apriltag_family_t *tf);

#ifdef __cplusplus
}
#endif

#endif
