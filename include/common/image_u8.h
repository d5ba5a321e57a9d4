/* The image_u8 functions an upstream caller needs around detection:
 * creation (rows padded to the reference's 96-byte alignment), PNM input
 * and output, destruction. */
#ifndef ATG_COMMON_IMAGE_U8_H
#define ATG_COMMON_IMAGE_U8_H

#include "image_types.h"

#ifdef __cplusplus
extern "C" {
#endif

image_u8_t *image_u8_create(unsigned int width, unsigned int height);
image_u8_t *image_u8_create_stride(unsigned int width, unsigned int height, unsigned int stride);
image_u8_t *image_u8_create_alignment(unsigned int width, unsigned int height, unsigned int alignment);
image_u8_t *image_u8_copy(const image_u8_t *in);
image_u8_t *image_u8_create_from_pnm(const char *path);  /* NULL on failure */
int image_u8_write_pnm(const image_u8_t *im, const char *path);  /* 0 on success */
void image_u8_destroy(image_u8_t *im);

#ifdef __cplusplus
}
#endif

#endif
