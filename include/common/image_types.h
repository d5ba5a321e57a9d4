/* image_u8_t (common/image_types.h:37-44 in the reference). */
#ifndef ATG_COMMON_IMAGE_TYPES_H
#define ATG_COMMON_IMAGE_TYPES_H

#include <stdint.h>

typedef struct image_u8 image_u8_t;
struct image_u8 {
    const int32_t width;
    const int32_t height;
    const int32_t stride;
    uint8_t *buf;
};

#endif
