/* zarray_t: the resizable array of the upstream AprilTag API
 * (common/zarray.h:43-50 in the reference: element size, count, capacity,
 * data). The functions are header-inline, as there, so that code written
 * against the upstream API compiles unchanged. */
#ifndef ATG_COMMON_ZARRAY_H
#define ATG_COMMON_ZARRAY_H

#include <assert.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct zarray zarray_t;
struct zarray {
    size_t el_sz;
    int size;
    int alloc;
    char *data;
};

static inline zarray_t *zarray_create(size_t el_sz) {
    zarray_t *za = (zarray_t *)calloc(1, sizeof(zarray_t));
    za->el_sz = el_sz;
    return za;
}

static inline void zarray_destroy(zarray_t *za) {
    if (!za) return;
    free(za->data);
    free(za);
}

static inline int zarray_size(const zarray_t *za) { return za->size; }

static inline void zarray_ensure_capacity(zarray_t *za, int capacity) {
    if (capacity <= za->alloc) return;
    int n = za->alloc > 8 ? za->alloc : 8;
    while (n < capacity) n *= 2;
    za->data = (char *)realloc(za->data, (size_t)n * za->el_sz);
    za->alloc = n;
}

static inline void zarray_add(zarray_t *za, const void *p) {
    zarray_ensure_capacity(za, za->size + 1);
    memcpy(za->data + (size_t)za->size * za->el_sz, p, za->el_sz);
    za->size++;
}

static inline void zarray_get(const zarray_t *za, int idx, void *p) {
    assert(idx >= 0 && idx < za->size);
    memcpy(p, za->data + (size_t)idx * za->el_sz, za->el_sz);
}

static inline void zarray_get_volatile(const zarray_t *za, int idx, void *p) {
    assert(idx >= 0 && idx < za->size);
    *(void **)p = za->data + (size_t)idx * za->el_sz;
}

static inline void zarray_set(zarray_t *za, int idx, const void *p, void *outp) {
    assert(idx >= 0 && idx < za->size);
    if (outp) memcpy(outp, za->data + (size_t)idx * za->el_sz, za->el_sz);
    memcpy(za->data + (size_t)idx * za->el_sz, p, za->el_sz);
}

/* shuffle != 0: the last element takes the removed one's place (order not
 * kept); else the later elements move down. */
static inline void zarray_remove_index(zarray_t *za, int idx, int shuffle) {
    assert(idx >= 0 && idx < za->size);
    if (shuffle) {
        if (idx < za->size - 1)
            memcpy(za->data + (size_t)idx * za->el_sz, za->data + (size_t)(za->size - 1) * za->el_sz, za->el_sz);
    } else {
        memmove(za->data + (size_t)idx * za->el_sz, za->data + (size_t)(idx + 1) * za->el_sz,
                (size_t)(za->size - idx - 1) * za->el_sz);
    }
    za->size--;
}

static inline void zarray_insert(zarray_t *za, int idx, const void *p) {
    assert(idx >= 0 && idx <= za->size);
    zarray_ensure_capacity(za, za->size + 1);
    memmove(za->data + (size_t)(idx + 1) * za->el_sz, za->data + (size_t)idx * za->el_sz,
            (size_t)(za->size - idx) * za->el_sz);
    memcpy(za->data + (size_t)idx * za->el_sz, p, za->el_sz);
    za->size++;
}

static inline void zarray_clear(zarray_t *za) { za->size = 0; }

static inline int zarray_index_of(const zarray_t *za, const void *p) {
    for (int i = 0; i < za->size; i++)
        if (!memcmp(p, za->data + (size_t)i * za->el_sz, za->el_sz)) return i;
    return -1;
}

static inline int zarray_contains(const zarray_t *za, const void *p) { return zarray_index_of(za, p) >= 0; }

static inline int zarray_remove_value(zarray_t *za, const void *p, int shuffle) {
    int i = zarray_index_of(za, p);
    if (i < 0) return 0;
    zarray_remove_index(za, i, shuffle);
    return 1;
}

static inline zarray_t *zarray_copy(const zarray_t *za) {
    zarray_t *c = zarray_create(za->el_sz);
    zarray_ensure_capacity(c, za->size);
    if (za->size) memcpy(c->data, za->data, (size_t)za->size * za->el_sz);
    c->size = za->size;
    return c;
}

static inline void zarray_map(zarray_t *za, void (*f)(void *)) {
    for (int i = 0; i < za->size; i++) f(*(void **)(za->data + (size_t)i * za->el_sz));
}

static inline void zarray_sort(zarray_t *za, int (*compar)(const void *, const void *)) {
    if (za->size > 1) qsort(za->data, (size_t)za->size, za->el_sz, compar);
}

#ifdef __cplusplus
}
#endif

#endif
