/* timeprofile_t: the per-stage stamps of apriltag_detector_t.tp
 * (common/timeprofile.h:41-52 in the reference); header-inline. */
#ifndef ATG_COMMON_TIMEPROFILE_H
#define ATG_COMMON_TIMEPROFILE_H

#include <stdint.h>
#include <stdio.h>
#include <time.h>

#include "zarray.h"

#ifdef __cplusplus
extern "C" {
#endif

static inline int64_t utime_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_REALTIME, &ts);
    return (int64_t)ts.tv_sec * 1000000 + ts.tv_nsec / 1000;
}

struct timeprofile_entry {
    char name[32];
    int64_t utime;
};

typedef struct timeprofile timeprofile_t;
struct timeprofile {
    int64_t utime;
    zarray_t *stamps;
};

static inline timeprofile_t *timeprofile_create(void) {
    timeprofile_t *tp = (timeprofile_t *)calloc(1, sizeof(timeprofile_t));
    tp->stamps = zarray_create(sizeof(struct timeprofile_entry));
    tp->utime = utime_now();
    return tp;
}

static inline void timeprofile_destroy(timeprofile_t *tp) {
    if (!tp) return;
    zarray_destroy(tp->stamps);
    free(tp);
}

static inline void timeprofile_clear(timeprofile_t *tp) {
    zarray_clear(tp->stamps);
    tp->utime = utime_now();
}

static inline void timeprofile_stamp(timeprofile_t *tp, const char *name) {
    struct timeprofile_entry e;
    memset(&e, 0, sizeof e);
    strncpy(e.name, name, sizeof e.name - 1);
    e.utime = utime_now();
    zarray_add(tp->stamps, &e);
}

static inline uint64_t timeprofile_total_utime(timeprofile_t *tp) {
    if (zarray_size(tp->stamps) == 0) return 0;
    struct timeprofile_entry last;
    zarray_get(tp->stamps, zarray_size(tp->stamps) - 1, &last);
    return (uint64_t)(last.utime - tp->utime);
}

static inline void timeprofile_display(timeprofile_t *tp) {
    int64_t prev = tp->utime;
    for (int i = 0; i < zarray_size(tp->stamps); i++) {
        struct timeprofile_entry e;
        zarray_get(tp->stamps, i, &e);
        printf("%2d %32s %15f ms %15f ms\n", i, e.name, (e.utime - prev) / 1000.0, (e.utime - tp->utime) / 1000.0);
        prev = e.utime;
    }
}

#ifdef __cplusplus
}
#endif

#endif
