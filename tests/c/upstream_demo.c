/* A caller of the upstream AprilTag API, written as a user's program is
 * (example/apriltag_demo.c of the reference uses the same calls): detect
 * in each PNM given and print one line per detection in the golden hex
 * format of the reference's tests (image id hamming margin c p H, %a),
 * then "# image nquads N". Options follow the reference's test/cases.txt:
 * -t threads, -x decimate, -b sigma, -f family, -a bits corrected (1 by
 * default, as there), -0 no edge refinement, --deglitch. With -s it prints
 * the sizes of the public structs and exits without touching the GPU. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "apriltag.h"
#include "apriltag_pose.h"
#include "common/homography.h"
#include "tag16h5.h"
#include "tag25h9.h"
#include "tag36h10.h"
#include "tag36h11.h"
#include "tagCircle21h7.h"
#include "tagCircle49h12.h"
#include "tagCustom48h12.h"
#include "tagStandard41h12.h"
#include "tagStandard52h13.h"

static apriltag_family_t *make_family(const char *name) {
    if (!strcmp(name, "tag16h5")) return tag16h5_create();
    if (!strcmp(name, "tag25h9")) return tag25h9_create();
    if (!strcmp(name, "tag36h10")) return tag36h10_create();
    if (!strcmp(name, "tag36h11")) return tag36h11_create();
    if (!strcmp(name, "tagCircle21h7")) return tagCircle21h7_create();
    if (!strcmp(name, "tagCircle49h12")) return tagCircle49h12_create();
    if (!strcmp(name, "tagCustom48h12")) return tagCustom48h12_create();
    if (!strcmp(name, "tagStandard41h12")) return tagStandard41h12_create();
    if (!strcmp(name, "tagStandard52h13")) return tagStandard52h13_create();
    return NULL;
}

int main(int argc, char **argv) {
    const char *fname = "tag36h11";
    int bits = 1, first = 1, threads = 1, refine = 1, deglitch = 0;
    float decimate = 2.0f, sigma = 0.0f;
    for (; first < argc && argv[first][0] == '-'; first++) {
        const char *a = argv[first];
        if (!strcmp(a, "-s")) {
            printf("apriltag_detector_t %zu\napriltag_detection_t %zu\napriltag_family_t %zu\nimage_u8_t %zu\n"
                   "zarray_t %zu\nmatd_t %zu\napriltag_detection_info_t %zu\napriltag_pose_t %zu\n",
                   sizeof(apriltag_detector_t), sizeof(apriltag_detection_t), sizeof(apriltag_family_t),
                   sizeof(image_u8_t), sizeof(zarray_t), sizeof(matd_t), sizeof(apriltag_detection_info_t),
                   sizeof(apriltag_pose_t));
            return 0;
        } else if (!strcmp(a, "-0")) {
            refine = 0;
        } else if (!strcmp(a, "--deglitch")) {
            deglitch = 1;
        } else if (first + 1 < argc && !strcmp(a, "-t")) {
            threads = atoi(argv[++first]);
        } else if (first + 1 < argc && !strcmp(a, "-x")) {
            decimate = (float)atof(argv[++first]);
        } else if (first + 1 < argc && !strcmp(a, "-b")) {
            sigma = (float)atof(argv[++first]);
        } else if (first + 1 < argc && !strcmp(a, "-f")) {
            fname = argv[++first];
        } else if (first + 1 < argc && !strcmp(a, "-a")) {
            bits = atoi(argv[++first]);
        } else {
            return 2;
        }
    }
    apriltag_family_t *tf = make_family(fname);
    if (!tf) return 2;
    apriltag_detector_t *td = apriltag_detector_create();
    apriltag_detector_add_family_bits(td, tf, bits);
    td->nthreads = threads;
    td->quad_decimate = decimate;
    td->quad_sigma = sigma;
    td->refine_edges = refine != 0;
    td->qtp.deglitch = deglitch;
    for (int a = first; a < argc; a++) {
        image_u8_t *im = image_u8_create_from_pnm(argv[a]);
        if (!im) return 3;
        zarray_t *dets = apriltag_detector_detect(td, im);
        for (int i = 0; i < zarray_size(dets); i++) {
            apriltag_detection_t *d;
            zarray_get(dets, i, &d);
            printf("%s %d %d %a %a %a", argv[a], d->id, d->hamming, (double)d->decision_margin, d->c[0], d->c[1]);
            for (int k = 0; k < 4; k++) printf(" %a %a", d->p[k][0], d->p[k][1]);
            for (int k = 0; k < 9; k++) printf(" %a", MATD_EL(d->H, k / 3, k % 3));
            printf("\n");
        }
        printf("# %s nquads %u\n", argv[a], td->nquads);
        for (int i = 0; i < zarray_size(td->tp->stamps); i++) {
            struct timeprofile_entry e;
            zarray_get(td->tp->stamps, i, &e);
            printf("# tp %s|%lld\n", e.name, (long long)(e.utime - td->tp->utime));
        }
        apriltag_detections_destroy(dets);
        image_u8_destroy(im);
    }
    apriltag_detector_destroy(td);
    apriltag_family_t *again = make_family(fname);
    image_u8_t *tagim = apriltag_to_image(again, 0);
    printf("# to_image %d %d %d\n", tagim->width, tagim->height, tagim->buf[tagim->stride + 1]);
    image_u8_destroy(tagim);
    return 0;
}
