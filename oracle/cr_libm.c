/* Correctly rounded single-precision atan2f / sinf / cosf / sincosf for the
 * reference build in oracle/_ref.
 *
 * The reference's hex goldens (test/golden/std41.hex, std52.hex) were made
 * with a libm whose float atan2f is correctly rounded; glibc 2.35's is not.
 * The reference calls these functions only in refine_edges
 * (fast/decode.c:740-744). Each wrapper evaluates the function in double
 * precision and rounds once to float. Linked with -Wl,--wrap=<name>, so the
 * reference sources are compiled unchanged.
 *
 * Test infrastructure only: nothing in the product links this file. */
#include <math.h>

float __wrap_atan2f(float y, float x) { return (float)atan2((double)y, (double)x); }
float __wrap_sinf(float x) { return (float)sin((double)x); }
float __wrap_cosf(float x) { return (float)cos((double)x); }
void __wrap_sincosf(float x, float *s, float *c) {
    *s = (float)sin((double)x);
    *c = (float)cos((double)x);
}
