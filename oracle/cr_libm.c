/* Correctly rounded single-precision atan2f / sinf / cosf / sincosf for the
 * oracle and the reference build in oracle/_ref.
 *
 * The reference's hex goldens (test/golden/std41.hex, std52.hex) were made
 * with a libm whose float atan2f is correctly rounded; glibc 2.35's is not.
 * The reference calls these functions only in refine_edges
 * (fast/decode.c:740-744). Each function here is evaluated in IEEE quad
 * precision (113-bit significand, libquadmath, within an ulp of quad) and
 * rounded once to float: the result is the correctly rounded float unless
 * the exact value lies within about 2^-88 of a float rounding boundary,
 * which no input of the tests comes near (tests/test_trig.py checks these
 * functions against 200-bit mpmath values, at inputs chosen next to such
 * boundaries). The reference build links them with -Wl,--wrap=<name>, so
 * its sources are compiled unchanged.
 *
 * Test infrastructure only: nothing in the product links this file. */
#include <quadmath.h>

float cr_atan2f(float y, float x) { return (float)atan2q((__float128)y, (__float128)x); }
float cr_sinf(float x) { return (float)sinq((__float128)x); }
float cr_cosf(float x) { return (float)cosq((__float128)x); }

float __wrap_atan2f(float y, float x) { return cr_atan2f(y, x); }
float __wrap_sinf(float x) { return cr_sinf(x); }
float __wrap_cosf(float x) { return cr_cosf(x); }
void __wrap_sincosf(float x, float *s, float *c) {
    *s = cr_sinf(x);
    *c = cr_cosf(x);
}
