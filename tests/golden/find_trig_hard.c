/* Inputs on which rounding a double-precision atan2f / cosf / sinf result
 * to float is hardest: the double result (glibc) lies within 2^-50 of its
 * magnitude from a float rounding boundary (a midpoint between floats).
 * Test-vector search for make_trig_golden.py (test infrastructure).
 *
 *   find_trig_hard cos|sin          every float in (0, pi/2]
 *   find_trig_hard atan2 N SEED     N seeded pairs of floats, magnitudes
 *                                   2^-12 .. 2^12, both signs
 * prints one input a line (hex floats). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int near_boundary(double r) {
    float f = (float)r;
    double fd = f, up = nextafterf(f, INFINITY), dn = nextafterf(f, -INFINITY);
    double eb = fabs(r) * 0x1p-50;
    return fabs(r - 0.5 * (fd + up)) <= eb || fabs(r - 0.5 * (fd + dn)) <= eb;
}

static uint64_t splitmix(uint64_t *s) {
    uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static float rnd_float(uint64_t *s) {
    uint64_t v = splitmix(s);
    float m = 1.0f + (float)(v & 0xffffff) * 0x1p-24f;  /* [1, 2) */
    int e = (int)((v >> 24) % 25) - 12;
    float x = ldexpf(m, e);
    return (v >> 40) & 1 ? -x : x;
}

int main(int argc, char **argv) {
    if (argc >= 2 && (!strcmp(argv[1], "cos") || !strcmp(argv[1], "sin"))) {
        const int use_cos = !strcmp(argv[1], "cos");
        const float lim = 1.57079637f;
        uint32_t lo, hi;
        float one = 0x1p-149f;
        memcpy(&lo, &one, 4);
        memcpy(&hi, &lim, 4);
        #pragma omp parallel for schedule(static, 1 << 20)
        for (int64_t b = lo; b <= (int64_t)hi; b++) {
            uint32_t u = (uint32_t)b;
            float x;
            memcpy(&x, &u, 4);
            double r = use_cos ? cos((double)x) : sin((double)x);
            if (r != 0.0 && near_boundary(r)) {
                #pragma omp critical
                printf("%a\n", (double)x);
            }
        }
        return 0;
    }
    if (argc >= 4 && !strcmp(argv[1], "atan2")) {
        const int64_t n = atoll(argv[2]);
        const uint64_t seed = strtoull(argv[3], 0, 10);
        #pragma omp parallel
        {
            #pragma omp for schedule(static, 1 << 20)
            for (int64_t i = 0; i < n; i++) {
                uint64_t s = seed * 0x100000001b3ull + (uint64_t)i * 0x9e3779b97f4a7c15ull;
                float y = rnd_float(&s), x = rnd_float(&s);
                double r = atan2((double)y, (double)x);
                if (near_boundary(r)) {
                    #pragma omp critical
                    printf("%a %a\n", (double)y, (double)x);
                }
            }
        }
        return 0;
    }
    return 2;
}
