/* Exhaustive check behind the device's table-position division: for every
 * float x in [lo, hi], x / 5 (IEEE, correctly rounded) equals the
 * multiply-and-two-FMA form q = x * RN(1/5); r = fma(-q, 5, x);
 * fma(r, RN(1/5), q).  Prints the mismatch count. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv) {
    float lo = argc > 1 ? strtof(argv[1], NULL) : 0.0f, hi = argc > 2 ? strtof(argv[2], NULL) : 350.0f;
    uint32_t a, b;
    memcpy(&a, &lo, 4);
    memcpy(&b, &hi, 4);
    unsigned long long bad = 0, n = 0;
#pragma omp parallel for reduction(+ : bad, n) schedule(static)
    for (int64_t u = a; u <= (int64_t)b; ++u) {
        uint32_t uu = (uint32_t)u;
        float x;
        memcpy(&x, &uu, 4);
        volatile float five = 5.0f;
        float q1 = x / five;
        float q = x * 0.2f;
        float r = fmaf(-q, 5.0f, x);
        float q2 = fmaf(r, 0.2f, q);
        if (memcmp(&q1, &q2, 4)) ++bad;
        ++n;
    }
    printf("%llu %llu\n", n, bad);
    return 0;
}
