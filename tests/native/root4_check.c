/* Compares fvc_root4_rn (flowvc-1_amd/csrc/fvc_root4.h) with the C library's
 * pow(x, 0.25): log-uniform samples over [2^-60, 2^60], values next to exact
 * fourth powers, and values next to fourth powers of midpoints between
 * doubles.  Prints the first differing inputs with both results (hex floats),
 * then the sample count and the number of differing results. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "fvc_root4.h"

static uint64_t s = 0x9e3779b97f4a7c15ull;
static uint64_t next(void) { /* splitmix64 */
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static long n, bad;
static void check(double x) {
    if (!(x > 0.0)) return;
    n++;
    double a = fvc_root4_rn(x), b = pow(x, 0.25);
    if (a != b) {
        if (bad < 200) printf("differs %a %a %a\n", x, a, b);
        bad++;
    }
}

int main(int argc, char **argv) {
    long m = argc > 1 ? atol(argv[1]) : 20000000;
    for (long i = 0; i < m; i++) {
        double u = (double)(next() >> 11) * 0x1p-53;
        check(exp2(-60.0 + 120.0 * u));
    }
    for (long k = 1; k < 2000000; k++) { /* near k^4 and (k + 1/2)^4 */
        double r = (double)k, x = r * r * r * r;
        check(x);
        check(nextafter(x, 0.0));
        check(nextafter(x, INFINITY));
        double h = r + 0.5, y = h * h * h * h;
        check(y);
        check(nextafter(y, 0.0));
        check(nextafter(y, INFINITY));
    }
    for (long i = 0; i < m / 4; i++) { /* fourth powers of midpoints between neighbouring doubles */
        double y = exp2(-10.0 + 20.0 * (double)(next() >> 11) * 0x1p-53);
        double mid = y + 0.5 * (nextafter(y, INFINITY) - y);
        double x = mid * mid * mid * mid;
        check(x);
        check(nextafter(x, 0.0));
        check(nextafter(x, INFINITY));
    }
    printf("%ld samples, %ld differ\n", n, bad);
    return 0;
}
