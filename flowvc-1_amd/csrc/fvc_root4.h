/* fvc_root4.h - x^(1/4) rounded to nearest, for the step-size control of the
 * Runge-Kutta-Fehlberg integrator (ompVC/integration.c:615-618 takes
 * pow(x, 0.25)).  The C library's pow rounds to nearest except within about
 * 0.02 ulp of a tie, so a correctly rounded root gives the same step sizes.
 *
 * y = sqrt(sqrt(x)) lies within one ulp of the root.  y^4 is formed as a
 * double-double with fused multiply-adds (y^2 = p + pe and p^2 = q + qe
 * exactly), the residual x - y^4 divided by the derivative 4 y^3 gives the
 * root's offset from y, and y moves one ulp when that offset passes half the
 * spacing to the neighbour.  Compiled for the device and, by the tests, for
 * the host. */
#ifndef FVC_ROOT4_H
#define FVC_ROOT4_H

#include <math.h>

#ifdef __HIPCC__
#define FVC_HOST_DEVICE __host__ __device__
#else
#define FVC_HOST_DEVICE
#endif

static inline FVC_HOST_DEVICE double fvc_root4_rn(double x) {
#pragma clang fp contract(off)
    if (!(x >= 0x1p-960 && x <= 0x1p960)) return pow(x, 0.25); /* zero, subnormal range, huge, inf, nan */
    double y = sqrt(sqrt(x));
    const double p = y * y, pe = fma(y, y, -p);
    const double q = p * p, qe = fma(p, p, -q);
    const double lo = qe + 2.0 * p * pe;      /* y^4 = q + lo, to about 2^-104 relative (pe^2 dropped) */
    const double d = (x - q) - lo;            /* x - q is exact: q is within a few ulp of x */
    const double off = d / (4.0 * y * y * y); /* root - y, to first order */
    const double up = nextafter(y, INFINITY), dn = nextafter(y, 0.0);
    if (off > 0.5 * (up - y))
        y = up;
    else if (off < -0.5 * (y - dn))
        y = dn;
    return y;
}

#endif
