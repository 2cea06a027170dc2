// vx3_powf.h -- pow of single-precision arguments near 1, rounded to single
// precision, for CVX_Voxel::strain's Poisson term pow(1 + add, -nu)
// (Voxelyze3/src/old/VX_Voxel.cpp:299-333).
//
// The reference evaluates pow of two floats in double precision and rounds once
// (glibc); the full double pow costs ~150 instructions per voxel per step.  For
// |base - 1| <= 1/4 and |nu * ln(base)| <= 1/5 this evaluates
//   pow(1 + t, y) = exp(y * ln(1 + t)),  ln(1 + t) = 2 atanh(t / (2 + t)),
// both by their Taylor series in double (relative error below 1e-15), and keeps
// the result only when every double within 2e-15 relative of it rounds to the same
// float -- then the correctly rounded value rounds to that float too, and the
// result equals (float)pow(double, double).  Otherwise (outside the range, or a
// value within 2e-15 of a float rounding boundary, about 1 in 3e7) it reports
// failure and the caller falls back to the full pow.  Header-only so the host test
// can build the same code with g++.
#ifndef VX3_POWF_H
#define VX3_POWF_H

#include <math.h>

#ifdef __HIPCC__
#define VX3_HD __host__ __device__ __forceinline__
#else
#define VX3_HD inline
#endif

namespace vx3 {

VX3_HD bool powNearOne(float base, float negExp, float& out) {
    const double t = (double)base - 1.0;   // exact: base is a float near 1
    const double y = (double)negExp;
    if (!(fabs(t) <= 0.25)) return false;
    const double u = t / (2.0 + t), u2 = u * u;   // |u| <= 1/7
    double s = 1.0 / 21;
    s = s * u2 + 1.0 / 19; s = s * u2 + 1.0 / 17; s = s * u2 + 1.0 / 15; s = s * u2 + 1.0 / 13;
    s = s * u2 + 1.0 / 11; s = s * u2 + 1.0 / 9;  s = s * u2 + 1.0 / 7;  s = s * u2 + 1.0 / 5;
    s = s * u2 + 1.0 / 3;  s = s * u2 + 1.0;
    const double z = y * (2.0 * u * s);   // y * ln(1 + t)
    if (!(fabs(z) <= 0.2)) return false;
    double e = 1.0 / 87178291200.0;       // 1/14!
    e = e * z + 1.0 / 6227020800.0; e = e * z + 1.0 / 479001600.0; e = e * z + 1.0 / 39916800.0;
    e = e * z + 1.0 / 3628800.0;     e = e * z + 1.0 / 362880.0;    e = e * z + 1.0 / 40320.0;
    e = e * z + 1.0 / 5040.0;        e = e * z + 1.0 / 720.0;       e = e * z + 1.0 / 120.0;
    e = e * z + 1.0 / 24.0;          e = e * z + 1.0 / 6.0;         e = e * z + 0.5;
    e = e * z + 1.0;                 e = e * z + 1.0;
    const float lo = (float)(e * (1.0 - 2e-15)), hi = (float)(e * (1.0 + 2e-15));
    if (lo != hi) return false;
    out = lo;
    return true;
}

}  // namespace vx3
#endif
