// Double-precision sin, cos and acos of this library.
//
// The FP64 step calls acos and sin in the curvature (InvRotateDivide,
// InvRotateDivide/Scalar.hpp:76-103), sin and cos in the director update
// (SO3PrimitiveAddAssign/Scalar.hpp:25-96) and sin in the sin ramp of an end
// force (Rampup.hpp:302-330).  Library implementations of these functions
// (glibc on the host, the device library on gfx950) differ from each other in
// the last bit for a share of their arguments, and an FP64 rod step carries
// such a difference forward like any other rounding.  These versions use
// only additions, multiplications, divisions and square roots, each rounded
// once (IEEE round to nearest), in a fixed order, so a CPU build that uses
// them and the device kernels compute the same bits for the same argument;
// the units that include this header are built without FMA contraction.
//
// Method
//   sin, cos: Cody-Waite reduction (W. J. Cody and W. Waite, Software Manual
//   for the Elementary Functions, 1980): n = nearest integer to x 2/pi and
//   r = ((x - n c1) - n c2) - n c3, with pi/2 = c1 + c2 + c3, c1 and c2 of 33
//   significant bits (n c1 and n c2 are exact for |n| < 2^20, that is
//   |x| < 1.6e6; beyond that the reduction loses accuracy but stays the same
//   sequence of operations), r kept as a head and the rounding errors of the
//   two subtractions (Knuth's two-sum); then the Taylor series of sin and cos
//   on |r| <= pi/4, through r^19 and r^18 (the first omitted terms are below
//   2e-21 there), by Horner's rule in r^2, the leading terms added last, plus
//   the first-order terms of the tail.  cos keeps the rounding error of
//   1 - r^2/2 (the subtraction is exact when r^2/2 <= 1/2).  Quadrant n mod 4
//   selects +-sin(r) or +-cos(r).
//   acos: the Taylor series of asin, asin(y) = sum_k C(2k, k) / (4^k (2k + 1))
//   y^(2k+1), through k = 23 on |y| <= 1/2 (the tail is below 1e-17 of y);
//   acos(x) = pi/2 - asin(x) for |x| <= 1/2, and, with y = sqrt((1 - |x|) / 2)
//   (y^2 = (1 - |x|) / 2 exactly), acos(x) = 2 asin(y) for x > 1/2 and
//   pi - 2 asin(y) for x < -1/2; y is carried as a 26-bit head plus the
//   correction (y^2 - head^2) / (y + head), the head's square being exact.
//   |x| > 1 and NaN give NaN (0 / 0, raising the invalid flag), like the
//   library function.
// Accuracy (against glibc, tests/test_math_functions.py): within 1 ulp.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define EROD_MATH_FN __host__ __device__ inline
#else
#define EROD_MATH_FN inline
#endif

namespace erod {
  namespace math {

    // pi/2 in three parts (c1, c2 of 33 significant bits), 2/pi, and pi/2 and
    // pi as a double plus the double nearest to the rest
    constexpr double kPio2C1 = 0x1.921fb54400000p+0;
    constexpr double kPio2C2 = 0x1.0b4611a600000p-34;
    constexpr double kPio2C3 = 0x1.3198a2e037073p-69;
    constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
    constexpr double kPio2Hi = 0x1.921fb54442d18p+0;
    constexpr double kPio2Lo = 0x1.1a62633145c07p-54;

    // (-1)^k / (2k + 1)!, k = 1 ... 9, and (-1)^k / (2k)!, k = 2 ... 9
    constexpr double kSin[9] = {-0x1.5555555555555p-3,  0x1.1111111111111p-7,  -0x1.a01a01a01a01ap-13,
                                0x1.71de3a556c734p-19,  -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33,
                                -0x1.ae7f3e733b81fp-41, 0x1.952c77030ad4ap-49,  -0x1.2f49b46814157p-57};
    constexpr double kCos[8] = {0x1.5555555555555p-5,   -0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-16,
                                -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29,  -0x1.93974a8c07c9dp-37,
                                0x1.ae7f3e733b81fp-45,  -0x1.6827863b97d97p-53};

    // C(2k, k) / (4^k (2k + 1)), k = 0 ... 23: numerator and denominator are
    // integers below 2^53, so one division rounds the rational once
    constexpr double asin_coef(int k) {
      double c = 1.0, p4 = 1.0;
      for (int i = 1; i <= k; ++i) {
        c = c * double(k + i) / double(i);  // C(k + i, i): exact integers at every stage for k <= 23
        p4 *= 4.0;
      }
      return c / (p4 * double(2 * k + 1));
    }
    struct AsinCoefs {
      double c[24];
      constexpr AsinCoefs() : c() {
        for (int k = 0; k < 24; ++k) c[k] = asin_coef(k);
      }
    };
    constexpr AsinCoefs kAsin{};

    EROD_MATH_FN double bits_head26(double v) {
      // v with its 27 low significand bits cleared (a 26-bit head whose square is exact)
      uint64_t u;
      __builtin_memcpy(&u, &v, sizeof(u));
      u &= ~((uint64_t(1) << 27) - 1);
      double h;
      __builtin_memcpy(&h, &u, sizeof(h));
      return h;
    }

    // sin and cos of a reduced argument r = hi + lo, |r| <= pi/4 (slightly
    // more after rounding), lo below half an ulp of hi: the series in hi plus
    // the first-order terms of lo (cos(hi) lo = lo (1 - hi^2 / 2) and
    // -sin(hi) lo = -hi lo, to the order that reaches the last bit)
    EROD_MATH_FN double sin_kernel(double hi, double lo) {
      const double z = hi * hi;
      double p = kSin[8];
      for (int i = 7; i >= 0; --i) p = p * z + kSin[i];
      return hi + (hi * (z * p) + lo * (1.0 - 0.5 * z));
    }
    EROD_MATH_FN double cos_kernel(double hi, double lo) {
      const double z = hi * hi;
      double q = kCos[7];
      for (int i = 6; i >= 0; --i) q = q * z + kCos[i];
      const double hz = 0.5 * z;
      const double w = 1.0 - hz;
      return w + (((1.0 - w) - hz) + ((z * z) * q - hi * lo));
    }

    // a + b = s + err exactly (Knuth's two-sum)
    EROD_MATH_FN double two_sum(double a, double b, double* err) {
      const double s = a + b;
      const double bb = s - a;
      *err = (a - (s - bb)) + (b - bb);
      return s;
    }

    // x = n pi/2 + hi + lo; quadrant n mod 4 in *q.  x - n c1 and n c2 are
    // exact (|n| < 2^20); the two subtractions that follow keep their rounding
    // errors in lo
    EROD_MATH_FN double reduce(double x, double* lo, int* q) {
      const double n = __builtin_rint(x * kTwoOverPi);
      const double t = x - n * kPio2C1;
      double e1, e2;
      const double r1 = two_sum(t, -(n * kPio2C2), &e1);
      const double r2 = two_sum(r1, -(n * kPio2C3), &e2);
      const double l = e1 + e2;
      const double hi = r2 + l;
      *lo = l - (hi - r2);
      const double m = n - 4.0 * __builtin_floor(n * 0.25);  // n mod 4, exact for every finite n
      *q = int(m);
      return hi;
    }

    EROD_MATH_FN bool finite(double x) { return x - x == 0.0; }

    // Quadrant n mod 4 selects the result: sin x = s, c, -s, -c and
    // cos x = c, -s, -c, s for n = 0 ... 3 (both kernels are evaluated and
    // one is selected, so a wave takes no divergent branch).
    // below 2^-27 the series' first correction is under half an ulp: sin x
    // rounds to x (signed zeros kept) and cos x to 1
    constexpr double kTiny = 0x1p-27;

    EROD_MATH_FN double sin(double x) {
      if (!finite(x)) return x - x;  // NaN for +-inf and NaN
      if ((x < 0.0 ? -x : x) < kTiny) return x;
      int q;
      double lo;
      const double hi = reduce(x, &lo, &q);
      const double s = sin_kernel(hi, lo), c = cos_kernel(hi, lo);
      const double v = (q & 1) ? c : s;
      return (q & 2) ? -v : v;
    }

    EROD_MATH_FN double cos(double x) {
      if (!finite(x)) return x - x;
      if ((x < 0.0 ? -x : x) < kTiny) return 1.0;
      int q;
      double lo;
      const double hi = reduce(x, &lo, &q);
      const double s = sin_kernel(hi, lo), c = cos_kernel(hi, lo);
      const double v = (q & 1) ? s : c;
      return ((q + 1) & 2) ? -v : v;
    }

    // sin and cos of one argument, one reduction (equal to sin(x) and cos(x))
    EROD_MATH_FN void sincos(double x, double* sp, double* cp) {
      if (!finite(x)) {
        *sp = *cp = x - x;
        return;
      }
      if ((x < 0.0 ? -x : x) < kTiny) {
        *sp = x;
        *cp = 1.0;
        return;
      }
      int q;
      double lo;
      const double hi = reduce(x, &lo, &q);
      const double s = sin_kernel(hi, lo), c = cos_kernel(hi, lo);
      const double vs = (q & 1) ? c : s, vc = (q & 1) ? s : c;
      *sp = (q & 2) ? -vs : vs;
      *cp = ((q + 1) & 2) ? -vc : vc;
    }

    // asin(y) - y for |y| <= 1/2, from z = y^2
    EROD_MATH_FN double asin_tail(double y, double z) {
      double p = kAsin.c[23];
      for (int k = 22; k >= 1; --k) p = p * z + kAsin.c[k];
      return y * (z * p);
    }

    EROD_MATH_FN double acos(double x) {
      const double ax = x < 0.0 ? -x : x;
      if (!(ax <= 1.0)) return (x - x) / (x - x);  // |x| > 1 or NaN: invalid
      if (ax <= 0.5) {
        // pi/2 - asin(x) = pi/2 - x - (asin(x) - x)
        const double t = asin_tail(x, x * x);
        return kPio2Hi - (x - (kPio2Lo - t));
      }
      const double z = (1.0 - ax) * 0.5;  // exact: 1 - ax by Sterbenz, then a halving
      const double y = __builtin_sqrt(z);
      const double t = asin_tail(y, z);
      if (x > 0.0) {
        // 2 asin(y) with y = h + c, h^2 exact, c = (z - h^2) / (y + h);
        // acos(1) = 0 (y = 0, no correction)
        if (!(y > 0.0)) return 0.0;
        const double h = bits_head26(y);
        const double c = (z - h * h) / (y + h);
        return 2.0 * (h + (c + t));
      }
      // pi - 2 asin(y) = 2 (pi/2 - asin(y))
      return 2.0 * (kPio2Hi - (y + (t - kPio2Lo)));
    }

  }  // namespace math
}  // namespace erod
