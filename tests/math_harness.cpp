// Test harness (built by tests/test_math_functions.py with g++): the library's
// double sin, cos and acos and their constant tables, as a C interface.
#include "../elasticapp_amd/csrc/erod_math.hpp"

extern "C" {
void h_sin(const double* x, double* y, long n) {
  for (long i = 0; i < n; ++i) y[i] = erod::math::sin(x[i]);
}
void h_cos(const double* x, double* y, long n) {
  for (long i = 0; i < n; ++i) y[i] = erod::math::cos(x[i]);
}
void h_sincos(const double* x, double* s, double* c, long n) {
  for (long i = 0; i < n; ++i) erod::math::sincos(x[i], s + i, c + i);
}
void h_acos(const double* x, double* y, long n) {
  for (long i = 0; i < n; ++i) y[i] = erod::math::acos(x[i]);
}
// 0-8: sin coefficients, 9-16: cos coefficients, 17-40: asin coefficients,
// 41-46: c1, c2, c3, 2/pi, pi/2 hi, pi/2 lo
void h_tables(double* out) {
  int k = 0;
  for (double v : erod::math::kSin) out[k++] = v;
  for (double v : erod::math::kCos) out[k++] = v;
  for (double v : erod::math::kAsin.c) out[k++] = v;
  out[k++] = erod::math::kPio2C1;
  out[k++] = erod::math::kPio2C2;
  out[k++] = erod::math::kPio2C3;
  out[k++] = erod::math::kTwoOverPi;
  out[k++] = erod::math::kPio2Hi;
  out[k++] = erod::math::kPio2Lo;
}
}
