// Band-integrated Planck function and its temperature derivative, for the
// device kernels and for host code of the C ABI (same arithmetic on both).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace dsa {
namespace planck {

#define HD __host__ __device__
constexpr double kPi = 3.14159265358979323846;
constexpr double kEps = 2.220446049250313e-16;   // DBL_EPSILON
constexpr double kLogDblMax = 709.782712893384;  // log(DBL_MAX)

// ---------------------------------------------------------------------------
// Band-integrated Planck function (W m^-2 sr^-1 over [lo, hi] cm^-1), the
// algorithm of src/Planck.cpp:35-169 (Simpson refinement for narrow bands,
// power series below v = 1.5, exponential series above), with the constants
// of src/Planck.hpp:10-22.  Arguments are validated on the host.
// ---------------------------------------------------------------------------
HD inline double planck_shape(double v) { return v * v * v / (exp(v) - 1.0); }

HD inline double planck_band(double lo, double hi, double t) {
  const double c2 = 1.438786, sigma = 5.67032e-8;
  if (t < 1.e-4) return 0.0;
  if (hi == lo) {
    const double arg = exp(-c2 * hi / t);
    return 1.1911e-8 * hi * hi * hi * arg / (1.0 - arg);
  }
  const double sigdpi = sigma / kPi;
  const double conc = 15.0 / (kPi * kPi * kPi * kPi);
  const double t4 = t * t * t * t;
  const double v0 = c2 * lo / t, v1 = c2 * hi / t;
  if (v0 > kEps && v1 < kLogDblMax && (hi - lo) / hi < 1.e-2) {
    const double span = v1 - v0;
    const double ends = planck_shape(v0) + planck_shape(v1);
    double prev = 0.0;
    for (int n = 1; n <= 10; ++n) {
      const double h = span / (2.0 * n);
      double acc = ends;
      for (int k = 1; k <= 2 * n - 1; ++k) acc += 2.0 * (1 + k % 2) * planck_shape(v0 + k * h);
      acc *= h * (1.0 / 3.0);
      if (fabs((acc - prev) / acc) <= 1.e-6) return sigdpi * t4 * conc * acc;
      prev = acc;
    }
    return sigdpi * t4 * conc * prev;
  }
  const double a1 = 1.0 / 3.0, a2 = -1.0 / 8.0, a3 = 1.0 / 60.0;
  const double a4 = -1.0 / 5040.0, a5 = 1.0 / 272160.0, a6 = -1.0 / 13305600.0;
  const double vcp[7] = {10.25, 5.7, 3.9, 2.9, 2.3, 1.9, 0.0};
  double p[2] = {0.0, 0.0}, d[2] = {0.0, 0.0};
  const double v[2] = {v0, v1};
  int nsmall = 0;
  for (int i = 0; i < 2; ++i) {
    if (v[i] < 1.5) {
      ++nsmall;
      const double vs = v[i] * v[i];
      p[i] = conc * vs * v[i] * (a1 + v[i] * (a2 + v[i] * (a3 + vs * (a4 + vs * (a5 + vs * a6)))));
    } else {
      int terms = 1;
      while (v[i] < vcp[terms - 1]) ++terms;
      const double ex = exp(-v[i]);
      double exm = 1.0, s = 0.0;
      for (int m = 1; m <= terms; ++m) {
        const double mv = m * v[i];
        exm *= ex;
        s += exm * (6.0 + mv * (6.0 + mv * (3.0 + mv))) / (double)(m * m * m * m);
      }
      d[i] = s * conc;
    }
  }
  double ans;
  if (nsmall == 2) ans = p[1] - p[0];
  else if (nsmall == 1) ans = 1.0 - p[0] - d[1];
  else ans = d[0] - d[1];
  return ans * (sigdpi * t4);
}

// Temperature derivative dB/dT of the band-integrated Planck function
// (src/Planck.cpp:171-214): with B = (sigma/pi)(15/pi^4) T^4 I,
// I = int_{v0}^{v1} f, f(v) = v^3 / (e^v - 1), v = c2 nu / T,
// dB/dT = 4 B / T - (sigma/pi)(15/pi^4) T^3 (v1 f(v1) - v0 f(v0));
// monochromatic (hi == lo): db/dT = b (x/T) / (1 - e^-x).
HD inline double planck_kernel(double v) {
  if (v < 1.e-6) return v * v;  // v^3 / (e^v - 1) -> v^2
  const double em = exp(-v);
  return v * v * v * em / (1.0 - em);
}

HD inline double planck_band_deriv(double lo, double hi, double t) {
  const double c2 = 1.438786, sigma = 5.67032e-8;
  if (t < 1.e-4) return 0.0;
  if (hi == lo) {
    const double x = c2 * hi / t, em = exp(-x);
    const double b = 1.1911e-8 * hi * hi * hi * em / (1.0 - em);
    return b * (x / t) / (1.0 - em);
  }
  const double sigdpi = sigma / kPi;
  const double conc = 15.0 / pow(kPi, 4.0);
  const double v0 = c2 * lo / t, v1 = c2 * hi / t;
  const double b = planck_band(lo, hi, t);
  return 4.0 * b / t - sigdpi * conc * t * t * t * (v1 * planck_kernel(v1) - v0 * planck_kernel(v0));
}
#undef HD

}  // namespace planck
using planck::planck_band;
using planck::planck_band_deriv;
}  // namespace dsa
