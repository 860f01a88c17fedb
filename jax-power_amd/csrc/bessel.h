// Configuration-space form of a |k| shell for the window matrix with an infinite-resolution
// theory (jaxpower/mesh2.py:1078-1092, BesselIntegral of jaxpower/utils.py:470-501):
//   xi_ell(s) = (-1)^(ell/2) / (2 pi)^3 4 pi [F_ell(k1 s) k1^3 - F_ell(k0 s) k0^3],
//   F_ell(x) = x^-3 int_0^x t^2 j_ell(t) dt,
// with the closed forms of F_ell for x > 0.1 and their Taylor series below (the split of
// jaxpower/utils.py:403-415, 453-462), in float64.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace mp {

// Sine integral Si(x), x >= 0: its power series sum (-1)^n x^(2n+1) / ((2n+1) (2n+1)!) up to
// x = 20 (the largest term, ~4e7 at x = 20, bounds the cancellation to ~5e-9), and beyond it
// Si = pi/2 - f cos x - g sin x with the asymptotic series of the auxiliary functions
// f ~ x^-1 sum (-1)^n (2n)! x^-2n and g ~ x^-2 sum (-1)^n (2n+1)! x^-2n (Abramowitz & Stegun
// 5.2.8-5.2.9, 5.2.34-5.2.35), cut before the terms start to grow (error ~1e-9 at x = 20).
__host__ __device__ inline double sine_integral(double x) {
  if (x <= 20.) {
    const double x2 = x * x;
    double term = x, sum = x;
    for (int n = 1; n < 80; ++n) {
      term *= -x2 / ((2. * n) * (2. * n + 1.));
      const double add = term / (2. * n + 1.);
      sum += add;
      if (fabs(add) <= 1e-17 * fabs(sum)) break;
    }
    return sum;
  }
  const double ix2 = 1. / (x * x);
  double f = 1., g = 1., tf = 1., tg = 1.;
  for (int n = 1; n < 40; ++n) {
    const double nf = tf * (2. * n - 1.) * (2. * n) * ix2, ng = tg * (2. * n) * (2. * n + 1.) * ix2;
    if (nf >= tf || ng >= tg) break;  // optimal truncation of the asymptotic series
    tf = nf;
    tg = ng;
    f += (n & 1) ? -tf : tf;
    g += (n & 1) ? -tg : tg;
  }
  return 1.5707963267948966 - f / x * cos(x) - g * ix2 * sin(x);
}

// F_ell(x) = x^-3 int_0^x t^2 j_ell(t) dt for ell = 0..5.
__host__ __device__ inline double tophat_bessel(int ell, double x) {
  if (x > 0.1) {
    const double s = sin(x), c = cos(x);
    const double x2 = x * x, x3 = x2 * x;
    switch (ell) {
      case 0: return (-c / x + s / x2) / x;
      case 1: return (-s - 2. * c / x) / x2 + 2. / x3;
      case 2: return (x * c - 4. * s + 3. * sine_integral(x)) / x3;
      case 3: return 8. / x3 + (x2 * s + 7. * x * c - 15. * s) / (x2 * x2);
      case 4: return (-x3 * c + 11. * x2 * s + 7.5 * x2 * sine_integral(x) + 52.5 * x * c - 52.5 * s) / (x3 * x2);
      default: return 16. / x3 + (-x2 * x2 * s - 16. * x3 * c + 105. * x2 * s + 315. * x * c - 315. * s) / (x3 * x3);
    }
  }
  const double x2 = x * x;
  switch (ell) {
    case 0: return 1. / 3. + x2 * (-1. / 30. + x2 * (1. / 840. + x2 * (-1. / 45360. + x2 * (1. / 3991680. - x2 / 518918400.))));
    case 1: return x * (1. / 12. + x2 * (-1. / 180. + x2 * (1. / 6720. + x2 * (-1. / 453600. + x2 / 47900160.))));
    case 2: return x2 * (1. / 75. + x2 * (-1. / 1470. + x2 * (1. / 68040. + x2 * (-1. / 5488560. + x2 / 674593920.))));
    case 3: return x * x2 * (1. / 630. + x2 * (-1. / 15120. + x2 * (1. / 831600. - x2 / 77837760.)));
    case 4: return x2 * x2 * (1. / 6615. + x2 * (-1. / 187110. + x2 * (1. / 11891880. - x2 / 1264863600.)));
    default: return x * x2 * x2 * (1. / 83160. + x2 * (-1. / 2702700. + x2 / 194594400.));
  }
}

// Spherical Bessel function j_ell(x), ell = 0..5: the closed forms for x > 0.1 and the Taylor
// series below, the split of jaxpower/utils.py:418-450.
__host__ __device__ inline double spherical_jn(int ell, double x) {
  if (x > 0.1) {
    const double s = sin(x), c = cos(x);
    const double x2 = x * x, x3 = x2 * x, x4 = x2 * x2;
    switch (ell) {
      case 0: return s / x;
      case 1: return -c / x + s / x2;
      case 2: return (-1. / x + 3. / x3) * s - 3. * c / x2;
      case 3: return (-6. / x2 + 15. / x4) * s + (1. / x - 15. / x3) * c;
      case 4: return (10. / x2 - 105. / x4) * c + (1. / x - 45. / x3 + 105. / (x4 * x)) * s;
      default: return (15. / x2 - 420. / x4 + 945. / (x4 * x2)) * s + (-1. / x + 105. / x3 - 945. / (x4 * x)) * c;
    }
  }
  const double x2 = x * x;
  switch (ell) {
    case 0: return 1. + x2 * (-1. / 6. + x2 * (1. / 120. + x2 * (-1. / 5040. + x2 * (1. / 362880. - x2 / 39916800.))));
    case 1: return x * (1. / 3. + x2 * (-1. / 30. + x2 * (1. / 840. + x2 * (-1. / 45360. + x2 / 3991680.))));
    case 2: return x2 * (1. / 15. + x2 * (-1. / 210. + x2 * (1. / 7560. + x2 * (-1. / 498960. + x2 / 51891840.))));
    case 3: return x * x2 * (1. / 105. + x2 * (-1. / 1890. + x2 * (1. / 83160. - x2 / 6486480.)));
    case 4: return x2 * x2 * (1. / 945. + x2 * (-1. / 20790. + x2 * (1. / 1081080. - x2 / 97297200.)));
    default: return x * x2 * x2 * (1. / 10395. + x2 * (-1. / 270270. + x2 / 16216200.));
  }
}

// xi_ell(s) of the shell [k0, k1), without the volume factor.
__host__ __device__ inline double tophat_shell_xi(int ell, double k0, double k1, double s) {
  const double sign = ((ell / 2) & 1) ? -1. : 1.;
  const double f1 = tophat_bessel(ell, k1 * s) * k1 * k1 * k1, f0 = tophat_bessel(ell, k0 * s) * k0 * k0 * k0;
  return sign * 4. * 3.14159265358979323846 / (8. * 3.14159265358979323846 * 3.14159265358979323846 * 3.14159265358979323846) * (f1 - f0);
}

}  // namespace mp
