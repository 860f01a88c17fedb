// Legendre polynomials and real spherical harmonics on the device.
//
// Real Y_lm follow the normalisation of jaxpower/utils.py:227-238:
//   sqrt((2l+1)/4pi) sqrt((l-|m|)!/(l+|m|)!) sqrt(2) P_l^|m|(cos theta) {cos m phi | sin |m| phi},
// written as polynomials of the unit vector; the value at the zero vector is the reference's
// (ylm_at_zero below).
// Only sum_m Y_lm(a) Y_lm(b) = (2l+1)/(4pi) P_l(a.b) enters the estimator, so the sign of
// each m term is a convention that cancels.
#pragma once

#include <hip/hip_runtime.h>

namespace mp {

constexpr int MAX_ELL = 4;

// P_l(mu): the closed forms of utils.py:341-360 for l <= 4 (the multipoles measured in
// practice), the three-term recurrence with tabulated 1/(n+1) above.
__host__ __device__ __forceinline__ float legendre(int ell, float mu) {
  const float m2 = mu * mu;
  switch (ell) {
    case 0: return 1.f;
    case 1: return mu;
    case 2: return 1.5f * m2 - 0.5f;
    case 3: return mu * (2.5f * m2 - 1.5f);
    case 4: return (4.375f * m2 - 3.75f) * m2 + 0.375f;
    default: break;
  }
  constexpr float inv[11] = {1.f, 0.5f, 1.f / 3.f, 0.25f, 0.2f, 1.f / 6.f, 1.f / 7.f, 0.125f, 1.f / 9.f, 0.1f, 1.f / 11.f};
  float p0 = 1.f, p1 = mu;
  for (int n = 1; n < ell; ++n) {
    float p2 = ((2 * n + 1) * mu * p1 - n * p0) * inv[n < 10 ? n : 10];
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

// P_l(mu) for a fixed list of multipoles without a branch per mode: for l <= 5,
// P_l(mu) = mu^(l mod 2) (c0 + c1 mu^2 + c2 mu^4), the closed forms above in Horner form;
// higher l take legendre().
constexpr int LEG_MAX = 6;
struct LegendreTable {
  float c0[LEG_MAX], c1[LEG_MAX], c2[LEG_MAX];
  int odd[LEG_MAX], ell[LEG_MAX];
};

inline void legendre_table_set(LegendreTable& t, int p, int ell) {
  static const float coef[6][3] = {{1.f, 0.f, 0.f},     {1.f, 0.f, 0.f},     {-0.5f, 1.5f, 0.f},
                                   {-1.5f, 2.5f, 0.f},  {0.375f, -3.75f, 4.375f}, {1.875f, -8.75f, 7.875f}};
  const int i = ell <= 5 ? ell : 0;
  t.c0[p] = coef[i][0];
  t.c1[p] = coef[i][1];
  t.c2[p] = coef[i][2];
  t.odd[p] = ell & 1;
  t.ell[p] = ell;
}

// m2 = mu * mu
__device__ __forceinline__ float legendre_eval(const LegendreTable& t, int p, float mu, float m2) {
  if (t.ell[p] > 5) return legendre(t.ell[p], mu);
  const float q = fmaf(fmaf(t.c2[p], m2, t.c1[p]), m2, t.c0[p]);
  return t.odd[p] ? mu * q : q;
}

__host__ __device__ __forceinline__ float ylm_unit(int ell, int m, float x, float y, float z) {
  switch (ell) {
    case 0:
      return 0.28209479177387814f;  // 1/(2 sqrt(pi))
    case 1: {
      const float a = 0.4886025119029199f;  // sqrt(3/(4pi))
      return m < 0 ? a * y : (m == 0 ? a * z : a * x);
    }
    case 2: {
      const float c1 = 1.0925484305920792f;  // (1/2) sqrt(15/pi)
      const float c0 = 0.31539156525252005f; // (1/4) sqrt(5/pi)
      const float c2 = 0.5462742152960396f;  // (1/4) sqrt(15/pi)
      switch (m) {
        case -2: return c1 * x * y;
        case -1: return c1 * y * z;
        case 0: return c0 * (2.f * z * z - x * x - y * y);
        case 1: return c1 * x * z;
        default: return c2 * (x * x - y * y);
      }
    }
    case 3: {
      const float c3 = 0.5900435899266435f;   // (1/4) sqrt(35/(2pi))
      const float c2 = 2.890611442640554f;    // (1/2) sqrt(105/pi)
      const float c1 = 0.4570457994644658f;   // (1/4) sqrt(21/(2pi))
      const float c0 = 0.3731763325901154f;   // (1/4) sqrt(7/pi)
      const float c2b = 1.445305721320277f;   // (1/4) sqrt(105/pi)
      float x2 = x * x, y2 = y * y, z2 = z * z;
      switch (m) {
        case -3: return c3 * y * (3.f * x2 - y2);
        case -2: return c2 * x * y * z;
        case -1: return c1 * y * (4.f * z2 - x2 - y2);
        case 0: return c0 * z * (2.f * z2 - 3.f * x2 - 3.f * y2);
        case 1: return c1 * x * (4.f * z2 - x2 - y2);
        case 2: return c2b * z * (x2 - y2);
        default: return c3 * x * (x2 - 3.f * y2);
      }
    }
    default: {  // ell == 4
      const float c4 = 2.5033429417967046f;   // (3/4) sqrt(35/pi)
      const float c3 = 1.7701307697799304f;   // (3/4) sqrt(35/(2pi))
      const float c2 = 0.9461746957575601f;   // (3/4) sqrt(5/pi)
      const float c1 = 0.6690465435572892f;   // (3/4) sqrt(5/(2pi))
      const float c0 = 0.10578554691520431f;  // (3/16) sqrt(1/pi)
      const float c2b = 0.47308734787878004f; // (3/8) sqrt(5/pi)
      const float c4b = 0.6258357354491761f;  // (3/16) sqrt(35/pi)
      float x2 = x * x, y2 = y * y, z2 = z * z, r2 = x2 + y2 + z2;
      switch (m) {
        case -4: return c4 * x * y * (x2 - y2);
        case -3: return c3 * (3.f * x2 - y2) * y * z;
        case -2: return c2 * x * y * (7.f * z2 - r2);
        case -1: return c1 * y * z * (7.f * z2 - 3.f * r2);
        case 0: return c0 * (35.f * z2 * z2 - 30.f * z2 * r2 + 3.f * r2 * r2);
        case 1: return c1 * x * z * (7.f * z2 - 3.f * r2);
        case 2: return c2b * (x2 - y2) * (7.f * z2 - r2);
        case 3: return c3 * (x2 - 3.f * y2) * x * z;
        default: return c4b * (x2 * (x2 - 3.f * y2) - y2 * (3.f * x2 - y2));
      }
    }
  }
}

// Y_lm at the zero vector as jax-power evaluates it: utils.py:252-254 maps the zero vector to
// xhat = 0 and evaluates the sympy polynomial of utils.py:299-327 there, i.e. its constant
// term, which is Y_00 itself and, for l = 4, m = 0, 3/8 sqrt(9/(4pi)) (the 3 r^4 term of
// 35 z^4 - 30 z^2 r^2 + 3 r^4 is left as the constant 3); every other l <= 4 gives 0.
__host__ __device__ __forceinline__ float ylm_at_zero(int ell, int m) {
  if (ell == 0) return 0.28209479177387814f;
  if (ell == 4 && m == 0) return 0.31735664074561293f;
  return 0.f;
}

// Y_lm of the direction of (x, y, z); ylm_at_zero at the origin.
__host__ __device__ __forceinline__ float ylm_of_vector(int ell, int m, float x, float y, float z) {
  float r2 = x * x + y * y + z * z;
  if (r2 == 0.f) return ylm_at_zero(ell, m);
  float inv = 1.f / sqrtf(r2);
  return ylm_unit(ell, m, x * inv, y * inv, z * inv);
}

// The same with 1/|v| from rsqrtf (within an ulp or two of 1/sqrtf): for the passes that
// evaluate Y_lm at every mode or cell for each m, where the correctly rounded square root
// and division were most of the per-m arithmetic.
__device__ __forceinline__ float ylm_of_vector_rsq(int ell, int m, float x, float y, float z) {
  const float r2 = x * x + y * y + z * z;
  if (r2 == 0.f) return ylm_at_zero(ell, m);
  const float inv = rsqrtf(r2);
  return ylm_unit(ell, m, x * inv, y * inv, z * inv);
}

}  // namespace mp
