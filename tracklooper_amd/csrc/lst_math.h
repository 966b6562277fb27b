// Device/host float math for the LST engine.
//
// Parity policy: float transcendentals are evaluated in double precision and
// rounded once to float, so the result is the correctly rounded float value
// except with probability ~1e-9 per call (when the exact value lies within one
// double ulp of a float rounding boundary).  sqrt and the four basic
// operations are IEEE correctly rounded; the library is compiled with
// -ffp-contract=off so no FMA is formed behind the source's back.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

#define LST_HD __host__ __device__ __forceinline__

namespace lst {

constexpr float kPi = 3.14159265358979323846f;  // float(M_PI)

LST_HD float m_sin(float x) { return (float)::sin((double)x); }
LST_HD float m_cos(float x) { return (float)::cos((double)x); }
LST_HD float m_tan(float x) { return (float)::tan((double)x); }
LST_HD float m_asin(float x) { return (float)::asin((double)x); }
LST_HD float m_acos(float x) { return (float)::acos((double)x); }
LST_HD float m_atan(float x) { return (float)::atan((double)x); }
LST_HD float m_atan2(float y, float x) { return (float)::atan2((double)y, (double)x); }
LST_HD float m_acosh(float x) { return (float)::acosh((double)x); }
LST_HD float m_sinh(float x) { return (float)::sinh((double)x); }
LST_HD float m_exp(float x) { return (float)::exp((double)x); }
LST_HD float m_log(float x) { return (float)::log((double)x); }
LST_HD float m_sqrt(float x) { return ::sqrtf(x); }
LST_HD double m_dsqrt(double x) { return ::sqrt(x); }
LST_HD float m_abs(float x) { return ::fabsf(x); }
LST_HD float m_min(float a, float b) { return ::fminf(a, b); }
LST_HD float m_max(float a, float b) { return ::fmaxf(a, b); }

// phi in (-pi, pi]: SDL/Hit.h:125-132
LST_HD float phi_mpi_pi(float x) {
  if (::fabsf(x) <= kPi) return x;
  const float o2pi = 1.f / (2.f * kPi);
  float n = ::roundf(x * o2pi);
  return x - n * float(2.f * kPi);
}
// SDL/Hit.h:135-137
LST_HD float phi_of(float x, float y) { return phi_mpi_pi(kPi + m_atan2(-y, -x)); }
// SDL/Hit.h:140-149
LST_HD float delta_phi(float x1, float y1, float x2, float y2) {
  return phi_mpi_pi(phi_of(x2, y2) - phi_of(x1, y1));
}
LST_HD float delta_phi_change(float x1, float y1, float x2, float y2) { return delta_phi(x1, y1, x2 - x1, y2 - y1); }
// SDL/Hit.h:152-156: sign(a) * sign(b) * a with sign(0) = +1
LST_HD float sign_copy(float a, float b) {
  int sa = (a < 0) ? -1 : 1;
  int sb = (b < 0) ? -1 : 1;
  return sa * sb * a;
}
// SDL/Hit.h:158-170
LST_HD float wrap_dphi(float p1, float p2) {
  float d = p1 - p2;
  if (d > kPi) d -= 2 * kPi;
  else if (d < -kPi) d += 2 * kPi;
  return d;
}

}  // namespace lst
