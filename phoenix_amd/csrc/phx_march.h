// Device functions of the polarized GRRT march (gfx950), templated on the
// scalar type T (float or double).  The physics is that of the reference's
// src/core/ (cited per function); the layout is MI355X-first: per-lane state in
// VGPRs, the sparse Kerr-Schild metric (7 non-zero covariant components), one
// AMR lookup per geodesic step, a flattened finest-level leaf map, 32-byte cell
// records gathered with 16-byte loads, and a fully unrolled Gauss-Jordan
// inverse whose partial pivoting is done with selects (no scratch memory).
#pragma once
#include <type_traits>

#include "phx_target.h"
#include "phx_bessel_table.h"

namespace phxd {

// a copy of a struct in global memory, loaded dword by dword through the global view
typedef unsigned __attribute__((may_alias)) dword_alias;
template <class S>
PHX_DEV S load_global(const S* p) {
  static_assert(sizeof(S) % 4 == 0, "whole dwords");
  const gptr<unsigned> w = (gptr<unsigned>)p;
  S out;
  dword_alias* d = reinterpret_cast<dword_alias*>(&out);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(S) / 4); i++) d[i] = w[i];
  return out;
}

// ------------------------------------------------------------------ constants
// CGS constants of the reference (src/core/numerics/constants.hpp:22-81)
constexpr double MP = 1.6726219e-24, ME = 9.1093829e-28, QE = 4.80320425e-10;
constexpr double CL = 2.99792458e10, HP = 6.62606885e-27;
constexpr double H_OVER_MEC2 = HP / (ME * CL * CL);
constexpr double NU_CRIT = 3. * QE / (4. * M_PI * ME * CL);
constexpr double JNU = QE * QE / (2. * 1.73205080756887729353 * CL);
constexpr double WP2 = 4. * M_PI * QE * QE / ME;
constexpr double GYRO = QE / (ME * CL);
constexpr double PI_D = 3.14159265358979323846;

template <class T> struct Lim;
template <> struct Lim<float> {  // src/core/numerics/precision.hpp:46-51
  static constexpr float tiny = 1e-30f, big = 1e30f, tol_series = 1e-9f, tol_asym = 1e-8f;
  static constexpr bool fp32 = true;
};
template <> struct Lim<double> {  // precision.hpp:53-58
  static constexpr double tiny = 1e-160, big = 1e100, tol_series = 1e-18, tol_asym = 1e-17;
  static constexpr bool fp32 = false;
};

// Divisions the optimizer rewrites (a negation folded into the divisor) or speculates out of a
// branch lose their fast-math marking, and an FP32 division then expands to the correctly rounded
// sequence (a dozen instructions and two mode switches).  These helpers state the FP32 form
// directly; FP64 divides.
// x / c for a constant c: FP32 multiplies by the rounded reciprocal (within an ulp of the quotient)
template <class T>
PHX_DEV T div_const(T x, double c) {
  if constexpr (Lim<T>::fp32)
    return x * float(1. / c);
  else
    return x / T(c);
}
// a / b: FP32 by the hardware reciprocal (1 ulp) and a multiply that is not fused with what
// follows.  For |b| < 2^96 this is the value of the compiler's own fast FP32 division (which adds a
// range scaling of b: four more instructions); every denominator it is used for stays far below.
PHX_DEV float div_rcp(float a, float b) {
#pragma clang fp contract(off)
  return a * hw_rcp(b);
}
PHX_DEV double div_rcp(double a, double b) { return a / b; }
// 1 / x and a / b to about half an ulp in FP32: the hardware reciprocal (1 ulp) refined by one
// Newton step, and the quotient refined by its residual.  The geodesic's right-hand side and step
// rule use these: with the library's fast FP32 division (about 2.5 ulp) the errors of the
// connection's reciprocals reach the ray's trajectory, and photon-ring pixels of the FP32 march
// then leave the per-pixel gate several times as often as the reference's FP32 build (whose
// divisions are correctly rounded).  FP64 divides.
PHX_DEV float recip_nr(float x) {
  const float r = hw_rcp(x);
  return fmaf(fmaf(-x, r, 1.f), r, r);
}
PHX_DEV double recip_nr(double x) { return 1. / x; }
PHX_DEV float div_nr(float a, float b) {
  const float r = recip_nr(b), q = a * r;
  return fmaf(fmaf(-b, q, a), r, q);
}
PHX_DEV double div_nr(double a, double b) { return a / b; }

PHX_DEV void scs(float x, float* s, float* c) { sincosf(x, s, c); }
PHX_DEV void scs(double x, double* s, double* c) { sincos(x, s, c); }

// sin and cos of a float |x| < ~100: Cody-Waite reduction by pi/2 (two-part constant,
// fma), Taylor polynomials on [-pi/4, pi/4] (sin to x^9, cos to x^10: truncation below
// 2e-9), quadrant by the low bits of the quotient.
PHX_DEV void sincos_small(float x, float* s, float* c) {
  const float q = rintf(x * 0.63661977236758134f);
  const int qi = (int)q;
  float r = fmaf(-q, 1.57079637050628662109375f, x);
  r = fmaf(-q, -4.37113900018624283e-8f, r);
  const float z = r * r;
  float ps = fmaf(z, 2.7557319223985893e-6f, -1.9841269841269841e-4f);
  ps = fmaf(ps, z, 8.3333333333333333e-3f);
  ps = fmaf(ps, z, -1.6666666666666667e-1f);
  ps = fmaf(ps * z, r, r);
  float pc = fmaf(z, -2.7557319223985888e-7f, 2.4801587301587302e-5f);
  pc = fmaf(pc, z, -1.3888888888888889e-3f);
  pc = fmaf(pc, z, 4.1666666666666667e-2f);
  pc = fmaf(pc, z, -0.5f);
  pc = fmaf(pc, z, 1.0f);
  float sn = (qi & 1) ? pc : ps, cs = (qi & 1) ? ps : pc;
  if (qi & 2) sn = -sn;
  if ((qi + 1) & 2) cs = -cs;
  *s = sn;
  *c = cs;
}

// fmod(x, y) for y > 0 and |x| < 2^22 y: the quotient estimated with the reciprocal, the
// remainder formed by one fma and corrected once when the estimate is off by one.  The true
// remainder is representable and the fma rounds it exactly, so this is the library's fmod.
template <class T>
PHX_DEV T fmod_exact(T x, T y, T inv_y) {
  const T n = trunc(x * inv_y);
  T r = fma(-n, y, x);
  if (x >= T(0)) {
    if (r < T(0))
      r = fma(-(n - T(1)), y, x);
    else if (r >= y)
      r = fma(-(n + T(1)), y, x);
  } else {
    if (r > T(0))
      r = fma(-(n + T(1)), y, x);
    else if (r <= -y)
      r = fma(-(n - T(1)), y, x);
  }
  return r;
}

// the library exp for r = exp(x1) (a reduction around the hardware exp2 was measured: no gain
// at this instruction mix, and ~2 ulp instead of ~1)
template <class T>
PHX_DEV T exp_r(T x) {
  return exp(x);
}

// sin/cos of 2 x2 and of theta = x2 + (h/2) sin 2x2 (the MKS-BHAC stretching).  FP32: one
// sincos of x2, the double-angle identities, and theta = x2 + delta by angle addition with
// |delta| <= h/2 <= 0.5 expanded in Taylor series (to delta^9 / delta^10, error < 1e-9);
// FP64: the library sincos of both angles.
PHX_DEV void theta_trig(float x2, float hs, float* s2x, float* c2x, float* s, float* c) {
  float sx, cx;
  sincos_small(x2, &sx, &cx);
  *s2x = 2.0f * sx * cx;
  *c2x = (cx - sx) * (cx + sx);
  const float d = 0.5f * hs * *s2x, z = d * d;
  float sd = fmaf(z, 2.7557319223985893e-6f, -1.9841269841269841e-4f);
  sd = fmaf(sd, z, 8.3333333333333333e-3f);
  sd = fmaf(sd, z, -1.6666666666666667e-1f);
  sd = fmaf(sd * z, d, d);
  float cd = fmaf(z, -2.7557319223985888e-7f, 2.4801587301587302e-5f);
  cd = fmaf(cd, z, -1.3888888888888889e-3f);
  cd = fmaf(cd, z, 4.1666666666666667e-2f);
  cd = fmaf(cd, z, -0.5f);
  cd = fmaf(cd, z, 1.0f);
  *s = sx * cd + cx * sd;
  *c = cx * cd - sx * sd;
}
PHX_DEV void theta_trig(double x2, double hs, double* s2x, double* c2x, double* s, double* c) {
  sincos(2. * x2, s2x, c2x);
  sincos(x2 + 0.5 * hs * *s2x, s, c);
}
// Taylor coefficients of sin r / r and cos r in z = r^2 (Horner order): sin to r^19, cos to r^20.
// Kept in constant memory and loaded where a polynomial is evaluated (taylor_coef): as FP64
// instruction operands they would otherwise be materialized in scalar registers and held across
// the whole march loop, which spills the march's other uniform values.
PHX_CONSTANT double kTaylorSin[9] = {-1. / 121645100408832000., 1. / 355687428096000., -1. / 1307674368000.,
                                     1. / 6227020800.,          -1. / 39916800.,       1. / 362880.,
                                     -1. / 5040.,               1. / 120.,             -1. / 6.};
PHX_CONSTANT double kTaylorCos[10] = {1. / 2432902008176640000., -1. / 6402373705728000., 1. / 20922789888000.,
                                      -1. / 87178291200.,        1. / 479001600.,         -1. / 3628800.,
                                      1. / 40320.,               -1. / 720.,              1. / 24.,
                                      -0.5};
// Taylor coefficients of (e^x - 1 - x) / x^2 in x (Horner order): 1/10! ... 1/2
PHX_CONSTANT double kTaylorExp[9] = {1. / 3628800., 1. / 362880., 1. / 40320., 1. / 5040., 1. / 720.,
                                    1. / 120.,     1. / 24.,     1. / 6.,     0.5};
struct TaylorSC {
  double s[9], c[10], e[9];
};
// the coefficients, loaded (scalar loads) at the call through an opaque copy of the table address,
// so that the loads are not hoisted out of an enclosing loop
PHX_DEV TaylorSC taylor_coef() {
  const double* ps = opaque_scalar<const double*>(kTaylorSin);
  const double* pc = opaque_scalar<const double*>(kTaylorCos);
  const double* pe = opaque_scalar<const double*>(kTaylorExp);
  TaylorSC t;
#pragma unroll
  for (int i = 0; i < 9; i++) t.s[i] = load_constant(ps, i);
#pragma unroll
  for (int i = 0; i < 10; i++) t.c[i] = load_constant(pc, i);
#pragma unroll
  for (int i = 0; i < 9; i++) t.e[i] = load_constant(pe, i);
  return t;
}
// sin(r) and cos(r) by the Taylor polynomials above (|r| <= pi/4: truncation below 1e-19)
PHX_DEV void taylor_sc(const TaylorSC& t, double r, double* sr, double* cr) {
  const double z = r * r;
  double ps = t.s[0];
#pragma unroll
  for (int i = 1; i < 9; i++) ps = fma(ps, z, t.s[i]);
  *sr = fma(ps * z, r, r);
  double pc = t.c[0];
#pragma unroll
  for (int i = 1; i < 10; i++) pc = fma(pc, z, t.c[i]);
  *cr = fma(pc, z, 1.);
}
// FP64 sin and cos of a small |x| (< ~1e3): Cody-Waite reduction by pi/2 with a two-part
// constant (fma), the Taylor polynomials on [-pi/4, pi/4], quadrant from the low bits of the
// quotient.  Compact, so that the FP64 steps of the mixed march keep few registers (the library's
// sincos carries a large-argument path).
PHX_DEV void sincos_small(const TaylorSC& t, double x, double* s, double* c) {
  const double q = rint(x * 0.63661977236758134308);
  const int qi = (int)q;
  double r = fma(-q, 1.5707963267948966192, x);
  r = fma(-q, 6.123233995736766036e-17, r);
  double sr, cr;
  taylor_sc(t, r, &sr, &cr);
  double sn = (qi & 1) ? cr : sr, cs = (qi & 1) ? sr : cr;
  if (qi & 2) sn = -sn;
  if ((qi + 1) & 2) cs = -cs;
  *s = sn;
  *c = cs;
}
PHX_DEV void sincos_small(double x, double* s, double* c) { sincos_small(taylor_coef(), x, s, c); }

// FP64 with one compact sincos: sin/cos of x2, the double-angle identities, and theta = x2 +
// delta by angle addition with sin/cos of |delta| <= h/2 <= 0.5 by the same Taylor polynomials
// (truncation below 2e-23).  Used by the FP64 steps of the mixed march.
PHX_DEV void theta_trig_1(double x2, double hs, double* s2x, double* c2x, double* s, double* c) {
  const TaylorSC t = taylor_coef();
  double sx, cx;
  sincos_small(t, x2, &sx, &cx);
  *s2x = 2. * sx * cx;
  *c2x = (cx - sx) * (cx + sx);
  const double d = 0.5 * hs * *s2x;
  double sd, cd;
  taylor_sc(t, d, &sd, &cd);
  *s = sx * cd + cx * sd;
  *c = cx * cd - sx * sd;
}
// The point quantities of an FP64 RK4 stage: r, sin/cos x2, sin/cos 2x2, sin/cos theta
struct Pt64 {
  double r, sx, cx, s2x, c2x, s, c;
};
PHX_DEV Pt64 pt64_at(const TaylorSC& t, double x1, double x2, double hs) {
  Pt64 p;
  p.r = exp(x1);
  sincos_small(t, x2, &p.sx, &p.cx);
  p.s2x = 2. * p.sx * p.cx;
  p.c2x = (p.cx - p.sx) * (p.cx + p.sx);
  double sd, cd;
  taylor_sc(t, 0.5 * hs * p.s2x, &sd, &cd);
  p.s = p.sx * cd + p.cx * sd;
  p.c = p.cx * cd - p.sx * sd;
  return p;
}
// sin and cos of a small |e| <= 0.03 to FP64 rounding: the tails of the Taylor polynomials above
// (sin to e^9, cos to e^10: truncation below 5e-25)
PHX_DEV void sincos_tiny(const TaylorSC& t, double e, double* se, double* ce) {
  const double z = e * e;
  double ps = t.s[5];
#pragma unroll
  for (int i = 6; i < 9; i++) ps = fma(ps, z, t.s[i]);
  *se = fma(ps * z, e, e);
  double pc = t.c[5];
#pragma unroll
  for (int i = 6; i < 10; i++) pc = fma(pc, z, t.c[i]);
  *ce = fma(pc, z, 1.);
}
// The FP64 point quantities at (x1 + e1, x2 + e2) from those at (x1, x2) for |e1|, |e2| <= 0.02 (an
// RK4 stage of one step): e^{e1} by its Taylor polynomial to e1^10 (truncation below 1e-24), the
// angles by addition with sincos_tiny; theta's offset is e2 + (h/2)(sin 2(x2 + e2) - sin 2x2).
// The offsets are applied to the base point exactly, as the FP32 stages do (pt32_offset).
PHX_DEV Pt64 pt64_offset(const TaylorSC& t, const Pt64& b, double e1, double e2, double hs) {
  Pt64 p;
  double pe = t.e[0];
#pragma unroll
  for (int i = 1; i < 9; i++) pe = fma(pe, e1, t.e[i]);
  p.r = fma(b.r, fma(pe * e1, e1, e1), b.r);
  double se, ce;
  sincos_tiny(t, e2, &se, &ce);
  p.sx = b.sx * ce + b.cx * se;
  p.cx = b.cx * ce - b.sx * se;
  p.s2x = 2. * p.sx * p.cx;
  p.c2x = (p.cx - p.sx) * (p.cx + p.sx);
  double st, ct;
  sincos_tiny(t, fma(0.5 * hs, p.s2x - b.s2x, e2), &st, &ct);
  p.s = b.s * ct + b.c * st;
  p.c = b.c * ct - b.s * st;
  return p;
}
template <bool ONE_SINCOS>
PHX_DEV void theta_trig_sel(double x2, double hs, double* s2x, double* c2x, double* s, double* c) {
  if (ONE_SINCOS)
    theta_trig_1(x2, hs, s2x, c2x, s, c);
  else
    theta_trig(x2, hs, s2x, c2x, s, c);
}
template <bool ONE_SINCOS>
PHX_DEV void theta_trig_sel(float x2, float hs, float* s2x, float* c2x, float* s, float* c) {
  theta_trig(x2, hs, s2x, c2x, s, c);
}

// ------------------------------------------------------------------ metric
// Kerr-Schild in MKS-BHAC coordinates (t, ln r, x2, phi), theta = x2 + h/2 sin 2x2
// (src/core/geometry/metric.hpp:26-142).  Only the 7 non-zero covariant and 6
// non-zero contravariant components are kept.
template <class T>
struct Met {
  T d00, d01, d03, d11, d13, d22, d33;  // g_mu nu
  T r, s, c, hf;                        // r, sin/cos theta, d theta / d x2
};
template <class T>
PHX_DEV Met<T> metric(T a, T hs, T x1, T x2) {
  Met<T> g;
  const T r = exp_r(x1);
  T s2x, c2x, s, c;
  theta_trig(x2, hs, &s2x, &c2x, &s, &c);
  const T s2 = s * s, rho2 = r * r + a * a * c * c, hf = T(1) + hs * c2x;
  const T z = T(2) * r / rho2;
  g.d00 = -T(1) + z;
  g.d01 = z * r;
  g.d03 = -T(2) * a * r * s2 / rho2;
  g.d11 = (T(1) + z) * r * r;
  g.d13 = -a * s2 * (T(1) + z) * r;
  g.d22 = rho2 * hf * hf;
  g.d33 = s2 * (rho2 + a * a * s2 * (T(1) + z));
  g.r = r;
  g.s = s;
  g.c = c;
  g.hf = hf;
  return g;
}
template <class T>
struct MetU {  // contravariant part (metric.hpp:287-292)
  T u00, u01, u11, u13, u22, u33;
};
template <class T>
PHX_DEV MetU<T> metric_up(T a, const Met<T>& g) {
  MetU<T> u;
  const T r = g.r, rho2 = r * r + a * a * g.c * g.c, ir = T(1) / rho2;
  u.u00 = -T(1) - T(2) * r / rho2;
  u.u01 = T(2) * ir;
  u.u11 = ir * (r * (r - T(2)) + a * a) / (r * r);
  u.u13 = a * ir / r;
  u.u22 = ir / (g.hf * g.hf);
  u.u33 = ir / (g.s * g.s);
  return u;
}
template <class T>
PHX_DEV T dot(const Met<T>& g, const T A[4], const T B[4]) {
  return g.d00 * A[0] * B[0] + g.d01 * (A[0] * B[1] + A[1] * B[0]) + g.d03 * (A[0] * B[3] + A[3] * B[0]) +
         g.d11 * A[1] * B[1] + g.d13 * (A[1] * B[3] + A[3] * B[1]) + g.d22 * A[2] * B[2] + g.d33 * A[3] * B[3];
}
template <class T>
PHX_DEV void lower(const Met<T>& g, const T V[4], T out[4]) {
  out[0] = g.d00 * V[0] + g.d01 * V[1] + g.d03 * V[3];
  out[1] = g.d01 * V[0] + g.d11 * V[1] + g.d13 * V[3];
  out[2] = g.d22 * V[2];
  out[3] = g.d03 * V[0] + g.d13 * V[1] + g.d33 * V[3];
}

// ------------------------------------------------------------------ geodesic
// The connection of the MKS-BHAC Kerr-Schild metric, from its derivatives.  The metric depends on
// x1 = ln r and x2 only, so with A = d g / d x1 and B = d g / d x2 (both with the sparsity of g)
//   Gamma_nu(U, V) = 1/2 [U^1 (A V)_nu + U^2 (B V)_nu + V^1 (A U)_nu + V^2 (B U)_nu]
//                    - 1/2 delta_nu^1 U.A.V - 1/2 delta_nu^2 U.B.V
// (Gamma_{nu ab} U^a V^b with Gamma_{nu ab} = (g_{nu a, b} + g_{nu b, a} - g_{ab, nu}) / 2), and the
// geodesic / parallel-transport right-hand side is -g^{mu nu} Gamma_nu.  With r = e^{x1},
// theta = x2 + h/2 sin 2x2, theta' = 1 + h cos 2x2, theta'' = -2 h sin 2x2, s = sin theta,
// c = cos theta, Sigma = r^2 + a^2 c^2 and z = 2 r / Sigma, the covariant metric is
//   g00 = z - 1, g01 = z r, g03 = -a z s^2, g11 = (1 + z) r^2, g13 = -a (1 + z) r s^2,
//   g22 = Sigma theta'^2, g33 = s^2 Sigma + a^2 s^4 (1 + z),
// and, writing z1 = dz/dx1 = z (a^2 c^2 - r^2) / Sigma (so that z + z1 = 2 z a^2 c^2 / Sigma exactly,
// kept in that form: the two terms cancel at large r), z2 = dz/dx2 = 2 a^2 s c theta' z / Sigma,
// S2 = dSigma/dx2 = -2 a^2 s c theta' and q2 = d(s^2)/dx2 = 2 s c theta':
//   A00 = z1, A01 = r (z + z1), A03 = -a s^2 z1, A11 = r^2 (2 + z + (z + z1)), A13 = -a r s^2 (1 + (z + z1)),
//   A22 = 2 r^2 theta'^2, A33 = s^2 (2 r^2 + a^2 s^2 z1);
//   B00 = z2, B01 = r z2, B03 = -a (s^2 z2 + z q2), B11 = r^2 z2, B13 = -a r (q2 (1 + z) + s^2 z2),
//   B22 = theta' (S2 theta' + 2 Sigma theta''), B33 = q2 (Sigma + 2 a^2 s^2 (1 + z)) + s^2 (S2 + a^2 s^2 z2).
// The contravariant metric: g^00 = -1 - z, g^01 = 2 / Sigma, g^11 = (r^2 - 2 r + a^2) / (Sigma r^2),
// g^13 = a / (Sigma r), g^22 = 1 / (Sigma theta'^2), g^33 = 1 / (Sigma s^2) (metric.hpp:64-91).
// This is the same connection as the reference's closed form (geodesic.hpp:81-169), written from
// the metric's derivatives; tests/test_cpu_lane.py holds the march to the reference's FP64 build.
template <class T>
struct SymD {  // a symmetric 4x4 with the sparsity of the Kerr-Schild metric
  T m00, m01, m03, m11, m13, m22, m33;
};
template <class T>
PHX_DEV void symd_vec(const SymD<T>& D, const T v[4], T out[4]) {
  out[0] = D.m00 * v[0] + D.m01 * v[1] + D.m03 * v[3];
  out[1] = D.m01 * v[0] + D.m11 * v[1] + D.m13 * v[3];
  out[2] = D.m22 * v[2];
  out[3] = D.m03 * v[0] + D.m13 * v[1] + D.m33 * v[3];
}
template <class T>
struct Geo {
  SymD<T> A, B;                          // d g / d x1, d g / d x2
  T u00, u01, u11, u13, u22, u33;        // g^{mu nu}
};
template <class T, bool ONE_SINCOS = false>
PHX_DEV Geo<T> geo_at(T a, T hs, T x1, T x2) {
  Geo<T> G;
  const T r = exp_r(x1), r2 = r * r, a2 = a * a;
  T s2x, c2x, s, c;
  theta_trig_sel<ONE_SINCOS>(x2, hs, &s2x, &c2x, &s, &c);
  const T tp = T(1) + hs * c2x, tpp = T(-2) * hs * s2x;
  const T ss = s * s, sc = s * c, a2c2 = a2 * c * c;
  const T Sig = r2 + a2c2, iS = recip_nr(Sig);
  const T z = T(2) * r * iS;
  const T z1 = z * (a2c2 - r2) * iS, zz1 = T(2) * z * a2c2 * iS;  // dz/dx1 and z + dz/dx1
  const T S2 = T(-2) * a2 * sc * tp, q2 = T(2) * sc * tp;
  const T z2 = -z * S2 * iS;
  const T opz = T(1) + z;
  G.A.m00 = z1;
  G.A.m01 = r * zz1;
  G.A.m03 = -a * ss * z1;
  G.A.m11 = r2 * (T(2) + z + zz1);
  G.A.m13 = -a * r * ss * (T(1) + zz1);
  G.A.m22 = T(2) * r2 * tp * tp;
  G.A.m33 = ss * (T(2) * r2 + a2 * ss * z1);
  G.B.m00 = z2;
  G.B.m01 = r * z2;
  G.B.m03 = -a * (ss * z2 + z * q2);
  G.B.m11 = r2 * z2;
  G.B.m13 = -a * r * (q2 * opz + ss * z2);
  G.B.m22 = tp * (S2 * tp + T(2) * Sig * tpp);
  G.B.m33 = q2 * (Sig + T(2) * a2 * ss * opz) + ss * (S2 + a2 * ss * z2);
  const T ir = recip_nr(r), itp = recip_nr(tp), is = recip_nr(s);
  G.u00 = -opz;
  G.u01 = T(2) * iS;
  G.u11 = iS * (r2 - T(2) * r + a2) * ir * ir;
  G.u13 = a * iS * ir;
  G.u22 = iS * itp * itp;
  G.u33 = iS * is * is;
  return G;
}
// -g^{mu nu} Gamma_nu for the covariant connection contraction w (Gamma_nu above)
template <class T>
PHX_DEV void raise_neg(const Geo<T>& G, const T w[4], T out[4]) {
  out[0] = -(G.u00 * w[0] + G.u01 * w[1]);
  out[1] = -(G.u01 * w[0] + G.u11 * w[1] + G.u13 * w[3]);
  out[2] = -(G.u22 * w[2]);
  out[3] = -(G.u13 * w[1] + G.u33 * w[3]);
}
// the geodesic acceleration -Gamma^mu(k, k) and the parallel transport -Gamma^mu(k, f) at one point
template <class T>
PHX_DEV void connection(const Geo<T>& G, const T k[4], const T f[4], T acc_k[4], T acc_f[4]) {
  T Ak[4], Bk[4], Af[4], Bf[4];
  symd_vec(G.A, k, Ak);
  symd_vec(G.B, k, Bk);
  symd_vec(G.A, f, Af);
  symd_vec(G.B, f, Bf);
  const T kAk = k[0] * Ak[0] + k[1] * Ak[1] + k[2] * Ak[2] + k[3] * Ak[3];
  const T kBk = k[0] * Bk[0] + k[1] * Bk[1] + k[2] * Bk[2] + k[3] * Bk[3];
  const T kAf = k[0] * Af[0] + k[1] * Af[1] + k[2] * Af[2] + k[3] * Af[3];
  const T kBf = k[0] * Bf[0] + k[1] * Bf[1] + k[2] * Bf[2] + k[3] * Bf[3];
  T wk[4], wf[4];
  const T h = T(0.5);
#pragma unroll
  for (int n = 0; n < 4; n++) {
    wk[n] = k[1] * Ak[n] + k[2] * Bk[n];
    wf[n] = h * (k[1] * Af[n] + k[2] * Bf[n] + f[1] * Ak[n] + f[2] * Bk[n]);
  }
  wk[1] -= h * kAk;
  wk[2] -= h * kBk;
  wf[1] -= h * kAf;
  wf[2] -= h * kBf;
  raise_neg(G, wk, acc_k);
  raise_neg(G, wf, acc_f);
}

// Ray state: (hi, lo) pairs of x, k, F1; lo is only carried in FP32 (geodesic.hpp:11-14)
template <class T>
struct Ray {
  T hi[12], lo[12];
};
template <class T>
PHX_DEV void two_sum_add(T& hi, T& lo, T d) {  // TwoSum + Fast2Sum (precision.hpp:65-73)
  const T s = hi + d;
  const T bb = s - hi;
  const T err = (hi - (s - bb)) + (d - bb);
  lo += err;
  const T s2 = s + lo;
  lo = lo - (s2 - s);
  hi = s2;
}
// RK4 of x, k (geodesic) and F1 (parallel transport) with the FP32 south-hemisphere
// fold (geodesic.hpp:222-256), the right-hand side from connection().  The weighted stage sum k1 + 2 k2 + 2 k3 + k4 is
// accumulated as the stages are formed, so only it and the next stage's offset are kept.
template <class T, bool ONE_SINCOS = false>
PHX_DEV void rk4(T a, T hs, Ray<T>& g, T dt) {
  constexpr bool F32 = Lim<T>::fp32;
  const T pi_hi = T(PI_D), pi_lo = T(PI_D - (double)T(PI_D));
  const bool south = F32 && g.hi[2] > T(0.5 * PI_D);
  T sum[12], inc[12];
#pragma unroll
  for (int i = 0; i < 12; i++) inc[i] = T(0);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    T y[12];
#pragma unroll
    for (int i = 0; i < 12; i++) y[i] = F32 ? g.hi[i] + (g.lo[i] + inc[i]) : g.hi[i] + inc[i];
    if (south) {
      y[2] = (pi_hi - g.hi[2]) + ((pi_lo - g.lo[2]) - inc[2]);
      y[6] = -y[6];
      y[10] = -y[10];
    }
    T d[12];
    const Geo<T> G = geo_at<T, ONE_SINCOS>(a, hs, y[1], y[2]);
#pragma unroll
    for (int i = 0; i < 4; i++) d[i] = y[4 + i];
    connection(G, y + 4, y + 8, d + 4, d + 8);
    if (south) {
      d[2] = -d[2];
      d[6] = -d[6];
      d[10] = -d[10];
    }
    const T wsum = (q == 0 || q == 3) ? T(1) : T(2);
#pragma unroll
    for (int i = 0; i < 12; i++) sum[i] = q == 0 ? d[i] : sum[i] + wsum * d[i];
    if (q == 3) break;
    const T wdt = (q == 2 ? T(1.0) : T(0.5)) * dt;
#pragma unroll
    for (int i = 0; i < 12; i++) inc[i] = wdt * d[i];
  }
  const T dt6 = div_const(dt, 6.);
#pragma unroll
  for (int i = 0; i < 12; i++) {
    const T incr = sum[i] * dt6;
    if (F32)
      two_sum_add(g.hi[i], g.lo[i], incr);
    else
      g.hi[i] += incr;
  }
}
// ------------------------------------------------------------------ Hamiltonian form
// The same null geodesic as Hamilton's equations of H = g^{mu nu} p_mu p_nu / 2, with the
// contravariant MKS-BHAC Kerr-Schild metric of src/core/geometry/metric.hpp:64-91.  H depends
// on neither t nor phi, so p_t and p_phi are constants of the motion and the state reduces to
// (x1, x2, x3, p_1, p_2).  With Sigma = r^2 + a^2 cos^2 theta and theta' = d theta / d x2,
//   2H = -p_t^2 + Q / Sigma,
//   Q  = -2 r p_t^2 + 4 p_t p_1 + (Delta / r^2) p_1^2 + (2 a / r) p_1 p_phi + p_2^2 / theta'^2
//        + p_phi^2 / sin^2 theta,
// so only d/dr and d/dtheta of Q and Sigma enter the force.  d[0..2] = k^{1,2,3} = dH/dp_i,
// d[3..4] = dp_{1,2}/dlambda = -dH/dx^{1,2} (x1 = ln r: d/dx1 = r d/dr; the x2 derivative adds
// the explicit theta'(x2) of the p_2 term), *k0 = k^t = g^{tt} p_t + g^{t1} p_1.
// RECIP: S, 1/r, 1/sin theta and 1/theta' from one reciprocal of their product (FP64, where a
// division is a long sequence); otherwise four reciprocals (FP32, hardware rcp).
// ham_core: the same from r, sin/cos 2x2 and sin/cos theta at the point.
template <class T, bool RECIP = false>
PHX_DEV void ham_core(T a, T hs, T pt, T pp, T r, T s2x, T c2x, T s, T c, T p1, T p2, T d[5], T* k0);
template <class T, bool ONE_SINCOS = false, bool RECIP = false>
PHX_DEV void ham_rhs(T a, T hs, T pt, T pp, const T y[5], T d[5], T* k0) {
  const T r = exp_r(y[0]);
  T s2x, c2x, s, c;
  theta_trig_sel<ONE_SINCOS>(y[1], hs, &s2x, &c2x, &s, &c);
  ham_core<T, RECIP>(a, hs, pt, pp, r, s2x, c2x, s, c, y[3], y[4], d, k0);
}
// The point quantities an FP32 RK4 stage needs: r, sin/cos x2, sin/cos 2x2, sin/cos theta.
struct Pt32 {
  float r, sx, cx, s2x, c2x, s, c;
};
// r = e^{x1} already known (the pass's view computes it)
PHX_DEV Pt32 pt32_at_r(float r, float x2, float hs) {
  Pt32 t;
  t.r = r;
  sincos_small(x2, &t.sx, &t.cx);
  t.s2x = 2.0f * t.sx * t.cx;
  t.c2x = (t.cx - t.sx) * (t.cx + t.sx);
  const float d = 0.5f * hs * t.s2x, z = d * d;
  float sd = fmaf(z, 2.7557319223985893e-6f, -1.9841269841269841e-4f);
  sd = fmaf(sd, z, 8.3333333333333333e-3f);
  sd = fmaf(sd, z, -1.6666666666666667e-1f);
  sd = fmaf(sd * z, d, d);
  float cd = fmaf(z, -2.7557319223985888e-7f, 2.4801587301587302e-5f);
  cd = fmaf(cd, z, -1.3888888888888889e-3f);
  cd = fmaf(cd, z, 4.1666666666666667e-2f);
  cd = fmaf(cd, z, -0.5f);
  cd = fmaf(cd, z, 1.0f);
  t.s = t.sx * cd + t.cx * sd;
  t.c = t.cx * cd - t.sx * sd;
  return t;
}
PHX_DEV Pt32 pt32_at(float x1, float x2, float hs) { return pt32_at_r(exp_r(x1), x2, hs); }
// sin and cos of a small |e| <= 0.03 (Taylor to e^3 / e^4: truncation below 7e-9 of sin e and
// 1e-12 of cos e, under FP32 rounding)
PHX_DEV void sincos_tiny(float e, float* s, float* c) {
  const float z = e * e;
  *s = fmaf(e * z, -1.f / 6.f, e);
  *c = fmaf(z, fmaf(z, 1.f / 24.f, -0.5f), 1.f);
}
// The point quantities at (x1 + e1, x2 + e2) from those at (x1, x2), for offsets |e1|, |e2| <=
// 0.02 (an RK4 stage of a step, whose rule keeps |dl k^1| and |dl k^2| at or below the 0.01
// stepsize; the theta offset is at most (1 + h) |e2|): exp by its Taylor polynomial to e1^4
// (truncation below 3e-11), sin/cos by angle addition with sincos_tiny, theta' offset =
// e2 + (h/2)(sin 2(x2 + e2) - sin 2x2).  Errors stay at FP32 rounding; the offsets are applied to
// the base point exactly, not to a rounded stage position.
PHX_DEV Pt32 pt32_offset(const Pt32& b, float e1, float e2, float hs) {
  Pt32 t;
  const float ex = e1 * fmaf(e1, fmaf(e1, fmaf(e1, 1.f / 24.f, 1.f / 6.f), 0.5f), 1.f);
  t.r = fmaf(b.r, ex, b.r);
  float se, ce;
  sincos_tiny(e2, &se, &ce);
  t.sx = b.sx * ce + b.cx * se;
  t.cx = b.cx * ce - b.sx * se;
  t.s2x = 2.0f * t.sx * t.cx;
  t.c2x = (t.cx - t.sx) * (t.cx + t.sx);
  const float dth = fmaf(0.5f * hs, t.s2x - b.s2x, e2);
  float st, ct;
  sincos_tiny(dth, &st, &ct);
  t.s = b.s * ct + b.c * st;
  t.c = b.c * ct - b.s * st;
  return t;
}
template <class T, bool RECIP>
PHX_DEV void ham_core(T a, T hs, T pt, T pp, T r, T s2x, T c2x, T s, T c, T p1, T p2, T d[5], T* k0) {
  const T tp = T(1) + hs * c2x, tpp = T(-2) * hs * s2x;
  const T a2 = a * a;
  const T Sig = r * r + a2 * c * c;
  T S, ir, is, itp;
  if (RECIP) {
    const T rs = r * s, tS = tp * Sig;
    const T inv = T(1) / (rs * tS);
    S = inv * rs * tp;
    ir = inv * s * tS;
    is = inv * r * tS;
    itp = inv * rs * Sig;
  } else {
    S = T(1) / Sig;
    ir = T(1) / r;
    is = T(1) / s;
    itp = T(1) / tp;
  }
  const T is2 = is * is, itp2 = itp * itp;
  const T dr2 = T(1) - T(2) * ir + a2 * ir * ir;  // Delta / r^2
  const T ptt = pt * pt, q22 = p2 * p2 * itp2, q33 = pp * pp * is2;
  const T Q = T(-2) * r * ptt + T(4) * pt * p1 + dr2 * p1 * p1 + T(2) * a * ir * p1 * pp + q22 + q33;
  const T dQr = T(-2) * ptt + T(2) * ir * ir * (p1 * p1 * (T(1) - a2 * ir) - a * p1 * pp);
  const T dQth = T(-2) * q33 * c * is;
  const T SQ = S * Q;
  d[0] = S * (T(2) * pt + dr2 * p1 + a * ir * pp);
  d[1] = S * p2 * itp2;
  d[2] = S * (a * ir * p1 + pp * is2);
  d[3] = T(-0.5) * r * S * (dQr - T(2) * r * SQ);
  d[4] = T(-0.5) * S * (tp * (dQth + T(2) * a2 * s * c * SQ) - T(2) * q22 * tpp * itp);
  *k0 = -pt - T(2) * S * (r * pt - p1);
}
// The Walker-Penrose constant kappa = (A - i B)(r - i a cos theta) of a vector f at a point with
// wavevector k, from the Kerr-Schild components W^{mu nu} = k^mu f^nu - k^nu f^mu in
// (t, r, theta, phi) (k^r = r k^1, k^theta = theta' k^2 for the MKS components given):
//   A = W^{tr} + a sin^2 theta W^{r phi},  B = sin theta [(r^2 + a^2) W^{phi theta} - a (W^{t theta} + W^{r theta})].
// This is the Boyer-Lindquist form (Walker & Penrose 1970; Chandrasekhar 1983) with dt and dphi
// shifted by the Kerr-Schild terms 2r/Delta dr and a/Delta dr, which cancel to the regular form
// above.  kappa is constant along a null geodesic for a parallel-transported f, and it is
// real-linear in f, with kappa(k) = 0.
template <class T>
PHX_DEV void wp_kappa(T a, T r, T s, T c, T tp, const T k[4], const T f[4], T* q1, T* q2) {
  const T kr = r * k[1], kth = tp * k[2], fr = r * f[1], fth = tp * f[2];
  const T Wtr = k[0] * fr - kr * f[0], Wrp = kr * f[3] - k[3] * fr, Wpth = k[3] * fth - kth * f[3];
  const T Wtth = k[0] * fth - kth * f[0], Wrth = kr * fth - kth * fr;
  const T A = Wtr + a * s * s * Wrp;
  const T B = s * ((r * r + a * a) * Wpth - a * (Wtth + Wrth));
  *q1 = A * r - B * a * c;
  *q2 = -(A * a * c + B * r);
}
// cos 2 psi, sin 2 psi of the polarization vector that carries the ray's Walker-Penrose constant
// kap, in the plasma-frame basis (perp, para) at the sample (both orthogonal to k): with
// f = f1 perp + f2 para + (multiple of k), kap = f1 kappa(perp) + f2 kappa(para), a 2 x 2 real
// system whose solution (up to a common factor) is (f1, f2); 2 psi follows as in chi_from_F1
// (src/core/geometry/tetrad.hpp:160-182).
template <class T>
PHX_DEV void wp_angle(T a, const Met<T>& g, const T k[4], const T perp[4], const T para[4], T kap1, T kap2, T* c2,
                      T* s2) {
  T e1r, e1i, e2r, e2i;
  wp_kappa(a, g.r, g.s, g.c, g.hf, k, perp, &e1r, &e1i);
  wp_kappa(a, g.r, g.s, g.c, g.hf, k, para, &e2r, &e2i);
  const T f1 = kap1 * e2i - e2r * kap2, f2 = e1r * kap2 - kap1 * e1i;
  const T nn = f1 * f1 + f2 * f2;
  *c2 = div_rcp(f1 * f1 - f2 * f2, nn);
  *s2 = div_rcp(T(2) * f1 * f2, nn);
}
template <class T>
PHX_DEV T pole_x2(const Ray<T>& g) {  // geodesic.hpp:200-205
  constexpr bool F32 = Lim<T>::fp32;
  const T pi_hi = T(PI_D), pi_lo = T(PI_D - (double)T(PI_D));
  if (F32 && g.hi[2] > T(0.5 * PI_D)) return (pi_hi - g.hi[2]) + (pi_lo - g.lo[2]);
  return F32 ? g.hi[2] + g.lo[2] : g.hi[2];
}
// RAPTOR step rule (geodesic.hpp:267-292): negative step, capped at 1/4 cell inside the grid
template <class T>
PHX_DEV T step_len(T stepsize, T x2, const T k[4], bool have_dxc, const T dxc[3]) {
  const T S2 = Lim<T>::tiny;
  const T ak1 = fabs(k[1]) + S2, ak2 = fabs(k[2]) + S2, ak3 = fabs(k[3]) + S2;
  const T d1 = div_nr(stepsize, ak1);
  const T d2 = div_nr(stepsize * fmin(x2, T(PI_D) - x2), ak2);
  const T d3 = div_nr(stepsize, ak3);
  const T i1 = recip_nr(fabs(d1) + S2), i2 = recip_nr(fabs(d2) + S2), i3 = recip_nr(fabs(d3) + S2);
  T grid = Lim<T>::big;
  if (have_dxc) {
    const T s1 = div_nr(dxc[0], ak1), s2 = div_nr(dxc[1], ak2), s3 = div_nr(dxc[2], ak3);
    const T j1 = recip_nr(fabs(s1) + S2), j2 = recip_nr(fabs(s2) + S2), j3 = recip_nr(fabs(s3) + S2);
    grid = div_nr(recip_nr(j1 + j2 + j3), T(4));
  }
  return -fmin(recip_nr(i1 + i2 + i3), grid);
}
// The same rule for the mixed march, with the stepsize's and the cell sizes' reciprocals
// (inv_dxc) in place of the divisions by them: 1/|stepsize/k| = |k|/stepsize and 1/|dxc/k| =
// |k|/dxc, exact in real arithmetic and equal to FP32 rounding (the guard against a zero x2 term
// at the pole is kept as it is)
PHX_DEV float step_len_mixed(float stepsize, float inv_step, float x2, const float k[4], bool have_dxc,
                             const float inv_dxc[3]) {
  const float S2 = Lim<float>::tiny;
  const float ak1 = fabsf(k[1]) + S2, ak2 = fabsf(k[2]) + S2, ak3 = fabsf(k[3]) + S2;
  const float d2 = stepsize * fminf(x2, float(PI_D) - x2) / ak2;
  const float i1 = ak1 * inv_step, i2 = 1.f / (fabsf(d2) + S2), i3 = ak3 * inv_step;
  float grid = Lim<float>::big;
  if (have_dxc) grid = 0.25f / (ak1 * inv_dxc[0] + ak2 * inv_dxc[1] + ak3 * inv_dxc[2]);
  return -fminf(1.f / (i1 + i2 + i3), grid);
}

// ------------------------------------------------------------------ AMR lookup
// The snapshot's device tables: one 12-scalar record per leaf (lower corner, upper
// corner = lo + nx * dxc formed on the host exactly as the reference forms it,
// cell size), 8-scalar cell records, and the finest-level leaf map in which each
// entry already holds the first leaf of the reference's coarse-to-fine walk
// (src/core/model/sampler.hpp:55-82).
template <class T>
struct Grid {
  const T* blk;    // [nleafs][12]: lo[3], hi[3], dxc[3], 1/dxc[3]
  const T* meta;   // [nleafs]: the largest rho of a field-free leaf (B^i = 0, rho >= 0 and finite
                   // primitives in every cell), else -1
  const T* prim;   // [nleafs * cells][8]
  const int* map;  // [fd0][fd1][fd2]
  int nleafs, cells, nx0, nx1, nx2, fd0, fd1, fd2;
  T w0, w1, w2, up;  // base-level block widths, 2^(nlevels-1)
  T lo1, lo2, lo3, hi1, hi2, hi3;
};
// fmod of the wrap: fmod_exact in FP32 (the same values, a fraction of the library's cost); in
// FP64 x itself when |x| < y (fmod's exact value there; x2 and x3 almost always lie in range),
// else the library fmod
PHX_DEV float wrap_mod(float x, float y) { return fmod_exact(x, y, 1.f / y); }
PHX_DEV double wrap_mod(double x, double y) { return fabs(x) < y ? x : fmod(x, y); }
template <class T>
PHX_DEV void wrap(T x[4]) {  // sampler.hpp:29-38
  const T TP = T(2. * PI_D), PI = T(PI_D);
  // the common case, in which both fmods return their argument and no reflection follows
  if (x[3] >= T(0) && x[3] < TP && x[2] >= T(1e-6) && x[2] < PI) {
    x[2] = x[2] - T(1e-6);
    return;
  }
  x[3] = wrap_mod(x[3], TP);
  x[2] = wrap_mod(x[2], PI) - T(1e-6);
  if (x[3] < T(0)) x[3] = TP + x[3];
  if (x[2] < T(0)) {
    x[2] = -x[2];
    x[3] = PI + x[3];
  }
}
template <class T>
PHX_DEV bool off_domain(const Grid<T>& G, const T x[4]) {  // sampler.hpp:43-50
  return x[1] > G.hi1 || x[1] < G.lo1 || x[2] < G.lo2 || x[2] > G.hi2 || x[3] < G.lo3 || x[3] > G.hi3;
}
// The batched memory forms: box tests that load the six bounds first and combine the comparisons
// without short circuit (a short-circuit chain makes each load conditional on the previous
// comparison, and the compiler then waits for each bound in turn: six memory round trips instead
// of one), a sample's leaf record loaded in one round trip before its domain tests, and corner
// gathers whose loads are issued a layer at a time.  The FP32-cell marches (FP32, mixed) always
// use them; the FP64 march uses them when PHX_FP64_BATCHED is 1, and the short-circuit,
// interleaved forms otherwise.  Both forms compute the same values in the same order.
#ifndef PHX_FP64_BATCHED
#define PHX_FP64_BATCHED 1
#endif
template <class T>
constexpr bool batched_loads() {
  return Lim<T>::fp32 || PHX_FP64_BATCHED;
}
// the leaf's box lo[3], hi[3]
template <class T>
PHX_DEV void leaf_box(gptr<T> blk, int q, T bx[6]) {
  const gptr<T> b = blk + 12 * q;
#pragma unroll
  for (int i = 0; i < 6; i++) bx[i] = b[i];
}
template <class T>
PHX_DEV bool in_box_half_open(const T bx[6], const T x[4], T small) {
  return !((x[1] + small < bx[0]) | (x[1] + small >= bx[3]) | (x[2] + small < bx[1]) | (x[2] + small >= bx[4]) |
           (x[3] + small < bx[2]) | (x[3] + small >= bx[5]));
}
template <class T>
PHX_DEV bool in_box_closed(const T bx[6], const T x[4]) {
  return !((x[1] < bx[0]) | (x[1] > bx[3]) | (x[2] < bx[1]) | (x[2] > bx[4]) | (x[3] < bx[2]) | (x[3] > bx[5]));
}
template <class T>
PHX_DEV bool in_leaf(const Grid<T>& G, int q, const T x[4], T small) {  // half-open (sampler.hpp:87-101)
  if constexpr (batched_loads<T>()) {
    T bx[6];
    leaf_box(glb(G.blk), q, bx);
    return in_box_half_open(bx, x, small);
  } else {
    const gptr<T> b = glb(G.blk) + 12 * q;
    return !(x[1] + small < b[0] || x[1] + small >= b[3] || x[2] + small < b[1] || x[2] + small >= b[4] ||
             x[3] + small < b[2] || x[3] + small >= b[5]);
  }
}
template <class T>
PHX_DEV bool in_leaf_closed(const Grid<T>& G, int q, const T x[4]) {  // inclusive (sampler.hpp:109-122)
  if constexpr (batched_loads<T>()) {
    T bx[6];
    leaf_box(glb(G.blk), q, bx);
    return in_box_closed(bx, x);
  } else {
    const gptr<T> b = glb(G.blk) + 12 * q;
    return !(x[1] < b[0] || x[1] > b[3] || x[2] < b[1] || x[2] > b[4] || x[3] < b[2] || x[3] > b[5]);
  }
}
// The table a sample's cell is located in: the block records and the domain bounds, in the
// locating precision L (FP64 in the mixed-precision march, else the grid's own table)
template <class L>
struct Loc {
  const L* blk;
  L lo1, lo2, lo3, hi1, hi2, hi3;
};
template <class T>
PHX_DEV Loc<T> loc_of(const Grid<T>& G) {
  return Loc<T>{G.blk, G.lo1, G.lo2, G.lo3, G.hi1, G.hi2, G.hi3};
}
template <class L>
PHX_DEV bool off_domain(const Loc<L>& G, const L x[4]) {  // sampler.hpp:43-50
  return x[1] > G.hi1 || x[1] < G.lo1 || x[2] < G.lo2 || x[2] > G.hi2 || x[3] < G.lo3 || x[3] > G.hi3;
}
// BATCH: the batched form (the marches with FP32 cells, whose locating table may be FP64)
template <bool BATCH, class L>
PHX_DEV bool in_leaf_closed(const Loc<L>& G, int q, const L x[4]) {  // inclusive (sampler.hpp:109-122)
  if constexpr (BATCH) {
    L bx[6];
    leaf_box(glb(G.blk), q, bx);
    return in_box_closed(bx, x);
  } else {
    const gptr<L> b = glb(G.blk) + 12 * q;
    return !(x[1] < b[0] || x[1] > b[3] || x[2] < b[1] || x[2] > b[4] || x[3] < b[2] || x[3] > b[5]);
  }
}
// find_leaf (sampler.hpp:143-180) on an already wrapped position
template <class T>
PHX_DEV int find_leaf(const Grid<T>& G, const T x[4]);
// find_leaf of an FP64 position on the FP32 grid's map and the FP64 block table GL (the mixed
// march): the map candidate is computed from the FP64 offset rounded to FP32, and the candidate is
// validated, and the linear fallback scan run, on the FP64 boxes, so the result is the leaf whose
// FP64 box holds the point (a candidate moved by the rounding fails validation and falls back to
// the scan).  Keeps the FP64 grid's map parameters out of the launch arguments.
template <class L>
PHX_DEV int find_leaf_located(const Grid<float>& G, const Loc<L>& GL, const L x[4]) {
  const L small = x[2] > L(PI_D / 2.) ? L(-1e-9) : L(1e-9);
  int cand = -1;
  {
    const float fi = float(x[1] + small - GL.lo1) / G.w0 * G.up;
    const float fj = float(x[2] + small - GL.lo2) / G.w1 * G.up;
    const float fk = float(x[3] + small - GL.lo3) / G.w2 * G.up;
    if (!(fi < 0.f || fj < 0.f || fk < 0.f)) {
      const int i = (int)fi, j = (int)fj, k = (int)fk;
      if (i < G.fd0 && j < G.fd1 && k < G.fd2) cand = glb(G.map)[(i * G.fd1 + j) * G.fd2 + k];
    }
  }
  auto inside = [&](int q) {
    L bx[6];
    leaf_box(glb(GL.blk), q, bx);
    return in_box_half_open(bx, x, small);
  };
  if (cand >= 0 && inside(cand)) return cand;
  const L eps = L(1e-6);
  if (cand < 0 && (x[1] < GL.lo1 - eps || x[1] > GL.hi1 + eps || x[2] < GL.lo2 - eps || x[2] > GL.hi2 + eps ||
                   x[3] < GL.lo3 - eps || x[3] > GL.hi3 + eps))
    return -1;
  for (int q = 0; q < G.nleafs; q++)
    if (inside(q)) return q;
  return -1;
}
template <class T>
PHX_DEV int find_leaf(const Grid<T>& G, const T x[4]) {
  const T small = x[2] > T(PI_D / 2.) ? T(-1e-9) : T(1e-9);
  int cand = -1;
  {
    const T fi = (x[1] + small - G.lo1) / G.w0 * G.up;
    const T fj = (x[2] + small - G.lo2) / G.w1 * G.up;
    const T fk = (x[3] + small - G.lo3) / G.w2 * G.up;
    if (!(fi < T(0) || fj < T(0) || fk < T(0))) {
      const int i = (int)fi, j = (int)fj, k = (int)fk;
      if (i < G.fd0 && j < G.fd1 && k < G.fd2) cand = glb(G.map)[((size_t)i * G.fd1 + j) * G.fd2 + k];
    }
  }
  if (cand >= 0 && in_leaf(G, cand, x, small)) return cand;
  const T eps = T(1e-6);
  if (cand < 0 && (x[1] < G.lo1 - eps || x[1] > G.hi1 + eps || x[2] < G.lo2 - eps || x[2] > G.hi2 + eps ||
                   x[3] < G.lo3 - eps || x[3] > G.hi3 + eps))
    return -1;
  for (int q = 0; q < G.nleafs; q++)  // the reference's linear fallback scan
    if (in_leaf(G, q, x, small)) return q;
  return -1;
}
// A lane's record of its last lookup: the leaf and its box and cell size.  A ray stays in one leaf
// for dozens of steps, and while its position lies in that leaf's (half-open, nudged) box the
// lookup needs no memory access (and no grid parameter).  Leaves tile the domain, so the leaf whose box holds a point is
// the one find_leaf returns; the two differ only where rounding makes two boxes overlap by an ulp
// and the map names the other leaf.
template <class T>
struct LeafCache {
  int leaf = -1;
  T box[6], dxc[3];  // dxc: the cell size (the mixed march keeps its reciprocal)
};
// whether the lane's cached leaf holds x (half-open box with find_leaf's nudge)
template <class T>
PHX_DEV bool leaf_cache_hit(const LeafCache<T>& c, const T x[4]) {
  const T small = x[2] > T(PI_D / 2.) ? T(-1e-9) : T(1e-9);
  return c.leaf >= 0 && !(x[1] + small < c.box[0] || x[1] + small >= c.box[3] || x[2] + small < c.box[1] ||
                          x[2] + small >= c.box[4] || x[3] + small < c.box[2] || x[3] + small >= c.box[5]);
}
// find_leaf on a miss of the lane's cache: the map candidate, once validated, refills the cache
// (*cached then tells that c.dxc holds the leaf's cell size); the fallbacks are find_leaf's
// DXC: the block-record slot the cache takes the cell size from (6: dxc; 9: 1/dxc, the mixed march)
template <class T, int DXC = 6>
PHX_DEV int find_leaf_fill(const Grid<T>& G, const T x[4], LeafCache<T>& c, bool* cached) {
  const T small = x[2] > T(PI_D / 2.) ? T(-1e-9) : T(1e-9);
  *cached = false;
  int cand = -1;
  {
    const T fi = (x[1] + small - G.lo1) / G.w0 * G.up;
    const T fj = (x[2] + small - G.lo2) / G.w1 * G.up;
    const T fk = (x[3] + small - G.lo3) / G.w2 * G.up;
    if (!(fi < T(0) || fj < T(0) || fk < T(0))) {
      const int i = (int)fi, j = (int)fj, k = (int)fk;
      if (i < G.fd0 && j < G.fd1 && k < G.fd2) cand = glb(G.map)[(i * G.fd1 + j) * G.fd2 + k];
    }
  }
  if (cand >= 0) {
    const gptr<T> b = glb(G.blk) + 12 * cand;
    T box[6];
#pragma unroll
    for (int q = 0; q < 6; q++) box[q] = b[q];
    const bool inside = batched_loads<T>() ? in_box_half_open(box, x, small)
                                     : !(x[1] + small < box[0] || x[1] + small >= box[3] || x[2] + small < box[1] ||
                                         x[2] + small >= box[4] || x[3] + small < box[2] || x[3] + small >= box[5]);
    if (inside) {
      c.leaf = cand;
#pragma unroll
      for (int q = 0; q < 6; q++) c.box[q] = box[q];
#pragma unroll
      for (int q = 0; q < 3; q++) c.dxc[q] = b[DXC + q];
      *cached = true;
      return cand;
    }
  }
  const T eps = T(1e-6);
  if (cand < 0 && (x[1] < G.lo1 - eps || x[1] > G.hi1 + eps || x[2] < G.lo2 - eps || x[2] > G.hi2 + eps ||
                   x[3] < G.lo3 - eps || x[3] > G.hi3 + eps))
    return -1;
  for (int q = 0; q < G.nleafs; q++)  // the reference's linear fallback scan
    if (in_leaf(G, q, x, small)) return q;
  return -1;
}

// ------------------------------------------------------------------ fluid
template <class T>
struct Fluid {
  T U[4], b[4], ne, B, te;
  T rho, uu, Bsq;  // the interpolated rho and u, and b.b + 1e-6: the unit-free inputs of ne, B and Theta_e
  T x[4];          // the (wrapped) position the sampler used
  Met<T> g;
};
template <class T>
struct Plasma {  // plasma parameters shared by every image of a launch (src/core/model/bhac.hpp:150-151)
  T a, hs, gam, rlow, ibc, sigma_cut, te_min, te_max, rt_outer;
};
// clang vector types: loaded through a global pointer as one wide load each
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
PHX_DEV void load8(gptr<float> p, float v[8]) {
  const f4v a = *reinterpret_cast<gptr<f4v>>(p), b = *reinterpret_cast<gptr<f4v>>(p + 4);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
PHX_DEV void load2(gptr<float> p, float v[2]) {
  const f2v a = *reinterpret_cast<gptr<f2v>>(p);
  v[0] = a.x; v[1] = a.y;
}
PHX_DEV void load2(gptr<double> p, double v[2]) {
  const d2v a = *reinterpret_cast<gptr<d2v>>(p);
  v[0] = a.x; v[1] = a.y;
}
PHX_DEV void load8(gptr<double> p, double v[8]) {
  const d2v a = *reinterpret_cast<gptr<d2v>>(p), b = *reinterpret_cast<gptr<d2v>>(p + 2);
  const d2v c = *reinterpret_cast<gptr<d2v>>(p + 4), d = *reinterpret_cast<gptr<d2v>>(p + 6);
  v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
  v[4] = c.x; v[5] = c.y; v[6] = d.x; v[7] = d.y;
}
// n_e, |B| and Theta_e of a sample (sampler.hpp:314-358) from rho, u and b.b + 1e-6 and the
// request's plasma units: the part of derive_fluid that depends on m_unit and R_high
template <class T>
PHX_DEV void plasma_units(const Plasma<T>& P, T Ne, T Bu, T rhigh, T rho, T uu, T Bsq, T* ne, T* B, T* te) {
  const T tiny = T(1.e-6);
  *ne = rho * Ne + tiny;
  *B = sqrt(Bsq) * Bu;
  const T beta = div_rcp(uu * (P.gam - T(1)), T(0.5) * (Bsq + tiny));
  const T bb = beta * P.ibc;
  const T b2 = bb * bb;
  const T trat = div_rcp(rhigh * b2, T(1) + b2) + div_rcp(P.rlow, T(1) + b2);
  const T te_unit = div_rcp(T(1. / 3. * (MP / ME)), trat + T(1));
  *te = div_rcp(uu, rho) * te_unit;
}

// get_fluid_params + derive_fluid (src/core/model/sampler.hpp:305-427): wrap, cached-leaf test,
// re-wrap and lookup on a miss, cell by truncation, weights from the cell's lower corner,
// block-local trilinear gather of the 8 primitives, then U^mu, b^mu, |B|, Theta_e and the cuts.
// The cell is located in precision L (the block table GL) and the gather and the fluid
// derivation run in T (the cell table G); L == T except in the mixed-precision march, which
// locates the cell in FP64.  Ne, Bu, rhigh are the request's (per image) plasma units.  UNITS
// false: the Theta_e cut, the only cut that depends on them, is left to the caller (rho, uu and Bsq
// are kept in fl for it).
template <class T, class L, bool UNITS = true>
PHX_DEV int sample_fluid(const Grid<T>& G, const Loc<L>& GL, const Plasma<T>& P, T Ne, T Bu, T rhigh, const L X[4],
                         int cache, Fluid<T>& fl) {
  constexpr bool F32 = Lim<L>::fp32;
  constexpr bool MIXED = !std::is_same<T, L>::value;
  // Batched form: the hinted leaf's whole record (box, cell size and its reciprocal) and its
  // field-free mark are loaded in one memory round trip, before the domain tests; they serve the
  // box test and the cell location (a failed test reloads them for the leaf the lookup finds).
  // Otherwise the box is loaded for the test and the rest after it.
  constexpr bool BATCH = batched_loads<T>();
  L x[4] = {X[0], X[1], X[2], X[3]};
  wrap(x);
  int q = cache;
  L rec[12];
  T rho_ff = T(-1);
  auto load_record = [&](int leaf) {
    const gptr<L> b = glb(GL.blk) + 12 * leaf;
#pragma unroll
    for (int i = 0; i < 12; i++) rec[i] = b[i];
    rho_ff = glb(G.meta)[leaf];
  };
  if constexpr (BATCH) {
    if (q != -1) load_record(q);
  }
  // r < 1: exp(x1) < 1 (the FP64 locate of the mixed march tests x1 < 0, the same cut)
  const bool inside_one = std::is_same<T, L>::value ? (exp(x[1]) < L(1)) : (x[1] < L(0));
  if (inside_one || off_domain(GL, x)) return 0;
  bool hint_ok;
  if constexpr (BATCH)
    hint_ok = q != -1 && in_box_closed(rec, x);
  else
    hint_ok = q != -1 && in_leaf_closed<false>(GL, q, x);
  if (!hint_ok) {
    wrap(x);
    if constexpr (std::is_same<T, L>::value)
      q = find_leaf(G, x);
    else
      q = find_leaf_located(G, GL, x);
    if constexpr (BATCH) {
      if (q != -1) load_record(q);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; i++) fl.x[i] = T(x[i]);
  if (q == -1) return 0;
  // A field-free leaf (B^i = 0, rho >= 0 and finite primitives in every cell): every sample in it has b = 0
  // exactly, and needs only rho and u.  Where its largest rho is so low that even that rho fails
  // the sigma cut (sigma = (0 + 1e-6) / (rho + 1e-20) > sigma_cut; the interpolated rho is a convex
  // combination of the corners, so it lies in [0, max]; the margin covers its rounding), every
  // sample in it is cut: nothing emits and nothing is counted, as the full derivation would find.
  if constexpr (!BATCH) load_record(q);  // unbatched: the record and the mark after the test
  L blo[3], bdx[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    blo[d] = rec[d];
    bdx[d] = MIXED ? rec[9 + d] : rec[6 + d];  // the mixed march: 1/dx
  }
  const bool ffleaf = rho_ff >= T(0);
  if (ffleaf && T(1e-6) > P.sigma_cut * (rho_ff + T(1e-20)) * T(1.00001)) {
    fl.ne = T(0);
    return 0;
  }
  // The mixed-precision march stops at any sample in a field-free leaf: it can only be cut or
  // dropped (b = 0), it changes no image, and the mixed march does not keep the reference's
  // plasma-segment and dropped-segment counts (its samples are not counted there).
  if (MIXED && ffleaf) {
    fl.ne = T(0);
    return 0;
  }
  const int nxs[3] = {G.nx0, G.nx1, G.nx2};
  int c[3];
  T del[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const L lb = blo[d], dx = bdx[d];
    // the mixed march (FP64 location, FP32 fluid) multiplies by the reciprocal cell width: a
    // last-ulp change of u only moves a sample across a cell face, where the block-local trilinear
    // weights are continuous; FP32 and FP64 divide as the reference does
    const L u = MIXED ? (x[1 + d] - lb) * dx : (x[1 + d] - lb) / dx;
    c[d] = (int)u;
    if (c[d] >= nxs[d]) c[d] = nxs[d] - 1;
    const int ix = F32 ? (int)floor(u) : (int)(u + L(1000)) - 1000;
    del[d] = (ix < 0)             ? T(0)
             : (ix > nxs[d] - 2)  ? T(1)
             : MIXED              ? T(u - L(ix))
                                  : T((x[1 + d] - (L(ix) * dx + lb)) / dx);
  }
  const int ip = c[0] + 1 < G.nx0 ? c[0] + 1 : c[0];
  const int jp = c[1] + 1 < G.nx1 ? c[1] + 1 : c[1];
  const int kp = c[2] + 1 < G.nx2 ? c[2] + 1 : c[2];
  const T w1 = T(1) - del[0], w2 = T(1) - del[1], w3 = T(1) - del[2];
  const int ci[4] = {c[0], c[0], ip, ip}, cj[4] = {c[1], jp, c[1], jp};
  const T wq[4] = {w1 * w2, w1 * del[1], del[0] * w2, del[0] * del[1]};
  const gptr<T> base = glb(G.prim) + (size_t)q * G.cells * 8;
  const gptr<T> prim = glb(G.prim);               // FP32 cells: 32-bit offsets from the table base
  const unsigned qc = (unsigned)q * (unsigned)G.cells;  // the leaf's first cell
  T qv[8];
#pragma unroll
  for (int a = 0; a < 8; a++) qv[a] = T(0);
  // Batched form: the corner loads of a layer (4 corners; the field-free gather: all 8) are issued
  // before the first is used: the scheduling barriers keep the compiler from interleaving them with
  // the sums, which under register pressure waits for each corner in turn (8 memory round trips a
  // sample).  Otherwise the loads and sums interleave.
  if constexpr (!BATCH) {
    if (ffleaf) {
#pragma unroll
      for (int z = 0; z < 2; z++) {
        const int ck = z ? kp : c[2];
        const T wz = z ? del[2] : w3;
#pragma unroll
        for (int s = 0; s < 4; s++) {
          T v[2];
          load2(base + (size_t)(ci[s] + cj[s] * G.nx0 + ck * G.nx0 * G.nx1) * 8, v);
          const T w = wq[s] * wz;
#pragma unroll
          for (int a = 0; a < 2; a++) qv[a] += v[a] * w;
        }
      }
    } else {
#pragma unroll
      for (int z = 0; z < 2; z++) {
        const int ck = z ? kp : c[2];
        const T wz = z ? del[2] : w3;
#pragma unroll
        for (int s = 0; s < 4; s++) {
          T v[8];
          load8(base + (size_t)(ci[s] + cj[s] * G.nx0 + ck * G.nx0 * G.nx1) * 8, v);
          const T w = wq[s] * wz;
#pragma unroll
          for (int a = 0; a < 8; a++) qv[a] += v[a] * w;
        }
      }
    }
  } else if (ffleaf) {
    // a field-free leaf: only rho and u of the 8 corners, interpolated as below
    T v[8][2];
#pragma unroll
    for (int z = 0; z < 2; z++) {
      const int ck = z ? kp : c[2];
#pragma unroll
      for (int s = 0; s < 4; s++) load2(at32(prim, (qc + (unsigned)(ci[s] + cj[s] * G.nx0 + ck * G.nx0 * G.nx1)) * 8u), v[4 * z + s]);
    }
    PHX_SCHED_BARRIER();
#pragma unroll
    for (int z = 0; z < 2; z++) {
      const T wz = z ? del[2] : w3;
#pragma unroll
      for (int s = 0; s < 4; s++) {
        const T w = wq[s] * wz;
#pragma unroll
        for (int a = 0; a < 2; a++) qv[a] += v[4 * z + s][a] * w;
      }
    }
  } else {
#pragma unroll
    for (int z = 0; z < 2; z++) {
      const int ck = z ? kp : c[2];
      T v8[4][8];
#pragma unroll
      for (int s = 0; s < 4; s++) load8(at32(prim, (qc + (unsigned)(ci[s] + cj[s] * G.nx0 + ck * G.nx0 * G.nx1)) * 8u), v8[s]);
      PHX_SCHED_BARRIER();
      const T wz = z ? del[2] : w3;
#pragma unroll
      for (int s = 0; s < 4; s++) {
        const T* v = v8[s];
        const T w = wq[s] * wz;
#pragma unroll
        for (int a = 0; a < 8; a++) qv[a] += v[a] * w;
      }
      PHX_SCHED_BARRIER();
    }
  }
  // derive_fluid
  const T tiny = T(1.e-6);
  if (ffleaf ||
      (qv[5] == T(0) && qv[6] == T(0) && qv[7] == T(0) && isfinite(qv[2]) && isfinite(qv[3]) && isfinite(qv[4]))) {
    // a field-free sample (B^i = 0 at all 8 corners): b^mu = 0 exactly, whatever the velocity and the
    // metric (alpha and the Lorentz factor are finite: g^tt < 0 and the Kerr-Schild spatial metric is
    // positive definite), so |B|, beta, Theta_e and the cuts follow from rho and u alone, with the
    // same expressions as below; U^mu and the metric are not needed by any caller of a sample with b = 0
    const T rho = qv[0], uu = qv[1];
#pragma unroll
    for (int i = 0; i < 4; i++) fl.b[i] = T(0);
    const T Bsq = tiny;
    fl.rho = rho;
    fl.uu = uu;
    fl.Bsq = Bsq;
    plasma_units(P, Ne, Bu, rhigh, rho, uu, Bsq, &fl.ne, &fl.B, &fl.te);
    if ((div_rcp(Bsq, rho + T(1e-20)) > P.sigma_cut) || exp_r(fl.x[1]) > P.rt_outer ||
        (UNITS && (fl.te > P.te_max || fl.te < P.te_min))) {
      fl.ne = T(0);
      return 0;
    }
    return 1;
  }
  const Met<T> g = metric(P.a, P.hs, fl.x[1], fl.x[2]);
  const MetU<T> u = metric_up(P.a, g);
  fl.g = g;
  const T rho = qv[0], uu = qv[1];
  const T gV[4] = {T(0), qv[2], qv[3], qv[4]};
  const T Bp[4] = {T(0), qv[5], qv[6], qv[7]};
  const T shift1 = div_rcp(u.u01, -u.u00);  // shift^2 = shift^3 = 0 in Kerr-Schild
  const T alpha = T(1) / sqrt(-u.u00);
  const T vv = g.d11 * gV[1] * gV[1] + T(2) * g.d13 * gV[1] * gV[3] + g.d22 * gV[2] * gV[2] + g.d33 * gV[3] * gV[3];
  const T lf = sqrt(vv + T(1));
  const T V1 = div_rcp(gV[1], lf), V2 = div_rcp(gV[2], lf), V3 = div_rcp(gV[3], lf);
  fl.U[0] = div_rcp(lf, alpha);
  fl.U[1] = V1 * lf - div_rcp(shift1 * lf, alpha);
  fl.U[2] = V2 * lf;
  fl.U[3] = V3 * lf;
  const T Vl1 = g.d11 * V1 + g.d13 * V3, Vl2 = g.d22 * V2, Vl3 = g.d13 * V1 + g.d33 * V3;
  fl.b[0] = div_rcp(lf * (Bp[1] * Vl1 + Bp[2] * Vl2 + Bp[3] * Vl3), alpha);
#pragma unroll
  for (int i = 1; i < 4; i++) fl.b[i] = div_rcp(Bp[i] + alpha * fl.b[0] * fl.U[i], lf);
  T bd[4];
  lower(g, fl.b, bd);
  const T Bsq = fabs(fl.b[0] * bd[0] + fl.b[1] * bd[1] + fl.b[2] * bd[2] + fl.b[3] * bd[3]) + tiny;
  fl.rho = rho;
  fl.uu = uu;
  fl.Bsq = Bsq;
  plasma_units(P, Ne, Bu, rhigh, rho, uu, Bsq, &fl.ne, &fl.B, &fl.te);
  if ((div_rcp(Bsq, rho + T(1e-20)) > P.sigma_cut) || g.r > P.rt_outer ||
      (UNITS && (fl.te > P.te_max || fl.te < P.te_min))) {
    fl.ne = T(0);
    return 0;
  }
  return 1;
}

// ------------------------------------------------------------------ tetrad
// Plasma-frame tetrad columns e_(1) (perp) and e_(2) (para) (src/core/geometry/tetrad.hpp:51-101);
// sqrt(-g) of Kerr-Schild/MKS in closed form, Sigma |sin theta| r (1 + h cos 2x2).
template <class T>
PHX_DEV void frame(const Met<T>& g, T a, const T k[4], const T U[4], const T b[4], T perp[4], T para[4]) {
  const T om = -dot(g, k, U);
  T eK[4];
#pragma unroll
  for (int i = 0; i < 4; i++) eK[i] = div_rcp(k[i], om) - U[i];
  const T Be = dot(g, U, b);
  const T Cc = div_rcp(dot(g, k, b), om) - Be;
  const T b2 = dot(g, b, b);
  const T Nc = sqrt(b2 + Be * Be - Cc * Cc);
#pragma unroll
  for (int i = 0; i < 4; i++) para[i] = div_rcp(b[i] + Be * U[i] - Cc * eK[i], Nc);
  T Ud[4], kd[4], bd[4];
  lower(g, U, Ud);
  lower(g, k, kd);
  lower(g, b, bd);
  const T sq = (g.r * g.r + a * a * g.c * g.c) * fabs(g.s) * g.r * g.hf;
  const T s = div_rcp(T(-1), sq * om * Nc);
  // Levi-Civita contraction eps^{i..} U_a k_c b_d (tetrad.hpp:70-80)
  const T m01 = kd[0] * bd[1] - kd[1] * bd[0], m02 = kd[0] * bd[2] - kd[2] * bd[0], m03 = kd[0] * bd[3] - kd[3] * bd[0];
  const T m12 = kd[1] * bd[2] - kd[2] * bd[1], m13 = kd[1] * bd[3] - kd[3] * bd[1], m23 = kd[2] * bd[3] - kd[3] * bd[2];
  perp[0] = s * (Ud[1] * m23 - Ud[2] * m13 + Ud[3] * m12);
  perp[1] = -s * (Ud[0] * m23 - Ud[2] * m03 + Ud[3] * m02);
  perp[2] = s * (Ud[0] * m13 - Ud[1] * m03 + Ud[3] * m01);
  perp[3] = -s * (Ud[0] * m12 - Ud[1] * m02 + Ud[2] * m01);
}
template <class T>
PHX_DEV T pitch_angle(const Met<T>& g, const T k[4], const T b[4], const T U[4]) {  // tetrad.hpp:138-154
  const T bn = sqrt(fabs(dot(g, b, b)));
  if (bn == T(0)) return T(PI_D / 2.);
  const T kk = fabs(dot(g, k, U));
  T mu = div_rcp(dot(g, k, b), kk * bn);
  if (fabs(mu) > T(1)) mu = copysign(T(1), mu);  // mu / |mu|
  return acos(mu);
}

// ------------------------------------------------------------------ coefficients
// Dexter (2016) thermal synchrotron fits and Faraday coefficients (src/core/transfer/coeffs.hpp:44-138)
template <class T>
struct Coef {
  T jI, jQ, jU, jV, aI, aQ, aU, aV, rQ, rU, rV;
};
template <class T>
PHX_DEV T bes_c(int which, int p, int j);
template <>
PHX_DEV float bes_c<float>(int which, int p, int j) { return which ? PHX_BES_K0K2_F[p][j] : PHX_BES_K1K2_F[p][j]; }
template <>
PHX_DEV double bes_c<double>(int which, int p, int j) { return which ? PHX_BES_K0K2[p][j] : PHX_BES_K1K2[p][j]; }
template <class T>
PHX_DEV T bes_ln(int p, int e);
template <>
PHX_DEV float bes_ln<float>(int p, int e) { return PHX_BES_LNEDGE_F[p][e]; }
template <>
PHX_DEV double bes_ln<double>(int p, int e) { return PHX_BES_LNEDGE[p][e]; }

template <class T>
PHX_DEV T hankel(int nu, T x) {  // src/core/numerics/bessel.hpp:99-108
  const T mu = T(4 * nu * nu);
  T term = T(1), sum = T(1);
  for (int k = 1; k <= 30; k++) {
    term = term * ((mu - T(2 * k - 1) * T(2 * k - 1)) / (T(k) * T(8) * x));
    sum += term;
    if (fabs(term) < Lim<T>::tol_asym) break;
  }
  return sum;
}
template <class T>
PHX_DEV void bessel_ratios(T x, T* r10, T* r02) {  // K1/K2 and K0/K2 (bessel.hpp:118-148)
  if (x >= T(30)) {
    const T s2 = hankel(2, x);
    *r10 = hankel(1, x) / s2;
    *r02 = hankel(0, x) / s2;
    return;
  }
  // x = 1 / Theta_e >= 0.2 here (kratios takes the asymptotic forms for Theta_e > 5), so the
  // reference's numerical-integral branch for x < 0.01 (bessel.hpp:65-88, 138-141) is never reached;
  // below the table's first edge the ratios are NaN, and a segment built from them is dropped
  if (x < T(0.01)) {
    *r10 = *r02 = T(NAN);
    return;
  }
  const int p = (x < T(0.1)) ? 0 : (x < T(1.0)) ? 1 : 2;
  const T la = bes_ln<T>(p, 0), lb = bes_ln<T>(p, 1);
  const T z = (T(2) * log(x) - (lb + la)) / (lb - la);
  T b1a = T(0), b2a = T(0), b1b = T(0), b2b = T(0);
  if constexpr (Lim<T>::fp32) {
    // FP32: all 24 table entries of the piece (zero past the fit's length, where the recurrence
    // stays exactly 0 from its zero start, so the result is the same), in 3 batches of 8 per fit
    // whose loads are issued together (a per-term load would wait for each coefficient in turn)
#pragma unroll
    for (int j0 = 23; j0 >= 1; j0 -= 8) {
      T ca[8], cb[8];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        ca[i] = j0 - i >= 0 ? bes_c<T>(0, p, j0 - i) : T(0);
        cb[i] = j0 - i >= 0 ? bes_c<T>(1, p, j0 - i) : T(0);
      }
      PHX_SCHED_BARRIER();
#pragma unroll
      for (int i = 0; i < 8; i++) {
        if (j0 - i < 1) break;
        const T ta = T(2) * z * b1a - b2a + ca[i];
        b2a = b1a;
        b1a = ta;
        const T tb = T(2) * z * b1b - b2b + cb[i];
        b2b = b1b;
        b1b = tb;
      }
      PHX_SCHED_BARRIER();
    }
  } else {
    for (int j = PHX_BES_NTERM[p] - 1; j >= 1; j--) {  // Clenshaw, both fits at once
      const T ta = T(2) * z * b1a - b2a + bes_c<T>(0, p, j);
      b2a = b1a;
      b1a = ta;
      const T tb = T(2) * z * b1b - b2b + bes_c<T>(1, p, j);
      b2b = b1b;
      b1b = tb;
    }
  }
  *r10 = bes_c<T>(0, p, 0) + z * b1a - b2a;
  *r02 = bes_c<T>(1, p, 0) + z * b1b - b2b;
}
template <class T>
PHX_DEV T fm_conv(T X) {  // f_m(X) (coeffs.hpp:63-70)
  const T lx = log(X);
  const T ex = exp(-div_const(X, 47.2));
  return T(2.011) * exp(-div_const(exp(T(1.035) * lx), 4.7)) - cos(X * T(0.5)) * exp(-div_const(exp(T(1.2) * lx), 2.73)) -
         T(0.011) * ex +
         (T(0.011) * ex - T(369.69342869090815) * exp(T(-8. / 3.) * lx)) * T(0.5) *
             (T(1) + tanh(T(10) * (lx - T(4.787491742782046))));
}
// plasma-frame coefficients of one frequency (coeffs.hpp:95-138)
template <class T>
PHX_DEV Coef<T> thermal(T te, T ne, T nu, T B, T sth, T cth, T kr1, T kr0) {
  Coef<T> r;
  r.jU = r.aU = r.rU = T(0);
  const T ite = div_rcp(T(1), te);
  const T nuc = T(NU_CRIT) * B * sth * te * te;
  const T x = div_rcp(nu, nuc);
  const T pref = T(JNU) * ne * nu * ite * ite;
  const T cx = cbrt(x), ic = div_rcp(T(1), cx), ex = exp(T(-1.8899) * cx);
  r.jI = pref * (T(2.5651) * (T(1) + T(1.92) * ic + T(0.9977) * ic * ic) * ex);
  r.jQ = pref * (T(2.5651) * (T(1) + T(0.93193) * ic + T(0.499873) * ic * ic) * ex);
  r.jV = T(4. / 3.) * pref * ite * div_rcp(cth, sth) *
         ((T(1.81348) * ic * ic * ic + T(3.42319) * ic * ic + div_rcp(T(0.0292545), sqrt(x)) + T(2.03773) * ic) * ex);
  const T pre = T(2. * HP) * nu * nu * nu / T(CL * CL);
  const T xx = div_rcp(T(H_OVER_MEC2) * nu, te);
  const T Bnu = Lim<T>::fp32 ? div_rcp(pre, expm1(xx)) : pre / (exp(xx) - T(1));
  const T iB = div_rcp(T(1), Bnu);
  r.aI = r.jI * iB;
  r.aQ = r.jQ * iB;
  r.aV = r.jV * iB;
  const T iw = div_rcp(T(1), T(2 * PI_D) * nu);
  const T wp2 = T(WP2) * ne;
  const T om0 = T(GYRO) * B;
  const T Xe = te * sqrt(T(1.4142135623730951) * sth * div_rcp(T(1.e3 / (2. * PI_D)) * om0, nu));
  const T q = om0 * iw;
  const T wpw = wp2 * iw;
  r.rQ = (T(0.5) / T(CL)) * wpw * q * q * fm_conv(Xe) * (kr1 + T(6) * te) * sth * sth;
  const T shgm = T(1) - T(0.11) * log(T(1) + T(0.035) * Xe);
  r.rV = (T(1) / T(CL)) * wpw * q * (kr0 * shgm) * cth;
  return r;
}
// K1/K2 and K0/K2 at 1/Theta_e (frequency independent; coeffs.hpp:122-131)
template <class T>
PHX_DEV void kratios(T te, T* kr1, T* kr0) {
  const T ite = T(1) / te;
  if (ite < T(1. / 5.)) {
    const T k0 = -log(ite / T(2)) - T(0.5772), k1 = div_rcp(T(1), ite), k2 = div_rcp(div_rcp(T(2), ite), ite);
    *kr1 = div_rcp(k1, k2);
    *kr0 = (k2 > T(0)) ? div_rcp(k0, k2) : T(1);
  } else {
    bessel_ratios(ite, kr1, kr0);
  }
}
// Lorentz-invariant form + polarization-fraction clamp + rotation into the camera basis
// (coeffs.hpp:147-182)
template <class T>
PHX_DEV Coef<T> invariant_rotated(Coef<T> c, T nup, T jscale, T c2, T s2) {
  const T js = div_rcp(div_rcp(jscale, nup), nup);
  c.jI = c.jI * js;
  c.jQ = c.jQ * js;
  c.jV = c.jV * js;
  c.aI = c.aI * nup;
  c.aQ = c.aQ * nup;
  c.aV = c.aV * nup;
  c.rQ = c.rQ * nup;
  c.rV = c.rV * nup;
  const T pf = div_rcp(sqrt(c.jQ * c.jQ + c.jV * c.jV), c.jI);
  if (pf > T(1)) {
    c.jQ = div_rcp(c.jQ, pf + T(0.005));
    c.jU = div_rcp(c.jU, pf + T(0.005));
    c.jV = div_rcp(c.jV, pf + T(0.005));
  }
  Coef<T> r = c;
  r.jQ = c2 * c.jQ + s2 * c.jU;
  r.jU = -s2 * c.jQ + c2 * c.jU;
  r.aQ = c2 * c.aQ + s2 * c.aU;
  r.aU = -s2 * c.aQ + c2 * c.aU;
  r.rQ = c2 * c.rQ + s2 * c.rU;
  r.rU = -s2 * c.rQ + c2 * c.rU;
  return r;
}

// ------------------------------------------------------------------ operator
// O_A = (I + h/2 K)^-1 (I - h/2 K), O_S = h (I + h/2 K)^-1 j (src/core/transfer/operators.hpp:158-176),
// Gauss-Jordan with the reference's successive-swap partial pivoting (operators.hpp:91-136)
template <class T>
PHX_DEV void segment_op(const Coef<T>& c, T h, T OA[16], T OS[4]) {
  const T K[4][4] = {{c.aI, c.aQ, c.aU, c.aV}, {c.aQ, c.aI, c.rV, -c.rU}, {c.aU, -c.rV, c.aI, c.rQ},
                     {c.aV, c.rU, -c.rQ, c.aI}};
  const T hh = T(0.5) * h;
  T A[4][8];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      A[i][j] = (i == j ? T(1) : T(0)) + hh * K[i][j];
      A[i][4 + j] = (i == j) ? T(1) : T(0);
    }
#pragma unroll
  for (int c0 = 0; c0 < 4; c0++) {
#pragma unroll
    for (int r = c0 + 1; r < 4; r++) {
      const bool sw = fabs(A[r][c0]) > fabs(A[c0][c0]);
#pragma unroll
      for (int j = c0; j < 8; j++) {
        const T t = A[c0][j];
        A[c0][j] = sw ? A[r][j] : t;
        A[r][j] = sw ? t : A[r][j];
      }
    }
    const T d = A[c0][c0];
    if constexpr (Lim<T>::fp32) {  // one reciprocal per pivot (the row's divisions share it)
      const T id = div_rcp(T(1), d);
#pragma unroll
      for (int j = c0; j < 8; j++) A[c0][j] = A[c0][j] * id;
    } else {
#pragma unroll
      for (int j = c0; j < 8; j++) A[c0][j] = A[c0][j] / d;
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (r == c0) continue;
      const T f = A[r][c0];
#pragma unroll
      for (int j = c0; j < 8; j++) A[r][j] = A[r][j] - f * A[c0][j];
    }
  }
  // O_A = P^-1 N with N = I - h/2 K = 2 I - P, i.e. O_A = 2 P^-1 - I (no product needed)
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) OA[4 * i + j] = T(2) * A[i][4 + j] - (i == j ? T(1) : T(0));
  const T jv[4] = {c.jI, c.jQ, c.jU, c.jV};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    T s = T(0);
#pragma unroll
    for (int k = 0; k < 4; k++) s += A[i][4 + k] * jv[k];
    OS[i] = h * s;
  }
}

// ------------------------------------------------------------------ camera
// RAPTOR camera photon in double (src/core/geometry/camera.hpp:37-135)
struct Cam {
  double a, hs, incl, r_cam, t_init, phi_cam, fov_x, fov_y, cos_pa, sin_pa;
  int W, H;
};
PHX_DEV double sgnd(double x) { return (x < 0.) ? -1. : (x > 0.) ? 1. : 0.; }
PHX_DEV void camera_photon(const Cam& C, double alpha, double beta, double ph[8]) {
  const double a = C.a, rcam = C.r_cam;
  const double mu0 = cos(C.incl / 180. * PI_D);
  const double th0 = acos(mu0);
  const double phi0 = C.phi_cam / 180. * PI_D;
  const double ll = -alpha * sqrt(1. - mu0 * mu0);
  const double qq = beta * beta + mu0 * mu0 * (alpha * alpha - 1.);
  const double sint = sin(th0), sinc2 = sint * sint, cost = cos(th0), cosc2 = cost * cost;
  const double kd0 = -1., kd3 = ll;
  const double kd2 = sgnd(beta) * sqrt(fabs(qq - ll * ll * (cosc2 / sinc2) + cosc2));
  const double r = rcam;
  const double sig = r * r + a * a * cost * cost;
  const double del = r * r + a * a - 2. * r;
  const double A_ = (r * r + a * a) * (r * r + a * a) - del * a * a * sint * sint;
  const double g11 = sig / del * r * r;
  const double g00 = -A_ / (sig * del), g03 = -2. * a * r / (sig * del);
  const double g33 = (del - a * a * sint * sint) / (sig * del * sint * sint), g22 = 1. / sig;
  const double ku0 = g00 * kd0 + g03 * kd3;
  const double ku3 = g33 * kd3 + g03 * kd0;
  const double ku2 = g22 * kd2;
  const double ku1 = sqrt((-ku0 * kd0 - ku2 * kd2 - ku3 * kd3) / g11);
  const double lr = log(rcam), rp = exp(lr);
  const double Dl = rp * rp - 2. * rp + a * a;
  const double sq = sqrt(1. - a * a);
  const double lg = log((rp - (1. + sq)) / (rp - (1. - sq)));
  ph[0] = C.t_init + (log(Dl) + lg / sq);
  ph[1] = lr;
  ph[3] = phi0 + a / (2. * sq) * lg;
  ph[4] = ku0 + 2. * rp / Dl * rp * ku1;
  ph[5] = ku1;
  ph[7] = ku3 + a / Dl * rp * ku1;
  // theta -> x2 and k^theta -> k^x2 by Newton (camera.hpp:37-61)
  double cur = th0, prev = 1.e-15;
  for (int it = 0; it < 100 && fabs(cur - prev) > 1.e-15; it++) {
    prev = cur;
    cur = cur - (cur + 0.5 * C.hs * sin(2. * cur) - th0) / (1. + C.hs * cos(2. * cur));
  }
  ph[2] = cur;
  const double dd = 1. + C.hs * cos(2. * cur);
  double ck = ku2;
  prev = 1.e-15;
  for (int it = 0; it < 100 && fabs(ck - prev) > 1.e-15; it++) {
    prev = ck;
    ck = ck - (ck * dd - ku2) / dd;
  }
  ph[6] = ck;
}
// camera polarization basis F1 = e_(1) of the camera observer with up = (0,0,0,-1) (tetrad.hpp:27-133)
template <class T>
PHX_DEV void camera_F1(T a, T hs, const double ph[8], T F1[4]) {
  T k[4];
#pragma unroll
  for (int i = 0; i < 4; i++) k[i] = T(ph[4 + i]);
  const Met<T> g = metric(a, hs, T(ph[1]), T(ph[2]));
  const MetU<T> u = metric_up(a, g);
  const T B = T(0);  // -g^{03} U_0 / g^{33} with g^{03} = 0 in Kerr-Schild
  const T C = -(T(1) + u.u00) / u.u33;
  const T U3d = B + sqrt(B * B + C);
  const T U[4] = {-u.u00, -u.u01 + u.u13 * U3d, T(0), u.u33 * U3d};  // raise (-1, 0, 0, U3d)
  const T up[4] = {T(0), T(0), T(0), T(-1)};
  T para[4];
  frame(g, a, k, U, up, F1, para);
}

}  // namespace phxd
