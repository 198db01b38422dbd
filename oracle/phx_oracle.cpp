// ============================================================================
// ORACLE — test infrastructure only.  A CPU restatement of PHOENIX's polarized
// GRRT raymarch (the reference's `core/` + `host/io/bhac_io.hpp` + the CPU
// pixel loop of `host/render/jobs.hpp`), used by tests/, by the smoke check and
// by bench.py's CPU-baseline leg as the *checker*.  Nothing in the product
// (phoenix_amd/) links, loads or calls this file.
//
// Arithmetic order follows the reference expression by expression, so that a
// build with the reference's own flags (g++ -O3, x86-64 baseline, glibc libm)
// reproduces its FP32 and FP64 CPU results bit for bit.  Every function cites
// the reference file:line it restates.
//
// The scalar type is a template parameter: float (FP32 build), double (FP64
// build) or Counted<B>, which runs the B arithmetic while counting operations
// per phase (ray setup / geodesic step / plasma sample / sample x frequency).
// The counts give the algorithmic work per unit of the roofline.
//
// Pins (tests/test_oracle_*.py): bit/1e-12 agreement with the reference CPU
// build (oracle/_ref) on the synthetic snapshots; metric inverse, Christoffel
// symbols against finite differences of the metric, null/Carter conservation
// along rays, Bessel ratios against scipy.special.kve, thermal coefficients
// against closed forms, the trapezoid operator against scipy.linalg.expm and
// closed forms, AMR lookup against brute force, decode round trip.
// ============================================================================
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "oracle_bessel_table.inc"

namespace orc {

// ----------------------------------------------------------------------------
// operation counting
// ----------------------------------------------------------------------------
enum { PH_SETUP = 0, PH_STEP = 1, PH_SAMPLE = 2, PH_FREQ = 3, NPH = 4 };
enum { OP_ADD = 0, OP_MUL, OP_DIV, OP_SQRT, OP_TRANS, OP_CMP, NOPS };
thread_local long long tl_ops[NPH][NOPS];
thread_local int tl_phase = 0;

template <class B>
struct Counted {
  B v;
  Counted() = default;
  Counted(double x) : v((B)x) {}
  explicit operator double() const { return (double)v; }
  explicit operator float() const { return (float)v; }
  explicit operator int() const { return (int)v; }
};
template <class B> static inline void tick(int op) { tl_ops[tl_phase][op]++; }
#define CNT_BIN(OPR, KIND)                                                                              \
  template <class B> static inline Counted<B> operator OPR(Counted<B> a, Counted<B> b) {                \
    tick<B>(KIND); Counted<B> r; r.v = a.v OPR b.v; return r; }                                         \
  template <class B> static inline Counted<B>& operator OPR##=(Counted<B>& a, Counted<B> b) {           \
    tick<B>(KIND); a.v = a.v OPR b.v; return a; }
CNT_BIN(+, OP_ADD)
CNT_BIN(-, OP_ADD)
CNT_BIN(*, OP_MUL)
CNT_BIN(/, OP_DIV)
#undef CNT_BIN
template <class B> static inline Counted<B> operator-(Counted<B> a) { Counted<B> r; r.v = -a.v; return r; }
#define CNT_CMP(OPR)                                                                                    \
  template <class B> static inline bool operator OPR(Counted<B> a, Counted<B> b) {                      \
    tick<B>(OP_CMP); return a.v OPR b.v; }
CNT_CMP(<) CNT_CMP(>) CNT_CMP(<=) CNT_CMP(>=) CNT_CMP(==) CNT_CMP(!=)
#undef CNT_CMP

// math entry points: float/double go to glibc, Counted<B> counts and forwards
#define NM_UN(NAME, KIND)                                                                               \
  static inline float NAME(float x) { return std::NAME(x); }                                           \
  static inline double NAME(double x) { return std::NAME(x); }                                         \
  template <class B> static inline Counted<B> NAME(Counted<B> x) {                                      \
    tick<B>(KIND); Counted<B> r; r.v = std::NAME(x.v); return r; }
NM_UN(exp, OP_TRANS) NM_UN(log, OP_TRANS) NM_UN(sin, OP_TRANS) NM_UN(cos, OP_TRANS)
NM_UN(cbrt, OP_TRANS) NM_UN(expm1, OP_TRANS) NM_UN(tanh, OP_TRANS) NM_UN(acos, OP_TRANS)
NM_UN(acosh, OP_TRANS) NM_UN(sinh, OP_TRANS) NM_UN(sqrt, OP_SQRT) NM_UN(floor, OP_CMP)
#undef NM_UN
static inline float fabs_(float x) { return std::fabs(x); }
static inline double fabs_(double x) { return std::fabs(x); }
template <class B> static inline Counted<B> fabs_(Counted<B> x) { Counted<B> r; r.v = std::fabs(x.v); return r; }
static inline float fmin_(float a, float b) { return std::fmin(a, b); }
static inline double fmin_(double a, double b) { return std::fmin(a, b); }
template <class B> static inline Counted<B> fmin_(Counted<B> a, Counted<B> b) {
  tick<B>(OP_CMP); Counted<B> r; r.v = std::fmin(a.v, b.v); return r; }
static inline float fmod_(float a, float b) { return std::fmod(a, b); }
static inline double fmod_(double a, double b) { return std::fmod(a, b); }
template <class B> static inline Counted<B> fmod_(Counted<B> a, Counted<B> b) {
  tick<B>(OP_TRANS); Counted<B> r; r.v = std::fmod(a.v, b.v); return r; }
static inline bool finite_(float x) { return std::isfinite(x); }
static inline bool finite_(double x) { return std::isfinite(x); }
template <class B> static inline bool finite_(Counted<B> x) { return std::isfinite(x.v); }
template <class T> static inline void sincos_(T x, T* s, T* c) { *s = sin(x); *c = cos(x); }

template <class T> struct base_of { using type = T; };
template <class B> struct base_of<Counted<B>> { using type = B; };
template <class T> constexpr bool is_fp32() { return sizeof(typename base_of<T>::type) == sizeof(float); }
template <class T> static inline double dbl(T x) { return (double)x; }

template <class T> struct lim;
template <class T> struct lim {
  static T step_tiny() { return is_fp32<T>() ? T(1e-30f) : T(1e-160); }
  static T step_big() { return is_fp32<T>() ? T(1e30f) : T(1e100); }
  static T tol_series() { return is_fp32<T>() ? T(1e-9f) : T(1e-18); }
  static T tol_asym() { return is_fp32<T>() ? T(1e-8f) : T(1e-17); }
};

// constants (src/core/numerics/constants.hpp:22-81)
namespace cgs {
constexpr double MP = 1.6726219e-24, ME = 9.1093829e-28, QE = 4.80320425e-10, KB = 1.3806488e-16;
constexpr double C = 2.99792458e10, H = 6.62606885e-27, G = 6.674e-8, MSUN = 1.989e33, KPC = 3.086e21;
constexpr double JY = 1e23;
constexpr double H_OVER_MEC2 = H / (ME * C * C);
constexpr double NU_CRIT = 3. * QE / (4. * M_PI * ME * C);
constexpr double JNU = QE * QE / (2. * 1.73205080756887729353 * C);
constexpr double WP2 = 4. * M_PI * QE * QE / ME;
constexpr double GYRO = QE / (ME * C);
}  // namespace cgs

// ----------------------------------------------------------------------------
// model
// ----------------------------------------------------------------------------
constexpr int MAXF = 32;
struct Params {  // what config.toml + the grid file give (src/core/model/bhac.hpp:50-129)
  double inclination, fov_x, fov_y, r_cam, t_init, pa, phi_cam;
  int img_width, img_height, nfreqs;
  double freqs[MAXF];
  double mass, distance;
  double stepsize;
  int max_steps;
  double horizon_margin;
  int max_order;
  double m_unit, r_low, r_high, beta_c, sigma_cut, thetae_min, thetae_max, rt_outer;
  int nspin, bpol, nlevels, sfc;
};

template <class T>
struct Model {
  Params up;
  // grid
  int nxlone[3];
  double xprobmin[3], xprobmax[3], hslope_d;
  int nleafs, ndim, nwini, nws, cells, nx[3], ng[3], nlevels;
  int level_dims[16][3];
  std::vector<int> level_map[16];
  T lvl_w[16][3];
  T x_lo[4], x_hi[4];
  std::vector<T> lb, dxc, prim;
  std::vector<int> leaf_level, leaf_ind;
  // scalars
  T a, hslope, gam, Ne, Bu, rhigh, rlow, ibc;
  T stepsize, sigma_cut, thetae_min, thetae_max, rt_outer;
  double L, B_unit_d, Ne_unit_d, cos_pa, sin_pa;
  T r_out, cut_in;
  T freq[MAXF], hconst[MAXF], jscale[MAXF];
};

// grid file: last token of each non-empty line (src/host/io/bhac_io.hpp:143-174)
template <class T>
static void load_grid(Model<T>* m, const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open grid file " + path);
  std::vector<double> v;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string tok, last;
    while (ss >> tok) last = tok;
    if (!last.empty()) v.push_back(std::atof(last.c_str()));
  }
  if (v.size() < 10) throw std::runtime_error(path + ": expected 10 values");
  for (int d = 0; d < 3; d++) {
    m->nxlone[d] = (int)v[d];
    m->xprobmin[d] = v[3 + d];
    m->xprobmax[d] = v[6 + d];
  }
  m->hslope_d = v[9];
}

// units and per-run scalars (src/host/io/bhac_io.hpp:183-251)
template <class T>
static void finalize(Model<T>* m, const std::vector<double>& neqpar) {
  const Params& u = m->up;
  const double spin = neqpar[u.nspin];
  if (std::fabs(spin) > 1.) throw std::runtime_error("spin unphysical");
  m->a = T(spin);
  m->hslope = T(m->hslope_d);
  m->gam = T(neqpar[0]);
  for (int d = 1; d < 3; d++) {
    m->xprobmin[d] *= 2. * M_PI;
    m->xprobmax[d] *= 2. * M_PI;
  }
  m->cells = m->nx[0] * m->nx[1] * m->nx[2];
  for (int d = 0; d < 3; d++) {
    if (m->nxlone[d] % m->nx[d] != 0) throw std::runtime_error("nxlone is not a multiple of the block size");
    m->ng[d] = m->nxlone[d] / m->nx[d];
    m->x_lo[1 + d] = T(m->xprobmin[d]);
    m->x_hi[1 + d] = T(m->xprobmax[d]);
  }
  m->x_lo[0] = m->x_hi[0] = T(0);
  for (int l = 0; l < 16; l++)
    for (int d = 0; d < 3; d++)
      m->lvl_w[l][d] = T((m->xprobmax[d] - m->xprobmin[d]) / m->ng[d] / (double)(1 << l));
  m->L = cgs::G * u.mass * cgs::MSUN / cgs::C / cgs::C;
  const double rho = u.m_unit / std::pow(m->L, 3);
  m->B_unit_d = cgs::C * std::sqrt(4. * M_PI * rho);
  m->Ne_unit_d = rho / (cgs::MP + cgs::ME);
  m->Bu = T(m->B_unit_d);
  m->Ne = T(m->Ne_unit_d);
  m->rhigh = T(u.r_high);
  m->rlow = T(u.r_low);
  m->ibc = T(1. / u.beta_c);
  m->cos_pa = std::cos(u.pa / 180. * M_PI);
  m->sin_pa = std::sin(u.pa / 180. * M_PI);
  m->cut_in = T((1. + std::sqrt(1. - spin * spin)) * (1. + u.horizon_margin));
  m->r_out = T(1.1 * u.r_cam);
  const double cc = m->L * cgs::H / (cgs::ME * cgs::C * cgs::C);
  for (int f = 0; f < u.nfreqs; f++) {
    const double nu = u.freqs[f];
    m->freq[f] = T(nu);
    m->hconst[f] = T((cgs::ME * cgs::C * cgs::C) / (cgs::H * nu) * cc);
    m->jscale[f] = T(nu * nu * nu);
  }
  m->stepsize = T(u.stepsize);
  m->sigma_cut = T(u.sigma_cut);
  m->thetae_min = T(u.thetae_min);
  m->thetae_max = T(u.thetae_max);
  m->rt_outer = T(u.rt_outer);
}

// ----------------------------------------------------------------------------
// metric (src/core/geometry/metric.hpp)
// ----------------------------------------------------------------------------
template <class T>
struct Trig {  // r, sin/cos of theta and of 2 x2 shared by the three metric forms
  T r, s2x, c2x, s, c;
};
template <class T>
static inline Trig<T> trig(T a, T hs, const T X[4]) {
  (void)a;
  Trig<T> t;
  t.r = exp(X[1]);
  sincos_(T(2) * X[2], &t.s2x, &t.c2x);
  const T th = X[2] + T(0.5) * hs * t.s2x;
  sincos_(th, &t.s, &t.c);
  return t;
}
template <class T>
static inline void zero44(T g[4][4]) {
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) g[i][j] = T(0);
}
// covariant metric (metric.hpp:26-54)
template <class T>
static void gcov(T a, T hs, const T X[4], T g[4][4]) {
  const Trig<T> t = trig(a, hs, X);
  const T one = T(1), two = T(2), r = t.r;
  const T s2 = t.s * t.s, rho2 = r * r + a * a * t.c * t.c, hf = one + hs * t.c2x;
  zero44(g);
  g[0][0] = -one + two * r / rho2;
  g[0][1] = g[1][0] = (two * r / rho2) * r;
  g[0][3] = g[3][0] = -two * a * r * s2 / rho2;
  g[1][1] = (one + two * r / rho2) * r * r;
  g[1][3] = g[3][1] = -a * s2 * (one + two * r / rho2) * r;
  g[2][2] = rho2 * hf * hf;
  g[3][3] = s2 * (rho2 + a * a * s2 * (one + two * r / rho2));
}
// contravariant metric (metric.hpp:64-91)
template <class T>
static void gcon(T a, T hs, const T X[4], T g[4][4]) {
  const Trig<T> t = trig(a, hs, X);
  const T one = T(1), two = T(2), r = t.r;
  const T s2 = t.s * t.s, rho2 = r * r + a * a * t.c * t.c, hf = one + hs * t.c2x, ir = one / rho2;
  zero44(g);
  g[0][0] = -one - two * r * ir;
  g[0][1] = g[1][0] = two * ir;
  g[1][1] = ir * (r * (r - two) + a * a) / (r * r);
  g[1][3] = g[3][1] = a * ir / r;
  g[2][2] = ir / (hf * hf);
  g[3][3] = ir / s2;
}
// both at once (metric.hpp:103-142); note g^00 is formed with a division here
template <class T>
static void gboth(T a, T hs, const T X[4], T gd[4][4], T gu[4][4]) {
  const Trig<T> t = trig(a, hs, X);
  const T one = T(1), two = T(2), r = t.r;
  const T s2 = t.s * t.s, rho2 = r * r + a * a * t.c * t.c, ir = one / rho2, hf = one + hs * t.c2x;
  zero44(gd);
  zero44(gu);
  gd[0][0] = -one + two * r / rho2;
  gd[0][1] = gd[1][0] = (two * r / rho2) * r;
  gd[0][3] = gd[3][0] = -two * a * r * s2 / rho2;
  gd[1][1] = (one + two * r / rho2) * r * r;
  gd[1][3] = gd[3][1] = -a * s2 * (one + two * r / rho2) * r;
  gd[2][2] = rho2 * hf * hf;
  gd[3][3] = s2 * (rho2 + a * a * s2 * (one + two * r / rho2));
  gu[0][0] = -one - two * r / rho2;
  gu[0][1] = gu[1][0] = two * ir;
  gu[1][1] = ir * (r * (r - two) + a * a) / (r * r);
  gu[1][3] = gu[3][1] = a * ir / r;
  gu[2][2] = ir / (hf * hf);
  gu[3][3] = ir / s2;
}
template <class T>
static inline T dot(const T g[4][4], const T* A, const T* B) {  // metric.hpp:147-160
  T s = T(0);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) s += g[i][j] * A[i] * B[j];
  return s;
}
template <class T>
static inline void lower(const T g[4][4], const T V[4], T out[4]) {  // metric.hpp:165-176
  for (int i = 0; i < 4; i++) {
    out[i] = T(0);
    for (int j = 0; j < 4; j++) out[i] += g[i][j] * V[j];
  }
}
template <class T>
static inline T det4(const T m[4][4]) {  // 2x2 minors of rows (0,1) and (2,3); metric.hpp:197-213
  const T s0 = m[2][0] * m[3][1] - m[2][1] * m[3][0];
  const T s1 = m[2][0] * m[3][2] - m[2][2] * m[3][0];
  const T s2 = m[2][0] * m[3][3] - m[2][3] * m[3][0];
  const T s3 = m[2][1] * m[3][2] - m[2][2] * m[3][1];
  const T s4 = m[2][1] * m[3][3] - m[2][3] * m[3][1];
  const T s5 = m[2][2] * m[3][3] - m[2][3] * m[3][2];
  const T c0 = m[0][0] * m[1][1] - m[0][1] * m[1][0];
  const T c1 = m[0][0] * m[1][2] - m[0][2] * m[1][0];
  const T c2 = m[0][0] * m[1][3] - m[0][3] * m[1][0];
  const T c3 = m[0][1] * m[1][2] - m[0][2] * m[1][1];
  const T c4 = m[0][1] * m[1][3] - m[0][3] * m[1][1];
  const T c5 = m[0][2] * m[1][3] - m[0][3] * m[1][2];
  return c0 * s5 - c1 * s4 + c2 * s3 + c3 * s2 - c4 * s1 + c5 * s0;
}

// ----------------------------------------------------------------------------
// decode: conserved -> primitive in FP64 (src/core/model/bhac.hpp:178-346)
// ----------------------------------------------------------------------------
static double sqrt_det_gamma(double a, double hs, double x, double y, double z) {  // bhac.hpp:178-192
  const double X[4] = {0., x, y, z};
  double g[4][4];
  gcov<double>(a, hs, X, g);
  const double det = g[1][1] * (g[2][2] * g[3][3] - g[3][2] * g[2][3]) -
                     g[1][2] * (g[2][1] * g[3][3] - g[3][1] * g[2][3]) +
                     g[1][3] * (g[2][1] * g[3][2] - g[2][2] * g[3][1]);
  return sqrt(det);
}
static void centroid(double a, double hs, const double X[3], const double dx[3], double out[3]) {  // bhac.hpp:202-235
  static const double wt[3] = {1., 4., 1.}, off[3] = {-1., 0., 1.};
  double nrm = 0., acc[3] = {0., 0., 0.};
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        const double w = wt[i] * wt[j] * wt[k];
        const double p[3] = {X[0] + dx[0] / 2. * off[i], X[1] + dx[1] / 2. * off[j], X[2] + dx[2] / 2. * off[k]};
        const double dg = sqrt_det_gamma(a, hs, p[0], p[1], p[2]);
        nrm += dg * w;
        for (int d = 0; d < 3; d++) acc[d] += dg * w * p[d];
      }
  const double vol = (dx[0] / 6.0) * (dx[1] / 6.0) * (dx[2] / 6.0);
  nrm *= vol;
  for (int d = 0; d < 3; d++) out[d] = acc[d] * (vol / nrm);
}
template <class T>
static void cons_to_prim(const Model<T>* m, int leaf, const double* cons, int cell, T* prim) {
  const int cells = m->cells, ilf = m->nwini - 2, ixi = m->nwini - 1;
  auto CV = [&](int nw) { return cons[(size_t)nw * cells + cell]; };
  const int loc[3] = {cell % m->nx[0], (cell / m->nx[0]) % m->nx[1], cell / (m->nx[0] * m->nx[1])};
  double X[3], dx[3];
  for (int d = 0; d < 3; d++) {
    dx[d] = dbl(m->dxc[3 * leaf + d]);
    X[d] = dbl(m->lb[3 * leaf + d]) + (loc[d] + 0.5) * dbl(m->dxc[3 * leaf + d]);
  }
  for (int v = 0; v < 8; v++) prim[v] = T(0);
  if (!(std::exp(X[0]) > 1.0)) return;  // bhac.hpp:277
  double Xb[3];
  centroid(dbl(m->a), dbl(m->hslope), X, dx, Xb);
  const double Xu[4] = {0., Xb[0], Xb[1], Xb[2]};
  double gd[4][4], gu[4][4];
  gboth<double>(dbl(m->a), dbl(m->hslope), Xu, gd, gu);
  const double BS = CV(1) * CV(5) + CV(2) * CV(6) + CV(3) * CV(7);
  double Bsq = 0., Bd[4] = {0, 0, 0, 0}, Su[4] = {0, 0, 0, 0}, gam3[4][4];
  for (int i = 1; i < 4; i++)
    for (int j = 1; j < 4; j++) gam3[i][j] = gu[i][j] + gu[0][i] * gu[0][j] / (-gu[0][0]);
  for (int j = 1; j < 4; j++)
    for (int i = 1; i < 4; i++) {
      Su[j] += gam3[i][j] * CV(1 + i - 1);
      Bd[j] += gd[i][j] * CV(5 + i - 1);
    }
  for (int i = 1; i < 4; i++) Bsq += Bd[i] * CV(5 + i - 1);
  const double g = dbl(m->gam), lf = CV(ilf), xi = CV(ixi);
  const double rho = CV(0) / lf;
  double uu = (g - 1.) / g * (xi / std::pow(lf, 2.) - rho) / (g - 1.);
  if (uu < 0) {  // entropy fallback
    uu = (CV(8) / CV(0)) * std::pow(rho, g - 1) / (g - 1.);
    if (uu < 0) uu = 1e-14;
  }
  const double den = xi + Bsq;
  double out[8];
  out[0] = rho;
  out[1] = uu;
  for (int i = 0; i < 3; i++) out[2 + i] = (Su[1 + i] / den + CV(5 + i) * BS / (xi * den)) * lf;
  for (int i = 0; i < 3; i++) out[5 + i] = m->up.bpol * CV(5 + i);
  for (int v = 0; v < 8; v++) prim[v] = T(out[v]);
}

// ----------------------------------------------------------------------------
// BHAC dump reader (src/host/io/bhac_io.hpp:83-130, 257-499)
// ----------------------------------------------------------------------------
static void child_of(int c, int ip, int jp, int kp, int* i, int* j, int* k) {  // bhac_io.hpp:39-66 (3D)
  *i = 2 * ip + c % 2;
  *j = 2 * jp + ((int)(c / 2.)) % 2;
  *k = 2 * kp + (int)(c / 4.);
}
static void walk(FILE* f, int level, int i, int j, int k, std::vector<int>& ind, std::vector<int>& lev, bool& ok) {
  if (!ok) return;
  int leaf = 0;
  if (std::fread(&leaf, 4, 1, f) != 1) {
    ok = false;
    return;
  }
  if (leaf) {
    ind.push_back(i);
    ind.push_back(j);
    ind.push_back(k);
    lev.push_back(level);
    return;
  }
  for (int c = 0; c < 8; c++) {
    int ci, cj, ck;
    child_of(c, i, j, k, &ci, &cj, &ck);
    walk(f, level + 1, ci, cj, ck, ind, lev, ok);
  }
}
static uint64_t interleave3(unsigned a, unsigned b, unsigned c) {  // bhac_io.hpp:123-130
  uint64_t z = 0;
  for (uint64_t i = 0; i < 21; ++i)
    z |= (((uint64_t)a & (1ull << i)) << (2 * i)) | (((uint64_t)b & (1ull << i)) << (2 * i + 1)) |
         (((uint64_t)c & (1ull << i)) << (2 * i + 2));
  return z;
}

template <class T>
static void load_model(Model<T>* m, const char* dump, const char* grid) {
  load_grid(m, grid);
  FILE* f = std::fopen(dump, "rb");
  if (!f) throw std::runtime_error(std::string("cannot open ") + dump);
  struct Closer {
    FILE* f;
    ~Closer() { std::fclose(f); }
  } closer{f};
  int hdr[8];
  double tdump;
  std::fseek(f, -40, SEEK_END);
  if (std::fread(hdr, 4, 8, f) != 8 || std::fread(&tdump, 8, 1, f) != 1) throw std::runtime_error("bad header");
  m->nleafs = hdr[0];
  m->ndim = hdr[2];
  m->nwini = hdr[4];
  m->nws = hdr[5];
  const int neq = hdr[6];
  if (m->ndim != 3) throw std::runtime_error("only 3D dumps");
  std::fseek(f, -40 - (m->ndim * 4 + neq * 8), SEEK_END);
  std::vector<double> neqpar(neq);
  if (std::fread(m->nx, 4, 3, f) != 3 || std::fread(neqpar.data(), 8, neq, f) != (size_t)neq)
    throw std::runtime_error("short header");
  if (m->up.nspin < 0 || m->up.nspin >= neq) throw std::runtime_error("bad nspin");
  finalize(m, neqpar);
  const long blk = (long)m->cells * m->nwini * 8;
  const long stag = (long)(m->nx[0] + 1) * (m->nx[1] + 1) * (m->nx[2] + 1) * m->nws * 8;
  std::fseek(f, (long)m->nleafs * (blk + stag), SEEK_SET);
  std::vector<int> ind, lev;
  bool ok = true;
  if (m->up.sfc) {
    std::vector<std::pair<uint64_t, int>> order;
    for (int i = 0; i < m->ng[0]; i++)
      for (int j = 0; j < m->ng[1]; j++)
        for (int k = 0; k < m->ng[2]; k++) order.push_back({interleave3(i, j, k), (i * m->ng[1] + j) * m->ng[2] + k});
    std::sort(order.begin(), order.end());
    for (auto& o : order) {
      const int n = o.second;
      walk(f, 1, n / (m->ng[1] * m->ng[2]), (n / m->ng[2]) % m->ng[1], n % m->ng[2], ind, lev, ok);
    }
  } else {
    for (int k = 0; k < m->ng[2]; k++)
      for (int j = 0; j < m->ng[1]; j++)
        for (int i = 0; i < m->ng[0]; i++) walk(f, 1, i, j, k, ind, lev, ok);
  }
  if (!ok) throw std::runtime_error("short read in the tree");
  if ((int)lev.size() != m->nleafs) throw std::runtime_error("grid mismatch");
  m->leaf_level = lev;
  m->leaf_ind = ind;
  m->lb.resize(3 * (size_t)m->nleafs);
  m->dxc.resize(3 * (size_t)m->nleafs);
  double dx1[3], dxc0[3];
  for (int d = 0; d < 3; d++) {
    dx1[d] = (m->xprobmax[d] - m->xprobmin[d]) / m->ng[d];
    dxc0[d] = (m->xprobmax[d] - m->xprobmin[d]) / m->nxlone[d];
  }
  for (int q = 0; q < m->nleafs; q++) {
    const double inv = 1.0 / (double)(1 << (lev[q] - 1));
    for (int d = 0; d < 3; d++) {
      m->lb[3 * q + d] = T(m->xprobmin[d] + ind[3 * q + d] * dx1[d] * inv);
      m->dxc[3 * q + d] = T(dxc0[d] * inv);
    }
  }
  // level map (bhac_io.hpp:257-309)
  m->nlevels = m->up.nlevels;
  if (m->nlevels == -1) {
    for (int q = 0; q < m->nleafs; q++) m->nlevels = std::max(m->nlevels, lev[q]);
    if (m->nlevels > 16) throw std::runtime_error("too many levels");
  }
  for (int l = 0; l < m->nlevels; l++) {
    for (int d = 0; d < 3; d++) m->level_dims[l][d] = m->ng[d] * (1 << l);
    const size_t n = (size_t)m->level_dims[l][0] * m->level_dims[l][1] * m->level_dims[l][2];
    m->level_map[l].assign(n, -1);
    for (int q = 0; q < m->nleafs; q++) {
      if (lev[q] - 1 != l) continue;
      const int i = ind[3 * q], j = ind[3 * q + 1], k = ind[3 * q + 2];
      if (i < 0 || i >= m->level_dims[l][0] || j < 0 || j >= m->level_dims[l][1] || k < 0 ||
          k >= m->level_dims[l][2])
        continue;
      m->level_map[l][((size_t)i * m->level_dims[l][1] + j) * m->level_dims[l][2] + k] = q;
    }
  }
  // host decode (bhac_io.hpp:456-491)
  m->prim.assign((size_t)m->nleafs * m->cells * 8, T(0));
  const size_t per_leaf = (size_t)m->nwini * m->cells;
  std::vector<double> raw(per_leaf * (size_t)m->nleafs);
  std::fseek(f, 0, SEEK_SET);
  for (int q = 0; q < m->nleafs; q++) {
    if (std::fread(raw.data() + (size_t)q * per_leaf, 8, per_leaf, f) != per_leaf)
      throw std::runtime_error("short read in the data");
    std::fseek(f, stag, SEEK_CUR);
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (int q = 0; q < m->nleafs; q++)
    for (int c = 0; c < m->cells; c++)
      cons_to_prim(m, q, raw.data() + (size_t)q * per_leaf, c, &m->prim[((size_t)q * m->cells + c) * 8]);
}

// ----------------------------------------------------------------------------
// AMR sampler (src/core/model/sampler.hpp)
// ----------------------------------------------------------------------------
template <class T>
static inline void wrap(T x[4]) {  // sampler.hpp:29-38
  const T TP = T(2. * M_PI), PI = T(M_PI);
  x[3] = fmod_(x[3], TP);
  x[2] = fmod_(x[2], PI) - T(1e-6);
  if (x[3] < T(0)) x[3] = TP + x[3];
  if (x[2] < T(0)) {
    x[2] = -x[2];
    x[3] = PI + x[3];
  }
}
template <class T>
static inline bool off_domain(const Model<T>* m, const T x[4]) {  // sampler.hpp:43-50
  return x[1] > m->x_hi[1] || x[1] < m->x_lo[1] || x[2] < m->x_lo[2] || x[2] > m->x_hi[2] || x[3] < m->x_lo[3] ||
         x[3] > m->x_hi[3];
}
template <class T>
static inline T nudge(const T x[4]) {  // +-1e-9 hemisphere nudge of the lookups
  T s = T(1e-9);
  if (x[2] > T(M_PI / 2.)) s = -s;
  return s;
}
template <class T>
static int map_lookup(const Model<T>* m, const T x[4]) {  // sampler.hpp:55-82
  const T small = nudge(x);
  const int top = m->nlevels - 1;
  const T up = T(1 << top);
  const T fi = (x[1] + small - m->x_lo[1]) / m->lvl_w[0][0] * up;
  const T fj = (x[2] + small - m->x_lo[2]) / m->lvl_w[0][1] * up;
  const T fk = (x[3] + small - m->x_lo[3]) / m->lvl_w[0][2] * up;
  if (fi < T(0) || fj < T(0) || fk < T(0)) return -1;
  const int i = (int)fi, j = (int)fj, k = (int)fk;
  if (i >= m->level_dims[top][0] || j >= m->level_dims[top][1] || k >= m->level_dims[top][2]) return -1;
  for (int l = 0; l < m->nlevels; l++) {
    const int s = top - l;
    const int q = m->level_map[l][((size_t)(i >> s) * m->level_dims[l][1] + (j >> s)) * m->level_dims[l][2] + (k >> s)];
    if (q >= 0) return q;
  }
  return -1;
}
template <class T>
static inline void leaf_box(const Model<T>* m, int q, T lo[3], T hi[3]) {
  for (int d = 0; d < 3; d++) {
    lo[d] = m->lb[3 * q + d];
    hi[d] = lo[d] + T(m->nx[d]) * m->dxc[3 * q + d];
  }
}
template <class T>
static inline bool in_leaf(const Model<T>* m, int q, const T x[4], T small) {  // half-open; sampler.hpp:87-101
  T lo[3], hi[3];
  leaf_box(m, q, lo, hi);
  return !(x[1] + small < lo[0] || x[1] + small >= hi[0] || x[2] + small < lo[1] || x[2] + small >= hi[1] ||
           x[3] + small < lo[2] || x[3] + small >= hi[2]);
}
template <class T>
static inline bool in_leaf_closed(const Model<T>* m, int q, const T x[4]) {  // inclusive; sampler.hpp:109-122
  T lo[3], hi[3];
  leaf_box(m, q, lo, hi);
  return !(x[1] < lo[0] || x[1] > hi[0] || x[2] < lo[1] || x[2] > hi[1] || x[3] < lo[2] || x[3] > hi[2]);
}
template <class T>
static int find_leaf(const Model<T>* m, const T x[4]) {  // sampler.hpp:143-180
  const T small = nudge(x);
  const int cand = map_lookup(m, x);
  if (cand >= 0 && in_leaf(m, cand, x, small)) return cand;
  const T eps = T(1e-6);
  if (cand < 0 && (x[1] < m->x_lo[1] - eps || x[1] > m->x_hi[1] + eps || x[2] < m->x_lo[2] - eps ||
                   x[2] > m->x_hi[2] + eps || x[3] < m->x_lo[3] - eps || x[3] > m->x_hi[3] + eps))
    return -1;
  for (int q = 0; q < m->nleafs; q++)
    if (in_leaf(m, q, x, small)) return q;
  return -1;
}
template <class T>
static int leaf_at(const Model<T>* m, const T X[4]) {  // sampler.hpp:182-192
  T x[4] = {X[0], X[1], X[2], X[3]};
  wrap(x);
  if (exp(x[1]) < T(1)) return -1;
  if (off_domain(m, x)) return -1;
  return find_leaf(m, x);
}
template <class T>
static const T* cell_size_at(const Model<T>* m, const T X[4]) {  // sampler.hpp:198-206
  T x[4] = {X[0], X[1], X[2], X[3]};
  if (exp(x[1]) >= m->rt_outer) return nullptr;
  wrap(x);
  const int q = find_leaf(m, x);
  return q >= 0 ? &m->dxc[3 * q] : nullptr;
}

template <class T>
struct Fluid {
  T U[4], b[4], ne, B, te, X[4];
};

// derived fluid quantities (sampler.hpp:305-376)
template <class T>
static int fluid_from_prims(const Model<T>* m, const T x[4], T rho, T uu, const T Bp[4], const T gV[4], Fluid<T>* fl) {
  const T tiny = T(1.e-6);
  const T r = exp(x[1]);
  T gd[4][4], gu[4][4];
  gboth<T>(m->a, m->hslope, x, gd, gu);
  fl->ne = rho * m->Ne + tiny;
  T shift[4];
  for (int j = 1; j < 4; j++) shift[j] = gu[0][j] / (-gu[0][0]);
  const T alpha = T(1) / sqrt(-gu[0][0]);
  T vv = T(0);
  for (int i = 1; i < 4; i++)
    for (int j = 1; j < 4; j++) vv += gd[i][j] * gV[i] * gV[j];
  const T lf = sqrt(vv + T(1));
  T V[4];
  for (int i = 1; i < 4; i++) V[i] = gV[i] / lf;
  fl->U[0] = lf / alpha;
  for (int i = 1; i < 4; i++) fl->U[i] = V[i] * lf - shift[i] * lf / alpha;
  fl->b[0] = T(0);
  for (int i = 1; i < 4; i++)
    for (int l = 1; l < 4; l++) fl->b[0] += lf * Bp[i] * (gd[i][l] * V[l]) / alpha;
  for (int i = 1; i < 4; i++) fl->b[i] = (Bp[i] + alpha * fl->b[0] * fl->U[i]) / lf;
  T bd[4];
  lower(gd, fl->b, bd);
  const T Bsq = fabs_(fl->b[0] * bd[0] + fl->b[1] * bd[1] + fl->b[2] * bd[2] + fl->b[3] * bd[3]) + tiny;
  fl->B = sqrt(Bsq) * m->Bu;
  const T beta = uu * (m->gam - T(1)) / (T(0.5) * (Bsq + tiny));
  const T bb = beta * m->ibc;
  const T b2 = bb * bb;
  const T trat = m->rhigh * b2 / (T(1) + b2) + m->rlow / (T(1) + b2);
  const T te_unit = T(1. / 3. * (cgs::MP / cgs::ME)) / (trat + T(1));
  fl->te = (uu / rho) * te_unit;
  if ((Bsq / (rho + T(1e-20)) > m->sigma_cut) || r > m->rt_outer || fl->te > m->thetae_max || fl->te < m->thetae_min) {
    fl->ne = T(0);
    return 0;
  }
  return 1;
}

// sampler.hpp:387-427 with cell lookup (:127-137), weights (:211-236) and block-local trilinear (:260-291)
template <class T>
static int sample_fluid(const Model<T>* m, const T X[4], Fluid<T>* fl, int* cache) {
  T x[4] = {X[0], X[1], X[2], X[3]};
  wrap(x);
  for (int i = 0; i < 4; i++) fl->X[i] = x[i];
  const T r = exp(x[1]);
  if (r < T(1)) return 0;
  if (off_domain(m, x)) return 0;
  int q = cache ? *cache : -1;
  if (q == -1 || !in_leaf_closed(m, q, x)) {
    wrap(x);  // the second wrap of a cache miss is part of the reference's behaviour
    q = find_leaf(m, x);
    if (cache) *cache = q;
  }
  for (int i = 0; i < 4; i++) fl->X[i] = x[i];
  if (q == -1) return 0;
  int c[3];
  for (int d = 0; d < 3; d++) {
    c[d] = (int)((x[1 + d] - m->lb[3 * q + d]) / m->dxc[3 * q + d]);
    if (c[d] >= m->nx[d]) c[d] = m->nx[d] - 1;
  }
  T del[4];
  {
    T bs[3], dx[3];
    int ix[3];
    for (int d = 0; d < 3; d++) {
      bs[d] = m->lb[3 * q + d];
      dx[d] = m->dxc[3 * q + d];
    }
    for (int d = 0; d < 3; d++) {
      const T u = (x[1 + d] - bs[d]) / dx[d];
      if (!is_fp32<T>())
        ix[d] = (int)(u + T(1000)) - 1000;
      else
        ix[d] = (int)floor(u);
    }
    for (int d = 0; d < 3; d++)
      del[1 + d] = (ix[d] < 0) ? T(0) : (ix[d] > m->nx[d] - 2) ? T(1) : (x[1 + d] - (T(ix[d]) * dx[d] + bs[d])) / dx[d];
  }
  T qv[8];
  {
    const int nx0 = m->nx[0], nx1 = m->nx[1], nx2 = m->nx[2];
    int ip = c[0] + 1, jp = c[1] + 1, kp = c[2] + 1;
    if (ip >= nx0) ip = c[0];
    if (jp >= nx1) jp = c[1];
    if (kp >= nx2) kp = c[2];
    const T b1 = T(1) - del[1], b2 = T(1) - del[2], b3 = T(1) - del[3];
    const int ci[4] = {c[0], c[0], ip, ip}, cj[4] = {c[1], jp, c[1], jp};
    const T wq[4] = {b1 * b2, b1 * del[2], del[1] * b2, del[1] * del[2]};
    const size_t base = (size_t)q * m->cells;
    for (int a = 0; a < 8; a++) qv[a] = T(0);
    for (int z = 0; z < 2; z++) {
      const int ck = z ? kp : c[2];
      const T wz = z ? del[3] : b3;
      for (int s = 0; s < 4; s++) {
        const T* p = &m->prim[(base + (ci[s] + cj[s] * nx0 + ck * nx0 * nx1)) * 8];
        const T w = wq[s] * wz;
        for (int a = 0; a < 8; a++) qv[a] += p[a] * w;
      }
    }
  }
  const T Bp[4] = {T(0), qv[5], qv[6], qv[7]};
  const T gV[4] = {T(0), qv[2], qv[3], qv[4]};
  return fluid_from_prims(m, x, qv[0], qv[1], Bp, gV, fl);
}

// ----------------------------------------------------------------------------
// camera and tetrads (src/core/geometry/camera.hpp, tetrad.hpp)
// ----------------------------------------------------------------------------
static double sgn(double x) { return (x < 0.) ? -1. : (x > 0.) ? 1. : 0.; }
static double x2_of_theta(double hs, double th) {  // Newton inverse of th = x2 + hs/2 sin 2x2 (camera.hpp:37-46)
  double cur = th, prev = 1.e-15;
  for (int it = 0; it < 100 && std::fabs(cur - prev) > 1.e-15; it++) {
    prev = cur;
    cur = cur - (cur + 0.5 * hs * std::sin(2. * cur) - th) / (1. + hs * std::cos(2. * cur));
  }
  return cur;
}
static double kx2_of_ktheta(double hs, double kth, double x2) {  // camera.hpp:51-61
  double cur = kth, prev = 1.e-15;
  const double d = 1. + hs * std::cos(2. * x2);
  for (int it = 0; it < 100 && std::fabs(cur - prev) > 1.e-15; it++) {
    prev = cur;
    cur = cur - (cur * d - kth) / d;
  }
  return cur;
}
// RAPTOR-style camera photon (camera.hpp:71-135), always in double; spin and hslope
// are the model's (scalar-rounded) values
template <class T>
static void camera_photon(const Model<T>* m, double alpha, double beta, double ph[8]) {
  const double a = dbl(m->a), rcam = m->up.r_cam;
  const double mu0 = std::cos(m->up.inclination / 180. * M_PI);
  const double th0 = std::acos(mu0);
  const double phi0 = m->up.phi_cam / 180. * M_PI;
  const double Xc[4] = {m->up.t_init, std::log(rcam), th0, phi0};
  const double ll = -alpha * std::sqrt(1. - mu0 * mu0);
  const double qq = beta * beta + mu0 * mu0 * (alpha * alpha - 1.);
  const double sint = std::sin(th0), sinc2 = sint * sint, cost = std::cos(th0), cosc2 = cost * cost;
  const double kd0 = -1., kd3 = ll;
  const double kd2 = sgn(beta) * std::sqrt(std::fabs(qq - ll * ll * (cosc2 / sinc2) + cosc2));
  const double r = rcam, rf = r;
  const double sig = r * r + a * a * cost * cost;
  const double del = r * r + a * a - 2. * r;
  const double A_ = (r * r + a * a) * (r * r + a * a) - del * a * a * sint * sint;
  const double g11 = sig / del * rf * rf;
  const double g00 = -A_ / (sig * del), g03 = -2. * a * r / (sig * del);
  const double g33 = (del - a * a * sint * sint) / (sig * del * sint * sint), g22 = 1. / sig;
  const double ku0 = g00 * kd0 + g03 * kd3;
  const double ku3 = g33 * kd3 + g03 * kd0;
  const double ku2 = g22 * kd2;
  const double ku1 = std::sqrt((-ku0 * kd0 - ku2 * kd2 - ku3 * kd3) / g11);
  const double rp = std::exp(Xc[1]);
  const double Dl = rp * rp - 2. * rp + a * a;
  const double rpl = 1. + std::sqrt(1. - a * a), rmi = 1. - std::sqrt(1. - a * a);
  const double lg = std::log((rp - rpl) / (rp - rmi));
  ph[0] = Xc[0] + (std::log(Dl) + lg / std::sqrt(1. - a * a));
  ph[1] = Xc[1];
  ph[2] = Xc[2];
  ph[3] = Xc[3] + a / (2. * std::sqrt(1. - a * a)) * lg;
  ph[4] = ku0 + 2. * rp / Dl * rp * ku1;
  ph[5] = ku1;
  ph[6] = ku2;
  ph[7] = ku3 + a / Dl * rp * ku1;
  ph[2] = x2_of_theta(dbl(m->hslope), ph[2]);
  ph[6] = kx2_of_ktheta(dbl(m->hslope), ph[6], ph[2]);
}
// plasma/observer tetrad (tetrad.hpp:51-101); returns e_(1) and e_(2) columns
template <class T>
static void observer_frame(const Model<T>* m, const T X[4], const T k[4], const T U[4], const T b[4], T perp[4],
                           T para[4]) {
  T g[4][4];
  gcov<T>(m->a, m->hslope, X, g);
  const T om = -dot(g, k, U);
  T eK[4];
  for (int i = 0; i < 4; i++) eK[i] = k[i] / om - U[i];
  const T Be = dot(g, U, b);
  const T Cc = dot(g, k, b) / om - Be;
  const T b2 = dot(g, b, b);
  const T Nc = sqrt(b2 + Be * Be - Cc * Cc);
  for (int i = 0; i < 4; i++) para[i] = (b[i] + Be * U[i] - Cc * eK[i]) / Nc;
  T Ud[4], kd[4], bd[4];
  lower(g, U, Ud);
  lower(g, k, kd);
  lower(g, b, bd);
  const T sq = sqrt(-det4(g));
  const T s = T(-1) / (sq * om * Nc);
  static const int rest[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  for (int i = 0; i < 4; i++) {
    const int p = rest[i][0], q = rest[i][1], w = rest[i][2];
    const T det = Ud[p] * (kd[q] * bd[w] - kd[w] * bd[q]) - Ud[q] * (kd[p] * bd[w] - kd[w] * bd[p]) +
                  Ud[w] * (kd[p] * bd[q] - kd[q] * bd[p]);
    perp[i] = ((i & 1) ? -s : s) * det;
  }
}
// camera polarization basis F1 (tetrad.hpp:27-48, 111-133), in the scalar type
template <class T>
static void camera_F1(const Model<T>* m, const double X0[4], const double k0[4], double F1[4]) {
  T X[4], k[4], U[4], gu[4][4];
  for (int i = 0; i < 4; i++) {
    X[i] = T(X0[i]);
    k[i] = T(k0[i]);
  }
  gcon<T>(m->a, m->hslope, X, gu);
  const T Ud0 = T(-1);
  const T B = -gu[0][3] * Ud0 / gu[3][3];
  const T C = -(T(1) + gu[0][0] * Ud0 * Ud0) / gu[3][3];
  const T Ud[4] = {Ud0, T(0), T(0), B + sqrt(B * B + C)};
  for (int i = 0; i < 4; i++) {
    U[i] = T(0);
    for (int j = 0; j < 4; j++) U[i] += gu[i][j] * Ud[j];
  }
  const T up[4] = {T(0), T(0), T(0), T(-1)};
  T perp[4], para[4];
  observer_frame(m, X, k, U, up, perp, para);
  for (int i = 0; i < 4; i++) F1[i] = dbl(perp[i]);
}
template <class T>
static T pitch(const T g[4][4], const T k[4], const T b[4], const T U[4]) {  // tetrad.hpp:138-154
  const T bn = sqrt(fabs_(dot(g, b, b)));
  if (bn == T(0)) return T(M_PI / 2.);
  const T kk = fabs_(dot(g, k, U));
  T mu = dot(g, k, b) / (kk * bn);
  if (fabs_(mu) > T(1)) mu = mu / fabs_(mu);
  return acos(mu);
}

// ----------------------------------------------------------------------------
// geodesic (src/core/geometry/geodesic.hpp)
// ----------------------------------------------------------------------------
template <class T>
struct Ray {  // compensated (hi, lo) state of x^mu, k^mu, F1^mu (geodesic.hpp:11-14)
  T hi[12], lo[12];
};
template <class T>
struct Pt {
  int i;
  T x[4], k[4], f[4], dl;
};
template <class T>
struct Shape {  // per-point metric quantities shared by the RHS (geodesic.hpp:34-39, 81-109)
  T a, a2, r, r2, tp, tp2, tpp, s, c, s2, s2th, cot, Sig, Sb, Delta, A, B, C;
};
template <class T>
static inline Shape<T> shape_at(T a, T hs, T x1, T x2) {
  Shape<T> G;
  G.a = a;
  G.a2 = a * a;
  G.r = exp(x1);
  G.r2 = G.r * G.r;
  T s2x, c2x;
  sincos_(T(2) * x2, &s2x, &c2x);
  const T th = x2 + T(0.5) * hs * s2x;
  G.tp = T(1) + hs * c2x;
  G.tp2 = G.tp * G.tp;
  G.tpp = -T(2) * hs * s2x;
  sincos_(th, &G.s, &G.c);
  G.s2 = G.s * G.s;
  G.s2th = T(2) * G.s * G.c;
  G.cot = G.c / G.s;
  G.Sig = G.r2 + G.a2 * G.c * G.c;
  G.Sb = T(1) / G.Sig;
  G.Delta = G.r2 - T(2) * G.r + G.a2;
  G.A = G.Sig * G.Delta + T(2) * G.r * (G.r2 + G.a2);
  G.B = (T(2) * G.r2 - G.Sig) * G.Sb * G.Sb * G.Sb;
  G.C = G.r * G.Sb - G.a2 * G.B * G.s2;
  return G;
}
// -Gamma^i_jk U^j V^k for the MKS-BHAC Kerr-Schild metric (geodesic.hpp:114-169)
template <class T>
static void christoffel_uv(const Shape<T>& G, const T U[4], const T V[4], T out[4]) {
  const T h = T(0.5), two = T(2), four = T(4);
  const T P00 = U[0] * V[0];
  const T P01 = h * (U[0] * V[1] + U[1] * V[0]);
  const T P02 = h * (U[0] * V[2] + U[2] * V[0]);
  const T P03 = h * (U[0] * V[3] + U[3] * V[0]);
  const T P11 = U[1] * V[1];
  const T P12 = h * (U[1] * V[2] + U[2] * V[1]);
  const T P13 = h * (U[1] * V[3] + U[3] * V[1]);
  const T P22 = U[2] * V[2];
  const T P23 = h * (U[2] * V[3] + U[3] * V[2]);
  const T P33 = U[3] * V[3];
  const T a = G.a, a2 = G.a2, r = G.r, r2 = G.r2, rp = G.r, rp2 = G.r2;
  const T tp = G.tp, tp2 = G.tp2, s = G.s, c = G.c, s2 = G.s2, s2th = G.s2th;
  const T Sig = G.Sig, Sb = G.Sb, Sb2 = Sb * Sb, Sb3 = Sb2 * Sb;
  const T Del = G.Delta, A = G.A, B = G.B, C = G.C, cot = G.cot;
  const T K = a2 * r * s2th * Sb2 * tp;
  out[0] = -two * (-r2 * Sb * tp2 * P22 - r * C * s2 * P33 + K * (-P02 - rp * P12 + a * s2 * P23) +
                   B * ((Sig + two * r) * rp * P01 + r * P00 + (Sig + r) * rp2 * P11 - two * a * r * s2 * P03 -
                        a * s2 * (Sig + two * r) * rp * P13));
  out[1] = -((T(1) - (two * Sig - Del) * B * rp) * P11 +
             (Del / rp) * (B * P00 - two * a * B * s2 * P03 - r * Sb * tp2 * P22 - C * s2 * P33) +
             two * ((Del - Sig) * B * P01 - a2 * s * c * Sb * tp * P12 + a * (r * Sb + (Sig - Del) * B) * s2 * P13));
  out[2] = -((G.tpp / tp - a2 * s * c * Sb * tp) * P22 +
             Sb3 *
                 ((-a2 * r * s2th) * (P00 + rp2 * P11) -
                  s * c * (Del * Sig * Sig + two * r * (r2 + a2) * (r2 + a2)) * P33 - two * a2 * r * s2th * rp * P01 +
                  two * a * s * c * (A + Sig * (Sig - Del)) * rp * P13 + two * a * r * (r2 + a2) * s2th * P03) /
                 tp +
             two * r * Sb * rp * P12);
  out[3] = -(a * (B * P00 + B * rp2 * P11 - r * Sb * tp2 * P22 - C * s2 * P33) + two * a * B * rp * P01 -
             four * a * r * cot * Sb2 * tp * P02 - two * a * (Sig + two * r) * cot * Sb2 * rp * tp * P12 -
             two * a2 * B * s2 * P03 + two * (cot + a2 * r * s2th * Sb2) * tp * P23 + two * C * rp * P13);
}
template <class T>
static inline void transport_rhs(T a, T hs, const T y[8], const T f[4], T dy[8], T df[4]) {  // geodesic.hpp:180-195
  const Shape<T> G = shape_at(a, hs, y[1], y[2]);
  const T* U = y + 4;
  T ag[4];
  for (int i = 0; i < 4; i++) dy[i] = U[i];
  christoffel_uv(G, U, U, ag);
  christoffel_uv(G, U, f, df);
  for (int i = 0; i < 4; i++) dy[4 + i] = ag[i];
}
template <class T>
static inline void two_sum_acc(T* hi, T* lo, T d) {  // TwoSum + Fast2Sum (precision.hpp:65-73)
  const T s = *hi + d;
  const T bb = s - *hi;
  const T err = (*hi - (s - bb)) + (d - bb);
  *lo += err;
  const T s2 = s + *lo;
  *lo = *lo - (s2 - s);
  *hi = s2;
}
template <class T>
static inline void bump(Ray<T>* g, int i, T inc) {  // geodesic.hpp:211-217
  if (is_fp32<T>())
    two_sum_acc(&g->hi[i], &g->lo[i], inc);
  else
    g->hi[i] += inc;
}
template <class T>
static inline T pole_dist(const Ray<T>* g) {  // geodesic.hpp:200-205
  const bool south = is_fp32<T>() && g->hi[2] > T(0.5 * M_PI);
  const T pi_hi = T(M_PI), pi_lo = T(M_PI - (double)(typename base_of<T>::type)(M_PI));
  if (south) return (pi_hi - g->hi[2]) + (pi_lo - g->lo[2]);
  return g->hi[2] + g->lo[2];
}
// classical RK4 with the FP32 south-hemisphere fold (geodesic.hpp:222-256)
template <class T>
static void rk4(T a, T hs, Ray<T>* g, T dt) {
  const T pi_hi = T(M_PI), pi_lo = T(M_PI - (double)(typename base_of<T>::type)(M_PI));
  const bool south = is_fp32<T>() && g->hi[2] > T(0.5 * M_PI);
  T dy[4][8], dfr[4][4];
  T iy[8], iff[4];
  for (int i = 0; i < 8; i++) iy[i] = T(0);
  for (int i = 0; i < 4; i++) iff[i] = T(0);
  const T w[3] = {T(0.5), T(0.5), T(1.0)};
  for (int q = 0; q < 4; q++) {
    T ty[8], tf[4];
    for (int i = 0; i < 8; i++) ty[i] = g->hi[i] + (g->lo[i] + iy[i]);
    for (int i = 0; i < 4; i++) tf[i] = g->hi[8 + i] + (g->lo[8 + i] + iff[i]);
    if (south) {
      ty[2] = (pi_hi - g->hi[2]) + ((pi_lo - g->lo[2]) - iy[2]);
      ty[6] = -ty[6];
      tf[2] = -tf[2];
    }
    transport_rhs(a, hs, ty, tf, dy[q], dfr[q]);
    if (south) {
      dy[q][2] = -dy[q][2];
      dy[q][6] = -dy[q][6];
      dfr[q][2] = -dfr[q][2];
    }
    if (q == 3) break;
    const T wdt = w[q] * dt;
    for (int i = 0; i < 8; i++) iy[i] = wdt * dy[q][i];
    for (int i = 0; i < 4; i++) iff[i] = wdt * dfr[q][i];
  }
  const T two = T(2), dt6 = dt / T(6);
  for (int i = 0; i < 8; i++) bump(g, i, ((dy[0][i] + dy[3][i]) + two * (dy[1][i] + dy[2][i])) * dt6);
  for (int i = 0; i < 4; i++) bump(g, 8 + i, ((dfr[0][i] + dfr[3][i]) + two * (dfr[1][i] + dfr[2][i])) * dt6);
}
// RAPTOR step rule (geodesic.hpp:267-292); returns a negative step
template <class T>
static T step_len(T stepsize, const T X[4], const T k[4], const T* dxc) {
  const T S2 = lim<T>::step_tiny();
  const T d1 = stepsize / (fabs_(k[1]) + S2);
  const T d2 = stepsize * fmin_(X[2], T(M_PI) - X[2]) / (fabs_(k[2]) + S2);
  const T d3 = stepsize / (fabs_(k[3]) + S2);
  const T i1 = T(1) / (fabs_(d1) + S2), i2 = T(1) / (fabs_(d2) + S2), i3 = T(1) / (fabs_(d3) + S2);
  T grid = lim<T>::step_big();
  if (dxc) {
    const T s1 = dxc[0] / (fabs_(k[1]) + S2), s2 = dxc[1] / (fabs_(k[2]) + S2), s3 = dxc[2] / (fabs_(k[3]) + S2);
    const T j1 = T(1) / (fabs_(s1) + S2), j2 = T(1) / (fabs_(s2) + S2), j3 = T(1) / (fabs_(s3) + S2);
    grid = T(1) / (j1 + j2 + j3) / T(4);
  }
  return -fmin_(T(1) / (i1 + i2 + i3), grid);
}
template <class T>
static void null_fix(const T g[4][4], T k[4]) {  // re-solve k^t from k.k = 0 (geodesic.hpp:297-306)
  const T aa = g[0][0];
  const T k1 = k[1], k2 = k[2], k3 = k[3];
  const T bc = -(g[0][1] * k1) / (aa) - (g[0][2] * k2) / (aa) - (g[0][3] * k3) / (aa);
  const T gz = -(g[1][1] * k1 * k1) / (aa) - (g[1][2] * k1 * k2) / (aa) - (g[1][3] * k1 * k3) / (aa) -
               (g[2][1] * k2 * k1) / (aa) - (g[2][2] * k2 * k2) / (aa) - (g[2][3] * k2 * k3) / (aa) -
               (g[3][1] * k3 * k1) / (aa) - (g[3][2] * k3 * k2) / (aa) - (g[3][3] * k3 * k3) / (aa);
  k[0] = bc + sqrt(bc * bc + gz);
}

// ----------------------------------------------------------------------------
// transfer coefficients (src/core/transfer/coeffs.hpp, src/core/numerics/bessel.hpp)
// ----------------------------------------------------------------------------
template <class T>
struct Coef {
  T jI, jQ, jU, jV, aI, aQ, aU, aV, rQ, rU, rV;
};
template <class T>
static inline Coef<T> coef_zero() {
  Coef<T> c;
  c.jI = c.jQ = c.jU = c.jV = c.aI = c.aQ = c.aU = c.aV = c.rQ = c.rU = c.rV = T(0);
  return c;
}
// Dexter (2016) thermal synchrotron fitting functions (coeffs.hpp:44-70)
template <class T>
static T fit_I(T x) {
  const T cx = cbrt(x), ic = T(1) / cx;
  return T(2.5651) * (T(1) + T(1.92) * ic + T(0.9977) * ic * ic) * exp(T(-1.8899) * cx);
}
template <class T>
static T fit_Q(T x) {
  const T cx = cbrt(x), ic = T(1) / cx;
  return T(2.5651) * (T(1) + T(0.93193) * ic + T(0.499873) * ic * ic) * exp(T(-1.8899) * cx);
}
template <class T>
static T fit_V(T x) {
  const T cx = cbrt(x), ic = T(1) / cx;
  return (T(1.81348) * ic * ic * ic + T(3.42319) * ic * ic + T(0.0292545) / sqrt(x) + T(2.03773) * ic) *
         exp(T(-1.8899) * cx);
}
template <class T>
static T faraday_fm(T X) {  // Shcherbakov-type conversion factor f_m(X) (coeffs.hpp:63-70)
  const double C0 = 369.69342869090815, L120 = 4.787491742782046;
  const T lx = log(X);
  const T ex = exp(-X / T(47.2));
  return T(2.011) * exp(-exp(T(1.035) * lx) / T(4.7)) - cos(X * T(0.5)) * exp(-exp(T(1.2) * lx) / T(2.73)) -
         T(0.011) * ex + (T(0.011) * ex - T(C0) * exp(T(-8. / 3.) * lx)) * T(0.5) * (T(1) + tanh(T(10) * (lx - T(L120))));
}
template <class T>
static T planck(T nu, T te) {  // coeffs.hpp:75-80
  const T pre = T(2. * cgs::H) * nu * nu * nu / T(cgs::C * cgs::C);
  const T xx = T(cgs::H_OVER_MEC2) * nu / te;
  return is_fp32<T>() ? pre / expm1(xx) : pre / (exp(xx) - T(1));
}
template <class T>
static T hankel(int nu, T x) {  // Hankel asymptotic series without its prefactor (bessel.hpp:99-108)
  const T mu = T(4 * nu * nu);
  T term = T(1), sum = T(1);
  for (int k = 1; k <= 30; k++) {
    term = term * ((mu - T(2 * k - 1) * T(2 * k - 1)) / (T(k) * T(8) * x));
    sum += term;
    if (fabs_(term) < lim<T>::tol_asym()) break;
  }
  return sum;
}
template <class T>
static void bessel_K_scaled(T x, T* K0, T* K1, T* K2) {  // trapezoid in t of e^x K_n(x) (bessel.hpp:65-88)
  const T dt = fmin_(T(0.2), T(0.7) / sqrt(x));
  const T tmax = acosh(T(1) + T(42) / x) + T(2);
  const int nmax = (int)(tmax / dt) + 1;
  T s0 = T(0.5), s1 = T(0.5), s2 = T(0.5);
  for (int k = 1; k <= nmax; k++) {
    const T t = T(k) * dt;
    const T sh = sinh(T(0.5) * t);
    const T ch = T(1) + T(2) * sh * sh;
    const T e = exp(-T(2) * x * sh * sh);
    s0 += e;
    s1 += e * ch;
    s2 += e * (T(2) * ch * ch - T(1));
    if (e * (T(2) * ch * ch) < lim<T>::tol_series() * T(0.1) * s2) break;
  }
  *K0 = dt * s0;
  *K1 = dt * s1;
  *K2 = dt * s2;
}
template <class T>
static void bessel_ratio(T x, T* r10, T* r02) {  // K1/K2, K0/K2 (bessel.hpp:118-148)
  if (x >= T(30)) {
    const T s2 = hankel(2, x);
    *r10 = hankel(1, x) / s2;
    *r02 = hankel(0, x) / s2;
    return;
  }
  if (x < T(0.01)) {
    T k0, k1, k2;
    bessel_K_scaled(x, &k0, &k1, &k2);
    *r10 = k1 / k2;
    *r02 = k0 / k2;
    return;
  }
  const int p = (x < T(ORC_BES_EDGE[0][1])) ? 0 : (x < T(ORC_BES_EDGE[1][1])) ? 1 : 2;
  const T la = T(ORC_BES_LNEDGE[p][0]), lb = T(ORC_BES_LNEDGE[p][1]);
  const T z = (T(2) * log(x) - (lb + la)) / (lb - la);
  T b1a = T(0), b2a = T(0), b1b = T(0), b2b = T(0);
  for (int j = ORC_BES_NTERM[p] - 1; j >= 1; j--) {  // Clenshaw, both series at once
    const T ta = T(2) * z * b1a - b2a + T(ORC_BES_K1K2[p][j]);
    b2a = b1a;
    b1a = ta;
    const T tb = T(2) * z * b1b - b2b + T(ORC_BES_K0K2[p][j]);
    b2b = b1b;
    b1b = tb;
  }
  *r10 = T(ORC_BES_K1K2[p][0]) + z * b1a - b2a;
  *r02 = T(ORC_BES_K0K2[p][0]) + z * b1b - b2b;
}
// plasma-frame coefficients (coeffs.hpp:95-138)
template <class T>
static Coef<T> thermal(T te, T ne, T nu, T B, T thB) {
  Coef<T> r = coef_zero<T>();
  T sth, cth;
  sincos_(thB, &sth, &cth);
  const T ite = T(1) / te;
  const T nuc = T(cgs::NU_CRIT) * B * sth * te * te;
  const T x = nu / nuc;
  const T pref = T(cgs::JNU) * ne * nu * ite * ite;
  r.jI = pref * fit_I(x);
  r.jQ = pref * fit_Q(x);
  r.jV = T(4. / 3.) * pref * ite * (cth / sth) * fit_V(x);
  const T iB = T(1) / planck(nu, te);
  r.aI = r.jI * iB;
  r.aQ = r.jQ * iB;
  r.aV = r.jV * iB;
  const T iw = T(1) / (T(2 * M_PI) * nu);
  const T wp2 = T(cgs::WP2) * ne;
  const T om0 = T(cgs::GYRO) * B;
  const T Xe = te * sqrt(T(1.4142135623730951) * sth * (T(1.e3 / (2. * M_PI)) * om0 / nu));
  T kr1, kr0;
  if (ite < T(1. / 5.)) {  // Theta_e > 5: asymptotic ratios
    const T k0 = -log(ite / T(2)) - T(0.5772), k1 = T(1) / ite, k2 = T(2) / ite / ite;
    kr1 = k1 / k2;
    kr0 = (k2 > T(0)) ? k0 / k2 : T(1);
  } else {
    bessel_ratio(ite, &kr1, &kr0);
  }
  const T q = om0 * iw;
  const T wpw = wp2 * iw;
  r.rQ = (T(0.5) / T(cgs::C)) * wpw * q * q * faraday_fm(Xe) * (kr1 + T(6) * te) * sth * sth;
  const T shgm = T(1) - T(0.11) * log(T(1) + T(0.035) * Xe);
  r.rV = (T(1) / T(cgs::C)) * wpw * q * (kr0 * shgm) * cth;
  return r;
}
template <class T>
static Coef<T> to_invariant(Coef<T> c, T nup, T jscale) {  // coeffs.hpp:147-164
  const T js = jscale / nup / nup;
  c.jI = c.jI * js;
  c.jQ = c.jQ * js;
  c.jV = c.jV * js;
  c.aI = c.aI * nup;
  c.aQ = c.aQ * nup;
  c.aV = c.aV * nup;
  c.rQ = c.rQ * nup;
  c.rV = c.rV * nup;
  const T pf = sqrt(c.jQ * c.jQ + c.jV * c.jV) / c.jI;
  if (pf > T(1)) {
    c.jQ = c.jQ / (pf + T(0.005));
    c.jU = c.jU / (pf + T(0.005));
    c.jV = c.jV / (pf + T(0.005));
  }
  return c;
}
template <class T>
static Coef<T> rotate(const Coef<T>& c, T c2, T s2) {  // coeffs.hpp:173-182
  Coef<T> r = c;
  r.jQ = c2 * c.jQ + s2 * c.jU;
  r.jU = -s2 * c.jQ + c2 * c.jU;
  r.aQ = c2 * c.aQ + s2 * c.aU;
  r.aU = -s2 * c.aQ + c2 * c.aU;
  r.rQ = c2 * c.rQ + s2 * c.rU;
  r.rU = -s2 * c.rQ + c2 * c.rU;
  return r;
}

// ----------------------------------------------------------------------------
// single-pass segment operator (src/core/transfer/operators.hpp)
// ----------------------------------------------------------------------------
template <class T>
static void invert4(const T m[4][4], T out[4][4]) {  // Gauss-Jordan, successive-swap partial pivoting (:91-136)
  T A[4][8];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      A[i][j] = m[i][j];
      A[i][4 + j] = (i == j) ? T(1) : T(0);
    }
  for (int c = 0; c < 4; c++) {
    for (int r = c + 1; r < 4; r++)
      if (fabs_(A[r][c]) > fabs_(A[c][c]))
        for (int j = 0; j < 8; j++) std::swap(A[c][j], A[r][j]);
    const T d = A[c][c];
    for (int j = 0; j < 8; j++) A[c][j] = A[c][j] / d;
    for (int r = 0; r < 4; r++) {
      if (r == c) continue;
      const T f = A[r][c];
      for (int j = 0; j < 8; j++) A[r][j] = A[r][j] - f * A[c][j];
    }
  }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) out[i][j] = A[i][4 + j];
}
template <class T>
static inline void matmul4(const T A[4][4], const T B[4][4], T C[4][4]) {
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      T s = T(0);
      for (int k = 0; k < 4; k++) s += A[i][k] * B[k][j];
      C[i][j] = s;
    }
}
template <class T>
static inline void matvec4(const T A[4][4], const T v[4], T out[4]) {
  for (int i = 0; i < 4; i++) {
    out[i] = T(0);
    for (int j = 0; j < 4; j++) out[i] += A[i][j] * v[j];
  }
}
template <class T>
struct Op {
  T A[4][4], S[4];
};
// implicit trapezoid: O_A = (I + h/2 K)^-1 (I - h/2 K), O_S = h (I + h/2 K)^-1 j (operators.hpp:158-176)
template <class T>
static Op<T> segment_op(const Coef<T>& c, T h) {
  const T K[4][4] = {{c.aI, c.aQ, c.aU, c.aV}, {c.aQ, c.aI, c.rV, -c.rU}, {c.aU, -c.rV, c.aI, c.rQ},
                     {c.aV, c.rU, -c.rQ, c.aI}};
  const T j[4] = {c.jI, c.jQ, c.jU, c.jV};
  T P[4][4], N[4][4], Pi[4][4], tmp[4];
  for (int i = 0; i < 4; i++)
    for (int k = 0; k < 4; k++) {
      const T id = (i == k) ? T(1) : T(0);
      P[i][k] = id + T(0.5) * h * K[i][k];
      N[i][k] = id - T(0.5) * h * K[i][k];
    }
  invert4(P, Pi);
  Op<T> r;
  matmul4(Pi, N, r.A);
  matvec4(Pi, j, tmp);
  for (int i = 0; i < 4; i++) r.S[i] = h * tmp[i];
  return r;
}

// ----------------------------------------------------------------------------
// the ray (src/core/render/pixel.hpp)
// ----------------------------------------------------------------------------
template <class T>
struct Acc {
  T A[4][4], S[4], tau, tauF;
};
template <class T>
struct RayState {
  Acc<T> acc[MAXF];
  int nplasma, bad;
};
// one plasma sample (pixel.hpp:52-123)
template <class T>
static void sample(const Model<T>* m, RayState<T>* ps, const Pt<T>* sm, int next_leaf) {
  const int prev = tl_phase;
  tl_phase = PH_SAMPLE;
  T X[4], k[4], F1[4];
  for (int q = 0; q < 4; q++) {
    X[q] = sm->x[q];
    k[q] = sm->k[q];
    F1[q] = sm->f[q];
  }
  const T r = exp(X[1]);
  T g[4][4];
  gcov<T>(m->a, m->hslope, X, g);
  if (fabs_(dot(g, k, k)) > T(1e-6) && r > T(2)) null_fix(g, k);
  Fluid<T> fl;
  int cache = next_leaf;
  const int in_vol = sample_fluid(m, X, &fl, &cache);
  int emit = 0;
  T ratio = T(0), c2 = T(0), s2 = T(0), pa = T(0);
  if (in_vol && r < m->rt_outer) {
    T gs[4][4], perp[4], para[4];
    gcov<T>(m->a, m->hslope, fl.X, gs);
    pa = pitch(gs, k, fl.b, fl.U);
    if (!(pa == T(0) || pa == T(M_PI))) {
      observer_frame(m, fl.X, k, fl.U, fl.b, perp, para);
      const T Cp = dot(gs, perp, F1), Sp = dot(gs, para, F1);  // chi_from_F1 (tetrad.hpp:160-182)
      const T n = Cp * Cp + Sp * Sp;
      c2 = (Cp * Cp - Sp * Sp) / n;
      s2 = T(2) * Cp * Sp / n;
      ratio = -dot(gs, k, fl.U);
      emit = 1;
    }
  }
  if (!emit) {
    tl_phase = prev;
    return;
  }
  for (int f = 0; f < m->up.nfreqs; f++) {
    tl_phase = PH_FREQ;
    const T nup = m->freq[f] * ratio;
    const Coef<T> seg = rotate(to_invariant(thermal(fl.te, fl.ne, nup, fl.B, pa), nup, m->jscale[f]), c2, s2);
    const T h = sm->dl * m->hconst[f];
    const Op<T> op = segment_op(seg, h);
    bool fin = true;
    for (int a = 0; a < 4; a++) {
      fin = fin && finite_(op.S[a]);
      for (int b = 0; b < 4; b++) fin = fin && finite_(op.A[a][b]);
    }
    if (!fin) {
      ps->bad++;
      continue;
    }
    Acc<T>& acc = ps->acc[f];
    T Tc[4], Tn[4][4];
    matvec4(acc.A, op.S, Tc);
    for (int i = 0; i < 4; i++) acc.S[i] += Tc[i];
    matmul4(acc.A, op.A, Tn);
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) acc.A[i][j] = Tn[i][j];
    acc.tau += seg.aI * h;
    acc.tauF += fabs_(seg.rV * h);
  }
  ps->nplasma++;
  tl_phase = prev;
}

struct Stats {
  long long steps, nplasma, bad;
};
// one ray, camera to termination (pixel.hpp:134-224); out: nf x {I,Q,U,V,tau,tauF}
template <class T>
static Stats trace(const Model<T>* m, double alpha, double beta, double* out) {
  const int nf = m->up.nfreqs;
  Stats st = {0, 0, 0};
  for (int q = 0; q < nf * 6; q++) out[q] = 0.;
  tl_phase = PH_SETUP;
  double ph[8], F1[4];
  camera_photon(m, alpha, beta, ph);
  camera_F1(m, ph, ph + 4, F1);
  RayState<T> ps;
  for (int f = 0; f < nf; f++) {
    for (int i = 0; i < 4; i++) {
      for (int j = 0; j < 4; j++) ps.acc[f].A[i][j] = (i == j) ? T(1) : T(0);
      ps.acc[f].S[i] = T(0);
    }
    ps.acc[f].tau = ps.acc[f].tauF = T(0);
  }
  Ray<T> g;
  for (int i = 0; i < 8; i++) g.hi[i] = T(ph[i]);
  for (int i = 0; i < 4; i++) g.hi[8 + i] = T(F1[i]);
  for (int i = 0; i < 12; i++) g.lo[i] = T(0);
  if (is_fp32<T>()) {
    for (int i = 0; i < 8; i++) g.lo[i] = T(ph[i] - dbl(g.hi[i]));
    for (int i = 0; i < 4; i++) g.lo[8 + i] = T(F1[i] - dbl(g.hi[8 + i]));
  }
  ps.nplasma = ps.bad = 0;
  int steps = 0, turns = 0, stop = 0;
  T thdot_prev = T(0), dl_prev = T(0);
  const bool bneg = beta < 0., bpos = beta > 0.;
  Pt<T> pend;
  pend.i = -1;
  for (;;) {
    tl_phase = PH_STEP;
    Pt<T> cur;
    for (int i = 0; i < 4; i++) {
      cur.x[i] = g.hi[i];
      cur.k[i] = g.hi[4 + i];
      cur.f[i] = g.hi[8 + i];
      if (is_fp32<T>()) {
        cur.x[i] = cur.x[i] + g.lo[i];
        cur.k[i] = cur.k[i] + g.lo[4 + i];
        cur.f[i] = cur.f[i] + g.lo[8 + i];
      }
    }
    cur.i = steps;
    cur.dl = dl_prev;
    const T r = exp(cur.x[1]);
    const bool alive = r < m->r_out && r > m->cut_in && steps < m->up.max_steps && !stop;
    if (pend.i >= 1) {
      const int nl = alive ? leaf_at(m, cur.x) : -1;
      sample(m, &ps, &pend, nl);
    }
    pend = cur;
    if (!alive) break;
    if (thdot_prev * cur.k[2] < T(0) && steps > 2) turns++;
    thdot_prev = cur.k[2];
    if ((bneg && turns > m->up.max_order) || (bpos && turns > (m->up.max_order + 1))) stop = 1;
    const T* dxc = cell_size_at(m, cur.x);
    T X[4], k[4];
    for (int i = 0; i < 4; i++) {
      X[i] = cur.x[i];
      k[i] = cur.k[i];
    }
    X[2] = pole_dist(&g);
    const T dl = step_len(m->stepsize, X, k, dxc);
    rk4(m->a, m->hslope, &g, dl);
    dl_prev = fabs_(dl);
    steps++;
  }
  st.steps = steps;
  st.nplasma = ps.nplasma;
  st.bad = ps.bad;
  if (ps.nplasma > 0)
    for (int f = 0; f < nf; f++) {
      for (int q = 0; q < 4; q++) out[6 * f + q] = dbl(ps.acc[f].S[q]);
      out[6 * f + 4] = dbl(ps.acc[f].tau);
      out[6 * f + 5] = dbl(ps.acc[f].tauF);
    }
  return st;
}

// pixel (i, j) -> rotated impact parameters (camera.hpp:11-23, jobs.hpp:64-76)
template <class T>
static void impact(const Model<T>* m, int i, int j, double* alpha, double* beta) {
  const Params& u = m->up;
  const double a0 = -0.5 * u.fov_x + (i + 0.5) * u.fov_x / u.img_width;
  const double b0 = -0.5 * u.fov_y + (j + 0.5) * u.fov_y / u.img_height;
  *alpha = a0 * m->cos_pa + b0 * m->sin_pa;
  *beta = -a0 * m->sin_pa + b0 * m->cos_pa;
}

}  // namespace orc

// ============================================================================
// C interface for the tests (ctypes)
// ============================================================================
using namespace orc;

namespace {
struct Handle {
  int precision;  // 32, 64, or 0 = counted double, 1 = counted float
  void* model;
};
thread_local std::string g_err;
template <class T>
Model<T>* M(Handle* h) {
  return (Model<T>*)h->model;
}
template <class T>
int render_any(Handle* h, const long* pix, long npix, double* res, long long* stats, int nthreads) {
  Model<T>* m = M<T>(h);
  const int nf = m->up.nfreqs;
  long long s0 = 0, s1 = 0, s2 = 0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : s0, s1, s2)
  for (long n = 0; n < npix; n++) {
    const long idx = pix[n];
    const int i = (int)(idx / m->up.img_height), j = (int)(idx % m->up.img_height);
    double al, be;
    impact(m, i, j, &al, &be);
    const Stats st = trace(m, al, be, res + (size_t)n * nf * 6);
    s0 += st.steps;
    s1 += st.nplasma;
    s2 += st.bad;
  }
  stats[0] = s0;
  stats[1] = s1;
  stats[2] = s2;
  return 0;
}
}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

// precision: 32 (float), 64 (double), 164 (counted double), 132 (counted float)
void* orc_load(const char* dump, const char* grid, const Params* p, int precision) {
  Handle* h = new Handle{precision, nullptr};
  try {
    switch (precision) {
      case 32: { auto* m = new Model<float>(); m->up = *p; h->model = m; load_model(m, dump, grid); break; }
      case 64: { auto* m = new Model<double>(); m->up = *p; h->model = m; load_model(m, dump, grid); break; }
      case 164: { auto* m = new Model<Counted<double>>(); m->up = *p; h->model = m; load_model(m, dump, grid); break; }
      case 132: { auto* m = new Model<Counted<float>>(); m->up = *p; h->model = m; load_model(m, dump, grid); break; }
      default: throw std::runtime_error("precision must be 32, 64, 132 or 164");
    }
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
  return h;
}

void orc_free(void* hv) {
  Handle* h = (Handle*)hv;
  if (!h) return;
  switch (h->precision) {
    case 32: delete M<float>(h); break;
    case 64: delete M<double>(h); break;
    case 164: delete M<Counted<double>>(h); break;
    case 132: delete M<Counted<float>>(h); break;
  }
  delete h;
}

int orc_render(void* hv, const long* pix, long npix, double* res, long long* stats, int nthreads) {
  Handle* h = (Handle*)hv;
  switch (h->precision) {
    case 32: return render_any<float>(h, pix, npix, res, stats, nthreads);
    case 64: return render_any<double>(h, pix, npix, res, stats, nthreads);
    case 164: return render_any<Counted<double>>(h, pix, npix, res, stats, 1);
    case 132: return render_any<Counted<float>>(h, pix, npix, res, stats, 1);
  }
  return -1;
}

// op counts of the calling thread since the last reset: [phase][op]
void orc_counts(long long* out) { std::memcpy(out, tl_ops, sizeof tl_ops); }
void orc_counts_reset(void) { std::memset(tl_ops, 0, sizeof tl_ops); }

// model facts: nleafs, cells, nlevels, nx[3], ng[3]
void orc_model_info(void* hv, int* out, double* dout) {
  Handle* h = (Handle*)hv;
  auto fill = [&](auto* m) {
    out[0] = m->nleafs; out[1] = m->cells; out[2] = m->nlevels;
    for (int d = 0; d < 3; d++) { out[3 + d] = m->nx[d]; out[6 + d] = m->ng[d]; }
    dout[0] = dbl(m->a); dout[1] = dbl(m->hslope); dout[2] = m->L; dout[3] = dbl(m->cut_in); dout[4] = dbl(m->r_out);
    dout[5] = m->B_unit_d; dout[6] = m->Ne_unit_d;
  };
  if (h->precision == 32) fill(M<float>(h));
  else if (h->precision == 64) fill(M<double>(h));
}

// decoded primitives [nleafs][cells][8] and block table lb/dxc [nleafs][3] (as double)
int orc_tables(void* hv, double* prim, double* lb, double* dxc, int* level, int* ind) {
  Handle* h = (Handle*)hv;
  auto fill = [&](auto* m) {
    if (prim) for (size_t i = 0; i < m->prim.size(); i++) prim[i] = dbl(m->prim[i]);
    for (size_t i = 0; i < m->lb.size(); i++) { if (lb) lb[i] = dbl(m->lb[i]); if (dxc) dxc[i] = dbl(m->dxc[i]); }
    if (level) for (int q = 0; q < m->nleafs; q++) level[q] = m->leaf_level[q];
    if (ind) for (size_t q = 0; q < m->leaf_ind.size(); q++) ind[q] = m->leaf_ind[q];
  };
  if (h->precision == 32) fill(M<float>(h));
  else if (h->precision == 64) fill(M<double>(h));
  else return -1;
  return 0;
}

// AMR lookup of already-wrapped points: x [n][4] -> leaf (find_leaf) and the map candidate
void orc_find_leaf(void* hv, const double* x, long n, int* leaf, int* cand) {
  Handle* h = (Handle*)hv;
  auto run = [&](auto* m) {
    using T = typename std::remove_reference<decltype(m->a)>::type;
    for (long q = 0; q < n; q++) {
      T p[4] = {T(x[4 * q]), T(x[4 * q + 1]), T(x[4 * q + 2]), T(x[4 * q + 3])};
      leaf[q] = find_leaf(m, p);
      if (cand) cand[q] = map_lookup(m, p);
    }
  };
  if (h->precision == 32) run(M<float>(h));
  else if (h->precision == 64) run(M<double>(h));
}

// fluid at points X [n][4] (unwrapped): out [n][14] = emit, ne, B, te, U[4], b[4], leaf, pad
void orc_fluid(void* hv, const double* X, long n, double* out) {
  Handle* h = (Handle*)hv;
  auto run = [&](auto* m) {
    using T = typename std::remove_reference<decltype(m->a)>::type;
    for (long q = 0; q < n; q++) {
      T p[4] = {T(X[4 * q]), T(X[4 * q + 1]), T(X[4 * q + 2]), T(X[4 * q + 3])};
      Fluid<T> fl = {};
      int cache = -1;
      const int e = sample_fluid(m, p, &fl, &cache);
      double* o = out + 14 * q;
      o[0] = e; o[1] = dbl(fl.ne); o[2] = dbl(fl.B); o[3] = dbl(fl.te);
      for (int i = 0; i < 4; i++) { o[4 + i] = dbl(fl.U[i]); o[8 + i] = dbl(fl.b[i]); }
      o[12] = cache; o[13] = 0;
    }
  };
  if (h->precision == 32) run(M<float>(h));
  else if (h->precision == 64) run(M<double>(h));
}

// --- pure functions (double) for the unit pins
void orc_metric(double a, double hs, const double* X, double* gd, double* gu) {
  double A[4][4], B[4][4];
  gboth<double>(a, hs, X, A, B);
  std::memcpy(gd, A, sizeof A);
  std::memcpy(gu, B, sizeof B);
}
void orc_metric_split(double a, double hs, const double* X, double* gd, double* gu) {
  double A[4][4], B[4][4];
  gcov<double>(a, hs, X, A);
  gcon<double>(a, hs, X, B);
  std::memcpy(gd, A, sizeof A);
  std::memcpy(gu, B, sizeof B);
}
void orc_christoffel(double a, double hs, const double* X, const double* U, const double* V, double* out) {
  const Shape<double> G = shape_at(a, hs, X[1], X[2]);
  christoffel_uv(G, U, V, out);
}
void orc_bessel_ratios(const double* x, long n, double* r10, double* r02, int fp32) {
  for (long q = 0; q < n; q++) {
    if (fp32) {
      float a, b;
      bessel_ratio<float>((float)x[q], &a, &b);
      r10[q] = a;
      r02[q] = b;
    } else
      bessel_ratio<double>(x[q], &r10[q], &r02[q]);
  }
}
void orc_bessel_scaled(double x, double* k) { bessel_K_scaled<double>(x, &k[0], &k[1], &k[2]); }
// in: te, ne, nu, B, thB -> 11 plasma-frame coefficients
void orc_thermal(const double* in, double* out) {
  const Coef<double> c = thermal<double>(in[0], in[1], in[2], in[3], in[4]);
  const double v[11] = {c.jI, c.jQ, c.jU, c.jV, c.aI, c.aQ, c.aU, c.aV, c.rQ, c.rU, c.rV};
  std::memcpy(out, v, sizeof v);
}
void orc_invariant(const double* in, double nup, double jscale, double* out) {
  Coef<double> c = {in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8], in[9], in[10]};
  c = to_invariant(c, nup, jscale);
  const double v[11] = {c.jI, c.jQ, c.jU, c.jV, c.aI, c.aQ, c.aU, c.aV, c.rQ, c.rU, c.rV};
  std::memcpy(out, v, sizeof v);
}
void orc_segment_op(const double* coef11, double h, double* A16, double* S4) {
  const double* in = coef11;
  const Coef<double> c = {in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8], in[9], in[10]};
  const Op<double> op = segment_op(c, h);
  std::memcpy(A16, op.A, sizeof op.A);
  std::memcpy(S4, op.S, sizeof op.S);
}
void orc_invert4(const double* m16, double* out16) {
  double A[4][4], B[4][4];
  std::memcpy(A, m16, sizeof A);
  invert4(A, B);
  std::memcpy(out16, B, sizeof B);
}
// camera photon and F1 for a bare Kerr spacetime of spin a, stretching hs (FP64)
void orc_camera(double a, double hs, const Params* p, double alpha, double beta, double* ph8, double* F1) {
  Model<double> m;
  m.up = *p;
  m.a = a;
  m.hslope = hs;
  camera_photon(&m, alpha, beta, ph8);
  camera_F1(&m, ph8, ph8 + 4, F1);
}
// vacuum geodesic from a state: n RK4 steps of size dl (negative = backwards); states [n+1][12]
void orc_geodesic(double a, double hs, const double* y12, double dl, int n, double* states, int fp32) {
  auto run = [&](auto zero) {
    using T = decltype(zero);
    Ray<T> g;
    for (int i = 0; i < 12; i++) {
      g.hi[i] = T(y12[i]);
      g.lo[i] = is_fp32<T>() ? T(y12[i] - dbl(g.hi[i])) : T(0);
    }
    for (int s = 0; s <= n; s++) {
      for (int i = 0; i < 12; i++) states[12 * s + i] = dbl(g.hi[i]) + (is_fp32<T>() ? dbl(g.lo[i]) : 0.);
      if (s < n) rk4<T>(T(a), T(hs), &g, T(dl));
    }
  };
  if (fp32) run(0.f); else run(0.0);
}
double orc_step_len(double stepsize, const double* X, const double* k, const double* dxc) {
  return step_len<double>(stepsize, X, k, dxc);
}
}  // extern "C"
