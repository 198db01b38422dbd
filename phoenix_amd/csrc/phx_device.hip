// Device runtime of libphx: resident snapshots (device decode, block table,
// flattened AMR map), the persistent lane-refill march kernel and the C ABI of
// the device path (include/phx.h).  gfx950 only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <unordered_map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "phx_internal.h"
#include "phx_march.h"

using namespace phxd;

namespace phx {

#define HIPCHECK(call)                                                                                 \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess)                                                                              \
      throw Error(PHX_E_DEVICE, std::string(hipGetErrorString(e_)) + " (" #call ", " __FILE__ ":" +   \
                                    std::to_string(__LINE__) + ")");                                   \
  } while (0)

constexpr int BLOCK = 256;  // 4 wave64 per workgroup

// ============================================================================ kernels
// Precision of a march (the reference's PHX_ACCURACY, CMakeLists.txt:16-22, as a template):
//   M32  float cells and arithmetic, compensated (hi, lo) FP32 ray state (the FP32 build);
//   M64  double throughout (the FP64 build);
//   MIX  float cells and bulk arithmetic, FP64 ray state, FP64 RK4 steps near the horizon and
//        the poles, FP64 location of a sample's cell, Kahan-compensated Stokes sums.
constexpr int M32 = 0, M64 = 1, MIX = 2;
template <int M> struct ModeOf;
template <> struct ModeOf<M32> { using T = float; using L = float; };
template <> struct ModeOf<M64> { using T = double; using L = double; };
template <> struct ModeOf<MIX> { using T = float; using L = double; };

constexpr int CHUNK = 4;  // frequencies per task (per lane)

// A task is one image request times one chunk of at most CHUNK of its frequencies; a job is
// (task, pixel).  One launch marches the jobs of any number of tasks of one snapshot.
template <class T>
struct TaskDev {
  T Ne, Bu, rhigh, pad;                          // plasma units of the request (bhac_io.hpp:230-237)
  T freq[CHUNK], hconst[CHUNK], jscale[CHUNK];   // per frequency (bhac_io.hpp:244-249)
  phx_freqout* out;                              // result of (pixel 0, first frequency of the chunk)
  int nf, nf_total, stats, pad1;                 // frequencies in the chunk / in the request; count steps
};

// Launch parameters only the sample work needs.  They live in device memory and are loaded where a
// sample is processed (side_of), so that they do not hold scalar registers across the whole march:
// the march's launch arguments and uniform values otherwise exceed the SGPR file and are spilled to
// VGPR lanes, each use then costing a v_readlane.
template <int M>
struct SideArgs {
  Grid<typename ModeOf<M>::T> G;  // cell table and the lookup's block table and map
  Plasma<typename ModeOf<M>::T> P;
  Loc<double> GL;  // MIX: the FP64 block table a sample's cell is located in (else G's own table)
  const TaskDev<typename ModeOf<M>::T>* tasks;
  const int2* jobs;
  const void* start;  // per-job start records, in job order
  int* counter;
  unsigned long long* tot;
  int j0, j1;
};

template <int M>
struct MarchArgs {
  using T = typename ModeOf<M>::T;
  using L = typename ModeOf<M>::L;
  Grid<T> G;                // cell table and the lookup's block table and map (FP32 reads it here)
  const SideArgs<M>* side;  // MIX, FP64: the grid and sample parameters in device memory (FP32: G and P)
  Plasma<T> P;              // a, hs and rt_outer are read by every march from here
  const void* start;  // per-job start records, in job order
  T stepsize, r_out, cut_in;
  double a64, hs64, lnr_fp64, x2_pole;  // MIX: FP64 steps where ln r < lnr_fp64 or x2 within x2_pole of a pole
  int fp64_wave;                        // MIX: a lane's FP64 step takes its whole wave to FP64
  int refill;                           // idle lanes a wave gathers before they take new jobs (1: at once)
  int escape;                           // 1: a ray moving outward beyond rt_outer ends (it cannot emit again)
  int max_steps, max_order;
  const TaskDev<T>* tasks;
  const int2* jobs;
  int j0, j1;
  int* counter;
  unsigned long long* tot;
};

// ---- ray state of a lane.  StRef: the reference's state (compensated FP32 pair, or FP64);
// StHam: the mixed-precision march's Hamiltonian state.
template <class T>
struct StRef {
  static constexpr int REC = 24;  // start record: hi[12], lo[8] (FP32 compensation of x, k), sign of beta
  Ray<T> g;
  PHX_DEV void load(const T* st, T* bsign) {
#pragma unroll
    for (int q = 0; q < 12; q++) g.hi[q] = st[q];
#pragma unroll
    for (int q = 0; q < 8; q++) g.lo[q] = Lim<T>::fp32 ? st[12 + q] : T(0);
#pragma unroll
    for (int q = 8; q < 12; q++) g.lo[q] = T(0);
    *bsign = st[20];
  }
  PHX_DEV void view(T cx[4], T ck[4], T cf[4]) const {
    constexpr bool F32 = Lim<T>::fp32;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      cx[q] = F32 ? g.hi[q] + g.lo[q] : g.hi[q];
      ck[q] = F32 ? g.hi[4 + q] + g.lo[4 + q] : g.hi[4 + q];
      cf[q] = F32 ? g.hi[8 + q] + g.lo[8 + q] : g.hi[8 + q];
    }
  }
  PHX_DEV void lpos(const T cx[4], T pl[4]) const {
#pragma unroll
    for (int q = 0; q < 4; q++) pl[q] = cx[q];
  }
  PHX_DEV T x2p() const { return pole_x2(g); }
  template <class A>
  PHX_DEV bool step(const A& args, T dl) {
    rk4(args.P.a, args.P.hs, g, dl);
    return false;
  }
};
// Ray state of the mixed-precision march: the null geodesic in Hamiltonian form (ham_rhs), the
// FP64 state (x1, x2, x3, p_1, p_2) in VGPRs, the Killing constants p_t and p_phi, and the ray's
// Walker-Penrose constant in place of a transported polarization vector (the reference transports
// F1 with the Christoffel symbols, src/core/geometry/geodesic.hpp:180-195; kappa determines the
// same polarization angle at every sample, wp_angle).  A step runs in FP32 from the FP64 state,
// mirrored into the northern hemisphere when x2 > pi/2 as the FP32 march does, or in FP64 where
// r < 2 r_H or |cos theta| > 0.99.  The step's first RK4 stage is evaluated by begin(): its
// derivative is the wavevector the pass needs (turning points, step rule, the sample).
struct StHam {
  static constexpr int REC = 12;  // start record (doubles): x1, x2, x3, p_1, p_2, p_t, p_phi, kappa (2), sign of beta
  double X[5];
  double pt, pp;
  float kap1, kap2;
  double d1[5];  // first stage of the coming step: FP64, or the (mirrored) FP32 derivative widened
  Pt32 tb;       // FP32 step: the point quantities of its first stage, the base of the others
  bool f64 = false;
  PHX_DEV void load(const double* st, float* bsign) {
#pragma unroll
    for (int q = 0; q < 5; q++) X[q] = st[q];
    pt = st[5];
    pp = st[6];
    kap1 = float(st[7]);
    kap2 = float(st[8]);
    *bsign = float(st[9]);
  }
  PHX_DEV void view(float cx[4], float ck[4], float cf[4]) const {
    cx[0] = 0.f;  // t: nothing on the path depends on it (a stationary snapshot)
#pragma unroll
    for (int q = 0; q < 3; q++) cx[1 + q] = float(X[q]);
    cf[0] = kap1;
    cf[1] = kap2;
    cf[2] = cf[3] = 0.f;
    (void)ck;
  }
  PHX_DEV void lpos(const float*, double pl[4]) const {
    pl[0] = 0.;
#pragma unroll
    for (int q = 0; q < 3; q++) pl[1 + q] = X[q];
  }
  PHX_DEV float x2p() const { return float(X[1] > 0.5 * PI_D ? PI_D - X[1] : X[1]); }
  // FP32 copy of the state, mirrored to x2 < pi/2 (x2 -> pi - x2, p_2 -> -p_2, under which the
  // Hamiltonian is symmetric) when x2 > pi/2
  PHX_DEV bool base32(float b[5]) const {
    const bool south = X[1] > 0.5 * PI_D;
#pragma unroll
    for (int q = 0; q < 5; q++) b[q] = float(X[q]);
    if (south) {
      b[1] = float(PI_D - X[1]);
      b[4] = -b[4];
    }
    return south;
  }
  // precision of the coming step and its first stage; ck = the wavevector at the current point
  template <class A>
  PHX_DEV void begin(const A& args, bool alive, float ck[4]) {
    f64 = alive && (X[0] < args.lnr_fp64 || X[1] < args.x2_pole || X[1] > PI_D - args.x2_pole);
    if (args.fp64_wave) f64 = __ballot(f64) != 0ull;
    if (f64) {
      double k0;
      ham_rhs<double, true, true>(args.a64, args.hs64, pt, pp, X, d1, &k0);
      ck[0] = float(k0);
#pragma unroll
      for (int q = 0; q < 3; q++) ck[1 + q] = float(d1[q]);
    } else {
      float b[5], d[5], k0;
      const bool south = base32(b);
      tb = pt32_at(b[0], b[1], args.P.hs);
      ham_core<float>(args.P.a, args.P.hs, float(pt), float(pp), tb.r, tb.s2x, tb.c2x, tb.s, tb.c, b[3], b[4], d, &k0);
#pragma unroll
      for (int q = 0; q < 5; q++) d1[q] = d[q];
      ck[0] = k0;
      ck[1] = d[0];
      ck[2] = south ? -d[1] : d[1];
      ck[3] = d[2];
    }
  }
  // classical RK4 (weights 1/2, 1/2, 1; (k1 + 2 k2 + 2 k3 + k4) dt / 6) from the stage begin()
  // evaluated; the increment is added to the FP64 state.  Returns whether it ran in FP64.
  template <class A>
  PHX_DEV bool step(const A& args, float dl) {
    if (f64) {
      const double dt = dl;
      double sum[5], y[5];
#pragma unroll
      for (int i = 0; i < 5; i++) {
        sum[i] = d1[i];
        y[i] = X[i] + 0.5 * dt * d1[i];
      }
#pragma unroll 1
      for (int q = 1; q < 4; q++) {
        double d[5], k0;
        ham_rhs<double, true, true>(args.a64, args.hs64, pt, pp, y, d, &k0);
        const double w = q == 3 ? 1. : 2., h = q == 2 ? dt : 0.5 * dt;
#pragma unroll
        for (int i = 0; i < 5; i++) {
          sum[i] += w * d[i];
          y[i] = X[i] + h * d[i];
        }
      }
      const double dt6 = dt / 6.;
#pragma unroll
      for (int i = 0; i < 5; i++) X[i] += sum[i] * dt6;
    } else {
      float b[5], sum[5], y[5];
      const bool south = base32(b);
      const float pt32 = float(pt), pp32 = float(pp);
#pragma unroll
      for (int i = 0; i < 5; i++) {
        const float d = float(d1[i]);
        sum[i] = d;
        y[i] = b[i] + 0.5f * dl * d;
      }
      // stages 2-4 at the offsets h d of stage 1's point (|h d| <= 0.02; else evaluated afresh)
      float e1 = 0.5f * dl * float(d1[0]), e2 = 0.5f * dl * float(d1[1]);
#pragma unroll
      for (int q = 1; q < 4; q++) {
        float d[5], k0;
        const Pt32 t = (fabsf(e1) <= 0.02f && fabsf(e2) <= 0.02f) ? pt32_offset(tb, e1, e2, args.P.hs)
                                                                   : pt32_at(y[0], y[1], args.P.hs);
        ham_core<float>(args.P.a, args.P.hs, pt32, pp32, t.r, t.s2x, t.c2x, t.s, t.c, y[3], y[4], d, &k0);
        const float w = q == 3 ? 1.f : 2.f, h = q == 2 ? dl : 0.5f * dl;
#pragma unroll
        for (int i = 0; i < 5; i++) {
          sum[i] += w * d[i];
          if (q < 3) y[i] = b[i] + h * d[i];
        }
        e1 = h * d[0];
        e2 = h * d[1];
      }
      const float dt6 = div_const(dl, 6.);
      float inc[5];
#pragma unroll
      for (int i = 0; i < 5; i++) inc[i] = sum[i] * dt6;
      if (south) {
        inc[1] = -inc[1];
        inc[4] = -inc[4];
      }
#pragma unroll
      for (int i = 0; i < 5; i++) X[i] += double(inc[i]);
    }
    return f64;
  }
};
template <int M>
using StateOf = typename std::conditional<M == MIX, StHam, StRef<typename ModeOf<M>::T>>::type;
template <int M>
using RecOf = typename std::conditional<M == M32, float, double>::type;  // start record scalar

// the grid and sample parameters.  The mixed and FP64 marches load them (scalar loads) where a
// lookup miss or a sample needs them: the empty asm makes the pointer a new value at each use, so
// the loads are not hoisted out of the march loop.  The FP32 march reads the launch arguments,
// which measured faster there (162 vs 172 ms; FP64 349 -> 346 ms, mixed about -7%).
template <int M>
PHX_DEV SideArgs<M> side_of(const MarchArgs<M>& A) {
  if constexpr (M == M32) {
    SideArgs<M> out;
    out.G = A.G;
    out.P = A.P;
    out.tasks = A.tasks;
    out.jobs = A.jobs;
    out.start = A.start;
    out.counter = A.counter;
    out.tot = A.tot;
    out.j0 = A.j0;
    out.j1 = A.j1;
    return out;
  }
  const SideArgs<M>* p = A.side;
  asm volatile("" : "+s"(p));
  using W = const __attribute__((address_space(4))) unsigned*;
  static_assert(sizeof(SideArgs<M>) % 4 == 0, "whole dwords");
  const W w = (W)p;
  SideArgs<M> out;
  unsigned* d = reinterpret_cast<unsigned*>(&out);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(SideArgs<M>) / 4); i++) d[i] = w[i];
  return out;
}
// the table a sample's cell is located in
template <int M>
PHX_DEV Loc<typename ModeOf<M>::L> loc_for(const MarchArgs<M>& A, const SideArgs<M>& S) {
  if constexpr (M == MIX)
    return S.GL;
  else
    return loc_of(S.G);
}
// re-solve k^t for k.k = 0 when |k.k| > 1e-6 and r > 2 (geodesic.hpp:297-306, pixel.hpp:69-71)
template <class T>
PHX_DEV void null_resolve(const Met<T>& g, T k[4], T r) {
  if (fabs(dot(g, k, k)) > T(1e-6) && r > T(2)) {
    const T aa = g.d00;
    const T bc = div_rcp(-(g.d01 * k[1]), aa) - div_rcp(g.d03 * k[3], aa);
    const T gz = div_rcp(-(g.d11 * k[1] * k[1]), aa) - div_rcp(T(2) * g.d13 * k[1] * k[3], aa) -
                 div_rcp(g.d22 * k[2] * k[2], aa) - div_rcp(g.d33 * k[3] * k[3], aa);
    k[0] = bc + sqrt(bc * bc + gz);
  }
}
// ---- one plasma sample, split into its frequency-independent part and the per-frequency
// segment operator (src/core/render/pixel.hpp:52-123)
template <class T>
struct Prep {
  T te, ne, B, sth, cth, kr1, kr0, c2, s2, ratio, dl;
};
// What the second half of a sample needs from the first, a candidate record of 23 scalars (one
// queue slot): [0..1] the position the sampler used (x1, x2), [2..5] the (re-solved) wavevector,
// [6..9] U^mu, [10..13] b^mu, [14..17] the ray's polarization vector (or its Walker-Penrose
// constant), [18] Theta_e, [19] n_e, [20] |B|, [21] the step length, [22] the task.
constexpr int CREC = 23;
static_assert(CREC < 24, "a candidate record fits a queue slot (QW scalars) with the owner in slot 23");
// First half of a sample (src/core/render/pixel.hpp:52-79): null re-solve of k, fluid at the
// sample (wrap, cached leaf / lookup, cell, trilinear gather, derive_fluid, cuts).
// 0: no emission; 2: emits nothing finite (b^mu = 0, a field-free cell: the pitch angle is pi/2,
// tetrad.hpp:144, and the plasma-frame basis is 0/0, so every frequency's operator is non-finite);
// 3: a candidate for emission, written to c.
template <int M>
PHX_DEV int sample_fluid_stage(const MarchArgs<M>& A, const TaskDev<typename ModeOf<M>::T>& tk,
                               const typename ModeOf<M>::T px[4], const typename ModeOf<M>::L pl[4],
                               const typename ModeOf<M>::T pk[4], const typename ModeOf<M>::T pf[4],
                               typename ModeOf<M>::T dl, int leaf, int task, typename ModeOf<M>::T* c) {
  using T = typename ModeOf<M>::T;
  T k[4] = {pk[0], pk[1], pk[2], pk[3]};
  const SideArgs<M> S = side_of(A);
  if constexpr (M != MIX) null_resolve(metric(S.P.a, S.P.hs, px[1], px[2]), k, exp(px[1]));
  Fluid<T> fl;
  if (!sample_fluid(S.G, loc_for(A, S), S.P, tk.Ne, tk.Bu, tk.rhigh, pl, leaf, fl)) return 0;
  if (fl.b[0] == T(0) && fl.b[1] == T(0) && fl.b[2] == T(0) && fl.b[3] == T(0)) return 2;
  // the mixed march's wavevector k^mu = g^{mu nu} p_nu is null by construction up to FP32 rounding;
  // it is checked (and re-solved) with the metric at the sampler's position, the one every later
  // contraction with k uses, and only for samples that reach this point
  if constexpr (M == MIX) null_resolve(fl.g, k, fl.g.r);
  c[0] = fl.x[1];
  c[1] = fl.x[2];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    c[2 + q] = k[q];
    c[6 + q] = fl.U[q];
    c[10 + q] = fl.b[q];
    c[14 + q] = pf[q];
  }
  c[18] = fl.te;
  c[19] = fl.ne;
  c[20] = fl.B;
  c[21] = dl;
  c[22] = T(task);
  return 3;
}
// Second half (pixel.hpp:80-94) of the candidate c: pitch angle, plasma-frame tetrad,
// polarization angle, frequency ratio, Bessel ratios.  The metric at the sample is evaluated
// again (the same function of the same position as in the first half).  0: no emission (pitch
// angle 0 or pi); 1: emits; 2: the rotation or the frequency ratio is NaN, so every frequency's
// operator is non-finite and dropped.
template <int M>
PHX_DEV int sample_frame_stage(const MarchArgs<M>& A, const typename ModeOf<M>::T* c, Prep<typename ModeOf<M>::T>& sp) {
  using T = typename ModeOf<M>::T;
  const SideArgs<M> S = side_of(A);
  const Met<T> g = metric(S.P.a, S.P.hs, c[0], c[1]);
  T k[4], U[4], b[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    k[q] = c[2 + q];
    U[q] = c[6 + q];
    b[q] = c[10 + q];
  }
  const T pa = pitch_angle(g, k, b, U);
  if (pa == T(0) || pa == T(PI_D)) return 0;
  T perp[4], para[4];
  frame(g, S.P.a, k, U, b, perp, para);
  if constexpr (M == MIX) {
    // the record holds the ray's Walker-Penrose constant (StHam)
    wp_angle(S.P.a, g, k, perp, para, c[14], c[15], &sp.c2, &sp.s2);
  } else {
    const T f[4] = {c[14], c[15], c[16], c[17]};
    const T Cp = dot(g, perp, f), Sp = dot(g, para, f);  // chi_from_F1 (tetrad.hpp:160-182)
    const T nn = Cp * Cp + Sp * Sp;
    sp.c2 = (Cp * Cp - Sp * Sp) / nn;
    sp.s2 = T(2) * Cp * Sp / nn;
  }
  sp.ratio = -dot(g, k, U);
  if (__builtin_isnan(sp.c2) || __builtin_isnan(sp.s2) || __builtin_isnan(sp.ratio)) return 2;
  sp.te = c[18];
  sp.ne = c[19];
  sp.B = c[20];
  sp.dl = c[21];
  scs(pa, &sp.sth, &sp.cth);
  kratios(sp.te, &sp.kr1, &sp.kr0);
  return 1;
}
// Both halves at once, for the sample processed in place by its own lane (the reference
// kernel's way): 0 no emission, 1 emits, 2 every frequency's operator is dropped.
template <int M>
PHX_DEV int sample_prepare(const MarchArgs<M>& A, const TaskDev<typename ModeOf<M>::T>& tk,
                           const typename ModeOf<M>::T px[4], const typename ModeOf<M>::L pl[4],
                           const typename ModeOf<M>::T pk[4], const typename ModeOf<M>::T pf[4],
                           typename ModeOf<M>::T dl, int leaf, Prep<typename ModeOf<M>::T>& sp) {
  using T = typename ModeOf<M>::T;
  T k[4] = {pk[0], pk[1], pk[2], pk[3]};
  const SideArgs<M> S = side_of(A);
  if constexpr (M != MIX) null_resolve(metric(S.P.a, S.P.hs, px[1], px[2]), k, exp(px[1]));
  Fluid<T> fl;
  if (!sample_fluid(S.G, loc_for(A, S), S.P, tk.Ne, tk.Bu, tk.rhigh, pl, leaf, fl)) return 0;
  if (fl.b[0] == T(0) && fl.b[1] == T(0) && fl.b[2] == T(0) && fl.b[3] == T(0)) return 2;
  // the mixed march's wavevector k^mu = g^{mu nu} p_nu is null by construction up to FP32 rounding;
  // it is checked (and re-solved) with the metric at the sampler's position, the one every later
  // contraction with k uses, and only for samples that reach this point
  if constexpr (M == MIX) null_resolve(fl.g, k, fl.g.r);
  const T pa = pitch_angle(fl.g, k, fl.b, fl.U);
  if (pa == T(0) || pa == T(PI_D)) return 0;
  T perp[4], para[4];
  frame(fl.g, S.P.a, k, fl.U, fl.b, perp, para);
  if constexpr (M == MIX) {
    wp_angle(S.P.a, fl.g, k, perp, para, pf[0], pf[1], &sp.c2, &sp.s2);
  } else {
    const T Cp = dot(fl.g, perp, pf), Sp = dot(fl.g, para, pf);  // chi_from_F1 (tetrad.hpp:160-182)
    const T nn = Cp * Cp + Sp * Sp;
    sp.c2 = (Cp * Cp - Sp * Sp) / nn;
    sp.s2 = T(2) * Cp * Sp / nn;
  }
  sp.ratio = -dot(fl.g, k, fl.U);
  if (__builtin_isnan(sp.c2) || __builtin_isnan(sp.s2) || __builtin_isnan(sp.ratio)) return 2;
  sp.te = fl.te;
  sp.ne = fl.ne;
  sp.B = fl.B;
  sp.dl = dl;
  scs(pa, &sp.sth, &sp.cth);
  kratios(fl.te, &sp.kr1, &sp.kr0);
  return 1;
}
// segment operator of frequency f; false when it is not finite (the reference drops it)
template <class T>
PHX_DEV bool sample_op(const TaskDev<T>& tk, const Prep<T>& sp, int f, T OA[16], T OS[4], T* dtau, T* dtauF) {
  const T nup = tk.freq[f] * sp.ratio;
  const Coef<T> seg = invariant_rotated(thermal(sp.te, sp.ne, nup, sp.B, sp.sth, sp.cth, sp.kr1, sp.kr0), nup,
                                        tk.jscale[f], sp.c2, sp.s2);
  const T h = sp.dl * tk.hconst[f];
  segment_op(seg, h, OA, OS);
  bool fin = true;
#pragma unroll
  for (int q = 0; q < 16; q++) fin = fin && __builtin_isfinite(OA[q]);
#pragma unroll
  for (int q = 0; q < 4; q++) fin = fin && __builtin_isfinite(OS[q]);
  *dtau = seg.aI * h;
  *dtauF = fabs(seg.rV * h);
  return fin;
}
// S += A O_S; A <- A O_A (operators.hpp:74-86), one row at a time in place.  KAHAN: the
// Stokes sum carries a compensation term per component (the mixed-precision march).
template <class T, bool KAHAN>
PHX_DEV void accumulate(T AA[16], T SS[4], T SC[4], const T OA[16], const T OS[4]) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    T row[4];
    T s = T(0);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      s += AA[4 * i + j] * OS[j];
      T t = T(0);
#pragma unroll
      for (int kq = 0; kq < 4; kq++) t += AA[4 * i + kq] * OA[4 * kq + j];
      row[j] = t;
    }
    if (KAHAN) {
      const T y = s - SC[i];
      const T t = SS[i] + y;
      SC[i] = (t - SS[i]) - y;
      SS[i] = t;
    } else {
      SS[i] += s;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) AA[4 * i + j] = row[j];
  }
}

// ---------------------------------------------------------------------------- the march
// Persistent lane-refill march: each lane owns one ray at a time; a lane whose ray ends
// writes it out and takes the next job (wave-aggregated atomic on the job counter), so rays
// that hit the horizon or escape free their lane at once instead of idling until the wave's
// longest ray ends.  One pass of the loop is one pass of render_pixel's loop
// (src/core/render/pixel.hpp:179-211): view, the pending sample with the current position's
// leaf as its cache, turning points, step size, RK4.
//
// QUEUE: a per-wave sample queue in LDS.  A lane whose pending sample can emit pushes it
// (position, wavevector, polarization vector or Walker-Penrose constant, step length, cache
// leaf, task, owner lane) into its wave's queue instead of processing it at once.  When the
// queue holds 64 samples, or when a ray with samples still queued ends, all active lanes of the
// wave work through the queue together, one sample per lane: first the fluid half of every
// sample; then, on the candidates for emission (these are a minority: most samples lie outside
// the cuts or in field-free cells) in whole waves of 64, the tetrad half and the segment
// operators.  Each owning lane folds its own operators into its accumulator in the order it
// queued them, which is the reference's order (pixel.hpp:186-188).  The divergent, expensive
// parts of a sample (fluid, tetrad, coefficients, 4x4 inverse) thus run on full waves.  Without
// QUEUE the lane processes its sample in place.
constexpr int QCAP = 128;  // slots per wave (at most 63 + 64 are in use at a flush)
constexpr int QW = 24;     // scalars per slot: 21 in (x, k, F1, dl, leaf, FP64 position, task), then the
                           // prepared sample, then the operators; 23 = position in the emitting list
constexpr int LCAP = 64;   // leftover candidate slots per wave (fewer than a wave's lanes remain)

PHX_DEV void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#ifndef PHX_MIX_WAVES_PER_SIMD
#define PHX_MIX_WAVES_PER_SIMD 2  // occupancy the mixed march is compiled for (1: 512 VGPRs, 2: 256)
#endif
template <int M, int NF, bool QUEUE>
__global__ __launch_bounds__(BLOCK, M == M64 ? 1 : M == MIX ? PHX_MIX_WAVES_PER_SIMD : 2) void k_march(
    const MarchArgs<M> A) {
  using T = typename ModeOf<M>::T;
  using L = typename ModeOf<M>::L;
  using St = StateOf<M>;
  using R = RecOf<M>;
  constexpr bool KAHAN = M == MIX;
  constexpr int NW = QUEUE ? BLOCK / 64 : 1;  // per-wave LDS areas of the sample queue
  __shared__ __align__(16) T qbuf[NW][QUEUE ? QCAP * QW : 1];  // queued samples, then their candidates
  __shared__ __align__(16) T lqbuf[NW][QUEUE ? LCAP * QW : 1];  // candidates left over from earlier flushes
  __shared__ int clbuf[NW][QUEUE ? QCAP : 1];                  // slots of this flush's candidates, in order
  __shared__ unsigned long long ombuf[NW][QUEUE ? 64 : 1];     // per owner: the lanes holding its records
  __shared__ int pcbuf[NW][QUEUE ? 64 : 1];                    // per owner: candidates made in a flush
  __shared__ int qsbuf[NW][2];  // the wave's queue counts (qcount, nleft), for lanes that sat out passes
  const int wv = QUEUE ? (threadIdx.x >> 6) : 0;
  T* const Q = qbuf[wv];
  T* const LQ = lqbuf[wv];
  int* const cl = clbuf[wv];
  unsigned long long* const om = ombuf[wv];
  int* const pc = pcbuf[wv];
  const int lane = threadIdx.x & 63;
  int* const qs = qsbuf[wv];
  if constexpr (QUEUE) {
    om[lane] = 0ull;
    pc[lane] = 0;
    if (lane == 0) qs[0] = qs[1] = 0;
    wave_sync();
  }
  const unsigned long long lt = (1ull << lane) - 1ull;
  St st;
  T px[4] = {T(0), T(0), T(0), T(0)}, pk[4] = {T(0), T(0), T(0), T(0)}, pf[4] = {T(0), T(0), T(0), T(0)};
  L pl[4] = {L(0), L(0), L(0), L(0)};
  T pdl = T(0);
  T AA[NF][16], SS[NF][4], SC[NF][4], tau[NF], tauF[NF];
  int steps = 0, turns = 0, stop = 0, np = 0, bad = 0, pix = 0, task = 0, nf = 0;
  T thdot = T(0), dlp = T(0);
  T r_prev = T(0);  // r of the previous pass's position, the pending sample's (read from the third pass on)
  bool bneg = false, bpos = false, have = false, stats = false;
  unsigned s_steps = 0, s_np = 0, s_bad = 0;  // per-lane sums (a lane marches well under 2^32 steps)
  unsigned s_fp64 = 0;                        // MIX: steps taken in FP64
  // queue bookkeeping: qcount and nleft are the same in every active lane of the wave; nmine and
  // ncm count this lane's queued samples and its candidates not yet folded in
  int qcount = 0, nleft = 0, nmine = 0, ncm = 0;
  LeafCache<T> lcache;  // the lane's last leaf and its box (valid across rays: one snapshot per launch)

  for (;;) {
    // refill: a lane whose ray ended waits until at least A.refill lanes of its wave are idle (or
    // none is busy), and those lanes then take consecutive jobs together.  1 refills each lane at
    // once; 64 refills the wave only when all its lanes are done, which is how the reference
    // kernel's per-thread job loop (kernel.cu:80-96) runs on a wave whose lanes share one
    // instruction stream.  Rays started together stay coherent (neighbouring pixels).
    bool wait = false;
    if (A.refill > 1) {
      const unsigned long long busy = __ballot(have);
      wait = busy != 0ull && 64 - (int)__popcll(busy) < A.refill;
    }
    if (!have && !wait) {
      const SideArgs<M> S = side_of(A);
      const unsigned long long m = __ballot(1);
      const int leader = __ffsll((long long)m) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(S.counter, (int)__popcll(m));
      base = __shfl(base, leader);
      const int n = S.j0 + base + (int)__popcll(m & lt);
      if (n >= S.j1) break;
      const int2 jb = S.jobs[n];
      task = jb.x;
      pix = jb.y;
      {
        const TaskDev<T> tr = load_global(S.tasks + task);
        nf = tr.nf;
        stats = tr.stats != 0;
      }
      T bsign;
      st.load(static_cast<const R*>(S.start) + (size_t)n * St::REC, &bsign);
#pragma unroll
      for (int f = 0; f < NF; f++) {
#pragma unroll
        for (int q = 0; q < 16; q++) AA[f][q] = (q % 5 == 0) ? T(1) : T(0);
#pragma unroll
        for (int q = 0; q < 4; q++) SS[f][q] = SC[f][q] = T(0);
        tau[f] = tauF[f] = T(0);
      }
      steps = turns = stop = np = bad = 0;
      thdot = dlp = T(0);
      bneg = bsign < T(0);
      bpos = bsign > T(0);
      have = true;
    }
    if (!have) continue;  // waiting for the wave's refill
    if constexpr (QUEUE) {
      // a lane that waited for the refill skipped the queue updates of the passes it sat out
      qcount = qs[0];
      nleft = qs[1];
    }

    // ---- view of the current state, liveness, and the one AMR lookup of this pass: it
    // serves the pending sample's cache (leaf_of) and the step's cell cap (grid_dxc)
    T cx[4], ck[4], cf[4];
    st.view(cx, ck, cf);
    const T r = exp_r(cx[1]);
    // ln r of the pending sample (the state before the last step; MIX keeps it in FP64 only)
    const T lrs = M == MIX ? T(pl[1]) : px[1];
    const T rs = r_prev;  // radius of the pending sample: the previous pass's r (exp_r of the same lrs)
    // escape: a ray moving outward at r >= rt_outer, whose pending sample lies there too, can
    // never emit again (a null geodesic in Kerr has at most one radial turning point outside the
    // photon orbits, so it does not return), and the rest of its march changes none of its
    // outputs: it ends here instead of at r_out (the reference marches on to r_out,
    // pixel.hpp:163, and only its step count differs)
    const bool escaped = A.escape && steps >= 2 && r >= A.P.rt_outer && rs >= A.P.rt_outer && cx[1] > lrs;
    const bool alive = r < A.r_out && r > A.cut_in && steps < A.max_steps && !stop && !escaped;
    if constexpr (M == MIX) {
      // the pending sample's FP32 position and polarization constant follow from its FP64
      // position and the ray's constant: not carried from pass to pass
#pragma unroll
      for (int q = 0; q < 4; q++) {
        px[q] = T(pl[q]);
        pf[q] = cf[q];
      }
    }
    // the pending sample is the state before the last step, i.e. sample number steps - 1; the
    // camera sample (number 0) is never processed (pixel.hpp:186).  A sample at r >= rt_outer
    // cannot emit (pixel.hpp:80) and changes no ray state, so nothing of it is computed.
    const bool emit = steps >= 2 && rs < A.P.rt_outer;
    int leaf_next = -1;
    bool have_dxc = false;
    T dxc[3] = {T(0), T(0), T(0)};
    // far from the grid (r >= rt_outer, no emitting sample pending) nothing is looked up
    if (alive && (r < A.P.rt_outer || emit)) {
      T wx[4] = {cx[0], cx[1], cx[2], cx[3]};
      wrap(wx);
      const bool need_dxc = r < A.P.rt_outer;
      // leaf_of (sampler.hpp:184-190) is -1 off the domain.  The mixed march hands the sample its
      // leaf as a hint only: the sample tests the domain and the leaf itself, in FP64
      bool need_leaf = emit;
      if constexpr (M != MIX) need_leaf = need_leaf && !off_domain(A.G, wx);
      if (need_dxc || need_leaf) {
        bool cached = leaf_cache_hit(lcache, wx);
        int Lf = lcache.leaf;
        if (!cached) {
          const Grid<T> G = side_of(A).G;
          constexpr int DXC = M == MIX ? 9 : 6;  // the mixed march's step rule takes 1/dxc
          Lf = find_leaf_fill<T, DXC>(G, wx, lcache, &cached);
          if (need_dxc && Lf >= 0 && !cached) {  // a leaf of the fallback scan (not cached)
#pragma unroll
            for (int d = 0; d < 3; d++) lcache.dxc[d] = glb(G.blk)[12 * Lf + DXC + d];
            lcache.leaf = -1;
          }
        }
        if (need_leaf) leaf_next = Lf;
        if (need_dxc && Lf >= 0) {
          have_dxc = true;
#pragma unroll
          for (int d = 0; d < 3; d++) dxc[d] = lcache.dxc[d];
        }
      }
    }

    if constexpr (QUEUE) {
      // ---- queue the pending sample (every loop-resident lane takes part in the ballot)
      const unsigned long long pm = __ballot(emit);
      if (emit) {
        const int slot = qcount + (int)__popcll(pm & lt);
        T* e = Q + slot * QW;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          e[q] = px[q];
          e[4 + q] = pk[q];
          e[8 + q] = pf[q];
        }
        e[12] = pdl;
        e[13] = T(leaf_next);
        if constexpr (!std::is_same<T, L>::value) {
          double* e64 = reinterpret_cast<double*>(e + 14);
#pragma unroll
          for (int q = 0; q < 3; q++) e64[q] = pl[1 + q];
        }
        e[20] = T(task);
        e[21] = T(lane);
        nmine++;
      }
      qcount += (int)__popcll(pm);
    } else if (emit) {
      // ---- the pending sample, in place
      const TaskDev<T> tk = load_global(side_of(A).tasks + task);
      Prep<T> sp;
      const int status = sample_prepare<M>(A, tk, px, pl, pk, pf, pdl, leaf_next, sp);
      if (status) np++;
      if (status == 2) {
        // a NaN rotation or frequency ratio makes every frequency's operator non-finite: the
        // reference drops each of them (pixel.hpp:107-116); count them without computing them
        bad += nf;
      } else if (status == 1) {
#pragma unroll
        for (int f = 0; f < NF; f++) {
          if (f >= nf) break;
          T OA[16], OS[4], dtau, dtauF;
          if (!sample_op(tk, sp, f, OA, OS, &dtau, &dtauF)) {
            bad++;
            continue;
          }
          accumulate<T, KAHAN>(AA[f], SS[f], SC[f], OA, OS);
          tau[f] += dtau;
          tauF[f] += dtauF;
        }
      }
    }

    // the coming step's precision and first stage, and the wavevector at the current point (not
    // needed before: the pending sample carries its own); evaluated after the sample, so the stage
    // is not held in registers across it
    if constexpr (M == MIX) st.begin(A, alive, ck);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      pk[q] = ck[q];
      if constexpr (M != MIX) {
        px[q] = cx[q];
        pf[q] = cf[q];
      }
    }
    st.lpos(cx, pl);
    r_prev = r;
    pdl = dlp;
    if (alive) {
      // ---- turning points of theta (pixel.hpp:193-197), grid-aware step and RK4 (:200-210)
      if (thdot * ck[2] < T(0) && steps > 2) turns++;
      thdot = ck[2];
      if ((bneg && turns > A.max_order) || (bpos && turns > A.max_order + 1)) stop = 1;
      T dl;
      if constexpr (M == MIX)
        dl = step_len_mixed(A.stepsize, 1.f / A.stepsize, st.x2p(), ck, have_dxc, dxc);  // dxc: 1/cell size
      else
        dl = step_len(A.stepsize, st.x2p(), ck, have_dxc, dxc);
      if (st.step(A, dl) && stats) s_fp64++;
      dlp = fabs(dl);
      steps++;
    }

    if constexpr (QUEUE) {
      // ---- flush.  Pass 1: all active lanes take the queued samples, one per lane, and run their
      // first half (fluid, cuts), which most samples end in; the candidates for emission stay in
      // their slots, listed in queue order after the candidates left over from earlier flushes.
      // Pass 2: all active lanes take candidates, one per lane, in whole waves of them (the
      // remainder waits for the next flush, unless a ray that ended needs its own), run the second
      // half (tetrad, polarization angle, Bessel ratios) and each frequency's segment operator;
      // after each frequency the owners fold their operators in, in queue order.  Run statistics
      // are counted by the lane that computes the sample.
      const bool want = qcount >= QCAP - 64 || (!alive && (nmine > 0 || ncm > 0));
      if (__ballot(want) != 0ull) {
        const unsigned long long act = __ballot(1);
        const int nact = (int)__popcll(act), rank = (int)__popcll(act & lt);
        // whole waves of queued samples, oldest first, unless a ray that ended needs its own: the
        // newest samples (fewer than a wave) wait for the next flush, in queue order
        const bool force = __ballot(!alive && (nmine > 0 || ncm > 0)) != 0ull;
        const int n1 = force ? qcount : (qcount / nact) * nact;
        int ncomp = 0;  // the same in every active lane
        for (int base = 0; base < n1; base += nact) {
          const int e = base + rank;
          int status = 0;
          if (e < n1) {
            T* const in = Q + e * QW;
            T x[4], k[4], f1[4];
            L xl[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
              x[q] = in[q];
              k[q] = in[4 + q];
              f1[q] = in[8 + q];
              xl[q] = L(x[q]);
            }
            if constexpr (!std::is_same<T, L>::value) {
              const double* e64 = reinterpret_cast<const double*>(in + 14);
#pragma unroll
              for (int q = 0; q < 3; q++) xl[1 + q] = e64[q];
            }
            const int etask = (int)in[20], owner = (int)in[21];
            const TaskDev<T> tk = load_global(side_of(A).tasks + etask);
            // the candidate record overwrites the slot; everything of the queued sample is read first
            const T sdl = in[12];
            const int sleaf = (int)in[13];
            status = sample_fluid_stage<M>(A, tk, x, xl, k, f1, sdl, sleaf, etask, in);
            if (status == 2) {
              // every frequency's operator non-finite: dropped by the reference (pixel.hpp:107-116)
              if (tk.stats) s_np++;
              s_bad += tk.nf;
            } else if (status == 3) {
              in[23] = T(owner);
            }
            // tell the owner: one queued sample done (high half), one candidate made (low half)
            atomicAdd(&pc[owner], status == 3 ? 0x10001 : 0x10000);
          }
          const unsigned long long em = __ballot(status == 3);
          if (status == 3) cl[ncomp + (int)__popcll(em & lt)] = e;
          ncomp += (int)__popcll(em);
        }
        wave_sync();
        {
          const int v = pc[lane];
          pc[lane] = 0;
          ncm += v & 0xffff;
          nmine -= v >> 16;
        }
        const int ncand = nleft + ncomp;
        const bool force2 = __ballot(!alive && ncm > 0) != 0ull;
        const int nproc = force2 ? ncand : (ncand / nact) * nact;
        for (int cb = 0; cb < nproc; cb += nact) {
          const int pos = cb + rank;
          const bool mine = pos < nproc;
          T* const rec = pos < nleft ? LQ + pos * QW : Q + cl[mine ? pos - nleft : 0] * QW;
          const int owner = mine ? (int)rec[23] : 0;
          // the lanes holding each owner's records: intersect the ballots of the owner id's bits
          unsigned long long same = __ballot(mine);
#pragma unroll
          for (int bit = 0; bit < 6; bit++) {
            const bool on = (owner >> bit) & 1;
            const unsigned long long m = __ballot(mine && on);
            same &= on ? m : ~m;
          }
          if (mine && (same & lt) == 0ull) om[owner] = same;
          Prep<T> sp;
          int status = 0, enf = 0;
          int ctask = 0;
          if (mine) {
            ctask = (int)rec[22];
            status = sample_frame_stage<M>(A, rec, sp);
            const TaskDev<T> tk = load_global(side_of(A).tasks + ctask);
            enf = tk.nf;
            if (status != 0 && tk.stats) s_np++;
            if (status == 2) s_bad += enf;
          }
          wave_sync();
          const unsigned long long myrec = om[lane];
          om[lane] = 0ull;
#pragma unroll
          for (int f = 0; f < NF; f++) {
            if (mine && f < enf) {
              int flag = 0;  // 1: fold in; 0: nothing to fold (no emission, or dropped)
              if (status == 1) {
                T OA[16], OS[4], dtau, dtauF;
                if (sample_op(load_global(side_of(A).tasks + ctask), sp, f, OA, OS, &dtau, &dtauF)) {
#pragma unroll
                  for (int q = 0; q < 16; q++) rec[q] = OA[q];
#pragma unroll
                  for (int q = 0; q < 4; q++) rec[16 + q] = OS[q];
                  rec[20] = dtau;
                  rec[21] = dtauF;
                  flag = 1;
                } else {
                  s_bad++;
                }
              }
              rec[22] = T(flag);
            }
            wave_sync();
            unsigned long long mm = myrec;
            while (mm != 0ull) {
              const int hl = __ffsll((long long)mm) - 1;
              mm &= mm - 1ull;
              const int pp = cb + (int)__popcll(act & ((1ull << hl) - 1ull));
              const T* rr = pp < nleft ? LQ + pp * QW : Q + cl[pp - nleft] * QW;
              if (f < nf && rr[22] == T(1)) {
                T OA[16], OS[4];
#pragma unroll
                for (int q = 0; q < 16; q++) OA[q] = rr[q];
#pragma unroll
                for (int q = 0; q < 4; q++) OS[q] = rr[16 + q];
                accumulate<T, KAHAN>(AA[f], SS[f], SC[f], OA, OS);
                tau[f] += rr[20];
                tauF[f] += rr[21];
              }
            }
            wave_sync();
          }
          ncm -= (int)__popcll(myrec);
        }
        // the candidates not processed move to the front of the leftover area (sources lie at or
        // after their destinations, or in the queue slots, so the copy needs no ordering)
        const int nrem = ncand - nproc;
        if (rank < nrem) {
          const int pos = nproc + rank;
          if (pos >= nleft || nproc > 0) {
            const T* src = pos < nleft ? LQ + pos * QW : Q + cl[pos - nleft] * QW;
            T* dst = LQ + rank * QW;
#pragma unroll
            for (int q = 0; q < QW; q++) dst[q] = src[q];
          }
        }
        wave_sync();
        nleft = nrem;
        // the queued samples not processed move to the front of the queue (after the candidates in
        // those slots have been moved out; more slots were processed than remain, so the copy needs
        // no ordering)
        const int qrem = qcount - n1;
        if (rank < qrem) {
          const T* src = Q + (n1 + rank) * QW;
          T* dst = Q + rank * QW;
#pragma unroll
          for (int q = 0; q < QW; q++) dst[q] = src[q];
        }
        wave_sync();
        qcount = qrem;
      }
      qs[0] = qcount;
      qs[1] = nleft;
      wave_sync();
    }

    if (!alive) {
      // ---- ray done: write it out and free the lane (pixel.hpp:213-223)
      const TaskDev<T> tkw = load_global(side_of(A).tasks + task);
      phx_freqout* o0 = tkw.out + (size_t)pix * tkw.nf_total;
#pragma unroll
      for (int f = 0; f < NF; f++) {
        if (f >= nf) break;
        using GD2 = __attribute__((address_space(1))) d2v*;  // global stores (not flat)
        const GD2 o = (GD2)(o0 + f);
        // with the queue, np is counted where the samples are computed; a ray with no plasma
        // segment has folded nothing in, so its sums are the zeros written otherwise
        if (QUEUE || np > 0) {
          o[0] = d2v{(double)SS[f][0], (double)SS[f][1]};
          o[1] = d2v{(double)SS[f][2], (double)SS[f][3]};
          o[2] = d2v{(double)tau[f], (double)tauF[f]};
        } else {
          o[0] = o[1] = o[2] = d2v{0., 0.};
        }
      }
      if (stats) {
        s_steps += steps;
        s_np += np;
      }
      s_bad += bad;
      have = false;
    }
  }

  // wave reduction, then one atomic per wave; steps and segments are counted by the first
  // frequency chunk of each request only, dropped segments (per frequency) by every chunk
  unsigned long long w_steps = s_steps, w_np = s_np, w_bad = s_bad, w_fp64 = s_fp64;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    w_steps += __shfl_xor(w_steps, off);
    w_np += __shfl_xor(w_np, off);
    w_bad += __shfl_xor(w_bad, off);
    if (M == MIX) w_fp64 += __shfl_xor(w_fp64, off);
  }
  if (lane == 0) {
    unsigned long long* tot = side_of(A).tot;
    atomicAdd(&tot[0], w_steps);
    atomicAdd(&tot[1], w_np);
    atomicAdd(&tot[2], w_bad);
    if (M == MIX) atomicAdd(&tot[3], w_fp64);
  }
}

// Starting state of every job's ray: pixel -> impact parameters with the position-angle
// rotation (camera.hpp:11-23, jobs.hpp:64-76), the RAPTOR camera photon in FP64
// (camera.hpp:71-135), the camera polarization basis (tetrad.hpp:111-133: in the march's
// precision for M32 / M64, in FP64 for MIX), and for M32 the compensated (hi, lo) split of
// the FP64 photon (geodesic.hpp:48-60).  cams[task] is the task's camera.
template <int M>
__global__ __launch_bounds__(BLOCK) void k_rays(const Cam* __restrict__ cams, const int2* __restrict__ jobs,
                                                int njobs, void* __restrict__ out_) {
  using T = typename ModeOf<M>::T;
  using R = RecOf<M>;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= njobs) return;
  R* const out = static_cast<R*>(out_);
  const int2 jb = jobs[n];
  Cam C = cams[jb.x];
  if constexpr (M == M32) {  // the FP32 build's camera sees the spin and stretching rounded to float (camera.hpp:77)
    C.a = double(float(C.a));
    C.hs = double(float(C.hs));
  }
  const int idx = jb.y;
  const int i = idx / C.H, j = idx % C.H;
  const double a0 = -0.5 * C.fov_x + (i + 0.5) * C.fov_x / C.W;
  const double b0 = -0.5 * C.fov_y + (j + 0.5) * C.fov_y / C.H;
  const double alpha = a0 * C.cos_pa + b0 * C.sin_pa;
  const double beta = -a0 * C.sin_pa + b0 * C.cos_pa;
  double ph[8];
  camera_photon(C, alpha, beta, ph);
  const R bs = beta < 0. ? R(-1) : beta > 0. ? R(1) : R(0);
  if constexpr (M == MIX) {
    // Hamiltonian start: p_mu = g_mu nu k^nu and the Walker-Penrose constant of (k, F1), in FP64
    double F1[4], kd[4], q1, q2;
    camera_F1<double>(C.a, C.hs, ph, F1);
    const Met<double> g = metric<double>(C.a, C.hs, ph[1], ph[2]);
    lower(g, ph + 4, kd);
    wp_kappa<double>(C.a, g.r, g.s, g.c, g.hf, ph + 4, F1, &q1, &q2);
    double* o = out + (size_t)n * StHam::REC;
    o[0] = ph[1];
    o[1] = ph[2];
    o[2] = ph[3];
    o[3] = kd[1];
    o[4] = kd[2];
    o[5] = kd[0];
    o[6] = kd[3];
    o[7] = q1;
    o[8] = q2;
    o[9] = bs;
    o[10] = o[11] = 0.;
  } else {
    T F1[4];
    camera_F1<T>(T(C.a), T(C.hs), ph, F1);
    R* o = out + (size_t)n * StRef<T>::REC;
    for (int q = 0; q < 8; q++) {
      const T hi = T(ph[q]);
      o[q] = hi;
      o[12 + q] = Lim<T>::fp32 ? T(ph[q] - (double)hi) : T(0);
    }
    for (int q = 0; q < 4; q++) o[8 + q] = F1[q];
    o[20] = bs;
    o[21] = o[22] = o[23] = R(0);
  }
}

// Per leaf: the largest rho if every cell is field-free (B^i = 0), has rho >= 0 and finite
// primitives, else -1.  One workgroup per leaf.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_leafmeta(const T* __restrict__ prim, int cells, T* __restrict__ meta) {
  __shared__ T smax[BLOCK];
  __shared__ int sbad[BLOCK];
  const int q = blockIdx.x;
  const T* p = prim + (size_t)q * cells * 8;
  T mx = T(0);
  int bad = 0;
  for (int c = threadIdx.x; c < cells; c += BLOCK) {
    const T* v = p + (size_t)c * 8;
    for (int a = 0; a < 8; a++) bad |= !isfinite(v[a]);
    bad |= v[5] != T(0) || v[6] != T(0) || v[7] != T(0) || v[0] < T(0);
    mx = fmax(mx, v[0]);
  }
  smax[threadIdx.x] = mx;
  sbad[threadIdx.x] = bad;
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
      sbad[threadIdx.x] |= sbad[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) meta[q] = sbad[0] ? T(-1) : smax[0];
}

// cons -> prim of one cell in FP64 (src/core/model/bhac.hpp:178-346): metric at the
// sqrt(gamma)-weighted 27-point Simpson centroid, Valencia inversion, entropy fallback.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_decode(const double* __restrict__ raw, int leaf0, int nl, int cells,
                                                  int nwini, int nx0, int nx1, const T* __restrict__ blk, T a,
                                                  T hs, T gam, int bpol, T* __restrict__ prim) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nl * cells) return;
  const int q = (int)(t / cells), cell = (int)(t % cells), leaf = leaf0 + q;
  const double* cons = raw + (size_t)q * nwini * cells;
  auto CV = [&](int nw) { return cons[(size_t)nw * cells + cell]; };
  const int loc[3] = {cell % nx0, (cell / nx0) % nx1, cell / (nx0 * nx1)};
  double X[3], dx[3];
  for (int d = 0; d < 3; d++) {
    dx[d] = (double)blk[12 * leaf + 6 + d];
    X[d] = (double)blk[12 * leaf + d] + (loc[d] + 0.5) * dx[d];
  }
  T* out = prim + ((size_t)leaf * cells + cell) * 8;
  if (!(exp(X[0]) > 1.0)) {
    for (int v = 0; v < 8; v++) out[v] = T(0);
    return;
  }
  const double ad = (double)a, hd = (double)hs;
  const double wt[3] = {1., 4., 1.}, off[3] = {-1., 0., 1.};
  double nrm = 0., xb = 0., yb = 0.;  // the metric does not depend on x3
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        const double w = wt[i] * wt[j] * wt[k];
        const double x = X[0] + dx[0] / 2. * off[i], y = X[1] + dx[1] / 2. * off[j];
        const Met<double> g = metric<double>(ad, hd, x, y);
        const double dg = sqrt(g.d11 * (g.d22 * g.d33) - g.d13 * (g.d22 * g.d13));
        nrm += dg * w;
        xb += dg * w * x;
        yb += dg * w * y;
      }
  const double vol = (dx[0] / 6.0) * (dx[1] / 6.0) * (dx[2] / 6.0);
  nrm *= vol;
  const double Xb1 = xb * (vol / nrm), Xb2 = yb * (vol / nrm);
  const Met<double> g = metric<double>(ad, hd, Xb1, Xb2);
  const MetU<double> u = metric_up<double>(ad, g);
  const double S1 = CV(1), S2 = CV(2), S3 = CV(3), B1 = CV(5), B2 = CV(6), B3 = CV(7);
  const double BS = S1 * B1 + S2 * B2 + S3 * B3;
  // gamma^ij = g^ij + g^0i g^0j / (-g^00): only the (1,1) entry picks up the shift term
  const double y11 = u.u11 + u.u01 * u.u01 / (-u.u00), y13 = u.u13, y22 = u.u22, y33 = u.u33;
  const double Su1 = y11 * S1 + y13 * S3, Su2 = y22 * S2, Su3 = y13 * S1 + y33 * S3;
  const double Bd1 = g.d11 * B1 + g.d13 * B3, Bd2 = g.d22 * B2, Bd3 = g.d13 * B1 + g.d33 * B3;
  const double Bsq = Bd1 * B1 + Bd2 * B2 + Bd3 * B3;
  const double gm = (double)gam, lf = CV(nwini - 2), xi = CV(nwini - 1);
  const double rho = CV(0) / lf;
  double uu = (gm - 1.) / gm * (xi / (lf * lf) - rho) / (gm - 1.);
  if (uu < 0) {
    uu = (CV(8) / CV(0)) * pow(rho, gm - 1) / (gm - 1.);
    if (uu < 0) uu = 1e-14;
  }
  const double den = xi + Bsq;
  out[0] = T(rho);
  out[1] = T(uu);
  out[2] = T((Su1 / den + B1 * BS / (xi * den)) * lf);
  out[3] = T((Su2 / den + B2 * BS / (xi * den)) * lf);
  out[4] = T((Su3 / den + B3 * BS / (xi * den)) * lf);
  out[5] = T(bpol * B1);
  out[6] = T(bpol * B2);
  out[7] = T(bpol * B3);
}

// ============================================================================ host runtime
static double now_s() {
  using clk = std::chrono::steady_clock;
  return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

template <class T>
struct Tables {
  T* blk = nullptr;   // [nleafs][12] block records
  T* prim = nullptr;  // [nleafs * cells][8] primitives
  T* meta = nullptr;  // [nleafs] field-free leaves' largest rho, else -1 (k_leafmeta)
};

// Launch knobs (picked by the autotuner or from the environment)
struct Tune {
  int queue = -1;         // -1: default (samples in place), 0: samples in place, 1: LDS queue
  int blocks_per_cu = 0;  // 0: the occupancy limit
  int order = 0;          // batch job order: 0 request after request, 1 requests interleaved per wave
  int fp64_wave = 1;      // mixed march: 1 = a lane's FP64 step takes its whole wave to FP64
  int refill = 64;        // idle lanes a wave gathers before they take new jobs (64: the whole wave)
  int escape = 0;         // 0: by precision (MIXED ends escaping rays at rt_outer, FP32 / FP64 march them
                          // to r_out as the reference does), 1: march to r_out, 2: end at rt_outer
};
// grow-only device buffer
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

}  // namespace phx

struct phx_snapshot {
  int tables = 0;
  phx_model geo{};  // grid description (no host pointers)
  phx::Tables<float> t32;
  phx::Tables<double> t64;
  int* map = nullptr;
  int fd[3] = {0, 0, 0};
  int64_t bytes = 0;  // resident tables
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  phx::DevBuf res, jobs, pix, start, tasks, cams, side;  // per-render work buffers (grow only)
  int* d_counter = nullptr;
  unsigned long long* d_tot = nullptr;
  double decode_s = 0., upload_s = 0.;
  int64_t last_nres = 0;
  int last_jobs = 0;
  int pix_w = 0, pix_h = 0;  // image whose Morton pixel list `pix` holds (0: none)
  int device = 0;
  std::vector<void*> allocs;
  phx::Tune tune;  // launch configuration of batch renders (set by the autotuner)
  std::mutex mu;   // one render at a time per snapshot
};

namespace phx {

static int g_device = 0;
static std::thread* g_warm = nullptr;
static std::mutex g_warm_mu;

static void warm_wait() {
  std::lock_guard<std::mutex> lk(g_warm_mu);
  if (g_warm) {
    g_warm->join();
    delete g_warm;
    g_warm = nullptr;
  }
}

template <class P>
static P* dalloc(phx_snapshot* s, size_t bytes) {
  void* d = nullptr;
  HIPCHECK(hipMalloc(&d, bytes ? bytes : 16));
  s->allocs.push_back(d);
  s->bytes += (int64_t)bytes;
  return (P*)d;
}

static void* ensure(DevBuf& b, size_t bytes) {
  bytes = std::max<size_t>(bytes, 256);
  if (bytes > b.cap) {
    if (b.p) HIPCHECK(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
    HIPCHECK(hipMalloc(&b.p, bytes));
    b.cap = bytes;
  }
  return b.p;
}

static void snapshot_free(phx_snapshot* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (void* p : s->allocs) (void)hipFree(p);
  for (DevBuf* b : {&s->res, &s->jobs, &s->pix, &s->start, &s->tasks, &s->cams, &s->side})
    if (b->p) (void)hipFree(b->p);
  if (s->e0) (void)hipEventDestroy(s->e0);
  if (s->e1) (void)hipEventDestroy(s->e1);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

// block records: lower corner, upper corner (lo + nx * dxc in the table's precision, as
// src/core/model/sampler.hpp:93-97 forms it), cell size, reciprocal cell size
template <class T>
static std::vector<T> block_records(const phx_model* m) {
#pragma clang fp contract(off)
  std::vector<T> rec((size_t)m->nleafs * 12, T(0));
  for (int q = 0; q < m->nleafs; q++)
    for (int d = 0; d < 3; d++) {
      const T lo = T(m->lb[3 * q + d]), dx = T(m->dxc[3 * q + d]);
      const T span = T(m->nx[d]) * dx;
      rec[12 * q + d] = lo;
      rec[12 * q + 3 + d] = lo + span;
      rec[12 * q + 6 + d] = dx;
      rec[12 * q + 9 + d] = T(1) / dx;  // the mixed march locates cells with it (FP64 table)
    }
  return rec;
}

// finest-level map whose entries are the first leaf of the coarse-to-fine walk of
// find_leaf_map (src/core/model/sampler.hpp:76-80)
static std::vector<int> flat_map(const phx_model* m, int fd[3]) {
  const int top = m->nlevels - 1;
  for (int d = 0; d < 3; d++) fd[d] = m->level_dims[top][d];
  const size_t n = (size_t)fd[0] * fd[1] * fd[2];
  if (n > ((size_t)1 << 28)) throw Error(PHX_E_RANGE, "flattened AMR map above 2^28 entries");
  std::vector<int> out(n, -1);
  for (int i = 0; i < fd[0]; i++)
    for (int j = 0; j < fd[1]; j++)
      for (int k = 0; k < fd[2]; k++) {
        int leaf = -1;
        for (int L = 0; L <= top && leaf < 0; L++) {
          const int s = top - L;
          leaf = m->level_map[L][((size_t)(i >> s) * m->level_dims[L][1] + (j >> s)) * m->level_dims[L][2] + (k >> s)];
        }
        out[((size_t)i * fd[1] + j) * fd[2] + k] = leaf;
      }
  return out;
}

template <class T>
static void upload_blocks(phx_snapshot* s, const phx_model* m, Tables<T>& tb) {
  const std::vector<T> rec = block_records<T>(m);
  tb.blk = dalloc<T>(s, rec.size() * sizeof(T));
  HIPCHECK(hipMemcpyAsync(tb.blk, rec.data(), rec.size() * sizeof(T), hipMemcpyHostToDevice, s->stream));
  HIPCHECK(hipStreamSynchronize(s->stream));
}
template <class T>
static void upload_prim(phx_snapshot* s, const phx_model* m, Tables<T>& tb) {
  const size_t nprim = (size_t)m->nleafs * m->cells * 8;
  // the FP32 marches address the FP32 cell table by 32-bit byte offsets
  if (sizeof(T) == 4 && nprim * sizeof(T) >= (size_t(1) << 32))
    throw Error(PHX_E_ARG, "the FP32 cell table of this snapshot exceeds 4 GiB");
  tb.prim = dalloc<T>(s, nprim * sizeof(T));
  if (m->prim) {  // host primitives given: convert and upload; otherwise the dump is decoded on the device
    std::vector<T> p(nprim);
    for (size_t i = 0; i < nprim; i++) p[i] = T(m->prim[i]);
    HIPCHECK(hipMemcpyAsync(tb.prim, p.data(), nprim * sizeof(T), hipMemcpyHostToDevice, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
  }
}

// Streams the dump's leaf blocks through two pinned buffers and decodes them on the device
// (the reference's decode_to_device, src/device/kernel.cu:200-260).
static void decode_dump(phx_snapshot* s, const phx_model* m, const char* dump) {
  const double t0 = now_s();
  const size_t per_leaf = (size_t)m->nwini * m->cells;
  const long stag = (long)(m->nx[0] + 1) * (m->nx[1] + 1) * (m->nx[2] + 1) * m->nws * 8;
  int per = (int)((64u << 20) / (per_leaf * sizeof(double)));
  per = std::max(1, std::min(per, m->nleafs));
  FILE* f = std::fopen(dump, "rb");
  if (!f) throw Error(PHX_E_IO, std::string("cannot open ") + dump);
  struct Closer {
    FILE* f;
    ~Closer() { std::fclose(f); }
  } closer{f};
  double* h_raw[2] = {nullptr, nullptr};
  double* d_raw[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(s->stream);
    for (int b = 0; b < 2; b++) {
      if (h_raw[b]) (void)hipHostFree(h_raw[b]);
      if (d_raw[b]) (void)hipFree(d_raw[b]);
      if (ev[b]) (void)hipEventDestroy(ev[b]);
    }
  };
  try {
    for (int b = 0; b < 2; b++) {
      HIPCHECK(hipHostMalloc((void**)&h_raw[b], (size_t)per * per_leaf * sizeof(double), hipHostMallocDefault));
      HIPCHECK(hipMalloc((void**)&d_raw[b], (size_t)per * per_leaf * sizeof(double)));
      HIPCHECK(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
      HIPCHECK(hipEventRecord(ev[b], s->stream));
    }
    int nb = 0;
    for (int l0 = 0; l0 < m->nleafs; l0 += per, nb ^= 1) {
      const int nl = std::min(per, m->nleafs - l0);
      HIPCHECK(hipEventSynchronize(ev[nb]));  // buffer nb free again
      for (int q = 0; q < nl; q++) {
        if (std::fread(h_raw[nb] + (size_t)q * per_leaf, sizeof(double), per_leaf, f) != per_leaf)
          throw Error(PHX_E_IO, "short read in the data, leaf " + std::to_string(l0 + q));
        std::fseek(f, stag, SEEK_CUR);
      }
      HIPCHECK(hipMemcpyAsync(d_raw[nb], h_raw[nb], (size_t)nl * per_leaf * sizeof(double), hipMemcpyHostToDevice,
                              s->stream));
      const size_t cells = (size_t)nl * m->cells;
      const unsigned grid = (unsigned)((cells + BLOCK - 1) / BLOCK);
      if (s->tables & PHX_TABLE_FP32)
        k_decode<float><<<grid, BLOCK, 0, s->stream>>>(d_raw[nb], l0, nl, m->cells, m->nwini, m->nx[0], m->nx[1],
                                                       s->t32.blk, (float)m->a, (float)m->hslope, (float)m->gam,
                                                       m->bpol, s->t32.prim);
      if (s->tables & PHX_TABLE_FP64)
        k_decode<double><<<grid, BLOCK, 0, s->stream>>>(d_raw[nb], l0, nl, m->cells, m->nwini, m->nx[0], m->nx[1],
                                                        s->t64.blk, m->a, m->hslope, m->gam, m->bpol, s->t64.prim);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipEventRecord(ev[nb], s->stream));
    }
    HIPCHECK(hipStreamSynchronize(s->stream));
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
  s->decode_s = now_s() - t0;
}

static void check_model(const phx_model* m) {
  if (!m) throw Error(PHX_E_ARG, "null model");
  if (m->nleafs <= 0 || m->cells <= 0 || !m->lb || !m->dxc) throw Error(PHX_E_ARG, "model has no block tables");
  if (m->nlevels < 1 || m->nlevels > PHX_MAX_LEVELS) throw Error(PHX_E_ARG, "model nlevels out of range");
  for (int l = 0; l < m->nlevels; l++)
    if (!m->level_map[l]) throw Error(PHX_E_ARG, "model level map missing");
  if (m->nx[0] * m->nx[1] * m->nx[2] != m->cells) throw Error(PHX_E_ARG, "cells != nx0 nx1 nx2");
  if (m->nwini < 9) throw Error(PHX_E_ARG, "nwini < 9");
}

static phx_snapshot* snapshot_open(const phx_model* m, const char* dump, int tables) {
  check_model(m);
  if (!(tables & (PHX_TABLE_FP32 | PHX_TABLE_FP64))) throw Error(PHX_E_ARG, "no table precision requested");
  if (!m->prim && !dump) throw Error(PHX_E_ARG, "no primitives and no dump file");
  warm_wait();
  HIPCHECK(hipSetDevice(g_device));
  std::unique_ptr<phx_snapshot, void (*)(phx_snapshot*)> s(new phx_snapshot(), snapshot_free);
  const double t0 = now_s();
  s->device = g_device;
  s->tables = tables;
  s->geo = *m;
  s->geo.lb = s->geo.dxc = s->geo.prim = nullptr;
  s->geo.leaf_level = s->geo.leaf_index = nullptr;
  for (int l = 0; l < PHX_MAX_LEVELS; l++) s->geo.level_map[l] = nullptr;
  HIPCHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  HIPCHECK(hipEventCreate(&s->e0));
  HIPCHECK(hipEventCreate(&s->e1));
  s->d_counter = dalloc<int>(s.get(), 64 * sizeof(int));
  s->d_tot = dalloc<unsigned long long>(s.get(), 4 * sizeof(unsigned long long));
  const std::vector<int> fm = flat_map(m, s->fd);
  s->map = dalloc<int>(s.get(), fm.size() * sizeof(int));
  HIPCHECK(hipMemcpyAsync(s->map, fm.data(), fm.size() * sizeof(int), hipMemcpyHostToDevice, s->stream));
  // both block tables are always resident (a few KB): the mixed march locates cells in FP64
  upload_blocks<float>(s.get(), m, s->t32);
  upload_blocks<double>(s.get(), m, s->t64);
  if (tables & PHX_TABLE_FP32) upload_prim<float>(s.get(), m, s->t32);
  if (tables & PHX_TABLE_FP64) upload_prim<double>(s.get(), m, s->t64);
  HIPCHECK(hipStreamSynchronize(s->stream));
  s->upload_s = now_s() - t0;
  if (!m->prim) decode_dump(s.get(), m, dump);
  // the field-free leaves of each resident cell table (PHX_FIELD_FREE_LEAVES=0: none is marked, every
  // sample takes the full derivation)
  const char* ffe = std::getenv("PHX_FIELD_FREE_LEAVES");
  const bool mark = !(ffe && std::atoi(ffe) == 0);
  if (tables & PHX_TABLE_FP32) {
    s->t32.meta = dalloc<float>(s.get(), (size_t)m->nleafs * sizeof(float));
    if (mark) {
      k_leafmeta<float><<<m->nleafs, BLOCK, 0, s->stream>>>(s->t32.prim, m->cells, s->t32.meta);
    } else {
      const std::vector<float> none(m->nleafs, -1.f);
      HIPCHECK(hipMemcpyAsync(s->t32.meta, none.data(), none.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
      HIPCHECK(hipStreamSynchronize(s->stream));
    }
  }
  if (tables & PHX_TABLE_FP64) {
    s->t64.meta = dalloc<double>(s.get(), (size_t)m->nleafs * sizeof(double));
    if (mark) {
      k_leafmeta<double><<<m->nleafs, BLOCK, 0, s->stream>>>(s->t64.prim, m->cells, s->t64.meta);
    } else {
      const std::vector<double> none(m->nleafs, -1.);
      HIPCHECK(hipMemcpyAsync(s->t64.meta, none.data(), none.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
      HIPCHECK(hipStreamSynchronize(s->stream));
    }
  }
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(s->stream));
  return s.release();
}

template <class T>
static Grid<T> grid_of(const phx_snapshot* s, const Tables<T>& tb) {
  const phx_model& m = s->geo;
  Grid<T> G;
  G.blk = tb.blk;
  G.meta = tb.meta;
  G.prim = tb.prim;
  G.map = s->map;
  G.nleafs = m.nleafs;
  G.cells = m.cells;
  G.nx0 = m.nx[0];
  G.nx1 = m.nx[1];
  G.nx2 = m.nx[2];
  G.fd0 = s->fd[0];
  G.fd1 = s->fd[1];
  G.fd2 = s->fd[2];
  G.w0 = T(m.lvl_w[0][0]);
  G.w1 = T(m.lvl_w[0][1]);
  G.w2 = T(m.lvl_w[0][2]);
  G.up = T(1 << (m.nlevels - 1));
  G.lo1 = T(m.x_lo[1]);
  G.lo2 = T(m.x_lo[2]);
  G.lo3 = T(m.x_lo[3]);
  G.hi1 = T(m.x_hi[1]);
  G.hi2 = T(m.x_hi[2]);
  G.hi3 = T(m.x_hi[3]);
  return G;
}
template <class T>
static const Tables<T>& tables_of(const phx_snapshot* s);
template <>
const Tables<float>& tables_of<float>(const phx_snapshot* s) {
  return s->t32;
}
template <>
const Tables<double>& tables_of<double>(const phx_snapshot* s) {
  return s->t64;
}

// x2 at which |cos theta| = 0.99 on the northern side (theta = x2 + h/2 sin 2 x2), by Newton
static double pole_x2_limit(double hs) {
  const double th = std::acos(0.99);
  double x = th;
  for (int it = 0; it < 60; it++) {
    const double f = x + 0.5 * hs * std::sin(2. * x) - th, df = 1. + hs * std::cos(2. * x);
    const double nx = x - f / df;
    if (std::fabs(nx - x) < 1e-15) {
      x = nx;
      break;
    }
    x = nx;
  }
  return x;
}

static Tune tune_from_env(Tune t) {
  if (const char* e = std::getenv("PHX_MARCH")) {
    if (!std::strcmp(e, "direct")) t.queue = 0;
    if (!std::strcmp(e, "queued")) t.queue = 1;
  }
  if (const char* e = std::getenv("PHX_BLOCKS_PER_CU"))
    if (std::atoi(e) > 0) t.blocks_per_cu = std::atoi(e);
  if (const char* e = std::getenv("PHX_FP64_WAVE")) t.fp64_wave = std::atoi(e) != 0;
  if (const char* e = std::getenv("PHX_ESCAPE"))
    t.escape = !std::strcmp(e, "full") ? 1 : !std::strcmp(e, "exit") ? 2 : t.escape;
  if (const char* e = std::getenv("PHX_REFILL"))
    t.refill = !std::strcmp(e, "wave") ? 64 : !std::strcmp(e, "lane") ? 1 : std::atoi(e);
  return t;
}

// what one march launch needs beyond the device buffers of the snapshot
struct LaunchInfo {
  int njobs = 0, ntasks = 0, nfmax = 1;
  const phx_model* uni = nullptr;  // the request whose launch-wide parameters (stepping, cuts) apply
};

template <int M, int NF, bool Q>
static void launch_march(phx_snapshot* s, const LaunchInfo& li, int j0, int j1, const Tune& tn, phx_info* info) {
  using T = typename ModeOf<M>::T;
  MarchArgs<M> A;
  std::memset(&A, 0, sizeof A);
  A.G = grid_of<T>(s, tables_of<T>(s));
  SideArgs<M> side;
  std::memset(&side, 0, sizeof side);
  side.G = A.G;
  if constexpr (M == MIX) {
    const Grid<double> g64 = grid_of<double>(s, tables_of<double>(s));
    side.GL = Loc<double>{g64.blk, g64.lo1, g64.lo2, g64.lo3, g64.hi1, g64.hi2, g64.hi3};
  }
  const phx_model& m = s->geo;
  const phx_model* rq = li.uni;
  A.P.a = T(m.a);
  A.P.hs = T(m.hslope);
  A.P.gam = T(m.gam);
  A.P.rlow = T(rq->rlow);
  A.P.ibc = T(rq->ibc);
  A.P.sigma_cut = T(rq->sigma_cut);
  A.P.te_min = T(rq->thetae_min);
  A.P.te_max = T(rq->thetae_max);
  A.P.rt_outer = T(rq->rt_outer);
  side.P = A.P;
  A.start = s->start.p;
  A.stepsize = T(rq->stepsize);
  A.r_out = T(rq->r_out);
  A.cut_in = T(rq->cut_in);
  A.a64 = m.a;
  A.hs64 = m.hslope;
  A.lnr_fp64 = std::log(2. * (1. + std::sqrt(std::max(0., 1. - m.a * m.a))));  // r < 2 r_H
  A.x2_pole = pole_x2_limit(m.hslope);                                          // |cos theta| > 0.99
  A.fp64_wave = tn.fp64_wave;
  A.refill = std::min(64, std::max(1, tn.refill));
  A.escape = tn.escape == 2 || (tn.escape == 0 && M == MIX);
  A.max_steps = rq->max_steps;
  A.max_order = rq->max_order;
  A.tasks = (const TaskDev<T>*)s->tasks.p;
  A.jobs = (const int2*)s->jobs.p;
  A.j0 = j0;
  A.j1 = j1;
  A.counter = s->d_counter;
  A.tot = s->d_tot;
  side.tasks = A.tasks;
  side.jobs = A.jobs;
  side.start = A.start;
  side.counter = A.counter;
  side.tot = A.tot;
  side.j0 = A.j0;
  side.j1 = A.j1;
  // from pageable host memory: the call returns once the source has been staged
  A.side = static_cast<const SideArgs<M>*>(ensure(s->side, sizeof side));
  HIPCHECK(hipMemcpyAsync(const_cast<SideArgs<M>*>(A.side), &side, sizeof side, hipMemcpyHostToDevice, s->stream));
  const void* kfn = reinterpret_cast<const void*>(&k_march<M, NF, Q>);
  int cus = 0, per_cu = 0;
  HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device));
  HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, BLOCK, 0));
  per_cu = std::max(per_cu, 1);
  if (tn.blocks_per_cu > 0) per_cu = std::min(per_cu, tn.blocks_per_cu);
  int grid = per_cu * cus;
  const int njobs = j1 - j0;
  grid = std::max(1, std::min(grid, (njobs + BLOCK - 1) / BLOCK));
  hipFuncAttributes fa;
  HIPCHECK(hipFuncGetAttributes(&fa, kfn));
  // a kernel whose private segment exceeds the per-lane stack limit does not run at all; raise the
  // limit to what the kernel needs (the reference sets it explicitly too, kernel.cu:305)
  size_t limit = 0;
  HIPCHECK(hipDeviceGetLimit(&limit, hipLimitStackSize));
  if (fa.localSizeBytes > limit) HIPCHECK(hipDeviceSetLimit(hipLimitStackSize, (fa.localSizeBytes + 255) & ~size_t(255)));
  HIPCHECK(hipMemsetAsync(s->d_counter, 0, sizeof(int), s->stream));
  k_march<M, NF, Q><<<grid, BLOCK, 0, s->stream>>>(A);
  HIPCHECK(hipGetLastError());
  s->last_jobs = njobs;
  if (info) {
    info->vgprs = fa.numRegs;
    info->scratch_bytes = (int32_t)fa.localSizeBytes;
    info->lds_bytes = (int32_t)fa.sharedSizeBytes;
    info->cus = cus;
    info->blocks = grid;
    info->block = BLOCK;
    info->waves_per_cu = per_cu * BLOCK / 64;
  }
}

template <int M, int NF>
static void launch_nf(phx_snapshot* s, const LaunchInfo& li, int j0, int j1, const Tune& tn, phx_info* info) {
  // default: samples in place with whole-wave refill (measured faster in every precision: rays of
  // one 8 x 8 pixel tile stay coherent; the LDS queue pays off only with per-lane refill)
  const bool q = tn.queue == 1;
  if (q)
    launch_march<M, NF, true>(s, li, j0, j1, tn, info);
  else
    launch_march<M, NF, false>(s, li, j0, j1, tn, info);
}

// Ray start records for the jobs in s->jobs (cameras in s->cams), then the march over them,
// in chunks of PHX_CHUNK jobs when that is set (the reference's chunking, kernel.cu:177-180).
template <int M>
static void march_jobs(phx_snapshot* s, const LaunchInfo& li, const Tune& tn, phx_info* info,
                       phx_progress_fn progress, void* user) {
  using R = RecOf<M>;
  const int rec = M == MIX ? StHam::REC : StRef<typename ModeOf<M>::T>::REC;
  ensure(s->start, (size_t)std::max(li.njobs, 1) * rec * sizeof(R));
  if (li.njobs > 0)
    k_rays<M><<<(li.njobs + BLOCK - 1) / BLOCK, BLOCK, 0, s->stream>>>((const Cam*)s->cams.p,
                                                                      (const int2*)s->jobs.p, li.njobs,
                                                                      s->start.p);
  HIPCHECK(hipGetLastError());
  int chunk = std::max(li.njobs, 1);
  if (const char* e = std::getenv("PHX_CHUNK"))
    if (std::atoi(e) > 0) chunk = std::atoi(e);
  for (int j0 = 0; j0 < li.njobs; j0 += chunk) {
    const int j1 = std::min(li.njobs, j0 + chunk);
    if (li.nfmax == 1)
      launch_nf<M, 1>(s, li, j0, j1, tn, info);
    else if (li.nfmax == 2)
      launch_nf<M, 2>(s, li, j0, j1, tn, info);
    else
      launch_nf<M, 4>(s, li, j0, j1, tn, info);
    if (progress) {
      HIPCHECK(hipStreamSynchronize(s->stream));
      progress(j1 - j0, user);
    }
  }
}

static void check_precision(const phx_snapshot* s, int precision) {
  if (precision != PHX_FP32 && precision != PHX_FP64 && precision != PHX_MIXED)
    throw Error(PHX_E_ARG, "precision must be PHX_FP32, PHX_FP64 or PHX_MIXED");
  if (precision != PHX_FP64 && !(s->tables & PHX_TABLE_FP32)) throw Error(PHX_E_ARG, "snapshot has no FP32 table");
  if (precision == PHX_FP64 && !(s->tables & PHX_TABLE_FP64)) throw Error(PHX_E_ARG, "snapshot has no FP64 table");
}

static void check_request(const phx_model* rq) {
  if (!rq) throw Error(PHX_E_ARG, "null request");
  if (rq->img_width <= 0 || rq->img_height <= 0) throw Error(PHX_E_ARG, "image size must be positive");
  if ((int64_t)rq->img_width * rq->img_height > (int64_t)1 << 30) throw Error(PHX_E_RANGE, "image too large");
  if (rq->nfreqs < 1 || rq->nfreqs > PHX_MAX_FREQS) throw Error(PHX_E_ARG, "nfreqs must be in 1..32");
}

// the parameters one launch shares: a batch may differ only in camera, image, frequencies,
// m_unit and R_high
static void check_uniform(const phx_model* a, const phx_model* b) {
  const bool same = a->stepsize == b->stepsize && a->r_out == b->r_out && a->cut_in == b->cut_in &&
                    a->max_steps == b->max_steps && a->max_order == b->max_order && a->rlow == b->rlow &&
                    a->ibc == b->ibc && a->sigma_cut == b->sigma_cut && a->thetae_min == b->thetae_min &&
                    a->thetae_max == b->thetae_max && a->rt_outer == b->rt_outer;
  if (!same)
    throw Error(PHX_E_ARG, "requests of one batch must share stepsize, r_cam, horizon margin, max_steps, max_order, "
                           "r_low, beta_c, sigma_cut, thetae bounds and rt_outer");
}

static Cam cam_of(const phx_snapshot* s, const phx_model* rq) {
  Cam C;
  C.a = s->geo.a;
  C.hs = s->geo.hslope;
  C.incl = rq->inclination;
  C.r_cam = rq->r_cam;
  C.t_init = rq->t_init;
  C.phi_cam = rq->phi_cam;
  C.fov_x = rq->fov_x;
  C.fov_y = rq->fov_y;
  C.cos_pa = rq->cos_pa;
  C.sin_pa = rq->sin_pa;
  C.W = rq->img_width;
  C.H = rq->img_height;
  return C;
}

// Tasks of one request: chunks of CHUNK frequencies; `out` is the request's result block.
template <class T>
static int add_tasks(const phx_model* rq, phx_freqout* out, std::vector<TaskDev<T>>& tasks) {
  const int first = (int)tasks.size();
  for (int f0 = 0; f0 < rq->nfreqs; f0 += CHUNK) {
    TaskDev<T> t;
    std::memset(&t, 0, sizeof t);
    t.Ne = T(rq->Ne);
    t.Bu = T(rq->Bu);
    t.rhigh = T(rq->rhigh);
    t.nf = std::min(CHUNK, rq->nfreqs - f0);
    for (int f = 0; f < t.nf; f++) {
      t.freq[f] = T(rq->freq[f0 + f]);
      t.hconst[f] = T(rq->hconst[f0 + f]);
      t.jscale[f] = T(rq->jscale[f0 + f]);
    }
    t.out = out + f0;
    t.nf_total = rq->nfreqs;
    t.stats = f0 == 0;
    tasks.push_back(t);
  }
  return (int)tasks.size() - first;
}

__global__ void k_expand(const int* __restrict__ pix, int npix, int task0, int ntasks, int2* __restrict__ jobs) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= npix * ntasks) return;
  const int t = n / npix, i = n % npix;
  jobs[n] = make_int2(task0 + t, pix[i]);
}

// pixel order of a full image: Morton order of (i, j) (the reference's GPU schedule,
// src/host/render/jobs.hpp:18-46); a square power-of-two image is enumerated directly
static void morton_pixels(const phx_model* rq, std::vector<int32_t>& out) {
  const int W = rq->img_width, H = rq->img_height;
  const int64_t n = (int64_t)W * H;
  out.resize(n);
  if (W == H && (W & (W - 1)) == 0 && W <= 65536) {
    for (int64_t k = 0; k < n; k++) {
      uint32_t i = 0, j = 0;
      for (int b = 0; b < 16; b++) {  // i in the even bits, j in the odd bits (jobs.hpp:33-35)
        i |= (uint32_t)((k >> (2 * b)) & 1) << b;
        j |= (uint32_t)((k >> (2 * b + 1)) & 1) << b;
      }
      out[k] = (int32_t)((int64_t)i * H + j);
    }
    return;
  }
  if (phx_schedule(rq, 1, out.data(), (int32_t)n) != PHX_OK) throw Error(PHX_E_ARG, phx_last_error());
}

static void fill_info_device(const phx_snapshot* s, phx_info* info) {
  hipDeviceProp_t prop;
  HIPCHECK(hipGetDeviceProperties(&prop, s->device));
  std::snprintf(info->device, sizeof info->device, "%s (%s)", prop.name, prop.gcnArchName);
}

// the march of the jobs already in s->jobs; fills timing and statistics
template <int M>
static void run_march(phx_snapshot* s, const LaunchInfo& li, const Tune& tn, phx_info* info, phx_progress_fn progress,
                      void* user) {
  HIPCHECK(hipMemsetAsync(s->d_tot, 0, 4 * sizeof(unsigned long long), s->stream));
  HIPCHECK(hipEventRecord(s->e0, s->stream));
  march_jobs<M>(s, li, tn, info, progress, user);
  HIPCHECK(hipEventRecord(s->e1, s->stream));
  HIPCHECK(hipEventSynchronize(s->e1));
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, s->e0, s->e1));
  unsigned long long tot[4] = {0, 0, 0, 0};
  HIPCHECK(hipMemcpy(tot, s->d_tot, sizeof tot, hipMemcpyDeviceToHost));
  int taken = 0;
  HIPCHECK(hipMemcpy(&taken, s->d_counter, sizeof taken, hipMemcpyDeviceToHost));
  if (li.njobs > 0 && taken < s->last_jobs)  // every lane takes jobs until the list is exhausted
    throw Error(PHX_E_DEVICE, "march kernel did not run to completion (" + std::to_string(taken) + " of " +
                                  std::to_string(s->last_jobs) + " jobs taken)");
  info->kernel_s = ms * 1e-3;
  info->steps = (int64_t)tot[0];
  info->nplasma = (int64_t)tot[1];
  info->bad = (int64_t)tot[2];
  info->fp64_steps = (int64_t)tot[3];
  info->decode_s = s->decode_s;
  info->upload_s = s->upload_s;
}

static void run_precision(phx_snapshot* s, int precision, const LaunchInfo& li, const Tune& tn, phx_info* info,
                          phx_progress_fn progress, void* user) {
  if (precision == PHX_FP32)
    run_march<M32>(s, li, tn, info, progress, user);
  else if (precision == PHX_FP64)
    run_march<M64>(s, li, tn, info, progress, user);
  else
    run_march<MIX>(s, li, tn, info, progress, user);
}

template <class T>
static void upload_tasks(phx_snapshot* s, const std::vector<TaskDev<T>>& tasks, const std::vector<Cam>& cams) {
  ensure(s->tasks, tasks.size() * sizeof(TaskDev<T>));
  ensure(s->cams, cams.size() * sizeof(Cam));
  HIPCHECK(hipMemcpyAsync(s->tasks.p, tasks.data(), tasks.size() * sizeof(TaskDev<T>), hipMemcpyHostToDevice,
                          s->stream));
  HIPCHECK(hipMemcpyAsync(s->cams.p, cams.data(), cams.size() * sizeof(Cam), hipMemcpyHostToDevice, s->stream));
}

// One request: its frequency chunks become tasks over the same pixel list (the full image in
// Morton order, cached on the device per image size, or the caller's jobs).
static void render(phx_snapshot* s, const phx_model* rq, int precision, const int32_t* jobs, int32_t njobs,
                   phx_freqout* res, phx_info* info, phx_progress_fn progress, void* user, const Tune& tn0) {
  check_precision(s, precision);
  check_request(rq);
  std::lock_guard<std::mutex> lk(s->mu);
  HIPCHECK(hipSetDevice(s->device));
  const Tune tn = tune_from_env(tn0);
  const int64_t npix = (int64_t)rq->img_width * rq->img_height;
  if (jobs) {
    if (njobs < 0 || njobs > npix) throw Error(PHX_E_ARG, "njobs out of range");
    for (int32_t n = 0; n < njobs; n++)
      if (jobs[n] < 0 || jobs[n] >= npix) throw Error(PHX_E_ARG, "job index outside the image");
  } else {
    njobs = (int32_t)npix;
  }
  const size_t nres = (size_t)npix * rq->nfreqs;
  phx_freqout* d_res = (phx_freqout*)ensure(s->res, nres * sizeof(phx_freqout));
  // pixel list
  if (jobs) {
    ensure(s->pix, (size_t)njobs * sizeof(int32_t));
    HIPCHECK(hipMemcpyAsync(s->pix.p, jobs, (size_t)njobs * sizeof(int32_t), hipMemcpyHostToDevice, s->stream));
    s->pix_w = s->pix_h = 0;
  } else if (!(s->pix_w == rq->img_width && s->pix_h == rq->img_height)) {
    std::vector<int32_t> order;
    morton_pixels(rq, order);
    ensure(s->pix, (size_t)npix * sizeof(int32_t));
    HIPCHECK(hipMemcpyAsync(s->pix.p, order.data(), (size_t)npix * sizeof(int32_t), hipMemcpyHostToDevice, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    s->pix_w = rq->img_width;
    s->pix_h = rq->img_height;
  }
  // tasks, cameras, jobs
  const Cam cam = cam_of(s, rq);
  int ntasks = 0;
  if (precision == PHX_FP64) {
    std::vector<TaskDev<double>> tk;
    ntasks = add_tasks<double>(rq, d_res, tk);
    upload_tasks(s, tk, std::vector<Cam>(ntasks, cam));
  } else {
    std::vector<TaskDev<float>> tk;
    ntasks = add_tasks<float>(rq, d_res, tk);
    upload_tasks(s, tk, std::vector<Cam>(ntasks, cam));
  }
  const int64_t ntot = (int64_t)njobs * ntasks;
  if (ntot > INT32_MAX) throw Error(PHX_E_RANGE, "too many jobs in one render");
  ensure(s->jobs, (size_t)std::max<int64_t>(ntot, 1) * sizeof(int2));
  if (ntot > 0)
    k_expand<<<(unsigned)((ntot + BLOCK - 1) / BLOCK), BLOCK, 0, s->stream>>>((const int*)s->pix.p, njobs, 0, ntasks,
                                                                              (int2*)s->jobs.p);
  HIPCHECK(hipGetLastError());
  if ((int64_t)njobs < npix) HIPCHECK(hipMemsetAsync(d_res, 0, nres * sizeof(phx_freqout), s->stream));
  phx_info local;
  if (!info) info = &local;
  std::memset(info, 0, sizeof *info);
  fill_info_device(s, info);
  LaunchInfo li;
  li.njobs = (int)ntot;
  li.ntasks = ntasks;
  li.nfmax = std::min(CHUNK, rq->nfreqs) <= 1 ? 1 : std::min(CHUNK, rq->nfreqs) <= 2 ? 2 : 4;
  li.uni = rq;
  run_precision(s, precision, li, tn, info, progress, user);
  info->rays = (int64_t)njobs * rq->nfreqs;
  s->last_nres = (int64_t)nres;
  if (res) HIPCHECK(hipMemcpy(res, d_res, nres * sizeof(phx_freqout), hipMemcpyDeviceToHost));
}


// Many requests in one launch: every request's frequency chunks become tasks, and the job list
// holds all their pixels (each request's in Morton order, requests one after another, or
// interleaved 64 jobs at a time).  Results go to one device block per request.
struct BatchPlan {
  std::vector<int64_t> off;  // result offset of each request in s->res (entries)
  int64_t nres = 0, rays = 0;
  int njobs = 0, ntasks = 0, nfmax = 1;
};
template <class T>
static BatchPlan plan_batch(phx_snapshot* s, const phx_model* const* rqs, int n, int order, int sample_stride) {
  BatchPlan bp;
  bp.off.resize(n);
  for (int r = 0; r < n; r++) {
    check_request(rqs[r]);
    check_uniform(rqs[0], rqs[r]);
    bp.off[r] = bp.nres;
    bp.nres += (int64_t)rqs[r]->img_width * rqs[r]->img_height * rqs[r]->nfreqs;
  }
  phx_freqout* d_res = (phx_freqout*)ensure(s->res, (size_t)bp.nres * sizeof(phx_freqout));
  std::vector<TaskDev<T>> tasks;
  std::vector<Cam> cams;
  std::vector<std::vector<int2>> per(n);
  std::vector<int32_t> pix;
  for (int r = 0; r < n; r++) {
    const int t0 = (int)tasks.size();
    const int nt = add_tasks<T>(rqs[r], d_res + bp.off[r], tasks);
    const Cam c = cam_of(s, rqs[r]);
    for (int t = 0; t < nt; t++) cams.push_back(c);
    bp.nfmax = std::max(bp.nfmax, std::min(CHUNK, rqs[r]->nfreqs));
    morton_pixels(rqs[r], pix);
    std::vector<int2>& jl = per[r];
    for (int t = 0; t < nt; t++)
      for (size_t i = 0; i < pix.size(); i += 64 * (size_t)sample_stride)
        for (size_t k = i; k < std::min(pix.size(), i + 64); k++) jl.push_back(make_int2(t0 + t, pix[k]));
    bp.rays += (int64_t)(jl.size() / nt) * rqs[r]->nfreqs;
  }
  std::vector<int2> jobs;
  if (order == 1) {  // round-robin over requests, 64 jobs (one wave) at a time
    std::vector<size_t> pos(n, 0);
    for (bool any = true; any;) {
      any = false;
      for (int r = 0; r < n; r++)
        if (pos[r] < per[r].size()) {
          const size_t e = std::min(per[r].size(), pos[r] + 64);
          jobs.insert(jobs.end(), per[r].begin() + pos[r], per[r].begin() + e);
          pos[r] = e;
          any = true;
        }
    }
  } else {
    for (int r = 0; r < n; r++) jobs.insert(jobs.end(), per[r].begin(), per[r].end());
  }
  if (jobs.size() > (size_t)INT32_MAX) throw Error(PHX_E_RANGE, "too many jobs in one batch");
  bp.njobs = (int)jobs.size();
  bp.ntasks = (int)tasks.size();
  bp.nfmax = bp.nfmax <= 1 ? 1 : bp.nfmax <= 2 ? 2 : 4;
  upload_tasks(s, tasks, cams);
  ensure(s->jobs, jobs.size() * sizeof(int2));
  HIPCHECK(hipMemcpyAsync(s->jobs.p, jobs.data(), jobs.size() * sizeof(int2), hipMemcpyHostToDevice, s->stream));
  if (sample_stride > 1) HIPCHECK(hipMemsetAsync(d_res, 0, (size_t)bp.nres * sizeof(phx_freqout), s->stream));
  HIPCHECK(hipStreamSynchronize(s->stream));
  return bp;
}

static BatchPlan run_batch(phx_snapshot* s, const phx_model* const* rqs, int n, int precision, const Tune& tn,
                           int sample_stride, phx_info* info) {
  check_precision(s, precision);
  if (n < 1) throw Error(PHX_E_ARG, "empty batch");
  HIPCHECK(hipSetDevice(s->device));
  const BatchPlan bp = precision == PHX_FP64 ? plan_batch<double>(s, rqs, n, tn.order, sample_stride)
                                             : plan_batch<float>(s, rqs, n, tn.order, sample_stride);
  std::memset(info, 0, sizeof *info);
  fill_info_device(s, info);
  LaunchInfo li;
  li.njobs = bp.njobs;
  li.ntasks = bp.ntasks;
  li.nfmax = bp.nfmax;
  li.uni = rqs[0];
  run_precision(s, precision, li, tn, info, nullptr, nullptr);
  info->rays = bp.rays;
  return bp;
}

static void render_batch(phx_snapshot* s, const phx_model* const* rqs, int n, int precision, phx_freqout* const* res,
                         phx_info* info) {
  std::lock_guard<std::mutex> lk(s->mu);
  phx_info local;
  if (!info) info = &local;
  const BatchPlan bp = run_batch(s, rqs, n, precision, tune_from_env(s->tune), 1, info);
  s->last_nres = bp.nres;
  for (int r = 0; r < n; r++)
    if (res && res[r])
      HIPCHECK(hipMemcpy(res[r], (phx_freqout*)s->res.p + bp.off[r],
                         (size_t)rqs[r]->img_width * rqs[r]->img_height * rqs[r]->nfreqs * sizeof(phx_freqout),
                         hipMemcpyDeviceToHost));
}

// Autotuner: times every launch configuration (LDS queue or in place, occupancy, job order) on
// a strided sample of the batch (every sample_stride-th wave of 64 jobs), keeps the fastest in
// the snapshot and returns it with its sampled rate.
static void autotune(phx_snapshot* s, const phx_model* const* rqs, int n, int precision, int sample_stride,
                     phx_tune* best, double* best_rate) {
  std::lock_guard<std::mutex> lk(s->mu);
  sample_stride = std::max(1, sample_stride);
  int max_bpc = 0;
  {
    int cus = 0;
    HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device));
    (void)cus;
    max_bpc = precision == PHX_FP64 ? 1 : 2;
  }
  double rate_best = -1.;
  Tune tb;
  for (int q = 0; q <= 1; q++)
    for (int rf = 0; rf <= 1; rf++)
    for (int bpc = 1; bpc <= max_bpc; bpc++)
      for (int order = 0; order <= (n > 1 ? 1 : 0); order++)
        for (int fw = 0; fw <= (precision == PHX_MIXED ? 1 : 0); fw++) {
        Tune t;
        t.queue = q;
        t.refill = rf ? 1 : 64;
        t.blocks_per_cu = bpc;
        t.order = order;
        t.fp64_wave = fw;
        phx_info inf;
        const BatchPlan bp = run_batch(s, rqs, n, precision, t, sample_stride, &inf);
        const double rate = (double)bp.rays / std::max(inf.kernel_s, 1e-9);
        if (rate > rate_best) {
          rate_best = rate;
          tb = t;
        }
      }
  s->tune = tb;
  if (best) {
    best->queue = tb.queue;
    best->blocks_per_cu = tb.blocks_per_cu;
    best->order = tb.order;
    best->fp64_wave = tb.fp64_wave;
    best->refill = tb.refill;
    best->escape = tb.escape;
  }
  if (best_rate) *best_rate = rate_best;
}

// ============================================================================ serving loop
// A worker thread keeps the snapshots resident and serves render requests with continuous
// batching: a request is admitted when its result buffer fits the HBM budget, its jobs (pixels
// x frequency chunks) join a FIFO of active work, and every launch takes the next `quantum`
// jobs across all active requests of a snapshot, so requests that arrive while others are on
// the GPU share the following launches with them.  Snapshots with pending work are launched on
// their own streams and overlap on the device.  A request is done when its last job has been
// marched; its result is copied to the caller's buffer and its latency (submit -> copied)
// recorded.  No counterpart in the reference (one image per process run, main.cpp:143-230).
struct Request {
  int64_t ticket = 0;
  int snap = 0;
  phx_model rq{};
  phx_freqout* host_res = nullptr;
  phx_freqout* d_res = nullptr;
  int64_t bytes = 0;
  std::vector<int2> jobs;  // (chunk, pixel), chunk-major
  size_t next = 0, done = 0;
  double t_submit = 0., t_done = 0.;
  int state = 0;  // 0 waiting, 1 active, 2 done, 3 failed
  std::string error;
};

struct Server {
  std::vector<phx_snapshot*> snaps;
  phx_server_opts opts{};
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  int64_t next_ticket = 1;
  std::deque<std::shared_ptr<Request>> waiting;
  std::vector<std::shared_ptr<Request>> active;
  std::unordered_map<int64_t, std::shared_ptr<Request>> all;
  std::vector<double> latencies;
  std::vector<phx_model> uni;  // per snapshot: the launch-wide parameters of its first request
  std::vector<char> have_uni;
  int64_t submitted = 0, completed = 0, failed = 0, rays = 0, launches = 0, admitted_bytes = 0, peak_bytes = 0;
  double busy_s = 0., t_start = 0.;
  std::string fatal;
};

template <class T>
static void serve_tick(Server* sv, std::vector<std::shared_ptr<Request>>& act) {
  const int M = sv->opts.precision == PHX_FP32 ? M32 : sv->opts.precision == PHX_FP64 ? M64 : MIX;
  const int nsn = (int)sv->snaps.size();
  struct Slice {
    Request* r;
    size_t a, b;
  };
  std::vector<std::vector<Slice>> slices(nsn);
  std::vector<int> njobs(nsn, 0);
  const size_t quantum = (size_t)std::max(1, sv->opts.quantum_jobs);
  for (auto& r : act) {  // FIFO: the oldest request's remaining jobs first
    const int k = r->snap;
    const size_t left = r->jobs.size() - r->next;
    const size_t room = quantum - (size_t)njobs[k];
    if (!left || !room) continue;
    const size_t take = std::min(left, room);
    slices[k].push_back({r.get(), r->next, r->next + take});
    njobs[k] += (int)take;
    r->next += take;
  }
  std::vector<int> launched;
  std::vector<std::unique_lock<std::mutex>> held;  // the snapshots' own renders wait for this tick
  for (int k = 0; k < nsn; k++) {
    if (!njobs[k]) continue;
    phx_snapshot* s = sv->snaps[k];
    held.emplace_back(s->mu);
    std::vector<TaskDev<T>> tasks;
    std::vector<Cam> cams;
    std::vector<int2> jobs;
    jobs.reserve(njobs[k]);
    int nfmax = 1;
    for (const Slice& sl : slices[k]) {
      const int t0 = (int)tasks.size();
      const int nt = add_tasks<T>(&sl.r->rq, sl.r->d_res, tasks);
      const Cam c = cam_of(s, &sl.r->rq);
      for (int t = 0; t < nt; t++) cams.push_back(c);
      nfmax = std::max(nfmax, std::min(CHUNK, sl.r->rq.nfreqs));
      for (size_t j = sl.a; j < sl.b; j++) jobs.push_back(make_int2(t0 + sl.r->jobs[j].x, sl.r->jobs[j].y));
    }
    HIPCHECK(hipSetDevice(s->device));
    upload_tasks(s, tasks, cams);
    ensure(s->jobs, jobs.size() * sizeof(int2));
    HIPCHECK(hipMemcpyAsync(s->jobs.p, jobs.data(), jobs.size() * sizeof(int2), hipMemcpyHostToDevice, s->stream));
    LaunchInfo li;
    li.njobs = njobs[k];
    li.ntasks = (int)tasks.size();
    li.nfmax = nfmax <= 1 ? 1 : nfmax <= 2 ? 2 : 4;
    li.uni = &slices[k].front().r->rq;
    HIPCHECK(hipMemsetAsync(s->d_tot, 0, 4 * sizeof(unsigned long long), s->stream));
    HIPCHECK(hipEventRecord(s->e0, s->stream));
    const Tune tn = tune_from_env(s->tune);
    if (M == M32)
      march_jobs<M32>(s, li, tn, nullptr, nullptr, nullptr);
    else if (M == M64)
      march_jobs<M64>(s, li, tn, nullptr, nullptr, nullptr);
    else
      march_jobs<MIX>(s, li, tn, nullptr, nullptr, nullptr);
    HIPCHECK(hipEventRecord(s->e1, s->stream));
    launched.push_back(k);
  }
  double busy = 0.;
  for (int k : launched) {
    phx_snapshot* s = sv->snaps[k];
    HIPCHECK(hipEventSynchronize(s->e1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, s->e0, s->e1));
    busy = std::max(busy, ms * 1e-3);
    int taken = 0;
    HIPCHECK(hipMemcpy(&taken, s->d_counter, sizeof taken, hipMemcpyDeviceToHost));
    if (taken < s->last_jobs) throw Error(PHX_E_DEVICE, "serving launch did not run to completion");
    for (const Slice& sl : slices[k]) sl.r->done += sl.b - sl.a;
  }
  held.clear();
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->busy_s += busy;
  sv->launches += (int64_t)launched.size();
}

static void serve_loop(Server* sv) {
  std::vector<std::shared_ptr<Request>> act;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_work.wait(lk, [&] { return sv->stop || !sv->waiting.empty() || !sv->active.empty(); });
      if (sv->stop) break;
      // admission by HBM budget, in arrival order (one request is always admitted when none is active)
      while (!sv->waiting.empty()) {
        auto& r = sv->waiting.front();
        if (!sv->active.empty() && sv->admitted_bytes + r->bytes > sv->opts.hbm_budget) break;
        if (sv->opts.max_active > 0 && (int)sv->active.size() >= sv->opts.max_active) break;
        sv->admitted_bytes += r->bytes;
        sv->peak_bytes = std::max(sv->peak_bytes, sv->admitted_bytes);
        r->state = 1;
        sv->active.push_back(r);
        sv->waiting.pop_front();
      }
      act = sv->active;
    }
    // device buffers and job lists of newly admitted requests (outside the lock)
    for (auto& r : act)
      if (!r->d_res && r->state == 1) {
        try {
          phx_snapshot* s = sv->snaps[r->snap];
          HIPCHECK(hipSetDevice(s->device));
          HIPCHECK(hipMalloc((void**)&r->d_res, (size_t)r->bytes));
          std::vector<int32_t> pix;
          morton_pixels(&r->rq, pix);
          const int nch = (r->rq.nfreqs + CHUNK - 1) / CHUNK;
          r->jobs.reserve(pix.size() * nch);
          for (int c = 0; c < nch; c++)
            for (int32_t p : pix) r->jobs.push_back(make_int2(c, p));
        } catch (const std::exception& e) {
          std::lock_guard<std::mutex> lk(sv->mu);
          r->state = 3;
          r->error = e.what();
        }
      }
    try {
      std::vector<std::shared_ptr<Request>> run;
      for (auto& r : act)
        if (r->state == 1) run.push_back(r);
      if (!run.empty()) {
        if (sv->opts.precision == PHX_FP64)
          serve_tick<double>(sv, run);
        else
          serve_tick<float>(sv, run);
      }
    } catch (const std::exception& e) {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->fatal = e.what();
      for (auto& r : act) {
        r->state = 3;
        r->error = e.what();
      }
    }
    // completed requests: copy out, free, record
    for (auto& r : act) {
      const bool finished = r->state == 1 && r->done == r->jobs.size() && !r->jobs.empty();
      std::string copy_error;
      if (finished) {
        try {
          const size_t n = (size_t)r->rq.img_width * r->rq.img_height * r->rq.nfreqs;
          if (r->host_res) HIPCHECK(hipMemcpy(r->host_res, r->d_res, n * sizeof(phx_freqout), hipMemcpyDeviceToHost));
        } catch (const std::exception& e) {
          copy_error = e.what();
        }
      }
      bool failed_now = false;
      {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (!copy_error.empty()) {
          r->state = 3;
          r->error = copy_error;
        }
        failed_now = r->state == 3;
      }
      if (finished || failed_now) {
        if (r->d_res) (void)hipFree(r->d_res);
        r->d_res = nullptr;
        r->jobs.clear();
        r->jobs.shrink_to_fit();
        std::lock_guard<std::mutex> lk(sv->mu);
        r->t_done = now_s();
        sv->admitted_bytes -= r->bytes;
        if (r->state == 3) {
          sv->failed++;
        } else {
          r->state = 2;
          sv->completed++;
          sv->rays += (int64_t)r->rq.img_width * r->rq.img_height * r->rq.nfreqs;
          sv->latencies.push_back(r->t_done - r->t_submit);
        }
        sv->active.erase(std::remove(sv->active.begin(), sv->active.end(), r), sv->active.end());
      }
    }
    sv->cv_done.notify_all();
  }
}

static double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t i = (size_t)pos;
  const double f = pos - (double)i;
  return i + 1 < v.size() ? v[i] * (1. - f) + v[i + 1] * f : v[i];
}

}  // namespace phx

using namespace phx;

extern "C" {

int phx_set_device(int32_t device) {
  return guarded([&] {
    int n = 0;
    HIPCHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw Error(PHX_E_ARG, "no such device");
    g_device = device;
  });
}

int phx_dev_warmup(void) {
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_warm_mu);
    if (!g_warm) g_warm = new std::thread([] { (void)hipFree(nullptr); });
  });
}

int phx_dev_render(const phx_model* model, const char* dump, int32_t precision, const int32_t* jobs, int32_t njobs,
                   phx_freqout* res, phx_info* info, phx_progress_fn progress, void* user) {
  return guarded([&] {
    if (!res) throw Error(PHX_E_ARG, "res must not be NULL");
    const int tables = precision == PHX_FP64 ? PHX_TABLE_FP64 : PHX_TABLE_FP32;
    std::unique_ptr<phx_snapshot, void (*)(phx_snapshot*)> s(snapshot_open(model, dump, tables), snapshot_free);
    render(s.get(), model, precision, jobs, njobs, res, info, progress, user, Tune());
  });
}

int phx_snapshot_open(const phx_model* model, const char* dump, int32_t tables, phx_snapshot** out) {
  return guarded([&] {
    if (!out) throw Error(PHX_E_ARG, "out must not be NULL");
    *out = snapshot_open(model, dump, tables);
  });
}

int phx_snapshot_close(phx_snapshot* s) {
  return guarded([&] { snapshot_free(s); });
}

int64_t phx_snapshot_bytes(const phx_snapshot* s) { return s ? s->bytes : 0; }

int phx_snapshot_render(phx_snapshot* s, const phx_model* rq, int32_t precision, const int32_t* jobs, int32_t njobs,
                        phx_freqout* res, phx_info* info) {
  return guarded([&] {
    if (!s) throw Error(PHX_E_ARG, "null snapshot");
    render(s, rq, precision, jobs, njobs, res, info, nullptr, nullptr, s->tune);
  });
}

int phx_snapshot_render_batch(phx_snapshot* s, const phx_model* const* requests, int32_t n, int32_t precision,
                              phx_freqout* const* res, phx_info* info) {
  return guarded([&] {
    if (!s || !requests) throw Error(PHX_E_ARG, "null snapshot or request list");
    render_batch(s, requests, n, precision, res, info);
  });
}

int phx_autotune(phx_snapshot* s, const phx_model* const* requests, int32_t n, int32_t precision,
                 int32_t sample_stride, phx_tune* best, double* rate) {
  return guarded([&] {
    if (!s || !requests) throw Error(PHX_E_ARG, "null snapshot or request list");
    autotune(s, requests, n, precision, sample_stride, best, rate);
  });
}

int phx_snapshot_set_tune(phx_snapshot* s, const phx_tune* t) {
  return guarded([&] {
    if (!s || !t) throw Error(PHX_E_ARG, "null snapshot or tune");
    if (t->queue < -1 || t->queue > 1 || t->blocks_per_cu < 0 || t->order < 0 || t->order > 1 || t->fp64_wave < 0 ||
        t->fp64_wave > 1 || t->refill < 0 || t->refill > 64 || t->escape < 0 || t->escape > 2)
      throw Error(PHX_E_ARG, "tune value out of range");
    s->tune.queue = t->queue;
    s->tune.blocks_per_cu = t->blocks_per_cu;
    s->tune.order = t->order;
    s->tune.fp64_wave = t->fp64_wave;
    s->tune.refill = t->refill > 0 ? t->refill : 64;
    s->tune.escape = t->escape;
  });
}

}  // extern "C"

struct phx_server {
  phx::Server sv;
};

extern "C" {

int phx_server_start(phx_snapshot* const* snaps, int32_t nsnaps, const phx_server_opts* opts, phx_server** out) {
  return guarded([&] {
    if (!snaps || nsnaps < 1 || !out) throw Error(PHX_E_ARG, "server needs at least one snapshot");
    for (int i = 0; i < nsnaps; i++)
      if (!snaps[i]) throw Error(PHX_E_ARG, "null snapshot");
    auto srv = std::make_unique<phx_server>();
    Server& sv = srv->sv;
    sv.snaps.assign(snaps, snaps + nsnaps);
    sv.uni.assign(nsnaps, phx_model{});
    sv.have_uni.assign(nsnaps, 0);
    if (opts) sv.opts = *opts;
    if (sv.opts.precision == 0) sv.opts.precision = PHX_FP32;
    for (phx_snapshot* s : sv.snaps) check_precision(s, sv.opts.precision);
    if (sv.opts.hbm_budget <= 0) sv.opts.hbm_budget = (int64_t)8 << 30;
    if (sv.opts.quantum_jobs <= 0) sv.opts.quantum_jobs = 1 << 18;
    sv.t_start = now_s();
    sv.worker = std::thread(serve_loop, &sv);
    *out = srv.release();
  });
}

int phx_server_submit(phx_server* srv, int32_t snap, const phx_model* request, phx_freqout* res, int64_t* ticket) {
  return guarded([&] {
    if (!srv || !request) throw Error(PHX_E_ARG, "null server or request");
    Server& sv = srv->sv;
    if (snap < 0 || snap >= (int)sv.snaps.size()) throw Error(PHX_E_ARG, "no such snapshot");
    check_request(request);
    auto r = std::make_shared<Request>();
    r->snap = snap;
    r->rq = *request;
    r->host_res = res;
    r->bytes = (int64_t)request->img_width * request->img_height * request->nfreqs * (int64_t)sizeof(phx_freqout);
    if (r->bytes > sv.opts.hbm_budget) throw Error(PHX_E_RANGE, "request larger than the server's HBM budget");
    {
      std::lock_guard<std::mutex> lk(sv.mu);
      if (sv.stop) throw Error(PHX_E_ARG, "server stopped");
      // a launch shares stepping and cuts: the requests of one snapshot must agree with its first
      if (sv.have_uni[snap])
        check_uniform(&sv.uni[snap], request);
      else
        sv.uni[snap] = *request;
      sv.have_uni[snap] = true;
      r->ticket = sv.next_ticket++;
      r->t_submit = now_s();
      sv.waiting.push_back(r);
      sv.all[r->ticket] = r;
      sv.submitted++;
      if (ticket) *ticket = r->ticket;
    }
    sv.cv_work.notify_one();
  });
}

int phx_server_wait(phx_server* srv, int64_t ticket, double timeout_s, int32_t* done, double* latency_s) {
  return guarded([&] {
    if (!srv) throw Error(PHX_E_ARG, "null server");
    Server& sv = srv->sv;
    std::unique_lock<std::mutex> lk(sv.mu);
    auto it = sv.all.find(ticket);
    if (it == sv.all.end()) throw Error(PHX_E_ARG, "unknown ticket");
    std::shared_ptr<Request> r = it->second;
    auto ready = [&] { return r->state >= 2; };
    if (timeout_s < 0)
      sv.cv_done.wait(lk, ready);
    else
      sv.cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
    if (done) *done = r->state >= 2;
    if (latency_s) *latency_s = r->state >= 2 ? r->t_done - r->t_submit : 0.;
    if (r->state == 3) throw Error(PHX_E_DEVICE, "request failed: " + r->error);
    if (r->state >= 2) sv.all.erase(it);
  });
}

int phx_server_stats(phx_server* srv, phx_serve_stats* st) {
  return guarded([&] {
    if (!srv || !st) throw Error(PHX_E_ARG, "null server or stats");
    Server& sv = srv->sv;
    std::lock_guard<std::mutex> lk(sv.mu);
    std::memset(st, 0, sizeof *st);
    st->submitted = sv.submitted;
    st->completed = sv.completed;
    st->failed = sv.failed;
    st->active = (int64_t)sv.active.size();
    st->waiting = (int64_t)sv.waiting.size();
    st->rays = sv.rays;
    st->launches = sv.launches;
    st->busy_s = sv.busy_s;
    st->wall_s = now_s() - sv.t_start;
    st->peak_bytes = sv.peak_bytes;
    st->p50_s = percentile(sv.latencies, 0.50);
    st->p99_s = percentile(sv.latencies, 0.99);
    st->max_s = sv.latencies.empty() ? 0. : *std::max_element(sv.latencies.begin(), sv.latencies.end());
    double sum = 0.;
    for (double v : sv.latencies) sum += v;
    st->mean_s = sv.latencies.empty() ? 0. : sum / (double)sv.latencies.size();
  });
}

int phx_server_reset_stats(phx_server* srv) {
  return guarded([&] {
    if (!srv) throw Error(PHX_E_ARG, "null server");
    Server& sv = srv->sv;
    std::lock_guard<std::mutex> lk(sv.mu);
    sv.latencies.clear();
    sv.completed = sv.failed = sv.rays = sv.launches = sv.submitted = 0;
    sv.busy_s = 0.;
    sv.t_start = now_s();
  });
}

int phx_server_stop(phx_server* srv) {
  return guarded([&] {
    if (!srv) return;
    {
      std::lock_guard<std::mutex> lk(srv->sv.mu);
      srv->sv.stop = true;
    }
    srv->sv.cv_work.notify_all();
    if (srv->sv.worker.joinable()) srv->sv.worker.join();
    for (auto& r : srv->sv.active)
      if (r->d_res) (void)hipFree(r->d_res);
    delete srv;
  });
}

int phx_snapshot_result(phx_snapshot* s, phx_freqout* res, int64_t nres) {
  return guarded([&] {
    if (!s || !res || nres != s->last_nres) throw Error(PHX_E_ARG, "result size mismatch");
    HIPCHECK(hipMemcpy(res, s->res.p, (size_t)nres * sizeof(phx_freqout), hipMemcpyDeviceToHost));
  });
}

int phx_snapshot_prims(phx_snapshot* s, int32_t table, double* out, int64_t n) {
  return guarded([&] {
    const int64_t want = (int64_t)s->geo.nleafs * s->geo.cells * 8;
    if (!s || !out || n != want) throw Error(PHX_E_ARG, "prims size mismatch");
    if (table == PHX_TABLE_FP64 && s->t64.prim) {
      HIPCHECK(hipMemcpy(out, s->t64.prim, n * sizeof(double), hipMemcpyDeviceToHost));
    } else if (table == PHX_TABLE_FP32 && s->t32.prim) {
      std::vector<float> tmp(n);
      HIPCHECK(hipMemcpy(tmp.data(), s->t32.prim, n * sizeof(float), hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < n; i++) out[i] = tmp[i];
    } else {
      throw Error(PHX_E_ARG, "snapshot has no such table");
    }
  });
}

}  // extern "C"
