// The per-lane code of the march: launch parameters, the ray state of each precision, one
// plasma sample, the segment accumulation, and a lane's ray (Lane: start, one pass of the
// reference's per-ray loop, write-out), shared by the gfx950 march kernel (phx_device.hip) and the
// CPU test build of the same code (tests/cpu/lane_host.cpp, PHX_HOST_BUILD).
#pragma once
#include <type_traits>

#include "../../include/phx.h"
#include "phx_march.h"

namespace phxd {

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
  const T* fq;  // two-phase renders: the request's freq[nf], hconst[nf], jscale[nf] (nf = nf_total)
};

// Two-phase renders (requests with many frequencies): the march (phase 1) marches each ray once
// and appends the frequency-independent part of each emitting sample (a Prep record) to the ray's
// chain of record blocks; the transfer kernel (phase 2) then folds every frequency's segment
// operators over the ray's records, in march order, one lane per (ray, frequency).  Blocks of
// RB records are taken from a per-launch pool with one atomic per wave; the jobs are marched in
// chunks whose records the pool holds, and a ray whose block the pool cannot hold is counted, and
// the launch is run again with smaller chunks.
constexpr int RB = 32;    // records per block
constexpr int RECW = 12;  // scalars per record (a Prep record, padded)

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
  // two-phase renders: the record pool (blocks of RB records), its block links, the count of blocks
  // taken, the count of rays the pool could not hold, and per job the ray's first block and record
  // count (count -1: the pool ran out)
  typename ModeOf<M>::T* rec;
  int* rec_next;
  int* rec_used;
  int* rec_ovf;
  int2* rray;
  int rec_cap, pad;
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
  int fetch;                            // steps per fetch: passes between a wave's job fetches (1: every pass)
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
  Pt64 tb64;     // FP64 step: the same in FP64
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
  // precision of the coming step and its first stage; ck = the wavevector at the current point;
  // r32 = e^{x1} of the FP32 position, which the pass's view has computed
  template <class A>
  PHX_DEV void begin(const A& args, bool alive, float ck[4], float r32) {
    f64 = alive && (X[0] < args.lnr_fp64 || X[1] < args.x2_pole || X[1] > PI_D - args.x2_pole);
    if (args.fp64_wave) f64 = ballot(f64) != 0ull;
    if (f64) {
      double k0;
      tb64 = pt64_at(taylor_coef(), X[0], X[1], args.hs64);
      ham_core<double, true>(args.a64, args.hs64, pt, pp, tb64.r, tb64.s2x, tb64.c2x, tb64.s, tb64.c, X[3], X[4], d1,
                             &k0);
      ck[0] = float(k0);
#pragma unroll
      for (int q = 0; q < 3; q++) ck[1 + q] = float(d1[q]);
    } else {
      float b[5], d[5], k0;
      const bool south = base32(b);
      tb = pt32_at_r(r32, b[1], args.P.hs);  // b[0] = float(X[0]): the view's exp of it
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
      const TaylorSC t = taylor_coef();
      double sum[5], y[5];
#pragma unroll
      for (int i = 0; i < 5; i++) {
        sum[i] = d1[i];
        y[i] = X[i] + 0.5 * dt * d1[i];
      }
      // stages 2-4 at the offsets h d of stage 1's point (|h d| <= 0.02; else evaluated afresh)
      double e1 = 0.5 * dt * d1[0], e2 = 0.5 * dt * d1[1];
#pragma unroll 1
      for (int q = 1; q < 4; q++) {
        double d[5], k0;
        const Pt64 p = (fabs(e1) <= 0.02 && fabs(e2) <= 0.02) ? pt64_offset(t, tb64, e1, e2, args.hs64)
                                                               : pt64_at(t, y[0], y[1], args.hs64);
        ham_core<double, true>(args.a64, args.hs64, pt, pp, p.r, p.s2x, p.c2x, p.s, p.c, y[3], y[4], d, &k0);
        const double w = q == 3 ? 1. : 2., h = q == 2 ? dt : 0.5 * dt;
#pragma unroll
        for (int i = 0; i < 5; i++) {
          sum[i] += w * d[i];
          y[i] = X[i] + h * d[i];
        }
        e1 = h * d[0];
        e2 = h * d[1];
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
PHX_DEV SideArgs<M> side_mem(const MarchArgs<M>& A) {
  const unsigned* w = opaque_scalar(reinterpret_cast<const unsigned*>(A.side));
  static_assert(sizeof(SideArgs<M>) % 4 == 0, "whole dwords");
  SideArgs<M> out;
  unsigned* d = reinterpret_cast<unsigned*>(&out);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(SideArgs<M>) / 4); i++) d[i] = load_constant(w, i);
  return out;
}
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
  return side_mem(A);
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
  T rho, uu, Bsq;  // the unit-free inputs of te, ne and B (a sample prepared without plasma units)
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
// kernel's way): 0 no emission, 1 emits, 2 every frequency's operator is dropped.  UNITS false: the
// sample is prepared for any plasma units (a sweep over m_unit and R_high): the Theta_e cut, te, ne,
// B and the Bessel ratios are left out, and sp keeps rho, uu and Bsq instead.
template <int M, bool UNITS = true>
PHX_DEV int sample_prepare(const MarchArgs<M>& A, const TaskDev<typename ModeOf<M>::T>& tk,
                           const typename ModeOf<M>::T px[4], const typename ModeOf<M>::L pl[4],
                           const typename ModeOf<M>::T pk[4], const typename ModeOf<M>::T pf[4],
                           typename ModeOf<M>::T dl, int leaf, Prep<typename ModeOf<M>::T>& sp) {
  using T = typename ModeOf<M>::T;
  T k[4] = {pk[0], pk[1], pk[2], pk[3]};
  const SideArgs<M> S = side_of(A);
  if constexpr (M != MIX) null_resolve(metric(S.P.a, S.P.hs, px[1], px[2]), k, exp(px[1]));
  Fluid<T> fl;
  if (!sample_fluid<T, typename ModeOf<M>::L, UNITS>(S.G, loc_for(A, S), S.P, tk.Ne, tk.Bu, tk.rhigh, pl, leaf, fl))
    return 0;
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
  sp.dl = dl;
  scs(pa, &sp.sth, &sp.cth);
  if constexpr (UNITS) {
    sp.te = fl.te;
    sp.ne = fl.ne;
    sp.B = fl.B;
    kratios(fl.te, &sp.kr1, &sp.kr0);
  } else {
    sp.rho = fl.rho;
    sp.uu = fl.uu;
    sp.Bsq = fl.Bsq;
  }
  return 1;
}
// segment operator at frequency freq (jscale = freq^3, hconst = L / freq); false when it is not
// finite (the reference drops it)
template <class T>
PHX_DEV bool sample_op_at(T freq, T jscale, T hconst, const Prep<T>& sp, T OA[16], T OS[4], T* dtau, T* dtauF) {
  const T nup = freq * sp.ratio;
  const Coef<T> seg = invariant_rotated(thermal(sp.te, sp.ne, nup, sp.B, sp.sth, sp.cth, sp.kr1, sp.kr0), nup,
                                        jscale, sp.c2, sp.s2);
  const T h = sp.dl * hconst;
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
// segment operator of frequency f of a task's chunk
template <class T>
PHX_DEV bool sample_op(const TaskDev<T>& tk, const Prep<T>& sp, int f, T OA[16], T OS[4], T* dtau, T* dtauF) {
  return sample_op_at(tk.freq[f], tk.jscale[f], tk.hconst[f], sp, OA, OS, dtau, dtauF);
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

// One FreqOut (pixel.hpp:26-29) as three 16-byte global stores: the Stokes sums and the optical
// depths when sums, else exact zeros (a ray with no plasma segment, pixel.hpp:216-222).
template <class T>
PHX_DEV void store_freqout(phx_freqout* out, bool sums, const T S[4], T tau, T tauF) {
  const gmut<d2v> o = glb_mut(reinterpret_cast<d2v*>(out));
  if (sums) {
    o[0] = d2v{(double)S[0], (double)S[1]};
    o[1] = d2v{(double)S[2], (double)S[3]};
    o[2] = d2v{(double)tau, (double)tauF};
  } else {
    o[0] = o[1] = o[2] = d2v{0., 0.};
  }
}

// ---------------------------------------------------------------------------- a lane's ray
// What one pass of the loop found out before the pending sample and the step: liveness, whether
// the pending sample can emit, and the one AMR lookup's answers.
template <class T>
struct PassView {
  T cx[4], ck[4], cf[4];  // the current state's position, wavevector, polarization (MIX: kappa)
  T r;                    // its radius
  bool alive, emit, have_dxc;
  int leaf_next;          // the pending sample's cache leaf (leaf_of of the current position), or -1
  T dxc[3];               // the current leaf's cell size (MIX: its reciprocal) for the step rule
};

// One lane's ray: the state of the reference's per-ray loop (src/core/render/pixel.hpp:131-224)
// with its one-sample delay (pixel.hpp:179-211): the pending sample (position, wavevector,
// polarization vector or Walker-Penrose constant, step length), the accumulators of each
// frequency, the turning-point and step counters, and the lane's leaf cache.
template <int M, int NF>
struct Lane {
  using T = typename ModeOf<M>::T;
  using L = typename ModeOf<M>::L;
  using St = StateOf<M>;
  using R = RecOf<M>;
  static constexpr bool KAHAN = M == MIX;
  static constexpr bool REC = NF <= 0;     // phase 1 of a two-phase render: records, no accumulators
  static constexpr bool SWEEP = NF < 0;    // records for any plasma units (a sweep over m_unit, R_high)
  static constexpr int NA = REC ? 1 : NF;  // accumulator sets held
  St st;
  T px[4] = {T(0), T(0), T(0), T(0)}, pk[4] = {T(0), T(0), T(0), T(0)}, pf[4] = {T(0), T(0), T(0), T(0)};
  L pl[4] = {L(0), L(0), L(0), L(0)};
  T pdl = T(0);
  T AA[NA][16], SS[NA][4], SC[NA][4], tau[NA], tauF[NA];
  int steps = 0, turns = 0, stop = 0, np = 0, bad = 0, pix = 0, task = 0, nf = 0;
  int jobn = 0;                           // the ray's job (REC: its entry of the per-job record table)
  int rhead = -1, rcur = -1, rn = 0;      // REC: first and current record block, records appended
  bool rovf = false;                      // REC: the pool had no block left for this ray
  T thdot = T(0), dlp = T(0);
  T r_prev = T(0);  // r of the previous pass's position, the pending sample's (read from the third pass on)
  bool bneg = false, bpos = false, stats = false;
  LeafCache<T> lcache;  // the lane's last leaf and its box (valid across rays: one snapshot per launch)

  // a new ray: job n of the launch (its task, pixel and start record); zero accumulators
  PHX_DEV void start(const SideArgs<M>& S, int n) {
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
    if (pix < 0) stop = 1;  // a padding job (no pixel): the ray ends at its first pass, writing nothing
    jobn = n;
    rhead = rcur = -1;
    rn = 0;
    rovf = false;
    thdot = dlp = T(0);
    bneg = bsign < T(0);
    bpos = bsign > T(0);
  }

  // view of the current state, liveness, and the one AMR lookup of this pass: it serves the
  // pending sample's cache (leaf_of, sampler.hpp:184-190) and the step's cell cap (grid_dxc)
  PHX_DEV PassView<T> look(const MarchArgs<M>& A) {
    PassView<T> v;
    st.view(v.cx, v.ck, v.cf);
    const T r = exp_r(v.cx[1]);
    v.r = r;
    // ln r of the pending sample (the state before the last step; MIX keeps it in FP64 only)
    const T lrs = M == MIX ? T(pl[1]) : px[1];
    const T rs = r_prev;  // radius of the pending sample: the previous pass's r (exp_r of the same lrs)
    // escape: a ray moving outward at r >= rt_outer, whose pending sample lies there too, can never
    // emit again when rt_outer lies outside the photon orbits (a null geodesic in Kerr then has no
    // radial turning point left, so it does not return), and the rest of its march changes none of
    // its outputs: it ends here instead of at r_out (the reference marches on to r_out,
    // pixel.hpp:163, and only its step count differs).  The host enables it only for such rt_outer.
    const bool escaped = A.escape && steps >= 2 && r >= A.P.rt_outer && rs >= A.P.rt_outer && v.cx[1] > lrs;
    v.alive = r < A.r_out && r > A.cut_in && steps < A.max_steps && !stop && !escaped;
    if constexpr (M == MIX) {
      // the pending sample's FP32 position and polarization constant follow from its FP64
      // position and the ray's constant: not carried from pass to pass
#pragma unroll
      for (int q = 0; q < 4; q++) {
        px[q] = T(pl[q]);
        pf[q] = v.cf[q];
      }
    }
    // the pending sample is the state before the last step, i.e. sample number steps - 1; the
    // camera sample (number 0) is never processed (pixel.hpp:186).  A sample at r >= rt_outer
    // cannot emit (pixel.hpp:80) and changes no ray state, so nothing of it is computed.
    v.emit = steps >= 2 && rs < A.P.rt_outer;
    v.leaf_next = -1;
    v.have_dxc = false;
    v.dxc[0] = v.dxc[1] = v.dxc[2] = T(0);
    // far from the grid (r >= rt_outer, no emitting sample pending) nothing is looked up
    if (v.alive && (r < A.P.rt_outer || v.emit)) {
      T wx[4] = {v.cx[0], v.cx[1], v.cx[2], v.cx[3]};
      wrap(wx);
      const bool need_dxc = r < A.P.rt_outer;
      // leaf_of is -1 off the domain.  The mixed march hands the sample its leaf as a hint only:
      // the sample tests the domain and the leaf itself, in FP64
      bool need_leaf = v.emit;
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
        if (need_leaf) v.leaf_next = Lf;
        if (need_dxc && Lf >= 0) {
          v.have_dxc = true;
#pragma unroll
          for (int d = 0; d < 3; d++) v.dxc[d] = lcache.dxc[d];
        }
      }
    }
    return v;
  }

  // REC: appends the frequency-independent record of an emitting sample to the ray's block chain
  // (a new block from the pool every RB records; none left: the ray is marked and records nothing)
  PHX_DEV void record(const MarchArgs<M>& A, const Prep<T>& sp) {
    const SideArgs<M> S = side_mem(A);
    const bool need = !rovf && rn % RB == 0;
    const int b = wave_take(need, S.rec_used);
    if (need) {
      if (b >= S.rec_cap) {
        rovf = true;
        count_one(S.rec_ovf);
      } else {
        if (rcur >= 0) glb_mut(S.rec_next)[rcur] = b;
        else rhead = b;
        rcur = b;
      }
    }
    if (rovf) return;
    T r[RECW];
    if constexpr (SWEEP) {
      const T q[RECW] = {sp.rho, sp.uu, sp.Bsq, sp.sth, sp.cth, sp.c2, sp.s2, sp.ratio, sp.dl, T(0), T(0), T(0)};
#pragma unroll
      for (int i = 0; i < RECW; i++) r[i] = q[i];
    } else {
      const T q[RECW] = {sp.te, sp.ne, sp.B, sp.sth, sp.cth, sp.kr1, sp.kr0, sp.c2, sp.s2, sp.ratio, sp.dl, T(0)};
#pragma unroll
      for (int i = 0; i < RECW; i++) r[i] = q[i];
    }
    const gmut<T> o = glb_mut(S.rec + ((size_t)rcur * RB + rn % RB) * RECW);
#pragma unroll
    for (int q = 0; q < RECW; q++) o[q] = r[q];
    rn++;
  }

  // the pending sample, in place by its own lane (pixel.hpp:52-123, 186-188)
  PHX_DEV void sample_here(const MarchArgs<M>& A, const PassView<T>& v) {
    const TaskDev<T> tk = load_global(side_of(A).tasks + task);
    Prep<T> sp;
    const int status = sample_prepare<M, !SWEEP>(A, tk, px, pl, pk, pf, pdl, v.leaf_next, sp);
    if constexpr (REC) {
      // a sweep's segments and drops are counted per parameter point by the transfer
      if (status && !SWEEP) np++;
      if (status == 2 && !SWEEP) bad += nf;  // the task holds every frequency of the request
      if (status == 1) record(A, sp);
      return;
    }
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

  // the current state becomes the pending sample; then the turning points of theta
  // (pixel.hpp:193-197), the grid-aware step rule and the RK4 step (:200-210).  Returns whether
  // the step ran in FP64 (MIX).
  PHX_DEV bool advance(const MarchArgs<M>& A, PassView<T>& v) {
    // the coming step's precision and first stage, and the wavevector at the current point (not
    // needed before: the pending sample carries its own); evaluated after the sample, so the stage
    // is not held in registers across it
    if constexpr (M == MIX) st.begin(A, v.alive, v.ck, v.r);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      pk[q] = v.ck[q];
      if constexpr (M != MIX) {
        px[q] = v.cx[q];
        pf[q] = v.cf[q];
      }
    }
    st.lpos(v.cx, pl);
    r_prev = v.r;
    pdl = dlp;
    bool f64 = false;
    if (v.alive) {
      if (thdot * v.ck[2] < T(0) && steps > 2) turns++;
      thdot = v.ck[2];
      if ((bneg && turns > A.max_order) || (bpos && turns > A.max_order + 1)) stop = 1;
      T dl;
      if constexpr (M == MIX)
        dl = step_len_mixed(A.stepsize, 1.f / A.stepsize, st.x2p(), v.ck, v.have_dxc, v.dxc);  // dxc: 1/cell size
      else
        dl = step_len(A.stepsize, st.x2p(), v.ck, v.have_dxc, v.dxc);
      f64 = st.step(A, dl);
      dlp = fabs(dl);
      steps++;
    }
    return f64;
  }

  // the ray's outputs (pixel.hpp:213-223): Stokes sums and optical depths per frequency, exact
  // zeros when no plasma segment was processed.  np_counted: the sums were built by this lane
  // (otherwise a ray without segments has folded nothing in and its sums are the zeros anyway).
  PHX_DEV void write_out(const MarchArgs<M>& A, bool np_counted) const {
    if (pix < 0) return;  // a padding job
    if constexpr (REC) {  // the ray's records for the transfer kernel
      const gmut<int> o = glb_mut(reinterpret_cast<int*>(side_mem(A).rray) + 2 * (size_t)jobn);
      o[0] = rhead;
      o[1] = rovf ? -1 : rn;
      return;
    }
    const TaskDev<T> tkw = load_global(side_of(A).tasks + task);
    phx_freqout* o0 = tkw.out + (size_t)pix * tkw.nf_total;
#pragma unroll
    for (int f = 0; f < NF; f++) {
      if (f >= nf) break;
      store_freqout(o0 + f, (!np_counted || np > 0), SS[f], tau[f], tauF[f]);
    }
  }
};

// Phase 2 of a two-phase render for job `job` and frequency f of its task (nothing when f is not
// one of the task's frequencies): the frequency's segment operators folded over the ray's records
// in march order, exactly as the march folds them in place (sample_op, accumulate), and the ray's
// FreqOut stored; an operator that is not finite is dropped (pixel.hpp:107-116).  Returns the
// operators dropped.
template <int M>
PHX_DEV int transfer_one(const TaskDev<typename ModeOf<M>::T>* tasks, const int2* jobs, int job, int f,
                         const typename ModeOf<M>::T* rec, const int* rec_next, const int2* rray) {
  using T = typename ModeOf<M>::T;
  constexpr bool KAHAN = M == MIX;
  const int2 jb = jobs[job];
  if (jb.y < 0) return 0;  // a padding job (no pixel)
  const TaskDev<T> tk = load_global(tasks + jb.x);
  if (f >= tk.nf) return 0;
  const int2 rr = rray[job];
  const T freq = tk.fq[f], hconst = tk.fq[tk.nf + f], jscale = tk.fq[2 * tk.nf + f];
  T AA[16], SS[4], SC[4], tau = T(0), tauF = T(0);
#pragma unroll
  for (int q = 0; q < 16; q++) AA[q] = (q % 5 == 0) ? T(1) : T(0);
#pragma unroll
  for (int q = 0; q < 4; q++) SS[q] = SC[q] = T(0);
  int bad = 0, blk = rr.x;
  for (int i = 0; i < rr.y; i++) {
    if (i > 0 && i % RB == 0) blk = rec_next[blk];
    const T* r = rec + ((size_t)blk * RB + i % RB) * RECW;
    Prep<T> sp;
    sp.te = r[0];
    sp.ne = r[1];
    sp.B = r[2];
    sp.sth = r[3];
    sp.cth = r[4];
    sp.kr1 = r[5];
    sp.kr0 = r[6];
    sp.c2 = r[7];
    sp.s2 = r[8];
    sp.ratio = r[9];
    sp.dl = r[10];
    T OA[16], OS[4], dtau, dtauF;
    if (!sample_op_at(freq, jscale, hconst, sp, OA, OS, &dtau, &dtauF)) {
      bad++;
      continue;
    }
    accumulate<T, KAHAN>(AA, SS, SC, OA, OS);
    tau += dtau;
    tauF += dtauF;
  }
  store_freqout(tk.out + (size_t)jb.y * tk.nf_total + f, true, SS, tau, tauF);
  return bad;
}

// Phase 2 of a parameter sweep (requests that share the camera and differ in m_unit, R_high and
// frequencies; their rays are marched once, recorded without plasma units): for sweep point r (task
// r), job `job` and frequency f, the point's n_e, |B| and Theta_e of each record (plasma_units, as
// derive_fluid forms them), its Theta_e cut, the Bessel ratios, then the segment operators folded
// in as transfer_one does.  Returns the operators dropped; *np (frequency 0 only) the emitting
// segments.  The images equal the point's own render bit for bit; the counts leave out the samples
// with b = 0 or a non-finite rotation, which the march drops for every point without recording them.
template <int M>
PHX_DEV int transfer_sweep(const TaskDev<typename ModeOf<M>::T>* tasks, int r, const int2* jobs, int job, int f,
                           const Plasma<typename ModeOf<M>::T>& P, const typename ModeOf<M>::T* rec,
                           const int* rec_next, const int2* rray, int* np) {
  using T = typename ModeOf<M>::T;
  constexpr bool KAHAN = M == MIX;
  *np = 0;
  const int2 jb = jobs[job];
  if (jb.y < 0) return 0;  // a padding job (no pixel)
  const TaskDev<T> tk = load_global(tasks + r);
  if (f >= tk.nf) return 0;
  const int2 rr = rray[job];
  const T freq = tk.fq[f], hconst = tk.fq[tk.nf + f], jscale = tk.fq[2 * tk.nf + f];
  T AA[16], SS[4], SC[4], tau = T(0), tauF = T(0);
#pragma unroll
  for (int q = 0; q < 16; q++) AA[q] = (q % 5 == 0) ? T(1) : T(0);
#pragma unroll
  for (int q = 0; q < 4; q++) SS[q] = SC[q] = T(0);
  int bad = 0, emit = 0, blk = rr.x;
  for (int i = 0; i < rr.y; i++) {
    if (i > 0 && i % RB == 0) blk = rec_next[blk];
    const T* q = rec + ((size_t)blk * RB + i % RB) * RECW;
    Prep<T> sp;
    plasma_units(P, tk.Ne, tk.Bu, tk.rhigh, q[0], q[1], q[2], &sp.ne, &sp.B, &sp.te);
    if (sp.te > P.te_max || sp.te < P.te_min) continue;  // the Theta_e cut (sampler.hpp:358)
    emit++;
    sp.sth = q[3];
    sp.cth = q[4];
    sp.c2 = q[5];
    sp.s2 = q[6];
    sp.ratio = q[7];
    sp.dl = q[8];
    kratios(sp.te, &sp.kr1, &sp.kr0);
    T OA[16], OS[4], dtau, dtauF;
    if (!sample_op_at(freq, jscale, hconst, sp, OA, OS, &dtau, &dtauF)) {
      bad++;
      continue;
    }
    accumulate<T, KAHAN>(AA, SS, SC, OA, OS);
    tau += dtau;
    tauF += dtauF;
  }
  store_freqout(tk.out + (size_t)jb.y * tk.nf_total + f, true, SS, tau, tauF);
  if (f == 0) *np = emit;
  return bad;
}

// Starting state of job n's ray: pixel -> impact parameters with the position-angle rotation
// (camera.hpp:11-23, jobs.hpp:64-76), the RAPTOR camera photon in FP64 (camera.hpp:71-135), the
// camera polarization basis (tetrad.hpp:111-133: in the march's precision for M32 / M64, in FP64
// for MIX), and for M32 the compensated (hi, lo) split of the FP64 photon (geodesic.hpp:48-60).
template <int M>
PHX_DEV void ray_start(Cam C, int idx, RecOf<M>* out_rec) {
  using T = typename ModeOf<M>::T;
  using R = RecOf<M>;
  if constexpr (M == M32) {  // the FP32 build's camera sees the spin and stretching rounded to float (camera.hpp:77)
    C.a = double(float(C.a));
    C.hs = double(float(C.hs));
  }
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
    double* o = out_rec;
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
    R* o = out_rec;
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

// The jobs of a request's npix pixels x ntasks frequency chunks, in whole waves: each 64-pixel tile
// of the pixel list once per chunk (tile-major), so that a wave of whole-wave refill marches the same
// rays for every chunk, as a render with one chunk does (in mixed precision the rays of a wave share
// its FP64 steps).  With more than one chunk a partial last tile is padded to 64 jobs with jobs of
// no pixel (-1), which end at once; with one chunk the jobs are the pixels in order.
PHX_HD int64_t wave_jobs(int npix, int ntasks) {
  return ntasks <= 1 ? (int64_t)npix : (int64_t)((npix + 63) / 64) * 64 * ntasks;
}
// job n of that list: *t the chunk, return the index in the pixel list (-1: padding)
PHX_HD int wave_job(int n, int npix, int ntasks, int* t) {
  if (ntasks <= 1) {
    *t = 0;
    return n;
  }
  const int tile = n / (64 * ntasks), rem = n - tile * 64 * ntasks;
  *t = rem / 64;
  const int i = tile * 64 + (rem & 63);
  return i < npix ? i : -1;
}

}  // namespace phxd
