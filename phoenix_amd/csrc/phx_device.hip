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
#include <memory>
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
// per-ray start record: hi[12] (x, k, F1), lo[8] (FP32 compensation of x, k), sign of beta, padding
constexpr int RAY_REC = 24;

template <class T, int NF>
struct MarchArgs {
  Grid<T> G;
  Plasma<T> P;
  const T* start;  // [njobs][RAY_REC], in job order
  T stepsize, r_out, cut_in;
  int max_steps, max_order;
  int nf_total, f0, nf, count_stats;
  T freq[NF], hconst[NF], jscale[NF];
  const int* jobs;
  int j0, j1;
  int* counter;
  phx_freqout* res;
  unsigned long long* tot;
};

// ---- one plasma sample, split into its frequency-independent part and the per-frequency
// segment operator (src/core/render/pixel.hpp:52-123)
template <class T>
struct Prep {
  T te, ne, B, sth, cth, kr1, kr0, c2, s2, ratio, dl;
};
// 0: no emission; 1: emits; 2: emits, but the rotation or the frequency ratio is NaN so every
// frequency's operator will be non-finite and dropped
template <class T, int NF>
PHX_DEV int sample_prepare(const MarchArgs<T, NF>& A, const T px[4], const T pk[4], const T pf[4], T dl, int leaf,
                           Prep<T>& sp) {
  T k[4] = {pk[0], pk[1], pk[2], pk[3]};
  const T rs = exp(px[1]);
  const Met<T> gs = metric(A.P.a, A.P.hs, px[1], px[2]);
  if (fabs(dot(gs, k, k)) > T(1e-6) && rs > T(2)) {  // re-solve k^t for k.k = 0 (geodesic.hpp:297-306)
    const T aa = gs.d00;
    const T bc = -(gs.d01 * k[1]) / aa - (gs.d03 * k[3]) / aa;
    const T gz = -(gs.d11 * k[1] * k[1]) / aa - (T(2) * gs.d13 * k[1] * k[3]) / aa - (gs.d22 * k[2] * k[2]) / aa -
                 (gs.d33 * k[3] * k[3]) / aa;
    k[0] = bc + sqrt(bc * bc + gz);
  }
  Fluid<T> fl;
  if (!sample_fluid(A.G, A.P, px, leaf, fl)) return 0;
  const T pa = pitch_angle(fl.g, k, fl.b, fl.U);
  if (pa == T(0) || pa == T(PI_D)) return 0;
  T perp[4], para[4];
  frame(fl.g, A.P.a, k, fl.U, fl.b, perp, para);
  const T Cp = dot(fl.g, perp, pf), Sp = dot(fl.g, para, pf);
  const T nn = Cp * Cp + Sp * Sp;
  sp.c2 = (Cp * Cp - Sp * Sp) / nn;
  sp.s2 = T(2) * Cp * Sp / nn;
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
template <class T, int NF>
PHX_DEV bool sample_op(const MarchArgs<T, NF>& A, const Prep<T>& sp, int f, T OA[16], T OS[4], T* dtau, T* dtauF) {
  const T nup = A.freq[f] * sp.ratio;
  const Coef<T> seg = invariant_rotated(thermal(sp.te, sp.ne, nup, sp.B, sp.sth, sp.cth, sp.kr1, sp.kr0), nup,
                                        A.jscale[f], sp.c2, sp.s2);
  const T h = sp.dl * A.hconst[f];
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
// S += A O_S; A <- A O_A (operators.hpp:74-86), one row at a time in place
template <class T>
PHX_DEV void accumulate(T AA[16], T SS[4], const T OA[16], const T OS[4]) {
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
    SS[i] += s;
#pragma unroll
    for (int j = 0; j < 4; j++) AA[4 * i + j] = row[j];
  }
}

// Persistent march: each lane owns one ray at a time; a lane whose ray ends writes it
// out and takes the next job (wave-aggregated atomic on the job counter), so rays that
// hit the horizon or escape free their lane at once instead of idling until the wave's
// longest ray ends.  One pass of the loop is one pass of render_pixel's loop
// (src/core/render/pixel.hpp:179-211): view, process the pending sample with the
// current position's leaf as its cache, turning points, step size, RK4.
// FP32: two waves per SIMD (a 256-VGPR budget); FP64: one wave, its state does not fit in 256.
template <class T, int NF>
__global__ __launch_bounds__(BLOCK, Lim<T>::fp32 ? 2 : 1) void k_march(const MarchArgs<T, NF> A) {
  constexpr bool F32 = Lim<T>::fp32;
  const int lane = threadIdx.x & 63;
  Ray<T> g;
  T px[4] = {T(0), T(0), T(0), T(0)}, pk[4], pf[4], pdl = T(0);
  T AA[NF][16], SS[NF][4], tau[NF], tauF[NF];
  int steps = 0, turns = 0, stop = 0, np = 0, bad = 0, idx = 0;
  T thdot = T(0), dlp = T(0);
  bool bneg = false, bpos = false, have = false;
  unsigned s_steps = 0, s_np = 0, s_bad = 0;  // per-lane sums (a lane marches well under 2^32 steps)

  for (;;) {
    if (!have) {
      const unsigned long long m = __ballot(1);
      const int leader = __ffsll((long long)m) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(A.counter, (int)__popcll(m));
      base = __shfl(base, leader);
      const int n = A.j0 + base + (int)__popcll(m & ((1ull << lane) - 1ull));
      if (n >= A.j1) break;
      idx = A.jobs[n];
      // the ray's starting state, formed by k_rays (camera photon + polarization basis)
      const T* st = A.start + (size_t)n * RAY_REC;
#pragma unroll
      for (int q = 0; q < 12; q++) g.hi[q] = st[q];
#pragma unroll
      for (int q = 0; q < 8; q++) g.lo[q] = F32 ? st[12 + q] : T(0);
#pragma unroll
      for (int q = 8; q < 12; q++) g.lo[q] = T(0);
      const T bsign = st[20];
#pragma unroll
      for (int f = 0; f < NF; f++) {
#pragma unroll
        for (int q = 0; q < 16; q++) AA[f][q] = (q % 5 == 0) ? T(1) : T(0);
#pragma unroll
        for (int q = 0; q < 4; q++) SS[f][q] = T(0);
        tau[f] = tauF[f] = T(0);
      }
      steps = turns = stop = np = bad = 0;
      thdot = dlp = T(0);
      bneg = bsign < T(0);
      bpos = bsign > T(0);
      have = true;
    }

    // ---- view of the current state
    T cx[4], ck[4], cf[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      cx[q] = F32 ? g.hi[q] + g.lo[q] : g.hi[q];
      ck[q] = F32 ? g.hi[4 + q] + g.lo[4 + q] : g.hi[4 + q];
      cf[q] = F32 ? g.hi[8 + q] + g.lo[8 + q] : g.hi[8 + q];
    }
    const T r = exp(cx[1]);
    const bool alive = r < A.r_out && r > A.cut_in && steps < A.max_steps && !stop;
    const T rs = exp(px[1]);  // radius of the pending sample

    // ---- one AMR lookup serves the pending sample's cache (leaf_of) and the step's cell cap (grid_dxc)
    int leaf_next = -1;
    bool have_dxc = false;
    T dxc[3] = {T(0), T(0), T(0)};
    if (alive) {
      T wx[4] = {cx[0], cx[1], cx[2], cx[3]};
      wrap(wx);
      const bool need_dxc = r < A.P.rt_outer;
      // the cache leaf only matters to a pending sample that can emit (r < rt_outer)
      const bool need_leaf = steps >= 2 && rs < A.P.rt_outer && !off_domain(A.G, wx);
      if (need_dxc || need_leaf) {
        const int L = find_leaf(A.G, wx);
        if (need_leaf) leaf_next = L;
        if (need_dxc && L >= 0) {
          have_dxc = true;
#pragma unroll
          for (int d = 0; d < 3; d++) dxc[d] = A.G.blk[12 * L + 6 + d];
        }
      }
    }

    // ---- the pending sample (src/core/render/pixel.hpp:52-123).  A sample at r >= rt_outer
    // cannot emit (pixel.hpp:80) and changes no ray state, so nothing of it is computed.
    // the pending sample is the state before the last step, i.e. sample number steps - 1; the
    // camera sample (number 0) is never processed (pixel.hpp:186)
    if (steps >= 2 && rs < A.P.rt_outer) {
      Prep<T> sp;
      const int status = sample_prepare(A, px, pk, pf, pdl, leaf_next, sp);
      if (status) np++;
      if (status == 2) {
        // a NaN rotation or frequency ratio makes every frequency's operator non-finite: the
        // reference drops each of them (pixel.hpp:107-116); count them without computing them
        bad += A.nf;
      } else if (status == 1) {
#pragma unroll
        for (int f = 0; f < NF; f++) {
          if (f >= A.nf) break;
          T OA[16], OS[4], dtau, dtauF;
          if (!sample_op(A, sp, f, OA, OS, &dtau, &dtauF)) {
            bad++;
            continue;
          }
          accumulate(AA[f], SS[f], OA, OS);
          tau[f] += dtau;
          tauF[f] += dtauF;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      px[q] = cx[q];
      pk[q] = ck[q];
      pf[q] = cf[q];
    }
    pdl = dlp;

    if (!alive) {
      // ---- ray done: write it out and free the lane (pixel.hpp:213-223)
#pragma unroll
      for (int f = 0; f < NF; f++) {
        if (f >= A.nf) break;
        double2* o = reinterpret_cast<double2*>(A.res + ((size_t)idx * A.nf_total + A.f0 + f));
        if (np > 0) {
          o[0] = make_double2((double)SS[f][0], (double)SS[f][1]);
          o[1] = make_double2((double)SS[f][2], (double)SS[f][3]);
          o[2] = make_double2((double)tau[f], (double)tauF[f]);
        } else {
          o[0] = o[1] = o[2] = make_double2(0., 0.);
        }
      }
      s_steps += steps;
      s_np += np;
      s_bad += bad;
      have = false;
      continue;
    }

    // ---- turning points of theta (pixel.hpp:193-197)
    if (thdot * ck[2] < T(0) && steps > 2) turns++;
    thdot = ck[2];
    if ((bneg && turns > A.max_order) || (bpos && turns > A.max_order + 1)) stop = 1;

    // ---- grid-aware step and RK4 (pixel.hpp:200-210)
    const T dl = step_len(A.stepsize, pole_x2(g), ck, have_dxc, dxc);
    rk4(A.P.a, A.P.hs, g, dl);
    dlp = fabs(dl);
    steps++;
  }

  // wave reduction, then one atomic per wave; steps and segments are counted by the first
  // frequency chunk only, dropped segments (per frequency) by every chunk
  unsigned long long w_steps = s_steps, w_np = s_np, w_bad = s_bad;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    w_steps += __shfl_xor(w_steps, off);
    w_np += __shfl_xor(w_np, off);
    w_bad += __shfl_xor(w_bad, off);
  }
  if (lane == 0) {
    if (A.count_stats) {
      atomicAdd(&A.tot[0], w_steps);
      atomicAdd(&A.tot[1], w_np);
    }
    atomicAdd(&A.tot[2], w_bad);
  }
}

// Starting state of every job's ray: pixel -> impact parameters with the position-angle
// rotation (camera.hpp:11-23, jobs.hpp:64-76), the RAPTOR camera photon in FP64
// (camera.hpp:71-135), the camera polarization basis in the march's precision
// (tetrad.hpp:111-133), and the compensated (hi, lo) split of the FP64 photon (geodesic.hpp:48-60).
template <class T>
__global__ __launch_bounds__(BLOCK) void k_rays(const Cam C, T a, T hs, const int* __restrict__ jobs, int njobs,
                                                T* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= njobs) return;
  const int idx = jobs[n];
  const int i = idx / C.H, j = idx % C.H;
  const double a0 = -0.5 * C.fov_x + (i + 0.5) * C.fov_x / C.W;
  const double b0 = -0.5 * C.fov_y + (j + 0.5) * C.fov_y / C.H;
  const double alpha = a0 * C.cos_pa + b0 * C.sin_pa;
  const double beta = -a0 * C.sin_pa + b0 * C.cos_pa;
  double ph[8];
  camera_photon(C, alpha, beta, ph);
  T F1[4];
  camera_F1<T>(a, hs, ph, F1);
  T* o = out + (size_t)n * RAY_REC;
  for (int q = 0; q < 8; q++) {
    const T hi = T(ph[q]);
    o[q] = hi;
    o[12 + q] = Lim<T>::fp32 ? T(ph[q] - (double)hi) : T(0);
  }
  for (int q = 0; q < 4; q++) o[8 + q] = F1[q];
  o[20] = beta < 0. ? T(-1) : beta > 0. ? T(1) : T(0);
  o[21] = o[22] = o[23] = T(0);
}

// ---------------------------------------------------------------------------- queued march
// FP32 march with a per-wave sample queue in LDS.  Geodesic steps run on every lane each
// pass; a lane whose pending sample can emit pushes it (position, wavevector, polarization
// vector, step length, cache leaf) into its wave's queue instead of processing it at once.
// When the queue holds 64 samples, when a lane holds KMAX of them, or when a ray with queued
// samples ends, all active lanes of the wave compute the queued samples' segment operators
// together, one sample per lane, and each owning lane then folds its own operators into its
// accumulator in the order it queued them, which is the reference's order
// (src/core/render/pixel.hpp:186-188).  The divergent, expensive part of a sample (fluid,
// tetrad, coefficients, 4x4 inverse) thus runs on full waves.
constexpr int QCAP = 128;  // slots per wave (at most 63 + 64 are in use at a flush)
constexpr int QW = 24;     // scalars per slot: 14 in (x, k, F1, dl, leaf), 23 out (O_A, O_S, dtau, dtauF, flag, emits)
constexpr int KMAX = 4;    // queued samples a lane may hold

PHX_DEV void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T, int NF>
__global__ __launch_bounds__(BLOCK, 2) void k_march_q(const MarchArgs<T, NF> A) {
  __shared__ T qbuf[BLOCK / 64][QCAP * QW];
  T* const Q = qbuf[threadIdx.x >> 6];
  constexpr bool F32 = Lim<T>::fp32;
  const int lane = threadIdx.x & 63;
  const unsigned long long lt = (1ull << lane) - 1ull;
  Ray<T> g;
  T px[4] = {T(0), T(0), T(0), T(0)}, pk[4] = {T(0), T(0), T(0), T(0)}, pf[4] = {T(0), T(0), T(0), T(0)};
  T pdl = T(0);
  T AA[NF][16], SS[NF][4], tau[NF], tauF[NF];
  int steps = 0, turns = 0, stop = 0, np = 0, bad = 0, idx = 0;
  T thdot = T(0), dlp = T(0);
  bool bneg = false, bpos = false, have = false, waiting = false;
  unsigned s_steps = 0, s_np = 0, s_bad = 0;
  int qcount = 0, nmine = 0;  // qcount is the same in every active lane of the wave
  int myslot[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; j++) myslot[j] = 0;

  for (;;) {
    if (!have) {
      const unsigned long long m = __ballot(1);
      const int leader = __ffsll((long long)m) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(A.counter, (int)__popcll(m));
      base = __shfl(base, leader);
      const int n = A.j0 + base + (int)__popcll(m & lt);
      if (n >= A.j1) break;
      idx = A.jobs[n];
      const T* st = A.start + (size_t)n * RAY_REC;
#pragma unroll
      for (int q = 0; q < 12; q++) g.hi[q] = st[q];
#pragma unroll
      for (int q = 0; q < 8; q++) g.lo[q] = F32 ? st[12 + q] : T(0);
#pragma unroll
      for (int q = 8; q < 12; q++) g.lo[q] = T(0);
      const T bsign = st[20];
#pragma unroll
      for (int f = 0; f < NF; f++) {
#pragma unroll
        for (int q = 0; q < 16; q++) AA[f][q] = (q % 5 == 0) ? T(1) : T(0);
#pragma unroll
        for (int q = 0; q < 4; q++) SS[f][q] = T(0);
        tau[f] = tauF[f] = T(0);
      }
      steps = turns = stop = np = bad = 0;
      thdot = dlp = T(0);
      bneg = bsign < T(0);
      bpos = bsign > T(0);
      have = true;
    }

    // ---- view, liveness and the one AMR lookup of this pass
    T cx[4], ck[4], cf[4];
    bool alive = false, push = false, have_dxc = false;
    int leaf_next = -1;
    T dxc[3] = {T(0), T(0), T(0)};
    if (!waiting) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        cx[q] = F32 ? g.hi[q] + g.lo[q] : g.hi[q];
        ck[q] = F32 ? g.hi[4 + q] + g.lo[4 + q] : g.hi[4 + q];
        cf[q] = F32 ? g.hi[8 + q] + g.lo[8 + q] : g.hi[8 + q];
      }
      const T r = exp(cx[1]);
      alive = r < A.r_out && r > A.cut_in && steps < A.max_steps && !stop;
      const T rs = exp(px[1]);
      push = steps >= 2 && rs < A.P.rt_outer;
      if (alive) {
        T wx[4] = {cx[0], cx[1], cx[2], cx[3]};
        wrap(wx);
        const bool need_dxc = r < A.P.rt_outer;
        const bool need_leaf = push && !off_domain(A.G, wx);
        if (need_dxc || need_leaf) {
          const int L = find_leaf(A.G, wx);
          if (need_leaf) leaf_next = L;
          if (need_dxc && L >= 0) {
            have_dxc = true;
#pragma unroll
            for (int d = 0; d < 3; d++) dxc[d] = A.G.blk[12 * L + 6 + d];
          }
        }
      }
    }

    // ---- queue the pending sample (every loop-resident lane takes part in the ballot)
    const unsigned long long pm = __ballot(push);
    if (push) {
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
#pragma unroll
      for (int j = 0; j < KMAX; j++)
        if (j == nmine) myslot[j] = slot;
      nmine++;
    }
    qcount += (int)__popcll(pm);

    if (!waiting) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        px[q] = cx[q];
        pk[q] = ck[q];
        pf[q] = cf[q];
      }
      pdl = dlp;
      if (!alive) {
        waiting = true;  // written out once its queued samples are folded in
      } else {
        if (thdot * ck[2] < T(0) && steps > 2) turns++;
        thdot = ck[2];
        if ((bneg && turns > A.max_order) || (bpos && turns > A.max_order + 1)) stop = 1;
        const T dl = step_len(A.stepsize, pole_x2(g), ck, have_dxc, dxc);
        rk4(A.P.a, A.P.hs, g, dl);
        dlp = fabs(dl);
        steps++;
      }
    }

    // ---- flush: all active lanes compute the queued samples, owners accumulate in order
    const bool want = qcount >= QCAP - 64 || nmine >= KMAX || (waiting && nmine > 0);
    if (__ballot(want) != 0ull) {
      const unsigned long long act = __ballot(1);
      const int nact = (int)__popcll(act), rank = (int)__popcll(act & lt);
      for (int base = 0; base < qcount; base += nact) {
        const int e = base + rank;
        const bool mine = e < qcount;
        Prep<T> sp;
        int status = 0;
        if (mine) {
          const T* in = Q + e * QW;
          T x[4], k[4], f1[4];
#pragma unroll
          for (int q = 0; q < 4; q++) {
            x[q] = in[q];
            k[q] = in[4 + q];
            f1[q] = in[8 + q];
          }
          status = sample_prepare(A, x, k, f1, in[12], (int)in[13], sp);
        }
#pragma unroll
        for (int f = 0; f < NF; f++) {
          if (f >= A.nf) break;
          if (mine) {
            T* out = Q + e * QW;
            int flag = 0;  // 1: fold in, 2: dropped as non-finite
            if (status == 1) {
              T OA[16], OS[4], dtau, dtauF;
              flag = sample_op(A, sp, f, OA, OS, &dtau, &dtauF) ? 1 : 2;
              if (flag == 1) {
#pragma unroll
                for (int q = 0; q < 16; q++) out[q] = OA[q];
#pragma unroll
                for (int q = 0; q < 4; q++) out[16 + q] = OS[q];
                out[20] = dtau;
                out[21] = dtauF;
              }
            } else if (status == 2) {
              flag = 2;
            }
            out[22] = T(flag);
            out[23] = T(status != 0 ? 1 : 0);
          }
          wave_sync();
#pragma unroll
          for (int j = 0; j < KMAX; j++) {
            const int sl = myslot[j];
            if (j < nmine && sl >= base && sl < base + nact) {
              const T* rr = Q + sl * QW;
              const int flag = (int)rr[22];
              if (f == 0 && rr[23] != T(0)) np++;
              if (flag == 1) {
                T OA[16], OS[4];
#pragma unroll
                for (int q = 0; q < 16; q++) OA[q] = rr[q];
#pragma unroll
                for (int q = 0; q < 4; q++) OS[q] = rr[16 + q];
                accumulate(AA[f], SS[f], OA, OS);
                tau[f] += rr[20];
                tauF[f] += rr[21];
              } else if (flag == 2) {
                bad++;
              }
            }
          }
          wave_sync();
        }
      }
      qcount = 0;
      nmine = 0;
    }

    if (waiting) {
      // ---- ray done: write it out and free the lane (pixel.hpp:213-223)
#pragma unroll
      for (int f = 0; f < NF; f++) {
        if (f >= A.nf) break;
        double2* o = reinterpret_cast<double2*>(A.res + ((size_t)idx * A.nf_total + A.f0 + f));
        if (np > 0) {
          o[0] = make_double2((double)SS[f][0], (double)SS[f][1]);
          o[1] = make_double2((double)SS[f][2], (double)SS[f][3]);
          o[2] = make_double2((double)tau[f], (double)tauF[f]);
        } else {
          o[0] = o[1] = o[2] = make_double2(0., 0.);
        }
      }
      s_steps += steps;
      s_np += np;
      s_bad += bad;
      have = false;
      waiting = false;
    }
  }

  unsigned long long w_steps = s_steps, w_np = s_np, w_bad = s_bad;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    w_steps += __shfl_xor(w_steps, off);
    w_np += __shfl_xor(w_np, off);
    w_bad += __shfl_xor(w_bad, off);
  }
  if (lane == 0) {
    if (A.count_stats) {
      atomicAdd(&A.tot[0], w_steps);
      atomicAdd(&A.tot[1], w_np);
    }
    atomicAdd(&A.tot[2], w_bad);
  }
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
  T* blk = nullptr;
  T* prim = nullptr;
};

}  // namespace phx

struct phx_snapshot {
  int tables = 0;
  phx_model geo{};  // grid description (no host pointers)
  phx::Tables<float> t32;
  phx::Tables<double> t64;
  int* map = nullptr;
  int fd[3] = {0, 0, 0};
  int64_t bytes = 0;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  phx_freqout* d_res = nullptr;
  size_t res_cap = 0;
  int* d_jobs = nullptr;
  size_t jobs_cap = 0;
  void* d_start = nullptr;  // per-job ray starting states
  size_t start_cap = 0;
  int* d_counter = nullptr;
  unsigned long long* d_tot = nullptr;
  double decode_s = 0., upload_s = 0.;
  int64_t last_nres = 0;
  int last_chunk_jobs = 0;
  int morton_w = 0, morton_h = 0;  // image size whose Morton job list d_jobs holds (0: none)
  int device = 0;
  std::vector<void*> allocs;
};

namespace phx {

static int g_device = 0;
static std::thread* g_warm = nullptr;

static void warm_wait() {
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

static void snapshot_free(phx_snapshot* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (void* p : s->allocs) (void)hipFree(p);
  if (s->d_res) (void)hipFree(s->d_res);
  if (s->d_jobs) (void)hipFree(s->d_jobs);
  if (s->d_start) (void)hipFree(s->d_start);
  if (s->e0) (void)hipEventDestroy(s->e0);
  if (s->e1) (void)hipEventDestroy(s->e1);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

// block records: lower corner, upper corner (lo + nx * dxc in the table's precision, as
// src/core/model/sampler.hpp:93-97 forms it), cell size, padding
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
static void upload_tables(phx_snapshot* s, const phx_model* m, const char* dump, Tables<T>& tb,
                          const std::vector<double>* raw, int per, int nchunks) {
  const std::vector<T> rec = block_records<T>(m);
  tb.blk = dalloc<T>(s, rec.size() * sizeof(T));
  HIPCHECK(hipMemcpyAsync(tb.blk, rec.data(), rec.size() * sizeof(T), hipMemcpyHostToDevice, s->stream));
  const size_t nprim = (size_t)m->nleafs * m->cells * 8;
  tb.prim = dalloc<T>(s, nprim * sizeof(T));
  if (m->prim) {
    std::vector<T> p(nprim);
    for (size_t i = 0; i < nprim; i++) p[i] = T(m->prim[i]);
    HIPCHECK(hipMemcpyAsync(tb.prim, p.data(), nprim * sizeof(T), hipMemcpyHostToDevice, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    return;
  }
  (void)dump;
  (void)raw;
  (void)per;
  (void)nchunks;
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
  if (tables & PHX_TABLE_FP32) upload_tables<float>(s.get(), m, dump, s->t32, nullptr, 0, 0);
  if (tables & PHX_TABLE_FP64) upload_tables<double>(s.get(), m, dump, s->t64, nullptr, 0, 0);
  HIPCHECK(hipStreamSynchronize(s->stream));
  s->upload_s = now_s() - t0;
  if (!m->prim) decode_dump(s.get(), m, dump);
  return s.release();
}

template <class T>
static Grid<T> grid_of(const phx_snapshot* s, const Tables<T>& tb) {
  const phx_model& m = s->geo;
  Grid<T> G;
  G.blk = tb.blk;
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

template <class T, int NF>
static void launch_march(phx_snapshot* s, const Tables<T>& tb, const phx_model* rq, int f0, int nf,
                         const int* d_jobs, int j0, int j1, int count_stats, phx_info* info) {
  MarchArgs<T, NF> A;
  std::memset(&A, 0, sizeof A);
  A.G = grid_of<T>(s, tb);
  const phx_model& m = s->geo;
  A.P.a = T(m.a);
  A.P.hs = T(m.hslope);
  A.P.gam = T(m.gam);
  A.P.Ne = T(rq->Ne);
  A.P.Bu = T(rq->Bu);
  A.P.rhigh = T(rq->rhigh);
  A.P.rlow = T(rq->rlow);
  A.P.ibc = T(rq->ibc);
  A.P.sigma_cut = T(rq->sigma_cut);
  A.P.te_min = T(rq->thetae_min);
  A.P.te_max = T(rq->thetae_max);
  A.P.rt_outer = T(rq->rt_outer);
  A.start = (const T*)s->d_start;
  A.stepsize = T(rq->stepsize);
  A.r_out = T(rq->r_out);
  A.cut_in = T(rq->cut_in);
  A.max_steps = rq->max_steps;
  A.max_order = rq->max_order;
  A.nf_total = rq->nfreqs;
  A.f0 = f0;
  A.nf = nf;
  A.count_stats = count_stats;
  for (int f = 0; f < nf; f++) {
    A.freq[f] = T(rq->freq[f0 + f]);
    A.hconst[f] = T(rq->hconst[f0 + f]);
    A.jscale[f] = T(rq->jscale[f0 + f]);
  }
  A.jobs = d_jobs;
  A.j0 = j0;
  A.j1 = j1;
  A.counter = s->d_counter;
  A.res = s->d_res;
  A.tot = s->d_tot;
  // FP32 with up to 4 frequencies per lane marches with the per-wave sample queue (PHX_MARCH=direct
  // selects the direct kernel); FP64 and wider FP32 chunks march directly
  bool queued = false;
  const void* kfn = reinterpret_cast<const void*>(&k_march<T, NF>);
  if constexpr (Lim<T>::fp32 && NF <= 4) {
    const char* e = std::getenv("PHX_MARCH");
    queued = !(e && !std::strcmp(e, "direct"));
    if (queued) kfn = reinterpret_cast<const void*>(&k_march_q<T, NF>);
  }
  int cus = 0, per_cu = 0;
  HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device));
  HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, BLOCK, 0));
  per_cu = std::max(per_cu, 1);
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
  if constexpr (Lim<T>::fp32 && NF <= 4) {
    if (queued)
      k_march_q<T, NF><<<grid, BLOCK, 0, s->stream>>>(A);
    else
      k_march<T, NF><<<grid, BLOCK, 0, s->stream>>>(A);
  } else {
    k_march<T, NF><<<grid, BLOCK, 0, s->stream>>>(A);
  }
  HIPCHECK(hipGetLastError());
  s->last_chunk_jobs = njobs;
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

template <class T>
static void march_all(phx_snapshot* s, const Tables<T>& tb, const phx_model* rq, const int* d_jobs, int njobs,
                      phx_info* info, phx_progress_fn progress, void* user) {
  // ray starting states for every job (shared by all frequency chunks)
  const size_t need = (size_t)std::max(njobs, 1) * RAY_REC * sizeof(T);
  if (need > s->start_cap) {
    if (s->d_start) HIPCHECK(hipFree(s->d_start));
    s->d_start = nullptr;
    HIPCHECK(hipMalloc(&s->d_start, need));
    s->start_cap = need;
  }
  {
    const phx_model& m = s->geo;
    Cam C;
    C.a = (double)T(m.a);  // the camera uses the table-precision spin and stretching (camera.hpp:77)
    C.hs = (double)T(m.hslope);
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
    if (njobs > 0)
      k_rays<T><<<(njobs + BLOCK - 1) / BLOCK, BLOCK, 0, s->stream>>>(C, T(m.a), T(m.hslope), d_jobs, njobs,
                                                                      (T*)s->d_start);
    HIPCHECK(hipGetLastError());
  }
  int chunk = njobs;
  if (const char* e = std::getenv("PHX_CHUNK"))
    if (std::atoi(e) > 0) chunk = std::atoi(e);
  const int nfr = rq->nfreqs;
  const int per = Lim<T>::fp32 ? 8 : 4;  // frequencies per march: keeps the FP64 state out of scratch
  for (int f0 = 0; f0 < nfr; f0 += per) {
    const int nf = std::min(per, nfr - f0);
    for (int j0 = 0; j0 < njobs; j0 += chunk) {
      const int j1 = std::min(njobs, j0 + chunk);
      const int stats = f0 == 0;
      if (nf == 1)
        launch_march<T, 1>(s, tb, rq, f0, nf, d_jobs, j0, j1, stats, info);
      else if (nf == 2)
        launch_march<T, 2>(s, tb, rq, f0, nf, d_jobs, j0, j1, stats, info);
      else if (nf <= 4)
        launch_march<T, 4>(s, tb, rq, f0, nf, d_jobs, j0, j1, stats, info);
      else if constexpr (Lim<T>::fp32)
        launch_march<T, 8>(s, tb, rq, f0, nf, d_jobs, j0, j1, stats, info);
      if (progress) {
        HIPCHECK(hipStreamSynchronize(s->stream));
        if (f0 + nf >= nfr) progress(j1 - j0, user);
      }
    }
  }
}

static void check_request(const phx_snapshot* s, const phx_model* rq, int precision) {
  if (!rq) throw Error(PHX_E_ARG, "null request");
  if (rq->img_width <= 0 || rq->img_height <= 0) throw Error(PHX_E_ARG, "image size must be positive");
  if ((int64_t)rq->img_width * rq->img_height > (int64_t)1 << 30) throw Error(PHX_E_RANGE, "image too large");
  if (rq->nfreqs < 1 || rq->nfreqs > PHX_MAX_FREQS) throw Error(PHX_E_ARG, "nfreqs must be in 1..32");
  if (precision == PHX_FP32 && !(s->tables & PHX_TABLE_FP32)) throw Error(PHX_E_ARG, "snapshot has no FP32 table");
  if (precision == PHX_FP64 && !(s->tables & PHX_TABLE_FP64)) throw Error(PHX_E_ARG, "snapshot has no FP64 table");
  if (precision != PHX_FP32 && precision != PHX_FP64)
    throw Error(PHX_E_ARG, "precision must be PHX_FP32 or PHX_FP64");
}

static void render(phx_snapshot* s, const phx_model* rq, int precision, const int32_t* jobs, int32_t njobs,
                   phx_freqout* res, phx_info* info, phx_progress_fn progress, void* user) {
  check_request(s, rq, precision);
  HIPCHECK(hipSetDevice(s->device));
  const int64_t npix = (int64_t)rq->img_width * rq->img_height;
  std::vector<int32_t> order;
  // the full-image Morton list is built once per image size and stays on the device
  const bool reuse = !jobs && s->morton_w == rq->img_width && s->morton_h == rq->img_height;
  if (!jobs) {
    njobs = (int32_t)npix;
    if (!reuse) {
      order.resize(npix);
      if (phx_schedule(rq, 1, order.data(), (int32_t)npix) != PHX_OK) throw Error(PHX_E_ARG, phx_last_error());
      jobs = order.data();
    }
  } else {
    if (njobs < 0 || njobs > npix) throw Error(PHX_E_ARG, "njobs out of range");
    for (int32_t n = 0; n < njobs; n++)
      if (jobs[n] < 0 || jobs[n] >= npix) throw Error(PHX_E_ARG, "job index outside the image");
  }
  const size_t nres = (size_t)npix * rq->nfreqs;
  if (nres > s->res_cap) {
    if (s->d_res) HIPCHECK(hipFree(s->d_res));
    s->d_res = nullptr;
    HIPCHECK(hipMalloc((void**)&s->d_res, nres * sizeof(phx_freqout)));
    s->res_cap = nres;
  }
  if ((size_t)njobs > s->jobs_cap) {
    if (s->d_jobs) HIPCHECK(hipFree(s->d_jobs));
    s->d_jobs = nullptr;
    s->morton_w = s->morton_h = 0;
    HIPCHECK(hipMalloc((void**)&s->d_jobs, std::max<size_t>(njobs, 1) * sizeof(int32_t)));
    s->jobs_cap = std::max<size_t>(njobs, 1);
  }
  phx_info local;
  if (!info) info = &local;
  std::memset(info, 0, sizeof *info);
  hipDeviceProp_t prop;
  HIPCHECK(hipGetDeviceProperties(&prop, s->device));
  std::snprintf(info->device, sizeof info->device, "%s (%s)", prop.name, prop.gcnArchName);
  if (jobs) {
    HIPCHECK(hipMemcpyAsync(s->d_jobs, jobs, (size_t)njobs * sizeof(int32_t), hipMemcpyHostToDevice, s->stream));
    s->morton_w = s->morton_h = 0;
    if (!order.empty()) {
      s->morton_w = rq->img_width;
      s->morton_h = rq->img_height;
    }
  }
  if ((int64_t)njobs < npix) HIPCHECK(hipMemsetAsync(s->d_res, 0, nres * sizeof(phx_freqout), s->stream));
  HIPCHECK(hipMemsetAsync(s->d_tot, 0, 4 * sizeof(unsigned long long), s->stream));
  HIPCHECK(hipEventRecord(s->e0, s->stream));
  if (precision == PHX_FP32)
    march_all<float>(s, s->t32, rq, s->d_jobs, njobs, info, progress, user);
  else
    march_all<double>(s, s->t64, rq, s->d_jobs, njobs, info, progress, user);
  HIPCHECK(hipEventRecord(s->e1, s->stream));
  HIPCHECK(hipEventSynchronize(s->e1));
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, s->e0, s->e1));
  unsigned long long tot[4] = {0, 0, 0, 0};
  HIPCHECK(hipMemcpy(tot, s->d_tot, sizeof tot, hipMemcpyDeviceToHost));
  int taken = 0;
  HIPCHECK(hipMemcpy(&taken, s->d_counter, sizeof taken, hipMemcpyDeviceToHost));
  if (taken < s->last_chunk_jobs)  // every lane takes jobs until the list is exhausted
    throw Error(PHX_E_DEVICE, "march kernel did not run to completion (" + std::to_string(taken) + " of " +
                                  std::to_string(s->last_chunk_jobs) + " jobs taken)");
  info->kernel_s = ms * 1e-3;
  info->steps = (int64_t)tot[0];
  info->nplasma = (int64_t)tot[1];
  info->bad = (int64_t)tot[2];
  info->rays = (int64_t)njobs * rq->nfreqs;
  info->decode_s = s->decode_s;
  info->upload_s = s->upload_s;
  s->last_nres = (int64_t)nres;
  if (res) HIPCHECK(hipMemcpy(res, s->d_res, nres * sizeof(phx_freqout), hipMemcpyDeviceToHost));
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
    if (!g_warm) g_warm = new std::thread([] { (void)hipFree(nullptr); });
  });
}

int phx_dev_render(const phx_model* model, const char* dump, int32_t precision, const int32_t* jobs, int32_t njobs,
                   phx_freqout* res, phx_info* info, phx_progress_fn progress, void* user) {
  return guarded([&] {
    if (!res) throw Error(PHX_E_ARG, "res must not be NULL");
    const int tables = precision == PHX_FP64 ? PHX_TABLE_FP64 : PHX_TABLE_FP32;
    std::unique_ptr<phx_snapshot, void (*)(phx_snapshot*)> s(snapshot_open(model, dump, tables), snapshot_free);
    render(s.get(), model, precision, jobs, njobs, res, info, progress, user);
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
    render(s, rq, precision, jobs, njobs, res, info, nullptr, nullptr);
  });
}

int phx_snapshot_result(phx_snapshot* s, phx_freqout* res, int64_t nres) {
  return guarded([&] {
    if (!s || !res || nres != s->last_nres) throw Error(PHX_E_ARG, "result size mismatch");
    HIPCHECK(hipMemcpy(res, s->d_res, (size_t)nres * sizeof(phx_freqout), hipMemcpyDeviceToHost));
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
