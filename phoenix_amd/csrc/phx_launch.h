// Host side of a march launch, shared by the device runtime (phx_device.hip) and the CPU test
// build of the per-ray code (tests/cpu/lane_host.cpp): the device tables a snapshot is turned
// into (block records, flattened leaf map), the tasks of a request (frequency chunks), its
// camera, and the launch arguments of a precision.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "phx_internal.h"
#include "phx_lane.h"

namespace phx {
using namespace phxd;

// Launch knobs (picked by the autotuner or from the environment)
struct Tune {
  int queue = -1;         // -1: default (samples in place), 0: samples in place, 1: LDS queue
  int blocks_per_cu = 0;  // 0: the occupancy limit
  int order = 0;          // batch job order: 0 request after request, 1 requests interleaved per wave
  int fp64_wave = 1;      // mixed march: 1 = a lane's FP64 step takes its whole wave to FP64
  int refill = 64;        // idle lanes a wave gathers before they take new jobs (64: the whole wave)
  int escape = 0;         // 0: by precision (MIXED ends escaping rays at rt_outer, FP32 / FP64 march them
                          // to r_out as the reference does), 1: march to r_out, 2: end at rt_outer
  int block = 256;        // workgroup size of the in-place march: 64 or 256 lanes
  int fetch = 1;          // steps per fetch: a wave with rays in flight takes jobs every `fetch` passes
};

// block records: lower corner, upper corner (lo + nx * dxc in the table's precision, as
// src/core/model/sampler.hpp:93-97 forms it), cell size, reciprocal cell size
template <class T>
inline std::vector<T> block_records(const phx_model* m) {
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
inline std::vector<int> flat_map(const phx_model* m, int fd[3]) {
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

// the march's view of a snapshot's tables in precision T
template <class T>
inline Grid<T> grid_from(const phx_model& m, const T* blk, const T* meta, const T* prim, const int* map,
                         const int fd[3]) {
  Grid<T> G;
  G.blk = blk;
  G.meta = meta;
  G.prim = prim;
  G.map = map;
  G.nleafs = m.nleafs;
  G.cells = m.cells;
  G.nx0 = m.nx[0];
  G.nx1 = m.nx[1];
  G.nx2 = m.nx[2];
  G.fd0 = fd[0];
  G.fd1 = fd[1];
  G.fd2 = fd[2];
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

// x2 at which |cos theta| = 0.99 on the northern side (theta = x2 + h/2 sin 2 x2), by Newton
inline double pole_x2_limit(double hs) {
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

// Radius of the retrograde circular photon orbit of a Kerr black hole of spin a (M = 1), the
// outermost of the spherical photon orbits: r = 2 (1 + cos(2/3 arccos |a|)) (Bardeen, Press &
// Teukolsky 1972), 3 at a = 0 and 4 at |a| = 1.  Outside it a null geodesic has no spherical
// orbit to approach, and one moving outward has no radial turning point left.
inline double retrograde_photon_orbit(double a) {
  return 2. * (1. + std::cos(2. / 3. * std::acos(std::min(1., std::fabs(a)))));
}

// Whether escaping rays may end at rt_outer: asked for (tune 2, or by default in the mixed march)
// and rt_outer outside the photon orbits, so that a ray moving outward there cannot come back
// inside rt_outer and emit again.  Inside them a ray may still turn, and every ray is marched on.
inline bool escape_allowed(int tune_escape, bool mixed, double a, double rt_outer) {
  const bool asked = tune_escape == 2 || (tune_escape == 0 && mixed);
  return asked && rt_outer > retrograde_photon_orbit(a);
}

inline Cam cam_of(const phx_model& geo, const phx_model* rq) {
  Cam C;
  C.a = geo.a;
  C.hs = geo.hslope;
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
inline int add_tasks(const phx_model* rq, phx_freqout* out, std::vector<TaskDev<T>>& tasks) {
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

// The task of a request in a two-phase render: every frequency in one task (fq: the device copy of
// its frequency table, filled from freq_table)
template <class T>
inline int add_task_all(const phx_model* rq, phx_freqout* out, const T* fq, std::vector<TaskDev<T>>& tasks) {
  TaskDev<T> t;
  std::memset(&t, 0, sizeof t);
  t.Ne = T(rq->Ne);
  t.Bu = T(rq->Bu);
  t.rhigh = T(rq->rhigh);
  t.nf = rq->nfreqs;
  t.nf_total = rq->nfreqs;
  t.out = out;
  t.stats = 1;
  t.fq = fq;
  tasks.push_back(t);
  return 1;
}
// freq[nf], hconst[nf], jscale[nf] of a request, rounded to T as the chunked tasks round them
template <class T>
inline void freq_table(const phx_model* rq, std::vector<T>& out) {
  const int nf = rq->nfreqs;
  const size_t o = out.size();
  out.resize(o + 3 * (size_t)nf);
  for (int f = 0; f < nf; f++) {
    out[o + f] = T(rq->freq[f]);
    out[o + nf + f] = T(rq->hconst[f]);
    out[o + 2 * nf + f] = T(rq->jscale[f]);
  }
}

// The device buffers one launch reads besides the snapshot's tables
struct LaunchBufs {
  const void* tasks;
  const int2* jobs;
  const void* start;
  int* counter;
  unsigned long long* tot;
  const void* side;  // device copy of the SideArgs (MIX, FP64)
  // two-phase renders: the record pool, its block links and count of blocks taken, the per-job
  // record table, and the pool's capacity in blocks (0: not a two-phase launch)
  void* rec = nullptr;
  int* rec_next = nullptr;
  int* rec_used = nullptr;
  int* rec_ovf = nullptr;
  void* rray = nullptr;
  int rec_cap = 0;
};

// the plasma parameters a launch shares, in precision T: snapshot m's spin, stretching and
// adiabatic index, request rq's cuts and R_low
template <class T>
inline Plasma<T> plasma_of(const phx_model& m, const phx_model* rq) {
  Plasma<T> P;
  P.a = T(m.a);
  P.hs = T(m.hslope);
  P.gam = T(m.gam);
  P.rlow = T(rq->rlow);
  P.ibc = T(rq->ibc);
  P.sigma_cut = T(rq->sigma_cut);
  P.te_min = T(rq->thetae_min);
  P.te_max = T(rq->thetae_max);
  P.rt_outer = T(rq->rt_outer);
  return P;
}

// The launch arguments of precision M: the grid G (and the FP64 locating table GL of the mixed
// march), the launch-wide parameters of request rq on snapshot geo, the job range and the knobs.
template <int M>
inline void fill_args(MarchArgs<M>& A, SideArgs<M>& side, const Grid<typename ModeOf<M>::T>& G,
                      const Loc<double>& GL, const phx_model& m, const phx_model* rq, const Tune& tn,
                      const LaunchBufs& b, int j0, int j1) {
  using T = typename ModeOf<M>::T;
  std::memset(&A, 0, sizeof A);
  std::memset(&side, 0, sizeof side);
  A.G = G;
  side.G = G;
  side.GL = GL;
  A.P = plasma_of<T>(m, rq);
  side.P = A.P;
  A.start = b.start;
  A.stepsize = T(rq->stepsize);
  A.r_out = T(rq->r_out);
  A.cut_in = T(rq->cut_in);
  A.a64 = m.a;
  A.hs64 = m.hslope;
  A.lnr_fp64 = std::log(2. * (1. + std::sqrt(std::max(0., 1. - m.a * m.a))));  // r < 2 r_H
  A.x2_pole = pole_x2_limit(m.hslope);                                          // |cos theta| > 0.99
  A.fp64_wave = tn.fp64_wave;
  A.refill = std::min(64, std::max(1, tn.refill));
  A.escape = escape_allowed(tn.escape, M == MIX, m.a, rq->rt_outer);
  A.fetch = std::max(1, tn.fetch);
  A.max_steps = rq->max_steps;
  A.max_order = rq->max_order;
  A.tasks = static_cast<const TaskDev<T>*>(b.tasks);
  A.jobs = b.jobs;
  A.j0 = j0;
  A.j1 = j1;
  A.counter = b.counter;
  A.tot = b.tot;
  side.tasks = A.tasks;
  side.jobs = A.jobs;
  side.start = A.start;
  side.counter = A.counter;
  side.tot = A.tot;
  side.j0 = A.j0;
  side.j1 = A.j1;
  side.rec = static_cast<T*>(b.rec);
  side.rec_next = b.rec_next;
  side.rec_used = b.rec_used;
  side.rec_ovf = b.rec_ovf;
  side.rray = static_cast<int2*>(b.rray);
  side.rec_cap = b.rec_cap;
  A.side = static_cast<const SideArgs<M>*>(b.side);
}

}  // namespace phx
