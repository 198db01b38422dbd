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
#include <map>
#include <memory>
#include <array>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "phx_launch.h"

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

// an integer knob from the environment, or its default
static int env_int(const char* name, int dflt) {
  const char* e = std::getenv(name);
  return e && *e ? std::atoi(e) : dflt;
}

// ============================================================================ kernels
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
#ifndef PHX_FP64_WAVES_PER_SIMD
#define PHX_FP64_WAVES_PER_SIMD 1  // occupancy the FP64 march is compiled for
#endif
// BS: the workgroup size (64 or 256 lanes in place; 256 with the queue).  The occupancy bound is in
// waves per SIMD, whatever the workgroup size.
template <int M, int NF, bool QUEUE, int BS>
__global__ __launch_bounds__(BS, M == M64 ? PHX_FP64_WAVES_PER_SIMD : M == MIX ? PHX_MIX_WAVES_PER_SIMD : 2) void k_march(
    const MarchArgs<M> A) {
  using T = typename ModeOf<M>::T;
  using L = typename ModeOf<M>::L;
  constexpr bool KAHAN = M == MIX;
  constexpr int NW = QUEUE ? BS / 64 : 1;  // per-wave LDS areas of the sample queue
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
  Lane<M, NF> ln;
  bool have = false;
  unsigned s_steps = 0, s_np = 0, s_bad = 0;  // per-lane sums (a lane marches well under 2^32 steps)
  unsigned s_fp64 = 0;                        // MIX: steps taken in FP64
  // queue bookkeeping: qcount and nleft are the same in every active lane of the wave; nmine and
  // ncm count this lane's queued samples and its candidates not yet folded in
  int qcount = 0, nleft = 0, nmine = 0, ncm = 0;
  int since = A.fetch;  // passes since the wave last took jobs (the same in every lane of the wave)

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
    // steps per fetch: a wave with rays in flight takes jobs for its idle lanes only every A.fetch
    // passes (their atomic, job, task and start-record loads batched); 1 = at every pass
    if (A.fetch > 1) {
      if (since < A.fetch && __ballot(have) != 0ull) wait = true;
      if (__ballot(!have && !wait) != 0ull) since = 0;
      since++;
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
      ln.start(S, n);
      have = true;
    }
    if (!have) continue;  // waiting for the wave's refill
    if constexpr (QUEUE) {
      // a lane that waited for the refill skipped the queue updates of the passes it sat out
      qcount = qs[0];
      nleft = qs[1];
    }

    PassView<T> v = ln.look(A);
    const bool alive = v.alive, emit = v.emit;

    if constexpr (QUEUE) {
      // ---- queue the pending sample (every loop-resident lane takes part in the ballot)
      const unsigned long long pm = __ballot(emit);
      if (emit) {
        const int slot = qcount + (int)__popcll(pm & lt);
        T* e = Q + slot * QW;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          e[q] = ln.px[q];
          e[4 + q] = ln.pk[q];
          e[8 + q] = ln.pf[q];
        }
        e[12] = ln.pdl;
        e[13] = T(v.leaf_next);
        if constexpr (!std::is_same<T, L>::value) {
          double* e64 = reinterpret_cast<double*>(e + 14);
#pragma unroll
          for (int q = 0; q < 3; q++) e64[q] = ln.pl[1 + q];
        }
        e[20] = T(ln.task);
        e[21] = T(lane);
        nmine++;
      }
      qcount += (int)__popcll(pm);
    } else if (emit) {
      ln.sample_here(A, v);  // the pending sample, in place
    }

    if (ln.advance(A, v) && ln.stats) s_fp64++;

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
              if (f < ln.nf && rr[22] == T(1)) {
                T OA[16], OS[4];
#pragma unroll
                for (int q = 0; q < 16; q++) OA[q] = rr[q];
#pragma unroll
                for (int q = 0; q < 4; q++) OS[q] = rr[16 + q];
                accumulate<T, KAHAN>(ln.AA[f], ln.SS[f], ln.SC[f], OA, OS);
                ln.tau[f] += rr[20];
                ln.tauF[f] += rr[21];
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
      // ---- ray done: write it out and free the lane (pixel.hpp:213-223); with the queue, np is
      // counted where the samples are computed
      ln.write_out(A, !QUEUE);
      if (ln.stats) {
        s_steps += ln.steps;
        s_np += ln.np;
      }
      s_bad += ln.bad;
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

// Starting state of every job's ray (ray_start); cams[task] is the task's camera.
template <int M>
__global__ __launch_bounds__(BLOCK) void k_rays(const Cam* __restrict__ cams, const int2* __restrict__ jobs,
                                                int njobs, void* __restrict__ out_) {
  using R = RecOf<M>;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= njobs) return;
  const int2 jb = jobs[n];
  constexpr int REC = M == MIX ? StHam::REC : StRef<typename ModeOf<M>::T>::REC;
  R* const o = static_cast<R*>(out_) + (size_t)n * REC;
  if (jb.y < 0) {  // a padding job (no pixel): a zero record, which its lane never marches
    for (int q = 0; q < REC; q++) o[q] = R(0);
    return;
  }
  ray_start<M>(cams[jb.x], jb.y, o);
}

// zeroes a launch's job counter and (when tot is given) its run statistics: a kernel rather than a
// memset, so that a launch sequence stays valid inside a captured HIP graph
__global__ void k_reset(int* counter, unsigned long long* tot) {
  if (threadIdx.x == 0) counter[0] = 0;
  // two-phase renders: the record blocks taken and the rays the pool could not hold
  if (tot && (threadIdx.x == 16 || threadIdx.x == 17)) counter[threadIdx.x] = 0;
  if (tot && threadIdx.x < 4) tot[threadIdx.x] = 0ull;
}

// Phase 2 of a two-phase render: one lane per (job, frequency), lane n taking job n / F and
// frequency n % F (F: the largest frequency count of the launch's requests), running
// transfer_one; the dropped operators are counted per wave.
template <int M>
__global__ __launch_bounds__(256) void k_transfer(const TaskDev<typename ModeOf<M>::T>* __restrict__ tasks,
                                                  const int2* __restrict__ jobs, int j0, int j1, int F,
                                                  const typename ModeOf<M>::T* __restrict__ rec,
                                                  const int* __restrict__ rec_next, const int2* __restrict__ rray,
                                                  unsigned long long* tot) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long q = n / F;
  const int job = j0 + (int)q, f = (int)(n - q * F);
  unsigned long long w = job < j1 ? (unsigned long long)transfer_one<M>(tasks, jobs, job, f, rec, rec_next, rray) : 0ull;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off);
  if ((threadIdx.x & 63) == 0 && w) atomicAdd(&tot[2], w);
}

// Phase 2 of a parameter sweep: lane n takes frequency n % F, sweep point (n / F) % R and job
// j0 + n / (F R) (the points of one ray adjacent), running transfer_sweep; the emitting segments and
// the dropped operators are counted per wave.
template <int M>
__global__ __launch_bounds__(256) void k_transfer_sweep(const TaskDev<typename ModeOf<M>::T>* __restrict__ tasks, int R,
                                                        const int2* __restrict__ jobs, int j0, int j1, int F,
                                                        const Plasma<typename ModeOf<M>::T> P,
                                                        const typename ModeOf<M>::T* __restrict__ rec,
                                                        const int* __restrict__ rec_next,
                                                        const int2* __restrict__ rray, unsigned long long* tot) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long q = n / F, jq = q / R;
  const int job = j0 + (int)jq, f = (int)(n - q * F), r = (int)(q - jq * R);
  int np = 0;
  unsigned long long w = job < j1 ? (unsigned long long)transfer_sweep<M>(tasks, r, jobs, job, f, P, rec, rec_next, rray, &np)
                                  : 0ull;
  unsigned long long e = (unsigned long long)np;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    w += __shfl_xor(w, off);
    e += __shfl_xor(e, off);
  }
  if ((threadIdx.x & 63) == 0) {
    if (w) atomicAdd(&tot[2], w);
    if (e) atomicAdd(&tot[1], e);
  }
}

// one value to device memory (a launch's side arguments): a kernel rather than a copy from host
// memory, so that no host staging is involved and a launch sequence stays capturable
template <class V>
__global__ void k_store(const V v, V* dst) {
  *dst = v;
}

// The flux of each frequency of a result block (output.hpp:43-53, compute_spec): the sum over the
// pixels whose I + Q + U + V is finite of IQUV dOm, times the Jy scale; one workgroup per frequency.
// Each thread sums the pixels t, t + BLOCK, ... in index order, and the workgroup adds the partial
// sums in a fixed tree: the same order at every call (not the reference's pixel-serial order, from
// which it differs by double rounding only).
__global__ __launch_bounds__(BLOCK) void k_flux(const phx_freqout* __restrict__ res, long npix, int nf, double dOm,
                                                double jy, double* __restrict__ out) {
  __shared__ double sh[4][BLOCK];
  const int f = blockIdx.x;
  double acc[4] = {0., 0., 0., 0.};
  for (long p = threadIdx.x; p < npix; p += BLOCK) {
    const phx_freqout r = res[p * nf + f];
    if (!isfinite(r.IQUV[0] + r.IQUV[1] + r.IQUV[2] + r.IQUV[3])) continue;
    for (int q = 0; q < 4; q++) acc[q] += r.IQUV[q] * dOm;
  }
  for (int q = 0; q < 4; q++) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < 4; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[4 * f + threadIdx.x] = sh[threadIdx.x][0] * jy;
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

// grow-only device buffer; a fixed one (a window of preallocated memory) cannot grow
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  bool fixed = false;
};

// The per-launch device buffers of a march: jobs, ray start records, tasks, cameras, the launch's
// side arguments, the job counter and run statistics, and timing events.  A snapshot has one for
// its renders; the serving loop keeps two per snapshot, so that one launch is prepared and queued
// while the other runs.
struct Slot {
  DevBuf jobs, start, tasks, cams, side;
  DevBuf fq, rec, rnext, rray;  // two-phase renders: frequency tables, record pool, block links, per-job records
  int rec_cap = 0;              // record blocks the pool holds
  int rec_chunk = 0;            // jobs per recording chunk of the last two-phase launch
  int rec_per_job = 0;          // record blocks provided per job (0: PHX_REC_BLOCKS_PER_JOB, 8); doubled
                                // after a launch whose rays needed more, and kept for the next ones
  int rec_per_job_env = 0;      // the PHX_REC_BLOCKS_PER_JOB value it was set from (a change resets it)
  int* counter = nullptr;             // [0]: jobs taken by the launch in flight
  unsigned long long* tot = nullptr;  // steps, plasma segments, dropped segments, FP64 steps
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int last_jobs = 0;
  void* ctl = nullptr;  // the allocation holding counter and tot (null when they live elsewhere)
  hipStream_t stream = nullptr;  // the stream its launches run on (null: the snapshot's)
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
  phx::DevBuf res, pix;  // per-render results and Morton pixel list (grow only)
  phx::DevBuf spec;      // the device flux of the last render (4 doubles per frequency)
  phx_model last_rq{};   // the last single render's request (its image, frequencies and units)
  phx::Slot slot;        // the launch buffers of the snapshot's own renders
  double decode_s = 0., upload_s = 0.;
  int64_t last_nres = 0;
  int pix_w = 0, pix_h = 0;  // image whose Morton pixel list `pix` holds (0: none)
  int device = 0;
  std::vector<void*> allocs;
  phx::Tune tune;  // launch configuration of batch renders (set by the autotuner)
  std::mutex mu;   // one render at a time per snapshot
};

namespace phx {

// the stream a slot's launches run on
static hipStream_t stream_of(const phx_snapshot* s, const Slot& sl) { return sl.stream ? sl.stream : s->stream; }

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
    if (b.fixed) throw Error(PHX_E_RANGE, "launch needs " + std::to_string(bytes) + " bytes of a " +
                                               std::to_string(b.cap) + "-byte preallocated buffer");
    if (b.p) HIPCHECK(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
    HIPCHECK(hipMalloc(&b.p, bytes));
    b.cap = bytes;
  }
  return b.p;
}

static void slot_open(Slot& sl) {
  HIPCHECK(hipEventCreate(&sl.e0));
  HIPCHECK(hipEventCreate(&sl.e1));
  HIPCHECK(hipMalloc(&sl.ctl, 512));
  sl.counter = static_cast<int*>(sl.ctl);
  sl.tot = reinterpret_cast<unsigned long long*>(static_cast<char*>(sl.ctl) + 256);
}
static void slot_free(Slot& sl) {
  for (DevBuf* b : {&sl.jobs, &sl.start, &sl.tasks, &sl.cams, &sl.side, &sl.fq, &sl.rec, &sl.rnext, &sl.rray})
    if (b->p && !b->fixed) (void)hipFree(b->p);
  if (sl.ctl) (void)hipFree(sl.ctl);
  if (sl.e0) (void)hipEventDestroy(sl.e0);
  if (sl.e1) (void)hipEventDestroy(sl.e1);
  if (sl.stream) (void)hipStreamDestroy(sl.stream);
  sl = Slot();
}

static void snapshot_free(phx_snapshot* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (void* p : s->allocs) (void)hipFree(p);
  for (DevBuf* b : {&s->res, &s->pix, &s->spec})
    if (b->p) (void)hipFree(b->p);
  slot_free(s->slot);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
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
  // the batched gathers address a cell table by 32-bit byte offsets
  if (batched_loads<T>() && nprim * sizeof(T) >= (size_t(1) << 32))
    throw Error(PHX_E_ARG, std::string("the ") + (sizeof(T) == 4 ? "FP32" : "FP64") +
                               " cell table of this snapshot exceeds 4 GiB");
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
  slot_open(s->slot);
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
  return grid_from<T>(s->geo, tb.blk, tb.meta, tb.prim, s->map, s->fd);
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
  if (const char* e = std::getenv("PHX_BLOCK"))
    if (std::atoi(e) == 64 || std::atoi(e) == 256) t.block = std::atoi(e);
  if (const char* e = std::getenv("PHX_FETCH"))
    if (std::atoi(e) >= 1) t.fetch = std::atoi(e);
  if (const char* e = std::getenv("PHX_REFILL"))
    t.refill = !std::strcmp(e, "wave") ? 64 : !std::strcmp(e, "lane") ? 1 : std::atoi(e);
  return t;
}

// what one march launch needs beyond the device buffers of the snapshot
// jobs of [j0, j1) of such a list that march a pixel (padding excluded)
static int real_jobs(int j0, int j1, int npix, int ntasks) {
  if (ntasks <= 1 || npix <= 0) return j1 - j0;
  int n = 0, t;
  for (int j = j0; j < j1; j++) n += wave_job(j, npix, ntasks, &t) >= 0;
  return n;
}

struct LaunchInfo {
  int njobs = 0, ntasks = 0, nfmax = 1;
  const phx_model* uni = nullptr;  // the request whose launch-wide parameters (stepping, cuts) apply
  bool no_alloc = false;           // the allocation-free entry point: no device allocation may happen
  bool capture = false;            // being recorded into a HIP graph: no timing events
  bool two_phase = false;          // march once recording samples, then one transfer lane per frequency
  int sweep = 0;                   // two-phase: the number of parameter points sharing the rays (0: none)
  int npix = 0;                    // a single request's pixels (the jobs are wave_job's list of npix x ntasks)
};

template <int M, int NF, bool Q, int BS>
static void launch_march(phx_snapshot* s, Slot& sl, const LaunchInfo& li, int j0, int j1, const Tune& tn,
                         phx_info* info) {
  using T = typename ModeOf<M>::T;
  MarchArgs<M> A;
  SideArgs<M> side;
  const Grid<double> g64 = grid_of<double>(s, tables_of<double>(s));
  const Loc<double> GL{g64.blk, g64.lo1, g64.lo2, g64.lo3, g64.hi1, g64.hi2, g64.hi3};
  LaunchBufs bufs{sl.tasks.p, (const int2*)sl.jobs.p, sl.start.p, sl.counter, sl.tot, ensure(sl.side, sizeof side)};
  if (li.two_phase) {
    bufs.rec = sl.rec.p;
    bufs.rec_next = static_cast<int*>(sl.rnext.p);
    bufs.rec_used = sl.counter + 16;
    bufs.rec_ovf = sl.counter + 17;
    bufs.rray = sl.rray.p;
    bufs.rec_cap = sl.rec_cap;
  }
  fill_args<M>(A, side, grid_of<T>(s, tables_of<T>(s)), GL, s->geo, li.uni, tn, bufs, j0, j1);
  k_store<SideArgs<M>><<<1, 1, 0, stream_of(s, sl)>>>(side, static_cast<SideArgs<M>*>(const_cast<void*>(bufs.side)));
  const void* kfn = reinterpret_cast<const void*>(&k_march<M, NF, Q, BS>);
  int cus = 0, per_cu = 0;
  HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device));
  HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, BS, 0));
  per_cu = std::max(per_cu, 1);
  if (tn.blocks_per_cu > 0) per_cu = std::min(per_cu, tn.blocks_per_cu);
  int grid = per_cu * cus;
  const int njobs = j1 - j0;
  grid = std::max(1, std::min(grid, (njobs + BS - 1) / BS));
  hipFuncAttributes fa;
  HIPCHECK(hipFuncGetAttributes(&fa, kfn));
  // a kernel whose private segment exceeds the per-lane stack limit does not run at all; raise the
  // limit to what the kernel needs (the reference sets it explicitly too, kernel.cu:305)
  size_t limit = 0;
  HIPCHECK(hipDeviceGetLimit(&limit, hipLimitStackSize));
  if (li.no_alloc && fa.localSizeBytes > 0)  // the runtime allocates private-segment memory at launch
    throw Error(PHX_E_RANGE, "the march kernel of this request (" + std::to_string(li.nfmax) +
                                 " frequencies per lane) uses " + std::to_string(fa.localSizeBytes) +
                                 " bytes of per-lane scratch, which the allocation-free render cannot use");
  if (fa.localSizeBytes > limit) {
    HIPCHECK(hipDeviceSetLimit(hipLimitStackSize, (fa.localSizeBytes + 255) & ~size_t(255)));
  }
  k_reset<<<1, 64, 0, stream_of(s, sl)>>>(sl.counter, nullptr);
  k_march<M, NF, Q, BS><<<grid, BS, 0, stream_of(s, sl)>>>(A);
  HIPCHECK(hipGetLastError());
  sl.last_jobs = njobs;
  if (info) {
    info->launches++;
    info->vgprs = fa.numRegs;
    info->scratch_bytes = (int32_t)fa.localSizeBytes;
    info->lds_bytes = (int32_t)fa.sharedSizeBytes;
    info->cus = cus;
    info->blocks = grid;
    info->block = BS;
    info->waves_per_cu = per_cu * BS / 64;
  }
}

template <int M, int NF>
static void launch_nf(phx_snapshot* s, Slot& sl, const LaunchInfo& li, int j0, int j1, const Tune& tn,
                      phx_info* info) {
  // default: samples in place with whole-wave refill (measured faster in every precision: rays of
  // one 8 x 8 pixel tile stay coherent; the LDS queue pays off only with per-lane refill)
  // the recording march of a two-phase render (NF = 0; a sweep's, NF = -1) runs in place with
  // 256-lane workgroups
  if constexpr (NF <= 0) {
    launch_march<M, NF, false, 256>(s, sl, li, j0, j1, tn, info);
  } else {
    if (tn.queue == 1)
      launch_march<M, NF, true, 256>(s, sl, li, j0, j1, tn, info);
    else if (tn.block == 64)
      launch_march<M, NF, false, 64>(s, sl, li, j0, j1, tn, info);
    else
      launch_march<M, NF, false, 256>(s, sl, li, j0, j1, tn, info);
  }
}

// Ray start records for the jobs in sl.jobs (cameras in sl.cams), then the march over them,
// in chunks of PHX_CHUNK jobs when that is set (the reference's chunking, kernel.cu:177-180).
template <int M>
static void march_jobs(phx_snapshot* s, Slot& sl, const LaunchInfo& li, const Tune& tn, phx_info* info,
                       phx_progress_fn progress, void* user) {
  using R = RecOf<M>;
  const int rec = M == MIX ? StHam::REC : StRef<typename ModeOf<M>::T>::REC;
  ensure(sl.start, (size_t)std::max(li.njobs, 1) * rec * sizeof(R));
  if (li.njobs > 0)
    k_rays<M><<<(li.njobs + BLOCK - 1) / BLOCK, BLOCK, 0, stream_of(s, sl)>>>((const Cam*)sl.cams.p,
                                                                      (const int2*)sl.jobs.p, li.njobs,
                                                                      sl.start.p);
  HIPCHECK(hipGetLastError());
  if (li.two_phase) {
    // The record pool (grow-only, at most PHX_REC_POOL_MB MiB, 48 GiB by default): RB-record blocks,
    // sl.rec_per_job a job on average (8 at first: 256 records).  The jobs are marched and
    // transferred in chunks of as many jobs as it then holds; a launch in which a ray found no block
    // left is run again with twice the blocks per job (record_overflow).
    using T = typename ModeOf<M>::T;
    const size_t blk = (size_t)RB * RECW * sizeof(T);
    const int per_job_env = std::max(1, env_int("PHX_REC_BLOCKS_PER_JOB", 8));
    if (sl.rec_per_job <= 0 || per_job_env != sl.rec_per_job_env) sl.rec_per_job = per_job_env;
    sl.rec_per_job_env = per_job_env;
    const int64_t per_job = sl.rec_per_job;
    // at most PHX_REC_POOL_MB MiB, and never less than one block per lane of a wave (0: that minimum)
    const int64_t limit = std::max<int64_t>(64, (std::max<int64_t>(0, env_int("PHX_REC_POOL_MB", 49152)) << 20) /
                                                    (int64_t)blk);
    const int64_t want = std::max<int64_t>(64, std::min<int64_t>((int64_t)li.njobs * per_job + 64, limit));
    ensure(sl.rec, (size_t)want * blk);
    // the pool is the part of the (grow-only) buffer within the limit
    const int64_t cap = std::min<int64_t>({(int64_t)(sl.rec.cap / blk), limit, INT32_MAX});
    ensure(sl.rnext, (size_t)cap * sizeof(int));
    ensure(sl.rray, (size_t)std::max(li.njobs, 1) * sizeof(int2));
    sl.rec_cap = (int)cap;
    // chunks of whole waves of 64 jobs, so that the rays that share a wave (and, in mixed
    // precision, its FP64 steps) are the ones of an unchunked launch
    const int chunk = (int)std::max<int64_t>(64, std::min<int64_t>(li.njobs, (cap - 64) / per_job / 64 * 64));
    sl.rec_chunk = chunk;
    const int F = std::max(1, li.nfmax);
    for (int j0 = 0; j0 < li.njobs; j0 += chunk) {
      const int j1 = std::min(li.njobs, j0 + chunk);
      if (j0 > 0) k_reset<<<1, 1, 0, stream_of(s, sl)>>>(sl.counter + 16, nullptr);  // the pool, for this chunk
      if (li.sweep > 0)
        launch_nf<M, -1>(s, sl, li, j0, j1, tn, info);
      else
        launch_nf<M, 0>(s, sl, li, j0, j1, tn, info);
      // the transfer, in launches of at most 2^30 lanes (a dispatch counts its work-items in 32 bits)
      const int64_t per_job = (int64_t)F * std::max(1, li.sweep);
      const int step = (int)std::max<int64_t>(1, ((int64_t)1 << 30) / per_job);
      for (int t0 = j0; t0 < j1; t0 += step) {
        const int t1 = std::min(j1, t0 + step);
        const unsigned blocks = (unsigned)(((int64_t)(t1 - t0) * per_job + 255) / 256);
        if (li.sweep > 0)
          k_transfer_sweep<M><<<blocks, 256, 0, stream_of(s, sl)>>>(
              static_cast<const TaskDev<T>*>(sl.tasks.p), li.sweep, (const int2*)sl.jobs.p, t0, t1, F,
              plasma_of<T>(s->geo, li.uni), static_cast<const T*>(sl.rec.p), static_cast<const int*>(sl.rnext.p),
              static_cast<const int2*>(sl.rray.p), sl.tot);
        else
          k_transfer<M><<<blocks, 256, 0, stream_of(s, sl)>>>(
              static_cast<const TaskDev<T>*>(sl.tasks.p), (const int2*)sl.jobs.p, t0, t1, F,
              static_cast<const T*>(sl.rec.p), static_cast<const int*>(sl.rnext.p), static_cast<const int2*>(sl.rray.p),
              sl.tot);
        HIPCHECK(hipGetLastError());
      }
      if (progress) {
        HIPCHECK(hipStreamSynchronize(stream_of(s, sl)));
        progress(j1 - j0, user);
      }
    }
    return;
  }
  int chunk = std::max(li.njobs, 1);
  if (const char* e = std::getenv("PHX_CHUNK"))
    if (std::atoi(e) > 0) chunk = std::atoi(e);
  for (int j0 = 0; j0 < li.njobs; j0 += chunk) {
    const int j1 = std::min(li.njobs, j0 + chunk);
    if (li.nfmax == 1)
      launch_nf<M, 1>(s, sl, li, j0, j1, tn, info);
    else if (li.nfmax == 2)
      launch_nf<M, 2>(s, sl, li, j0, j1, tn, info);
    else
      launch_nf<M, 4>(s, sl, li, j0, j1, tn, info);
    if (progress) {
      HIPCHECK(hipStreamSynchronize(stream_of(s, sl)));
      progress(real_jobs(j0, j1, li.npix, li.ntasks), user);
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
  if (rq->nfreqs < 1 || rq->nfreqs > PHX_FREQS_MAX) throw Error(PHX_E_ARG, "nfreqs must be in 1..32");
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

__global__ void k_expand(const int* __restrict__ pix, int npix, int task0, int ntasks, int2* __restrict__ jobs) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= wave_jobs(npix, ntasks)) return;
  int t;
  const int i = wave_job(n, npix, ntasks, &t);
  jobs[n] = make_int2(task0 + t, i < 0 ? -1 : pix[i]);
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

// The march of the jobs already in sl.jobs, queued on the snapshot's stream between the slot's
// timing events (returns at once unless a progress callback needs the chunks to finish).
template <int M>
static void enqueue_march(phx_snapshot* s, Slot& sl, const LaunchInfo& li, const Tune& tn, phx_info* info,
                          phx_progress_fn progress, void* user) {
  k_reset<<<1, 64, 0, stream_of(s, sl)>>>(sl.counter, sl.tot);
  if (!li.capture) HIPCHECK(hipEventRecord(sl.e0, stream_of(s, sl)));
  march_jobs<M>(s, sl, li, tn, info, progress, user);
  if (!li.capture) HIPCHECK(hipEventRecord(sl.e1, stream_of(s, sl)));
}
static void enqueue_precision(phx_snapshot* s, Slot& sl, int precision, const LaunchInfo& li, const Tune& tn,
                              phx_info* info, phx_progress_fn progress, void* user) {
  if (precision == PHX_FP32)
    enqueue_march<M32>(s, sl, li, tn, info, progress, user);
  else if (precision == PHX_FP64)
    enqueue_march<M64>(s, sl, li, tn, info, progress, user);
  else
    enqueue_march<MIX>(s, sl, li, tn, info, progress, user);
}
// Waits for the slot's launch and reads its time and statistics into info (when given).  The
// counter check makes a launch that stopped early an error: every lane takes jobs until the list
// is exhausted.
static double finish_march(phx_snapshot* s, Slot& sl, int njobs, phx_info* info) {
  HIPCHECK(hipEventSynchronize(sl.e1));
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, sl.e0, sl.e1));
  unsigned long long tot[4] = {0, 0, 0, 0};
  int taken = 0;
  HIPCHECK(hipMemcpy(tot, sl.tot, sizeof tot, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(&taken, sl.counter, sizeof taken, hipMemcpyDeviceToHost));
  if (njobs > 0 && taken < sl.last_jobs)
    throw Error(PHX_E_DEVICE, "march kernel did not run to completion (" + std::to_string(taken) + " of " +
                                  std::to_string(sl.last_jobs) + " jobs taken)");
  if (info) {
    info->kernel_s = ms * 1e-3;
    info->steps = (int64_t)tot[0];
    info->nplasma = (int64_t)tot[1];
    info->bad = (int64_t)tot[2];
    info->fp64_steps = (int64_t)tot[3];
    info->decode_s = s->decode_s;
    info->upload_s = s->upload_s;
  }
  return ms * 1e-3;
}
// After a two-phase launch: whether some ray found no record block left (it recorded nothing, and
// the launch must run again).  The slot then provides twice the blocks per job: a larger pool, up
// to its limit, and smaller chunks beyond it; the next launches of the slot keep the new figure.
static bool record_overflow(Slot& sl, const LaunchInfo& li) {
  if (!li.two_phase) return false;
  int ovf = 0;
  HIPCHECK(hipMemcpy(&ovf, sl.counter + 17, sizeof ovf, hipMemcpyDeviceToHost));
  if (ovf == 0) return false;
  // chunks are whole waves: a chunk of one wave that ran out cannot be split further
  if (sl.rec_chunk <= 64)
    throw Error(PHX_E_RANGE, "two-phase render: the records of one wave of 64 rays exceed the record pool "
                             "(PHX_REC_POOL_MB)");
  sl.rec_per_job = std::min(sl.rec_per_job * 2, 1 << 16);
  return true;
}
static void run_precision(phx_snapshot* s, int precision, const LaunchInfo& li, const Tune& tn, phx_info* info,
                          phx_progress_fn progress, void* user) {
  double t = 0.0;
  for (;;) {
    enqueue_precision(s, s->slot, precision, li, tn, info, progress, user);
    t += finish_march(s, s->slot, li.njobs, info);
    if (!record_overflow(s->slot, li)) break;
    if (info) info->reruns++;
  }
  if (info) info->kernel_s = t;  // the launches that ran out of records included
}

template <class T>
static void upload_tasks(phx_snapshot* s, Slot& sl, const std::vector<TaskDev<T>>& tasks,
                         const std::vector<Cam>& cams) {
  ensure(sl.tasks, tasks.size() * sizeof(TaskDev<T>));
  ensure(sl.cams, cams.size() * sizeof(Cam));
  HIPCHECK(hipMemcpyAsync(sl.tasks.p, tasks.data(), tasks.size() * sizeof(TaskDev<T>), hipMemcpyHostToDevice,
                          stream_of(s, sl)));
  HIPCHECK(hipMemcpyAsync(sl.cams.p, cams.data(), cams.size() * sizeof(Cam), hipMemcpyHostToDevice, stream_of(s, sl)));
}

// Whether a request renders in two phases: each ray marched once, its emitting samples recorded,
// then every frequency folded over the records by the transfer kernel, with the same results bit
// for bit.  A march with 4 frequencies per lane holds 104 accumulator registers per lane and
// spills, and above 4 every chunk of 4 marches the ray again: at 1024^2 on the c2 snapshot, mixed
// precision, the two phases take 75.7 / 78.7 / 81.0 / 96.7 ms for 2 / 3 / 4 / 8 frequencies against
// 76.9 / 121.0 / 128.4 / 229.1 ms in place.  PHX_TWO_PHASE_MIN_NF sets the threshold (0: never).
static int two_phase_min_nf() {
  const char* e = std::getenv("PHX_TWO_PHASE_MIN_NF");
  return e ? std::atoi(e) : 2;
}
static bool two_phase(const phx_model* rq) { return two_phase_min_nf() > 0 && rq->nfreqs >= two_phase_min_nf(); }

// the tasks of the requests of one launch and, for two-phase requests, their frequency tables
// (uploaded to sl.fq; the tasks point into it)
template <class T>
static void build_tasks(phx_snapshot* s, Slot& sl, const phx_model* const* rqs, const phx_freqout* const* outs, int n,
                        bool two, std::vector<TaskDev<T>>& tasks, std::vector<int>& first) {
  first.assign(n, 0);
  if (!two) {
    for (int r = 0; r < n; r++) {
      first[r] = (int)tasks.size();
      add_tasks<T>(rqs[r], const_cast<phx_freqout*>(outs[r]), tasks);
    }
    return;
  }
  std::vector<T> fq;
  std::vector<size_t> at(n);
  for (int r = 0; r < n; r++) {
    at[r] = fq.size();
    freq_table<T>(rqs[r], fq);
  }
  const T* d_fq = static_cast<const T*>(ensure(sl.fq, fq.size() * sizeof(T)));
  // a pageable source: the copy is staged before the call returns, so fq may go
  HIPCHECK(hipMemcpyAsync(sl.fq.p, fq.data(), fq.size() * sizeof(T), hipMemcpyHostToDevice, stream_of(s, sl)));
  for (int r = 0; r < n; r++) {
    first[r] = (int)tasks.size();
    add_task_all<T>(rqs[r], const_cast<phx_freqout*>(outs[r]), d_fq + at[r], tasks);
  }
}

// One request: its frequency chunks become tasks over the same pixel list (the full image in
// Morton order, cached on the device per image size, or the caller's jobs); a two-phase request
// is one task over the pixel list.
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
  const Cam cam = cam_of(s->geo, rq);
  const bool two = two_phase(rq);
  int ntasks = 0;
  std::vector<int> first;
  const phx_freqout* outs[1] = {d_res};
  if (precision == PHX_FP64) {
    std::vector<TaskDev<double>> tk;
    build_tasks<double>(s, s->slot, &rq, outs, 1, two, tk, first);
    ntasks = (int)tk.size();
    upload_tasks(s, s->slot, tk, std::vector<Cam>(ntasks, cam));
  } else {
    std::vector<TaskDev<float>> tk;
    build_tasks<float>(s, s->slot, &rq, outs, 1, two, tk, first);
    ntasks = (int)tk.size();
    upload_tasks(s, s->slot, tk, std::vector<Cam>(ntasks, cam));
  }
  const int64_t ntot = wave_jobs(njobs, ntasks);
  if ((int64_t)njobs * ntasks > INT32_MAX / 2 || ntot > INT32_MAX) throw Error(PHX_E_RANGE, "too many jobs in one render");
  ensure(s->slot.jobs, (size_t)std::max<int64_t>(ntot, 1) * sizeof(int2));
  if (ntot > 0)
    k_expand<<<(unsigned)((ntot + BLOCK - 1) / BLOCK), BLOCK, 0, s->stream>>>((const int*)s->pix.p, njobs, 0, ntasks,
                                                                              (int2*)s->slot.jobs.p);
  HIPCHECK(hipGetLastError());
  if ((int64_t)njobs < npix) HIPCHECK(hipMemsetAsync(d_res, 0, nres * sizeof(phx_freqout), s->stream));
  phx_info local;
  if (!info) info = &local;
  std::memset(info, 0, sizeof *info);
  fill_info_device(s, info);
  LaunchInfo li;
  li.njobs = (int)ntot;
  li.ntasks = ntasks;
  li.nfmax = two ? rq->nfreqs : std::min(CHUNK, rq->nfreqs) <= 1 ? 1 : std::min(CHUNK, rq->nfreqs) <= 2 ? 2 : 4;
  li.uni = rq;
  li.two_phase = two;
  li.npix = njobs;
  run_precision(s, precision, li, tn, info, progress, user);
  info->rays = (int64_t)njobs * rq->nfreqs;
  s->last_nres = (int64_t)nres;
  s->last_rq = *rq;
  if (res) HIPCHECK(hipMemcpy(res, d_res, nres * sizeof(phx_freqout), hipMemcpyDeviceToHost));
}


// Many requests in one launch: every request's frequency chunks become tasks, and the job list
// holds all their pixels (each request's in Morton order, requests one after another, or
// interleaved 64 jobs at a time).  Results go to one device block per request.  Requests that
// render in two phases (two_phase) form a launch of their own, after the others.
struct BatchPlan {
  std::vector<int64_t> off;  // result offset of each request in s->res (entries)
  int64_t nres = 0, rays = 0;
  int njobs = 0, ntasks = 0, nfmax = 1;
};
// the launch of the requests idx (results at offsets off of d_res): tasks, cameras, jobs
template <class T>
static BatchPlan plan_batch(phx_snapshot* s, const phx_model* const* rqs, const std::vector<int>& idx,
                            const std::vector<int64_t>& off, phx_freqout* d_res, bool two, int order,
                            int sample_stride) {
  BatchPlan bp;
  const int n = (int)idx.size();
  std::vector<const phx_model*> sub(n);
  std::vector<const phx_freqout*> outs(n);
  for (int q = 0; q < n; q++) {
    sub[q] = rqs[idx[q]];
    outs[q] = d_res + off[idx[q]];
  }
  std::vector<TaskDev<T>> tasks;
  std::vector<int> first;
  build_tasks<T>(s, s->slot, sub.data(), outs.data(), n, two, tasks, first);
  std::vector<Cam> cams;
  std::vector<std::vector<int2>> per(n);
  std::vector<int32_t> pix;
  for (int q = 0; q < n; q++) {
    const int t0 = first[q];
    const int nt = (q + 1 < n ? first[q + 1] : (int)tasks.size()) - t0;
    const Cam c = cam_of(s->geo, sub[q]);
    for (int t = 0; t < nt; t++) cams.push_back(c);
    bp.nfmax = std::max(bp.nfmax, two ? sub[q]->nfreqs : std::min(CHUNK, sub[q]->nfreqs));
    morton_pixels(sub[q], pix);
    std::vector<int2>& jl = per[q];
    // whole waves of 64 pixels, each once per frequency chunk (the order of wave_job); a partial
    // tile is padded to a whole wave with jobs of no pixel (-1), which end at once, so that every
    // request's waves start on a multiple of 64 and march the rays they march when it renders alone
    for (size_t i = 0; i < pix.size(); i += 64 * (size_t)sample_stride)
      for (int t = 0; t < nt; t++) {
        const size_t e = std::min(pix.size(), i + 64);
        for (size_t k = i; k < e; k++) jl.push_back(make_int2(t0 + t, pix[k]));
        for (size_t k = e; k < i + 64; k++) jl.push_back(make_int2(t0 + t, -1));
      }
    int64_t real = 0;
    for (size_t i = 0; i < pix.size(); i += 64 * (size_t)sample_stride) real += (int64_t)std::min<size_t>(64, pix.size() - i);
    bp.rays += real * sub[q]->nfreqs;
  }
  std::vector<int2> jobs;
  if (order == 1) {  // round-robin over requests, 64 jobs (one wave) at a time
    std::vector<size_t> pos(n, 0);
    for (bool any = true; any;) {
      any = false;
      for (int q = 0; q < n; q++)
        if (pos[q] < per[q].size()) {
          const size_t e = std::min(per[q].size(), pos[q] + 64);
          jobs.insert(jobs.end(), per[q].begin() + pos[q], per[q].begin() + e);
          pos[q] = e;
          any = true;
        }
    }
  } else {
    for (int q = 0; q < n; q++) jobs.insert(jobs.end(), per[q].begin(), per[q].end());
  }
  if (jobs.size() > (size_t)INT32_MAX) throw Error(PHX_E_RANGE, "too many jobs in one batch");
  bp.njobs = (int)jobs.size();
  bp.ntasks = (int)tasks.size();
  if (!two) bp.nfmax = bp.nfmax <= 1 ? 1 : bp.nfmax <= 2 ? 2 : 4;
  upload_tasks(s, s->slot, tasks, cams);
  ensure(s->slot.jobs, jobs.size() * sizeof(int2));
  HIPCHECK(hipMemcpyAsync(s->slot.jobs.p, jobs.data(), jobs.size() * sizeof(int2), hipMemcpyHostToDevice, s->stream));
  HIPCHECK(hipStreamSynchronize(s->stream));
  return bp;
}

// whether two requests look through the same camera at the same image (their rays are the same)
static bool same_camera(const phx_model* a, const phx_model* b) {
  return a->img_width == b->img_width && a->img_height == b->img_height && a->inclination == b->inclination &&
         a->fov_x == b->fov_x && a->fov_y == b->fov_y && a->r_cam == b->r_cam && a->t_init == b->t_init &&
         a->phi_cam == b->phi_cam && a->cos_pa == b->cos_pa && a->sin_pa == b->sin_pa;
}
// A parameter sweep: two or more requests through one camera (differing only in m_unit, R_high and
// frequencies).  Their rays are marched once, each sample recorded without plasma units, and every
// point's images are folded from the records (transfer_sweep).  PHX_SWEEP=0 renders them as a batch
// of separate images instead.
static bool is_sweep(const phx_model* const* rqs, int n) {
  if (n < 2 || env_int("PHX_SWEEP", 1) == 0) return false;
  for (int r = 1; r < n; r++)
    if (!same_camera(rqs[0], rqs[r])) return false;
  return true;
}
// the sweep's launch: one task per point (every frequency), one job per pixel of the common image
template <class T>
static BatchPlan plan_sweep(phx_snapshot* s, const phx_model* const* rqs, int n, const std::vector<int64_t>& off,
                            phx_freqout* d_res, int sample_stride) {
  BatchPlan bp;
  std::vector<const phx_freqout*> outs(n);
  for (int r = 0; r < n; r++) outs[r] = d_res + off[r];
  std::vector<TaskDev<T>> tasks;
  std::vector<int> first;
  build_tasks<T>(s, s->slot, rqs, outs.data(), n, true, tasks, first);
  const std::vector<Cam> cams(tasks.size(), cam_of(s->geo, rqs[0]));
  std::vector<int32_t> pix;
  morton_pixels(rqs[0], pix);
  std::vector<int2> jobs;
  for (size_t i = 0; i < pix.size(); i += 64 * (size_t)sample_stride)
    for (size_t k = i; k < std::min(pix.size(), i + 64); k++) jobs.push_back(make_int2(0, pix[k]));
  int nfsum = 0;
  for (int r = 0; r < n; r++) {
    bp.nfmax = std::max(bp.nfmax, rqs[r]->nfreqs);
    nfsum += rqs[r]->nfreqs;
  }
  bp.njobs = (int)jobs.size();
  bp.ntasks = (int)tasks.size();
  bp.rays = (int64_t)jobs.size() * nfsum;
  upload_tasks(s, s->slot, tasks, cams);
  ensure(s->slot.jobs, jobs.size() * sizeof(int2));
  HIPCHECK(hipMemcpyAsync(s->slot.jobs.p, jobs.data(), jobs.size() * sizeof(int2), hipMemcpyHostToDevice, s->stream));
  HIPCHECK(hipStreamSynchronize(s->stream));
  return bp;
}

static BatchPlan run_batch(phx_snapshot* s, const phx_model* const* rqs, int n, int precision, const Tune& tn,
                           int sample_stride, phx_info* info) {
  check_precision(s, precision);
  if (n < 1) throw Error(PHX_E_ARG, "empty batch");
  HIPCHECK(hipSetDevice(s->device));
  BatchPlan all;
  all.off.resize(n);
  std::vector<int> groups[2];  // chunked, two-phase
  for (int r = 0; r < n; r++) {
    check_request(rqs[r]);
    check_uniform(rqs[0], rqs[r]);
    all.off[r] = all.nres;
    all.nres += (int64_t)rqs[r]->img_width * rqs[r]->img_height * rqs[r]->nfreqs;
    groups[two_phase(rqs[r]) ? 1 : 0].push_back(r);
  }
  phx_freqout* d_res = (phx_freqout*)ensure(s->res, (size_t)all.nres * sizeof(phx_freqout));
  if (sample_stride > 1) HIPCHECK(hipMemsetAsync(d_res, 0, (size_t)all.nres * sizeof(phx_freqout), s->stream));
  std::memset(info, 0, sizeof *info);
  fill_info_device(s, info);
  if (is_sweep(rqs, n)) {
    const BatchPlan bp = precision == PHX_FP64 ? plan_sweep<double>(s, rqs, n, all.off, d_res, sample_stride)
                                               : plan_sweep<float>(s, rqs, n, all.off, d_res, sample_stride);
    LaunchInfo li;
    li.njobs = bp.njobs;
    li.ntasks = bp.ntasks;
    li.nfmax = bp.nfmax;
    li.uni = rqs[0];
    li.two_phase = true;
    li.sweep = n;
    run_precision(s, precision, li, tn, info, nullptr, nullptr);
    all.njobs = bp.njobs;
    all.ntasks = bp.ntasks;
    all.nfmax = bp.nfmax;
    all.rays = bp.rays;
    info->rays = all.rays;
    return all;
  }
  for (int g = 0; g < 2; g++) {
    if (groups[g].empty()) continue;
    const bool two = g == 1;
    const BatchPlan bp = precision == PHX_FP64
                             ? plan_batch<double>(s, rqs, groups[g], all.off, d_res, two, tn.order, sample_stride)
                             : plan_batch<float>(s, rqs, groups[g], all.off, d_res, two, tn.order, sample_stride);
    phx_info gi;
    std::memset(&gi, 0, sizeof gi);
    LaunchInfo li;
    li.njobs = bp.njobs;
    li.ntasks = bp.ntasks;
    li.nfmax = bp.nfmax;
    li.uni = rqs[0];
    li.two_phase = two;
    run_precision(s, precision, li, tn, &gi, nullptr, nullptr);
    // the launch figures of the first launch; times and counts summed
    if (all.njobs == 0) {
      info->cus = gi.cus;
      info->blocks = gi.blocks;
      info->block = gi.block;
      info->vgprs = gi.vgprs;
      info->sgprs = gi.sgprs;
      info->scratch_bytes = gi.scratch_bytes;
      info->lds_bytes = gi.lds_bytes;
      info->waves_per_cu = gi.waves_per_cu;
      info->upload_s = gi.upload_s;
      info->decode_s = gi.decode_s;
    }
    info->kernel_s += gi.kernel_s;
    info->steps += gi.steps;
    info->nplasma += gi.nplasma;
    info->bad += gi.bad;
    info->fp64_steps += gi.fp64_steps;
    info->launches += gi.launches;
    info->reruns += gi.reruns;
    all.njobs += bp.njobs;
    all.ntasks += bp.ntasks;
    all.nfmax = std::max(all.nfmax, bp.nfmax);
    all.rays += bp.rays;
  }
  info->rays = all.rays;
  return all;
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
  double rate_best = -1.;
  Tune tb;
  // in place (workgroups of 64 or 256 lanes) or through the LDS queue (256); the whole wave or each
  // lane refilled, each lane's refill at every pass or every 4th (steps per fetch); job order;
  // whole-wave FP64 promotion (mixed)
  struct Shape {
    int queue, block, refill, fetch;
  };
  std::vector<Shape> shapes;
  for (int q = 0; q <= 1; q++)
    for (int blk : {64, 256}) {
      if (q == 1 && blk != 256) continue;
      shapes.push_back({q, blk, 64, 1});
      shapes.push_back({q, blk, 1, 1});
      shapes.push_back({q, blk, 1, 4});
    }
  for (const Shape& sh : shapes)
    for (int order = 0; order <= (n > 1 ? 1 : 0); order++)
      for (int fw = 0; fw <= (precision == PHX_MIXED ? 1 : 0); fw++) {
        Tune t;
        t.queue = sh.queue;
        t.block = sh.block;
        t.refill = sh.refill;
        t.fetch = sh.fetch;
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
    best->block = tb.block;
    best->fetch = tb.fetch;
  }
  if (best_rate) *best_rate = rate_best;
}

// the flux of every frequency of the device result block of request rq into host spec[nfreqs][4]
static void device_flux(hipStream_t stream, const phx_freqout* d_res, const phx_model* rq, double* d_spec,
                        double* spec) {
  double dOm = 0., jy = 0.;
  flux_scale(rq, &dOm, &jy);
  k_flux<<<rq->nfreqs, BLOCK, 0, stream>>>(d_res, (long)rq->img_width * rq->img_height, rq->nfreqs, dOm, jy, d_spec);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(spec, d_spec, (size_t)rq->nfreqs * 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
  HIPCHECK(hipStreamSynchronize(stream));
}

// ============================================================================ allocation-free render
// For callers that cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid it): the per-render buffers (job list, ray start records, tasks,
// cameras, side arguments, counters) live in `scratch`, a fixed array in device memory that each
// GPU the library runs on holds from module load until the process ends; the caller passes the
// output.  Its device address comes from hipGetSymbolAddress, once per device, at the first call
// on that device (in host code the array's name is not a device address).  Calls on one device
// share the array; the library serializes them (they would otherwise race on it), and the
// counters and statistics are cleared by kernels at every call.
constexpr size_t NOALLOC_HEAD = size_t(64) << 10;       // counters, side arguments, tasks, cameras
constexpr int NOALLOC_MAX_JOBS = 1 << 20;                // FP32 / mixed; FP64 start records are twice as large
constexpr size_t NOALLOC_JOB_BYTES = sizeof(int2) + 96;  // a job and its FP32 / mixed start record
constexpr size_t NOALLOC_BYTES = NOALLOC_HEAD + (size_t)NOALLOC_MAX_JOBS * NOALLOC_JOB_BYTES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

struct NoAllocDev {
  std::mutex mu;        // one allocation-free render at a time on this device
  void* ptr = nullptr;  // device address of `scratch` on this device
};
static NoAllocDev g_noalloc[64];
static std::mutex g_noalloc_init;

static void* scratch_on(int device) {
  std::lock_guard<std::mutex> lk(g_noalloc_init);
  NoAllocDev& d = g_noalloc[device];
  if (!d.ptr) {
    HIPCHECK(hipSetDevice(device));
    void* p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(PHX_E_DEVICE, "no device address for the allocation-free render's array");
    d.ptr = p;
  }
  return d.ptr;
}

// what the setup kernel writes into the array: the request's tasks and cameras, and the job list
// (every pixel of the image for every task, in the Morton order of the reference's GPU schedule,
// jobs.hpp:18-46: pixels sorted by the key with i in the even bits and j in the odd bits)
template <class T>
struct TaskArray {
  TaskDev<T> t[PHX_FREQS_MAX / CHUNK];
};
template <class T>
struct NoAllocSetup {
  TaskArray<T> tasks;
  Cam cams[PHX_FREQS_MAX / CHUNK];
  int ntasks, npix, W, H, levels;  // levels: log2 of the power-of-two square that holds the image
  int write_tasks;                 // 0: the tasks are stored by a launch of their own (a captured render)
};
// the n-th pixel (i, j) of a W x H image in Morton order, by descending the quadtree of the
// enclosing 2^levels square: at each level the four children in key order (i bit low, j bit high)
// hold the pixels of the image that fall in them, counted in closed form
__device__ inline int2 morton_nth(unsigned n, int W, int H, int levels) {
  int i0 = 0, j0 = 0;
  for (int l = levels - 1; l >= 0; l--) {
    const int s = 1 << l;
    for (int c = 0; c < 4; c++) {
      const int ii = i0 + (c & 1) * s, jj = j0 + (c >> 1) * s;
      const unsigned cnt = (unsigned)max(0, min(W, ii + s) - ii) * (unsigned)max(0, min(H, jj + s) - jj);
      if (n < cnt) {
        i0 = ii;
        j0 = jj;
        break;
      }
      n -= cnt;
    }
  }
  return make_int2(i0, j0);
}
template <class T>
__global__ __launch_bounds__(BLOCK) void k_noalloc_setup(const NoAllocSetup<T> a, TaskDev<T>* __restrict__ tasks,
                                                         Cam* __restrict__ cams, int2* __restrict__ jobs) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < a.ntasks) {
    if (a.write_tasks) tasks[n] = a.tasks.t[n];
    cams[n] = a.cams[n];
  }
  const int njobs = a.npix * a.ntasks;
  for (int j = n; j < njobs; j += gridDim.x * blockDim.x) {
    const int t = j / a.npix, k = j - t * a.npix;  // one task: the pixels in order
    const int2 ij = morton_nth((unsigned)k, a.W, a.H, a.levels);
    jobs[j] = make_int2(t, ij.x * a.H + ij.y);
  }
}

// The array's layout for a request: counter and statistics, side arguments, tasks, cameras; then
// jobs and start records.  Returns the launch slot over those windows (its timing events are the
// snapshot's own); refuses a request beyond the array's job capacity for its precision.
static Slot noalloc_slot(phx_snapshot* s, const phx_model* rq, int precision, int64_t njobs) {
  const size_t rec = precision == PHX_FP64 ? StRef<double>::REC * sizeof(double)
                     : precision == PHX_MIXED ? StHam::REC * sizeof(double)
                                              : StRef<float>::REC * sizeof(float);
  const int64_t max_jobs = (int64_t)((NOALLOC_BYTES - NOALLOC_HEAD) / (sizeof(int2) + rec));
  if (njobs > max_jobs)
    throw Error(PHX_E_RANGE, "allocation-free render: " + std::to_string(njobs) + " jobs (pixels x frequency "
                                 "chunks of 4) above its maximum of " + std::to_string(max_jobs) + " for this precision");
  char* base = static_cast<char*>(scratch_on(s->device));
  Slot sl;
  sl.counter = reinterpret_cast<int*>(base);
  sl.tot = reinterpret_cast<unsigned long long*>(base + 256);
  auto window = [](DevBuf& b, char* p, size_t bytes) {
    b.p = p;
    b.cap = bytes;
    b.fixed = true;
  };
  window(sl.side, base + 1024, 8192);
  window(sl.tasks, base + 9216, 16384);
  window(sl.cams, base + 25600, 16384);
  window(sl.jobs, base + NOALLOC_HEAD, (size_t)max_jobs * sizeof(int2));
  window(sl.start, base + NOALLOC_HEAD + (size_t)max_jobs * sizeof(int2), (size_t)max_jobs * rec);
  sl.e0 = s->slot.e0;
  sl.e1 = s->slot.e1;
  static_assert(sizeof(TaskDev<double>) * (PHX_FREQS_MAX / CHUNK) <= 16384, "tasks window");
  static_assert(sizeof(Cam) * (PHX_FREQS_MAX / CHUNK) <= 16384, "cameras window");
  (void)rq;
  return sl;
}

// the setup kernel's arguments for request rq: its tasks (results into d_res) and cameras
template <class T>
static NoAllocSetup<T> noalloc_args(phx_snapshot* s, const phx_model* rq, phx_freqout* d_res) {
  NoAllocSetup<T> a;
  std::memset(&a, 0, sizeof a);
  std::vector<TaskDev<T>> tasks;
  const int ntasks = add_tasks<T>(rq, d_res, tasks);
  const Cam cam = cam_of(s->geo, rq);
  for (int t = 0; t < ntasks; t++) {
    a.tasks.t[t] = tasks[t];
    a.cams[t] = cam;
  }
  a.ntasks = ntasks;
  a.write_tasks = 1;
  a.npix = rq->img_width * rq->img_height;
  a.W = rq->img_width;
  a.H = rq->img_height;
  a.levels = 0;
  while ((1 << a.levels) < std::max(a.W, a.H)) a.levels++;
  return a;
}

// The allocation-free render's launch sequence on the snapshot's stream: the setup kernel (tasks,
// cameras, Morton job list), then the counter resets, camera rays, side arguments and march.  With
// capture set it is being recorded into a HIP graph: no timing events are recorded.
template <class T>
static LaunchInfo enqueue_noalloc(phx_snapshot* s, Slot& sl, const phx_model* rq, int precision,
                                  const NoAllocSetup<T>& a, bool capture, phx_info* info) {
  const int64_t njobs = (int64_t)a.npix * a.ntasks;
  const int grid = (int)std::min<int64_t>(4096, (njobs + BLOCK - 1) / BLOCK);
  k_noalloc_setup<T><<<std::max(grid, 1), BLOCK, 0, s->stream>>>(a, static_cast<TaskDev<T>*>(sl.tasks.p),
                                                                  static_cast<Cam*>(sl.cams.p),
                                                                  static_cast<int2*>(sl.jobs.p));
  HIPCHECK(hipGetLastError());
  LaunchInfo li;
  li.njobs = (int)njobs;
  li.ntasks = a.ntasks;
  li.nfmax = std::min(CHUNK, rq->nfreqs) <= 1 ? 1 : std::min(CHUNK, rq->nfreqs) <= 2 ? 2 : 4;
  li.uni = rq;
  li.no_alloc = true;
  li.capture = capture;
  Tune tn = s->tune;
  tn.queue = 0;  // samples in place: the LDS-queue kernels carry per-lane scratch with several frequencies
  enqueue_precision(s, sl, precision, li, tn, info, nullptr, nullptr);
  return li;
}

template <class T>
static void render_noalloc_t(phx_snapshot* s, const phx_model* rq, int precision, phx_freqout* d_res,
                             phx_info* info) {
  const int64_t npix = (int64_t)rq->img_width * rq->img_height;
  const int64_t njobs = npix * ((rq->nfreqs + CHUNK - 1) / CHUNK);
  Slot sl = noalloc_slot(s, rq, precision, njobs);
  const NoAllocSetup<T> a = noalloc_args<T>(s, rq, d_res);
  phx_info local;
  if (!info) info = &local;
  std::memset(info, 0, sizeof *info);
  fill_info_device(s, info);
  const LaunchInfo li = enqueue_noalloc<T>(s, sl, rq, precision, a, false, info);
  finish_march(s, sl, li.njobs, info);
  info->rays = npix * rq->nfreqs;
}

static void render_noalloc(phx_snapshot* s, const phx_model* rq, int precision, phx_freqout* d_res, phx_info* info) {
  check_precision(s, precision);
  check_request(rq);
  if (!d_res) throw Error(PHX_E_ARG, "d_res must be a device buffer of img_width * img_height * nfreqs results");
  if (s->device < 0 || s->device >= 64) throw Error(PHX_E_RANGE, "device index above 63");
  std::lock_guard<std::mutex> lk(s->mu);
  std::lock_guard<std::mutex> dk(g_noalloc[s->device].mu);
  HIPCHECK(hipSetDevice(s->device));
  if (precision == PHX_FP64)
    render_noalloc_t<double>(s, rq, precision, d_res, info);
  else
    render_noalloc_t<float>(s, rq, precision, d_res, info);
}

// ============================================================================ captured renders
// A render captured once into a HIP graph and replayed.  The graph records the allocation-free
// render's launch sequence (kernels only: no memset or host copy nodes, so every replay starts from
// the same state: the setup kernel rewrites cameras and jobs, and kernels clear the counters) and is
// instantiated once.  The parameter point (plasma units and frequencies, the request's tasks) is
// not part of the graph: each replay is one small launch that stores the point's tasks into the
// device array, then one hipGraphLaunch, between two timing events; the graph itself never
// changes.  Replays share the allocation-free render's device array and are serialized with its
// calls on the device.
struct GraphRender {
  phx_snapshot* s = nullptr;
  int precision = 0;
  phx_model rq{};  // the request of the last replay (camera, image and launch-wide fields fixed)
  phx_freqout* d_res = nullptr;
  Slot sl;
  LaunchInfo li;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  phx_info kinfo{};  // the march kernel's launch figures (registers, grid), noted at capture
};

// a replay's request may change only the plasma units and the frequency values
static void check_sweep_point(const phx_model* a, const phx_model* b) {
  check_uniform(a, b);
  const bool same = a->img_width == b->img_width && a->img_height == b->img_height && a->nfreqs == b->nfreqs &&
                    a->inclination == b->inclination && a->fov_x == b->fov_x && a->fov_y == b->fov_y &&
                    a->r_cam == b->r_cam && a->t_init == b->t_init && a->phi_cam == b->phi_cam &&
                    a->cos_pa == b->cos_pa && a->sin_pa == b->sin_pa;
  if (!same)
    throw Error(PHX_E_ARG, "a replay of a captured render may change only m_unit, R_high and the frequency values "
                           "(not the camera, the image size or the number of frequencies)");
}

template <class T>
static void graph_capture(GraphRender* g) {
  phx_snapshot* s = g->s;
  const int64_t njobs = (int64_t)g->rq.img_width * g->rq.img_height * ((g->rq.nfreqs + CHUNK - 1) / CHUNK);
  g->sl = noalloc_slot(s, &g->rq, g->precision, njobs);
  NoAllocSetup<T> a = noalloc_args<T>(s, &g->rq, g->d_res);
  a.write_tasks = 0;
  HIPCHECK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed));
  try {
    g->li = enqueue_noalloc<T>(s, g->sl, &g->rq, g->precision, a, true, &g->kinfo);
  } catch (...) {
    hipGraph_t dropped = nullptr;
    (void)hipStreamEndCapture(s->stream, &dropped);
    if (dropped) (void)hipGraphDestroy(dropped);
    throw;
  }
  g->li.uni = nullptr;  // pointed at the caller's request
  HIPCHECK(hipStreamEndCapture(s->stream, &g->graph));
  HIPCHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
}

// stores request rq's tasks (its plasma units and frequencies; results into d_res) where the
// graph's march reads them
template <class T>
static void graph_store_point(GraphRender* g, const phx_model* rq) {
  const NoAllocSetup<T> a = noalloc_args<T>(g->s, rq, g->d_res);
  k_store<TaskArray<T>><<<1, 1, 0, g->s->stream>>>(a.tasks, static_cast<TaskArray<T>*>(g->sl.tasks.p));
  HIPCHECK(hipGetLastError());
}

static void graph_free(GraphRender* g) {
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  g->exec = nullptr;
  g->graph = nullptr;
}

static void graph_create(GraphRender* g, phx_snapshot* s, const phx_model* rq, int precision, phx_freqout* d_res) {
  check_precision(s, precision);
  check_request(rq);
  if (!d_res) throw Error(PHX_E_ARG, "d_res must be a device buffer of img_width * img_height * nfreqs results");
  if (s->device < 0 || s->device >= 64) throw Error(PHX_E_RANGE, "device index above 63");
  g->s = s;
  g->precision = precision;
  g->rq = *rq;
  g->d_res = d_res;
  std::lock_guard<std::mutex> lk(s->mu);
  std::lock_guard<std::mutex> dk(g_noalloc[s->device].mu);
  HIPCHECK(hipSetDevice(s->device));
  try {
    if (precision == PHX_FP64)
      graph_capture<double>(g);
    else
      graph_capture<float>(g);
  } catch (...) {
    graph_free(g);
    throw;
  }
}

static void graph_launch(GraphRender* g, const phx_model* rq, phx_info* info) {
  phx_snapshot* s = g->s;
  if (rq) check_sweep_point(&g->rq, rq);
  std::lock_guard<std::mutex> lk(s->mu);
  std::lock_guard<std::mutex> dk(g_noalloc[s->device].mu);
  HIPCHECK(hipSetDevice(s->device));
  if (rq) g->rq = *rq;
  phx_info local;
  if (!info) info = &local;
  *info = g->kinfo;
  fill_info_device(s, info);
  HIPCHECK(hipEventRecord(g->sl.e0, s->stream));
  if (g->precision == PHX_FP64)
    graph_store_point<double>(g, &g->rq);
  else
    graph_store_point<float>(g, &g->rq);
  HIPCHECK(hipGraphLaunch(g->exec, s->stream));
  HIPCHECK(hipEventRecord(g->sl.e1, s->stream));
  finish_march(s, g->sl, g->li.njobs, info);
  info->rays = (int64_t)g->rq.img_width * g->rq.img_height * g->rq.nfreqs;
}

// ============================================================================ serving loop
// A worker thread keeps the snapshots resident and serves render requests with continuous
// batching.  A request is admitted when its result block fits the HBM budget (a block of the
// server's result pool, allocated once); its jobs (pixels x frequency chunks) join the active
// work of its snapshot, and every launch takes the next `quantum` jobs across the active requests
// of a snapshot, so requests that arrive while others are on the GPU share the following launches
// with them.  Each snapshot has two launch slots, each on a stream of its own: while one launch
// runs, the next is built and queued, so the GPU does not wait for the host between launches, and
// the next launch's waves take the compute units that the running launch's last rays leave idle
// (PHX_SERVE_STREAMS=0 queues both on the snapshot's stream instead).  A request is done when
// its last job has been marched; its result is copied to the caller's buffer on a copy stream of
// its own (concurrently with the running launch) and its latency (submit -> copied) recorded.
// Jobs are taken oldest request first (order 0), or least attained service first (order 1: the
// request with the fewest jobs marched or in flight, so short requests pass long ones at every
// launch boundary).  No counterpart in the reference (one image per process run,
// main.cpp:143-230).
struct Request {
  int64_t ticket = 0;
  int snap = 0;
  phx_model rq{};
  phx_freqout* host_res = nullptr;
  double* host_spec = nullptr;   // receives the flux of each frequency (reduced on the device), or null
  phx_freqout* d_res = nullptr;  // the request's block of the result pool
  size_t pool_off = 0;
  int64_t bytes = 0;
  std::vector<int2> jobs;  // (chunk, pixel), chunk-major; two-phase: (0, pixel)
  bool two = false;        // renders in two phases (two_phase)
  size_t next = 0, done = 0;
  double t_submit = 0., t_done = 0.;
  int state = 0;  // 0 waiting, 1 active, 2 done, 3 failed
  std::string error;
};

// First-fit allocator over one device allocation (256-byte granules, adjacent free blocks merged):
// the result blocks of admitted requests, with no hipMalloc / hipFree per request (a hipFree
// synchronizes the device, stalling the launches of every snapshot).
struct ResultPool {
  void* base = nullptr;
  size_t cap = 0;
  std::map<size_t, size_t> free_;  // offset -> size
  void open(size_t bytes) {
    cap = (bytes + 255) & ~size_t(255);
    HIPCHECK(hipMalloc(&base, cap));
    free_[0] = cap;
  }
  void close() {
    if (base) (void)hipFree(base);
    base = nullptr;
    free_.clear();
  }
  bool take(size_t bytes, size_t* off) {
    bytes = std::max<size_t>(256, (bytes + 255) & ~size_t(255));
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes, at = it->first + bytes;
        free_.erase(it);
        if (rest) free_[at] = rest;
        return true;
      }
    return false;
  }
  void give(size_t off, size_t bytes) {
    bytes = std::max<size_t>(256, (bytes + 255) & ~size_t(255));
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

// one queued launch of a snapshot: the jobs [a, b) of each request in it
struct Slice {
  std::shared_ptr<Request> r;
  size_t a, b;
};
struct Flight {
  int snap = 0, slot = 0, njobs = 0;
  bool two = false;  // a two-phase launch
  LaunchInfo li;     // what it was launched with (li.uni: the first slice's request)
  std::vector<Slice> slices;
};

struct Server {
  std::vector<phx_snapshot*> snaps;
  phx_server_opts opts{};
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  int waiters = 0;  // threads inside phx_server_wait
  int64_t next_ticket = 1;
  std::deque<std::shared_ptr<Request>> waiting;
  std::vector<std::shared_ptr<Request>> active;
  std::unordered_map<int64_t, std::shared_ptr<Request>> all;
  std::vector<double> latencies;
  std::vector<phx_model> uni;  // per snapshot: the launch-wide parameters of its first request
  std::vector<char> have_uni;
  int64_t submitted = 0, completed = 0, failed = 0, rays = 0, launches = 0, admitted_bytes = 0, peak_bytes = 0;
  double busy_s = 0., t_start = 0.;
  std::string fatal;  // a device error: every request after it fails, and submissions are refused
  // worker-owned
  std::vector<std::array<Slot, 2>> slots;
  std::vector<std::array<char, 2>> slot_busy;
  ResultPool pool;
  hipStream_t copy_stream = nullptr;
  double* d_spec = nullptr;    // the flux of a finished request (4 doubles per frequency)
  std::deque<Flight> flights;  // queued launches, oldest first
};

// waits for everything the server queued: every snapshot's stream and every launch slot's
static void sync_launches(Server* sv) {
  for (int k = 0; k < (int)sv->snaps.size(); k++) {
    (void)hipStreamSynchronize(sv->snaps[k]->stream);
    for (Slot& sl : sv->slots[k])
      if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  }
}

static void fail_request(Server* sv, Request* r, const std::string& why) {  // sv->mu held
  r->state = 3;
  r->error = why;
}

// The jobs a launch takes from a request with `left` jobs left, given `room` in the launch: whole
// waves of 64 (its remaining jobs when they fit).  Every slice but a request's last then starts and
// ends on a multiple of 64 in its own job list and in the launch's, and a launch ends after a last
// slice that is not whole waves: a wave marches the same 64 consecutive rays of a request as when
// the request is rendered alone, so the rays that share its FP64 steps (mixed precision) are the
// same, and so are the results.
static size_t wave_take_of(size_t left, size_t room) {
  if (left <= room) return left;
  return room / 64 * 64;
}

// builds and queues one launch of snapshot k on a free slot from the active requests' remaining
// jobs; false when the snapshot has no job left to launch or no free slot
template <class T>
static bool launch_tick(Server* sv, int k, const std::vector<std::shared_ptr<Request>>& act) {
  int si = sv->slot_busy[k][0] ? (sv->slot_busy[k][1] ? -1 : 1) : 0;
  if (si < 0) return false;
  std::vector<std::shared_ptr<Request>> order;
  for (auto& r : act)
    if (r->snap == k && r->state == 1 && r->next < r->jobs.size()) order.push_back(r);
  if (order.empty()) return false;
  if (sv->opts.order == 1)  // least attained service first (ties: arrival order)
    std::stable_sort(order.begin(), order.end(),
                     [](const std::shared_ptr<Request>& x, const std::shared_ptr<Request>& y) { return x->next < y->next; });
  const size_t quantum = (size_t)std::max(1, sv->opts.quantum_jobs);
  phx_snapshot* s = sv->snaps[k];
  Slot& sl = sv->slots[k][si];
  Flight fl;
  fl.snap = k;
  fl.slot = si;
  // Parameter-sweep points: the active requests that look through the first request's camera and
  // have reached the same job (points of a sweep submitted together) share one march of the next
  // quantum of its pixels; each point's results are folded from the records (transfer_sweep).
  // Every request's job list is one job per pixel here (one frequency in place, or two-phase),
  // so their pixels line up.
  {
    const std::shared_ptr<Request>& lead = order.front();
    // a request whose jobs are one per pixel: two-phase, or in place with one frequency chunk
    auto one_per_pixel = [](const Request& r) { return r.two || r.rq.nfreqs <= CHUNK; };
    std::vector<std::shared_ptr<Request>> group;
    if (env_int("PHX_SWEEP", 1) != 0 && one_per_pixel(*lead))
      for (auto& r : order)
        if (r->next == lead->next && one_per_pixel(*r) && same_camera(&lead->rq, &r->rq)) group.push_back(r);
    if (group.size() >= 2) {
      const size_t take = wave_take_of(lead->jobs.size() - lead->next, quantum);
      std::vector<const phx_model*> rqs;
      std::vector<const phx_freqout*> outs;
      int nfmax = 1;
      for (auto& r : group) {
        rqs.push_back(&r->rq);
        outs.push_back(r->d_res);
        nfmax = std::max(nfmax, r->rq.nfreqs);
      }
      std::unique_lock<std::mutex> snap_lock(s->mu);
      HIPCHECK(hipSetDevice(s->device));
      std::vector<TaskDev<T>> tasks;
      std::vector<int> first;
      build_tasks<T>(s, sl, rqs.data(), outs.data(), (int)rqs.size(), true, tasks, first);
      const std::vector<Cam> cams(tasks.size(), cam_of(s->geo, &lead->rq));
      std::vector<int2> jobs(take);
      for (size_t j = 0; j < take; j++) jobs[j] = make_int2(0, lead->jobs[lead->next + j].y);
      for (auto& r : group) {
        fl.slices.push_back({r, r->next, r->next + take});
        r->next += take;
      }
      fl.njobs = (int)take;
      fl.two = true;
      upload_tasks(s, sl, tasks, cams);
      ensure(sl.jobs, jobs.size() * sizeof(int2));
      HIPCHECK(hipMemcpyAsync(sl.jobs.p, jobs.data(), jobs.size() * sizeof(int2), hipMemcpyHostToDevice,
                              stream_of(s, sl)));
      LaunchInfo li;
      li.njobs = fl.njobs;
      li.ntasks = (int)tasks.size();
      li.nfmax = nfmax;
      li.uni = &lead->rq;
      li.two_phase = true;
      li.sweep = (int)group.size();
      fl.li = li;
      enqueue_precision(s, sl, sv->opts.precision, li, tune_from_env(s->tune), nullptr, nullptr, nullptr);
      sv->slot_busy[k][si] = 1;
      sv->flights.push_back(std::move(fl));
      return true;
    }
  }
  // one launch marches requests of one kind: those of the first request in the order (chunked
  // or two-phase), the others wait for a launch of their kind
  const bool two = order.front()->two;
  order.erase(std::remove_if(order.begin(), order.end(),
                             [&](const std::shared_ptr<Request>& r) { return r->two != two; }),
              order.end());
  std::vector<const phx_model*> rqs;
  std::vector<const phx_freqout*> outs;
  std::vector<size_t> takes;
  size_t njob = 0;
  for (auto& r : order) {
    if (njob >= quantum) break;
    const size_t left = r->jobs.size() - r->next, take = wave_take_of(left, quantum - njob);
    if (take == 0) break;
    rqs.push_back(&r->rq);
    outs.push_back(r->d_res);
    takes.push_back(take);
    njob += take;
    if (take % 64) break;  // a request's last slice ends the launch's whole waves
  }
  std::unique_lock<std::mutex> snap_lock(s->mu);  // the snapshot's own renders wait while a launch is queued
  HIPCHECK(hipSetDevice(s->device));
  std::vector<TaskDev<T>> tasks;
  std::vector<int> first;
  build_tasks<T>(s, sl, rqs.data(), outs.data(), (int)rqs.size(), two, tasks, first);
  std::vector<Cam> cams;
  std::vector<int2> jobs;
  int nfmax = 1;
  for (size_t q = 0; q < rqs.size(); q++) {
    const std::shared_ptr<Request>& r = order[q];
    const int t0 = first[q];
    const int nt = (q + 1 < rqs.size() ? first[q + 1] : (int)tasks.size()) - t0;
    const Cam c = cam_of(s->geo, &r->rq);
    for (int t = 0; t < nt; t++) cams.push_back(c);
    nfmax = std::max(nfmax, two ? r->rq.nfreqs : std::min(CHUNK, r->rq.nfreqs));
    for (size_t j = r->next; j < r->next + takes[q]; j++) jobs.push_back(make_int2(t0 + r->jobs[j].x, r->jobs[j].y));
    fl.slices.push_back({r, r->next, r->next + takes[q]});
    r->next += takes[q];
  }
  fl.njobs = (int)jobs.size();
  fl.two = two;
  upload_tasks(s, sl, tasks, cams);
  ensure(sl.jobs, jobs.size() * sizeof(int2));
  HIPCHECK(hipMemcpyAsync(sl.jobs.p, jobs.data(), jobs.size() * sizeof(int2), hipMemcpyHostToDevice, stream_of(s, sl)));
  LaunchInfo li;
  li.njobs = fl.njobs;
  li.ntasks = (int)tasks.size();
  li.nfmax = two ? nfmax : nfmax <= 1 ? 1 : nfmax <= 2 ? 2 : 4;
  li.uni = &fl.slices.front().r->rq;
  li.two_phase = two;
  fl.li = li;
  enqueue_precision(s, sl, sv->opts.precision, li, tune_from_env(s->tune), nullptr, nullptr, nullptr);
  // the host job list and tasks were staged by the copies above (pageable sources): they may go
  sv->slot_busy[k][si] = 1;
  sv->flights.push_back(std::move(fl));
  return true;
}

// a finished request: copy its result out, release its pool block, record it (sv->mu not held)
static void complete_request(Server* sv, const std::shared_ptr<Request>& r) {
  std::string copy_error;
  if (r->state == 1) {
    try {
      const size_t n = (size_t)r->rq.img_width * r->rq.img_height * r->rq.nfreqs;
      if (r->host_spec) device_flux(sv->copy_stream, r->d_res, &r->rq, sv->d_spec, r->host_spec);
      if (r->host_res) {
        HIPCHECK(hipMemcpyAsync(r->host_res, r->d_res, n * sizeof(phx_freqout), hipMemcpyDeviceToHost,
                                sv->copy_stream));
        HIPCHECK(hipStreamSynchronize(sv->copy_stream));
      }
    } catch (const std::exception& e) {
      copy_error = e.what();
    }
  }
  std::lock_guard<std::mutex> lk(sv->mu);
  if (r->d_res) sv->pool.give(r->pool_off, (size_t)r->bytes);
  r->d_res = nullptr;
  r->jobs.clear();
  r->jobs.shrink_to_fit();
  r->t_done = now_s();
  sv->admitted_bytes -= r->bytes;
  if (!copy_error.empty()) fail_request(sv, r.get(), copy_error);
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

static void serve_loop(Server* sv) {
  for (;;) {
    std::vector<std::shared_ptr<Request>> act, fresh;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_work.wait(lk, [&] {
        return sv->stop || !sv->waiting.empty() || !sv->active.empty() || !sv->flights.empty();
      });
      if (sv->stop) break;
      // admission by HBM budget, in arrival order (one request is always admitted when none is
      // active); the block comes from the result pool
      while (!sv->waiting.empty()) {
        std::shared_ptr<Request> r = sv->waiting.front();
        if (!sv->fatal.empty()) {
          fail_request(sv, r.get(), sv->fatal);
          r->t_done = now_s();
          sv->failed++;
          sv->waiting.pop_front();
          continue;
        }
        if (!sv->active.empty() && sv->admitted_bytes + r->bytes > sv->opts.hbm_budget) break;
        if (sv->opts.max_active > 0 && (int)sv->active.size() >= sv->opts.max_active) break;
        size_t off = 0;
        if (!sv->pool.take((size_t)r->bytes, &off)) break;  // fragmented: wait for a block to be given back
        r->pool_off = off;
        r->d_res = reinterpret_cast<phx_freqout*>(static_cast<char*>(sv->pool.base) + off);
        sv->admitted_bytes += r->bytes;
        sv->peak_bytes = std::max(sv->peak_bytes, sv->admitted_bytes);
        r->state = 1;
        sv->active.push_back(r);
        fresh.push_back(r);
        sv->waiting.pop_front();
      }
      act = sv->active;
    }
    sv->cv_done.notify_all();
    // job lists of newly admitted requests (outside the lock)
    for (auto& r : fresh) {
      std::vector<int32_t> pix;
      morton_pixels(&r->rq, pix);
      r->two = two_phase(&r->rq);
      const int nch = r->two ? 1 : (r->rq.nfreqs + CHUNK - 1) / CHUNK;
      r->jobs.resize((size_t)wave_jobs((int)pix.size(), nch));
      for (size_t n = 0; n < r->jobs.size(); n++) {  // whole waves of 64 pixels, each once per chunk
        int t;
        const int i = wave_job((int)n, (int)pix.size(), nch, &t);
        r->jobs[n] = make_int2(t, i < 0 ? -1 : pix[i]);
      }
    }
    try {
      // keep both slots of every snapshot with work queued
      for (int k = 0; k < (int)sv->snaps.size(); k++)
        for (int rep = 0; rep < 2; rep++) {
          const bool q = sv->opts.precision == PHX_FP64 ? launch_tick<double>(sv, k, act)
                                                        : launch_tick<float>(sv, k, act);
          if (!q) break;
        }
      if (sv->flights.empty()) continue;
      // the oldest queued launch: wait for it, then account its jobs
      Flight fl = std::move(sv->flights.front());
      sv->flights.pop_front();
      Slot& sl = sv->slots[fl.snap][fl.slot];
      phx_snapshot* fs = sv->snaps[fl.snap];
      double t = finish_march(fs, sl, fl.njobs, nullptr);
      // a two-phase launch whose record pool ran out runs again, at once, with more blocks per job
      while (record_overflow(sl, fl.li)) {
        std::unique_lock<std::mutex> snap_lock(fs->mu);
        HIPCHECK(hipSetDevice(fs->device));
        enqueue_precision(fs, sl, sv->opts.precision, fl.li, tune_from_env(fs->tune), nullptr, nullptr, nullptr);
        snap_lock.unlock();
        t += finish_march(fs, sl, fl.njobs, nullptr);
      }
      sv->slot_busy[fl.snap][fl.slot] = 0;
      {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->busy_s += t;
        sv->launches++;
      }
      for (const Slice& x : fl.slices) x.r->done += x.b - x.a;
      for (const Slice& x : fl.slices)
        if (x.r->state == 1 && x.r->done == x.r->jobs.size() && !x.r->jobs.empty() &&
            std::find(sv->active.begin(), sv->active.end(), x.r) != sv->active.end())
          complete_request(sv, x.r);
    } catch (const std::exception& e) {
      // a device error: drain what is queued, fail every admitted request, refuse new ones
      sync_launches(sv);
      sv->flights.clear();
      for (auto& b : sv->slot_busy) b[0] = b[1] = 0;
      std::vector<std::shared_ptr<Request>> failed;
      {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->fatal = e.what();
        for (auto& r : sv->active) {
          fail_request(sv, r.get(), e.what());
          failed.push_back(r);
        }
      }
      for (auto& r : failed) complete_request(sv, r);
    }
    sv->cv_done.notify_all();
  }
  // stopping: the queued launches finish before their buffers go away
  sync_launches(sv);
  sv->flights.clear();
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

int phx_snapshot_flux(phx_snapshot* s, double* spec, int32_t nfreqs) {
  return guarded([&] {
    if (!s || !spec) throw Error(PHX_E_ARG, "null snapshot or spectrum");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->last_nres == 0 || s->last_rq.nfreqs != nfreqs ||
        s->last_nres != (int64_t)s->last_rq.img_width * s->last_rq.img_height * nfreqs)
      throw Error(PHX_E_ARG, "no single render with this many frequencies to reduce");
    HIPCHECK(hipSetDevice(s->device));
    double* d_spec = static_cast<double*>(ensure(s->spec, PHX_FREQS_MAX * 4 * sizeof(double)));
    device_flux(s->stream, static_cast<const phx_freqout*>(s->res.p), &s->last_rq, d_spec, spec);
  });
}

int phx_snapshot_render_noalloc(phx_snapshot* s, const phx_model* rq, int32_t precision, phx_freqout* d_res,
                                phx_info* info) {
  return guarded([&] {
    if (!s) throw Error(PHX_E_ARG, "null snapshot");
    render_noalloc(s, rq, precision, d_res, info);
  });
}

}  // extern "C"

struct phx_graph {
  phx::GraphRender g;
};

extern "C" {

int phx_graph_create(phx_snapshot* s, const phx_model* rq, int32_t precision, phx_freqout* d_res, phx_graph** out) {
  return guarded([&] {
    if (!s || !out) throw Error(PHX_E_ARG, "null snapshot or output handle");
    *out = nullptr;
    std::unique_ptr<phx_graph> h(new phx_graph());
    graph_create(&h->g, s, rq, precision, d_res);
    *out = h.release();
  });
}

int phx_graph_launch(phx_graph* h, const phx_model* rq, phx_info* info) {
  return guarded([&] {
    if (!h) throw Error(PHX_E_ARG, "null graph");
    graph_launch(&h->g, rq, info);
  });
}

int phx_graph_destroy(phx_graph* h) {
  return guarded([&] {
    if (!h) return;
    {
      std::lock_guard<std::mutex> lk(h->g.s->mu);
      HIPCHECK(hipSetDevice(h->g.s->device));
      graph_free(&h->g);
    }
    delete h;
  });
}

int phx_snapshot_set_tune(phx_snapshot* s, const phx_tune* t) {
  return guarded([&] {
    if (!s || !t) throw Error(PHX_E_ARG, "null snapshot or tune");
    if (t->queue < -1 || t->queue > 1 || t->blocks_per_cu < 0 || t->order < 0 || t->order > 1 || t->fp64_wave < 0 ||
        t->fp64_wave > 1 || t->refill < 0 || t->refill > 64 || t->escape < 0 || t->escape > 2 ||
        (t->block != 0 && t->block != 64 && t->block != 256) || t->fetch < 0 || t->fetch > 1024)
      throw Error(PHX_E_ARG, "tune value out of range");
    s->tune.queue = t->queue;
    s->tune.blocks_per_cu = t->blocks_per_cu;
    s->tune.order = t->order;
    s->tune.fp64_wave = t->fp64_wave;
    s->tune.refill = t->refill > 0 ? t->refill : 64;
    s->tune.escape = t->escape;
    s->tune.block = t->block ? t->block : 256;
    s->tune.fetch = std::max(1, t->fetch);
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
    sv.opts.quantum_jobs = std::max(sv.opts.quantum_jobs, 64);  // at least one whole wave of jobs
    if (sv.opts.order < 0 || sv.opts.order > 1) throw Error(PHX_E_ARG, "order must be 0 (oldest first) or 1 (least attained service)");
    HIPCHECK(hipSetDevice(sv.snaps[0]->device));
    sv.slots.resize(nsnaps);
    sv.slot_busy.assign(nsnaps, std::array<char, 2>{0, 0});
    try {
      const char* ps = std::getenv("PHX_SERVE_STREAMS");
      const bool own_streams = !ps || std::atoi(ps) != 0;
      for (auto& pair : sv.slots)
        for (Slot& sl : pair) {
          slot_open(sl);
          // each launch slot on a stream of its own: the next launch's waves fill the compute units
          // that the running launch's last rays leave idle
          if (own_streams) HIPCHECK(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
        }
      sv.pool.open((size_t)sv.opts.hbm_budget);
      HIPCHECK(hipStreamCreateWithFlags(&sv.copy_stream, hipStreamNonBlocking));
      HIPCHECK(hipMalloc(&sv.d_spec, PHX_FREQS_MAX * 4 * sizeof(double)));
    } catch (...) {
      for (auto& pair : sv.slots)
        for (Slot& sl : pair) slot_free(sl);
      sv.pool.close();
      if (sv.copy_stream) (void)hipStreamDestroy(sv.copy_stream);
      if (sv.d_spec) (void)hipFree(sv.d_spec);
      throw;
    }
    sv.t_start = now_s();
    sv.worker = std::thread(serve_loop, &sv);
    *out = srv.release();
  });
}

int phx_server_submit_spec(phx_server* srv, int32_t snap, const phx_model* request, phx_freqout* res, double* spec,
                           int64_t* ticket) {
  return guarded([&] {
    if (!srv || !request) throw Error(PHX_E_ARG, "null server or request");
    Server& sv = srv->sv;
    if (snap < 0 || snap >= (int)sv.snaps.size()) throw Error(PHX_E_ARG, "no such snapshot");
    check_request(request);
    auto r = std::make_shared<Request>();
    r->snap = snap;
    r->rq = *request;
    r->host_res = res;
    r->host_spec = spec;
    r->bytes = (int64_t)request->img_width * request->img_height * request->nfreqs * (int64_t)sizeof(phx_freqout);
    if (r->bytes > sv.opts.hbm_budget) throw Error(PHX_E_RANGE, "request larger than the server's HBM budget");
    {
      std::lock_guard<std::mutex> lk(sv.mu);
      if (sv.stop) throw Error(PHX_E_ARG, "server stopped");
      if (!sv.fatal.empty()) throw Error(PHX_E_DEVICE, "server failed: " + sv.fatal);
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

int phx_server_submit(phx_server* srv, int32_t snap, const phx_model* request, phx_freqout* res, int64_t* ticket) {
  return phx_server_submit_spec(srv, snap, request, res, nullptr, ticket);
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
    sv.waiters++;
    if (timeout_s < 0)
      sv.cv_done.wait(lk, ready);
    else
      sv.cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
    sv.waiters--;
    if (sv.waiters == 0) sv.cv_done.notify_all();  // phx_server_stop waits for the last waiter to leave
    it = sv.all.find(ticket);
    if (done) *done = r->state >= 2;
    if (latency_s) *latency_s = r->state >= 2 ? r->t_done - r->t_submit : 0.;
    if (r->state == 3) throw Error(PHX_E_DEVICE, "request failed: " + r->error);
    if (r->state >= 2 && it != sv.all.end()) sv.all.erase(it);
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
    Server& sv = srv->sv;
    {
      std::lock_guard<std::mutex> lk(sv.mu);
      sv.stop = true;
    }
    sv.cv_work.notify_all();
    if (sv.worker.joinable()) sv.worker.join();  // returns after the queued launches have finished
    {
      // requests not completed fail with "server stopped"; waiters wake and leave before the
      // server (its mutex and condition variables) is deleted
      std::unique_lock<std::mutex> lk(sv.mu);
      for (auto& r : sv.active)
        if (r->state < 2) fail_request(&sv, r.get(), "server stopped");
      for (auto& r : sv.waiting)
        if (r->state < 2) fail_request(&sv, r.get(), "server stopped");
      sv.cv_done.notify_all();
      sv.cv_done.wait(lk, [&] { return sv.waiters == 0; });
    }
    for (auto& pair : sv.slots)
      for (Slot& sl : pair) slot_free(sl);
    sv.pool.close();
    if (sv.copy_stream) (void)hipStreamDestroy(sv.copy_stream);
    if (sv.d_spec) (void)hipFree(sv.d_spec);
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
    if (!s || !out) throw Error(PHX_E_ARG, "null snapshot or output");
    const int64_t want = (int64_t)s->geo.nleafs * s->geo.cells * 8;
    if (n != want) throw Error(PHX_E_ARG, "prims size mismatch");
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
