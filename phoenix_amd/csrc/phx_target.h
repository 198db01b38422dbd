// The few target primitives the march code uses, for its two compilations:
//
// * the library (hipcc, gfx950): device functions, global-address-space views of table
//   pointers, the hardware reciprocal, scheduling barriers, constant-memory tables, wave
//   ballots;
// * the CPU test build of the same per-ray code (clang++ -DPHX_HOST_BUILD, one ray per call,
//   run with the address and undefined-behaviour sanitizers by tests/cpu/): plain inline host
//   functions, plain pointers, 1/x, no barriers, const tables, and a one-lane "wave" whose
//   ballot is the lane's own predicate.
#pragma once

#ifdef PHX_HOST_BUILD
#include <cmath>
#include <cstddef>
#include <cstdint>

#define PHX_DEV inline
#define PHX_HD inline
#define PHX_CONSTANT static const
#define PHX_SCHED_BARRIER() ((void)0)

namespace phxd {
using std::acos;
using std::acosh;
using std::cbrt;
using std::copysign;
using std::cos;
using std::exp;
using std::expm1;
using std::fabs;
using std::floor;
using std::fma;
using std::fmin;
using std::fmod;
using std::isfinite;
using std::log;
using std::rint;
using std::sin;
using std::sqrt;
using std::tanh;
using std::trunc;
template <class T>
using gptr = const T*;
template <class T>
PHX_DEV gptr<T> glb(const T* p) {
  return p;
}
template <class T>
PHX_DEV gptr<T> at32(gptr<T> base, unsigned i) {
  return base + i;
}
PHX_DEV float hw_rcp(float x) { return 1.0f / x; }
PHX_DEV void sincosf(float x, float* s, float* c) {
  *s = std::sin(x);
  *c = std::cos(x);
}
PHX_DEV void sincos(double x, double* s, double* c) {
  *s = std::sin(x);
  *c = std::cos(x);
}
struct int2 {
  int x, y;
};
PHX_DEV int2 make_int2(int x, int y) { return int2{x, y}; }
template <class T>
using gmut = T*;
template <class T>
PHX_DEV gmut<T> glb_mut(T* p) {
  return p;
}
// one lane: the ballot of a predicate is the lane's own bit
PHX_DEV unsigned long long ballot(bool p) { return p ? 1ull : 0ull; }
// the next index of a counter for a lane that needs one, else -1
PHX_DEV int wave_take(bool need, int* counter) { return need ? (*counter)++ : -1; }
PHX_DEV void count_one(int* counter) { (*counter)++; }
// a pointer whose loads the device build keeps from hoisting (an opaque copy); the host has no such concern
template <class P>
PHX_DEV P opaque_scalar(P p) {
  return p;
}
template <class T>
PHX_DEV T load_constant(const T* p, int i) {
  return p[i];
}
}  // namespace phxd

#else
#include <hip/hip_runtime.h>

#define PHX_DEV __device__ __forceinline__
#define PHX_HD __host__ __device__ inline  // host and device (job lists)
#define PHX_CONSTANT __constant__ static const
#define PHX_SCHED_BARRIER() __builtin_amdgcn_sched_barrier(0)

namespace phxd {
// The global-memory view of a pointer.  The march reaches its tables through pointers that it
// loads from memory, whose address space the compiler cannot infer: loads through them are flat
// instructions, which also count against the scalar-load counter, so that every wait for a scalar
// load waits for them as well.  Loads through this view are global loads.
template <class T>
using gptr = const __attribute__((address_space(1))) T*;
template <class T>
PHX_DEV gptr<T> glb(const T* p) {
  return (gptr<T>)p;
}
// Element i of a table at a uniform base, by a 32-bit byte offset: the load then takes the base
// from scalar registers and the offset from one vector register, with no 64-bit address arithmetic
// per lane.  The table must be smaller than 4 GiB (the snapshot loader checks the cell table).
template <class T>
PHX_DEV gptr<T> at32(gptr<T> base, unsigned i) {
  using C = const __attribute__((address_space(1))) char*;
  return (gptr<T>)((C)base + i * (unsigned)sizeof(T));
}
// the same for stores (global stores instead of flat ones)
template <class T>
using gmut = __attribute__((address_space(1))) T*;
template <class T>
PHX_DEV gmut<T> glb_mut(T* p) {
  return (gmut<T>)p;
}
PHX_DEV float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
PHX_DEV unsigned long long ballot(bool p) { return __ballot(p); }
// The next index of a device counter for every active lane that needs one (consecutive indices in
// lane order), with one atomic per wave; -1 for the others.  May be called in divergent code: the
// ballot and the shuffle involve only the active lanes.
PHX_DEV int wave_take(bool need, int* counter) {
  const unsigned long long m = __ballot(need);
  if (m == 0ull) return -1;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, (int)__popcll(m));
  base = __shfl(base, leader);
  return need ? base + (int)__popcll(m & ((1ull << lane) - 1ull)) : -1;
}
PHX_DEV void count_one(int* counter) { atomicAdd(counter, 1); }
// A copy of a uniform pointer that the compiler treats as a new value (an empty asm on a scalar
// register): loads through it stay where they are written instead of being hoisted out of the
// march loop, which would hold their results in scalar registers across the whole loop.
template <class P>
PHX_DEV P opaque_scalar(P p) {
  asm volatile("" : "+s"(p));
  return p;
}
// element i of a constant-memory table, by a scalar load through the constant address space
template <class T>
PHX_DEV T load_constant(const T* p, int i) {
  using W = const __attribute__((address_space(4))) T*;
  return ((W)p)[i];
}
}  // namespace phxd
#endif
