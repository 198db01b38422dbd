// ============================================================================
// HAMILTONIAN TWIN — test infrastructure only.  A CPU statement of the
// algorithm the product's mixed-precision march runs (the null geodesic in
// Hamiltonian form, the polarization angle from the Walker-Penrose constant),
// on the oracle's model, sampler, transfer and accumulation, which restate the
// reference (phx_oracle.cpp).  It is used by tests/ to show, at FP64 and away
// from the GPU, that this formulation reproduces the reference's FP64 build on
// the golden cases, and by tests/golden/count_ops.py to count the formulation's
// operations for the roofline.  Nothing in the product (phoenix_amd/) links,
// loads or calls this file.
//
// Differences from the reference's ray loop (src/core/render/pixel.hpp:134-224):
//  * state (x1, x2, x3, p_1, p_2) with p_t, p_phi constant, integrated by the
//    same classical RK4 (geodesic.hpp:222-256) under the same step rule
//    (geodesic.hpp:267-292); k^mu = g^{mu nu} p_nu wherever the reference uses
//    its integrated k^mu (turning points, step rule, the sample);
//  * no transported F1 (geodesic.hpp:180-195): at a sample, 2 psi follows from
//    the Walker-Penrose constant of the camera's (k, F1) instead of projecting
//    F1 on the plasma-frame basis (tetrad.hpp:160-182);
//  * t is not tracked (nothing on the path depends on it).
// ============================================================================
#include "phx_oracle.cpp"

namespace orc {

// Hamilton's equations of H = g^{mu nu} p_mu p_nu / 2 with the contravariant
// MKS-BHAC Kerr-Schild metric (metric.hpp:64-91): 2H = -p_t^2 + Q / Sigma,
// Q = -2 r p_t^2 + 4 p_t p_1 + (Delta/r^2) p_1^2 + (2a/r) p_1 p_phi + p_2^2/theta'^2
//     + p_phi^2 / sin^2 theta.
// y = (x1, x2, x3, p_1, p_2); d = dy/dlambda; k = k^mu.
template <class T>
static void ham_rhs(T a, T hs, T pt, T pp, const T y[5], T d[5], T k[4]) {
  const T r = exp(y[0]);
  T s2x, c2x, s, c;
  sincos_(T(2) * y[1], &s2x, &c2x);
  sincos_(y[1] + T(0.5) * hs * s2x, &s, &c);
  const T tp = T(1) + hs * c2x, tpp = T(-2) * hs * s2x;
  const T a2 = a * a;
  const T S = T(1) / (r * r + a2 * c * c);
  const T ir = T(1) / r, is = T(1) / s, itp = T(1) / tp;
  const T is2 = is * is, itp2 = itp * itp;
  const T p1 = y[3], p2 = y[4];
  const T dr2 = T(1) - T(2) * ir + a2 * ir * ir;
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
  if (k) {
    k[0] = -pt - T(2) * S * (r * pt - p1);
    k[1] = d[0];
    k[2] = d[1];
    k[3] = d[2];
  }
}
template <class T>
static void ham_rk4(T a, T hs, T pt, T pp, T y[5], T dt) {
  T k1[5], k2[5], k3[5], k4[5], t[5];
  ham_rhs<T>(a, hs, pt, pp, y, k1, nullptr);
  for (int i = 0; i < 5; i++) t[i] = y[i] + T(0.5) * dt * k1[i];
  ham_rhs<T>(a, hs, pt, pp, t, k2, nullptr);
  for (int i = 0; i < 5; i++) t[i] = y[i] + T(0.5) * dt * k2[i];
  ham_rhs<T>(a, hs, pt, pp, t, k3, nullptr);
  for (int i = 0; i < 5; i++) t[i] = y[i] + dt * k3[i];
  ham_rhs<T>(a, hs, pt, pp, t, k4, nullptr);
  for (int i = 0; i < 5; i++) y[i] += dt / T(6) * (k1[i] + T(2) * (k2[i] + k3[i]) + k4[i]);
}
// Walker-Penrose constant kappa = (A - iB)(r - i a cos theta) of (k, f) at X, from the
// Kerr-Schild components W^{mu nu} = k^mu f^nu - k^nu f^mu in (t, r, theta, phi):
// A = W^{tr} + a sin^2 W^{r phi}, B = sin [(r^2 + a^2) W^{phi theta} - a (W^{t theta} + W^{r theta})]
template <class T>
static void wp_kappa(T a, T hs, const T X[4], const T k[4], const T f[4], T* q1, T* q2) {
  const T r = exp(X[1]);
  T s2x, c2x, s, c;
  sincos_(T(2) * X[2], &s2x, &c2x);
  sincos_(X[2] + T(0.5) * hs * s2x, &s, &c);
  const T tp = T(1) + hs * c2x;
  const T kr = r * k[1], kth = tp * k[2], fr = r * f[1], fth = tp * f[2];
  const T Wtr = k[0] * fr - kr * f[0], Wrp = kr * f[3] - k[3] * fr, Wpth = k[3] * fth - kth * f[3];
  const T Wtth = k[0] * fth - kth * f[0], Wrth = kr * fth - kth * fr;
  const T A = Wtr + a * s * s * Wrp;
  const T B = s * ((r * r + a * a) * Wpth - a * (Wtth + Wrth));
  *q1 = A * r - B * a * c;
  *q2 = -(A * a * c + B * r);
}

// process_sample (pixel.hpp:52-123) with 2 psi from the Walker-Penrose constant kap
template <class T>
static void sample_wp(const Model<T>* m, RayState<T>* ps, const Pt<T>* sm, int next_leaf, const T kap[2]) {
  const int prev = tl_phase;
  tl_phase = PH_SAMPLE;
  T X[4], k[4];
  for (int q = 0; q < 4; q++) {
    X[q] = sm->x[q];
    k[q] = sm->k[q];
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
      T e1r, e1i, e2r, e2i;
      wp_kappa(m->a, m->hslope, fl.X, k, perp, &e1r, &e1i);
      wp_kappa(m->a, m->hslope, fl.X, k, para, &e2r, &e2i);
      // kap = f1 kappa(perp) + f2 kappa(para): (f1, f2) up to a common factor
      const T f1 = kap[0] * e2i - e2r * kap[1], f2 = e1r * kap[1] - kap[0] * e1i;
      const T n = f1 * f1 + f2 * f2;
      c2 = (f1 * f1 - f2 * f2) / n;
      s2 = T(2) * f1 * f2 / n;
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

// one ray, camera to termination, in Hamiltonian form (pixel.hpp:134-224 otherwise)
template <class T>
static Stats trace_ham(const Model<T>* m, double alpha, double beta, double* out) {
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
  T X0[4], k0[4], f0[4], g0[4][4], p[4];
  for (int i = 0; i < 4; i++) {
    X0[i] = T(ph[i]);
    k0[i] = T(ph[4 + i]);
    f0[i] = T(F1[i]);
  }
  gcov<T>(m->a, m->hslope, X0, g0);
  lower(g0, k0, p);
  T kap[2];
  wp_kappa(m->a, m->hslope, X0, k0, f0, &kap[0], &kap[1]);
  const T pt = p[0], pp = p[3];
  T y[5] = {X0[1], X0[2], X0[3], p[1], p[2]};
  ps.nplasma = ps.bad = 0;
  int steps = 0, turns = 0, stop = 0;
  T thdot_prev = T(0), dl_prev = T(0);
  const bool bneg = beta < 0., bpos = beta > 0.;
  Pt<T> pend;
  pend.i = -1;
  for (;;) {
    tl_phase = PH_STEP;
    Pt<T> cur;
    cur.x[0] = T(0);
    cur.x[1] = y[0];
    cur.x[2] = y[1];
    cur.x[3] = y[2];
    T d[5];
    ham_rhs<T>(m->a, m->hslope, pt, pp, y, d, cur.k);
    for (int i = 0; i < 4; i++) cur.f[i] = T(0);
    cur.i = steps;
    cur.dl = dl_prev;
    const T r = exp(cur.x[1]);
    const bool alive = r < m->r_out && r > m->cut_in && steps < m->up.max_steps && !stop;
    if (pend.i >= 1) {
      const int nl = alive ? leaf_at(m, cur.x) : -1;
      sample_wp(m, &ps, &pend, nl, kap);
    }
    pend = cur;
    if (!alive) break;
    if (thdot_prev * cur.k[2] < T(0) && steps > 2) turns++;
    thdot_prev = cur.k[2];
    if ((bneg && turns > m->up.max_order) || (bpos && turns > (m->up.max_order + 1))) stop = 1;
    const T* dxc = cell_size_at(m, cur.x);
    const T dl = step_len(m->stepsize, cur.x, cur.k, dxc);
    ham_rk4<T>(m->a, m->hslope, pt, pp, y, dl);
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

}  // namespace orc

namespace {
template <class T>
int render_ham_any(Handle* h, const long* pix, long npix, double* res, long long* stats, int nthreads) {
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
    const Stats st = trace_ham(m, al, be, res + (size_t)n * nf * 6);
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
// the Hamiltonian twin on a handle loaded at precision 64 or 164 (counted double)
int orc_render_ham(void* hv, const long* pix, long npix, double* res, long long* stats, int nthreads) {
  Handle* h = (Handle*)hv;
  switch (h->precision) {
    case 64: return render_ham_any<double>(h, pix, npix, res, stats, nthreads);
    case 164: return render_ham_any<Counted<double>>(h, pix, npix, res, stats, 1);
  }
  return -1;
}
// the Walker-Penrose constant of (k, f) at X, for the tests' closed-form checks
void orc_wp_kappa(double a, double hs, const double* X, const double* k, const double* f, double* q) {
  wp_kappa<double>(a, hs, X, k, f, &q[0], &q[1]);
}
// the Hamiltonian right-hand side at y, for the tests' finite-difference checks
void orc_ham_rhs(double a, double hs, double pt, double pp, const double* y, double* d, double* k) {
  ham_rhs<double>(a, hs, pt, pp, y, d, k);
}
}  // extern "C"
