// CPU build of the march's per-ray code (test infrastructure).
//
// phoenix_amd/csrc/phx_lane.h compiled as host C++ (PHX_HOST_BUILD): the same Lane (start, one
// pass of the reference's per-ray loop, write-out), ray_start, sample and transfer functions and
// launch arguments (phx_launch.h) that the gfx950 march kernel runs, driven here one lane at a
// time over a job list, with the lane's state (leaf cache included) carried from ray to ray as a
// kernel lane carries it.  Built by tests/cpu/Makefile, plain and under
// -fsanitize=address,undefined, once per memory form of the FP64 march (PHX_FP64_BATCHED).
//
// Usage: lane_host in.bin out.bin [two | sweep points.bin]
//   two: the two-phase render (recording march + transfer); sweep: the rays marched once without
//   plasma units and every point of points.bin (int32 K, K phx_model) folded from the records
//   in.bin: int32 {magic 0x50485831, sizeof(phx_model), precision, escape, npix, field_free}
//           phx_model bytes (pointer fields ignored)
//           f64 lb[nleafs*3], f64 dxc[nleafs*3], int32 level_map[L][level_dims[L] product] per level,
//           f64 prim[nleafs*cells*8] (the decoded primitives), int32 pix[npix]
//   out.bin: int64 {steps, nplasma, bad, fp64_steps}, f64 res[npix][nfreqs][6] (FreqOut)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "phx_launch.h"

namespace phx {
void set_error(const std::string&) {}
}  // namespace phx

using namespace phx;

namespace {

template <class V>
void read_into(FILE* f, V* p, size_t n) {
  if (std::fread(p, sizeof(V), n, f) != n) throw std::runtime_error("short read");
}

// per leaf: the largest rho if every cell is field-free with finite primitives and rho >= 0, else -1
template <class T>
std::vector<T> leaf_meta(const std::vector<T>& prim, int nleafs, int cells, bool mark) {
  std::vector<T> meta(nleafs, T(-1));
  if (!mark) return meta;
  for (int q = 0; q < nleafs; q++) {
    T mx = T(0);
    bool bad = false;
    for (int c = 0; c < cells; c++) {
      const T* v = &prim[((size_t)q * cells + c) * 8];
      for (int a = 0; a < 8; a++) bad |= !std::isfinite(v[a]);
      bad |= v[5] != T(0) || v[6] != T(0) || v[7] != T(0) || v[0] < T(0);
      mx = std::fmax(mx, v[0]);
    }
    meta[q] = bad ? T(-1) : mx;
  }
  return meta;
}

template <class T>
T* aligned_copy(const std::vector<T>& v) {
  const size_t bytes = ((v.size() * sizeof(T) + 255) / 256) * 256;
  T* p = static_cast<T*>(std::aligned_alloc(256, bytes ? bytes : 256));
  std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

struct Input {
  int precision = 0, escape = 1, npix = 0, field_free = 1;
  phx_model m{};
  std::vector<double> lb, dxc, prim;
  std::vector<std::vector<int32_t>> maps;
  std::vector<int32_t> pix;
};

Input read_input(const char* path) {
  Input in;
  FILE* f = std::fopen(path, "rb");
  if (!f) throw std::runtime_error("cannot open input");
  int32_t hdr[6];
  read_into(f, hdr, 6);
  if (hdr[0] != 0x50485831 || hdr[1] != (int32_t)sizeof(phx_model)) throw std::runtime_error("bad input header");
  in.precision = hdr[2];
  in.escape = hdr[3];
  in.npix = hdr[4];
  in.field_free = hdr[5];
  read_into(f, reinterpret_cast<char*>(&in.m), sizeof(phx_model));
  const phx_model& m = in.m;
  in.lb.resize((size_t)m.nleafs * 3);
  in.dxc.resize((size_t)m.nleafs * 3);
  read_into(f, in.lb.data(), in.lb.size());
  read_into(f, in.dxc.data(), in.dxc.size());
  in.maps.resize(m.nlevels);
  for (int L = 0; L < m.nlevels; L++) {
    in.maps[L].resize((size_t)m.level_dims[L][0] * m.level_dims[L][1] * m.level_dims[L][2]);
    read_into(f, in.maps[L].data(), in.maps[L].size());
  }
  in.prim.resize((size_t)m.nleafs * m.cells * 8);
  read_into(f, in.prim.data(), in.prim.size());
  in.pix.resize(in.npix);
  read_into(f, in.pix.data(), in.pix.size());
  std::fclose(f);
  in.m.lb = in.lb.data();
  in.m.dxc = in.dxc.data();
  in.m.prim = nullptr;
  for (int L = 0; L < PHX_MAX_LEVELS; L++) in.m.level_map[L] = L < m.nlevels ? in.maps[L].data() : nullptr;
  in.m.leaf_level = in.m.leaf_index = nullptr;
  return in;
}

// One lane marching the launch's jobs in order, as a kernel lane does between refills.
template <int M, int NF>
void march(const MarchArgs<M>& A, int njobs, unsigned long long tot[4]) {
  Lane<M, NF> ln;
  for (int n = 0; n < njobs; n++) {
    ln.start(side_of(A), n);
    for (;;) {
      PassView<typename ModeOf<M>::T> v = ln.look(A);
      if (v.emit) ln.sample_here(A, v);
      if (ln.advance(A, v) && ln.stats) tot[3]++;
      if (!v.alive) break;
    }
    ln.write_out(A, true);
    if (ln.stats) {
      tot[0] += ln.steps;
      tot[1] += ln.np;
    }
    tot[2] += ln.bad;
  }
}

// A two-phase render: the recording march of every ray (Lane<M, 0>), then the transfer of every
// (ray, frequency) over the records (transfer_one), as the device's two launches run them; a pool
// too small for the rays' records is grown and the render run again, as the device runtime does.
template <int M>
void render_two_phase(const MarchArgs<M>& A0, SideArgs<M>& side, const Input& in, std::vector<phx_freqout>& res,
                      std::vector<TaskDev<typename ModeOf<M>::T>>& tasks, const std::vector<int2>& jobs,
                      unsigned long long tot[4]) {
  using T = typename ModeOf<M>::T;
  const int njobs = (int)jobs.size();
  int cap = njobs * 6 + 64;
  if (const char* e = std::getenv("LANE_REC_BLOCKS")) cap = std::atoi(e);  // a small pool: the overflow path
  for (;;) {
    std::vector<T> rec((size_t)cap * RB * RECW);
    std::vector<int> next(cap, -1);
    std::vector<int2> rray(njobs);
    int used = 0, ovf = 0;
    side.rec = rec.data();
    side.rec_next = next.data();
    side.rec_used = &used;
    side.rec_ovf = &ovf;
    side.rray = rray.data();
    side.rec_cap = cap;
    *side.counter = 0;
    for (int q = 0; q < 4; q++) tot[q] = 0;
    march<M, 0>(A0, njobs, tot);
    if (ovf > 0) {
      cap = used + used / 4 + 64;
      continue;
    }
    for (int n = 0; n < njobs; n++)
      for (int f = 0; f < in.m.nfreqs; f++)
        tot[2] += transfer_one<M>(tasks.data(), jobs.data(), n, f, rec.data(), next.data(), rray.data());
    return;
  }
}

// A parameter sweep: the recording march of every ray without plasma units (Lane<M, -1>), then each
// point's transfer (transfer_sweep) into its own result array, as the device's sweep launches run them.
template <int M>
void render_sweep(const MarchArgs<M>& A0, SideArgs<M>& side, const Input& in, const std::vector<phx_model>& points,
                  std::vector<TaskDev<typename ModeOf<M>::T>>& tasks, const std::vector<int2>& jobs,
                  unsigned long long tot[4]) {
  using T = typename ModeOf<M>::T;
  const int njobs = (int)jobs.size(), K = (int)points.size();
  int cap = njobs * 8 + 64;
  for (;;) {
    std::vector<T> rec((size_t)cap * RB * RECW);
    std::vector<int> next(cap, -1);
    std::vector<int2> rray(njobs);
    int used = 0, ovf = 0;
    side.rec = rec.data();
    side.rec_next = next.data();
    side.rec_used = &used;
    side.rec_ovf = &ovf;
    side.rray = rray.data();
    side.rec_cap = cap;
    *side.counter = 0;
    for (int q = 0; q < 4; q++) tot[q] = 0;
    march<M, -1>(A0, njobs, tot);
    if (ovf > 0) {
      cap = 2 * cap;
      continue;
    }
    const Plasma<T> P = plasma_of<T>(in.m, &in.m);
    for (int n = 0; n < njobs; n++)
      for (int r = 0; r < K; r++)
        for (int f = 0; f < points[r].nfreqs; f++) {
          int np = 0;
          tot[2] += transfer_sweep<M>(tasks.data(), r, jobs.data(), n, f, P, rec.data(), next.data(), rray.data(), &np);
          tot[1] += np;
        }
    return;
  }
}

template <int M>
void render(const Input& in, std::vector<phx_freqout>& res, unsigned long long tot[4], bool two,
            const std::vector<phx_model>* points = nullptr, std::vector<std::vector<phx_freqout>>* outs = nullptr) {
  using T = typename ModeOf<M>::T;
  using R = RecOf<M>;
  const phx_model& m = in.m;
  // tables in the march's precision, and the FP64 block table the mixed march locates cells in
  const std::vector<T> rec = block_records<T>(&m);
  const std::vector<double> rec64 = block_records<double>(&m);
  int fd[3];
  const std::vector<int> map = flat_map(&m, fd);
  std::vector<T> prim(in.prim.size());
  for (size_t i = 0; i < prim.size(); i++) prim[i] = T(in.prim[i]);
  const std::vector<T> meta = leaf_meta<T>(prim, m.nleafs, m.cells, in.field_free != 0);
  T* d_rec = aligned_copy(rec);
  double* d_rec64 = aligned_copy(rec64);
  T* d_prim = aligned_copy(prim);
  T* d_meta = aligned_copy(meta);
  int* d_map = aligned_copy(map);
  const Grid<T> G = grid_from<T>(m, d_rec, d_meta, d_prim, d_map, fd);
  const Loc<double> GL{d_rec64, m.x_lo[1], m.x_lo[2], m.x_lo[3], m.x_hi[1], m.x_hi[2], m.x_hi[3]};
  // tasks over the pixel list, job = (task, pixel index in the output), the output indexed by pixel
  std::vector<TaskDev<T>> tasks;
  std::vector<T> fq;
  freq_table<T>(&m, fq);
  std::vector<std::vector<T>> fqs;
  int ntasks = 0;  // tasks whose jobs the march takes
  if (points) {    // one task per sweep point; the march takes the first point's jobs
    fqs.resize(points->size());
    for (size_t r = 0; r < points->size(); r++) {
      freq_table<T>(&(*points)[r], fqs[r]);
      add_task_all<T>(&(*points)[r], (*outs)[r].data(), fqs[r].data(), tasks);
    }
    ntasks = 1;
  } else {
    ntasks = two ? add_task_all<T>(&m, res.data(), fq.data(), tasks) : add_tasks<T>(&m, res.data(), tasks);
  }
  std::vector<int2> jobs;
  for (int t = 0; t < ntasks; t++)
    for (int p : in.pix) jobs.push_back(make_int2(t, p));
  const Cam cam = cam_of(m, &m);
  constexpr int REC = M == MIX ? StHam::REC : StRef<T>::REC;
  std::vector<R> start(jobs.size() * REC);
  for (size_t n = 0; n < jobs.size(); n++) ray_start<M>(cam, jobs[n].y, start.data() + n * REC);
  int counter = 0;
  SideArgs<M> side;
  MarchArgs<M> A;
  Tune tn;
  tn.escape = in.escape;
  const LaunchBufs bufs{tasks.data(), jobs.data(), start.data(), &counter, tot, &side};
  fill_args<M>(A, side, G, GL, m, &m, tn, bufs, 0, (int)jobs.size());
  const int nfmax = std::min(CHUNK, m.nfreqs);
  if (points)
    render_sweep<M>(A, side, in, *points, tasks, jobs, tot);
  else if (two)
    render_two_phase<M>(A, side, in, res, tasks, jobs, tot);
  else if (nfmax <= 1)
    march<M, 1>(A, (int)jobs.size(), tot);
  else if (nfmax <= 2)
    march<M, 2>(A, (int)jobs.size(), tot);
  else
    march<M, 4>(A, (int)jobs.size(), tot);
  std::free(d_rec);
  std::free(d_rec64);
  std::free(d_prim);
  std::free(d_meta);
  std::free(d_map);
}

}  // namespace

int main(int argc, char** argv) {
  const bool two = argc == 4 && std::strcmp(argv[3], "two") == 0;
  const bool sweep = argc == 5 && std::strcmp(argv[3], "sweep") == 0;
  if (argc != 3 && !two && !sweep) {
    std::fprintf(stderr, "usage: %s in.bin out.bin [two | sweep points.bin]\n", argv[0]);
    return 2;
  }
  try {
    const Input in = read_input(argv[1]);
    const phx_model& m = in.m;
    std::vector<phx_freqout> res((size_t)m.img_width * m.img_height * m.nfreqs);
    std::memset(res.data(), 0, res.size() * sizeof(phx_freqout));
    unsigned long long tot[4] = {0, 0, 0, 0};
    // sweep: points.bin = int32 K, then K phx_model (the points' units and frequencies; the same
    // camera, image and frequency count as in.bin's model)
    std::vector<phx_model> points;
    std::vector<std::vector<phx_freqout>> outs;
    if (sweep) {
      FILE* pf = std::fopen(argv[4], "rb");
      if (!pf) throw std::runtime_error("cannot open points");
      int32_t K = 0;
      read_into(pf, &K, 1);
      points.resize(K);
      read_into(pf, reinterpret_cast<char*>(points.data()), (size_t)K * sizeof(phx_model));
      std::fclose(pf);
      for (const phx_model& p : points)
        if (p.nfreqs != m.nfreqs) throw std::runtime_error("sweep points must share the frequency count");
      outs.assign(K, std::vector<phx_freqout>(res.size()));
    }
    if (in.precision == PHX_FP32)
      render<M32>(in, res, tot, two, sweep ? &points : nullptr, &outs);
    else if (in.precision == PHX_FP64)
      render<M64>(in, res, tot, two, sweep ? &points : nullptr, &outs);
    else
      render<MIX>(in, res, tot, two, sweep ? &points : nullptr, &outs);
    FILE* f = std::fopen(argv[2], "wb");
    if (!f) throw std::runtime_error("cannot write output");
    const long long st[4] = {(long long)tot[0], (long long)tot[1], (long long)tot[2], (long long)tot[3]};
    std::fwrite(st, sizeof st, 1, f);
    if (sweep) {
      for (const auto& o : outs)
        for (int32_t p : in.pix) std::fwrite(&o[(size_t)p * m.nfreqs], sizeof(phx_freqout), m.nfreqs, f);
    } else {
      for (int32_t p : in.pix) std::fwrite(&res[(size_t)p * m.nfreqs], sizeof(phx_freqout), m.nfreqs, f);
    }
    std::fclose(f);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "lane_host: %s\n", e.what());
    return 1;
  }
  return 0;
}
