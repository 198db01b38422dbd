// phoenix: command-line driver of libphx, the counterpart of the reference's src/main.cpp for the
// MI355X device path.  Same arguments (`phoenix [config.toml] [section.key=value ...]`), same
// flow (load config, grid and dump; schedule; device render; statistics; write_output; flux per
// frequency) and the same messages without the terminal colours and the progress bar.
//
// Beyond the reference:
//   PHX_ACCURACY=FP32|FP64|MIXED   arithmetic of the render (the reference's build option, run time)
//   --batch FILE                   one request per non-empty line of FILE (whitespace-separated
//                                  section.key=value overrides on top of the command line); all of
//                                  them are rendered by one persistent launch on the resident snapshot
//                                  and written to outdir/<line number>/
//   --serve FILE                   one request per non-empty line of FILE: an arrival time in seconds
//                                  after the start, then overrides as in --batch (a line may name another
//                                  grmhd.dump_file / grid_file: each dump is loaded once and kept resident);
//                                  the requests are submitted to the serving loop at their times, written
//                                  to outdir/<line number>/ as they complete, and their latencies reported.
//                                  PHX_SERVE_QUANTUM (jobs per launch), PHX_SERVE_ORDER (0 oldest first,
//                                  1 least attained service) and PHX_SERVE_BUDGET (bytes) tune the loop
#include <sys/stat.h>

#include <atomic>
#include <cstdarg>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>
#include <algorithm>

#include "phx.h"

namespace {

double now_s() {
  using clk = std::chrono::steady_clock;
  return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

void info(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void info(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vfprintf(stderr, fmt, ap);
  va_end(ap);
  std::fputc('\n', stderr);
}

int fail(const char* what) {
  std::fprintf(stderr, "FAIL: %s: %s\n", what, phx_last_error());
  return EXIT_FAILURE;
}

void count_progress(int32_t n, void* user) { static_cast<std::atomic<long>*>(user)->fetch_add(n); }

int precision_from_env() {
  const char* e = std::getenv("PHX_ACCURACY");
  if (!e || !std::strcmp(e, "FP32")) return PHX_FP32;
  if (!std::strcmp(e, "FP64")) return PHX_FP64;
  if (!std::strcmp(e, "MIXED")) return PHX_MIXED;
  std::fprintf(stderr, "FAIL: PHX_ACCURACY must be FP32, FP64 or MIXED\n");
  std::exit(EXIT_FAILURE);
}

const char* precision_name(int p) { return p == PHX_FP64 ? "FP64" : p == PHX_MIXED ? "MIXED" : "FP32"; }

// write_output of main.cpp:52-86 plus the flux lines
int write_all(const phx_model* m, const phx_config* cfg, const phx_freqout* res) {
  char log[8192];
  const int rc = phx_write_output(m, cfg, res, log, sizeof log);
  std::fputs(log, stderr);
  if (rc != PHX_OK) return fail("output");
  for (int f = 0; f < m->nfreqs; f++) {
    double S[4];
    if (phx_compute_spec(m, res, f, S) != PHX_OK) return fail("flux");
    info("Flux at %.4e Hz (Jy): I %+.9e Q %+.9e U %+.9e V %+.9e", m->frequencies[f], S[0], S[1], S[2], S[3]);
  }
  return EXIT_SUCCESS;
}

void print_run(const phx_model& m, int precision) {
  info("Precision: %s", precision_name(precision));
  info("Geometry:\n     a       = %g\n     hslope  = %g\n     Horizon = %g Rg", m.a, m.hslope, m.cut_in);
  info("Black Hole & Plasma:\n     Mass    = %.4e M_sun\n     Dist    = %.4e kpc\n     M_unit  = %.2e grams\n"
       "     r_low   = %g\n     r_high  = %g\n     beta_c  = %g",
       m.mass, m.distance, m.m_unit, m.r_low, m.r_high, m.beta_c);
  info("Camera:\n     Inclination = %g deg\n     Azimuth     = %g deg\n     Pos. Angle  = %g deg\n"
       "     Resolution  = %d x %d px\n     FOV         = %g x %g Rg\n     r_cam       = %g Rg\n     Frequencies = %d",
       m.inclination, m.phi_cam, m.pa, m.img_width, m.img_height, m.fov_x, m.fov_y, m.r_cam, m.nfreqs);
}

std::vector<std::vector<std::string>> read_batch(const char* path) {
  std::ifstream in(path);
  if (!in) {
    std::fprintf(stderr, "FAIL: cannot open batch file %s\n", path);
    std::exit(EXIT_FAILURE);
  }
  std::vector<std::vector<std::string>> out;
  std::string line;
  while (std::getline(in, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    std::istringstream ss(line);
    std::vector<std::string> ov;
    std::string tok;
    while (ss >> tok) ov.push_back(tok);
    if (!ov.empty()) out.push_back(ov);
  }
  return out;
}

// The user fields of request k (config + overrides) on the grid and dump of `base`
int request_on(const phx_model& base, const phx_model& mk, phx_model* out) {
  phx_model r = base;
  r.inclination = mk.inclination;
  r.fov_x = mk.fov_x;
  r.fov_y = mk.fov_y;
  r.r_cam = mk.r_cam;
  r.t_init = mk.t_init;
  r.pa = mk.pa;
  r.phi_cam = mk.phi_cam;
  r.img_width = mk.img_width;
  r.img_height = mk.img_height;
  r.nfreqs = mk.nfreqs;
  std::memcpy(r.frequencies, mk.frequencies, sizeof r.frequencies);
  r.m_unit = mk.m_unit;
  r.r_high = mk.r_high;
  if (phx_model_derive(&r) != PHX_OK) return fail("request");
  *out = r;
  return EXIT_SUCCESS;
}

long env_long(const char* name, long dflt) {
  const char* e = std::getenv(name);
  return e ? std::atol(e) : dflt;
}

// --serve: the requests of FILE submitted to the serving loop at their arrival times
int serve(const char* path, const std::string& cfgpath, const std::vector<const char*>& ov, const phx_config& cfg,
          int precision, double t_start) {
  const auto lines = read_batch(path);
  if (lines.empty()) {
    std::fprintf(stderr, "FAIL: empty serve file\n");
    return EXIT_FAILURE;
  }
  struct Snap {
    std::string dump;
    phx_model model;
    phx_snapshot* h = nullptr;
  };
  std::vector<Snap> snaps;
  struct Req {
    double at = 0.;
    int snap = 0;
    phx_model rq;
    phx_config cfg;
    std::vector<phx_freqout> res;
    int64_t ticket = 0;
  };
  std::vector<Req> reqs(lines.size());
  for (size_t k = 0; k < lines.size(); k++) {
    char* end = nullptr;
    reqs[k].at = std::strtod(lines[k][0].c_str(), &end);
    if (end == lines[k][0].c_str() || *end || reqs[k].at < 0) {
      std::fprintf(stderr, "FAIL: serve line %zu must start with an arrival time in seconds\n", k + 1);
      return EXIT_FAILURE;
    }
    std::vector<const char*> ovk(ov);
    for (size_t t = 1; t < lines[k].size(); t++) ovk.push_back(lines[k][t].c_str());
    phx_model mk;
    std::memset(&mk, 0, sizeof mk);
    if (phx_load_config(cfgpath.c_str(), ovk.data(), (int32_t)ovk.size(), &reqs[k].cfg, &mk) != PHX_OK)
      return fail("serve request");
    size_t si = 0;
    while (si < snaps.size() && snaps[si].dump != reqs[k].cfg.dump_file) si++;
    if (si == snaps.size()) {  // a dump not yet resident: grid, tables, device decode
      Snap sn;
      sn.dump = reqs[k].cfg.dump_file;
      sn.model = mk;
      info("Load GRMHD data       \"%s\" (resident snapshot %zu)", sn.dump.c_str(), snaps.size());
      if (phx_load_grid(reqs[k].cfg.grid_file, &sn.model) != PHX_OK) return fail("grid");
      if (phx_load_dump(&sn.model, sn.dump.c_str()) != PHX_OK) return fail("dump");
      if (phx_snapshot_open(&sn.model, sn.dump.c_str(), precision == PHX_FP64 ? PHX_TABLE_FP64 : PHX_TABLE_FP32,
                            &sn.h) != PHX_OK)
        return fail("snapshot");
      snaps.push_back(sn);
    }
    reqs[k].snap = (int)si;
    if (request_on(snaps[si].model, mk, &reqs[k].rq) != EXIT_SUCCESS) return EXIT_FAILURE;
    reqs[k].res.resize((size_t)reqs[k].rq.img_width * reqs[k].rq.img_height * reqs[k].rq.nfreqs);
    std::snprintf(reqs[k].cfg.outdir, sizeof reqs[k].cfg.outdir, "%s/%zu", cfg.outdir, k);
  }
  std::vector<phx_snapshot*> hs;
  for (const Snap& sn : snaps) hs.push_back(sn.h);
  phx_server_opts opts;
  std::memset(&opts, 0, sizeof opts);
  opts.precision = precision;
  opts.quantum_jobs = (int32_t)env_long("PHX_SERVE_QUANTUM", 0);
  opts.order = (int32_t)env_long("PHX_SERVE_ORDER", 0);
  opts.hbm_budget = (int64_t)env_long("PHX_SERVE_BUDGET", 0);
  phx_server* srv = nullptr;
  if (phx_server_start(hs.data(), (int32_t)hs.size(), &opts, &srv) != PHX_OK) return fail("server");
  std::vector<size_t> order(reqs.size());
  for (size_t k = 0; k < order.size(); k++) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return reqs[x].at < reqs[y].at; });
  const double t0 = now_s();
  for (size_t k : order) {
    const double dt = t0 + reqs[k].at - now_s();
    if (dt > 0) std::this_thread::sleep_for(std::chrono::duration<double>(dt));
    if (phx_server_submit(srv, reqs[k].snap, &reqs[k].rq, reqs[k].res.data(), &reqs[k].ticket) != PHX_OK)
      return fail("submit");
  }
  ::mkdir(cfg.outdir, 0755);
  int rc = EXIT_SUCCESS;
  for (size_t k = 0; k < reqs.size(); k++) {
    int32_t done = 0;
    double lat = 0.;
    if (phx_server_wait(srv, reqs[k].ticket, -1., &done, &lat) != PHX_OK) {
      rc = fail("request");
      continue;
    }
    info("Request %zu (snapshot %d, %d x %d x %d, arrived at %.3f s): latency %.3f s", k, reqs[k].snap,
         reqs[k].rq.img_width, reqs[k].rq.img_height, reqs[k].rq.nfreqs, reqs[k].at, lat);
    if (write_all(&reqs[k].rq, &reqs[k].cfg, reqs[k].res.data()) != EXIT_SUCCESS) rc = EXIT_FAILURE;
  }
  phx_serve_stats st;
  if (phx_server_stats(srv, &st) == PHX_OK)
    info("Served %lld requests in %.2f s: %.4g rays*freq/s, latency p50 %.3f s, p99 %.3f s, %lld launches",
         (long long)st.completed, st.wall_s, (double)st.rays / st.wall_s, st.p50_s, st.p99_s, (long long)st.launches);
  phx_server_stop(srv);
  for (Snap& sn : snaps) {
    phx_snapshot_close(sn.h);
    phx_free_model(&sn.model);
  }
  info("Execution: %.2f s", now_s() - t_start);
  return rc;
}

}  // namespace

int main(int argc, char* argv[]) {
  const double t_start = now_s();
  std::string cfgpath = "input/config.toml";
  const char* batch_file = nullptr;
  const char* serve_file = nullptr;
  std::vector<std::string> overrides;
  int a = 1;
  if (argc > 1 && !std::strchr(argv[1], '=')) {
    if (!std::strcmp(argv[1], "-h") || !std::strcmp(argv[1], "--help")) {
      info("usage: %s [input/config.toml] [section.key=value ...] [--batch FILE | --serve FILE]\n"
           "       PHX_ACCURACY=FP32|FP64|MIXED selects the arithmetic (default FP32)",
           argv[0]);
      return 0;
    }
    if (std::strcmp(argv[1], "--batch") && std::strcmp(argv[1], "--serve")) {
      cfgpath = argv[1];
      a = 2;
    }
  }
  for (; a < argc; a++) {
    if (!std::strcmp(argv[a], "--batch") && a + 1 < argc) {
      batch_file = argv[++a];
      continue;
    }
    if (!std::strcmp(argv[a], "--serve") && a + 1 < argc) {
      serve_file = argv[++a];
      continue;
    }
    overrides.push_back(argv[a]);
  }
  const int precision = precision_from_env();
  phx_dev_warmup();  // HIP context on a background thread while the config and the dump are read

  std::vector<const char*> ov;
  for (const std::string& s : overrides) ov.push_back(s.c_str());
  phx_config cfg;
  phx_model model;
  std::memset(&cfg, 0, sizeof cfg);
  std::memset(&model, 0, sizeof model);
  if (phx_load_config(cfgpath.c_str(), ov.data(), (int32_t)ov.size(), &cfg, &model) != PHX_OK) return fail("config");
  if (serve_file) return serve(serve_file, cfgpath, ov, cfg, precision, t_start);
  info("Input Filepaths");
  info("Load config file      \"%s\"", cfgpath.c_str());
  info("Load grid information \"%s\"", cfg.grid_file);
  if (phx_load_grid(cfg.grid_file, &model) != PHX_OK) return fail("grid");
  info("Load GRMHD data       \"%s\"", cfg.dump_file);
  const double t_model = now_s();
  if (phx_load_dump(&model, cfg.dump_file) != PHX_OK) return fail("dump");
  info(" Loaded in %.2f s   (decoding on the GPU)", now_s() - t_model);
  print_run(model, precision);

  if (batch_file) {
    // ---- batch: every line is a request on the same resident snapshot
    const auto lines = read_batch(batch_file);
    if (lines.empty()) {
      std::fprintf(stderr, "FAIL: empty batch file\n");
      return EXIT_FAILURE;
    }
    phx_snapshot* snap = nullptr;
    if (phx_snapshot_open(&model, cfg.dump_file, precision == PHX_FP64 ? PHX_TABLE_FP64 : PHX_TABLE_FP32, &snap) !=
        PHX_OK)
      return fail("snapshot");
    std::vector<phx_model> reqs(lines.size());
    std::vector<phx_config> cfgs(lines.size());
    std::vector<std::vector<phx_freqout>> res(lines.size());
    for (size_t k = 0; k < lines.size(); k++) {
      std::vector<const char*> ovk(ov);
      for (const std::string& s : lines[k]) ovk.push_back(s.c_str());
      phx_model mk;
      std::memset(&mk, 0, sizeof mk);
      if (phx_load_config(cfgpath.c_str(), ovk.data(), (int32_t)ovk.size(), &cfgs[k], &mk) != PHX_OK)
        return fail("batch request");
      // the request's user fields on the loaded grid and dump
      phx_model r;
      if (request_on(model, mk, &r) != EXIT_SUCCESS) return EXIT_FAILURE;
      reqs[k] = r;
      res[k].resize((size_t)r.img_width * r.img_height * r.nfreqs);
      std::snprintf(cfgs[k].outdir, sizeof cfgs[k].outdir, "%s/%zu", cfg.outdir, k);
    }
    std::vector<const phx_model*> rp;
    std::vector<phx_freqout*> op;
    for (size_t k = 0; k < reqs.size(); k++) {
      rp.push_back(&reqs[k]);
      op.push_back(res[k].data());
    }
    const double t_ray = now_s();
    phx_info gi;
    if (phx_snapshot_render_batch(snap, rp.data(), (int32_t)rp.size(), precision, op.data(), &gi) != PHX_OK)
      return fail("batch render");
    info("Raymarch (batch of %zu): %.2f s, kernel %.2f s, %.4g rays*freq/s", reqs.size(), now_s() - t_ray,
         gi.kernel_s, (double)gi.rays / gi.kernel_s);
    ::mkdir(cfg.outdir, 0755);
    for (size_t k = 0; k < reqs.size(); k++)
      if (write_all(&reqs[k], &cfgs[k], res[k].data()) != EXIT_SUCCESS) return EXIT_FAILURE;
    phx_snapshot_close(snap);
    phx_free_model(&model);
    info("Execution: %.2f s", now_s() - t_start);
    return EXIT_SUCCESS;
  }

  // ---- one image (main.cpp:143-230)
  const size_t npix = (size_t)model.img_width * model.img_height;
  std::vector<phx_freqout> res(npix * model.nfreqs);
  std::vector<int32_t> jobs(npix);
  if (phx_schedule(&model, 1, jobs.data(), (int32_t)npix) != PHX_OK) return fail("schedule");
  std::atomic<long> ndone(0);
  phx_info gi;
  const double t_ray = now_s();
  if (phx_dev_render(&model, cfg.dump_file, precision, jobs.data(), (int32_t)npix, res.data(), &gi, count_progress,
                     &ndone) != PHX_OK) {
    fail("render");
    phx_free_model(&model);
    return EXIT_FAILURE;
  }
  info("Raymarch: %.2f s", now_s() - t_ray);
  info(" Model upload %.2f s, GPU decode: %.2f s, kernel %.2f s", gi.upload_s, gi.decode_s, gi.kernel_s);
  long badpix = 0;
  for (size_t p = 0; p < npix; p++) {
    int nonfinite = 0;
    for (int f = 0; f < model.nfreqs; f++)
      for (int q = 0; q < 4; q++) nonfinite |= !std::isfinite(res[p * model.nfreqs + f].IQUV[q]);
    badpix += nonfinite;
  }
  info("%.0f geodesic samples/pixel, %.0f plasma segments/pixel (mean)", (double)gi.steps / npix,
       (double)gi.nplasma / npix);
  info("Non-finite segment operators: %lld, pixels with non-finite IQUV: %ld", (long long)gi.bad, badpix);
  info("Device: %s, %d CUs", gi.device, gi.cus);
  info("Launch: %d blocks x %d threads, %d VGPRs/lane, %d B scratch/lane", gi.blocks, gi.block, gi.vgprs,
       gi.scratch_bytes);
  info("Create output directory \"%s\"", cfg.outdir);
  const double t_out = now_s();
  if (write_all(&model, &cfg, res.data()) != EXIT_SUCCESS) {
    phx_free_model(&model);
    return EXIT_FAILURE;
  }
  info("Output: %.2f s", now_s() - t_out);
  phx_free_model(&model);
  info("Execution: %.2f s", now_s() - t_start);
  return EXIT_SUCCESS;
}
