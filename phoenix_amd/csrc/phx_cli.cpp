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
#include <vector>

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

}  // namespace

int main(int argc, char* argv[]) {
  const double t_start = now_s();
  std::string cfgpath = "input/config.toml";
  const char* batch_file = nullptr;
  std::vector<std::string> overrides;
  int a = 1;
  if (argc > 1 && !std::strchr(argv[1], '=')) {
    if (!std::strcmp(argv[1], "-h") || !std::strcmp(argv[1], "--help")) {
      info("usage: %s [input/config.toml] [section.key=value ...] [--batch FILE]\n"
           "       PHX_ACCURACY=FP32|FP64|MIXED selects the arithmetic (default FP32)",
           argv[0]);
      return 0;
    }
    if (std::strcmp(argv[1], "--batch")) {
      cfgpath = argv[1];
      a = 2;
    }
  }
  for (; a < argc; a++) {
    if (!std::strcmp(argv[a], "--batch") && a + 1 < argc) {
      batch_file = argv[++a];
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
      phx_model r = model;
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
      if (phx_model_derive(&r) != PHX_OK) return fail("batch request");
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
