// Test-infrastructure harness around the reference's own CPU path.
//
// Compiled (oracle/Makefile) against the reference's header-only sources where
// they lie; output goes to oracle/_ref/ only.  It does what the reference's
// main() does for the CPU build up to the render (src/main.cpp:110-205):
//   host::config::load_config          (src/host/io/config.hpp:229)
//   host::bhac_io::load_bhac_grid      (src/host/io/bhac_io.hpp:143)
//   host::bhac_io::load, gpu_decode=0  (src/host/io/bhac_io.hpp:331)
//   per pixel: pixel_impact + PA rotation + core::pixel::render_pixel,
//   OpenMP schedule(dynamic,1)         (src/host/render/jobs.hpp:50-92)
// and writes the raw FreqOut array instead of HDF5 (src/host/io/output.hpp
// needs hdf5.h, which is not part of the path).
//
// Usage: phx_ref_fpXX config.toml out.bin [section.key=value ...]
//        phx_ref_fpXX config.toml --config-only [section.key=value ...]   (parsed keys or "ERROR msg")
//        phx_ref_fpXX config.toml out.bin --pixels first,count[,stride] ...
//        (pixel subset first + k*stride, k < count)
// out.bin: int32 {W, H, nf, 0} int64 {steps, nplasma, bad} f64 {raymarch_s}
//          then W*H*nf*6 float64 (IQUV, tau, tauF), res[(i*H + j)*nf + f].
#include <omp.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "core/render/headers.hpp"
#include "host/io/bhac_io.hpp"
#include "host/io/config.hpp"

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s config.toml out.bin [--pixels first,count] [section.key=value ...]\n", argv[0]);
    return 2;
  }
  const std::string cfg = argv[1], outp = argv[2];
  std::vector<std::string> ov;
  long first = 0, count = -1, stride = 1;
  for (int a = 3; a < argc; a++) {
    if (!std::strcmp(argv[a], "--pixels") && a + 1 < argc) {
      std::sscanf(argv[++a], "%ld,%ld,%ld", &first, &count, &stride);
      continue;
    }
    ov.push_back(argv[a]);
  }
  core::bhac::Model model = {};
  host::config::Params params;
  if (outp == "--config-only") {  // host::config::load_config alone: the parsed keys, or the error
    try {
      host::config::load_config(cfg, ov, params, &model.uparams);
    } catch (const std::exception& e) {
      std::printf("ERROR %s\n", e.what());
      return 1;
    }
    const core::bhac::UserParams& u = model.uparams;
    std::printf("inclination %.17g\nimg_width %d\nimg_height %d\nfov_x %.17g\nfov_y %.17g\nr_cam %.17g\n",
                u.inclination, u.img_width, u.img_height, u.fov_x, u.fov_y, u.r_cam);
    std::printf("t_init %.17g\npa %.17g\nphi_cam %.17g\nnfreqs %d\n", u.t_init, u.pa, u.phi_cam, u.nfreqs);
    for (int f = 0; f < u.nfreqs; f++) std::printf("frequency%d %.17g\n", f, u.frequencies[f]);
    std::printf("mass %.17g\ndistance %.17g\nstepsize %.17g\nmax_steps %d\nhorizon_margin %.17g\nmax_order %d\n",
                u.mass, u.distance, (double)u.stepsize, u.max_steps, u.horizon_margin, u.max_order);
    std::printf("m_unit %.17g\nr_low %.17g\nr_high %.17g\nbeta_c %.17g\nsigma_cut %.17g\n", u.m_unit, u.r_low,
                u.r_high, u.beta_c, (double)u.sigma_cut);
    std::printf("thetae_min %.17g\nthetae_max %.17g\nrt_outer %.17g\n", (double)u.thetae_min,
                (double)u.thetae_max, (double)u.rt_outer);
    std::printf("nspin %d\nbpol %d\nnlevels %d\nsfc %d\n", u.nspin, u.bpol, u.nlevels, (int)u.sfc);
    std::printf("outdir %s\nimage %s\nblock_size %d\ndump_file %s\ngrid_file %s\ngpu_decode %d\nshow_header %d\n",
                params.outdir.c_str(), params.image.c_str(), params.block_size, params.dump_file.c_str(),
                params.grid_file.c_str(), (int)params.gpu_decode, (int)params.show_header);
    return 0;
  }
  try {
    host::config::load_config(cfg, ov, params, &model.uparams);
    host::bhac_io::load_bhac_grid(&model.gparams, params.grid_file);
    host::bhac_io::load(&model, params.dump_file, false);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "ref: %s\n", e.what());
    return 1;
  }
  const core::bhac::UserParams& up = model.uparams;
  const long npix = (long)up.img_width * up.img_height;
  if (stride < 1) stride = 1;
  if (count < 0 || first + (count - 1) * stride >= npix) count = (npix - first + stride - 1) / stride;
  std::vector<core::pixel::FreqOut> res((size_t)npix * up.nfreqs);
  std::memset(res.data(), 0, res.size() * sizeof(res[0]));
  long steps = 0, segs = 0, bad = 0;
  const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : steps, segs, bad)
  for (long kk = 0; kk < count; kk++) {
    const long n = first + kk * stride;
    const int i = (int)(n / up.img_height), j = (int)(n % up.img_height);
    double a0, b0;
    core::camera::pixel_impact(i, j, up.img_width, up.img_height, up.fov_x, up.fov_y, &a0, &b0);
    const double alpha = a0 * model.cos_pa + b0 * model.sin_pa;
    const double beta = -a0 * model.sin_pa + b0 * model.cos_pa;
    const core::pixel::RunStats st = core::pixel::render_pixel(&model, alpha, beta, &res[(size_t)n * up.nfreqs]);
    steps += st.steps;
    segs += st.nplasma;
    bad += st.bad;
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::fprintf(stderr, "Raymarch: %.3f s (%d threads), %ld steps, %ld segments, %ld bad, %ld pixels\n", secs,
               omp_get_max_threads(), steps, segs, bad, count);
  FILE* f = std::fopen(outp.c_str(), "wb");
  if (!f) {
    std::fprintf(stderr, "ref: cannot write %s\n", outp.c_str());
    return 1;
  }
  const int hdr[4] = {up.img_width, up.img_height, up.nfreqs, 0};
  const long long st3[3] = {steps, segs, bad};
  std::fwrite(hdr, sizeof hdr, 1, f);
  std::fwrite(st3, sizeof st3, 1, f);
  std::fwrite(&secs, sizeof secs, 1, f);
  std::fwrite(res.data(), sizeof(res[0]), res.size(), f);
  std::fclose(f);
  host::bhac_io::free_model(&model);
  return 0;
}
