// Test-infrastructure harness around the reference's own output writers.
//
// Compiled (oracle/Makefile, target ref_output) against the reference's sources where they lie
// and the image's HDF5 headers; output goes to oracle/_ref/ only.  It loads the run as the
// reference's main() does (src/main.cpp:110-150: load_config, load_bhac_grid, load), reads a
// result array, and writes it with the reference's host::output functions
// (src/host/io/output.hpp:77-167) as write_output does (src/main.cpp:52-86):
//   outdir/img_data.h5 (write_image_hdf5), outdir/image[_f].dat (write_image), outdir/spectrum.dat.
//
// Usage: phx_ref_output_fpXX config.toml results.bin outdir [section.key=value ...]
// results.bin: W*H*nf FreqOut records (6 float64 each), res[(i*H + j)*nf + f].
#include <sys/stat.h>

#include <cstdio>
#include <string>
#include <vector>

#include "core/render/headers.hpp"
#include "host/io/bhac_io.hpp"
#include "host/io/config.hpp"
#include "host/io/output.hpp"

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s config.toml results.bin outdir [section.key=value ...]\n", argv[0]);
    return 2;
  }
  const std::string cfg = argv[1], resp = argv[2], outdir = argv[3];
  std::vector<std::string> ov(argv + 4, argv + argc);
  core::bhac::Model model = {};
  host::config::Params params;
  try {
    host::config::load_config(cfg, ov, params, &model.uparams);
    host::bhac_io::load_bhac_grid(&model.gparams, params.grid_file);
    host::bhac_io::load(&model, params.dump_file, false);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "ref: %s\n", e.what());
    return 1;
  }
  const core::bhac::UserParams& up = model.uparams;
  std::vector<core::pixel::FreqOut> res((size_t)up.img_width * up.img_height * up.nfreqs);
  FILE* f = std::fopen(resp.c_str(), "rb");
  if (!f || std::fread(res.data(), sizeof(res[0]), res.size(), f) != res.size()) {
    std::fprintf(stderr, "ref: cannot read %s\n", resp.c_str());
    return 1;
  }
  std::fclose(f);
  mkdir(outdir.c_str(), 0755);
  int rc = 0;
  if (!host::output::write_image_hdf5(outdir + "/img_data.h5", &model, params.block_size, res)) rc |= 1;
  for (int fi = 0; fi < up.nfreqs; fi++)
    if (!host::output::write_image(outdir + "/image" + (up.nfreqs == 1 ? "" : "_" + std::to_string(fi)) + ".dat",
                                   &model, res, fi))
      rc |= 2;
  if (!host::output::write_spectrum(outdir + "/spectrum.dat", &model, res)) rc |= 4;
  host::bhac_io::free_model(&model);
  return rc;
}
