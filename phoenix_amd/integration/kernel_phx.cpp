// device::warmup_start, device::count_progress and device::render (declared in the reference's
// src/device/kernel.cuh:36-66) on libphx: the translation unit that takes the place of
// src/device/kernel.cu when the reference's own driver, src/main.cpp, is built for MI355X.
// main.cpp and the reference headers stay unchanged; this file maps core::bhac::Model onto
// phx_model, renders through phx_dev_render and reports device::Info.
//
// Precision: the build's ACCURACY (FP32 -> PHX_FP32, FP64 -> PHX_FP64), or PHX_ACCURACY=FP32|FP64|MIXED
// in the environment, as for libphx's own driver.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "device/kernel.cuh"
#include "phx.h"

static_assert(sizeof(core::pixel::FreqOut) == sizeof(phx_freqout), "FreqOut and phx_freqout share one layout");

namespace device {

static void check(int rc, const char* what) {
  if (rc != PHX_OK) throw std::runtime_error(std::string(what) + ": " + phx_last_error());
}

void warmup_start() { check(phx_dev_warmup(), "device warmup"); }

void count_progress(int n, void* user) { static_cast<std::atomic<int>*>(user)->fetch_add(n); }

static int precision_of_build() {
  const char* e = std::getenv("PHX_ACCURACY");
  if (e && !std::strcmp(e, "MIXED")) return PHX_MIXED;
  if (e && !std::strcmp(e, "FP64")) return PHX_FP64;
  if (e && !std::strcmp(e, "FP32")) return PHX_FP32;
  if (e) throw std::runtime_error("PHX_ACCURACY must be FP32, FP64 or MIXED");
  return ACCURACY == FP64 ? PHX_FP64 : PHX_FP32;
}

void render(const Model* m, const char* dump, const int* jobs, int njobs, FreqOut* res, Info* info,
            Progress progress, void* user) {
  const core::bhac::UserParams& u = m->uparams;
  if (u.nfreqs < 1 || u.nfreqs > PHX_FREQS_MAX) throw std::runtime_error("libphx renders 1..32 frequencies per call");
  phx_model p;
  std::memset(&p, 0, sizeof p);
  // [camera] [black_hole] [integration] [plasma] [grmhd]
  p.inclination = u.inclination;
  p.fov_x = u.fov_x;
  p.fov_y = u.fov_y;
  p.r_cam = u.r_cam;
  p.t_init = u.t_init;
  p.pa = u.pa;
  p.phi_cam = u.phi_cam;
  p.img_width = u.img_width;
  p.img_height = u.img_height;
  p.nfreqs = u.nfreqs;
  for (int f = 0; f < u.nfreqs; f++) p.frequencies[f] = u.frequencies[f];
  p.mass = u.mass;
  p.distance = u.distance;
  p.stepsize = u.stepsize;
  p.horizon_margin = u.horizon_margin;
  p.max_steps = u.max_steps;
  p.max_order = u.max_order;
  p.m_unit = u.m_unit;
  p.r_low = u.r_low;
  p.r_high = u.r_high;
  p.beta_c = u.beta_c;
  p.sigma_cut = u.sigma_cut;
  p.thetae_min = u.thetae_min;
  p.thetae_max = u.thetae_max;
  p.rt_outer = u.rt_outer;
  p.nspin = u.nspin;
  p.bpol = u.bpol;
  p.nlevels = u.nlevels;
  p.sfc = u.sfc;
  // grid file, dump header and the grid scalars the loader derived (bhac_io.hpp)
  for (int d = 0; d < 3; d++) {
    p.nxlone[d] = m->gparams.nxlone[d];
    p.nx[d] = m->nx[d];
    p.ng[d] = m->ng[d];
    p.xprobmin[d] = m->gparams.xprobmin[d];
    p.xprobmax[d] = m->gparams.xprobmax[d];
  }
  p.hslope_grid = m->gparams.hslope;
  p.nleafs = m->nleafs;
  p.ndim = m->ndim;
  p.nwini = m->nwini;
  p.nws = m->nws;
  p.cells = m->cells;
  p.a = m->a;
  p.hslope = m->hslope;
  p.gam = m->gam;
  for (int l = 0; l < PHX_MAX_LEVELS; l++) {
    for (int d = 0; d < 3; d++) {
      p.lvl_w[l][d] = m->lvl_w[l][d];
      p.level_dims[l][d] = m->level_dims[l][d];
    }
    p.level_map[l] = m->level_map[l];
  }
  for (int d = 0; d < 4; d++) {
    p.x_lo[d] = m->x_lo[d];
    p.x_hi[d] = m->x_hi[d];
  }
  // the block tables and primitives in double (the reference holds them as `scalar`)
  const size_t n3 = (size_t)3 * m->nleafs;
  std::vector<double> lb(m->lb, m->lb + n3), dxc(m->dxc, m->dxc + n3), prim;
  p.lb = lb.data();
  p.dxc = dxc.data();
  if (m->prim) {  // decoded on the host; NULL: libphx decodes `dump` on the device (grmhd.gpu_decode)
    prim.assign(m->prim, m->prim + (size_t)m->nleafs * m->cells * core::bhac::NPRIM);
    p.prim = prim.data();
  }
  check(phx_model_derive(&p), "model");  // units, plasma, camera rotation, cut_in, r_out, hconst / jscale

  phx_info pi;
  std::memset(&pi, 0, sizeof pi);
  check(phx_dev_render(&p, dump, precision_of_build(), jobs, njobs, reinterpret_cast<phx_freqout*>(res), &pi,
                       progress, user),
        "render");
  std::snprintf(info->device, sizeof info->device, "%s", pi.device);
  info->sms = pi.cus;
  info->blocks = pi.blocks;
  info->block = pi.block;
  info->regs = pi.vgprs;
  info->local_bytes = pi.scratch_bytes;
  info->upload_s = pi.upload_s;
  info->kernel_s = pi.kernel_s;
  info->decode_s = pi.decode_s;
  info->steps = pi.steps;
  info->nplasma = pi.nplasma;
  info->bad = pi.bad;
}

}  // namespace device
