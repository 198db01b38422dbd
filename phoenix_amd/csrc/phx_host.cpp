// Host layer of libphx: config.toml + CLI overrides, BHAC grid/dump reader, job
// schedule and flux reduction.  These keep the reference's behaviour (keys,
// file formats, error cases); the raymarch itself is device-only
// (phx_device.hip).
#include <dlfcn.h>
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "phx_internal.h"

namespace phx {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

namespace {

constexpr double K_MP = 1.6726219e-24, K_ME = 9.1093829e-28, K_C = 2.99792458e10, K_H = 6.62606885e-27;
constexpr double K_G = 6.674e-8, K_MSUN = 1.989e33, K_KPC = 3.086e21, K_JY = 1e23;

std::string trim(const std::string& t, const char* ws = " \t\r") {
  const size_t p = t.find_first_not_of(ws), q = t.find_last_not_of(ws);
  return p == std::string::npos ? std::string() : t.substr(p, q - p + 1);
}

// TOML subset of the reference (src/host/io/config.hpp:162-218): '#' comments outside
// strings, [section] headers, key = value with optional quotes, duplicates are errors.
void parse_toml(const std::string& text, std::map<std::string, std::string>& kv) {
  std::istringstream in(text);
  std::string line, section;
  int ln = 0;
  while (std::getline(in, line)) {
    ln++;
    bool quoted = false;
    for (size_t i = 0; i < line.size(); i++) {
      if (line[i] == '"') quoted = !quoted;
      if (line[i] == '#' && !quoted) {
        line.resize(i);
        break;
      }
    }
    line = trim(line);
    if (line.empty()) continue;
    if (line[0] == '[') {
      if (line.back() != ']') throw Error(PHX_E_ARG, "Line " + std::to_string(ln) + ": bad section header");
      section = trim(line.substr(1, line.size() - 2));
      continue;
    }
    const size_t eq = line.find('=');
    if (eq == std::string::npos) throw Error(PHX_E_ARG, "Line " + std::to_string(ln) + ": expected key = value");
    const std::string key = trim(line.substr(0, eq));
    std::string val = trim(line.substr(eq + 1));
    if (val.size() >= 2 && val.front() == '"' && val.back() == '"') val = val.substr(1, val.size() - 2);
    const std::string full = section.empty() ? key : section + "." + key;
    if (kv.count(full)) throw Error(PHX_E_ARG, "Line " + std::to_string(ln) + ": duplicate key " + full);
    kv[full] = val;
  }
}

// Exactly one of frequency / frequency_list / frequency_linspace / frequency_logspace
// (src/host/io/config.hpp:61-153).
std::vector<double> read_frequencies(const std::map<std::string, std::string>& kv, std::set<std::string>& used) {
  static const char* forms[4] = {"camera.frequency", "camera.frequency_list", "camera.frequency_linspace",
                                 "camera.frequency_logspace"};
  int given = 0, form = 0;
  for (int i = 0; i < 4; i++)
    if (kv.count(forms[i])) {
      given++;
      form = i;
    }
  if (given != 1) throw Error(PHX_E_ARG, "Too many (or no) frequency forms given!");
  const std::string key = forms[form];
  const std::string& text = kv.at(key);
  used.insert(key);
  auto number = [&](const std::string& t) {
    char* end = nullptr;
    const double v = std::strtod(t.c_str(), &end);
    if (end == t.c_str() || *end != '\0') throw Error(PHX_E_ARG, key + ": not a number: " + t);
    return v;
  };
  auto list = [&]() {
    const size_t a = text.find('['), b = text.rfind(']');
    if (a == std::string::npos || b == std::string::npos || b < a)
      throw Error(PHX_E_ARG, key + " must be a list on ONE line: [a, b, ...]");
    std::vector<double> v;
    std::stringstream ss(text.substr(a + 1, b - a - 1));
    std::string tok;
    while (std::getline(ss, tok, ',')) {
      const std::string t = trim(tok, " \t");
      if (t.empty()) continue;
      v.push_back(number(t));
    }
    return v;
  };
  std::vector<double> out;
  if (form == 0) {
    out.push_back(number(text));
  } else {
    const std::vector<double> v = list();
    if (form == 1) {
      if (v.empty()) throw Error(PHX_E_ARG, key + " is empty!");
      out = v;
    } else {
      if (v.size() != 3)
        throw Error(PHX_E_ARG, key + " has " + std::to_string(v.size()) + "arguments! (Expected 3)");
      const double lo = v[0], hi = v[1];
      if (!(v[2] >= 2. && v[2] <= 1e6) || v[2] != std::floor(v[2]))
        throw Error(PHX_E_ARG, key + ": n must be an integer >= 2! (use frequency = nu for a single frequency)");
      if (!(lo > 0.) || !(hi > lo)) throw Error(PHX_E_ARG, key + ": need 0 < start < stop!");
      const int n = (int)v[2];
      for (int i = 0; i < n; i++) {
        const double x = double(i) / (n - 1);
        out.push_back(i == n - 1 ? hi : (form == 3 ? lo * std::pow(hi / lo, x) : lo + (hi - lo) * x));
      }
    }
  }
  std::set<std::string> seen;
  for (double v : out) {
    if (!(v > 0.) || !std::isfinite(v)) throw Error(PHX_E_ARG, "frequencies must be positive and finite");
    char b[64];
    std::snprintf(b, sizeof b, "%e", v);
    if (!seen.insert(b).second)
      throw Error(PHX_E_ARG, std::string("duplicate frequency ") + b + " (they are used to name the HDF5 datasets)");
  }
  return out;
}

void copy_str(char* dst, size_t n, const std::string& s) {
  if (s.size() >= n) throw Error(PHX_E_RANGE, "string too long: " + s);
  std::memcpy(dst, s.c_str(), s.size() + 1);
}

}  // namespace

// src/host/io/config.hpp:229-381
static void load_config(const std::string& path, const std::vector<std::string>& overrides, phx_config* cfg,
                        phx_model* m) {
  std::ifstream file(path);
  if (!file) throw Error(PHX_E_IO, "Cannot open file " + path + ". Perhaps it's a malformed TOML key-value pair?");
  std::stringstream buf;
  buf << file.rdbuf();
  std::map<std::string, std::string> kv;
  parse_toml(buf.str(), kv);
  bool freq_override = false;
  for (const std::string& a : overrides) {
    const size_t eq = a.find('=');
    if (eq == std::string::npos) throw Error(PHX_E_ARG, "Override must be section.key=value: " + a);
    if (a.compare(0, 15, "camera.frequenc") == 0 && !freq_override) {  // replaces every frequency form
      freq_override = true;
      for (auto it = kv.begin(); it != kv.end();)
        it = (it->first.compare(0, 15, "camera.frequenc") == 0) ? kv.erase(it) : std::next(it);
    }
    kv[a.substr(0, eq)] = a.substr(eq + 1);
  }
  std::set<std::string> used;
  std::string missing;
  auto str = [&](const char* k, std::string& dst, bool required) {
    if (kv.count(k)) {
      dst = kv[k];
      used.insert(k);
    } else if (required) {
      missing += std::string(" ") + k;
    }
  };
  auto dbl = [&](const char* k, double& dst, bool required) {
    std::string t;
    str(k, t, required);
    if (!t.empty()) dst = std::atof(t.c_str());
  };
  auto integer = [&](const char* k, int32_t& dst, bool required) {
    double v = dst;
    dbl(k, v, required);
    dst = (int32_t)v;
  };
  auto boolean = [&](const char* k, int32_t& dst, bool required) {
    std::string t;
    str(k, t, required);
    if (!t.empty()) dst = (t == "true" || t == "1");
  };
  std::string s;
  boolean("visual.show_header", cfg->show_header, true);
  dbl("black_hole.mass", m->mass, true);
  dbl("black_hole.distance", m->distance, true);
  dbl("integration.stepsize", m->stepsize, true);
  integer("integration.max_steps", m->max_steps, true);
  dbl("integration.horizon_margin", m->horizon_margin, true);
  integer("integration.max_order", m->max_order, true);
  dbl("camera.inclination", m->inclination, true);
  integer("camera.img_width", m->img_width, true);
  integer("camera.img_height", m->img_height, true);
  dbl("camera.fov_x", m->fov_x, true);
  dbl("camera.fov_y", m->fov_y, true);
  dbl("camera.r_cam", m->r_cam, true);
  dbl("camera.t_init", m->t_init, true);
  m->pa = 0.;
  m->phi_cam = 0.;
  dbl("camera.pa", m->pa, false);
  dbl("camera.phi", m->phi_cam, false);
  try {
    const std::vector<double> f = read_frequencies(kv, used);
    if (f.size() > (size_t)PHX_FREQS_MAX)
      throw Error(PHX_E_RANGE, "at most " + std::to_string(PHX_FREQS_MAX) + " frequencies (PHX_MAX_FREQS), got " +
                                   std::to_string(f.size()));
    for (size_t i = 0; i < f.size(); i++) m->frequencies[i] = f[i];
    m->nfreqs = (int32_t)f.size();
  } catch (const Error& e) {
    throw Error(e.code, path + ": " + e.what());
  }
  std::snprintf(cfg->image, sizeof cfg->image, "hdf5");
  cfg->block_size = 10;
  str("output.dir", s, true);
  copy_str(cfg->outdir, sizeof cfg->outdir, s);
  s = cfg->image;
  str("output.image", s, false);
  if (s != "hdf5" && s != "text" && s != "both" && s != "none")
    throw Error(PHX_E_ARG, path + ": output.image must be hdf5, text, both or none");
  copy_str(cfg->image, sizeof cfg->image, s);
  integer("output.block_size", cfg->block_size, false);
  if (cfg->block_size < 1) throw Error(PHX_E_ARG, path + ": output.block_size must be >= 1");
  s.clear();
  str("grmhd.dump_file", s, true);
  copy_str(cfg->dump_file, sizeof cfg->dump_file, s);
  s.clear();
  str("grmhd.grid_file", s, true);
  copy_str(cfg->grid_file, sizeof cfg->grid_file, s);
  boolean("grmhd.gpu_decode", cfg->gpu_decode, true);
  integer("grmhd.nspin", m->nspin, true);
  boolean("grmhd.sfc", m->sfc, true);
  integer("grmhd.nlevels", m->nlevels, true);
  integer("grmhd.bpol", m->bpol, true);
  dbl("plasma.m_unit", m->m_unit, true);
  dbl("plasma.r_low", m->r_low, true);
  dbl("plasma.r_high", m->r_high, true);
  dbl("plasma.sigma_cut", m->sigma_cut, true);
  dbl("plasma.thetae_min", m->thetae_min, true);
  dbl("plasma.thetae_max", m->thetae_max, true);
  dbl("plasma.rt_outer", m->rt_outer, true);
  m->beta_c = 1.;
  dbl("plasma.beta_c", m->beta_c, false);
  if (!(m->beta_c > 0.)) throw Error(PHX_E_ARG, path + ": plasma.beta_c must be > 0");
  if (!missing.empty()) throw Error(PHX_E_ARG, path + ": missing required key(s):" + missing);
  for (const auto& e : kv)
    if (!used.count(e.first)) throw Error(PHX_E_ARG, path + ": unknown key " + e.first);
}

// src/host/io/bhac_io.hpp:143-174: ten values, the last token of each non-empty line
static void load_grid(const std::string& path, phx_model* m) {
  std::ifstream in(path);
  if (!in) throw Error(PHX_E_IO, "cannot open grid file " + path);
  std::vector<double> v;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string tok, last;
    while (ss >> tok) last = tok;
    if (!last.empty()) v.push_back(std::atof(last.c_str()));
  }
  if (v.size() < 10)
    throw Error(PHX_E_IO, path + ": expected 10 lines (nxlone x3, xprobmin x3, xprobmax x3, hslope)");
  for (int d = 0; d < 3; d++) {
    m->nxlone[d] = (int32_t)v[d];
    m->xprobmin[d] = v[3 + d];
    m->xprobmax[d] = v[6 + d];
  }
  m->hslope_grid = v[9];
  for (int d = 0; d < 3; d++)
    if (m->nxlone[d] <= 0 || m->xprobmax[d] <= m->xprobmin[d])
      throw Error(PHX_E_IO, path + ": nxlone / xprobmin / xprobmax not set (or empty domain)");
}

// per-request scalars of finalize_model: units, plasma, camera, per-frequency
// constants (src/host/io/bhac_io.hpp:224-250)
static void derive_request(phx_model* m) {
  if (m->nfreqs < 1 || m->nfreqs > PHX_FREQS_MAX) throw Error(PHX_E_ARG, "nfreqs must be in 1..32");
  if (!(m->beta_c > 0.)) throw Error(PHX_E_ARG, "plasma.beta_c must be > 0");
  m->L = K_G * m->mass * K_MSUN / K_C / K_C;
  m->T_unit = m->L / K_C;
  m->rho_unit = m->m_unit / std::pow(m->L, 3);
  m->B_unit = K_C * std::sqrt(4. * M_PI * m->rho_unit);
  m->Ne_unit = m->rho_unit / (K_MP + K_ME);
  m->U_unit = m->rho_unit * K_C * K_C;
  m->Bu = m->B_unit;
  m->Ne = m->Ne_unit;
  m->rhigh = m->r_high;
  m->rlow = m->r_low;
  m->ibc = 1. / m->beta_c;
  m->cos_pa = std::cos(m->pa / 180. * M_PI);
  m->sin_pa = std::sin(m->pa / 180. * M_PI);
  m->cut_in = (1. + std::sqrt(1. - m->a * m->a)) * (1. + m->horizon_margin);
  m->r_out = 1.1 * m->r_cam;
  const double cc = m->L * K_H / (K_ME * K_C * K_C);
  for (int f = 0; f < m->nfreqs; f++) {
    const double nu = m->frequencies[f];
    if (!(nu > 0.) || !std::isfinite(nu)) throw Error(PHX_E_ARG, "frequencies must be positive and finite");
    m->freq[f] = nu;
    m->hconst[f] = (K_ME * K_C * K_C) / (K_H * nu) * cc;
    m->jscale[f] = nu * nu * nu;
  }
}

// units and per-run scalars (src/host/io/bhac_io.hpp:183-251)
static void finalize(phx_model* m, const std::vector<double>& neqpar) {
  const double spin = neqpar[m->nspin];
  if (std::fabs(spin) > 1.) throw Error(PHX_E_ARG, "Spin " + std::to_string(spin) + " unphysical (wrong grmhd.nspin?)");
  m->a = spin;
  m->hslope = m->hslope_grid;
  m->gam = neqpar[0];
  for (int d = 1; d < 3; d++) {
    m->xprobmin[d] *= 2. * M_PI;
    m->xprobmax[d] *= 2. * M_PI;
  }
  m->cells = m->nx[0] * m->nx[1] * m->nx[2];
  for (int d = 0; d < 3; d++) {
    if (m->nx[d] <= 0 || m->nxlone[d] % m->nx[d] != 0)
      throw Error(PHX_E_IO, "nxlone[" + std::to_string(d) + "] = " + std::to_string(m->nxlone[d]) +
                                " is not a multiple of the dump's block size " + std::to_string(m->nx[d]) +
                                ": wrong grid file");
    m->ng[d] = m->nxlone[d] / m->nx[d];
    m->x_lo[1 + d] = m->xprobmin[d];
    m->x_hi[1 + d] = m->xprobmax[d];
  }
  m->x_lo[0] = m->x_hi[0] = 0.;
  for (int l = 0; l < PHX_MAX_LEVELS; l++)
    for (int d = 0; d < 3; d++) m->lvl_w[l][d] = (m->xprobmax[d] - m->xprobmin[d]) / m->ng[d] / (double)(1 << l);
  derive_request(m);
}

static void walk(FILE* f, int level, int i, int j, int k, std::vector<int>& ind, std::vector<int>& lev, bool& ok) {
  if (!ok) return;
  int32_t leaf = 0;
  if (std::fread(&leaf, 4, 1, f) != 1) {
    ok = false;
    return;
  }
  if (leaf) {
    ind.insert(ind.end(), {i, j, k});
    lev.push_back(level);
    return;
  }
  for (int c = 0; c < 8; c++)  // children (2i + c%2, 2j + (c/2)%2, 2k + c/4) (bhac_io.hpp:39-66)
    walk(f, level + 1, 2 * i + c % 2, 2 * j + (c / 2) % 2, 2 * k + c / 4, ind, lev, ok);
}

static uint64_t morton3(unsigned a, unsigned b, unsigned c) {  // bhac_io.hpp:123-130
  uint64_t z = 0;
  for (uint64_t i = 0; i < 21; ++i)
    z |= (((uint64_t)a >> i & 1) << (3 * i)) | (((uint64_t)b >> i & 1) << (3 * i + 1)) |
         (((uint64_t)c >> i & 1) << (3 * i + 2));
  return z;
}

// src/host/io/bhac_io.hpp:331-455 (tables only; the decode runs on the device)
static void load_dump(phx_model* m, const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw Error(PHX_E_IO, "Can't open " + path);
  struct Closer {
    FILE* f;
    ~Closer() { std::fclose(f); }
  } closer{f};
  int32_t hdr[8];
  double tdump;
  std::fseek(f, -40, SEEK_END);
  if (std::fread(hdr, 4, 8, f) != 8 || std::fread(&tdump, 8, 1, f) != 1)
    throw Error(PHX_E_IO, "Bad read in the dump header. Is " + path + " a BHAC dump?");
  m->nleafs = hdr[0];
  m->ndim = hdr[2];
  m->nwini = hdr[4];
  m->nws = hdr[5];
  const int neq = hdr[6];
  if (m->ndim != 3) throw Error(PHX_E_IO, "Only 3D dumps are supported! (ndim = " + std::to_string(m->ndim) + ")");
  if (neq <= 0 || neq > 1024 || m->nleafs <= 0 || m->nwini < 9) throw Error(PHX_E_IO, "implausible dump header");
  std::fseek(f, -40 - (m->ndim * 4 + neq * 8), SEEK_END);
  std::vector<double> neqpar(neq);
  bool ok = std::fread(m->nx, 4, 3, f) == 3 && std::fread(neqpar.data(), 8, neq, f) == (size_t)neq;
  if (!ok || m->nspin < 0 || m->nspin >= neq) throw Error(PHX_E_IO, "Short read in the dump header! (wrong grmhd.nspin?)");
  finalize(m, neqpar);
  const long blk = (long)m->cells * m->nwini * 8;
  const long stag = (long)(m->nx[0] + 1) * (m->nx[1] + 1) * (m->nx[2] + 1) * m->nws * 8;
  std::fseek(f, (long)m->nleafs * (blk + stag), SEEK_SET);
  std::vector<int> ind, lev;
  ok = true;
  if (m->sfc) {
    std::vector<std::pair<uint64_t, int>> order;
    for (int i = 0; i < m->ng[0]; i++)
      for (int j = 0; j < m->ng[1]; j++)
        for (int k = 0; k < m->ng[2]; k++) order.push_back({morton3(i, j, k), (i * m->ng[1] + j) * m->ng[2] + k});
    std::sort(order.begin(), order.end());
    for (const auto& o : order) {
      const int n = o.second;
      walk(f, 1, n / (m->ng[1] * m->ng[2]), (n / m->ng[2]) % m->ng[1], n % m->ng[2], ind, lev, ok);
    }
  } else {
    for (int k = 0; k < m->ng[2]; k++)
      for (int j = 0; j < m->ng[1]; j++)
        for (int i = 0; i < m->ng[0]; i++) walk(f, 1, i, j, k, ind, lev, ok);
  }
  if (!ok) throw Error(PHX_E_IO, "short read in the tree (wrong nxlone / grmhd.sfc for this dump?)");
  if ((int)lev.size() != m->nleafs)
    throw Error(PHX_E_IO, "grid mismatch: tree has " + std::to_string(lev.size()) + " leaves, header says " +
                              std::to_string(m->nleafs) + " (grmhd.sfc / grid file?)");
  const size_t n3 = 3 * (size_t)m->nleafs;
  m->lb = (double*)std::malloc(n3 * sizeof(double));
  m->dxc = (double*)std::malloc(n3 * sizeof(double));
  m->leaf_level = (int32_t*)std::malloc(m->nleafs * sizeof(int32_t));
  m->leaf_index = (int32_t*)std::malloc(n3 * sizeof(int32_t));
  m->prim = nullptr;
  double dx1[3], dxc0[3];
  for (int d = 0; d < 3; d++) {
    dx1[d] = (m->xprobmax[d] - m->xprobmin[d]) / m->ng[d];
    dxc0[d] = (m->xprobmax[d] - m->xprobmin[d]) / m->nxlone[d];
  }
  for (int q = 0; q < m->nleafs; q++) {
    const double inv = 1.0 / (double)(1 << (lev[q] - 1));
    m->leaf_level[q] = lev[q];
    for (int d = 0; d < 3; d++) {
      m->lb[3 * q + d] = m->xprobmin[d] + ind[3 * q + d] * dx1[d] * inv;
      m->dxc[3 * q + d] = dxc0[d] * inv;
      m->leaf_index[3 * q + d] = ind[3 * q + d];
    }
  }
  // level map (bhac_io.hpp:257-309)
  if (m->nlevels == -1) {
    for (int q = 0; q < m->nleafs; q++) m->nlevels = std::max<int32_t>(m->nlevels, lev[q]);
    if (m->nlevels > PHX_MAX_LEVELS)
      throw Error(PHX_E_RANGE, std::to_string(m->nlevels) + " AMR levels unsupported (max 16)");
  }
  if (m->nlevels < 1 || m->nlevels > PHX_MAX_LEVELS) throw Error(PHX_E_RANGE, "grmhd.nlevels out of range");
  for (int l = 0; l < m->nlevels; l++) {
    for (int d = 0; d < 3; d++) m->level_dims[l][d] = m->ng[d] * (1 << l);
    const size_t n = (size_t)m->level_dims[l][0] * m->level_dims[l][1] * m->level_dims[l][2];
    m->level_map[l] = (int32_t*)std::malloc(n * sizeof(int32_t));
    for (size_t i = 0; i < n; i++) m->level_map[l][i] = -1;
    for (int q = 0; q < m->nleafs; q++) {
      if (lev[q] - 1 != l) continue;
      const int i = ind[3 * q], j = ind[3 * q + 1], k = ind[3 * q + 2];
      if (i < 0 || i >= m->level_dims[l][0] || j < 0 || j >= m->level_dims[l][1] || k < 0 || k >= m->level_dims[l][2]) {
        std::fprintf(stderr, "[ INFO ] Leaf %d index (%d,%d,%d) outside level-%d dims\n", q, i, j, k, l + 1);
        continue;
      }
      int32_t& slot = m->level_map[l][((size_t)i * m->level_dims[l][1] + j) * m->level_dims[l][2] + k];
      if (slot != -1) std::fprintf(stderr, "[ FAIL ] Level-map collision at leaf %d\n", q);
      slot = q;
    }
  }
}

}  // namespace phx

using namespace phx;

namespace phx {
void flux_scale(const phx_model* m, double* dOm, double* jy) {
  const double dist = m->distance * K_KPC;
  *dOm = (m->fov_x / m->img_width * m->L / dist) * (m->fov_y / m->img_height * m->L / dist);
  *jy = K_JY;
}
}  // namespace phx

extern "C" {

const char* phx_last_error(void) { return g_err.c_str(); }
const char* phx_version(void) { return "phoenix-mi355x 0.1 (gfx950)"; }

int phx_load_config(const char* path, const char* const* overrides, int32_t n, phx_config* cfg, phx_model* model) {
  return guarded([&] {
    if (!path || !cfg || !model || n < 0) throw Error(PHX_E_ARG, "phx_load_config: null argument");
    std::vector<std::string> ov;
    for (int i = 0; i < n; i++) ov.emplace_back(overrides[i]);
    load_config(path, ov, cfg, model);
  });
}

int phx_load_grid(const char* grid, phx_model* model) {
  return guarded([&] {
    if (!grid || !model) throw Error(PHX_E_ARG, "phx_load_grid: null argument");
    load_grid(grid, model);
  });
}

int phx_load_dump(phx_model* model, const char* dump) {
  return guarded([&] {
    if (!dump || !model) throw Error(PHX_E_ARG, "phx_load_dump: null argument");
    try {
      load_dump(model, dump);
    } catch (...) {
      phx_free_model(model);
      throw;
    }
  });
}

int phx_model_derive(phx_model* model) {
  return guarded([&] {
    if (!model) throw Error(PHX_E_ARG, "phx_model_derive: null model");
    derive_request(model);
  });
}

void phx_free_model(phx_model* m) {
  if (!m) return;
  std::free(m->lb);
  std::free(m->dxc);
  std::free(m->prim);
  std::free(m->leaf_level);
  std::free(m->leaf_index);
  for (int l = 0; l < PHX_MAX_LEVELS; l++) std::free(m->level_map[l]);
  m->lb = m->dxc = m->prim = nullptr;
  m->leaf_level = m->leaf_index = nullptr;
  for (int l = 0; l < PHX_MAX_LEVELS; l++) m->level_map[l] = nullptr;
}

int phx_schedule(const phx_model* m, int32_t morton, int32_t* jobs, int32_t njobs) {
  return guarded([&] {
    const int64_t n = (int64_t)m->img_width * m->img_height;
    if (njobs != n) throw Error(PHX_E_ARG, "phx_schedule: njobs must be img_width * img_height");
    std::vector<std::pair<uint64_t, int32_t>> keyed;
    keyed.reserve(n);
    for (int i = 0; i < m->img_width; i++)
      for (int j = 0; j < m->img_height; j++) {
        uint64_t z = 0;
        if (morton)
          for (int b = 0; b < 16; b++)
            z |= ((uint64_t)((i >> b) & 1) << (2 * b)) | ((uint64_t)((j >> b) & 1) << (2 * b + 1));
        keyed.push_back({z, i * m->img_height + j});
      }
    if (morton) std::sort(keyed.begin(), keyed.end());
    for (int64_t q = 0; q < n; q++) jobs[q] = keyed[q].second;
  });
}

int phx_compute_spec(const phx_model* m, const phx_freqout* res, int32_t fi, double S[4]) {
  return guarded([&] {
    if (fi < 0 || fi >= m->nfreqs) throw Error(PHX_E_ARG, "frequency index out of range");
    const double dist = m->distance * K_KPC;
    const double dOm = (m->fov_x / m->img_width * m->L / dist) * (m->fov_y / m->img_height * m->L / dist);
    for (int q = 0; q < 4; q++) S[q] = 0.;
    for (int i = 0; i < m->img_width; i++)
      for (int j = 0; j < m->img_height; j++) {
        const phx_freqout& r = res[((size_t)i * m->img_height + j) * m->nfreqs + fi];
        if (!std::isfinite(r.IQUV[0] + r.IQUV[1] + r.IQUV[2] + r.IQUV[3])) continue;
        for (int q = 0; q < 4; q++) S[q] += r.IQUV[q] * dOm;
      }
    for (int q = 0; q < 4; q++) S[q] *= K_JY;
  });
}


// ---------------------------------------------------------------- output (src/host/io/output.hpp)
}  // extern "C"

namespace {

const phx_freqout& result_at(const phx_model* m, const phx_freqout* res, int i, int j, int f) {
  return res[((size_t)i * m->img_height + j) * m->nfreqs + f];
}
// camera::pixel_impact (src/core/geometry/camera.hpp:11-23)
void impact(const phx_model* m, int i, int j, double* a, double* b) {
  *a = -0.5 * m->fov_x + (i + 0.5) * m->fov_x / m->img_width;
  *b = -0.5 * m->fov_y + (j + 0.5) * m->fov_y / m->img_height;
}
// camera::pixel_solid_angle (camera.hpp:28-32)
double solid_angle(const phx_model* m) {
  const double dist = m->distance * K_KPC;
  return (m->fov_x / m->img_width * m->L / dist) * (m->fov_y / m->img_height * m->L / dist);
}

// The HDF5 C library, loaded at run time (the image's conda HDF5 >= 1.10, where hid_t is 64-bit),
// so that libphx itself has no link dependency on it.
struct H5 {
  typedef int64_t hid_t;
  typedef unsigned long long hsize_t;
  int (*open)() = nullptr;
  int (*get_libversion)(unsigned*, unsigned*, unsigned*) = nullptr;
  hid_t (*Fcreate)(const char*, unsigned, hid_t, hid_t) = nullptr;
  int (*Fclose)(hid_t) = nullptr;
  hid_t (*Screate_simple)(int, const hsize_t*, const hsize_t*) = nullptr;
  int (*Sclose)(hid_t) = nullptr;
  hid_t (*Dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t) = nullptr;
  int (*Dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*) = nullptr;
  int (*Dclose)(hid_t) = nullptr;
  hid_t* native_double = nullptr;
  bool ok = false;
  std::string why;
};

const H5& hdf5() {
  static H5 h;
  static bool tried = false;
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  if (tried) return h;
  tried = true;
  const char* names[] = {std::getenv("PHX_HDF5_LIB"), "libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so",
                         "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so"};
  void* lib = nullptr;
  for (const char* n : names)
    if (n && (lib = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
  if (!lib) {
    h.why = "no HDF5 library found (set PHX_HDF5_LIB)";
    return h;
  }
  auto sym = [&](const char* n) { return dlsym(lib, n); };
  h.open = (int (*)())sym("H5open");
  h.get_libversion = (int (*)(unsigned*, unsigned*, unsigned*))sym("H5get_libversion");
  h.Fcreate = (H5::hid_t(*)(const char*, unsigned, H5::hid_t, H5::hid_t))sym("H5Fcreate");
  h.Fclose = (int (*)(H5::hid_t))sym("H5Fclose");
  h.Screate_simple = (H5::hid_t(*)(int, const H5::hsize_t*, const H5::hsize_t*))sym("H5Screate_simple");
  h.Sclose = (int (*)(H5::hid_t))sym("H5Sclose");
  h.Dcreate2 = (H5::hid_t(*)(H5::hid_t, const char*, H5::hid_t, H5::hid_t, H5::hid_t, H5::hid_t, H5::hid_t))sym(
      "H5Dcreate2");
  h.Dwrite = (int (*)(H5::hid_t, H5::hid_t, H5::hid_t, H5::hid_t, H5::hid_t, const void*))sym("H5Dwrite");
  h.Dclose = (int (*)(H5::hid_t))sym("H5Dclose");
  h.native_double = (H5::hid_t*)sym("H5T_NATIVE_DOUBLE_g");
  if (!h.open || !h.get_libversion || !h.Fcreate || !h.Fclose || !h.Screate_simple || !h.Sclose || !h.Dcreate2 ||
      !h.Dwrite || !h.Dclose || !h.native_double) {
    h.why = "HDF5 library lacks a needed symbol";
    return h;
  }
  unsigned maj = 0, min = 0, rel = 0;
  if (h.open() < 0 || h.get_libversion(&maj, &min, &rel) < 0 || maj != 1 || min < 10) {
    h.why = "HDF5 1.10 or newer needed (64-bit hid_t)";
    return h;
  }
  h.ok = true;
  return h;
}

enum { W_I, W_Q, W_U, W_V, W_TAU, W_TAUF, W_ALPHA, W_BETA };
// output.hpp:117-126
double h5_value(const phx_model* m, const phx_freqout* res, int what, int f, int i, int j) {
  if (what <= W_V) return result_at(m, res, i, j, f).IQUV[what] * K_JY;
  if (what == W_TAU) return result_at(m, res, i, j, f).tau;
  if (what == W_TAUF) return result_at(m, res, i, j, f).tauF;
  double a, b;
  impact(m, i, j, &a, &b);
  return what == W_ALPHA ? a : b;
}
// one [nblocks][bs*bs] dataset in the RAPTOR block layout (output.hpp:128-142): block b = bi * nbj + bj,
// pixel p = pi * bs + pj, image pixel (i, j) = (bi * bs + pi, bj * bs + pj)
void h5_put(const H5& h, H5::hid_t file, const std::string& name, const phx_model* m, int bs, const phx_freqout* res,
            int what, int f, std::vector<double>& buf) {
  const int nbj = m->img_height / bs, nblocks = (m->img_width / bs) * nbj, npb = bs * bs;
  for (int b = 0; b < nblocks; b++)
    for (int p = 0; p < npb; p++) {
      const int i = (b / nbj) * bs + p / bs, j = (b % nbj) * bs + p % bs;
      buf[(size_t)b * npb + p] = h5_value(m, res, what, f, i, j);
    }
  const H5::hsize_t dims[2] = {(H5::hsize_t)nblocks, (H5::hsize_t)npb};
  const H5::hid_t sp = h.Screate_simple(2, dims, nullptr);
  if (sp < 0) throw Error(PHX_E_IO, "H5Screate_simple failed");
  const H5::hid_t ds = h.Dcreate2(file, name.c_str(), *h.native_double, sp, 0, 0, 0);
  if (ds < 0) {
    h.Sclose(sp);
    throw Error(PHX_E_IO, "H5Dcreate2 failed for " + name);
  }
  const int rc = h.Dwrite(ds, *h.native_double, 0, 0, 0, buf.data());
  h.Dclose(ds);
  h.Sclose(sp);
  if (rc < 0) throw Error(PHX_E_IO, "H5Dwrite failed for " + name);
}

void check_out_args(const phx_model* m, const phx_freqout* res, const char* path) {
  if (!m || !res || !path) throw Error(PHX_E_ARG, "null model, result or path");
  if (m->img_width <= 0 || m->img_height <= 0 || m->nfreqs < 1) throw Error(PHX_E_ARG, "empty image");
}

}  // namespace

extern "C" {

int phx_compute_tau(const phx_model* m, const phx_freqout* res, int32_t fi, double T[2]) {
  return guarded([&] {
    check_out_args(m, res, "");
    if (fi < 0 || fi >= m->nfreqs) throw Error(PHX_E_ARG, "frequency index out of range");
    double s0 = 0., s1 = 0.;
    size_t n = 0;
    for (int i = 0; i < m->img_width; i++)
      for (int j = 0; j < m->img_height; j++) {
        const phx_freqout& r = result_at(m, res, i, j, fi);
        if (!std::isfinite(r.IQUV[0] + r.IQUV[1] + r.IQUV[2] + r.IQUV[3] + r.tau + r.tauF)) continue;
        s0 += r.tau;
        s1 += r.tauF;
        n++;
      }
    T[0] = n ? s0 / (double)n : 0.;
    T[1] = n ? s1 / (double)n : 0.;
  });
}

int phx_write_image_text(const phx_model* m, const phx_freqout* res, int32_t fi, const char* path) {
  return guarded([&] {
    check_out_args(m, res, path);
    if (fi < 0 || fi >= m->nfreqs) throw Error(PHX_E_ARG, "frequency index out of range");
    FILE* f = std::fopen(path, "w");
    if (!f) throw Error(PHX_E_IO, std::string("cannot open ") + path);
    const double unit = K_JY * solid_angle(m);
    const double arcsec = 206265.0 * m->L / (m->distance * K_KPC);
    for (int i = 0; i < m->img_width; i++)
      for (int j = 0; j < m->img_height; j++) {
        double a, b;
        impact(m, i, j, &a, &b);
        const phx_freqout& r = result_at(m, res, i, j, fi);
        std::fprintf(f, "%+.15e\t%+.15e\t%+.15e\t%+.15e\t%+.15e\t%+.15e\t%+.15e\t%+.15e\n", a * arcsec, b * arcsec,
                     r.IQUV[0] * unit, r.IQUV[1] * unit, r.IQUV[2] * unit, r.IQUV[3] * unit, r.tau, r.tauF);
      }
    if (std::fclose(f) != 0) throw Error(PHX_E_IO, std::string("write failed: ") + path);
  });
}

int phx_write_spectrum(const phx_model* m, const phx_freqout* res, const char* path) {
  return guarded([&] {
    check_out_args(m, res, path);
    FILE* f = std::fopen(path, "w");
    if (!f) throw Error(PHX_E_IO, std::string("cannot open ") + path);
    for (int fi = 0; fi < m->nfreqs; fi++) {
      double S[4], T[2];
      if (phx_compute_spec(m, res, fi, S) != PHX_OK || phx_compute_tau(m, res, fi, T) != PHX_OK) {
        std::fclose(f);
        throw Error(PHX_E_ARG, phx_last_error());
      }
      std::fprintf(f, "%+.15e\t%+.15e\t%+.15e\t%+.15e\t%+.15e\t%+.15e\t%+.15e\n", m->frequencies[fi], S[0], S[1],
                   S[2], S[3], T[0], T[1]);
    }
    if (std::fclose(f) != 0) throw Error(PHX_E_IO, std::string("write failed: ") + path);
  });
}

int phx_write_image_hdf5(const phx_model* m, const phx_freqout* res, int32_t bs, const char* path) {
  return guarded([&] {
    check_out_args(m, res, path);
    if (bs <= 0 || m->img_width % bs || m->img_height % bs)
      throw Error(PHX_E_ARG, "output.block_size " + std::to_string(bs) + " must divide the image size " +
                                 std::to_string(m->img_width) + "x" + std::to_string(m->img_height));
    const H5& h = hdf5();
    if (!h.ok) throw Error(PHX_E_IO, h.why);
    const H5::hid_t file = h.Fcreate(path, 0x0002u /* H5F_ACC_TRUNC */, 0, 0);
    if (file < 0) throw Error(PHX_E_IO, std::string("cannot create ") + path);
    try {
      std::vector<double> buf((size_t)(m->img_width / bs) * (m->img_height / bs) * bs * bs);
      for (int fi = 0; fi < m->nfreqs; fi++) {
        char fb[64];
        std::snprintf(fb, sizeof fb, "%e", m->frequencies[fi]);
        const std::string fs = fb;
        h5_put(h, file, "I" + fs, m, bs, res, W_I, fi, buf);
        h5_put(h, file, "tau" + fs, m, bs, res, W_TAU, fi, buf);
        h5_put(h, file, "Q" + fs, m, bs, res, W_Q, fi, buf);
        h5_put(h, file, "U" + fs, m, bs, res, W_U, fi, buf);
        h5_put(h, file, "V" + fs, m, bs, res, W_V, fi, buf);
        h5_put(h, file, "tauF" + fs, m, bs, res, W_TAUF, fi, buf);
      }
      h5_put(h, file, "alpha", m, bs, res, W_ALPHA, 0, buf);
      h5_put(h, file, "beta", m, bs, res, W_BETA, 0, buf);
    } catch (...) {
      h.Fclose(file);
      throw;
    }
    if (h.Fclose(file) < 0) throw Error(PHX_E_IO, std::string("H5Fclose failed: ") + path);
  });
}

int phx_write_output(const phx_model* m, const phx_config* cfg, const phx_freqout* res, char* log, int32_t log_len) {
  std::string msgs;
  const int rc = guarded([&] {
    if (!cfg) throw Error(PHX_E_ARG, "null config");
    check_out_args(m, res, "");
    const std::string outdir = cfg->outdir, image = cfg->image;
    if (::mkdir(outdir.c_str(), 0755) != 0 && errno != EEXIST) throw Error(PHX_E_IO, "cannot create " + outdir);
    bool want_text = image == "text" || image == "both";
    const bool want_h5 = image == "hdf5" || image == "both";
    if (want_h5) {
      const std::string hp = outdir + "/img_data.h5";
      if (phx_write_image_hdf5(m, res, cfg->block_size, hp.c_str()) == PHX_OK) {
        msgs += "Wrote " + hp + "\n";
      } else {
        msgs += "Failed to write \"" + hp + "\" (" + phx_last_error() + ") writing text instead...\n";
        want_text = true;
      }
    }
    if (want_text)
      for (int f = 0; f < m->nfreqs; f++) {
        const std::string ip = outdir + "/image" + (m->nfreqs == 1 ? "" : "_" + std::to_string(f)) + ".dat";
        if (phx_write_image_text(m, res, f, ip.c_str()) != PHX_OK) throw Error(PHX_E_IO, phx_last_error());
        msgs += "Wrote " + ip + "\n";
      }
    const std::string sp = outdir + "/spectrum.dat";
    if (phx_write_spectrum(m, res, sp.c_str()) != PHX_OK) throw Error(PHX_E_IO, phx_last_error());
    msgs += "Wrote " + sp + "\n";
  });
  if (log && log_len > 0) std::snprintf(log, (size_t)log_len, "%s", msgs.c_str());
  return rc;
}
}  // extern "C"
