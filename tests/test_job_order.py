"""The job list of a request marched in place with several frequency chunks (wave_jobs / wave_job,
phoenix_amd/csrc/phx_lane.h), built for the CPU and checked over many sizes: every (chunk, pixel)
pair appears exactly once; with one chunk the jobs are the pixels in order; with more, every
aligned block of 64 jobs holds one chunk and one 64-pixel tile, in order, then padding (pixel -1),
so that a wave of 64 lanes marches the same rays for every chunk."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r"""
#include <cstdio>
#include <vector>
#include "phx_lane.h"
using namespace phxd;

static int check(int npix, int ntasks) {
  const long long total = wave_jobs(npix, ntasks);
  std::vector<int> seen((size_t)npix * ntasks, 0);
  long long pad = 0;
  for (long long n = 0; n < total; n++) {
    int t = -1;
    const int i = wave_job((int)n, npix, ntasks, &t);
    if (t < 0 || t >= ntasks) return 1;
    if (i < 0) { pad++; continue; }
    if (i >= npix) return 2;
    seen[(size_t)t * npix + i]++;
    if (ntasks == 1 && (i != n || t != 0)) return 3;
    if (ntasks > 1) {
      const long long block = n / 64, lane = n % 64;
      const long long tile = block / ntasks;
      if (t != (int)(block % ntasks)) return 4;            // one chunk per block, chunks in order
      if (i != tile * 64 + lane) return 5;                  // the block's tile, pixels in order
    }
  }
  for (int v : seen) if (v != 1) return 6;
  if (ntasks > 1 && total % 64) return 7;
  if (pad != total - (long long)npix * ntasks) return 8;
  if (ntasks == 1 && pad) return 9;
  return 0;
}

int main() {
  int cases = 0;
  const int sizes[] = {1, 2, 5, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 390, 1000, 4096, 4097};
  for (int npix : sizes)
    for (int ntasks = 1; ntasks <= 8; ntasks++) {
      const int r = check(npix, ntasks);
      if (r) { std::printf("FAIL npix %d ntasks %d: %d\n", npix, ntasks, r); return 1; }
      cases++;
    }
  std::printf("OK %d\n", cases);
  return 0;
}
"""


def test_wave_job_order(tmp_path):
    src, exe = tmp_path / "jobs.cpp", tmp_path / "jobs"
    src.write_text(SRC)
    subprocess.run(["/opt/rocm/llvm/bin/clang++", "-std=c++17", "-DPHX_HOST_BUILD", "-O1", "-Wno-unused-function",
                    "-I", os.path.join(ROOT, "phoenix_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("OK 136")
