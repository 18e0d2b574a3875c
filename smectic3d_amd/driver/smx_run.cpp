// smx_run.cpp -- C++ driver over libsmectic: the reference drivers' main()
// (periodicFFT/pf3d.cpp, pf3dMorph.cpp, cosFFT/mesabi/cosNoAdvMorph.cpp,
// cosFFT/comet/cosNoAdvMorphCurv.cpp) with runtime parameters instead of
// compile-time constants, writing the reference's text files.
//
//   smx_run [--preset NAME] [--n N | --nx X --ny Y --nz Z] [--nw NW] [--eps E] [--beta B]
//           [--gamma G] [--alpha A] [--q0 Q] [--dt DT] [--cubic|--quintic] [--neumann]
//           [--fp32] [--advect 0|1] [--force sh|mu] [--wmid W] [--ic noise|hyperboloid|pf3d|morph|dct|file:PATH]
//           [--amp A] [--seed S] [--steps N] [--l1-limit L] [--l1-every K] [--save-every K]
//           [--surf-every K] [--arm A] [--vsign S] [--out DIR] [--max-steps N] [--ic-only]
//           [--raw-ic] [--resume CKPT] [--checkpoint CKPT] [--psi-avg NW]
// --raw-ic writes the initial condition as psiIC.raw (native float64, dump order) instead of
// the text psiIC.dat
// --psi-avg writes psiAvg.dat, the local amplitude field of the final state over windows
// of NW points (cosFFT/comet/morphSurfAmp.cpp:1005-1134), in the psi dump format
// --resume continues bit for bit from an exact binary checkpoint written by --checkpoint
// (at the end of the run) instead of re-bootstrapping from a text dump.
//
// Files (one value per line, `os << v << "\n "`, default 6 significant digits, loop
// order i, j, k as pf3d.cpp:346-352): psiIC.dat, L1.dat (appended every l1-every
// steps, pf3d.cpp:510-513), psi.dat / psiB.dat (alternating on Neumann runs,
// cosNoAdvMorph.cpp:578-589), surfPsi.dat velSurf.dat curvH.dat curvK.dat, and with them
// psiMid.dat, the (x, z) slice at y = ny/2 appended at each surface output
// (cosNoAdvMorphCurv.cpp:981-1001, which fixes j = 256, the middle of its 512-point runs).
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "smectic.h"

namespace {

struct Opts {
  smx_config c;
  std::string ic = "noise", out = ".";
  double amp = 0.01;
  uint64_t seed = 0;
  long steps = -1;         // fixed step count (-1: run to the L1 limit)
  double l1_limit = 1e-6;  // pf3d.cpp:387
  long l1_every = 100;     // pf3d.cpp:489
  long save_every = 100;   // in L1 checks, pf3d.cpp:524
  long surf_every = 0;     // in L1 checks (0: off)
  long max_steps = 1000000;
  bool ic_only = false;  // write psiIC.dat and stop (no device needed)
  bool raw_ic = false;   // write the IC as psiIC.raw (native float64, same order) instead of text
  int psi_avg_nw = 0;    // > 0: write psiAvg.dat of the final state
  std::string resume, checkpoint;  // exact binary checkpoint to continue from / to write at the end
  double arm = -1, vsign = 1.0;
  int nw = 16;
};

double splitmix_uniform(uint64_t seed, uint64_t idx) {
  uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// Initial conditions on the dense [x][y][z] grid.
std::vector<double> make_ic(const Opts& o) {
  const smx_config& c = o.c;
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  const double dx = c.lx / nx, dy = c.ly / ny, dz = c.lz / nz;
  std::vector<double> psi((size_t)nx * ny * nz, 0.0);
  auto at = [&](int i, int j, int k) -> double& { return psi[((size_t)i * ny + j) * nz + k]; };
  if (o.ic == "noise") {
    for (size_t i = 0; i < psi.size(); ++i) psi[i] = o.amp * (2.0 * splitmix_uniform(o.seed, i) - 1.0);
  } else if (o.ic == "hyperboloid") {  // 0.4 cos(q0 sqrt(max(x^2+y^2-z^2,0)+4)) + U[-amp, amp]
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          const double x = (i - nx / 2) * dx, y = (j - ny / 2) * dy, z = (k - nz / 2) * dz;
          const double r2 = std::max(x * x + y * y - z * z, 0.0);
          const size_t id = ((size_t)i * ny + j) * nz + k;
          at(i, j, k) = 0.4 * std::cos(c.q0 * std::sqrt(r2 + 4.0)) + o.amp * (2.0 * splitmix_uniform(o.seed, id) - 1.0);
        }
  } else if (o.ic == "pf3d") {  // pf3d.cpp:302-324 (band 24 < k < 232 at 256, scaled to nz)
    const double Amp = 1.2441, Qi = 0.5;
    const int k0 = 24 * nz / 256, k1 = 232 * nz / 256;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k)
          if (k > k0 && k < k1)
            at(i, j, k) = Amp * std::cos(c.q0 * k * dz) +
                          Amp * 0.1 * std::sin(c.q0 * k * dz) * (std::cos(Qi * i * dx) + std::cos(Qi * j * dy));
  } else if (o.ic == "morph" || o.ic == "dct") {
    // pf3dMorph.cpp:331-383 (layers about the mid plane; aE = 245, bE = 180 at 512,
    // pf3dMorph.cpp:94-96) or the bottom-anchored Neumann variant
    // cosFFT/mesabi/cosNoAdvMorph.cpp:289-332 (zs = k; aE = 312, bE = 376 at 512,
    // cosNoAdvMorph.cpp:81-83); both scaled by nz/512 on other grids.
    const double Amp = 1.328, mid = nz / 2;
    const bool morph = o.ic == "morph";
    const double aE = (morph ? 245.0 : 312.0) * nz / 512, bE = (morph ? 180.0 : 376.0) * nz / 512;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          const bool band = morph ? (k > 62 * nz / 512 && k < 450 * nz / 512) : (k < bE + 1);
          if (!band) continue;
          const double xs = i - mid, ys = j - mid, zs = morph ? k - mid : (double)k;
          const double ds = std::sqrt(xs * xs + ys * ys);
          double v = 0.0;
          if (ds < mid) {
            if (std::sqrt(std::pow((ds - mid) / aE, 2) + std::pow(zs / bE, 2)) <= 1)
              v = Amp * std::cos(c.q0 * dz * std::sqrt(std::pow((bE / aE) * (ds - mid), 2) + zs * zs));
          } else if (std::fabs(zs) < bE) {
            v = Amp * std::cos(c.q0 * zs * dz);
          }
          at(i, j, k) = v;
        }
  } else if (o.ic.rfind("file:", 0) == 0) {  // pf3dMorph.cpp:416-447, checked for length
    std::ifstream in(o.ic.substr(5));
    if (!in) throw std::runtime_error("cannot open " + o.ic.substr(5));
    for (size_t i = 0; i < psi.size(); ++i)
      if (!(in >> psi[i])) throw std::runtime_error("short or malformed IC file " + o.ic.substr(5));
  } else {
    throw std::runtime_error("unknown --ic " + o.ic);
  }
  return psi;
}

void write_field(const std::string& path, const std::vector<double>& psi) {
  std::ofstream os(path);
  if (!os) throw std::runtime_error("cannot write " + path);
  for (double v : psi) os << v << "\n ";  // pf3d.cpp:351
}

// psi[(i*ny + j)*nz + k] for every i, k at j = ny/2, one value a line (cosNoAdvMorphCurv.cpp:989-1001)
void append_mid_slice(const std::string& path, const std::vector<double>& psi, int nx, int ny, int nz) {
  std::ofstream os(path, std::ios_base::app);
  if (!os) throw std::runtime_error("cannot write " + path);
  const int j = ny / 2;
  for (int i = 0; i < nx; ++i)
    for (int k = 0; k < nz; ++k) os << psi[((size_t)i * ny + j) * nz + k] << "\n";
}

void append_line(const std::string& path, double v) {
  std::ofstream os(path, std::ios_base::app);
  os << v << "\n";
}

void preset(Opts& o, const std::string& name, int n) {
  smx_config_default(&o.c, n);
  if (name == "pf3d") return;  // pf3d.cpp:35-75 (advected, quintic, Nw = 16)
  if (name == "pf3dmorph") {   // pf3dMorph.cpp:45-102
    o.c.dt = 0.0005;
    o.c.w_mid = 1.0;
    o.c.mid_plane = n / 2;
    o.ic = "morph";
    o.l1_limit = 1e-4;
    return;
  }
  if (name == "cosnoadv") {  // cosFFT/mesabi/cosNoAdvMorph.cpp:86-112
    o.c.bc = SMX_BC_NEUMANN;
    o.c.advect = 0;
    o.c.dt = 0.0005;
    o.ic = "dct";
    o.l1_limit = 1e-4;
    o.save_every = 10;
    return;
  }
  // cubic-model presets (beta = 0, gamma = 1) with 8 grid points per wavelength
  const double L = n / 8.0 * 2 * M_PI;
  o.c.lx = o.c.ly = o.c.lz = L;
  o.c.beta = 0.0;
  o.c.gamma = 1.0;
  o.c.nl_model = SMX_NL_CUBIC;
  o.c.advect = 0;
  if (name == "plain") { o.c.eps = 0.5; o.c.dt = 0.05; o.ic = "noise"; o.amp = 0.01; o.steps = 1000; }
  else if (name == "hyper") { o.c.eps = 0.3; o.c.dt = 0.1; o.ic = "hyperboloid"; o.amp = 1e-3; o.steps = 2000; }
  else if (name == "neumann-hyper") {
    o.c.bc = SMX_BC_NEUMANN; o.c.eps = 0.5; o.c.dt = 0.05; o.ic = "hyperboloid"; o.amp = 1e-3; o.steps = 4000;
  } else if (name == "advected") {
    o.c.eps = 0.5; o.c.dt = 0.02; o.c.advect = 1; o.c.force_form = SMX_FORCE_MU; o.ic = "noise"; o.amp = 0.05;
    o.steps = 1000;
  } else {
    throw std::runtime_error("unknown preset " + name);
  }
}

int run(int argc, char** argv) {
  Opts o;
  std::string pre = "pf3d";
  int n = 64, nx = -1, ny = -1, nz = -1;
  for (int a = 1; a < argc; ++a)
    if (!std::strcmp(argv[a], "--preset") && a + 1 < argc) pre = argv[a + 1];
    else if (!std::strcmp(argv[a], "--n") && a + 1 < argc) n = std::atoi(argv[a + 1]);
  preset(o, pre, n);
  auto num = [&](int& a) { return std::atof(argv[++a]); };
  for (int a = 1; a < argc; ++a) {
    std::string k = argv[a];
    bool has = a + 1 < argc;
    if (k == "--preset" || k == "--n") { ++a; }
    else if (k == "--nx" && has) nx = (int)num(a);
    else if (k == "--ny" && has) ny = (int)num(a);
    else if (k == "--nz" && has) nz = (int)num(a);
    else if (k == "--nw" && has) o.nw = (int)num(a);
    else if (k == "--eps" && has) o.c.eps = num(a);
    else if (k == "--beta" && has) o.c.beta = num(a);
    else if (k == "--gamma" && has) o.c.gamma = num(a);
    else if (k == "--alpha" && has) o.c.alpha = num(a);
    else if (k == "--q0" && has) o.c.q0 = num(a);
    else if (k == "--dt" && has) o.c.dt = num(a);
    else if (k == "--cubic") o.c.nl_model = SMX_NL_CUBIC;
    else if (k == "--quintic") o.c.nl_model = SMX_NL_QUINTIC;
    else if (k == "--neumann") { o.c.bc = SMX_BC_NEUMANN; o.c.advect = 0; }
    else if (k == "--fp32") o.c.precision = SMX_FP32_MIXED;
    else if (k == "--advect" && has) o.c.advect = (int)num(a);
    else if (k == "--force" && has) o.c.force_form = (std::string(argv[++a]) == "mu") ? SMX_FORCE_MU : SMX_FORCE_SH;
    else if (k == "--wmid" && has) { o.c.w_mid = num(a); }
    else if (k == "--ic" && has) o.ic = argv[++a];
    else if (k == "--amp" && has) o.amp = num(a);
    else if (k == "--seed" && has) o.seed = (uint64_t)std::strtoull(argv[++a], nullptr, 10);
    else if (k == "--steps" && has) o.steps = (long)num(a);
    else if (k == "--l1-limit" && has) o.l1_limit = num(a);
    else if (k == "--l1-every" && has) o.l1_every = (long)num(a);
    else if (k == "--save-every" && has) o.save_every = (long)num(a);
    else if (k == "--surf-every" && has) o.surf_every = (long)num(a);
    else if (k == "--max-steps" && has) o.max_steps = (long)num(a);
    else if (k == "--arm" && has) o.arm = num(a);
    else if (k == "--vsign" && has) o.vsign = num(a);
    else if (k == "--out" && has) o.out = argv[++a];
    else if (k == "--ic-only") o.ic_only = true;
    else if (k == "--raw-ic") o.raw_ic = true;
    else if (k == "--resume" && has) o.resume = argv[++a];
    else if (k == "--checkpoint" && has) o.checkpoint = argv[++a];
    else if (k == "--psi-avg" && has) o.psi_avg_nw = (int)num(a);
    else throw std::runtime_error("unknown or incomplete option " + k);
  }
  if (nx > 0 || ny > 0 || nz > 0) {
    const double L0 = o.c.lx / o.c.nx;  // keep dx
    o.c.nx = nx > 0 ? nx : o.c.nx;
    o.c.ny = ny > 0 ? ny : o.c.ny;
    o.c.nz = nz > 0 ? nz : o.c.nz;
    o.c.lx = L0 * o.c.nx;
    o.c.ly = L0 * o.c.ny;
    o.c.lz = L0 * o.c.nz;
  }
  if (o.c.w_mid != 0.0 && o.c.mid_plane < 0) o.c.mid_plane = o.c.nz / 2;
  if (o.arm < 0) o.arm = (o.c.bc == SMX_BC_NEUMANN) ? 0.7 : 0.1;

  const std::string dir = o.out + "/";
  std::vector<double> psi = make_ic(o);
  if (o.raw_ic) {
    std::ofstream os(dir + "psiIC.raw", std::ios::binary);
    os.write(reinterpret_cast<const char*>(psi.data()), (std::streamsize)(psi.size() * sizeof(double)));
    if (!os) throw std::runtime_error("cannot write " + dir + "psiIC.raw");
  } else {
    write_field(dir + "psiIC.dat", psi);
  }
  if (o.ic_only) return 0;
  smx_handle* h = nullptr;
  if (smx_create(&o.c, &h) != SMX_OK) throw std::runtime_error(std::string("smx_create: ") + smx_last_error());
  std::ofstream(dir + "L1.dat").close();
  if (o.surf_every > 0) std::ofstream(dir + "psiMid.dat").close();  // created empty, :482-486
  if (!o.resume.empty()) {
    if (smx_load_checkpoint(h, o.resume.c_str()) != SMX_OK) throw std::runtime_error(smx_last_error());
  } else if (smx_set_psi(h, psi.data(), o.c.nz) != SMX_OK) {
    throw std::runtime_error(smx_last_error());
  }

  const int64_t ncol = (int64_t)o.c.nx * o.c.ny;
  std::vector<double> sz(ncol), sv(ncol), sh(ncol), sk(ncol);
  const auto t0 = std::chrono::steady_clock::now();
  long done = 0, checks = 0;
  int swt = 0;
  double L1 = 1.0;
  while (true) {
    long chunk = o.l1_every;
    if (o.steps >= 0) chunk = std::min(chunk, o.steps - done);
    if (chunk <= 0 || done >= o.max_steps) break;
    if (smx_step(h, chunk) != SMX_OK) throw std::runtime_error(smx_last_error());
    done += chunk;
    if (smx_l1(h, &L1) != SMX_OK) throw std::runtime_error(smx_last_error());
    append_line(dir + "L1.dat", L1);
    ++checks;
    if (o.save_every > 0 && checks % o.save_every == 0) {
      smx_get_psi(h, psi.data(), o.c.nz);
      const bool alt = (o.c.bc == SMX_BC_NEUMANN);
      write_field(dir + (alt && swt ? "psiB.dat" : "psi.dat"), psi);
      if (alt) swt ^= 1;
    }
    if (o.surf_every > 0 && checks % o.surf_every == 0) {
      smx_surf_opts so = {o.arm, o.vsign};
      int64_t found = 0;
      if (smx_surftrack(h, &so, sz.data(), sv.data(), sh.data(), sk.data(), &found) != SMX_OK)
        throw std::runtime_error(smx_last_error());
      std::ofstream fz(dir + "surfPsi.dat", std::ios_base::app), fv(dir + "velSurf.dat", std::ios_base::app),
          fh(dir + "curvH.dat", std::ios_base::app), fk(dir + "curvK.dat", std::ios_base::app);
      for (int64_t q = 0; q < ncol; ++q) {  // cosNoAdvMorphCurv.cpp:1000-1052 write every column
        fz << sz[q] << "\n";
        fv << sv[q] << "\n";
        fh << sh[q] << "\n";
        fk << sk[q] << "\n";
      }
      if (smx_get_psi(h, psi.data(), o.c.nz) != SMX_OK) throw std::runtime_error(smx_last_error());
      append_mid_slice(dir + "psiMid.dat", psi, o.c.nx, o.c.ny, o.c.nz);
    }
    if (o.steps < 0 && !(L1 > o.l1_limit)) break;  // while (L1 > limL1), pf3d.cpp:390
  }
  smx_sync(h);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (!o.checkpoint.empty() && smx_save_checkpoint(h, o.checkpoint.c_str()) != SMX_OK)
    throw std::runtime_error(smx_last_error());
  smx_get_psi(h, psi.data(), o.c.nz);
  write_field(dir + "psi.dat", psi);
  if (o.psi_avg_nw > 0) {
    std::vector<double> avg(psi.size());
    if (smx_psi_avg(h, o.psi_avg_nw, avg.data(), o.c.nz) != SMX_OK) throw std::runtime_error(smx_last_error());
    write_field(dir + "psiAvg.dat", avg);  // morphSurfAmp.cpp:1120-1130
  }
  double F = 0;
  smx_energy(h, &F);
  std::printf("steps %ld  L1 %.6e  energy %.12e  wall %.3f s  (%.1f steps/s)\n", done, L1, F, secs,
              done / (secs > 0 ? secs : 1));
  smx_destroy(h);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "smx_run: %s\n", e.what());
    return 1;
  }
}
