// ./uniform: the uniform TGV driver over the C ABI. Flags, stdout lines and exit codes follow
// src/tgv3d_uniform.cu:142-203 (0 = VALID, 1 = CHECK, 2 = device error or N % 8 != 0); added: --bench
// (the reference has no uniform bench path), --bench-steps, --prec, --diag-every, --graph, --device.
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "../../include/jp.h"
#include "args.hpp"

using namespace jpcli;

static int die(jp_ctx* c, int rc) {
  std::fprintf(stderr, "jp error %d: %s\n", rc, jp_last_error(c));
  jp_destroy(c);
  return 2;
}

int main(int argc, char** argv) {
  jp_uniform_params p{};
  p.n = argi(argc, argv, "--n", 128);
  p.re = argd(argc, argv, "--re", 1600.0);
  p.u0 = argd(argc, argv, "--u0", 0.05);
  const double tstar = argd(argc, argv, "--tstar", 12.0);
  const char* vtk = args(argc, argv, "--vtk", nullptr);
  const bool bench = argd(argc, argv, "--bench", 0.0) > 0.5;
  const long bsteps = argi(argc, argv, "--bench-steps", 1000);
  const bool graph = argd(argc, argv, "--graph", 0.0) > 0.5;
  p.prec = prec_of(args(argc, argv, "--prec", "fp32"));
  p.device = argi(argc, argv, "--device", 0);
  if (p.n % 8) {
    std::fprintf(stderr, "N%%MB\n");
    return 2;
  }
  if (p.prec < 0) {
    std::fprintf(stderr, "--prec must be fp32, fp16, bf16 or fp64\n");
    return 2;
  }
  jp_ctx* c = nullptr;
  if (int rc = jp_uniform_create(&p, &c)) return die(c, rc);
  jp_derived d;
  jp_get_derived(c, &d);
  const long steps = long(tstar / (p.u0 * d.kf) + 0.5);
  const int nb = (p.n / 8) * (p.n / 8) * (p.n / 8);
  std::printf("[jp-solver uniform] TGV3D N=%d MB=8 blocks=%d Re=%g nu=%g tau=%.4f steps=%ld t*max=%.1f%s\n", p.n, nb, p.re,
              d.nu, d.tau0, steps, tstar, vtk ? "  (writing vti)" : "");
  if (int rc = jp_init(c)) return die(c, rc);
  if (bench) {
    jp_bench_result r;
    if (int rc = jp_bench(c, bsteps, 20, graph ? 1 : 0, &r)) return die(c, rc);
    int shape[4];
    jp_step_pair_shape(c, shape);
    std::printf("[bench%s] uniform N=%d: %.3f s  work=%.3e updates  %.3f GLUPS  %.1f GB/s at %.0f B/update (%.1f%% of 8 TB/s "
                "HBM)  pools %.3f GB  %s\n",
                graph ? "+graph" : "", p.n, r.ms_total / 1e3, r.work, r.glups, r.gbps, r.bytes_per_update, r.gbps / 80.0,
                r.pool_gb, shape[0] ? "(two fused steps per launch: the intermediate step stays on chip)" : "(one step per launch)");
    jp_destroy(c);
    return 0;
  }
  if (graph) jp_use_graph(c, 1);
  const long diagK = argi(argc, argv, "--diag-every", int(std::max(1L, steps / 240)));
  const int frames = argi(argc, argv, "--frames", 160);  // src/tgv3d_uniform.cu:146, 165
  const long frameK = vtk ? std::max(1L, steps / std::max(1, frames)) : 0;
  int fi = 0;
  double E0 = -1, peak_ens = -1, peak_ts = -1, prevE = 1e30, mass0 = -1, mass_drift = 0;
  bool mono = true;
  long s = 0;
  while (true) {
    if (s % diagK == 0) {
      double o[3];
      if (int rc = jp_diag(c, o)) return die(c, rc);
      const double E = o[0], ens = o[1], mass = o[2], ts = double(s) * p.u0 * d.kf;
      if (mass0 < 0) mass0 = mass;
      mass_drift = std::max(mass_drift, std::fabs(mass - mass0) / mass0);
      if (E0 < 0) E0 = E;
      if (E > prevE * 1.0000001) mono = false;
      prevE = E;
      if (ens > peak_ens) {
        peak_ens = ens;
        peak_ts = ts;
      }
      std::printf("  t*=%.3f  E/E0=%.5f  enstrophy=%.6e  mass_drift=%.2e\n", ts, E / E0, ens, mass_drift);
      std::fflush(stdout);
    }
    if (frameK && s % frameK == 0 && fi < frames) {
      char nm[512];
      std::snprintf(nm, sizeof nm, "%s_%04d.vti", vtk, fi++);
      if (int rc = jp_write_vti(c, nm)) return die(c, rc);
    }
    if (s >= steps) break;
    long next = std::min(steps, (s / diagK + 1) * diagK);
    if (frameK) next = std::min(next, (s / frameK + 1) * frameK);
    if (int rc = jp_step(c, next - s)) return die(c, rc);
    s = next;
  }
  if (int rc = jp_sync(c)) return die(c, rc);
  std::printf("  enstrophy peak at t*=%.3f (value %.6e)\n", peak_ts, peak_ens);
  std::printf("  energy monotonic: %s   mass drift %.2e\n", mono ? "yes" : "NO", mass_drift);
  const bool ok = std::isfinite(peak_ens) && peak_ts > 4.0 && peak_ts < 10.0 && mono && mass_drift < 5e-4;
  std::printf("  %s\n", ok ? "VALID (enstrophy peaks then falls, energy monotonic, mass held in FP32)" : "CHECK (see values)");
  if (vtk) std::printf("  wrote %d vti frames %s_0000.vti ..\n", fi, vtk);
  jp_destroy(c);
  return ok ? 0 : 1;
}
