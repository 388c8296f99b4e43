// ./amr: the 3-level AMR TGV driver over the C ABI. Flags, stdout lines and exit codes follow
// src/tgv3d_amr.cu:643-907 (0 = VALID, 1 = CHECK, 2 = device error, N % 8 != 0 or pool overflow);
// added: --prec, --diag-every, --bench-steps, --device.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../include/jp.h"
#include "args.hpp"

using namespace jpcli;

static int die(jp_ctx* c, int rc) {
  std::fprintf(stderr, "jp error %d: %s\n", rc, jp_last_error(c));
  jp_destroy(c);
  return 2;
}
static int mode_of(const char* s) {
  if (!std::strcmp(s, "all")) return JP_MODE_ALL;
  if (!std::strcmp(s, "slab")) return JP_MODE_SLAB;
  return JP_MODE_SENSOR;  // the reference treats any other name as the sensor mode (src/tgv3d_amr.cu:779, 794)
}

int main(int argc, char** argv) {
  jp_amr_params p{};
  p.n = argi(argc, argv, "--n", 32);
  p.re = argd(argc, argv, "--re", 800.0);
  p.u0 = argd(argc, argv, "--u0", 0.05);
  const double tstar = argd(argc, argv, "--tstar", 12.0);
  p.frac1 = argd(argc, argv, "--frac1", 0.35);
  p.frac2 = argd(argc, argv, "--frac2", 0.35);
  (void)argi(argc, argv, "--buffer", 0);  // parsed and unused, as in the reference (src/tgv3d_amr.cu:648)
  const char* l1 = args(argc, argv, "--l1", "all");
  const char* l2 = args(argc, argv, "--l2", "all");
  const char* vtk = args(argc, argv, "--vtk", nullptr);
  const bool bench = argd(argc, argv, "--bench", 0.0) > 0.5;
  p.skipprep = argd(argc, argv, "--skipprep", 0.0) > 0.5;
  p.skiprestr = argd(argc, argv, "--skiprestr", 0.0) > 0.5;
  const bool graph = argd(argc, argv, "--graph", 0.0) > 0.5;
  p.adapt = argi(argc, argv, "--adapt", 100);
  p.maxgb = argd(argc, argv, "--maxgb", 0.0);
  p.prec = prec_of(args(argc, argv, "--prec", "fp32"));
  p.device = argi(argc, argv, "--device", 0);
  p.l1_mode = mode_of(l1);
  p.l2_mode = mode_of(l2);
  if (p.n % 8) {
    std::fprintf(stderr, "N%%MB\n");
    return 2;
  }
  if (p.prec < 0) {
    std::fprintf(stderr, "--prec must be fp32, fp16, bf16 or fp64\n");
    return 2;
  }
  const bool dynamic = p.l1_mode == JP_MODE_SENSOR || p.l2_mode == JP_MODE_SENSOR;
  jp_ctx* c = nullptr;
  if (int rc = jp_amr_create(&p, &c)) return die(c, rc);
  if (int rc = jp_init(c)) return die(c, rc);
  jp_derived d;
  jp_get_derived(c, &d);
  jp_counts k;
  jp_get_counts(c, &k);
  const long steps = long(tstar / (p.u0 * d.kf) + 0.5);
  // src/tgv3d_amr.cu:818-819
  std::printf("[jp-solver amr] 3-level 3D. N0=%d N1=%d N2=%d Re=%g L1=%s(%ld blk) L2=%s(%ld blk) tau0=%.4f steps=%ld%s\n",
              p.n, 2 * p.n, 4 * p.n, p.re, l1, k.nf1, l2, k.nf2, d.tau0, steps, vtk ? " (vti+vtp)" : "");
  if (bench) {
    // src/tgv3d_amr.cu:839-864: 20 warm-up steps, min(steps, 600) timed steps
    const long bsteps = argi(argc, argv, "--bench-steps", int(std::min(steps, 600L)));
    jp_bench_result r;
    if (int rc = jp_bench(c, bsteps, 20, graph ? 1 : 0, &r)) return die(c, rc);
    jp_get_counts(c, &k);
    // src/tgv3d_amr.cu:861-862, then the roofline figures on a line of their own
    std::printf("[bench%s] AMR N0=%d(fine %d): %.3f s  work=%.3e updates  %.1f MLUPS(work)  pools %.3f GB  ctx+pools %.3f GB  "
                "(L1=%ld L2=%ld)\n",
                graph ? "+graph" : "", p.n, 4 * p.n, r.ms_total / 1e3, r.work, r.glups * 1e3, r.pool_gb, r.device_used_gb,
                k.nf1, k.nf2);
    std::printf("  %.1f GB/s at %.0f B/update (%.1f%% of 8 TB/s HBM)\n", r.gbps, r.bytes_per_update, r.gbps / 80.0);
    jp_destroy(c);
    return 0;
  }
  // --graph (src/tgv3d_amr.cu:872-881): pairs of coarse steps replayed from captured graphs. The regrids
  // run on device-resident counts, so a regrid needs no recapture; diagnostics and frames are taken
  // between replays.
  if (graph)
    if (int rc = jp_use_graph(c, 1)) return die(c, rc);
  const long sstep = graph ? 2 : 1;
  const long diagK = argi(argc, argv, "--diag-every", int(std::max(1L, steps / 160)));
  const int frames = argi(argc, argv, "--frames", 240);  // src/tgv3d_amr.cu:652
  const long frameK = vtk ? std::max(1L, steps / std::max(1, frames)) : 0;
  int fi = 0;
  double E0 = -1, peak_ens = -1, peak_ts = -1, prevE = 1e30, mass0 = -1, mass_drift = 0;
  bool mono = true;
  for (long s = 0; s <= steps; s += sstep) {
    if (dynamic && s > 0 && s % p.adapt == 0)
      if (int rc = jp_regrid(c)) return die(c, rc);
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
      std::printf("  t*=%.3f E/E0=%.5f enstrophy=%.6e mass_drift=%.2e\n", ts, E / E0, ens, mass_drift);
      std::fflush(stdout);
    }
    if (frameK && s % frameK == 0 && fi < frames) {  // src/tgv3d_amr.cu:889-895
      char nm[512];
      std::snprintf(nm, sizeof nm, "%s_%04d.vti", vtk, fi);
      if (int rc = jp_write_vti(c, nm)) return die(c, rc);
      std::snprintf(nm, sizeof nm, "%s_%04d.vtp", vtk, fi);
      if (int rc = jp_write_boxes(c, nm)) return die(c, rc);
      fi++;
    }
    if (s < steps)
      if (int rc = jp_step(c, sstep)) return die(c, rc);
  }
  if (int rc = jp_sync(c)) return die(c, rc);
  jp_get_counts(c, &k);
  std::printf("  enstrophy peak at t*=%.3f (value %.6e)\n", peak_ts, peak_ens);
  std::printf("  energy monotonic: %s  mass drift %.2e\n", mono ? "yes" : "NO", mass_drift);
  const long ng1 = long(p.n / 4) * (p.n / 4) * (p.n / 4);
  std::printf("  pool high-water: L1 %ld/%ld blocks (%.0f%%), L2 %ld/%ld (%.0f%%)  [full-refine would be %ld/%ld]\n", k.hw_nf1,
              k.cap1, 100.0 * k.hw_nf1 / k.cap1, k.hw_nf2, k.cap2, k.cap2 ? 100.0 * k.hw_nf2 / k.cap2 : 0.0, ng1, ng1 * 8);
  const bool ok = std::isfinite(peak_ens) && peak_ts > 4.0 && peak_ts < 10.0 && mono && mass_drift < 5e-4;
  std::printf("  %s\n", ok ? "VALID (breakdown present, energy monotonic, mass held)" : "CHECK");
  if (vtk) std::printf("  wrote %d frames %s_0000.vti/.vtp ..\n", fi, vtk);
  jp_destroy(c);
  return ok ? 0 : 1;
}
