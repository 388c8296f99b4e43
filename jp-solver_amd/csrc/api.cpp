// C-ABI layer (include/jp.h): contexts, device memory, orchestration of the time step, diagnostics,
// benchmarking and the test hooks. Host code only; kernels live in the .hip translation units.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/jp.h"
#include "ctx.hpp"

using namespace jp;

namespace jp {

int fail(jp_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int ctx_alloc(jp_ctx* c, void** p, size_t bytes, const char* what) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(c, e == hipErrorOutOfMemory ? JP_ERR_NOMEM : JP_ERR_HIP,
                std::string("hipMalloc(") + what + ", " + std::to_string(bytes) + " B): " + hipGetErrorString(e));
  }
  c->allocs.emplace_back(*p, bytes);
  c->pool_bytes += bytes;
  return JP_OK;
}

void ctx_free(jp_ctx* c, void* p) {
  if (!p) return;
  for (auto it = c->allocs.begin(); it != c->allocs.end(); ++it)
    if (it->first == p) {
      c->pool_bytes -= it->second;
      c->allocs.erase(it);
      break;
    }
  (void)hipFree(p);
}

// Population layout of a uniform lattice by default, from an A/B on MI355X at 512^3 (the same results bit
// for bit): z-layers (the 19 directions of a z-layer adjacent) for FP32 (fused pairs 60.5-62.5 against
// 58.2-61.9 GLUPS in direction planes, single steps even), FP16 (73.1-73.5 against 70.3-72.0) and BF16
// (68.7 against 67.9); direction planes for FP64 (17.9-18.3 against 15.3-16.0).
bool default_layers(int prec) { return prec != JP_PREC_FP64; }
// Elements after each direction row of a z-layer. 32 where the lattice runs fused pairs (FP32, FP16 or
// BF16 with a column shape that divides N): a sweep of the FP32 pad at 512^3 (128 x 4 / 64 x 8, GLUPS) gave 0:
// 58.8 / 51.4, 16: 60.5 / 59.7, 32: 64.6 / 64.2, 48: 61.4 / 60.8, 64-160: 63.5-64.2, 256: 62.4 / 52.0,
// 512-2048: 56.6-61.5 / 49.5-51.5 (the FP16 pair: 83.0 against 82.7 without a pad), while the single
// steps ran faster without a pad (FP32 34.9 against 34.3, FP16 80.2 against 75.4 in the same process), so
// every other lattice keeps 0.
uint64_t default_layer_pad(int prec, int N) {
  int shape[4];
  return step_pair_shape(prec, N, -1, shape) ? 32 : 0;
}

double default_scale(double u0) { return std::exp2(std::floor(std::log2(2048.0 / u0))); }
// Dense-step launch variant per storage form, from the 512^3 sweeps on MI355X: one cell per lane with
// cached loads; 512-lane groups and streaming stores for FP32 and the 16-bit forms (at 512^3 a group is one
// x-row, so the loads take scalar row bases: FP32 37.2 GLUPS against 34.0 for 8 cells per lane, FP16 79.1
// against 73.2 for 1024-lane groups), 256-lane groups and cached stores for FP64.
int default_variant(int prec) { return prec == 3 ? 12 : 6; }

int drop_graph(jp_ctx* c) {
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
  c->graphs.clear();
  return JP_OK;
}

}  // namespace jp

// ---------------------------------------------------------------------------------------------------
// uniform
// ---------------------------------------------------------------------------------------------------
static int uniform_step_once(jp_ctx* c) {
  const float enc = float(c->d.scale[0]), dec = float(1.0 / c->d.scale[0]);
  HCK(c, dense_step(c->prec, c->c0a, c->c0b, c->g0, c->d.omega0, enc, dec, c->variant, c->st));
  std::swap(c->c0a, c->c0b);
  return JP_OK;
}

// Partial sums of the diagnostic reductions: `slots` x 3 doubles, followed by the 3 results.
static int ensure_partials(jp_ctx* c, size_t slots) {
  if (c->partial_slots >= slots) return JP_OK;
  ctx_free(c, c->partial);
  c->partial = nullptr;
  c->partial_slots = 0;
  void* p;
  if (int r = ctx_alloc(c, &p, slots * 3 * 8 + 64, "diagnostic partials")) return r;
  c->partial = static_cast<double*>(p);
  c->partial_slots = slots;
  return JP_OK;
}

static int ensure_diag_buffers(jp_ctx* c, int nfield) {
  const size_t fb = size_t(nfield) * nfield * nfield * 4 * (c->prec == JP_PREC_FP64 ? 8 : 4);
  if (c->field_bytes < fb) {
    ctx_free(c, c->field);
    c->field = nullptr;
    c->field_bytes = 0;
    if (int r = ctx_alloc(c, &c->field, fb, "diagnostic field")) return r;
    c->field_bytes = fb;
  }
  return ensure_partials(c, kDiagBlocks);
}

extern "C" {

const char* jp_version(void) { return "jp-solver-mi355x 0.1 (gfx950)"; }

int jp_uniform_create(const jp_uniform_params* p, jp_ctx** out) {
  if (!p || !out) return JP_ERR_ARG;
  *out = nullptr;
  auto* c = new jp_ctx();
  c->kind = 0;
  if (p->n < 8 || p->n % MB != 0 || p->n > 1280) {
    *out = c;
    return fail(c, JP_ERR_ARG, "N%MB: n must be a positive multiple of 8 (and at most 1280)");
  }
  if (p->prec < 0 || p->prec > 3 || !(p->u0 > 0) || !(p->re > 0)) {
    *out = c;
    return fail(c, JP_ERR_ARG, "bad precision, u0 or Re");
  }
  c->device = p->device;
  c->prec = p->prec;
  c->u0 = p->u0;
  c->re = p->re;
  for (int i = 0; i < 3; ++i) c->phase[i] = p->phase[i];
  *out = c;
  HCK(c, hipSetDevice(c->device));
  HCK(c, hipStreamCreate(&c->st));
  const int N = p->n;
  // src/tgv3d_uniform.cu:151-152
  c->d.n = N;
  c->d.prec = p->prec;
  c->d.nu = p->u0 * double(N) / (2.0 * M_PI * p->re);
  c->d.tau0 = 3.0 * c->d.nu + 0.5;
  c->d.omega0 = (p->prec == JP_PREC_FP64) ? 1.0 / c->d.tau0 : double(float(1.0 / c->d.tau0));
  c->d.kf = 2.0 * M_PI / double(N);
  c->d.cells = long(N) * N * N;
  c->variant = default_variant(c->prec);
  const double s = (p->prec == JP_PREC_FP16 || p->prec == JP_PREC_BF16) ? (p->scale > 0 ? p->scale : default_scale(p->u0)) : 1.0;
  if (s != std::exp2(std::round(std::log2(s)))) return fail(c, JP_ERR_ARG, "scale must be a power of two");
  c->d.scale[0] = c->d.scale[1] = c->d.scale[2] = s;
  if (p->plane_pad < -1 || (p->plane_pad > 0 && p->plane_pad % 64 != 0))
    return fail(c, JP_ERR_ARG, "plane_pad must be a non-negative multiple of 64, 0 (default) or -1 (none)");
  // 0: the default layout; -1 or a pad: direction planes with that many elements appended to each
  c->g0 = p->plane_pad == 0 && default_layers(c->prec) ? make_dense_layers(N, default_layer_pad(c->prec, N))
                                                        : make_dense(N, p->plane_pad > 0 ? uint64_t(p->plane_pad) : 0);
  const size_t bytes = size_t(c->g0.elems) * prec_bytes(c->prec);
  if (int r = ctx_alloc(c, &c->c0a, bytes, "populations a")) return r;
  if (int r = ctx_alloc(c, &c->c0b, bytes, "populations b")) return r;
  return JP_OK;
}

int jp_amr_create(const jp_amr_params* p, jp_ctx** out) {
  if (!p || !out) return JP_ERR_ARG;
  auto* c = new jp_ctx();
  *out = c;
  c->kind = 1;
  if (p->n < 8 || p->n % MB != 0 || p->n > 320)
    return fail(c, JP_ERR_ARG, "N%MB: n must be a positive multiple of 8 (and at most 320)");
  if (p->prec < 0 || p->prec > 3 || !(p->u0 > 0) || !(p->re > 0))
    return fail(c, JP_ERR_ARG, "bad precision, u0 or Re");
  if (p->l1_mode < 0 || p->l1_mode > 2 || p->l2_mode < 0 || p->l2_mode > 2)
    return fail(c, JP_ERR_ARG, "refinement mode must be all, slab or sensor");
  if (!(p->frac1 > 0 && p->frac1 <= 1) || !(p->frac2 > 0 && p->frac2 <= 1))
    return fail(c, JP_ERR_ARG, "frac1 and frac2 must lie in (0, 1]");
  c->device = p->device;
  c->prec = p->prec;
  c->u0 = p->u0;
  c->re = p->re;
  for (int i = 0; i < 3; ++i) c->phase[i] = p->phase[i];
  HCK(c, hipSetDevice(c->device));
  HCK(c, hipStreamCreate(&c->st));
  return amr_create(p, c);
}

int jp_init_with_sensor(jp_ctx* c, const float* sen0, const float* sen1_by_pos) {
  if (!c) return JP_ERR_ARG;
  if (c->kind != 1) return fail(c, JP_ERR_STATE, "not an AMR context");
  HCK(c, hipSetDevice(c->device));
  return amr_init_with(c, sen0, sen1_by_pos);
}

int jp_regrid_with_sensor(jp_ctx* c, const float* sen0, const float* sen1_by_pos) {
  if (!c) return JP_ERR_ARG;
  if (c->kind == 2 && !sen0 && !sen1_by_pos) {
    if (!c->initialised) return fail(c, JP_ERR_STATE, "jp_regrid before jp_init");
    HCK(c, hipSetDevice(c->device));
    return ens_regrid(c);
  }
  if (c->kind != 1) return fail(c, JP_ERR_STATE, "not an AMR context");
  if (!c->initialised) return fail(c, JP_ERR_STATE, "jp_regrid before jp_init");
  HCK(c, hipSetDevice(c->device));
  return amr_regrid_with(c, sen0, sen1_by_pos);
}

int jp_regrid(jp_ctx* c) { return jp_regrid_with_sensor(c, nullptr, nullptr); }

int jp_get_counts(jp_ctx* c, jp_counts* out) {
  if (!c || !out) return JP_ERR_ARG;
  if (c->kind != 1) return fail(c, JP_ERR_STATE, "not an AMR context");
  return amr_counts(c, out);
}

int jp_get_table(jp_ctx* c, int which, void* dst, long cap, long* n) {
  if (!c || (!dst && cap > 0)) return JP_ERR_ARG;
  if (c->kind != 1) return fail(c, JP_ERR_STATE, "not an AMR context");
  HCK(c, hipSetDevice(c->device));
  return amr_table(c, which, dst, cap, n);
}

int jp_ensemble_create(const jp_member* members, int count, int prec, double maxgb, int device, jp_ctx** out) {
  if (!members || !out || count < 1) return JP_ERR_ARG;
  auto* c = new jp_ctx();
  *out = c;
  c->kind = 2;
  if (prec < 0 || prec > 3) return fail(c, JP_ERR_ARG, "bad precision");
  c->prec = prec;
  c->device = device;
  HCK(c, hipSetDevice(device));
  HCK(c, hipStreamCreate(&c->st));
  return ens_create(members, count, prec, maxgb, device, c);
}

int jp_ensemble_dense_shape(jp_ctx* ens, int set, int* current) {
  if (!ens) return JP_ERR_ARG;
  if (ens->kind != 2) return fail(ens, JP_ERR_STATE, "not an ensemble");
  if (!ens->initialised) return fail(ens, JP_ERR_STATE, "jp_ensemble_dense_shape before jp_init");
  return ens_dense_shape(ens, set, current);
}

int jp_ensemble_member(jp_ctx* ens, int index, jp_ctx** member) {
  if (!ens || !member) return JP_ERR_ARG;
  if (ens->kind != 2) return fail(ens, JP_ERR_STATE, "not an ensemble");
  *member = ens_member(ens, index);
  return *member ? JP_OK : fail(ens, JP_ERR_ARG, "member index out of range");
}

// Uniform lattice: times two steps in every form the lattice can take (single steps, and the fused pair
// with each column shape that divides the edge), each launch re-run from the current state into the other
// buffer so that the state is unchanged, and keeps the fastest (median of 3 after one untimed launch).
// Single-step launch variants the uniform autotuner times (jp_set_variant bits): cached loads; 256, 512 or
// 1024 lanes per workgroup; 1, 2 or 8 consecutive x-cells per lane; streaming or cached stores. Lanes per
// workgroup and registers per lane fix the waves a CU holds.
static const int kTuneVariants[] = {4, 6, 7, 12, 14, 15, 20, 22, 23, 28, 30, 31, 52, 54, 55, 60, 62, 63};

static int uniform_autotune(jp_ctx* c, int* out) {
  const float enc = float(c->d.scale[0]), dec = float(1.0 / c->d.scale[0]);
  hipEvent_t a, b;
  HCK(c, hipEventCreate(&a));
  HCK(c, hipEventCreate(&b));
  struct Cand {
    int pick, variant;  // pick >= 0: fused pair of column form pick; else single steps of launch variant
  };
  std::vector<Cand> cands;
  for (int v : kTuneVariants) cands.push_back({-1, v});
  for (int p = 0; p < step_pair_shape_count(); ++p) {
    int s[4];
    if (step_pair_shape(c->prec, c->d.n, p, s)) cands.push_back({p, c->variant});
  }
  // two steps of a candidate from the current state into the other buffer: the state is unchanged
  auto launch = [&](const Cand& k) -> hipError_t {
    if (k.pick >= 0) return dense_step_pair(c->prec, c->c0a, c->c0b, c->g0, c->d.omega0, enc, dec, k.pick, c->st);
    for (int r = 0; r < 2; ++r)
      if (hipError_t e = dense_step(c->prec, c->c0a, c->c0b, c->g0, c->d.omega0, enc, dec, k.variant, c->st)) return e;
    return hipSuccess;
  };
  float best = 1e30f, best_single = 1e30f;
  Cand keep{-1, c->variant};
  int single_variant = c->variant;
  for (const Cand& k : cands) {
    HCK(c, launch(k));
    float t[3];
    for (int r = 0; r < 3; ++r) {
      HCK(c, hipEventRecord(a, c->st));
      HCK(c, launch(k));
      HCK(c, hipEventRecord(b, c->st));
      HCK(c, hipEventSynchronize(b));
      HCK(c, hipEventElapsedTime(&t[r], a, b));
    }
    std::sort(t, t + 3);
    if (t[1] < best) {
      best = t[1];
      keep = k;
    }
    if (k.pick < 0 && t[1] < best_single) {
      best_single = t[1];
      single_variant = k.variant;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  c->step_pairs = keep.pick >= 0;
  c->pair_pick = keep.pick;
  c->variant = single_variant;  // also runs the odd step after the pairs
  drop_graph(c);
  if (out) {
    int shape[4] = {0, 0, 0, 0};
    if (keep.pick >= 0) step_pair_shape(c->prec, c->d.n, keep.pick, shape);
    for (int i = 0; i < 4; ++i) out[i] = shape[i];
    out[4] = c->variant;
    out[5] = dense_step_waves_per_cu(c->prec, c->g0, c->variant, keep.pick);
    out[6] = keep.pick >= 0 ? 2 : 1;  // steps per launch
  }
  return JP_OK;
}

int jp_autotune(jp_ctx* c, int* groups) {
  if (!c) return JP_ERR_ARG;
  if (c->kind == 1) return fail(c, JP_ERR_STATE, "autotuning applies to the batched ensemble and the uniform lattice");
  if (!c->initialised) return fail(c, JP_ERR_STATE, "jp_autotune before jp_init");
  HCK(c, hipSetDevice(c->device));
  return c->kind == 2 ? ens_autotune(c, groups) : uniform_autotune(c, groups);
}

int jp_init(jp_ctx* c) {
  if (!c) return JP_ERR_ARG;
  HCK(c, hipSetDevice(c->device));
  if (c->kind == 2) return ens_init(c);
  if (c->kind == 1) return amr_init(c);
  HCK(c, dense_init(c->prec, c->c0a, c->g0, c->u0, c->d.kf, c->phase, float(c->d.scale[0]), c->st));
  HCK(c, hipStreamSynchronize(c->st));
  c->initialised = true;
  return JP_OK;
}

// n uniform steps: two fused steps per launch where a column shape applies, a single step for an odd rest.
static int uniform_steps(jp_ctx* c, long n) {
  int shape[4];
  long s = 0;
  if (c->step_pairs && step_pair_shape(c->prec, c->d.n, c->pair_pick, shape)) {
    const float enc = float(c->d.scale[0]), dec = float(1.0 / c->d.scale[0]);
    for (; s + 2 <= n; s += 2) {
      HCK(c, dense_step_pair(c->prec, c->c0a, c->c0b, c->g0, c->d.omega0, enc, dec, c->pair_pick, c->st));
      std::swap(c->c0a, c->c0b);
    }
  }
  for (; s < n; ++s)
    if (int r = uniform_step_once(c)) return r;
  return JP_OK;
}

static int eager_steps(jp_ctx* c, long n) {
  if (c->kind == 0) return uniform_steps(c, n);
  for (long s = 0; s < n; ++s) {
    int r = c->kind == 2 ? ens_step_once(c) : c->kind == 1 ? amr_base_step(c, c->st) : uniform_step_once(c);
    if (r) return r;
  }
  return JP_OK;
}

// The buffers a run of steps starts from: after an even number of steps every fine ping-pong pair is back
// in its role, so a captured run applies exactly when the current roles equal those it was captured with.
static void graph_key(const jp_ctx* c, const void* key[3]) {
  if (c->kind == 2) {
    ens_graph_key(c, key);
    return;
  }
  key[0] = c->c0a;
  key[1] = c->kind == 1 ? c->amr.f1a : nullptr;
  key[2] = c->kind == 1 ? c->amr.f2a : nullptr;
}

// The captured run of `len` steps (even) for the current roles, captured on first use. A few runs are
// kept: per length, the two parities of the dense buffers and, with device-resident regrids (which swap
// both fine pairs), of the fine buffers. Capturing leaves the dense roles as the captured steps leave
// them; returning a kept run sets them the same way, before its replay is enqueued.
constexpr long kGraphSteps = 16;  // longest captured run

static int graph_for(jp_ctx* c, long len, hipGraphExec_t* out) {
  const void* key[3];
  graph_key(c, key);
  for (const auto& g : c->graphs)
    if (g.len == len && g.key[0] == key[0] && g.key[1] == key[1] && g.key[2] == key[2]) {
      *out = g.exec;
      c->c0a = g.c0[0];
      c->c0b = g.c0[1];
      return JP_OK;
    }
  if (c->graphs.size() >= 12) drop_graph(c);
  hipGraph_t g;
  HCK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  const int rc = eager_steps(c, len);
  hipError_t ec = hipStreamEndCapture(c->st, &g);
  if (rc) return rc;
  if (ec != hipSuccess) return fail(c, JP_ERR_HIP, std::string("capture: ") + hipGetErrorString(ec));
  hipGraphExec_t exec = nullptr;
  ec = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ec != hipSuccess) return fail(c, JP_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ec));
  c->graphs.push_back({{key[0], key[1], key[2]}, len, exec, {c->c0a, c->c0b}});
  HCK(c, hipGraphUpload(exec, c->st));
  *out = exec;
  return JP_OK;
}

int jp_step(jp_ctx* c, long n) {
  if (!c) return JP_ERR_ARG;
  if (!c->initialised) return fail(c, JP_ERR_STATE, "jp_step before jp_init");
  if (c->kind == 2)
    if (int r = ens_check(c)) return r;
  if (!c->use_graph) return eager_steps(c, n);
  // replays of captured runs of up to kGraphSteps steps (the gap between two replays costs more than
  // the launches inside one), then a single eager step for an odd rest
  long rem = n;
  while (rem >= 2) {
    const long len = std::min(rem & ~1L, kGraphSteps);
    hipGraphExec_t g;
    if (int r = graph_for(c, len, &g)) return r;
    HCK(c, hipGraphLaunch(g, c->st));
    rem -= len;
  }
  return rem ? eager_steps(c, 1) : JP_OK;
}

int jp_set_variant(jp_ctx* c, int variant) {
  if (!c) return JP_ERR_ARG;
  if (variant < 0 || variant > 63) return fail(c, JP_ERR_ARG, "variant must lie in [0, 63]");
  c->variant = variant;
  drop_graph(c);
  return JP_OK;
}

int jp_set_step_pairs(jp_ctx* c, int on) {
  if (!c) return JP_ERR_ARG;
  c->step_pairs = on != 0;
  c->pair_pick = -1;
  drop_graph(c);
  return JP_OK;
}

int jp_set_step_pair_shape(jp_ctx* c, int pick) {
  if (!c) return JP_ERR_ARG;
  if (c->kind != 0) return fail(c, JP_ERR_STATE, "column shapes apply to a uniform context");
  int shape[4];
  if (pick < -1 || pick >= step_pair_shape_count() || !step_pair_shape(c->prec, c->d.n, pick, shape))
    return fail(c, JP_ERR_ARG, "no such column shape for this lattice and storage form");
  c->step_pairs = true;
  c->pair_pick = pick;
  drop_graph(c);
  return JP_OK;
}

int jp_step_pair_shape(const jp_ctx* c, int shape[4]) {
  if (!c || !shape) return JP_ERR_ARG;
  for (int i = 0; i < 4; ++i) shape[i] = 0;
  if (c->kind == 0 && c->step_pairs) step_pair_shape(c->prec, c->d.n, c->pair_pick, shape);
  return JP_OK;
}

int jp_steps_per_launch(const jp_ctx* c, int* steps) {
  if (!c || !steps) return JP_ERR_ARG;
  int shape[4];
  *steps = c->kind == 0 && c->step_pairs && step_pair_shape(c->prec, c->d.n, c->pair_pick, shape) ? 2 : 1;
  return JP_OK;
}

int jp_use_graph(jp_ctx* c, int on) {
  if (!c) return JP_ERR_ARG;
  c->use_graph = on != 0;
  if (!c->use_graph) drop_graph(c);
  if (c->kind == 1) {
    HCK(c, hipSetDevice(c->device));
    return amr_set_resident(c, c->use_graph);
  }
  return JP_OK;
}

int jp_diag(jp_ctx* c, double out[3]) {
  if (!c || !out) return JP_ERR_ARG;
  if (c->kind == 2) return fail(c, JP_ERR_STATE, "diagnostics are per member (jp_ensemble_member)");
  if (!c->initialised) return fail(c, JP_ERR_STATE, "jp_diag before jp_init");
  HCK(c, hipSetDevice(c->device));
  if (c->kind == 1)  // a diagnostic point also reports a device-resident regrid that overflowed a pool
    if (int r = amr_sync_counts(c)) return r;
  double* res;
  if (c->kind == 1) {  // composite on the finest grid, then the reduction over it (src/tgv3d_amr.cu:871, 883)
    const int nf = c->amr.N2;
    if (int r = ensure_diag_buffers(c, nf)) return r;
    if (int r = amr_composite(c, c->field)) return r;
    res = c->partial + c->partial_slots * 3;
    HCK(c, field_diag(c->prec == JP_PREC_FP64, false, c->field, nf, c->partial, kDiagBlocks, res, c->st));
  } else {  // one fused pass over the populations, no macroscopic field
    if (int r = ensure_partials(c, dense_diag_items(c->d.n))) return r;
    res = c->partial + c->partial_slots * 3;
    HCK(c, dense_diag(c->prec, c->c0a, c->g0, float(1.0 / c->d.scale[0]), c->partial, res, c->st));
  }
  HCK(c, hipMemcpyAsync(out, res, 24, hipMemcpyDeviceToHost, c->st));
  HCK(c, hipStreamSynchronize(c->st));
  return JP_OK;
}

int jp_diag_noalloc(jp_ctx* c, double out[3]) {
  if (!c || !out) return JP_ERR_ARG;
  if (c->kind != 0) return fail(c, JP_ERR_STATE, "jp_diag_noalloc takes a uniform context (AMR, ensembles: jp_diag)");
  if (!c->initialised) return fail(c, JP_ERR_STATE, "jp_diag_noalloc before jp_init");
  HCK(c, hipSetDevice(c->device));
  double* sp = nullptr;
  size_t slots = 0;
  HCK(c, noalloc_scratch(c->device, &sp, &slots));
  const uint32_t items = dense_diag_items(c->d.n);
  if (items > slots)
    return fail(c, JP_ERR_ARG, "jp_diag_noalloc: the lattice needs " + std::to_string(items) +
                                   " partial sums, the fixed scratch array holds " + std::to_string(slots));
  double* res = sp + 3 * slots;
  HCK(c, dense_diag(c->prec, c->c0a, c->g0, float(1.0 / c->d.scale[0]), sp, res, c->st));
  HCK(c, hipMemcpyAsync(out, res, 24, hipMemcpyDeviceToHost, c->st));
  HCK(c, hipStreamSynchronize(c->st));
  return JP_OK;
}

long jp_field_edge(jp_ctx* c) {
  if (!c || c->kind == 2) return 0;
  return c->kind == 1 ? c->amr.N2 : c->d.n;
}

int jp_field(jp_ctx* c, float* Q, float* speed, float* vel, long n3) {
  if (!c || !Q || !speed) return JP_ERR_ARG;
  if (c->kind == 2) return fail(c, JP_ERR_STATE, "output fields are per member (jp_ensemble_member)");
  if (!c->initialised) return fail(c, JP_ERR_STATE, "jp_field before jp_init");
  HCK(c, hipSetDevice(c->device));
  const int nf = int(jp_field_edge(c));
  const size_t cells = size_t(nf) * nf * nf;
  if (n3 < long(cells)) return fail(c, JP_ERR_ARG, "output arrays hold fewer than edge^3 entries");
  if (int r = ensure_diag_buffers(c, nf)) return r;
  const size_t need = cells * 5 * sizeof(float);
  if (c->out_bytes < need) {
    ctx_free(c, c->out);
    c->out = nullptr;
    c->out_bytes = 0;
    void* p = nullptr;
    if (int r = ctx_alloc(c, &p, need, "output field")) return r;
    c->out = static_cast<float*>(p);
    c->out_bytes = need;
  }
  if (c->kind == 1) {
    if (int r = amr_composite(c, c->field)) return r;
  } else {
    HCK(c, dense_macro(c->prec, c->c0a, c->g0, float(1.0 / c->d.scale[0]), c->field, c->st));
  }
  float *dQ = c->out, *dS = c->out + cells, *dV = c->out + 2 * cells;
  HCK(c, field_qspeed(c->prec == JP_PREC_FP64, c->kind == 0, c->field, nf, dQ, dS, vel ? dV : nullptr, c->st));
  HCK(c, hipMemcpyAsync(Q, dQ, cells * 4, hipMemcpyDeviceToHost, c->st));
  HCK(c, hipMemcpyAsync(speed, dS, cells * 4, hipMemcpyDeviceToHost, c->st));
  if (vel) HCK(c, hipMemcpyAsync(vel, dV, cells * 12, hipMemcpyDeviceToHost, c->st));
  HCK(c, hipStreamSynchronize(c->st));
  return JP_OK;
}

int jp_get_derived(jp_ctx* c, jp_derived* d) {
  if (!c || !d) return JP_ERR_ARG;
  *d = c->d;
  return JP_OK;
}

int jp_population_bytes(jp_ctx* c, int level, size_t* bytes) {
  if (!c || !bytes) return JP_ERR_ARG;
  if (c->kind == 2) return fail(c, JP_ERR_STATE, "populations are per member (jp_ensemble_member)");
  if (c->kind == 1)
    if (int r = amr_sync_counts(c)) return r;
  long tiles;
  if (level == 0) tiles = c->d.cells / MB3;
  else if (c->kind == 1 && level == 1) tiles = c->amr.nf1;
  else if (c->kind == 1 && level == 2) tiles = c->amr.nf2;
  else return fail(c, JP_ERR_ARG, "bad level");
  *bytes = size_t(tiles) * Q * MB3 * prec_bytes(c->prec);
  return JP_OK;
}

static int level_buffer(jp_ctx* c, int level, void** p) {
  if (level == 0) *p = c->c0a;
  else if (c->kind == 1 && level == 1) *p = c->amr.f1a;
  else if (c->kind == 1 && level == 2) *p = c->amr.f2a;
  else return fail(c, JP_ERR_ARG, "bad level");
  return JP_OK;
}

int jp_get_populations(jp_ctx* c, int level, void* dst, size_t bytes) {
  if (!c || !dst) return JP_ERR_ARG;
  size_t need;
  if (int r = jp_population_bytes(c, level, &need)) return r;
  if (bytes != need) return fail(c, JP_ERR_ARG, "population buffer size mismatch");
  HCK(c, hipSetDevice(c->device));
  void* src;
  if (int r = level_buffer(c, level, &src)) return r;
  if (level == 0 && c->kind == 0) {  // the uniform lattice's planes or z-layers; AMR level 0 is tile-major
    void* tmp;
    if (int r = ctx_alloc(c, &tmp, need, "tile staging")) return r;
    hipError_t e = dense_to_tiles(c->prec, src, c->g0, tmp, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(dst, tmp, need, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    ctx_free(c, tmp);
    if (e != hipSuccess) return fail(c, JP_ERR_HIP, hipGetErrorString(e));
    return JP_OK;
  }
  if (need) {
    HCK(c, hipMemcpyAsync(dst, src, need, hipMemcpyDeviceToHost, c->st));
    HCK(c, hipStreamSynchronize(c->st));
  }
  return JP_OK;
}

int jp_set_populations(jp_ctx* c, int level, const void* srcp, size_t bytes) {
  if (!c || !srcp) return JP_ERR_ARG;
  size_t need;
  if (int r = jp_population_bytes(c, level, &need)) return r;
  if (bytes != need) return fail(c, JP_ERR_ARG, "population buffer size mismatch");
  HCK(c, hipSetDevice(c->device));
  void* dstb;
  if (int r = level_buffer(c, level, &dstb)) return r;
  drop_graph(c);
  if (level == 0 && c->kind == 0) {
    void* tmp;
    if (int r = ctx_alloc(c, &tmp, need, "tile staging")) return r;
    hipError_t e = hipMemcpyAsync(tmp, srcp, need, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = tiles_to_dense(c->prec, tmp, c->g0, dstb, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    ctx_free(c, tmp);
    if (e != hipSuccess) return fail(c, JP_ERR_HIP, hipGetErrorString(e));
  } else if (need) {
    HCK(c, hipMemcpyAsync(dstb, srcp, need, hipMemcpyHostToDevice, c->st));
    HCK(c, hipStreamSynchronize(c->st));
  }
  c->initialised = true;
  return JP_OK;
}

int jp_bench(jp_ctx* c, long steps, long warmup, int use_graph, jp_bench_result* res) {
  if (!c || !res || steps < 1 || warmup < 0) return JP_ERR_ARG;
  if (!c->initialised) return fail(c, JP_ERR_STATE, "jp_bench before jp_init");
  HCK(c, hipSetDevice(c->device));
  // src/tgv3d_amr.cu:839-864: a dynamic AMR run regrids every `adapt` steps, in warm-up and timed loop.
  // With graphs, an AMR context regrids on device-resident counts between replays, on the same schedule;
  // an ensemble replays pairs of steps without regrids, as the reference's graph bench does.
  const long adapt = c->kind == 1 ? ((amr_dynamic(c) && c->amr.adapt > 0) ? c->amr.adapt : 0)
                                   : (c->kind == 2 ? ens_adapt(c) : 0);
  const bool graphs = use_graph != 0;
  const bool regrids = adapt > 0 && !(graphs && c->kind == 2);
  const bool prev_graph = c->use_graph;
  if (graphs != prev_graph)
    if (int r = jp_use_graph(c, graphs ? 1 : 0)) return r;
  if (graphs && !regrids && steps % 2) steps += 1;  // whole pairs
  auto regrid = [&]() { return c->kind == 2 ? ens_regrid(c) : amr_regrid_with(c, nullptr, nullptr); };
  // Runs `total` steps with a regrid at every multiple of adapt; per segment between regrids it calls
  // seg(length) after the segment's steps are enqueued.
  auto run = [&](long total, auto&& seg) -> int {
    for (long s = 0; s < total;) {
      if (regrids && s % adapt == 0)
        if (int r = regrid()) return r;
      const long next = regrids ? std::min(total, (s / adapt + 1) * adapt) : total;
      if (int r = graphs ? jp_step(c, next - s) : eager_steps(c, next - s)) return r;
      if (int r = seg(next - s)) return r;
      s = next;
    }
    return JP_OK;
  };
  if (int r = run(warmup, [](long) { return JP_OK; })) return r;
  if (c->kind == 2)
    if (int r = ens_check(c)) return r;
  HCK(c, hipStreamSynchronize(c->st));
  if (c->kind == 1)
    if (int r = amr_sync_counts(c)) return r;
  size_t mem_free = 0, mem_total = 0;
  HCK(c, hipMemGetInfo(&mem_free, &mem_total));
  // Work of a segment in cell updates: the level counts it ran with. After a device-resident regrid the
  // host does not know them, so each segment's counts are copied to pinned memory behind its steps and
  // summed after the timed loop.
  const bool log_counts = c->kind == 1 && c->amr.resident;
  std::vector<long> seg_len;
  int* hlog = nullptr;
  const long max_segs = regrids ? steps / adapt + 2 : 1;
  if (log_counts) HCK(c, hipHostMalloc(reinterpret_cast<void**>(&hlog), size_t(max_segs) * 2 * sizeof(int), hipHostMallocDefault));
  hipEvent_t e0, e1;
  HCK(c, hipEventCreate(&e0));
  HCK(c, hipEventCreate(&e1));
  double work = 0;
  HCK(c, hipEventRecord(e0, c->st));
  int rc = run(steps, [&](long n) -> int {
    if (log_counts) {
      const hipError_t e = amr_log_counts(c, hlog + 2 * seg_len.size());
      if (e != hipSuccess) return fail(c, JP_ERR_HIP, hipGetErrorString(e));
    } else {
      work += double(n) * (c->kind == 2 ? ens_work(c) : c->kind == 1 ? amr_work(c) : double(c->d.cells));
    }
    seg_len.push_back(n);
    return JP_OK;
  });
  if (rc == JP_OK) {
    const hipError_t e = hipEventRecord(e1, c->st);
    rc = e == hipSuccess ? JP_OK : fail(c, JP_ERR_HIP, hipGetErrorString(e));
  }
  if (rc == JP_OK) {
    const hipError_t e = hipEventSynchronize(e1);
    rc = e == hipSuccess ? JP_OK : fail(c, JP_ERR_HIP, hipGetErrorString(e));
  }
  float ms = 0;
  if (rc == JP_OK) (void)hipEventElapsedTime(&ms, e0, e1);
  if (rc == JP_OK && log_counts)
    for (size_t k = 0; k < seg_len.size(); ++k)
      work += double(seg_len[k]) * double(c->amr.cnb0 + 2L * hlog[2 * k] + 4L * hlog[2 * k + 1]) * MB3;
  if (hlog) (void)hipHostFree(hlog);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (c->kind == 1)
    if (int r = amr_sync_counts(c)) return r;  // reports an overflowing resident regrid
  if (graphs != prev_graph)
    if (int r = jp_use_graph(c, prev_graph ? 1 : 0)) return r;
  res->steps = steps;
  res->ms_total = ms;
  res->ms_per_step = ms / double(steps);
  res->work = work;
  res->glups = work / (double(ms) * 1e6);
  res->bytes_per_update = 38.0 * double(prec_bytes(c->prec));
  res->gbps = res->glups * res->bytes_per_update;
  res->pool_gb = double(c->pool_bytes) / double(1ull << 30);
  res->device_used_gb = double(mem_total - mem_free) / double(1ull << 30);
  return JP_OK;
}

int jp_sync(jp_ctx* c) {
  if (!c) return JP_ERR_ARG;
  HCK(c, hipStreamSynchronize(c->st));
  return JP_OK;
}

const char* jp_last_error(const jp_ctx* c) { return c ? c->err.c_str() : "null context"; }

void jp_destroy(jp_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  drop_graph(c);
  if (c->st) (void)hipStreamSynchronize(c->st);
  ens_destroy(c);
  for (auto& a : c->allocs) (void)hipFree(a.first);
  if (c->amr.hcnt) (void)hipHostFree(c->amr.hcnt);
  if (c->host_out) (void)hipHostFree(c->host_out);
  if (c->amr.st2) {
    (void)hipStreamSynchronize(c->amr.st2);
    (void)hipStreamDestroy(c->amr.st2);
  }
  for (int i = 0; i < 2; ++i) {
    if (c->amr.ev_fork[i]) (void)hipEventDestroy(c->amr.ev_fork[i]);
    if (c->amr.ev_join[i]) (void)hipEventDestroy(c->amr.ev_join[i]);
  }
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

// Times `iters` launches of a plain bandwidth kernel over fresh buffers (mode -1: the streaming copy
// k_copy16 used to calibrate the traffic counters; 0 copy, 1 read, 2 write).
static int probe(int device, int mode, size_t bytes, int iters, double* gbps, double* ms_per_iter) {
  if (!gbps || iters < 1 || bytes < 4096) return JP_ERR_ARG;
  bytes &= ~size_t(4095);
  if (hipSetDevice(device) != hipSuccess) return JP_ERR_HIP;
  void *a = nullptr, *b = nullptr;
  hipStream_t st;
  if (hipStreamCreate(&st) != hipSuccess) return JP_ERR_HIP;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    (void)hipStreamDestroy(st);
    return JP_ERR_NOMEM;
  }
  int rc = JP_OK;
  auto launch = [&]() { return mode < 0 ? copy_probe(a, b, bytes, st) : bandwidth_probe(mode, a, b, bytes, st); };
  hipEvent_t e0, e1;
  (void)hipMemsetAsync(a, 0, bytes, st);
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  if (launch() != hipSuccess) rc = JP_ERR_HIP;
  (void)hipEventRecord(e0, st);
  for (int i = 0; i < iters && rc == JP_OK; ++i)
    if (launch() != hipSuccess) rc = JP_ERR_HIP;
  (void)hipEventRecord(e1, st);
  if (hipEventSynchronize(e1) != hipSuccess) rc = JP_ERR_HIP;
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  const double moved = (mode <= 0 ? 2.0 : 1.0) * double(bytes);
  *gbps = moved * iters / (double(ms) * 1e6);
  if (ms_per_iter) *ms_per_iter = ms / iters;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipStreamDestroy(st);
  return rc;
}

int jp_probe_copy(int device, size_t bytes, int iters, double* gbps, double* ms_per_iter) {
  return probe(device, -1, bytes, iters, gbps, ms_per_iter);
}

int jp_probe_bandwidth(int device, int mode, size_t bytes, int iters, double* gbps) {
  if (mode < 0 || mode > 2) return JP_ERR_ARG;
  return probe(device, mode, bytes, iters, gbps, nullptr);
}

}  // extern "C"
