// Batched ensemble (host side): members are ordinary contexts sharing the ensemble's stream; the
// ensemble keeps a device descriptor per member and one work list per batched kernel, and advances all
// members with one launch per substep kind in the order of the reference's base_step
// (src/tgv3d_amr.cu:821-837).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "ensemble.hpp"

namespace jp {

struct EnsState {
  std::vector<jp_ctx*> m;
  std::vector<int> adapt;
  MemberDesc* dmd = nullptr;
  std::vector<MemberDesc> hmd;
  int2* items[EK_COUNT] = {};
  size_t cap[EK_COUNT] = {};
  int n[EK_COUNT] = {};
  int groups[EK_COUNT] = {};
  // the dense step's second tile shape (8 cells per lane over 2048-cell chunks): its work list, whether
  // every member admits it, whether the autotuner chose it, and its workgroup count
  int2* items8 = nullptr;
  size_t cap8 = 0;
  int n8 = 0, groups8 = 0;
  bool dense8_ok = false, dense8 = false;
  int p0 = 0, p1 = 0, p2 = 0;  // parity of the current buffer of levels 0, 1, 2
  std::vector<long> sig;       // member counts and pool generation at upload
  bool uploaded = false;
  long steps = 0;
};

// Workgroups of a batched launch. A workgroup of the fused step + restriction kernels ends every item with
// one wave finishing 64 parents while the others wait, so a workgroup walking several items
// would serialise those tails; with one workgroup per item the dispatcher overlaps them with other
// items. The other kernels take their autotuned count.
static int ens_groups(const EnsState& e, int k) {
  if (k == EK_DENSE && e.dense8) return e.groups8;
  return (k == EK_STEP1R || k == EK_STEP2R) ? std::max(1, e.n[k]) : e.groups[k];
}


static MemberDesc describe(const jp_ctx* m) {
  MemberDesc d;
  std::memset(&d, 0, sizeof d);
  const AmrState& a = m->amr;
  d.c0[0] = m->c0a;
  d.c0[1] = m->c0b;
  d.f1[0] = a.f1a;
  d.f1[1] = a.f1b;
  d.f2[0] = a.f2a;
  d.f2[1] = a.f2b;
  d.gh1 = a.gh1;
  d.gh2 = a.gh2;
  d.ref0 = a.ref0;
  d.fof0 = a.fof0;
  d.rl0 = a.rl0;
  d.l1pcb = a.l1pcb;
  d.l1oct = a.l1oct;
  d.f1pos = a.f1pos;
  d.l1grid = a.l1grid;
  d.ref1 = a.ref1;
  d.fof1 = a.fof1;
  d.rl1 = a.rl1;
  d.l2pl1 = a.l2pl1;
  d.l2oct = a.l2oct;
  d.nb0 = m->kind == 1 ? a.nb0 : nullptr;
  d.nb1 = a.nb1;
  d.nb2 = a.nb2;
  d.slots = a.slots;
  d.g0 = m->g0;
  d.G.N0 = a.N;
  d.G.cgx = a.cgx;
  d.G.l1gx = a.l1gx;
  d.G.N1 = a.N1;
  d.G.N2 = a.N2;
  d.om[0] = m->d.omega0;
  d.om[1] = m->d.omega1;
  d.om[2] = m->d.omega2;
  for (int l = 0; l < 3; ++l) {
    d.enc[l] = float(m->d.scale[l]);
    d.dec[l] = float(1.0 / m->d.scale[l]);
  }
  d.nqs01 = float(m->d.nqs01);
  d.nqi10 = float(m->d.nqi10);
  d.nqs12 = float(m->d.nqs12);
  d.nqi21 = float(m->d.nqi21);
  return d;
}

static std::vector<long> signature(const EnsState& e) {
  std::vector<long> s;
  for (const jp_ctx* m : e.m) {
    const AmrState& a = m->amr;
    s.insert(s.end(), {long(a.nf1), long(a.nf2), long(a.nref0), long(a.nref1), long(a.pool_gen)});
  }
  return s;
}

// Descriptors and work lists from the members' current state; parities restart at 0.
static int upload(jp_ctx* c) {
  EnsState& e = *c->ens;
  e.hmd.clear();
  std::vector<int2> L[EK_COUNT], L8;
  e.dense8_ok = true;
  for (int i = 0; i < int(e.m.size()); ++i) {
    const jp_ctx* m = e.m[i];
    e.hmd.push_back(describe(m));
    const long chunks = m->d.cells / 256;
    for (long k = 0; k < chunks; ++k) L[EK_DENSE].push_back(make_int2(i, int(k)));
    // whole 2048-cell chunks only: a partial chunk would read and write past the member's lattice
    if (m->d.cells % kDense8Chunk) e.dense8_ok = false;
    if (m->kind == 1) e.dense8_ok = false;  // AMR level 0 is tile-major: one cell per lane only
    for (long k = 0; k < m->d.cells / kDense8Chunk; ++k) L8.push_back(make_int2(i, int(k)));
    if (m->kind != 1) continue;
    const AmrState& a = m->amr;
    if (a.nf1) {
      for (int b = 0; b < a.nf1; ++b) L[EK_FILL1].push_back(make_int2(i, b));
      for (int b = 0; b < a.nf1; ++b) L[EK_STEP1].push_back(make_int2(i, b));
      for (int b = 0; b < a.nf1; ++b) L[EK_STEP1R].push_back(make_int2(i, b));  // slab units, one per tile
    }
    if (a.nf1 && a.nf2) {
      for (int b = 0; b < a.nf2; ++b) L[EK_FILL2].push_back(make_int2(i, b));
      for (int b = 0; b < a.nf2; ++b) L[EK_STEP2].push_back(make_int2(i, b));
      for (int b = 0; b < a.nf2; ++b) L[EK_STEP2R].push_back(make_int2(i, b));
    }
  }
  HCK(c, hipStreamSynchronize(c->st));
  if (!e.dmd) {
    if (int r = ctx_alloc_t(c, &e.dmd, e.m.size(), "member descriptors")) return r;
  }
  HCK(c, hipMemcpy(e.dmd, e.hmd.data(), e.hmd.size() * sizeof(MemberDesc), hipMemcpyHostToDevice));
  for (int k = 0; k < EK_COUNT; ++k) {
    e.n[k] = int(L[k].size());
    if (L[k].size() > e.cap[k]) {
      ctx_free(c, e.items[k]);
      e.items[k] = nullptr;
      e.cap[k] = std::max<size_t>(L[k].size(), 1) * 5 / 4 + 16;
      if (int r = ctx_alloc_t(c, &e.items[k], e.cap[k], "work list")) return r;
    }
    if (!L[k].empty()) HCK(c, hipMemcpy(e.items[k], L[k].data(), L[k].size() * sizeof(int2), hipMemcpyHostToDevice));
  }
  if (!e.dense8_ok) L8.clear();
  e.n8 = int(L8.size());
  e.dense8 = e.dense8 && e.dense8_ok;
  if (L8.size() > e.cap8) {
    ctx_free(c, e.items8);
    e.items8 = nullptr;
    e.cap8 = std::max<size_t>(L8.size(), 1) * 5 / 4 + 16;
    if (int r = ctx_alloc_t(c, &e.items8, e.cap8, "work list")) return r;
  }
  if (!L8.empty()) HCK(c, hipMemcpy(e.items8, L8.data(), L8.size() * sizeof(int2), hipMemcpyHostToDevice));
  e.p0 = e.p1 = e.p2 = 0;
  e.sig = signature(e);
  e.uploaded = true;
  drop_graph(c);
  return JP_OK;
}

// Re-uploads when a member changed outside the batched path (its own jp_step, a regrid, set_populations).
int ens_check(jp_ctx* c) {
  EnsState& e = *c->ens;
  bool ok = e.uploaded && signature(e) == e.sig;
  for (size_t i = 0; ok && i < e.m.size(); ++i) {
    const jp_ctx* m = e.m[i];
    ok = m->c0a == e.hmd[i].c0[e.p0] && m->amr.f1a == e.hmd[i].f1[e.p1] && m->amr.f2a == e.hmd[i].f2[e.p2];
  }
  return ok ? JP_OK : upload(c);
}

void ens_graph_key(const jp_ctx* c, const void* key[3]) {
  const EnsState& e = *c->ens;
  const int p[3] = {e.p0, e.p1, e.p2};
  for (int l = 0; l < 3; ++l) key[l] = reinterpret_cast<const void*>(uintptr_t(1 + p[l]));
}

static void swap_all(EnsState& e, int level) {
  for (jp_ctx* m : e.m) {
    if (level == 0) std::swap(m->c0a, m->c0b);
    else if (level == 1) std::swap(m->amr.f1a, m->amr.f1b);
    else std::swap(m->amr.f2a, m->amr.f2b);
  }
  if (level == 0) e.p0 ^= 1;
  else if (level == 1) e.p1 ^= 1;
  else e.p2 ^= 1;
}

// One step of every member: the launch sequence of base_step with every launch batched.
int ens_step_once(jp_ctx* c) {
  EnsState& e = *c->ens;
  auto L = [&](EnsKernel k, int pa, int pb) -> hipError_t {
    return ens_launch(c->prec, k, e.dmd, e.items[k], e.n[k], ens_groups(e, k), pa, pb, c->st);
  };
  if (e.dense8) HCK(c, ens_launch_dense8(c->prec, e.dmd, e.items8, e.n8, e.groups8, e.p0, c->st));
  else HCK(c, L(EK_DENSE, e.p0, 0));
  swap_all(e, 0);
  HCK(c, L(EK_FILL1, e.p0, 0));
  for (int s1 = 0; s1 < 2; ++s1) {
    HCK(c, L(EK_FILL2, e.p1, 0));
    HCK(c, L(EK_STEP2, e.p2, 0));
    swap_all(e, 2);
    HCK(c, L(EK_STEP2R, e.p2, e.p1));  // second level-2 substep, restricted into level 1
    swap_all(e, 2);
    HCK(c, L(s1 ? EK_STEP1R : EK_STEP1, e.p1, e.p0));  // the second one restricts into level 0
    swap_all(e, 1);
  }
  ++e.steps;
  return JP_OK;
}

int ens_regrid(jp_ctx* c) {
  EnsState& e = *c->ens;
  for (jp_ctx* m : e.m)
    if (m->kind == 1 && amr_dynamic(m))
      if (int r = amr_regrid_with(m, nullptr, nullptr)) return fail(c, r, m->err);
  return upload(c);
}

int ens_adapt(const jp_ctx* c) {
  int a = 0;
  for (const jp_ctx* m : c->ens->m)
    if (m->kind == 1 && amr_dynamic(m) && m->amr.adapt > 0) a = a ? std::min(a, m->amr.adapt) : m->amr.adapt;
  return a;
}

bool ens_dynamic(const jp_ctx* c) {
  for (const jp_ctx* m : c->ens->m)
    if (m->kind == 1 && amr_dynamic(m)) return true;
  return false;
}

int ens_init(jp_ctx* c) {
  for (jp_ctx* m : c->ens->m) {
    if (int r = jp_init(m)) return fail(c, r, m->err);
  }
  if (int r = upload(c)) return r;
  c->initialised = true;
  return JP_OK;
}

double ens_work(const jp_ctx* c) {
  double w = 0;
  for (const jp_ctx* m : c->ens->m) w += m->kind == 1 ? amr_work(m) : double(m->d.cells);
  return w;
}

// Each candidate launch re-runs one batched kernel on unchanged inputs: the dense and fine steps write
// the non-current buffer, the fills their ghost records. The fused step + restriction launches also
// write the parent level's current buffer, so they are not timed and take their plain step's choice.
int ens_autotune(jp_ctx* c, int* out) {
  EnsState& e = *c->ens;
  if (int r = ens_check(c)) return r;
  hipDeviceProp_t prop;
  HCK(c, hipGetDeviceProperties(&prop, c->device));
  const int cu = prop.multiProcessorCount;
  const int cand[] = {1, 2, 4, 8, 16, 32};
  hipEvent_t a, b;
  HCK(c, hipEventCreate(&a));
  HCK(c, hipEventCreate(&b));
  // median-of-3 time of the fastest workgroup count for a launch over n items; launch(groups) enqueues it
  auto tune = [&](int n, auto&& launch, float* best_ms, int* best_groups) -> int {
    *best_ms = 1e30f;
    *best_groups = cu * 4;
    for (int g : cand) {
      const int groups = std::min(n, g * cu);
      float t[3];
      for (int r = 0; r < 3; ++r) {
        HCK(c, hipEventRecord(a, c->st));
        HCK(c, launch(groups));
        HCK(c, hipEventRecord(b, c->st));
        HCK(c, hipEventSynchronize(b));
        HCK(c, hipEventElapsedTime(&t[r], a, b));
      }
      std::sort(t, t + 3);
      if (t[1] < *best_ms) {
        *best_ms = t[1];
        *best_groups = groups;
      }
      if (groups == n) break;
    }
    return JP_OK;
  };
  for (int k = 0; k < EK_COUNT; ++k) {
    if (k == EK_STEP2R || k == EK_STEP1R) continue;  // they write the parent level: one workgroup per item
    if (e.n[k] == 0) {
      e.groups[k] = cu * 4;
      continue;
    }
    const int pa = (k == EK_DENSE || k == EK_FILL1) ? e.p0 : (k == EK_STEP2) ? e.p2 : e.p1;
    float best;
    if (int r = tune(e.n[k], [&](int groups) {
          return ens_launch(c->prec, EnsKernel(k), e.dmd, e.items[k], e.n[k], groups, pa, 0, c->st);
        }, &best, &e.groups[k]))
      return r;
    if (k == EK_DENSE) {  // the dense step's tile shape: 1 or 8 cells per lane, whichever is faster
      e.dense8 = false;
      float best8;
      if (e.dense8_ok && e.n8 > 0) {
        if (int r = tune(e.n8, [&](int groups) {
              return ens_launch_dense8(c->prec, e.dmd, e.items8, e.n8, groups, pa, c->st);
            }, &best8, &e.groups8))
          return r;
        e.dense8 = best8 < best;
      }
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  drop_graph(c);
  if (out)
    for (int k = 0; k < EK_COUNT; ++k) out[k] = ens_groups(e, k);
  return JP_OK;
}

// Sets (set = 1 or 8) or only reports (set = 0) the dense step's tile shape in cells per lane. 8 needs
// uniform members only (an AMR member's level 0 is tile-major), each a multiple of the 2048-cell chunk.
int ens_dense_shape(jp_ctx* c, int set, int* current) {
  EnsState& e = *c->ens;
  if (int r = ens_check(c)) return r;
  if (set == 8) {
    if (!e.dense8_ok)
      return fail(c, JP_ERR_ARG, "8 cells per lane needs uniform members only, each lattice a multiple of 2048 cells");
    e.dense8 = true;
  } else if (set == 1) {
    e.dense8 = false;
  } else if (set != 0) {
    return fail(c, JP_ERR_ARG, "the dense step takes 1 or 8 cells per lane");
  }
  if (set) drop_graph(c);
  if (current) *current = e.dense8 ? 8 : 1;
  return JP_OK;
}

void ens_destroy(jp_ctx* c) {
  if (!c->ens) return;
  for (jp_ctx* m : c->ens->m) {
    m->st = nullptr;  // the stream belongs to the ensemble
    jp_destroy(m);
  }
  delete c->ens;
  c->ens = nullptr;
}

int ens_create(const jp_member* mem, int count, int prec, double maxgb, int device, jp_ctx* c) {
  c->ens = new EnsState();
  EnsState& e = *c->ens;
  hipDeviceProp_t prop;
  HCK(c, hipGetDeviceProperties(&prop, device));
  for (int k = 0; k < EK_COUNT; ++k) e.groups[k] = prop.multiProcessorCount * 8;
  e.groups8 = prop.multiProcessorCount * 8;
  for (int i = 0; i < count; ++i) {
    const jp_member& p = mem[i];
    jp_ctx* m = nullptr;
    int r;
    if (p.amr) {
      jp_amr_params q{};
      q.n = p.n;
      q.re = p.re;
      q.u0 = p.u0;
      q.prec = prec;
      for (int k = 0; k < 3; ++k) q.phase[k] = p.phase[k];
      q.l1_mode = p.l1_mode;
      q.l2_mode = p.l2_mode;
      q.frac1 = p.frac1;
      q.frac2 = p.frac2;
      q.adapt = p.adapt;
      q.maxgb = maxgb;
      q.device = device;
      r = jp_amr_create(&q, &m);
    } else {
      jp_uniform_params q{};
      q.n = p.n;
      q.re = p.re;
      q.u0 = p.u0;
      q.prec = prec;
      for (int k = 0; k < 3; ++k) q.phase[k] = p.phase[k];
      q.device = device;
      r = jp_uniform_create(&q, &m);
    }
    if (r) {
      std::string msg = "member " + std::to_string(i) + ": " + (m ? m->err : std::string("?"));
      if (m) jp_destroy(m);
      return fail(c, r, msg);
    }
    (void)hipStreamDestroy(m->st);
    m->st = c->st;
    e.m.push_back(m);
    c->d.cells += m->d.cells;
  }
  return JP_OK;
}

jp_ctx* ens_member(jp_ctx* c, int i) {
  if (!c->ens || i < 0 || i >= int(c->ens->m.size())) return nullptr;
  return c->ens->m[i];
}

}  // namespace jp
