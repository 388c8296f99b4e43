// AMR orchestration (host side of the C ABI): pool sizing, the initial build, the regrid pipeline and
// the nested 1:2:4 subcycle (src/tgv3d_amr.cu:657-837). All data stays on the device; the host reads
// only the refinement counts it needs to size launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "amr_internal.hpp"
#include "ctx.hpp"

namespace jp {

static AmrGeo geo(const jp_ctx* c) {
  AmrGeo G;
  G.N0 = c->amr.N;
  G.cgx = c->amr.cgx;
  G.l1gx = c->amr.l1gx;
  G.N1 = c->amr.N1;
  G.N2 = c->amr.N2;
  return G;
}
static size_t real_bytes(int prec) { return prec == JP_PREC_FP64 ? 8 : 4; }
static float enc(const jp_ctx* c, int l) { return float(c->d.scale[l]); }
static float dec(const jp_ctx* c, int l) { return float(1.0 / c->d.scale[l]); }
// Launch width of the device-counted launches: the last count the host knew, in multiples of 8
// (workgroups beyond the live count return at once, at a cost of a few microseconds per launch when
// there are many; a count beyond the width gives some workgroups a second tile).
static int width8(long n) { return int(std::max(8L, (n + 7) / 8 * 8)); }
// A count beyond a pool leaves the grid incomplete: the context needs a new initial build before it
// steps again.
static int capacity_fail(jp_ctx* c, const std::string& msg) {
  c->initialised = false;
  return fail(c, JP_ERR_CAPACITY, msg);
}

// Tiles of the level-1 / level-2 launches: one workgroup per tile of the host's count, or, when regrids
// run on device-resident counts, the count read on the device and a launch width from the count the host
// last knew (width8; a launch recorded in a graph keeps its width, and the kernels loop over any tiles a
// later regrid adds beyond it).
static Tiles tiles(const jp_ctx* c, int level) {
  const AmrState& a = c->amr;
  if (!a.resident && !a.dev_counts) return {nullptr, level == 1 ? a.nf1 : a.nf2};
  return {a.dcnt + (level == 1 ? CNT_NF1 : CNT_NF2), level == 1 ? a.w1 : a.w2};
}
static Items items(const jp_ctx* c, int level) {
  const AmrState& a = c->amr;
  const int* dn = a.dcnt + (level == 1 ? CNT_NF1 : CNT_NF2);
  return {(a.resident || a.dev_counts) ? (level == 1 ? a.w1 : a.w2) : (level == 1 ? a.nf1 : a.nf2), dn};
}
static Tiles tiles0(const jp_ctx* c) { return {nullptr, c->amr.cnb0}; }
// One level-0 step c0a -> c0b (the fine-level step over all cnb0 tiles; no ghost record is read).
static int l0_step(jp_ctx* c, hipStream_t st) {
  HCK(c, amr_fine_step(c->prec, 0, c->c0a, c->c0b, nullptr, c->amr.nb0, tiles0(c), c->d.omega0, float(c->d.scale[0]),
                       float(1.0 / c->d.scale[0]), st));
  std::swap(c->c0a, c->c0b);
  return JP_OK;
}

// Fill table of the compact ghost record, one entry per slot (encoding in amr_internal.hpp). Every fine
// tile's origin is even on its own grid, so the box offset and the weights depend on the slot only.
static std::vector<uint32_t> ghost_fill_table() {
  std::vector<uint32_t> t(kGhostSlots, 0u);
  auto tri_local = [](int g) { return ((g & 1) ? (g >> 1) : (g >> 1) - 1) + 1; };  // parent base - (o/2 - 1)
  auto side = [](int s) { return s < 0 ? -1 : (s > 7 ? 1 : 0); };
  for (int q = 1; q < Q; ++q)
    for (int cell = 0; cell < MB3; ++cell) {
      const int i = cell & 7, j = (cell >> 3) & 7, k = cell >> 6;
      const int si = i - CX(q), sj = j - CY(q), sk = k - CZ(q);
      if (si >= 0 && si < 8 && sj >= 0 && sj < 8 && sk >= 0 && sk < 8) continue;
      const int dir = (side(sk) + 1) * 9 + (side(sj) + 1) * 3 + side(si) + 1;
      const int box = (tri_local(sk) * kFillBox + tri_local(sj)) * kFillBox + tri_local(si);
      t[ghost_slot(q, i, j, k)] = uint32_t(q) | uint32_t(dir) << 5 | uint32_t(box) << 10 | uint32_t(si & 1) << 18 |
                                  uint32_t(sj & 1) << 19 | uint32_t(sk & 1) << 20;
    }
  return t;
}

// ---- counts ----
static int read_counts(jp_ctx* c) {
  HCK(c, hipMemcpyAsync(c->amr.hcnt, c->amr.dcnt, sizeof(int) * CNT_N, hipMemcpyDeviceToHost, c->st));
  HCK(c, hipStreamSynchronize(c->st));
  return JP_OK;
}
static int fill(jp_ctx* c, int* p, int v, size_t n) {
  HCK(c, fill_i32(p, v, n, c->st));
  return JP_OK;
}
// Host copies of the counts after device-resident regrids, read when a caller needs them. The overflow
// flag is sticky: a resident regrid whose counts exceeded a pool left the grid incomplete.
// The host copies from hcnt, just read.
static int take_counts(jp_ctx* c) {
  AmrState& a = c->amr;
  a.counts_stale = false;
  a.nref0 = a.hcnt[CNT_NREF0];
  a.nf1 = a.hcnt[CNT_NF1];
  a.nref1 = a.hcnt[CNT_NREF1];
  a.nf2 = a.hcnt[CNT_NF2];
  a.hw_nf1 = std::max(a.hw_nf1, long(a.hcnt[CNT_HW1]));
  a.hw_nf2 = std::max(a.hw_nf2, long(a.hcnt[CNT_HW2]));
  a.w1 = width8(a.nf1);
  a.w2 = width8(a.nf2);
  if (a.hcnt[CNT_OVERFLOW])
    return capacity_fail(c, "a device-resident regrid exceeded the pool capacity (level-1 " +
                                        std::to_string(8L * a.nref0) + "/" + std::to_string(a.cap1) + ", level-2 " +
                                        std::to_string(8L * a.nref1) + "/" + std::to_string(a.cap2) + "); raise --maxgb");
  return JP_OK;
}
int amr_sync_counts(jp_ctx* c) {
  if (!c->amr.counts_stale) return JP_OK;
  if (int r = read_counts(c)) return r;
  return take_counts(c);
}

static bool pool_fits(const AmrState& a, long c1, long c2) {
  // src/tgv3d_amr.cu:687-688: (f x2 + ghost box) per fine tile in FP32, +20 % for tables; and at most
  // kMaxFineTiles tiles per level (32-bit element offsets in the fine steps)
  const double PB = double(Q) * 4 * (2 * MB3 + 1000);
  return c1 <= kMaxFineTiles && c2 <= kMaxFineTiles && double(c1 + c2) * PB * 6 / 5 < a.budget;
}

static int alloc_l2_pool(jp_ctx* c) {
  AmrState& a = c->amr;
  const size_t e = prec_bytes(c->prec), r = real_bytes(c->prec);
  const size_t tile = size_t(Q) * MB3;
  if (int rc = ctx_alloc(c, &a.f2a, size_t(a.cap2) * tile * e, "f2a")) return rc;
  if (int rc = ctx_alloc(c, &a.f2b, size_t(a.cap2) * tile * e, "f2b")) return rc;
  if (int rc = ctx_alloc(c, &a.gh2, size_t(a.cap2) * kGhostSlots * r, "gh2")) return rc;
  if (int rc = ctx_alloc_t(c, &a.l2pl1, a.cap2, "l2pl1")) return rc;
  if (int rc = ctx_alloc_t(c, &a.l2oct, a.cap2, "l2oct")) return rc;
  if (int rc = ctx_alloc_t(c, &a.nb2, size_t(a.cap2) * kNbr, "nb2")) return rc;
  a.cap2_alloc = a.cap2;
  ++a.pool_gen;
  return JP_OK;
}
static void free_l2_pool(jp_ctx* c) {
  AmrState& a = c->amr;
  for (void** p : {&a.f2a, &a.f2b, &a.gh2, reinterpret_cast<void**>(&a.l2pl1), reinterpret_cast<void**>(&a.l2oct),
                   reinterpret_cast<void**>(&a.nb2)}) {
    ctx_free(c, *p);
    *p = nullptr;
  }
}

// The 27-neighbour tables of both fine levels, which the steps and fills read. The skin lists
// (src/tgv3d_amr.cu:758-762) are part of the block tables but no kernel here reads them (the fills compute
// prep where they need it), so they are built on demand by ensure_skins.
static int finish_tables(jp_ctx* c) {
  AmrState& a = c->amr;
  HCK(c, rg_nbr_l1(a.f1pos, a.l1grid, a.l1gx, items(c, 1), a.nb1, c->st));
  HCK(c, rg_nbr_l2(a.l2pl1, a.l2oct, a.f1pos, a.l1grid, a.ref1, a.fof1, a.l1gx, items(c, 2), a.nb2, c->st));
  a.skins_stale = true;
  return JP_OK;
}
static int ensure_skins(jp_ctx* c) {
  AmrState& a = c->amr;
  if (int r = amr_sync_counts(c)) return r;
  if (!a.skins_stale) return JP_OK;
  HCK(c, rg_dilate_l0(a.ref0, a.cnb0, a.cgx, a.flags, c->st));
  HCK(c, rg_compact(a.flags, {a.cnb0, nullptr}, a.skin0, a.dcnt, CNT_SKIN0, c->st));
  if (a.nf1) {
    HCK(c, rg_dilate_l1(a.ref1, a.f1pos, a.l1grid, {a.nf1, nullptr}, a.l1gx, a.flags, c->st));
    HCK(c, rg_compact(a.flags, {a.nf1, nullptr}, a.skin1, a.dcnt, CNT_SKIN1, c->st));
  }
  if (int r = read_counts(c)) return r;
  a.n_skin0 = a.hcnt[CNT_SKIN0];
  a.n_skin1 = a.nf1 ? a.hcnt[CNT_SKIN1] : 0;
  a.skins_stale = false;
  return JP_OK;
}

static int alloc_l1_pool(jp_ctx* c, bool with_f1a) {
  AmrState& a = c->amr;
  const size_t e = prec_bytes(c->prec), r = real_bytes(c->prec), tile = size_t(Q) * MB3;
  ++a.pool_gen;
  if (with_f1a)
    if (int rc = ctx_alloc(c, &a.f1a, size_t(a.cap1) * tile * e, "f1a")) return rc;
  if (int rc = ctx_alloc(c, &a.f1b, size_t(a.cap1) * tile * e, "f1b")) return rc;
  if (int rc = ctx_alloc(c, &a.gh1, size_t(a.cap1) * kGhostSlots * r, "gh1")) return rc;
  return JP_OK;
}

// Ties at the level-0 threshold can exceed the level-1 pool (the reference writes out of bounds there;
// a symmetric TGV ties every tile at small N0). The pool grows when the budget allows; the refinement
// target cap_nref0 stays as the reference sizes it. keep_f1a: a regrid's migration still reads the old
// level-1 data, so f1a is kept at its old size and replaced after the migration.
static int grow_l1(jp_ctx* c, bool keep_f1a) {
  AmrState& a = c->amr;
  const long c2 = a.f2a ? a.cap2 : 0;
  int cn = std::min(a.cnb0, std::max(a.nref0, (a.cap1 / 8) * 3 / 2));
  while (cn > a.nref0 && !pool_fits(a, 8L * cn, c2)) cn--;
  if (!pool_fits(a, 8L * cn, c2))
    return capacity_fail(c, "level-1 count " + std::to_string(a.nf1) + " exceeds the pool capacity " +
                                        std::to_string(a.cap1) + " and the budget (threshold ties); raise --maxgb");
  HCK(c, hipStreamSynchronize(c->st));
  for (void* p : {a.f1b, a.gh1}) ctx_free(c, p);
  if (!keep_f1a) ctx_free(c, a.f1a);
  a.cap1 = 8 * cn;
  a.f1_stale = keep_f1a;
  return alloc_l1_pool(c, !keep_f1a);
}

// level-1 marks -> scan -> counts -> tables (common to init and regrid). ref0 is already marked.
static int scatter_l1(jp_ctx* c, bool clamp = false) {
  AmrState& a = c->amr;
  if (int r = fill(c, a.l1grid, -1, size_t(a.NG1))) return r;  // the scatter writes every fof0 entry
  HCK(c, rg_scatter_l1(a.ref0, a.rank, a.cnb0, a.cgx, a.l1gx, a.cap1, clamp, a.fof0, a.rl0, a.l1pcb, a.l1oct, a.l1grid, a.f1pos,
                       a.dcnt, c->st));
  return JP_OK;
}

// The level-1 tables are scattered before the host reads the count (the scatter needs only the ranks and
// refuses writes past the pool), so the launches after the read are not serialised behind it; in the rare
// case that the count exceeds the pool, the pool grows and the scatter is repeated.
static int build_l1_tables(jp_ctx* c, bool keep_f1a) {
  AmrState& a = c->amr;
  HCK(c, rg_scan(a.ref0, {a.cnb0, nullptr}, a.rank, a.dcnt, CNT_NREF0, CNT_NF1, a.cap1, c->st));
  if (int r = scatter_l1(c)) return r;
  if (int r = read_counts(c)) return r;
  a.nref0 = a.hcnt[CNT_NREF0];
  a.nf1 = 8 * a.nref0;
  a.hw_nf1 = std::max(a.hw_nf1, long(a.nf1));
  if (a.nf1 > a.cap1) {
    if (int r = grow_l1(c, keep_f1a)) return r;
    if (int r = fill(c, a.dcnt + CNT_OVERFLOW, 0, 1)) return r;
    if (int r = fill(c, a.dcnt + CNT_NF1, a.nf1, 1)) return r;  // the scan stopped the count at the old pool
    a.hcnt[CNT_OVERFLOW] = 0;
    if (int r = scatter_l1(c)) return r;
  }
  return JP_OK;
}

static int scatter_l2(jp_ctx* c, bool clamp = false) {
  AmrState& a = c->amr;
  if (int r = fill(c, a.fof1, -1, size_t(a.cap1))) return r;
  HCK(c, rg_scatter_l2(a.ref1, a.rank, items(c, 1), a.cap2, clamp, a.fof1, a.rl1, a.l2pl1, a.l2oct, a.dcnt, c->st));
  return JP_OK;
}

// Level-1 marks are in ref1: rank them, scatter the level-2 tables ahead of the count read (as above)
// when the pool exists, and read the level-2 count.
static int scan_l2(jp_ctx* c, bool scatter) {
  AmrState& a = c->amr;
  // before the pool exists (init) the count is not limited here; init checks it against the pool
  HCK(c, rg_scan(a.ref1, items(c, 1), a.rank, a.dcnt, CNT_NREF1, CNT_NF2, scatter ? a.cap2 : INT_MAX, c->st));
  if (scatter)
    if (int r = scatter_l2(c)) return r;
  if (int r = read_counts(c)) return r;
  a.nref1 = a.hcnt[CNT_NREF1];
  a.nf2 = 8 * a.nref1;
  a.hw_nf2 = std::max(a.hw_nf2, long(a.nf2));
  return JP_OK;
}

static int upload_sensor(jp_ctx* c, const float* sen0, int n) {
  HCK(c, hipMemcpyAsync(c->amr.sen0, sen0, size_t(n) * 4, hipMemcpyHostToDevice, c->st));
  return JP_OK;
}
static int sensor_l1(jp_ctx* c, const float* sen1_by_pos) {
  AmrState& a = c->amr;
  HCK(c, amr_sensor(c->prec, geo(c), a.f1a, tiles(c, 1), dec(c, 1), a.sen1, c->st));
  if (sen1_by_pos) {
    HCK(c, hipMemcpyAsync(a.sen1pos, sen1_by_pos, size_t(a.NG1) * 4, hipMemcpyHostToDevice, c->st));
    HCK(c, rg_gather_by_pos(a.sen1pos, a.f1pos, items(c, 1), a.sen1, c->st));
  }
  HCK(c, rg_deepmask(a.f1pos, a.l1grid, a.l1gx, items(c, 1), a.deep, c->st));
  return JP_OK;
}

static int apply_resident(jp_ctx* c);

// ---- initial build, src/tgv3d_amr.cu:776-815 ----
int amr_init_with(jp_ctx* c, const float* sen0, const float* sen1_by_pos) {
  AmrState& a = c->amr;
  const AmrGeo G = geo(c);
  drop_graph(c);
  HCK(c, amr_l0_init(c->prec, G, c->c0a, c->u0, c->d.kf, c->phase, enc(c, 0), c->st));
  if (int r = fill(c, a.dcnt, 0, CNT_N)) return r;
  a.hw_nf1 = a.hw_nf2 = 0;
  a.counts_stale = false;
  if (a.l1_mode == JP_MODE_ALL) HCK(c, rg_set(a.ref0, a.cnb0, 0, a.cgx, nullptr, c->st));
  else if (a.l1_mode == JP_MODE_SLAB) HCK(c, rg_set(a.ref0, a.cnb0, 1, a.cgx, nullptr, c->st));
  else {
    for (int s = 0; s < 300; ++s)  // develop(300): unrecorded coarse steps before the first marking
      if (int r = l0_step(c, c->st)) return r;
    HCK(c, amr_sensor(c->prec, G, c->c0a, tiles0(c), dec(c, 0), a.sen0, c->st));
    if (sen0)
      if (int r = upload_sensor(c, sen0, a.cnb0)) return r;
    HCK(c, rg_select(a.sen0, nullptr, {a.cnb0, nullptr}, a.frac1, a.cap_nref0, a.thr, a.dcnt, CNT_NSEL0, c->st));
    HCK(c, rg_mark_thr(a.sen0, a.thr, a.ref0, a.cnb0, c->st));
  }
  if (int r = build_l1_tables(c, false)) return r;
  HCK(c, amr_seed_l1(c->prec, G, c->c0a, a.f1a, a.l1pcb, a.l1oct, tiles(c, 1), float(c->d.nqs01), dec(c, 0), enc(c, 1), c->st));
  a.nref1 = a.nf2 = 0;
  if (a.nf1) {
    if (int r = sensor_l1(c, sen1_by_pos)) return r;
    if (int r = fill(c, a.ref1, 0, size_t(a.cap1))) return r;
    if (a.l2_mode == JP_MODE_ALL) HCK(c, rg_set(a.ref1, a.nf1, 2, a.cgx, a.deep, c->st));
    else {
      // init does not cap the level-2 target (src/tgv3d_amr.cu:795)
      HCK(c, rg_select(a.sen1, a.deep, items(c, 1), a.frac2, INT_MAX, a.thr, a.dcnt, CNT_NDEEP, c->st));
      HCK(c, rg_mark_ref1(a.sen1, a.thr, a.deep, a.dcnt, a.ref1, items(c, 1), c->st));
    }
    if (int r = scan_l2(c, false)) return r;
    a.cap_nref1 = std::max(1, int(std::ceil(a.nref1 * 2.5)));
    if (a.l2_mode == JP_MODE_ALL) a.cap_nref1 = std::max(a.cap_nref1, a.nf1);
    a.cap_nref1 = std::min(a.cap_nref1, a.cap1);
    while (a.cap_nref1 > 1 && !pool_fits(a, a.cap1, 8L * a.cap_nref1)) a.cap_nref1--;
    a.cap2 = 8 * a.cap_nref1;
  } else {
    a.cap_nref1 = 1;
    a.cap2 = 8;
  }
  if (a.f2a && a.cap2 > a.cap2_alloc) {  // a re-initialisation asks for more level-2 tiles than are held
    HCK(c, hipStreamSynchronize(c->st));
    free_l2_pool(c);
  }
  if (!a.f2a)
    if (int r = alloc_l2_pool(c)) return r;
  if (a.resident) a.cap2 = a.cap2_alloc;  // a re-initialisation keeps the pools sized for resident regrids
  if (a.nf2 > a.cap2)
    return capacity_fail(c, "level-2 count " + std::to_string(a.nf2) + " exceeds the pool capacity " +
                                        std::to_string(a.cap2) + "; raise --maxgb");
  if (int r = fill(c, a.fof1, -1, size_t(a.cap1))) return r;
  if (a.nf2) {
    HCK(c, rg_scatter_l2(a.ref1, a.rank, items(c, 1), a.cap2, false, a.fof1, a.rl1, a.l2pl1, a.l2oct, a.dcnt, c->st));
    HCK(c, amr_seed_l2(c->prec, G, a.f1a, a.f2a, a.l2pl1, a.l2oct, a.f1pos, a.l1grid, tiles(c, 2), float(c->d.nqs12),
                       dec(c, 1), enc(c, 2), c->st));
  }
  if (int r = finish_tables(c)) return r;
  if (a.hcnt[CNT_OVERFLOW]) return capacity_fail(c, "pool capacity exceeded while scattering tables");
  c->initialised = true;
  a.w1 = width8(a.nf1);
  a.w2 = width8(a.nf2);
  return a.resident_req && !a.resident ? apply_resident(c) : JP_OK;
}
int amr_init(jp_ctx* c) { return amr_init_with(c, nullptr, nullptr); }

// ---- regrid, src/tgv3d_amr.cu:722-764 ----
// snapshots of the tables the migration keys on, one launch
static int snapshot_tables(jp_ctx* c) {
  AmrState& a = c->amr;
  const int* src[5] = {a.ref0, a.fof0, a.l1grid, a.ref1, a.fof1};
  int* dst[5] = {a.o_ref0, a.o_fof0, a.o_l1grid, a.o_ref1, a.o_fof1};
  const long n[5] = {a.cnb0, a.cnb0, a.NG1, a.cap1, a.cap1};
  HCK(c, rg_copy_i32_multi(src, dst, n, 5, c->st));
  return JP_OK;
}

// The regrid on device-resident counts: the same pipeline with every level-1 and level-2 launch sized
// on the device, no count read back and no pool growth. Both fine levels always migrate, so each regrid
// swaps both buffer pairs and the captured steps of either parity stay valid. With clamp (regrids of a
// resident context, whose pools hold the largest counts the budget allows, see apply_resident) a parent
// beyond a pool is left unrefined; without it the tables keep such parents for the eager path to redo.
static int regrid_device(jp_ctx* c, bool clamp) {
  AmrState& a = c->amr;
  const AmrGeo G = geo(c);
  if (int r = snapshot_tables(c)) return r;
  HCK(c, amr_sensor(c->prec, G, c->c0a, tiles0(c), dec(c, 0), a.sen0, c->st));
  HCK(c, rg_select(a.sen0, nullptr, {a.cnb0, nullptr}, a.frac1, a.cap_nref0, a.thr, a.dcnt, CNT_NSEL0, c->st));
  HCK(c, rg_mark_thr(a.sen0, a.thr, a.ref0, a.cnb0, c->st));
  HCK(c, rg_scan(a.ref0, {a.cnb0, nullptr}, a.rank, a.dcnt, CNT_NREF0, CNT_NF1, a.cap1, c->st));
  if (int r = scatter_l1(c, clamp)) return r;
  HCK(c, amr_migrate_l1(c->prec, G, a.f1a, a.f1b, c->c0a, a.l1pcb, a.l1oct, a.o_ref0, a.o_fof0, tiles(c, 1),
                        float(c->d.nqs01), dec(c, 0), enc(c, 1), c->st));
  std::swap(a.f1a, a.f1b);
  if (int r = sensor_l1(c, nullptr)) return r;
  HCK(c, rg_select(a.sen1, a.deep, items(c, 1), a.frac2, a.cap_nref1, a.thr, a.dcnt, CNT_NDEEP, c->st));
  if (int r = fill(c, a.ref1, 0, size_t(a.cap1))) return r;
  HCK(c, rg_mark_ref1(a.sen1, a.thr, a.deep, a.dcnt, a.ref1, items(c, 1), c->st));
  HCK(c, rg_scan(a.ref1, items(c, 1), a.rank, a.dcnt, CNT_NREF1, CNT_NF2, a.cap2, c->st));
  if (int r = scatter_l2(c, clamp)) return r;
  HCK(c, amr_migrate_l2(c->prec, G, a.f2a, a.f2b, a.f1a, a.l2pl1, a.l2oct, a.f1pos, a.o_l1grid, a.o_ref1, a.o_fof1,
                        a.l1grid, tiles(c, 2), float(c->d.nqs12), dec(c, 1), enc(c, 2), c->st));
  std::swap(a.f2a, a.f2b);
  a.counts_stale = true;
  return finish_tables(c);
}

int amr_regrid_with(jp_ctx* c, const float* sen0, const float* sen1_by_pos) {
  AmrState& a = c->amr;
  const AmrGeo G = geo(c);
  if (a.resident && !sen0 && !sen1_by_pos) return regrid_device(c, true);
  if (int r = amr_sync_counts(c)) return r;
  drop_graph(c);
  bool snapshot = true;
  if (!sen0 && !sen1_by_pos) {
    // One count read instead of two: the regrid runs on device-resident counts against the current pools
    // and the counts are read once at the end. Only when a count exceeded a pool (threshold ties) is it
    // undone - the migrations wrote the second buffers, and the snapshots still hold the tables from
    // before - and run again below with the host reading each count and growing the pool.
    a.dev_counts = true;
    const int r = regrid_device(c, false);
    a.dev_counts = false;
    if (r) return r;
    if (int r2 = read_counts(c)) return r2;
    if (!a.hcnt[CNT_OVERFLOW]) return take_counts(c);
    std::swap(a.f1a, a.f1b);
    std::swap(a.f2a, a.f2b);
    if (int r2 = fill(c, a.dcnt + CNT_OVERFLOW, 0, 1)) return r2;
    a.hcnt[CNT_OVERFLOW] = 0;
    a.counts_stale = false;
    snapshot = false;
  }
  if (snapshot)
    if (int r = snapshot_tables(c)) return r;
  HCK(c, amr_sensor(c->prec, G, c->c0a, tiles0(c), dec(c, 0), a.sen0, c->st));
  if (sen0)
    if (int r = upload_sensor(c, sen0, a.cnb0)) return r;
  HCK(c, rg_select(a.sen0, nullptr, {a.cnb0, nullptr}, a.frac1, a.cap_nref0, a.thr, a.dcnt, CNT_NSEL0, c->st));
  HCK(c, rg_mark_thr(a.sen0, a.thr, a.ref0, a.cnb0, c->st));
  HCK(c, rg_differs(a.ref0, a.o_ref0, {a.cnb0, nullptr}, a.dcnt, CNT_DIFF0, c->st));
  if (int r = build_l1_tables(c, true)) return r;
  // Unchanged level-0 marks give the same level-1 tiles in the same order: the migration would be an
  // identity copy and is skipped (likewise for level 2 below).
  const bool same1 = !a.hcnt[CNT_DIFF0] && !a.f1_stale;
  if (a.nf1 && !same1) {
    HCK(c, amr_migrate_l1(c->prec, G, a.f1a, a.f1b, c->c0a, a.l1pcb, a.l1oct, a.o_ref0, a.o_fof0, tiles(c, 1),
                          float(c->d.nqs01), dec(c, 0), enc(c, 1), c->st));
    std::swap(a.f1a, a.f1b);
  }
  if (a.f1_stale) {  // the pool grew: the buffer that held the old level-1 data still has the old size
    HCK(c, hipStreamSynchronize(c->st));
    ctx_free(c, a.f1b);
    if (int r = ctx_alloc(c, &a.f1b, size_t(a.cap1) * Q * MB3 * prec_bytes(c->prec), "f1b")) return r;
    a.f1_stale = false;
    ++a.pool_gen;
  }
  a.nref1 = a.nf2 = 0;
  if (a.nf1) {
    if (int r = sensor_l1(c, sen1_by_pos)) return r;
    HCK(c, rg_select(a.sen1, a.deep, items(c, 1), a.frac2, a.cap_nref1, a.thr, a.dcnt, CNT_NDEEP, c->st));
    if (int r = fill(c, a.ref1, 0, size_t(a.cap1))) return r;
    HCK(c, rg_mark_ref1(a.sen1, a.thr, a.deep, a.dcnt, a.ref1, items(c, 1), c->st));
    HCK(c, rg_differs(a.ref1, a.o_ref1, items(c, 1), a.dcnt, CNT_DIFF1, c->st));
    if (int r = scan_l2(c, true)) return r;
    // Ties at the threshold can exceed the level-2 pool (the reference writes out of bounds there): the
    // pool grows when the budget allows; the refinement target cap_nref1 stays as measured at init.
    bool grown = false;
    if (a.nf2 > a.cap2) {
      int cn = std::max(a.nref1, (a.cap2 / 8) * 3 / 2);
      while (cn > a.nref1 && !pool_fits(a, a.cap1, 8L * cn)) cn--;
      if (!pool_fits(a, a.cap1, 8L * cn))
        return capacity_fail(c, "level-2 count " + std::to_string(a.nf2) + " exceeds the pool capacity " +
                                            std::to_string(a.cap2) + " and the budget (threshold ties); raise --maxgb");
      HCK(c, hipStreamSynchronize(c->st));
      void* keep = a.f2a;  // old level-2 data, read by the migration below
      for (void* p : {a.f2b, a.gh2, static_cast<void*>(a.l2pl1), static_cast<void*>(a.l2oct), static_cast<void*>(a.nb2)})
        ctx_free(c, p);
      a.cap2 = 8 * cn;
      if (int r = alloc_l2_pool(c)) return r;
      ctx_free(c, a.f2a);  // alloc_l2_pool made both buffers; the old data stays in `keep`
      a.f2a = keep;
      grown = true;
    }
    if (grown) {  // the scan and the scatter ahead of the count read stopped at the old pool
      if (int r = fill(c, a.dcnt + CNT_OVERFLOW, 0, 1)) return r;
      if (int r = fill(c, a.dcnt + CNT_NF2, a.nf2, 1)) return r;
      a.hcnt[CNT_OVERFLOW] = 0;
      if (int r = scatter_l2(c)) return r;
    }
    if (a.nf2 && !(same1 && !a.hcnt[CNT_DIFF1] && !grown)) {
      HCK(c, amr_migrate_l2(c->prec, G, a.f2a, a.f2b, a.f1a, a.l2pl1, a.l2oct, a.f1pos, a.o_l1grid, a.o_ref1, a.o_fof1,
                            a.l1grid, tiles(c, 2), float(c->d.nqs12), dec(c, 1), enc(c, 2), c->st));
      std::swap(a.f2a, a.f2b);
    }
    if (grown) {  // the other buffer still has the old size
      HCK(c, hipStreamSynchronize(c->st));
      ctx_free(c, a.f2b);
      if (int r = ctx_alloc(c, &a.f2b, size_t(a.cap2) * Q * MB3 * prec_bytes(c->prec), "f2b")) return r;
      ++a.pool_gen;
    }
  }
  if (int r = finish_tables(c)) return r;
  if (a.hcnt[CNT_OVERFLOW]) return capacity_fail(c, "pool capacity exceeded while scattering tables");
  a.w1 = width8(a.nf1);
  a.w2 = width8(a.nf2);
  return JP_OK;
}

// ---- device-resident regrids ----
// A buffer replaced by a larger one, its first `keep` bytes carried over.
static int realloc_keep(jp_ctx* c, void** p, size_t bytes, size_t keep, const char* what) {
  void* n = nullptr;
  if (int r = ctx_alloc(c, &n, bytes, what)) return r;
  if (keep) HCK(c, hipMemcpyAsync(n, *p, keep, hipMemcpyDeviceToDevice, c->st));
  HCK(c, hipStreamSynchronize(c->st));
  ctx_free(c, *p);
  *p = n;
  return JP_OK;
}
template <class T> static int realloc_keep_t(jp_ctx* c, T** p, size_t count, size_t keep, const char* what) {
  void* v = *p;
  int r = realloc_keep(c, &v, count * sizeof(T), keep * sizeof(T), what);
  *p = static_cast<T*>(v);
  return r;
}

// A regrid on device-resident counts cannot grow a pool (the host learns the counts only afterwards),
// so the fine pools are sized once for the largest counts a regrid can produce: every level-0 tile
// refined (8 cnb0 level-1 tiles) and every level-1 tile refined (8 level-2 tiles each), as far as the
// budget allows. The current populations and level-2 tables are carried over; ghost records and the
// second buffers are rewritten before they are read. Counts beyond a budget-limited pool still raise the
// sticky overflow flag, which the next count read reports.
static int apply_resident(jp_ctx* c) {
  AmrState& a = c->amr;
  if (int r = amr_sync_counts(c)) return r;
  a.resident = false;
  if (!a.resident_req || !amr_dynamic(c)) return JP_OK;
  const double per_tile = double(Q) * 4 * (2 * MB3 + 1000) * 6 / 5;  // pool_fits' bytes per fine tile
  const long room = std::max(0L, long(a.budget / per_tile) - 1);     // fine tiles of both levels
  const long c1 = std::max<long>(a.cap1, std::min<long>({8L * a.cnb0, (room - a.cap2_alloc) / 8 * 8, kMaxFineTiles}));
  const long c2 = std::max<long>(a.cap2_alloc, std::min<long>({8L * c1, (room - c1) / 8 * 8, kMaxFineTiles}));
  const size_t e = prec_bytes(c->prec), r = real_bytes(c->prec), tile = size_t(Q) * MB3;
  if (c1 > a.cap1) {
    if (int rc = realloc_keep(c, &a.f1a, size_t(c1) * tile * e, size_t(a.nf1) * tile * e, "f1a")) return rc;
    if (int rc = realloc_keep(c, &a.f1b, size_t(c1) * tile * e, 0, "f1b")) return rc;
    if (int rc = realloc_keep(c, &a.gh1, size_t(c1) * kGhostSlots * r, 0, "gh1")) return rc;
    a.cap1 = int(c1);
    ++a.pool_gen;
  }
  if (c2 > a.cap2_alloc) {
    if (int rc = realloc_keep(c, &a.f2a, size_t(c2) * tile * e, size_t(a.nf2) * tile * e, "f2a")) return rc;
    if (int rc = realloc_keep(c, &a.f2b, size_t(c2) * tile * e, 0, "f2b")) return rc;
    if (int rc = realloc_keep(c, &a.gh2, size_t(c2) * kGhostSlots * r, 0, "gh2")) return rc;
    if (int rc = realloc_keep_t(c, &a.l2pl1, size_t(c2), size_t(a.nf2), "l2pl1")) return rc;
    if (int rc = realloc_keep_t(c, &a.l2oct, size_t(c2), size_t(a.nf2), "l2oct")) return rc;
    if (int rc = realloc_keep_t(c, &a.nb2, size_t(c2) * kNbr, size_t(a.nf2) * kNbr, "nb2")) return rc;
    a.cap2_alloc = int(c2);
    ++a.pool_gen;
  }
  a.cap2 = a.cap2_alloc;
  a.resident = true;
  a.w1 = width8(a.nf1);
  a.w2 = width8(a.nf2);
  drop_graph(c);  // launches recorded before sized their level-2 work on the host's counts
  return JP_OK;
}
hipError_t amr_log_counts(jp_ctx* c, int* dst) {
  static_assert(CNT_NF2 == CNT_NF1 + 1, "the two tile counts are adjacent");
  return hipMemcpyAsync(dst, c->amr.dcnt + CNT_NF1, 2 * sizeof(int), hipMemcpyDeviceToHost, c->st);
}
int amr_set_resident(jp_ctx* c, bool on) {
  c->amr.resident_req = on;
  return c->initialised ? apply_resident(c) : JP_OK;  // else applied by the initial build
}

// ---- one coarse step of the nested 1:2:4 subcycle, src/tgv3d_amr.cu:821-837 ----
// The level-2 pair of one level-1 substep: ghost fill from level 1, two level-2 substeps, the second
// restricting into level 1 in its epilogue (src/tgv3d_amr.cu:828-833).
static int level2_pair(jp_ctx* c, const AmrGeo& G, hipStream_t st) {
  AmrState& a = c->amr;
  const int p = c->prec;
  const Tiles t2 = tiles(c, 2);
  if (a.skipprep)
    HCK(c, amr_fine_step(p, 2, a.f2a, a.f2b, a.gh2, a.nb2, t2, c->d.omega2, enc(c, 2), dec(c, 2), st));
  else  // fill and first substep in one launch
    HCK(c, amr_fill_step_l2(p, G, a.f1a, a.gh2, a.l2pl1, a.l2oct, a.f1pos, a.nb1, a.nb2, a.slots, t2, float(c->d.nqs12),
                            dec(c, 1), a.f2a, a.f2b, c->d.omega2, enc(c, 2), dec(c, 2), st));
  std::swap(a.f2a, a.f2b);
  if (a.skiprestr)
    HCK(c, amr_fine_step(p, 2, a.f2a, a.f2b, a.gh2, a.nb2, t2, c->d.omega2, enc(c, 2), dec(c, 2), st));
  else
    HCK(c, amr_fine_step_restrict(p, a.f2a, a.f2b, a.gh2, a.nb2, t2, c->d.omega2, enc(c, 2), dec(c, 2), a.f1a, a.l2pl1,
                                  float(c->d.nqi21), enc(c, 1), st));
  std::swap(a.f2a, a.f2b);
  return JP_OK;
}

// One coarse step (src/tgv3d_amr.cu:821-837). The first level-2 pair reads and writes only levels 1 and
// 2 (its restriction touches the interior of refined level-1 tiles), while the level-0 step and the
// level-1 fill touch only level 0 and the level-1 ghost record, so the two run concurrently: as two
// launches that each carry one part of both (the level-0 step with the pair's fused fill and first
// substep, then the pair's restricting substep with the level-1 fill; see amr.hip), or, when the
// profiling switches split the pair, on a second stream that the first level-1 substep waits for. Every
// buffer is written by at most one of the two and read by the other only after both, so the results are
// those of the sequential order. With device-resident regrids the host does not know whether a level is
// empty, so every launch is made (an empty level's workgroups read a zero count and return).
int amr_base_step(jp_ctx* c, hipStream_t st) {
  AmrState& a = c->amr;
  const AmrGeo G = geo(c);
  const int p = c->prec;
  const bool l2 = a.resident || a.nf2 > 0;
  const Tiles t1 = tiles(c, 1), t2 = tiles(c, 2);
  if (l2 && !a.skipprep && !a.skiprestr) {
    HCK(c, amr_dense_fill_step_l2(p, G, c->c0a, c->c0b, a.nb0, c->d.omega0, enc(c, 0), dec(c, 0), a.f1a, a.gh2, a.l2pl1, a.l2oct,
                                  a.f1pos, a.nb1, a.nb2, a.slots, t2, float(c->d.nqs12), dec(c, 1), a.f2a, a.f2b,
                                  c->d.omega2, enc(c, 2), dec(c, 2), st));
    std::swap(c->c0a, c->c0b);
    std::swap(a.f2a, a.f2b);
    // prep of level 0 (src/tgv3d_amr.cu:826) is computed inside the fill
    HCK(c, amr_step_r_l2_fill_l1(p, G, a.f2a, a.f2b, a.gh2, a.nb2, t2, c->d.omega2, enc(c, 2), dec(c, 2), a.f1a, a.l2pl1,
                                 float(c->d.nqi21), enc(c, 1), c->c0a, a.gh1, a.l1pcb, a.l1oct, a.nb1, a.slots, t1,
                                 float(c->d.nqs01), dec(c, 0), st));
    std::swap(a.f2a, a.f2b);
  } else {
    hipEvent_t ev_fork = a.ev_fork[a.evi], ev_join = a.ev_join[a.evi];
    a.evi ^= 1;  // two event pairs: consecutive steps in one captured graph do not reuse an event
    if (l2) {
      HCK(c, hipEventRecord(ev_fork, st));
      HCK(c, hipStreamWaitEvent(a.st2, ev_fork, 0));
      if (int r = level2_pair(c, G, a.st2)) return r;
      HCK(c, hipEventRecord(ev_join, a.st2));
    }
    if (int r = l0_step(c, st)) return r;
    if (!a.resident && !a.nf1) return JP_OK;
    if (!a.skipprep)
      HCK(c, amr_fill_l1(p, G, c->c0a, a.gh1, a.l1pcb, a.l1oct, a.nb1, a.slots, t1, float(c->d.nqs01), dec(c, 0), enc(c, 1), st));
    if (l2) HCK(c, hipStreamWaitEvent(st, ev_join, 0));
  }
  HCK(c, amr_fine_step(p, 1, a.f1a, a.f1b, a.gh1, a.nb1, t1, c->d.omega1, enc(c, 1), dec(c, 1), st));
  std::swap(a.f1a, a.f1b);
  if (l2)
    if (int r = level2_pair(c, G, st)) return r;
  if (a.skiprestr)
    HCK(c, amr_fine_step(p, 1, a.f1a, a.f1b, a.gh1, a.nb1, t1, c->d.omega1, enc(c, 1), dec(c, 1), st));
  else  // the second level-1 substep restricts into level 0
    HCK(c, amr_fine_step_restrict(p, a.f1a, a.f1b, a.gh1, a.nb1, t1, c->d.omega1, enc(c, 1), dec(c, 1), c->c0a, a.l1pcb,
                                  float(c->d.nqi10), enc(c, 0), st));
  std::swap(a.f1a, a.f1b);
  return JP_OK;
}

int amr_composite(jp_ctx* c, void* field) {
  AmrState& a = c->amr;
  const float d3[3] = {dec(c, 0), dec(c, 1), dec(c, 2)};
  HCK(c, amr_composite_field(c->prec, geo(c), c->c0a, a.f1a, a.f2a, a.ref1, a.fof1, a.l1grid, d3, field, c->st));
  return JP_OK;
}

// Creates the context: level parameters (src/tgv3d_amr.cu:658-664) and pools (675-709).
int amr_create(const jp_amr_params* p, jp_ctx* c) {
  AmrState& a = c->amr;
  const int N = p->n;
  a.N = N;
  a.cgx = N / MB;
  a.cnb0 = a.cgx * a.cgx * a.cgx;
  a.l1gx = 2 * a.cgx;
  a.N1 = 2 * N;
  a.N2 = 4 * N;
  a.NG1 = a.l1gx * a.l1gx * a.l1gx;
  a.l1_mode = p->l1_mode;
  a.l2_mode = p->l2_mode;
  a.frac1 = p->frac1;
  a.frac2 = p->frac2;
  a.adapt = p->adapt;
  a.skipprep = p->skipprep;
  a.skiprestr = p->skiprestr;
  jp_derived& d = c->d;
  d.n = N;
  d.prec = p->prec;
  d.nu = p->u0 * double(N) / (2.0 * M_PI * p->re);
  d.tau0 = 3.0 * d.nu + 0.5;
  d.tau1 = 2.0 * d.tau0 - 0.5;
  d.tau2 = 2.0 * d.tau1 - 0.5;
  const bool f64 = p->prec == JP_PREC_FP64;
  d.omega0 = f64 ? 1.0 / d.tau0 : double(float(1.0 / d.tau0));
  d.omega1 = f64 ? 1.0 / d.tau1 : double(float(1.0 / d.tau1));
  d.omega2 = f64 ? 1.0 / d.tau2 : double(float(1.0 / d.tau2));
  d.nqs01 = float((d.tau1 - 1.0) / (2.0 * (d.tau0 - 1.0)));
  d.nqi10 = float((2.0 * (d.tau0 - 1.0)) / (d.tau1 - 1.0));
  d.nqs12 = float((d.tau2 - 1.0) / (2.0 * (d.tau1 - 1.0)));
  d.nqi21 = float((2.0 * (d.tau1 - 1.0)) / (d.tau2 - 1.0));
  d.kf = 2.0 * M_PI / double(N);
  d.cells = long(N) * N * N;
  const bool shifted = p->prec == JP_PREC_FP16 || p->prec == JP_PREC_BF16;
  for (int l = 0; l < 3; ++l) {
    const double s = shifted ? (p->scale[l] > 0 ? p->scale[l] : default_scale(p->u0)) : 1.0;
    if (s != std::exp2(std::round(std::log2(s)))) return fail(c, JP_ERR_ARG, "scale must be a power of two");
    d.scale[l] = s;
  }
  size_t freeB = 0, totB = 0;
  HCK(c, hipMemGetInfo(&freeB, &totB));
  a.budget = p->maxgb > 0 ? p->maxgb * double(1ull << 30) : double(freeB) * 0.75;
  a.cap_nref0 = (a.l1_mode == JP_MODE_ALL) ? a.cnb0 : std::min(a.cnb0, std::max(1, int(std::ceil(a.frac1 * a.cnb0 * 1.4))));
  while (a.cap_nref0 > 1 && !pool_fits(a, 8L * a.cap_nref0, 0)) a.cap_nref0--;
  a.cap1 = 8 * a.cap_nref0;
  a.cap_nref1 = INT_MAX;
  HCK(c, hipStreamCreate(&a.st2));  // default priority, like every stream of the library
  for (int i = 0; i < 2; ++i) {
    HCK(c, hipEventCreateWithFlags(&a.ev_fork[i], hipEventDisableTiming));
    HCK(c, hipEventCreateWithFlags(&a.ev_join[i], hipEventDisableTiming));
  }
  // pools
  // Level 0 tile-major: cnb0 tiles of 19 x 512 elements, and its periodic neighbour table
  const size_t e = prec_bytes(c->prec);
  const size_t dense_b = size_t(a.cnb0) * Q * MB3;
  if (int rc = ctx_alloc(c, &c->c0a, dense_b * e, "c0a")) return rc;
  if (int rc = ctx_alloc(c, &c->c0b, dense_b * e, "c0b")) return rc;
  if (int rc = ctx_alloc_t(c, &a.nb0, size_t(a.cnb0) * kNbr, "nb0")) return rc;
  HCK(c, amr_l0_neighbours(a.nb0, a.cgx, c->st));
  if (int rc = alloc_l1_pool(c, true)) return rc;
  // level-1 tables are sized for full refinement (8 cnb0 tiles) so that only the pools ever grow
  const size_t n1 = size_t(8) * a.cnb0, n01 = n1;
  struct T {
    int** p;
    size_t n;
    const char* w;
  } tabs[] = {{&a.ref0, size_t(a.cnb0), "ref0"},     {&a.fof0, size_t(a.cnb0), "fof0"},     {&a.o_ref0, size_t(a.cnb0), "o_ref0"},
              {&a.o_fof0, size_t(a.cnb0), "o_fof0"}, {&a.skin0, size_t(a.cnb0), "skin0"},   {&a.l1pcb, n1, "l1pcb"},
              {&a.l1oct, n1, "l1oct"},   {&a.rl0, n1, "rl0"},       {&a.f1pos, n1, "f1pos"},
              {&a.ref1, n1, "ref1"},     {&a.fof1, n1, "fof1"},     {&a.o_ref1, n1, "o_ref1"},
              {&a.o_fof1, n1, "o_fof1"}, {&a.rl1, n1, "rl1"},       {&a.deep, n1, "deep"},
              {&a.skin1, n1, "skin1"},   {&a.l1grid, size_t(a.NG1), "l1grid"},  {&a.o_l1grid, size_t(a.NG1), "o_l1grid"},
              {&a.rank, n01, "rank"},                {&a.flags, n01, "flags"},               {&a.nb1, n1 * kNbr, "nb1"},
              {&a.dcnt, size_t(CNT_N), "counts"}};
  for (auto& t : tabs)
    if (int rc = ctx_alloc_t(c, t.p, t.n, t.w)) return rc;
  if (int rc = ctx_alloc_t(c, &a.sen0, size_t(a.cnb0), "sen0")) return rc;
  if (int rc = ctx_alloc_t(c, &a.sen1, n1, "sen1")) return rc;
  if (int rc = ctx_alloc_t(c, &a.sen1pos, size_t(a.NG1), "sen1 by position")) return rc;
  if (int rc = ctx_alloc_t(c, &a.thr, 4, "threshold")) return rc;
  if (int rc = ctx_alloc_t(c, &a.slots, size_t(kGhostSlots), "ghost fill table")) return rc;
  const std::vector<uint32_t> slots = ghost_fill_table();
  HCK(c, hipMemcpy(a.slots, slots.data(), slots.size() * 4, hipMemcpyHostToDevice));
  HCK(c, hipHostMalloc(reinterpret_cast<void**>(&a.hcnt), sizeof(int) * CNT_N, hipHostMallocDefault));
  std::memset(a.hcnt, 0, sizeof(int) * CNT_N);
  // tables start empty
  if (int rc = fill(c, a.ref1, 0, size_t(a.cap1))) return rc;
  if (int rc = fill(c, a.fof1, -1, size_t(a.cap1))) return rc;
  if (int rc = fill(c, a.l1grid, -1, size_t(a.NG1))) return rc;
  HCK(c, hipStreamSynchronize(c->st));
  return JP_OK;
}

int amr_counts(jp_ctx* c, jp_counts* o) {
  if (int r = ensure_skins(c)) return r;
  const AmrState& a = c->amr;
  o->cnb0 = a.cnb0;
  o->nref0 = a.nref0;
  o->nf1 = a.nf1;
  o->nref1 = a.nref1;
  o->nf2 = a.nf2;
  o->n_skin0 = a.n_skin0;
  o->n_skin1 = a.n_skin1;
  o->cap_nref0 = a.cap_nref0;
  o->cap1 = a.cap1;
  o->cap_nref1 = a.cap_nref1 == INT_MAX ? -1 : a.cap_nref1;
  o->cap2 = a.cap2;
  o->hw_nf1 = a.hw_nf1;
  o->hw_nf2 = a.hw_nf2;
  o->ndeep = a.hcnt ? a.hcnt[CNT_NDEEP] : 0;
  return JP_OK;
}

int amr_table(jp_ctx* c, int which, void* dst, long cap, long* n) {
  if (int r = amr_sync_counts(c)) return r;
  if (which == JP_TABLE_SKIN0 || which == JP_TABLE_SKIN1)
    if (int r = ensure_skins(c)) return r;
  const AmrState& a = c->amr;
  const void* src = nullptr;
  long len = 0;
  switch (which) {
    case JP_TABLE_REF0: src = a.ref0; len = a.cnb0; break;
    case JP_TABLE_FOF0: src = a.fof0; len = a.cnb0; break;
    case JP_TABLE_RL0: src = a.rl0; len = a.nref0; break;
    case JP_TABLE_L1PCB: src = a.l1pcb; len = a.nf1; break;
    case JP_TABLE_L1OCT: src = a.l1oct; len = a.nf1; break;
    case JP_TABLE_F1POS: src = a.f1pos; len = a.nf1; break;
    case JP_TABLE_L1GRID: src = a.l1grid; len = a.NG1; break;
    case JP_TABLE_REF1: src = a.ref1; len = a.nf1; break;
    case JP_TABLE_FOF1: src = a.fof1; len = a.nf1; break;
    case JP_TABLE_RL1: src = a.rl1; len = a.nref1; break;
    case JP_TABLE_L2PL1: src = a.l2pl1; len = a.nf2; break;
    case JP_TABLE_L2OCT: src = a.l2oct; len = a.nf2; break;
    case JP_TABLE_SKIN0: src = a.skin0; len = a.n_skin0; break;
    case JP_TABLE_SKIN1: src = a.skin1; len = a.n_skin1; break;
    case JP_TABLE_SEN0: src = a.sen0; len = a.cnb0; break;
    case JP_TABLE_SEN1: src = a.sen1; len = a.nf1; break;
    default: return fail(c, JP_ERR_ARG, "unknown table id");
  }
  if (n) *n = len;
  if (len > cap) return fail(c, JP_ERR_ARG, "table buffer too small");
  if (len > 0) {
    HCK(c, hipMemcpyAsync(dst, src, size_t(len) * 4, hipMemcpyDeviceToHost, c->st));
    HCK(c, hipStreamSynchronize(c->st));
  }
  return JP_OK;
}

// Work of one coarse step in cell updates, fine levels weighted by their subcycles (src/tgv3d_amr.cu:843).
double amr_work(const jp_ctx* c) { return double(c->amr.cnb0 + 2L * c->amr.nf1 + 4L * c->amr.nf2) * MB3; }
bool amr_dynamic(const jp_ctx* c) {
  return c->amr.l1_mode == JP_MODE_SENSOR || c->amr.l2_mode == JP_MODE_SENSOR;
}

}  // namespace jp
