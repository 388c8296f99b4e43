// Context object behind the opaque jp_ctx handle of include/jp.h.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/jp.h"
#include "internal.hpp"

namespace jp {

struct EnsState;

// Device state of the 3-level block AMR (tables follow src/tgv3d_amr.cu:692-709).
struct AmrState {
  int N = 0, cgx = 0, cnb0 = 0, l1gx = 0, N1 = 0, N2 = 0, NG1 = 0;
  int l1_mode = 0, l2_mode = 0, skipprep = 0, skiprestr = 0, adapt = 100;
  double frac1 = 0.35, frac2 = 0.35, budget = 0;
  int cap_nref0 = 0, cap1 = 0, cap_nref1 = 0, cap2 = 0;
  long hw_nf1 = 0, hw_nf2 = 0;
  int nref0 = 0, nf1 = 0, nref1 = 0, nf2 = 0, n_skin0 = 0, n_skin1 = 0;
  bool f1_stale = false;
  bool skins_stale = true;  // skin lists and their counts are rebuilt on demand after a regrid
  int cap2_alloc = 0;  // level-2 tiles the allocated pool holds
  int pool_gen = 0;    // bumped whenever a fine-level pool is reallocated (batched descriptors go stale)
  bool resident_req = false;  // regrids on device-resident counts requested (graphs on, sensor-driven grid)
  bool resident = false;      // ... and in effect: pools sized for them, every launch sized on the device
  bool counts_stale = false;  // a device-resident regrid ran since the host last read the counts
  bool dev_counts = false;    // launches read their tile counts on the device (during a regrid's pipeline)
  int w1 = 8, w2 = 8;         // launch widths of the level-1 / level-2 launches while resident
  hipStream_t st2 = nullptr;  // side stream of the coarse step (first level-2 pair)
  hipEvent_t ev_fork[2] = {}, ev_join[2] = {};
  int evi = 0;  // f1b still has a pre-growth size (until the regrid's migration completes)
  void *f1a = nullptr, *f1b = nullptr, *f2a = nullptr, *f2b = nullptr;  // stored populations
  void *gh1 = nullptr, *gh2 = nullptr;  // ghost records, working-type values
  int *ref0 = nullptr, *fof0 = nullptr, *rl0 = nullptr, *l1pcb = nullptr, *l1oct = nullptr, *f1pos = nullptr;
  int *l1grid = nullptr, *ref1 = nullptr, *fof1 = nullptr, *rl1 = nullptr, *l2pl1 = nullptr, *l2oct = nullptr;
  int *skin0 = nullptr, *skin1 = nullptr, *o_ref0 = nullptr, *o_fof0 = nullptr, *o_l1grid = nullptr;
  int *o_ref1 = nullptr, *o_fof1 = nullptr, *rank = nullptr, *deep = nullptr, *flags = nullptr;
  int *nb1 = nullptr, *nb2 = nullptr;  // 27 same-level neighbour tiles per fine tile (-1 = absent)
  int* nb0 = nullptr;                  // 27 neighbour tiles per level-0 tile (periodic, all present)
  float *sen0 = nullptr, *sen1 = nullptr, *sen1pos = nullptr, *thr = nullptr;
  uint32_t* slots = nullptr;  // ghost fill table
  int* dcnt = nullptr;  // device-resident counts and flags (see amr.hip)
  int* hcnt = nullptr;  // pinned host mirror of dcnt
};

}  // namespace jp

struct jp_ctx {
  int kind = 0;  // 0 uniform, 1 AMR, 2 batched ensemble
  int device = 0, prec = 0, variant = 0;
  bool initialised = false, use_graph = false;
  bool step_pairs = true;  // uniform lattice: two fused steps per launch where a column shape applies
  int pair_pick = -1;      // column shape of the fused pairs (-1: the first in preference order that fits)
  hipStream_t st = nullptr;
  std::string err;
  jp_derived d{};
  double u0 = 0.05, re = 1600, phase[3] = {0, 0, 0};
  jp::Dense g0{};
  void *c0a = nullptr, *c0b = nullptr;  // dense lattice (g0), or AMR level 0 (tile-major, cnb0 tiles)
  void* field = nullptr;
  size_t field_bytes = 0;
  float* out = nullptr;  // output-field arrays (Q, speed, velocity) of the last jp_field
  size_t out_bytes = 0;
  float* host_out = nullptr;  // pinned host copy of them for the frame writer
  size_t host_out_bytes = 0;
  double* partial = nullptr;  // diagnostic partial sums (3 doubles per slot), then the 3 results
  size_t partial_slots = 0;
  // Captured runs of an even number of steps, keyed by their length and the buffers they start from (the
  // ping-pong roles), so that an odd step or a device-resident regrid selects another run instead of
  // recapturing. c0 holds the dense pair's roles after the captured steps (two single steps restore them,
  // a fused pair swaps them once), which a replay applies.
  struct Graph {
    const void* key[3];
    long len;  // steps captured (even)
    hipGraphExec_t exec;
    void* c0[2];
  };
  std::vector<Graph> graphs;
  std::vector<std::pair<void*, size_t>> allocs;
  size_t pool_bytes = 0;
  jp::AmrState amr;
  jp::EnsState* ens = nullptr;  // batched ensemble (kind 2)
};

namespace jp {
int fail(jp_ctx* c, int code, const std::string& msg);
int ctx_alloc(jp_ctx* c, void** p, size_t bytes, const char* what);
template <class T> int ctx_alloc_t(jp_ctx* c, T** p, size_t count, const char* what) {
  void* v = nullptr;
  int r = ctx_alloc(c, &v, count * sizeof(T), what);
  *p = static_cast<T*>(v);
  return r;
}
void ctx_free(jp_ctx* c, void* p);
int drop_graph(jp_ctx* c);
double default_scale(double u0);
int default_variant(int prec);
// AMR orchestration (amr_host.cpp)
int amr_init(jp_ctx* c);
int amr_base_step(jp_ctx* c, hipStream_t st);
int amr_composite(jp_ctx* c, void* field);
int amr_create(const jp_amr_params* p, jp_ctx* c);
int amr_init_with(jp_ctx* c, const float* sen0, const float* sen1_by_pos);
int amr_regrid_with(jp_ctx* c, const float* sen0, const float* sen1_by_pos);
int amr_counts(jp_ctx* c, jp_counts* o);
int amr_sync_counts(jp_ctx* c);
hipError_t amr_log_counts(jp_ctx* c, int* dst);  // level-1 and level-2 tile counts to pinned dst[2], async
int amr_set_resident(jp_ctx* c, bool on);
int amr_table(jp_ctx* c, int which, void* dst, long cap, long* n);
double amr_work(const jp_ctx* c);
bool amr_dynamic(const jp_ctx* c);
// batched ensemble (ensemble_host.cpp)
int ens_create(const jp_member* mem, int count, int prec, double maxgb, int device, jp_ctx* c);
int ens_init(jp_ctx* c);
int ens_check(jp_ctx* c);
int ens_step_once(jp_ctx* c);
int ens_regrid(jp_ctx* c);
bool ens_dynamic(const jp_ctx* c);
int ens_adapt(const jp_ctx* c);
double ens_work(const jp_ctx* c);
int ens_autotune(jp_ctx* c, int* out);
int ens_dense_shape(jp_ctx* c, int set, int* current);  // the dense step's cells per lane (1 or 8)
void ens_destroy(jp_ctx* c);
void ens_graph_key(const jp_ctx* c, const void* key[3]);  // the batch's buffer parities
jp_ctx* ens_member(jp_ctx* c, int i);
}  // namespace jp

#define HCK(c, x)                                                                                 \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess) return ::jp::fail((c), JP_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
