// Launch interface of the AMR kernels (amr.hip, regrid.hip) used by the orchestration in amr_host.cpp.
//
// Level 0 is the coarse lattice, stored tile-major in 8^3 tiles like the fine levels (all cgx^3 tiles, in
// raster order): a restriction then stores its parents as whole 256 B blocks, where direction planes
// would take 32 B rows. Levels 1 and 2 are scattered tiles of 8^3 cells stored tile-major. Each fine tile keeps a compact ghost record of the 1824 (q, cell) pairs whose
// pull source lies outside the tile (64 per axis direction, 120 per diagonal direction), instead of the
// reference's 19 x 1000 ghost box (src/tgv3d_amr.cu:179-185): the step reads exactly those entries.
#pragma once
#include "internal.hpp"

namespace jp {

constexpr int kGhostSlots = 1824;
constexpr int kNbr = 27;
// Most tiles a fine level's pool holds: the fine steps address populations with 32-bit element offsets
// (tile * 19 * 512 + offset < 2^32), a multiple of 8 (whole octets).
constexpr long kMaxFineTiles = 441504;

// Ghost fills interpolate from the 6^3 parent-level box starting at parent coordinate o/2 - 1 of a tile
// at fine origin o. Fill-table entry of one ghost slot, source (si, sj, sk) in [-1, 8]^3 relative to the
// tile: q | dir << 5 | box << 10 | (si & 1) << 18 | (sj & 1) << 19 | (sk & 1) << 20, with dir the
// 27-neighbour index of the tile holding the source and box the local index of the lowest trilinear
// corner. Per axis the upper corner weighs t = 1/4 for an odd source coordinate and t = 3/4 for an even
// one, the lower corner 1 - t.
constexpr int kFillBox = 6;
// Slot of (q, destination cell) in a tile's ghost record; valid when the source cell - c_q is outside.
__host__ __device__ inline int ghost_base(int q) { return q < 7 ? (q - 1) * 64 : 384 + (q - 7) * 120; }
__host__ __device__ inline int ghost_slot(int q, int i, int j, int k) {
  const int cx = CX(q), cy = CY(q), cz = CZ(q);
  const int nz = (cx != 0) + (cy != 0) + (cz != 0);
  if (nz == 1) {
    if (cx != 0) return ghost_base(q) + k * 8 + j;
    if (cy != 0) return ghost_base(q) + k * 8 + i;
    return ghost_base(q) + j * 8 + i;
  }
  int ia, ib, it, ca;
  if (cz == 0) {
    ia = i; ib = j; it = k; ca = cx;
  } else if (cy == 0) {
    ia = i; ib = k; it = j; ca = cx;
  } else {
    ia = j; ib = k; it = i; ca = cy;
  }
  const int ea = ca > 0 ? 0 : 7;
  const int r = (ia == ea) ? ib : 8 + (ca > 0 ? ia - 1 : ia);
  return ghost_base(q) + it * 15 + r;
}

// Geometry shared by the AMR kernels.
struct AmrGeo {
  int N0, cgx, l1gx, N1, N2;
};

// Tiles of a per-tile launch. count points at the device-resident tile count (null: the width itself,
// one workgroup per tile); width is the number of workgroups. Launches recorded in a graph that a regrid
// must leave valid read the count and take a width fixed when they are recorded; the kernels loop when
// the count exceeds it.
struct Tiles {
  const int* count;
  int width;
};

// ---- amr.hip ----
hipError_t amr_fine_step(int prec, int level, const void* fa, void* fb, const void* gh, const int* nb, Tiles t,
                         double omega, float enc, float dec, hipStream_t st);
// fine step + restriction into the parent level cf (ptile: l1pcb or l2pl1); t covers whole octets of tiles
hipError_t amr_fine_step_restrict(int prec, const void* fa, void* fb, const void* gh, const int* nb, Tiles t, double omega,
                                  float enc, float dec, void* cf, const int* ptile, float nqi, float cenc, hipStream_t st);
// Level 0: TGV initial state, and the periodic 27-neighbour table that its step (amr_fine_step with gh
// null) reads.
hipError_t amr_l0_init(int prec, const AmrGeo& G, void* f, double u0, double kf, const double* phase, float enc,
                       hipStream_t st);
hipError_t amr_l0_neighbours(int* nb0, int cgx, hipStream_t st);
// Ghost records of level 1 (from level 0, c0): prep values of the parent box computed on the fly,
// trilinear interpolation of the slots whose source tile is absent, stored in the working representation
// of the form at level 1's scale enc1.
hipError_t amr_fill_l1(int prec, const AmrGeo& G, const void* c0, void* gh1, const int* l1pcb, const int* l1oct,
                       const int* nb1, const uint32_t* slots, Tiles t, float nqs, float dec, float enc1, hipStream_t st);
// The level-2 fill fused with the first level-2 substep (fa -> fb): the ghost record is written to gh2 for
// the second substep and read from LDS by this one.
hipError_t amr_fill_step_l2(int prec, const AmrGeo& G, const void* f1, void* gh2, const int* l2pl1, const int* l2oct,
                            const int* f1pos, const int* nb1, const int* nb2, const uint32_t* slots, Tiles t, float nqs,
                            float dec1, const void* fa, void* fb, double omega, float enc2, float dec2, hipStream_t st);
// The first two launches of a coarse step with level 2 present, each carrying two independent parts: the
// level-0 step (c0a -> c0b) with the fused level-2 fill and first substep (f2a -> f2b), then the second
// level-2 substep with its restriction into level 1 (f2a -> f2b, into f1) with the level-1 ghost fill
// from the new level 0 (c0).
hipError_t amr_dense_fill_step_l2(int prec, const AmrGeo& G, const void* c0a, void* c0b, const int* nb0, double om0,
                                  float enc0, float dec0, const void* f1, void* gh2, const int* l2pl1, const int* l2oct,
                                  const int* f1pos, const int* nb1, const int* nb2, const uint32_t* slots, Tiles t2,
                                  float nqs12, float dec1, const void* f2a, void* f2b, double om2, float enc2, float dec2,
                                  hipStream_t st);
hipError_t amr_step_r_l2_fill_l1(int prec, const AmrGeo& G, const void* f2a, void* f2b, const void* gh2, const int* nb2,
                                 Tiles t2, double om2, float enc2, float dec2, void* f1, const int* l2pl1,
                                 float nqi21, float enc1, const void* c0, void* gh1, const int* l1pcb, const int* l1oct,
                                 const int* nb1, const uint32_t* slots, Tiles t1, float nqs01, float dec0, hipStream_t st);
hipError_t amr_seed_l1(int prec, const AmrGeo& G, const void* c0, void* f1, const int* l1pcb, const int* l1oct,
                       Tiles t, float nqs, float dec0, float enc1, hipStream_t st);
hipError_t amr_seed_l2(int prec, const AmrGeo& G, const void* f1, void* f2, const int* l2pl1, const int* l2oct,
                       const int* f1pos, const int* l1grid, Tiles t, float nqs, float dec1, float enc2,
                       hipStream_t st);
hipError_t amr_migrate_l1(int prec, const AmrGeo& G, const void* oldf1, void* newf1, const void* c0,
                          const int* l1pcb, const int* l1oct, const int* o_ref0, const int* o_fof0, Tiles t, float nqs,
                          float dec0, float enc1, hipStream_t st);
hipError_t amr_migrate_l2(int prec, const AmrGeo& G, const void* oldf2, void* newf2, const void* newf1,
                          const int* l2pl1, const int* l2oct, const int* f1pos, const int* o_l1grid, const int* o_ref1,
                          const int* o_fof1, const int* l1grid, Tiles t, float nqs, float dec1, float enc2,
                          hipStream_t st);
hipError_t amr_composite_field(int prec, const AmrGeo& G, const void* c0, const void* f1, const void* f2,
                               const int* ref1, const int* fof1, const int* l1grid, const float* dec3, void* F,
                               hipStream_t st);
hipError_t amr_sensor(int prec, const AmrGeo& G, const void* f, Tiles t, float dec, float* sen, hipStream_t st);

// ---- regrid.hip ----
// Device counters (int slots of AmrState::dcnt).
enum {
  CNT_NREF0 = 0, CNT_NREF1, CNT_SKIN0, CNT_SKIN1, CNT_NDEEP, CNT_OVERFLOW, CNT_TGT, CNT_NSEL0,
  CNT_DIFF0, CNT_DIFF1,  // a regrid changed the level-0 / level-1 marks
  CNT_NF1, CNT_NF2,      // level-1 / level-2 tile counts the per-tile launches cover (at most the pools)
  CNT_HW1, CNT_HW2,      // their high-water marks since the initial build
  CNT_N
};
static_assert(CNT_HW1 - CNT_NF1 == 2 && CNT_HW2 - CNT_NF2 == 2, "a scan's high-water slot follows its count slot by 2");
// Items of a regrid launch: n from the host, or *dn when dn is given (n is then only the launch width's
// hint, and the kernels loop over what exceeds it).
struct Items {
  int n;
  const int* dn;
};
// thr = ascending-sorted[count - tgt] of sen (entries with mask != 0 when mask is given), tgt =
// min(cap, max(1, int(frac * count))); writes count to cnt[count_slot] and thr to *thr.
hipError_t rg_select(const float* sen, const int* mask, Items n, double frac, int cap, float* thr, int* cnt,
                     int count_slot, hipStream_t st);
hipError_t rg_mark_thr(const float* sen, const float* thr, int* ref, int n, hipStream_t st);
hipError_t rg_mark_ref1(const float* sen, const float* thr, const int* deep, const int* cnt, int* ref1, Items nf1,
                        hipStream_t st);
hipError_t rg_set(int* ref, int n, int mode, int cgx, const int* deep, hipStream_t st);  // 0 all, 1 slab, 2 deep
hipError_t rg_deepmask(const int* f1pos, const int* l1grid, int l1gx, Items nf1, int* deep, hipStream_t st);
// Exclusive scan of 0/1 flags into rank; total into cnt[slot]. With tiles_slot >= 0 also the tiles that
// the total refines, 8 per flag, limited to a pool of cap_tiles: cnt[tiles_slot] = 8 min(total, cap_tiles / 8),
// and their high-water mark in cnt[tiles_slot + 2].
hipError_t rg_scan(const int* flags, Items n, int* rank, int* cnt, int slot, int tiles_slot, int cap_tiles,
                   hipStream_t st);
// ascending indices i with flags[i] != 0 into out; count into cnt[slot]
hipError_t rg_compact(const int* flags, Items n, int* out, int* cnt, int slot, hipStream_t st);
// Block tables from the marks and their ranks; parents whose tiles exceed the pool raise the overflow
// flag and, with clamp, are unmarked (the tables then describe the grid that fits).
hipError_t rg_scatter_l1(int* ref, const int* rank, int cnb0, int cgx, int l1gx, int cap1, bool clamp, int* fof0, int* rl0,
                         int* l1pcb, int* l1oct, int* l1grid, int* f1pos, int* cnt, hipStream_t st);
hipError_t rg_scatter_l2(int* ref1, const int* rank, Items nf1, int cap2, bool clamp, int* fof1, int* rl1, int* l2pl1,
                         int* l2oct, int* cnt, hipStream_t st);
hipError_t rg_dilate_l0(const int* ref0, int cnb0, int cgx, int* flags, hipStream_t st);
hipError_t rg_dilate_l1(const int* ref1, const int* f1pos, const int* l1grid, Items nf1, int l1gx, int* flags,
                        hipStream_t st);
hipError_t rg_nbr_l1(const int* f1pos, const int* l1grid, int l1gx, Items nf1, int* nb1, hipStream_t st);
hipError_t rg_nbr_l2(const int* l2pl1, const int* l2oct, const int* f1pos, const int* l1grid, const int* ref1,
                     const int* fof1, int l1gx, Items nf2, int* nb2, hipStream_t st);
hipError_t rg_gather_by_pos(const float* by_pos, const int* f1pos, Items nf1, float* out, hipStream_t st);
// cnt[slot] = 1 if a[i] != b[i] for some i < n, else 0
hipError_t rg_differs(const int* a, const int* b, Items n, int* cnt, int slot, hipStream_t st);
// up to 5 copies src[k] -> dst[k] of n[k] ints in one launch
hipError_t rg_copy_i32_multi(const int* const* src, int* const* dst, const long* n, int count, hipStream_t st);

}  // namespace jp
