// Device bodies of the AMR kernels, shared by the per-context launches (amr.hip) and the batched
// ensemble launches (ensemble.hip). Each body computes one cell or one tile of one level.
#pragma once
#include "amr_internal.hpp"

namespace jp {

// Per-tile AMR kernels of the 4- and 2-byte forms ask for 8 waves per SIMD: their vector registers allow
// it (64 or fewer), and without the request the scalar registers of the looping and batched forms
// (100-106) held them at 7. FP64 asks for 4, which its 84-106 vector registers already give (at 8 it
// would spill).
#define JP_FULL_OCCUPANCY(S) __attribute__((amdgpu_waves_per_eu(sizeof(typename S::raw) == 8 ? 4 : 8)))
// The looping 16-bit forms of the restricting kernels ask for 7: within 64 registers the compiler spilled
// (k_fine_step_r) or loaded the gathered populations one at a time (k_step_r_l2_fill_l1).
#define JP_RESTRICT_OCCUPANCY(S, kLoop) \
  __attribute__((amdgpu_waves_per_eu(kLoop && sizeof(typename S::raw) == 2 ? 7 : (sizeof(typename S::raw) == 8 ? 4 : 8))))

__device__ __forceinline__ int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}
// Trilinear base and weight of the fine coordinate g on the parent grid: x = (g + 1/2)/2 - 1/2,
// base = floor(x), t = x - base (src/tgv3d_amr.cu:71-73). Exact in integers: g = 2m gives (m - 1, 3/4),
// g = 2m + 1 gives (m, 1/4).
__device__ __forceinline__ void tri(int g, int& base, float& t) {
  const int m = g >> 1;
  if (g & 1) {
    base = m;
    t = 0.25f;
  } else {
    base = m - 1;
    t = 0.75f;
  }
}
__device__ __forceinline__ size_t tidx(int b, int q, int cell) { return (size_t(b) * Q + q) * MB3 + cell; }
// Level 0 is stored tile-major like the fine levels: coarse cell (ci, cj, ck) is cell ((ck & 7) * 8 + (cj &
// 7)) * 8 + (ci & 7) of tile ((ck >> 3) * cgx + (cj >> 3)) * cgx + (ci >> 3).
__device__ __forceinline__ int l0_tile(const AmrGeo& G, int ci, int cj, int ck) {
  return ((ck >> 3) * G.cgx + (cj >> 3)) * G.cgx + (ci >> 3);
}
__device__ __forceinline__ int tile_cell(int i, int j, int k) { return ((k & 7) * MB + (j & 7)) * MB + (i & 7); }


// The lane index as a value the compiler cannot see through, taken once per iteration of a loop over
// tiles: address arithmetic derived from it is then computed inside the iteration instead of being
// hoisted out of the loop and held in registers across it.
__device__ __forceinline__ int opaque_lane() {
  int lane = threadIdx.x;
  asm volatile("" : "+v"(lane));
  return lane;
}

// Tile loop of the per-tile kernels. Without kLoop the launch has one workgroup per tile and workgroup w
// takes tile xcd_tile(w, width). With kLoop the tile count n is read on the device (*dn), so that a launch
// recorded in a graph stays valid when a regrid changes it: workgroup w takes the indices v = w, w +
// width, ... below n, each mapped to tile xcd_tile(v, n). The launch width is then a multiple of 8, so all
// indices of a workgroup fall in its XCD's share, and the shares are balanced over the live count (a
// mapping over the launch width would leave whole XCDs idle when the count is below it). With kSync, a
// barrier precedes every index but the first, so that a body may rewrite the LDS the previous one used.
// body(tile, lane) takes the lane index from the loop, in the looping form from opaque_lane(). The loop
// still costs registers, which is why the one-tile form is a separate instantiation.
// for_tiles_in: the same over a part of a launch, workgroups wg in [0, width) of it (width a multiple of 8
// and the part starting at a multiple of 8, so that wg keeps its XCD); the one-tile form then skips tiles
// from n up (width is n rounded up to a multiple of 8).
// per > 1: the unit is a group of per consecutive tiles, n and *dn still count tiles (a multiple of per).
template <bool kLoop, bool kSync = true, int per = 1, class F>
__device__ __forceinline__ void for_tiles_in(int n, const int* __restrict__ dn, int wg, int width, F&& body) {
  if constexpr (!kLoop) {
    const int b = xcd_tile(wg, width);
    if (b < n / per) body(b, int(threadIdx.x));
  } else {
    const int nd = *dn / per;
    for (int v = wg; v < nd; v += width) {
      if constexpr (kSync)
        if (v != wg) __syncthreads();
      body(xcd_tile(v, nd), opaque_lane());
    }
  }
}
template <bool kLoop, bool kSync = true, class F> __device__ __forceinline__ void for_tiles(const int* __restrict__ dn, F&& body) {
  if constexpr (!kLoop) body(xcd_tile(blockIdx.x, gridDim.x), int(threadIdx.x));
  else for_tiles_in<kLoop, kSync>(0, dn, blockIdx.x, gridDim.x, body);
}

template <class S> __device__ __forceinline__ void load_tile(const typename S::raw* __restrict__ f, int b, int cell, float dec,
                                                    typename S::real* out) {
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    out[q] = S::template load<q>(f[tidx(b, q, cell)], dec);
  });
}
// coarse cell (ci, cj, ck) of level 0
template <class S> __device__ __forceinline__ void load_l0(const typename S::raw* __restrict__ f, const AmrGeo& G, int ci, int cj,
                                                  int ck, float dec, typename S::real* out) {
  load_tile<S>(f, l0_tile(G, ci, cj, ck), tile_cell(ci, cj, ck), dec, out);
}

// BGK collision of pulled values (the form's working representation), stored to tile b
// (src/tgv3d_amr.cu:268-273). kept (optional, 19 registers): receives each stored population as it reads
// back, for the fused restriction.
template <class S, bool kKeep = false>
__device__ __forceinline__ void collide_store(typename S::real (&fin)[Q], typename S::real om, typename S::raw* __restrict__ fb,
                                     int b, int cell, float enc, typename S::real (&kept)[Q], float dec) {
  const Collider<S> col(fin, dec, enc);
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const typename S::raw v = S::template out<q>(col.template post<q>(fin[q], om), enc);
    fb[tidx(b, q, cell)] = v;
    if constexpr (kKeep) kept[q] = S::template load<q>(v, dec);
  });
}

// Restriction from the just-stored populations of fine tiles, fused into the last fine substep
// (src/tgv3d_amr.cu:275-298, 358-381): parent = feq + nqi * regularize(mean of the 8 children - feq). The
// children are summed in the reference's order c = 0..7 (di, dj, dk). The reference sums 0.125 f_c; this
// sums f_c and scales the total by 0.125 once. Scaling by a power of two is exact and commutes with
// rounding while values stay in the normal range (populations and their sums are far above 2^-126), so
// each partial sum is exactly 8 times the reference's and the mean is the same bits.
// Lane i receives lane i + N of its 16-lane row (DPP row_shl:N, VALU moves with no LDS round trip);
// lanes whose source lies past the row's end keep v.
template <int N> __device__ __forceinline__ double row_down(double v) {
  const int2 h = *reinterpret_cast<const int2*>(&v);
  int2 r;
  r.x = __builtin_amdgcn_update_dpp(h.x, h.x, 0x100 + N, 0xF, 0xF, false);
  r.y = __builtin_amdgcn_update_dpp(h.y, h.y, 0x100 + N, 0xF, 0xF, false);
  return *reinterpret_cast<const double*>(&r);
}
// acc + (v of lane i + N in the 16-lane row); lanes whose source lies past the row's end add 0. The
// zero-filled cross-lane read folds into the add (one v_add_f32 with a DPP operand).
template <int N> __device__ __forceinline__ float add_row_down(float acc, float v) {
  return acc + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 + N, 0xF, 0xF, true));
}
template <int N> __device__ __forceinline__ double add_row_down(double acc, double v) { return acc + row_down<N>(v); }
// (((s + v) + v[+1]) + v[+8]) + v[+9]: four children of a parent added to its running sum s, in the
// reference's child order, at the parent's corner lane (even i, even j; its sources +1, +8, +9 always lie
// in its own 16-lane row, which holds j = 2m, 2m + 1).
// The empty asm pins the sum outside the caller's corner-lane branch: sunk into it, the adds would run
// under a partial exec mask and the cross-lane reads could no longer fold into them.
template <class R> __device__ __forceinline__ R add_children(R s, R v) {
  R r = add_row_down<9>(add_row_down<8>(add_row_down<1>(s + v, v), v), v);
  asm volatile("" : "+v"(r));
  return r;
}
// Parent cell pc of tile pcb from its children's mean: feq + nqi * regularize(avg - feq).
template <class S>
__device__ __forceinline__ void restrict_finish(typename S::real (&avg)[Q], typename S::raw* __restrict__ cf, int pcb, int pc,
                                       float nqi, float cenc) {
  using R = typename S::real;
  R nq[Q];
  const Mom<R> m = moments(avg);
  R ux = m.ux, uy = m.uy, uz = m.uz;
  quot3(m.rho, ux, uy, uz);
  const EqPre<R> e = eq_pre<R>(m.rho, ux, uy, uz);
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    nq[q] = avg[q] - feq<q>(e);
  });
  regularize(nq);
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    cf[tidx(pcb, q, pc)] = S::template store<q>(feq<q>(e) + R(nqi) * nq[q], cenc);
  });
}

// What a lane of a fine step needs to locate its pulled values: the neighbour entries (dz, dy, dx) =
// (z ? ze : 0, y ? ye : 0, x ? xe : 0) as n<z><y><x>, where xe is -1 for i = 0, +1 for i = 7 and 0
// otherwise (likewise ye, ze), and the in-tile source coordinates of the x and y shifts c = +1, 0, -1.
struct PullLane {
  int n000, n001, n010, n011, n100, n101, n110, n111;
  int xe, ye, ze;
  uint32_t sxp, sx0, sxm, syp, sy0, sym;
};
__device__ __forceinline__ PullLane pull_lane(const int* snb, int i, int j, int k) {
  PullLane P;
  P.xe = i == 0 ? -1 : (i == 7 ? 1 : 0);
  P.ye = j == 0 ? -1 : (j == 7 ? 1 : 0);
  P.ze = k == 0 ? -1 : (k == 7 ? 1 : 0);
  const int x = P.xe + 1, y = (P.ye + 1) * 3, z = (P.ze + 1) * 9;
  P.n000 = snb[9 + 3 + 1];
  P.n001 = snb[9 + 3 + x];
  P.n010 = snb[9 + y + 1];
  P.n011 = snb[9 + y + x];
  P.n100 = snb[z + 3 + 1];
  P.n101 = snb[z + 3 + x];
  P.n110 = snb[z + y + 1];
  P.n111 = snb[z + y + x];
  // the entries as register values: otherwise the compiler turns each direction's choice among them into a
  // choice of address and keeps the struct in scratch memory
  asm volatile("" : "+v"(P.n000), "+v"(P.n001), "+v"(P.n010), "+v"(P.n011));
  asm volatile("" : "+v"(P.n100), "+v"(P.n101), "+v"(P.n110), "+v"(P.n111));
  P.sxp = uint32_t((i - 1) & 7);
  P.sx0 = uint32_t(i);
  P.sxm = uint32_t((i + 1) & 7);
  P.syp = uint32_t(((j - 1) & 7) * MB);
  P.sy0 = uint32_t(j * MB);
  P.sym = uint32_t(((j + 1) & 7) * MB);
  return P;
}
// Source tile t of direction q for the lane (-1: the ghost record) and the element offset src of its
// source cell within tile t. The lane crosses into the x-neighbour exactly when c_x = -xe (likewise y, z).
template <int q>
__device__ __forceinline__ void pull_source(PullLane P, int k, int& t, uint32_t& src) {
  constexpr int cx = CX(q), cy = CY(q), cz = CZ(q);
  const bool crx = cx != 0 && P.xe == -cx, cry = cy != 0 && P.ye == -cy, crz = cz != 0 && P.ze == -cz;
  t = crx ? (cry ? P.n011 : P.n001) : (cry ? P.n010 : P.n000);
  if constexpr (cz != 0)
    if (crz) t = crx ? (cry ? P.n111 : P.n101) : (cry ? P.n110 : P.n100);
  const uint32_t sx = cx > 0 ? P.sxp : (cx < 0 ? P.sxm : P.sx0), sy = cy > 0 ? P.syp : (cy < 0 ? P.sym : P.sy0);
  src = uint32_t(q * MB3 + ((k - cz) & 7) * (MB * MB)) + sy + sx;
}

// Fine pull stream-collide, levels 1 and 2 (src/tgv3d_amr.cu:241-274, 331-356). nb holds the tile's
// 27 same-level neighbours ((dz+1)*9 + (dy+1)*3 + dx+1, -1 where the neighbour position is not refined
// to this level); missing sources come from the ghost record.
// snb: the tile's neighbour entries, staged in LDS by the caller; g: the tile's ghost record (in HBM, or
// in LDS where the caller has just computed it), whose values are in the form's working representation. A wave must hold one z-plane of the tile (64 lanes of
// one k), as every caller's lane = cell mapping gives.
// The source of direction q lies in a neighbour tile only along the axes where the lane is on the tile's
// edge facing -c_q: a lane at x-edge xe (-1 for i = 0, +1 for i = 7, else 0) crosses in x exactly when
// c_x = -xe. So a lane needs at most the 8 entries (dz, dy, dx) in {0, ze} x {0, ye} x {0, xe}: it reads
// them once, and each direction selects among the ones its non-zero components allow (no per-direction
// LDS lookup and index arithmetic). Element offsets are 32-bit: a fine level holds at most kMaxFineTiles
// tiles. kBitFlags (16-bit forms): the per-direction ghost flags are bits of one vector register instead of
// lane masks in scalar registers, for the looping fused fill-and-step kernels, whose scalar registers are
// short (there the 18 masks spilled to vector lanes); a mask costs one select per direction, a bit three
// instructions.
template <class S, bool kKeep = false, bool kBitFlags = false>
__device__ __forceinline__ void fine_step_cell(const typename S::raw* __restrict__ fa, typename S::raw* __restrict__ fb,
                                      const typename S::real* __restrict__ g, const int* snb, typename S::real om,
                                      float enc, float dec, int b, int cell, typename S::real (&kept)[Q]) {
  using R = typename S::real;
  using raw = typename S::raw;
  const int i = cell & 7, j = (cell >> 3) & 7;
  const int k = __builtin_amdgcn_readfirstlane(cell >> 6);  // uniform: a wave is one z-plane of the tile
  const PullLane P = pull_lane(snb, i, j, k);
  R fin[Q];
  if constexpr (sizeof(raw) == sizeof(R)) {
    // populations and ghost values share a type: the compiler selects the address, one load per direction
    sfor<Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      if constexpr (q == 0) {
        fin[0] = S::template in<0>(fa[tidx(b, 0, cell)], dec);
      } else {
        int t;
        uint32_t src;
        pull_source<q>(P, k, t, src);
        if (t >= 0) fin[q] = S::template in<q>(fa[__umul24(uint32_t(t), uint32_t(Q * MB3)) + src], dec);
        else fin[q] = g[ghost_slot(q, i, j, k)];
      }
    });
  } else {
    // 16-bit populations, FP32 ghost values: every lane loads its population (a ghost lane its own cell's,
    // discarded), only ghost lanes load the ghost record, and nothing is decoded before all loads are
    // issued. Decoding inside the branch made each direction wait for its load: 19 round trips in series.
    raw v[Q];
    R gv[Q];
    bool from_ghost[Q];
    uint32_t ghost_bits = 0;
    sfor<Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      gv[q] = R(0);
      from_ghost[q] = false;
      if constexpr (q == 0) {
        v[0] = fa[tidx(b, 0, cell)];
      } else {
        int t;
        uint32_t src;
        pull_source<q>(P, k, t, src);
        // a ghost lane reads the same in-tile index of its own tile (in bounds, discarded)
        v[q] = fa[__umul24(uint32_t(t >= 0 ? t : b), uint32_t(Q * MB3)) + src];
        if constexpr (!kBitFlags) from_ghost[q] = t < 0;
        if (t < 0) {
          gv[q] = g[ghost_slot(q, i, j, k)];
          if constexpr (kBitFlags) ghost_bits |= 1u << q;
        }
      }
    });
    sfor<Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      // the empty asm takes the ghost value as loaded: it keeps the compiler from moving the decode
      // into the branch above, where it would wait for the load
      R gq = gv[q];
      asm volatile("" : "+v"(gq));
      const bool ghost = kBitFlags ? ((ghost_bits >> q) & 1u) != 0 : from_ghost[q];
      fin[q] = ghost ? gq : S::template in<q>(v[q], dec);
    });
  }
  collide_store<S, kKeep>(fin, om, fb, b, cell, enc, kept, dec);
}

// The same without keeping the stored values.
template <class S, bool kBitFlags = false>
__device__ __forceinline__ void fine_step_cell(const typename S::raw* __restrict__ fa, typename S::raw* __restrict__ fb,
                                      const typename S::real* __restrict__ g, const int* snb, typename S::real om,
                                      float enc, float dec, int b, int cell) {
  typename S::real none[Q];
  fine_step_cell<S, false, kBitFlags>(fa, fb, g, snb, om, enc, dec, b, cell, none);
}

// Last fine substep with the restriction fused in, over one slab of sibling tiles. The 8 fine tiles of a
// parent tile are 8r + oct (oct = qz * 4 + qy * 2 + qx: the octant they refine), and unit u of a launch over
// n tiles (n a multiple of 8) takes the four siblings 8r + 4qz + s (s = qy * 2 + qx) and their z-planes 2pk
// and 2pk + 1, with r = u / 8, qz = (u / 4) % 2, pk = u % 4: wave w steps plane 2pk + w % 2 of sibling w / 2.
// These planes hold all the children of the 8 x 8 parent cells of layer qz * 4 + pk of the parent tile.
// The even wave of a sibling adds each parent's four children of plane 2pk (lanes +0, +1, +8, +9 of a
// corner lane) and leaves the partial sum in LDS, the odd wave adds the four of plane 2pk + 1, and one wave
// finishes the 64 parents, a lane each in x-fastest order, so each direction is stored as one 8 x 8 block of
// 256 contiguous bytes of the (tile-major) parent level. One tile's own octant, 4 x 4 x 4 parents, would
// store 16 rows of 16 B per direction.
// ptile: the parent tile of each fine tile (l1pcb or l2pl1). Call with all 512 lanes.
template <class R> struct RestrictLds {
  int nb[4 * kNbr];     // the four siblings' neighbour entries
  R part[4][Q][16];     // partial sums per sibling, direction, parent of the sibling's 4 x 4 layer
  R avg[Q][64];         // complete child sums per direction and parent (py * 8 + px)
};
template <class S, bool kBitFlags = false>
__device__ __forceinline__ void fine_step_restrict(const typename S::raw* __restrict__ fa, typename S::raw* __restrict__ fb,
                                          const typename S::real* __restrict__ gh, const int* __restrict__ nb,
                                          const int* __restrict__ ptile, int u, int lane, typename S::real om, float enc,
                                          float dec, typename S::raw* __restrict__ cf, float nqi, float cenc,
                                          RestrictLds<typename S::real>& L) {
  using R = typename S::real;
  const int t0 = (u >> 3) * 8 + ((u >> 2) & 1) * 4, pz = ((u >> 2) & 1) * 4 + (u & 3);
  const int w = lane >> 6, l = lane & 63, sib = w >> 1;  // w and sib are uniform in a wave
  if (lane < 4 * kNbr) L.nb[lane] = nb[size_t(t0) * kNbr + lane];
  __syncthreads();
  const int b = t0 + sib, cell = (2 * (u & 3) + (w & 1)) * 64 + l;
  R kept[Q];
  fine_step_cell<S, true, kBitFlags>(fa, fb, gh + size_t(b) * kGhostSlots, L.nb + sib * kNbr, om, enc, dec, b, cell, kept);
  const int i = l & 7, j = l >> 3;
  const bool corner = !(i & 1) && !(j & 1);
  const int pp = (j >> 1) * 4 + (i >> 1);
  if (!(w & 1)) {  // children c = 0..3 of plane 2pk
    sfor<Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      const R s = add_children(R(0), kept[q]);
      if (corner) L.part[sib][q][pp] = s;
    });
  }
  __syncthreads();
  if (w & 1) {  // children c = 4..7 of plane 2pk + 1, continuing the running sum (every lane reads its parent's)
    const int pl = ((sib >> 1) * 4 + (j >> 1)) * 8 + (sib & 1) * 4 + (i >> 1);
    sfor<Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      const R s = add_children(L.part[sib][q][pp], kept[q]);
      if (corner) L.avg[q][pl] = s;
    });
  }
  __syncthreads();
  if (w == 0) {  // one full wave finishes the 64 parents of layer pz, lane = py * 8 + px
    R avg[Q];
    sfor<Q>([&](auto qc) { avg[decltype(qc)::value] = R(0.125) * L.avg[decltype(qc)::value][l]; });
    restrict_finish<S>(avg, cf, ptile[t0], pz * 64 + l, nqi, cenc);
  }
}

// prep = feq + nqs * regularize(f - feq) of one cell, in place (src/tgv3d_amr.cu:160-171).
template <class R> __device__ __forceinline__ void prep_values(R* v, float nqs) {
  R nq[Q];
  const Mom<R> m = moments(v);
  R ux = m.ux, uy = m.uy, uz = m.uz;
  quot3(m.rho, ux, uy, uz);
  const EqPre<R> e = eq_pre<R>(m.rho, ux, uy, uz);
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    nq[q] = v[q] - feq<q>(e);
  });
  regularize(nq);
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    v[q] = feq<q>(e) + R(nqs) * nq[q];
  });
}

// Ghost fill of one fine tile (src/tgv3d_amr.cu:120-202). The trilinear corners of every ghost source of
// a tile at fine origin o lie in the 6^3 parent-level box starting at o/2 - 1 (fine coordinate g has
// parent base floor((g - 1)/2) for g in [o - 1, o + 8]). The workgroup stages that box into LDS with the
// prep values computed in place (feq + nqs * regularize(f - feq) of the parent populations, exactly
// what the reference's prep pass stores for these cells), then interpolates each slot from LDS. A slot
// whose source lies in a same-level neighbour tile (nb >= 0) is not a ghost and is skipped, and only the
// box cells that some remaining slot reads are staged (a tile with all 26 neighbours stages nothing). The
// record holds each interpolated population in the working representation of the fine level (enc: its
// scale), so the fine steps read ghost values as they read decoded populations.
constexpr int kBox = kFillBox, kBox3 = kBox * kBox * kBox;
constexpr int kFillThreads = 256;
template <class R> struct FillLds {
  R box[Q * kBox3];
  int nb[kNbr];
  int nbp[kNbr];  // level 2: the 27 level-1 neighbours of the parent tile
  int ok[kBox3];  // level 2: the parent cell exists (level-1 tiles are scattered)
};
// Whether box cell (li, lj, lk) is a trilinear corner of some ghost slot. Slots toward a neighbour
// offset d reach box coordinates {0, 1} for d = -1, {4, 5} for d = +1 and 0..5 for d = 0 (per axis); a
// slot is a ghost only when that neighbour tile is absent (nb < 0).
__device__ __forceinline__ bool box_cell_needed(const int* nb, int li, int lj, int lk) {
  const int xl = li < 2 ? -1 : 0, xh = li > 3 ? 1 : 0, yl = lj < 2 ? -1 : 0, yh = lj > 3 ? 1 : 0;
  const int zl = lk < 2 ? -1 : 0, zh = lk > 3 ? 1 : 0;
  for (int dz = zl; dz <= zh; ++dz)
    for (int dy = yl; dy <= yh; ++dy)
      for (int dx = xl; dx <= xh; ++dx)
        if (nb[(dz + 1) * 9 + (dy + 1) * 3 + dx + 1] < 0) return true;
  return false;
}
// trilinear weight of corner c (bits di, dj, dk) for a fill-table entry's source parities
__device__ __forceinline__ float corner_weight(uint32_t e, int c) {
  const float tx = (e >> 18) & 1 ? 0.25f : 0.75f, ty = (e >> 19) & 1 ? 0.25f : 0.75f, tz = (e >> 20) & 1 ? 0.25f : 0.75f;
  return ((c & 1) ? tx : 1 - tx) * ((c & 2) ? ty : 1 - ty) * ((c & 4) ? tz : 1 - tz);
}

// Level 1 from level 0 (prep0 over the coarse box; every box cell exists); kThreads lanes.
template <class S, int kThreads = kFillThreads>
__device__ __forceinline__ void fill_l1_tile(const typename S::raw* __restrict__ c0, typename S::real* __restrict__ gh1,
                                    const int* __restrict__ l1pcb, const int* __restrict__ l1oct,
                                    const int* __restrict__ nb1, const uint32_t* __restrict__ slots, const AmrGeo& G,
                                    float nqs, float dec, float enc, int fb, int tid, FillLds<typename S::real>& L) {
  using R = typename S::real;
  constexpr int kSlotsPerLane = (kGhostSlots + kThreads - 1) / kThreads;
  const int cb = l1pcb[fb], oct = l1oct[fb];
  const int cbx = cb % G.cgx, cby = (cb / G.cgx) % G.cgx, cbz = cb / (G.cgx * G.cgx);
  const int ox = cbx * 2 * MB + (oct & 1) * MB, oy = cby * 2 * MB + ((oct >> 1) & 1) * MB, oz = cbz * 2 * MB + (oct >> 2) * MB;
  const int bx0 = ox / 2 - 1, by0 = oy / 2 - 1, bz0 = oz / 2 - 1;
  uint32_t ent[kSlotsPerLane];  // this lane's slot entries, loaded ahead of the staging (0: none)
#pragma unroll
  for (int k = 0; k < kSlotsPerLane; ++k) {
    const int s = tid + k * kThreads;
    ent[k] = s < kGhostSlots ? slots[s] : 0u;
  }
  __syncthreads();  // a previous tile of this workgroup is done with the LDS
  if (tid < kNbr) L.nb[tid] = nb1[size_t(fb) * kNbr + tid];
  __syncthreads();
  for (int c = tid; c < kBox3; c += kThreads) {
    const int li = c % kBox, lj = (c / kBox) % kBox, lk = c / (kBox * kBox);
    if (!box_cell_needed(L.nb, li, lj, lk)) continue;
    R v[Q];
    load_l0<S>(c0, G, pmod(bx0 + li, G.N0), pmod(by0 + lj, G.N0), pmod(bz0 + lk, G.N0), dec, v);
    prep_values(v, nqs);
    sfor<Q>([&](auto qc) { L.box[decltype(qc)::value * kBox3 + c] = v[decltype(qc)::value]; });
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kSlotsPerLane; ++k) {
    if (!ent[k]) continue;
    const uint32_t e = ent[k];
    if (L.nb[(e >> 5) & 31] >= 0) continue;
    const R* bq = L.box + (e & 31) * kBox3 + ((e >> 10) & 255);
    R out = R(0);
#pragma unroll
    for (int c = 0; c < 8; ++c) out += R(corner_weight(e, c)) * bq[((c >> 2) * kBox + ((c >> 1) & 1)) * kBox + (c & 1)];
    gh1[size_t(fb) * kGhostSlots + tid + k * kThreads] = ghost_value<S>(out, e & 31, enc);
  }
}

// Level 2 from the scattered level 1: box cells without a level-1 tile are dropped and the remaining
// weights renormalised (src/tgv3d_amr.cu:133-144, 187-202). kThreads lanes; lg (optional): an LDS copy
// of the tile's ghost record for a step fused behind the fill. The box lies within the parent level-1
// tile and its 26 neighbours, so a box cell's level-1 tile is read from the parent's neighbour entries
// (nb1, staged with the tile's own) rather than looked up in the level-1 grid per cell, which put one
// more dependent load before the box loads.
template <class S, int kThreads = kFillThreads>
__device__ __forceinline__ void fill_l2_tile(const typename S::raw* __restrict__ f1, typename S::real* __restrict__ gh2,
                                    const int* __restrict__ l2pl1, const int* __restrict__ l2oct,
                                    const int* __restrict__ f1pos, const int* __restrict__ nb1,
                                    const int* __restrict__ nb2, const uint32_t* __restrict__ slots, const AmrGeo& G,
                                    float nqs, float dec, float enc, int b, int tid, FillLds<typename S::real>& L,
                                    typename S::real* lg = nullptr) {
  using R = typename S::real;
  constexpr int kSlotsPerLane = (kGhostSlots + kThreads - 1) / kThreads;
  const int pl1 = l2pl1[b], oct = l2oct[b], p = f1pos[pl1];
  const int px = p % G.l1gx, py = (p / G.l1gx) % G.l1gx, pz = p / (G.l1gx * G.l1gx);
  const int ox = px * 2 * MB + (oct & 1) * MB, oy = py * 2 * MB + ((oct >> 1) & 1) * MB, oz = pz * 2 * MB + (oct >> 2) * MB;
  const int bx0 = ox / 2 - 1, by0 = oy / 2 - 1, bz0 = oz / 2 - 1;
  uint32_t ent[kSlotsPerLane];  // this lane's slot entries, loaded ahead of the staging (0: none)
#pragma unroll
  for (int k = 0; k < kSlotsPerLane; ++k) {
    const int s = tid + k * kThreads;
    ent[k] = s < kGhostSlots ? slots[s] : 0u;
  }
  __syncthreads();  // a previous tile of this workgroup is done with the LDS
  if (tid < kNbr) L.nb[tid] = nb2[size_t(b) * kNbr + tid];
  else if (tid < 2 * kNbr) L.nbp[tid - kNbr] = nb1[size_t(pl1) * kNbr + tid - kNbr];
  __syncthreads();
  for (int c = tid; c < kBox3; c += kThreads) {
    const int li = c % kBox, lj = (c / kBox) % kBox, lk = c / (kBox * kBox);
    L.ok[c] = 0;
    if (!box_cell_needed(L.nb, li, lj, lk)) continue;
    // level-1 coordinates, one cell outside the grid at most (-1 or N1: the periodic neighbour tile)
    const int ci = bx0 + li, cj = by0 + lj, ck = bz0 + lk;
    const int lb = L.nbp[((ck >> 3) - pz + 1) * 9 + ((cj >> 3) - py + 1) * 3 + (ci >> 3) - px + 1];
    if (lb < 0) continue;
    L.ok[c] = 1;
    R v[Q];
    load_tile<S>(f1, lb, ((ck & 7) * MB + (cj & 7)) * MB + (ci & 7), dec, v);
    prep_values(v, nqs);
    sfor<Q>([&](auto qc) { L.box[decltype(qc)::value * kBox3 + c] = v[decltype(qc)::value]; });
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kSlotsPerLane; ++k) {
    if (!ent[k]) continue;
    const uint32_t e = ent[k];
    if (L.nb[(e >> 5) & 31] >= 0) continue;
    const int base = (e >> 10) & 255;
    const R* bq = L.box + (e & 31) * kBox3 + base;
    float wsum = 0.f;
    R out = R(0);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int o = ((c >> 2) * kBox + ((c >> 1) & 1)) * kBox + (c & 1);
      if (!L.ok[base + o]) continue;
      const float wgt = corner_weight(e, c);
      wsum += wgt;
      out += R(wgt) * bq[o];
    }
    if (wsum > 0.f && wsum < 0.999f) out *= R(1) / R(wsum);
    const R w = ghost_value<S>(out, e & 31, enc);
    gh2[size_t(b) * kGhostSlots + tid + k * kThreads] = w;
    if (lg) lg[tid + k * kThreads] = w;
  }
}

}  // namespace jp
