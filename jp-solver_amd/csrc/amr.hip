// AMR kernels on gfx950: fine-level pull steps, coarse-fine interface (prep, ghost fill), restriction,
// seeding, migration, sensor and the composite macro field (src/tgv3d_amr.cu:69-440, 442-453, 515-536).
//
// One workgroup of 512 lanes (8 waves) per 8^3 tile for the per-tile kernels; lanes map to cells in
// x-fastest order so that each wave touches one contiguous 8x8 plane of a q slice. The fine step looks up
// the tile's 27 same-level neighbours (precomputed at regrid) once in LDS, then pulls each direction
// from the own tile, a neighbour tile, or the tile's compact ghost record.
#include "amr_bodies.hpp"
#include "dense_bodies.hpp"

namespace jp {

// ---- per-context kernels over the bodies above ----
template <class S, bool kLoop>
__global__ __launch_bounds__(512) JP_FULL_OCCUPANCY(S) void k_fine_step(const typename S::raw* __restrict__ fa, typename S::raw* __restrict__ fb,
                                                   const typename S::real* __restrict__ gh, const int* __restrict__ nb,
                                                   const int* __restrict__ dn, typename S::real om, float enc, float dec) {
  __shared__ int snb[kNbr];
  for_tiles<kLoop>(dn, [=](int b, int lane) {
    if (lane < kNbr) snb[lane] = nb[size_t(b) * kNbr + lane];
    __syncthreads();
    fine_step_cell<S>(fa, fb, gh + size_t(b) * kGhostSlots, snb, om, enc, dec, b, lane);
  });
}
// Last fine substep with the restriction into the parent level fused in (ptile: the parent tile of each
// fine tile, l2pl1 or l1pcb); one slab of sibling tiles per unit (fine_step_restrict).
template <class S, bool kLoop>
__global__ __launch_bounds__(512) JP_RESTRICT_OCCUPANCY(S, kLoop) void k_fine_step_r(const typename S::raw* __restrict__ fa, typename S::raw* __restrict__ fb,
                                                     const typename S::real* __restrict__ gh, const int* __restrict__ nb,
                                                     const int* __restrict__ dn, typename S::real om, float enc, float dec,
                                                     typename S::raw* __restrict__ cf, const int* __restrict__ ptile,
                                                     float nqi, float cenc) {
  __shared__ RestrictLds<typename S::real> L;
  for_tiles<kLoop>(dn, [=](int u, int lane) {
    fine_step_restrict<S>(fa, fb, gh, nb, ptile, u, lane, om, enc, dec, cf, nqi, cenc, L);
  });
}

// Ghost fills, one fine tile per workgroup (the parent box and prep values staged in LDS).
template <class S, bool kLoop>
__global__ __launch_bounds__(256) JP_FULL_OCCUPANCY(S) void k_fill_l1(const typename S::raw* __restrict__ c0, typename S::real* __restrict__ gh1,
                                                 const int* __restrict__ l1pcb, const int* __restrict__ l1oct,
                                                 const int* __restrict__ nb1, const uint32_t* __restrict__ slots,
                                                 const int* __restrict__ dn, AmrGeo G, float nqs, float dec, float enc1) {
  __shared__ FillLds<typename S::real> L;
  for_tiles<kLoop, false>(dn, [=](int b, int lane) { fill_l1_tile<S>(c0, gh1, l1pcb, l1oct, nb1, slots, G, nqs, dec, enc1, b, lane, L); });
}
// Level-2 ghost fill fused with the first level-2 substep of the pair that reads it: the workgroup fills
// its tile's ghost record (written to gh2 for the pair's second substep, and kept in LDS) and steps the tile
// with the ghost values read from LDS, one launch and one pass over the tile's neighbourhood less.
template <class S, bool kLoop>
__global__ __launch_bounds__(512) JP_FULL_OCCUPANCY(S) void k_fill_step_l2(const typename S::raw* __restrict__ f1, typename S::real* __restrict__ gh2,
                                                      const int* __restrict__ l2pl1, const int* __restrict__ l2oct,
                                                      const int* __restrict__ f1pos, const int* __restrict__ nb1,
                                                      const int* __restrict__ nb2, const uint32_t* __restrict__ slots,
                                                      const int* __restrict__ dn, AmrGeo G, float nqs, float dec1,
                                                      const typename S::raw* __restrict__ fa, typename S::raw* __restrict__ fb,
                                                      typename S::real om, float enc2, float dec2) {
  using R = typename S::real;
  __shared__ FillLds<R> L;
  __shared__ R ghost[kGhostSlots];
  for_tiles<kLoop, false>(dn, [=](int b, int lane) {
    fill_l2_tile<S, 512>(f1, gh2, l2pl1, l2oct, f1pos, nb1, nb2, slots, G, nqs, dec1, enc2, b, lane, L, ghost);
    __syncthreads();
    fine_step_cell<S, kLoop>(fa, fb, ghost, L.nb, om, enc2, dec2, b, lane);
  });
}

// The first launches of a coarse step when level 2 exists. The first level-2 pair touches only levels 1
// and 2 while the level-0 step and the level-1 fill touch only level 0 and the level-1 ghost record, so
// each launch below carries two independent parts in disjoint workgroup ranges (the first part first in
// dispatch order): no second stream, no cross-stream wait, and the parts share the GPU. Each part starts
// at a multiple of 8 workgroups so that workgroup indices keep their XCD.
//  - k_dense_fill_step_l2: the level-0 step (one coarse tile per workgroup, n0 tiles in nd workgroups) and
//    the fused level-2 fill and first substep of the pair;
//  - k_step_r_l2_fill_l1: the pair's second level-2 substep with the restriction into level 1 (w2
//    workgroups) and the level-1 ghost fill from the new level 0.
template <class S, bool kLoop>
__global__ __launch_bounds__(512) JP_FULL_OCCUPANCY(S) void k_dense_fill_step_l2(const typename S::raw* __restrict__ c0a,
                                                            typename S::raw* __restrict__ c0b, const int* __restrict__ nb0,
                                                            int n0, typename S::real om0, float enc0, float dec0, int nd,
                                                            const typename S::raw* __restrict__ f1,
                                                            typename S::real* __restrict__ gh2, const int* __restrict__ l2pl1,
                                                            const int* __restrict__ l2oct, const int* __restrict__ f1pos,
                                                            const int* __restrict__ nb1, const int* __restrict__ nb2,
                                                            const uint32_t* __restrict__ slots, int n2,
                                                            const int* __restrict__ dn2, AmrGeo G, float nqs, float dec1,
                                                            const typename S::raw* __restrict__ fa,
                                                            typename S::raw* __restrict__ fb, typename S::real om2,
                                                            float enc2, float dec2) {
  using R = typename S::real;
  __shared__ FillLds<R> L;
  __shared__ R ghost[kGhostSlots];
  if (int(blockIdx.x) < nd) {  // level 0: every neighbour tile exists (periodic), no ghost record is read
    const int b = xcd_tile(blockIdx.x, nd), lane = threadIdx.x;
    if (b >= n0) return;
    if (lane < kNbr) L.nb[lane] = nb0[size_t(b) * kNbr + lane];
    __syncthreads();
    fine_step_cell<S, kLoop>(c0a, c0b, ghost, L.nb, om0, enc0, dec0, b, lane);
    return;
  }
  for_tiles_in<kLoop, false>(n2, dn2, int(blockIdx.x) - nd, int(gridDim.x) - nd, [=](int b, int lane) {
    fill_l2_tile<S, 512>(f1, gh2, l2pl1, l2oct, f1pos, nb1, nb2, slots, G, nqs, dec1, enc2, b, lane, L, ghost);
    __syncthreads();
    fine_step_cell<S, kLoop>(fa, fb, ghost, L.nb, om2, enc2, dec2, b, lane);
  });
}
template <class S, bool kLoop>
__global__ __launch_bounds__(512) JP_RESTRICT_OCCUPANCY(S, kLoop) void k_step_r_l2_fill_l1(const typename S::raw* __restrict__ fa, typename S::raw* __restrict__ fb,
                                                           const typename S::real* __restrict__ gh2,
                                                           const int* __restrict__ nb2, int n2, const int* __restrict__ dn2,
                                                           int w2, typename S::real om2, float enc2, float dec2,
                                                           typename S::raw* __restrict__ f1, const int* __restrict__ l2pl1,
                                                           AmrGeo G, float nqi21, float enc1,
                                                           const typename S::raw* __restrict__ c0,
                                                           typename S::real* __restrict__ gh1, const int* __restrict__ l1pcb,
                                                           const int* __restrict__ l1oct, const int* __restrict__ nb1,
                                                           const uint32_t* __restrict__ slots, int n1,
                                                           const int* __restrict__ dn1, float nqs01, float dec0) {
  using R = typename S::real;
  // the two parts run in different workgroups: their LDS overlaps
  __shared__ union {
    RestrictLds<R> rl;
    FillLds<R> fl[2];
  } U;
  if (int(blockIdx.x) < w2) {
    for_tiles_in<kLoop>(n2, dn2, blockIdx.x, w2, [=](int u, int lane) {
      fine_step_restrict<S>(fa, fb, gh2, nb2, l2pl1, u, lane, om2, enc2, dec2, f1, nqi21, enc1, U.rl);
    });
    return;
  }
  // the level-1 fill takes two tiles per workgroup (x-siblings 2p, 2p + 1, whose parent boxes overlap), a
  // half of the lanes each: the fill is latency-bound, so twice the tiles in flight per workgroup
  for_tiles_in<kLoop, false, 2>(n1, dn1, int(blockIdx.x) - w2, int(gridDim.x) - w2, [=](int p, int lane) {
    const int h = lane >> 8;
    fill_l1_tile<S, 256>(c0, gh1, l1pcb, l1oct, nb1, slots, G, nqs01, dec0, enc1, 2 * p + h, lane & 255, U.fl[h]);
  });
}

// prolong_L0 (src/tgv3d_amr.cu:70-91): trilinear over 8 periodic coarse cells, each corner's
// non-equilibrium regularised and rescaled by nqs.
template <class S> __device__ inline void prolong_l0(const typename S::raw* __restrict__ c0, const AmrGeo& G, int gfi,
                                                     int gfj, int gfk, float nqs, float dec, typename S::real* out) {
  using R = typename S::real;
  int ib, jb, kb;
  float tx, ty, tz;
  tri(gfi, ib, tx);
  tri(gfj, jb, ty);
  tri(gfk, kb, tz);
  sfor<Q>([&](auto qc) { out[decltype(qc)::value] = R(0); });
  for (int c = 0; c < 8; ++c) {
    const int di = c & 1, dj = (c >> 1) & 1, dk = c >> 2;
    const float wgt = (di ? tx : 1 - tx) * (dj ? ty : 1 - ty) * (dk ? tz : 1 - tz);
    R fq[Q], nq[Q];
    load_l0<S>(c0, G, pmod(ib + di, G.N0), pmod(jb + dj, G.N0), pmod(kb + dk, G.N0), dec, fq);
    const Mom<R> m = moments(fq);
    const EqPre<R> e = eq_pre<R>(m.rho, m.ux / m.rho, m.uy / m.rho, m.uz / m.rho);
    sfor<Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      nq[q] = fq[q] - feq<q>(e);
    });
    regularize(nq);
    sfor<Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      out[q] += R(wgt) * (feq<q>(e) + R(nqs) * nq[q]);
    });
  }
}
// prolong_L1 (src/tgv3d_amr.cu:94-118): parents on the scattered level 1, absent corners dropped and
// the weights renormalised when 0 < wsum < 0.999.
template <class S> __device__ inline void prolong_l1(const typename S::raw* __restrict__ f1, const int* __restrict__ l1grid,
                                                     const AmrGeo& G, int gfi, int gfj, int gfk, float nqs, float dec,
                                                     typename S::real* out) {
  using R = typename S::real;
  int ib, jb, kb;
  float tx, ty, tz, wsum = 0.f;
  tri(gfi, ib, tx);
  tri(gfj, jb, ty);
  tri(gfk, kb, tz);
  sfor<Q>([&](auto qc) { out[decltype(qc)::value] = R(0); });
  for (int c = 0; c < 8; ++c) {
    const int di = c & 1, dj = (c >> 1) & 1, dk = c >> 2;
    const float wgt = (di ? tx : 1 - tx) * (dj ? ty : 1 - ty) * (dk ? tz : 1 - tz);
    const int ci = pmod(ib + di, G.N1), cj = pmod(jb + dj, G.N1), ck = pmod(kb + dk, G.N1);
    const int lb = l1grid[((ck >> 3) * G.l1gx + (cj >> 3)) * G.l1gx + (ci >> 3)];
    if (lb < 0) continue;
    const int cc = ((ck & 7) * MB + (cj & 7)) * MB + (ci & 7);
    R fq[Q], nq[Q];
    load_tile<S>(f1, lb, cc, dec, fq);
    const Mom<R> m = moments(fq);
    const EqPre<R> e = eq_pre<R>(m.rho, m.ux / m.rho, m.uy / m.rho, m.uz / m.rho);
    sfor<Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      nq[q] = fq[q] - feq<q>(e);
    });
    regularize(nq);
    wsum += wgt;
    sfor<Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      out[q] += R(wgt) * (feq<q>(e) + R(nqs) * nq[q]);
    });
  }
  if (wsum > 0.f && wsum < 0.999f) {
    const R inv = R(1) / R(wsum);
    sfor<Q>([&](auto qc) { out[decltype(qc)::value] *= inv; });
  }
}

template <class S> __device__ inline void store_tile(typename S::raw* __restrict__ f, int b, int cell,
                                                     const typename S::real* v, float enc) {
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    f[tidx(b, q, cell)] = S::template store<q>(v[q], enc);
  });
}

// Seeding (src/tgv3d_amr.cu:299-308, 382-393) and migration (397-430).
template <class S, bool kLoop>
__global__ __launch_bounds__(512) void k_seed_l1(const typename S::raw* __restrict__ c0, typename S::raw* __restrict__ f1,
                                                 const int* __restrict__ l1pcb, const int* __restrict__ l1oct,
                                                 const int* __restrict__ o_ref0, const int* __restrict__ o_fof0,
                                                 const typename S::raw* __restrict__ oldf1, const int* __restrict__ dn,
                                                 AmrGeo G, float nqs, float dec0, float enc1) {
  for_tiles<kLoop, false>(dn, [=](int fb, int cell) {
    const int cb = l1pcb[fb], oct = l1oct[fb];
    if (o_ref0 && o_ref0[cb]) {  // migration: the parent was refined before, copy its old tile
      const int ofb = o_fof0[cb] + oct;
      sfor<Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f1[tidx(fb, q, cell)] = oldf1[tidx(ofb, q, cell)];
      });
      return;
    }
    const int cbx = cb % G.cgx, cby = (cb / G.cgx) % G.cgx, cbz = cb / (G.cgx * G.cgx);
    typename S::real out[Q];
    prolong_l0<S>(c0, G, cbx * 2 * MB + (oct & 1) * MB + (cell & 7), cby * 2 * MB + ((oct >> 1) & 1) * MB + ((cell >> 3) & 7),
                  cbz * 2 * MB + (oct >> 2) * MB + (cell >> 6), nqs, dec0, out);
    store_tile<S>(f1, fb, cell, out, enc1);
  });
}
template <class S, bool kLoop>
__global__ __launch_bounds__(512) void k_seed_l2(const typename S::raw* __restrict__ f1, typename S::raw* __restrict__ f2,
                                                 const int* __restrict__ l2pl1, const int* __restrict__ l2oct,
                                                 const int* __restrict__ f1pos, const int* __restrict__ l1grid,
                                                 const int* __restrict__ o_l1grid, const int* __restrict__ o_ref1,
                                                 const int* __restrict__ o_fof1, const typename S::raw* __restrict__ oldf2,
                                                 const int* __restrict__ dn, AmrGeo G, float nqs, float dec1, float enc2) {
  for_tiles<kLoop, false>(dn, [=](int b, int cell) {
    const int pl1 = l2pl1[b], oct = l2oct[b], p = f1pos[pl1];
    if (o_l1grid) {  // migration keyed by the level-1 grid position
      const int ol = o_l1grid[p];
      if (ol >= 0 && o_ref1[ol]) {
        const int of2 = o_fof1[ol] + oct;
        sfor<Q>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          f2[tidx(b, q, cell)] = oldf2[tidx(of2, q, cell)];
        });
        return;
      }
    }
    const int px = p % G.l1gx, py = (p / G.l1gx) % G.l1gx, pz = p / (G.l1gx * G.l1gx);
    typename S::real out[Q];
    prolong_l1<S>(f1, l1grid, G, px * 2 * MB + (oct & 1) * MB + (cell & 7),
                  py * 2 * MB + ((oct >> 1) & 1) * MB + ((cell >> 3) & 7), pz * 2 * MB + (oct >> 2) * MB + (cell >> 6), nqs,
                  dec1, out);
    store_tile<S>(f2, b, cell, out, enc2);
  });
}

// Sensor (src/tgv3d_amr.cu:442-453): per tile, max over cells of sqrt(sum_q (f - feq)^2) / rho.
__device__ inline float wave_max(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_down(v, off, 64));
  return v;
}
template <class S, bool kLoop>
__global__ __launch_bounds__(512) void k_sensor(const typename S::raw* __restrict__ f, const int* __restrict__ dn, AmrGeo G,
                                                float dec, float* __restrict__ sen) {
  using R = typename S::real;
  __shared__ float sm[8];
  for_tiles<kLoop>(dn, [=](int b, int cell) {
    R v[Q];
    load_tile<S>(f, b, cell, dec, v);
    const Mom<R> m = moments(v);
    const EqPre<R> e = eq_pre<R>(m.rho, m.ux / m.rho, m.uy / m.rho, m.uz / m.rho);
    R s = R(0);
    sfor<Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      const R d2 = v[q] - feq<q>(e);
      s += d2 * d2;
    });
    float x = float(sqrt(s) / m.rho);
    x = wave_max(x);
    if ((cell & 63) == 0) sm[cell >> 6] = x;
    __syncthreads();
    if (cell == 0) {
      float r = sm[0];
      for (int w = 1; w < 8; ++w) r = fmaxf(r, sm[w]);
      sen[b] = r;
    }
  });
}

// Composite macro field at 4 N0 (src/tgv3d_amr.cu:515-536): level 2 where present, else level 1 at g/2,
// else level 0 at g/4.
template <class S>
__global__ __launch_bounds__(256) void k_composite(const typename S::raw* __restrict__ c0, const typename S::raw* __restrict__ f1,
                                                   const typename S::raw* __restrict__ f2, const int* __restrict__ ref1,
                                                   const int* __restrict__ fof1, const int* __restrict__ l1grid, AmrGeo G,
                                                   float dec0, float dec1, float dec2, V4<typename S::real>* __restrict__ F) {
  using R = typename S::real;
  const uint32_t idx = blockIdx.x * 256 + threadIdx.x;
  const uint32_t N2 = G.N2;
  if (idx >= N2 * N2 * N2) return;
  const int gi = idx % N2, gj = (idx / N2) % N2, gk = idx / (N2 * N2);
  const int lb = l1grid[((gk / (2 * MB)) * G.l1gx + (gj / (2 * MB))) * G.l1gx + (gi / (2 * MB))];
  R v[Q];
  if (lb >= 0 && ref1[lb]) {
    const int f2b = fof1[lb] + ((gk % (2 * MB)) / MB) * 4 + ((gj % (2 * MB)) / MB) * 2 + (gi % (2 * MB)) / MB;
    load_tile<S>(f2, f2b, ((gk % MB) * MB + (gj % MB)) * MB + (gi % MB), dec2, v);
  } else if (lb >= 0) {
    const int g1i = gi / 2, g1j = gj / 2, g1k = gk / 2;
    load_tile<S>(f1, lb, ((g1k % MB) * MB + (g1j % MB)) * MB + (g1i % MB), dec1, v);
  } else {
    load_l0<S>(c0, G, gi / 4, gj / 4, gk / 4, dec0, v);
  }
  const Mom<R> m = moments(v);
  V4<R> o;
  o.a = m.rho;
  o.b = m.ux / m.rho;
  o.c = m.uy / m.rho;
  o.d = m.uz / m.rho;
  F[idx] = o;
}

// ---------------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------------
#define JP_DISPATCH(prec, CALL)            \
  switch (prec) {                          \
    case 0: { using S = FormF32; CALL; } break;  \
    case 1: { using S = FormF16; CALL; } break;  \
    case 2: { using S = FormBF16; CALL; } break; \
    case 3: { using S = FormF64; CALL; } break;  \
    default: return hipErrorInvalidValue;  \
  }

// One-tile-per-workgroup instantiation for launches with a host count, looping one for device counts.
#define JP_TILES(t, CALL)                            \
  if ((t).count) {                                   \
    constexpr bool kLoop = true;                     \
    CALL;                                            \
  } else {                                           \
    constexpr bool kLoop = false;                    \
    CALL;                                            \
  }

#define RAW(p) static_cast<const typename S::raw*>(p)
#define WRAW(p) static_cast<typename S::raw*>(p)
#define REAL(p) static_cast<const typename S::real*>(p)
#define WREAL(p) static_cast<typename S::real*>(p)

hipError_t amr_fine_step(int prec, int, const void* fa, void* fb, const void* gh, const int* nb, Tiles t, double omega,
                         float enc, float dec, hipStream_t st) {
  if (t.width <= 0) return hipSuccess;
  JP_DISPATCH(prec, JP_TILES(t, hipLaunchKernelGGL((k_fine_step<S, kLoop>), dim3(t.width), dim3(512), 0, st, RAW(fa), WRAW(fb), REAL(gh), nb,
                                       t.count, typename S::real(omega), enc, dec)));
  return hipGetLastError();
}

hipError_t amr_fine_step_restrict(int prec, const void* fa, void* fb, const void* gh, const int* nb, Tiles t, double omega,
                                  float enc, float dec, void* cf, const int* ptile, float nqi, float cenc, hipStream_t st) {
  if (t.width <= 0) return hipSuccess;
  JP_DISPATCH(prec, JP_TILES(t, hipLaunchKernelGGL((k_fine_step_r<S, kLoop>), dim3(t.width), dim3(512), 0, st, RAW(fa), WRAW(fb),
                                                   REAL(gh), nb, t.count, typename S::real(omega), enc, dec, WRAW(cf), ptile,
                                                   nqi, cenc)));
  return hipGetLastError();
}

hipError_t amr_fill_l1(int prec, const AmrGeo& G, const void* c0, void* gh1, const int* l1pcb, const int* l1oct,
                       const int* nb1, const uint32_t* slots, Tiles t, float nqs, float dec, float enc1, hipStream_t st) {
  if (t.width <= 0) return hipSuccess;
  JP_DISPATCH(prec, JP_TILES(t, hipLaunchKernelGGL((k_fill_l1<S, kLoop>), dim3(t.width), dim3(256), 0, st, RAW(c0), WREAL(gh1), l1pcb, l1oct,
                                       nb1, slots, t.count, G, nqs, dec, enc1)));
  return hipGetLastError();
}

hipError_t amr_fill_step_l2(int prec, const AmrGeo& G, const void* f1, void* gh2, const int* l2pl1, const int* l2oct,
                            const int* f1pos, const int* nb1, const int* nb2, const uint32_t* slots, Tiles t, float nqs,
                            float dec1, const void* fa, void* fb, double omega, float enc2, float dec2, hipStream_t st) {
  if (t.width <= 0) return hipSuccess;
  JP_DISPATCH(prec, JP_TILES(t, hipLaunchKernelGGL((k_fill_step_l2<S, kLoop>), dim3(t.width), dim3(512), 0, st, RAW(f1),
                                                   WREAL(gh2), l2pl1, l2oct, f1pos, nb1, nb2, slots, t.count, G, nqs, dec1,
                                                   RAW(fa), WRAW(fb), typename S::real(omega), enc2, dec2)));
  return hipGetLastError();
}

// Part width: the launch width of a Tiles rounded up to a multiple of 8 (one-tile form: the count).
static inline int part8(Tiles t) { return t.width > 0 ? (t.width + 7) / 8 * 8 : 0; }

hipError_t amr_dense_fill_step_l2(int prec, const AmrGeo& G, const void* c0a, void* c0b, const int* nb0, double om0,
                                  float enc0, float dec0, const void* f1, void* gh2, const int* l2pl1, const int* l2oct,
                                  const int* f1pos, const int* nb1, const int* nb2, const uint32_t* slots, Tiles t2,
                                  float nqs12, float dec1, const void* f2a, void* f2b, double om2, float enc2, float dec2,
                                  hipStream_t st) {
  const int n0 = G.cgx * G.cgx * G.cgx, nd = (n0 + 7) / 8 * 8, w2 = part8(t2);
  JP_DISPATCH(prec, JP_TILES(t2, hipLaunchKernelGGL((k_dense_fill_step_l2<S, kLoop>), dim3(nd + w2), dim3(512), 0, st,
                                                    RAW(c0a), WRAW(c0b), nb0, n0, typename S::real(om0), enc0, dec0, nd,
                                                    RAW(f1), WREAL(gh2), l2pl1, l2oct, f1pos, nb1, nb2, slots, t2.width,
                                                    t2.count, G, nqs12, dec1, RAW(f2a), WRAW(f2b),
                                                    typename S::real(om2), enc2, dec2)));
  return hipGetLastError();
}
hipError_t amr_step_r_l2_fill_l1(int prec, const AmrGeo& G, const void* f2a, void* f2b, const void* gh2, const int* nb2,
                                 Tiles t2, double om2, float enc2, float dec2, void* f1, const int* l2pl1,
                                 float nqi21, float enc1, const void* c0, void* gh1, const int* l1pcb, const int* l1oct,
                                 const int* nb1, const uint32_t* slots, Tiles t1, float nqs01, float dec0, hipStream_t st) {
  // the level-1 fill's workgroups take tile pairs
  const int w2 = part8(t2), w1 = part8({t1.count, (t1.width + 1) / 2});
  if (w1 + w2 == 0) return hipSuccess;
  JP_DISPATCH(prec, JP_TILES(t2, hipLaunchKernelGGL((k_step_r_l2_fill_l1<S, kLoop>), dim3(w2 + w1), dim3(512), 0, st,
                                                    RAW(f2a), WRAW(f2b), REAL(gh2), nb2, t2.width, t2.count, w2,
                                                    typename S::real(om2), enc2, dec2, WRAW(f1), l2pl1, G, nqi21,
                                                    enc1, RAW(c0), WREAL(gh1), l1pcb, l1oct, nb1, slots, t1.width,
                                                    t1.count, nqs01, dec0)));
  return hipGetLastError();
}

hipError_t amr_seed_l1(int prec, const AmrGeo& G, const void* c0, void* f1, const int* l1pcb, const int* l1oct, Tiles t,
                       float nqs, float dec0, float enc1, hipStream_t st) {
  if (t.width <= 0) return hipSuccess;
  JP_DISPATCH(prec, JP_TILES(t, hipLaunchKernelGGL((k_seed_l1<S, kLoop>), dim3(t.width), dim3(512), 0, st, RAW(c0), WRAW(f1), l1pcb, l1oct,
                                       nullptr, nullptr, nullptr, t.count, G, nqs, dec0, enc1)));
  return hipGetLastError();
}
hipError_t amr_migrate_l1(int prec, const AmrGeo& G, const void* oldf1, void* newf1, const void* c0, const int* l1pcb,
                          const int* l1oct, const int* o_ref0, const int* o_fof0, Tiles t, float nqs, float dec0, float enc1,
                          hipStream_t st) {
  if (t.width <= 0) return hipSuccess;
  JP_DISPATCH(prec, JP_TILES(t, hipLaunchKernelGGL((k_seed_l1<S, kLoop>), dim3(t.width), dim3(512), 0, st, RAW(c0), WRAW(newf1), l1pcb, l1oct,
                                       o_ref0, o_fof0, RAW(oldf1), t.count, G, nqs, dec0, enc1)));
  return hipGetLastError();
}
hipError_t amr_seed_l2(int prec, const AmrGeo& G, const void* f1, void* f2, const int* l2pl1, const int* l2oct,
                       const int* f1pos, const int* l1grid, Tiles t, float nqs, float dec1, float enc2, hipStream_t st) {
  if (t.width <= 0) return hipSuccess;
  JP_DISPATCH(prec, JP_TILES(t, hipLaunchKernelGGL((k_seed_l2<S, kLoop>), dim3(t.width), dim3(512), 0, st, RAW(f1), WRAW(f2), l2pl1, l2oct,
                                       f1pos, l1grid, nullptr, nullptr, nullptr, nullptr, t.count, G, nqs, dec1, enc2)));
  return hipGetLastError();
}
hipError_t amr_migrate_l2(int prec, const AmrGeo& G, const void* oldf2, void* newf2, const void* newf1, const int* l2pl1,
                          const int* l2oct, const int* f1pos, const int* o_l1grid, const int* o_ref1, const int* o_fof1,
                          const int* l1grid, Tiles t, float nqs, float dec1, float enc2, hipStream_t st) {
  if (t.width <= 0) return hipSuccess;
  JP_DISPATCH(prec, JP_TILES(t, hipLaunchKernelGGL((k_seed_l2<S, kLoop>), dim3(t.width), dim3(512), 0, st, RAW(newf1), WRAW(newf2), l2pl1,
                                       l2oct, f1pos, l1grid, o_l1grid, o_ref1, o_fof1, RAW(oldf2), t.count, G, nqs, dec1,
                                       enc2)));
  return hipGetLastError();
}

hipError_t amr_composite_field(int prec, const AmrGeo& G, const void* c0, const void* f1, const void* f2, const int* ref1,
                               const int* fof1, const int* l1grid, const float* dec3, void* F, hipStream_t st) {
  const uint32_t n = uint32_t(G.N2) * G.N2 * G.N2;
  JP_DISPATCH(prec, hipLaunchKernelGGL((k_composite<S>), dim3((n + 255) / 256), dim3(256), 0, st, RAW(c0), RAW(f1), RAW(f2),
                                       ref1, fof1, l1grid, G, dec3[0], dec3[1], dec3[2],
                                       static_cast<V4<typename S::real>*>(F)));
  return hipGetLastError();
}

hipError_t amr_sensor(int prec, const AmrGeo& G, const void* f, Tiles t, float dec, float* sen, hipStream_t st) {
  if (t.width <= 0) return hipSuccess;
  JP_DISPATCH(prec, JP_TILES(t, hipLaunchKernelGGL((k_sensor<S, kLoop>), dim3(t.width), dim3(512), 0, st, RAW(f), t.count, G, dec, sen)));
  return hipGetLastError();
}

// TGV initial state of level 0, one coarse tile per workgroup (the dense lattice's k_dense_init per cell).
template <class S>
__global__ __launch_bounds__(512) void k_l0_init(typename S::raw* __restrict__ f, AmrGeo G, double u0, double kf, double px,
                                                 double py, double pz, float enc) {
  const int b = blockIdx.x, cell = threadIdx.x;
  const int bx = b % G.cgx, by = (b / G.cgx) % G.cgx, bz = b / (G.cgx * G.cgx);
  const EqPre<typename S::real> e = tgv_equilibrium<typename S::real>(
      uint32_t(bx * MB + (cell & 7)), uint32_t(by * MB + ((cell >> 3) & 7)), uint32_t(bz * MB + (cell >> 6)), u0, kf, px, py, pz);
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    f[tidx(b, q, cell)] = S::template store<q>(feq<q>(e), enc);
  });
}
hipError_t amr_l0_init(int prec, const AmrGeo& G, void* f, double u0, double kf, const double* phase, float enc, hipStream_t st) {
  const int n0 = G.cgx * G.cgx * G.cgx;
  JP_DISPATCH(prec, hipLaunchKernelGGL((k_l0_init<S>), dim3(n0), dim3(512), 0, st, WRAW(f), G, u0, kf, phase[0], phase[1], phase[2],
                                       enc));
  return hipGetLastError();
}

// The 27 neighbour tiles of every level-0 tile, periodic ((dz+1)*9 + (dy+1)*3 + dx+1 as for the fine levels).
__global__ void k_l0_neighbours(int* __restrict__ nb0, int cgx) {
  const int n0 = cgx * cgx * cgx;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n0 * kNbr) return;
  const int b = i / kNbr, d = i % kNbr;
  const int bx = b % cgx, by = (b / cgx) % cgx, bz = b / (cgx * cgx);
  const int x = (bx + d % 3 - 1 + cgx) % cgx, y = (by + (d / 3) % 3 - 1 + cgx) % cgx, z = (bz + d / 9 - 1 + cgx) % cgx;
  nb0[i] = (z * cgx + y) * cgx + x;
}
hipError_t amr_l0_neighbours(int* nb0, int cgx, hipStream_t st) {
  const int n = cgx * cgx * cgx * kNbr;
  hipLaunchKernelGGL(k_l0_neighbours, dim3((n + 255) / 256), dim3(256), 0, st, nb0, cgx);
  return hipGetLastError();
}

}  // namespace jp
