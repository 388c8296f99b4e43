// Device body of the dense-lattice pull stream-collide, shared by the per-context kernel (uniform.hip)
// and the batched ensemble kernel (ensemble.hip).
#pragma once
#include "internal.hpp"

namespace jp {

// Equilibrium factors of the TGV initial state at cell (x, y, z) (src/tgv3d_uniform.cu:65-77;
// src/tgv3d_amr.cu:431-440): FP64 trigonometry, the working type's equilibrium.
template <class R> __device__ inline EqPre<R> tgv_equilibrium(uint32_t x, uint32_t y, uint32_t z, double u0, double kf, double px,
                                                              double py, double pz) {
  const double X = (double(x) + 0.5) * kf + px, Y = (double(y) + 0.5) * kf + py, Z = (double(z) + 0.5) * kf + pz;
  const double ux = u0 * sin(X) * cos(Y) * cos(Z);
  const double uy = -u0 * cos(X) * sin(Y) * cos(Z);
  const double p = (u0 * u0 / 16.0) * (cos(2 * X) + cos(2 * Y)) * (cos(2 * Z) + 2.0);
  const double rho = 1.0 + 3.0 * p;
  return eq_pre<R>(R(rho), R(ux), R(uy), R(0));
}

// Source coordinate arrays for c in {-1, 0, +1} (index c + 1): source = destination - c, periodic; zs are
// offsets in an x-fastest N^3 field (the macro field), not in the population layout.
struct Nbr {
  uint32_t xs[3], ys[3], zs[3];
};
__device__ inline Nbr neighbours(uint32_t x, uint32_t y, uint32_t z, uint32_t N) {
  Nbr n;
  const uint32_t NN = N * N;
  n.xs[0] = x == N - 1 ? 0 : x + 1;
  n.xs[1] = x;
  n.xs[2] = x == 0 ? N - 1 : x - 1;
  n.ys[0] = (y == N - 1 ? 0 : y + 1) * N;
  n.ys[1] = y * N;
  n.ys[2] = (y == 0 ? N - 1 : y - 1) * N;
  n.zs[0] = (z == N - 1 ? 0 : z + 1) * NN;
  n.zs[1] = z * NN;
  n.zs[2] = (z == 0 ? N - 1 : z - 1) * NN;
  return n;
}

// Fused pull stream-collide over V consecutive x-cells per lane: each direction is one aligned vector
// load of the source row segment (a wave reads 64*V contiguous elements of a plane per instruction);
// the x-shifted directions take the element that crosses the lane boundary from the neighbouring lane
// by a cross-lane shuffle, and only lanes at a wave or row edge load it themselves (periodic wrap).
// Each cell's arithmetic is exactly that of dense_step_cell.
template <class T, int V> struct VecT {
  typedef T type __attribute__((ext_vector_type(V)));
};
template <class S, int V, bool kNtLoad, bool kNtStore>
__device__ inline void dense_step_vec(const typename S::raw* __restrict__ fa, typename S::raw* __restrict__ fb,
                                      const Dense& g, typename S::real omega, float enc, float dec, uint32_t lane_id) {
  using R = typename S::real;
  using raw = typename S::raw;
  using vec = typename VecT<raw, V>::type;
  const uint32_t N = g.N, NN = N * N;
  const uint32_t cell0 = lane_id * V;
  const uint32_t row = fdiv(cell0, g.divN), x0 = cell0 - row * N, z = fdiv(row, g.divN), y = row - z * N;
  const uint32_t ys[3] = {(y == N - 1 ? 0 : y + 1) * N, y * N, (y == 0 ? N - 1 : y - 1) * N};
  const uint64_t zs[3] = {uint64_t(z == N - 1 ? 0 : z + 1) * g.zs, uint64_t(z) * g.zs, uint64_t(z == 0 ? N - 1 : z - 1) * g.zs};
  const uint32_t xl = x0 == 0 ? N - 1 : x0 - 1, xr = x0 + V == N ? 0 : x0 + V;
  const int lane = threadIdx.x & 63;
  const bool left_edge = lane == 0 || x0 == 0, right_edge = lane == 63 || x0 + V == N;
  R f[Q][V];
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int cx = CX(q);
    const raw* base = fa + uint64_t(q) * g.qs + zs[CZ(q) + 1] + ys[CY(q) + 1];
    const vec* vp = reinterpret_cast<const vec*>(base + x0);
    vec w;
    if constexpr (kNtLoad) w = __builtin_nontemporal_load(vp);
    else w = *vp;
    R v[V];
#pragma unroll
    for (int j = 0; j < V; ++j) v[j] = S::template in<q>(w[j], dec);
    if constexpr (cx == 0) {
#pragma unroll
      for (int j = 0; j < V; ++j) f[q][j] = v[j];
    } else if constexpr (cx > 0) {  // source x - 1
      R l = __shfl_up(v[V - 1], 1, 64);
      if (left_edge) l = S::template in<q>(base[xl], dec);
      f[q][0] = l;
#pragma unroll
      for (int j = 1; j < V; ++j) f[q][j] = v[j - 1];
    } else {  // source x + 1
      R r = __shfl_down(v[0], 1, 64);
      if (right_edge) r = S::template in<q>(base[xr], dec);
#pragma unroll
      for (int j = 0; j < V - 1; ++j) f[q][j] = v[j + 1];
      f[q][V - 1] = r;
    }
  });
#pragma unroll
  for (int j = 0; j < V; ++j) {
    R fin[Q];
    sfor<Q>([&](auto qc) { fin[decltype(qc)::value] = f[decltype(qc)::value][j]; });
    const Collider<S> col(fin, dec, enc);
    sfor<Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      f[q][j] = col.template post<q>(fin[q], omega);
    });
  }
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    vec w;
#pragma unroll
    for (int j = 0; j < V; ++j) w[j] = S::template out<q>(f[q][j], enc);
    vec* p = reinterpret_cast<vec*>(fb + uint64_t(q) * g.qs + zs[1] + (cell0 - z * NN));
    if constexpr (kNtStore) __builtin_nontemporal_store(w, p);
    else *p = w;
  });
}

// Fused pull stream-collide (BGK), src/tgv3d_uniform.cu:41-61, one lattice cell per lane.
// kNtLoad / kNtStore select streaming (non-temporal) cache policy for the population reads / writes.
template <class S, bool kNtLoad = true, bool kNtStore = true>
__device__ inline void dense_step_cell(const typename S::raw* __restrict__ fa, typename S::raw* __restrict__ fb,
                                       const Dense& g, typename S::real omega, float enc, float dec, uint32_t cell) {
  using R = typename S::real;
  const uint32_t N = g.N, row = fdiv(cell, g.divN), x = cell - row * N, z = fdiv(row, g.divN), y = row - z * N;
  const Nbr n = neighbours(x, y, z, N);
  const uint64_t zs[3] = {uint64_t(z == N - 1 ? 0 : z + 1) * g.zs, uint64_t(z) * g.zs, uint64_t(z == 0 ? N - 1 : z - 1) * g.zs};
  R fin[Q];
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const typename S::raw* p = fa + uint64_t(q) * g.qs + zs[CZ(q) + 1] + (n.ys[CY(q) + 1] + n.xs[CX(q) + 1]);
    if constexpr (kNtLoad) fin[q] = S::template in<q>(__builtin_nontemporal_load(p), dec);
    else fin[q] = S::template in<q>(*p, dec);
  });
  const Collider<S> col(fin, dec, enc);
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const typename S::raw v = S::template out<q>(col.template post<q>(fin[q], omega), enc);
    typename S::raw* p = fb + uint64_t(q) * g.qs + zs[1] + (y * N + x);
    if constexpr (kNtStore) __builtin_nontemporal_store(v, p);
    else *p = v;
  });
}

}  // namespace jp
