// Device body of the dense-lattice pull stream-collide, shared by the per-context kernel (uniform.hip)
// and the batched ensemble kernel (ensemble.hip).
#pragma once
#include "internal.hpp"

namespace jp {

// Source coordinate arrays for c in {-1, 0, +1} (index c + 1): source = destination - c, periodic.
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

// Fused pull stream-collide (BGK), src/tgv3d_uniform.cu:41-61, one lattice cell per lane.
// kNtLoad / kNtStore select streaming (non-temporal) cache policy for the population reads / writes.
template <class S, bool kNtLoad = true, bool kNtStore = true>
__device__ inline void dense_step_cell(const typename S::raw* __restrict__ fa, typename S::raw* __restrict__ fb,
                                       const Dense& g, typename S::real omega, float enc, float dec, uint32_t cell) {
  using R = typename S::real;
  const uint32_t row = fdiv(cell, g.divN), x = cell - row * g.N, z = fdiv(row, g.divN), y = row - z * g.N;
  const Nbr n = neighbours(x, y, z, g.N);
  R fin[Q];
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const uint32_t src = n.zs[CZ(q) + 1] + n.ys[CY(q) + 1] + n.xs[CX(q) + 1];
    const typename S::raw* p = fa + uint64_t(q) * g.plane + src;
    if constexpr (kNtLoad) fin[q] = S::template load<q>(__builtin_nontemporal_load(p), dec);
    else fin[q] = S::template load<q>(*p, dec);
  });
  const Mom<R> m = moments(fin);
  const R ux = m.ux / m.rho, uy = m.uy / m.rho, uz = m.uz / m.rho;
  const EqPre<R> e = eq_pre<R>(m.rho, ux, uy, uz);
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const R fo = fin[q] - omega * (fin[q] - feq<q>(e));
    typename S::raw* p = fb + uint64_t(q) * g.plane + cell;
    if constexpr (kNtStore) __builtin_nontemporal_store(S::template store<q>(fo, enc), p);
    else *p = S::template store<q>(fo, enc);
  });
}

}  // namespace jp
