// Batched ensemble kernels: one persistent launch per substep advances the same substep of every
// member lattice (uniform lattices, AMR level 0, the fine levels of every AMR member). Work is a list of
// (member, unit) items; each workgroup walks the list with a grid stride, so the launch shape (groups
// per CU) is independent of the ensemble and is picked by the autotuner. The per-item work is the same
// device body the single-context kernels run, so a member advanced in the batch is bit-identical to the
// same member advanced alone.
#include "amr_bodies.hpp"
#include "dense_bodies.hpp"
#include "ensemble.hpp"

namespace jp {

template <class S> __device__ inline const typename S::raw* cr(const void* p) {
  return static_cast<const typename S::raw*>(p);
}
template <class S> __device__ inline typename S::raw* wr(void* p) { return static_cast<typename S::raw*>(p); }

// items: (member, 256-cell chunk). An AMR member's level 0 is tile-major: chunk k is half k % 2 of tile
// k / 2, stepped as the fine levels are (every neighbour tile exists, no ghost record is read).
template <class S>
__global__ __launch_bounds__(256) void kb_dense_step(const MemberDesc* __restrict__ md, const int2* __restrict__ items,
                                                     int n, int p0) {
  using R = typename S::real;
  __shared__ int snb[kNbr];
  for (int it = blockIdx.x; it < n; it += gridDim.x) {
    const int2 w = items[it];
    const MemberDesc& d = md[w.x];
    if (d.nb0) {
      const int b = w.y >> 1, lane = opaque_lane();
      __syncthreads();  // the previous item's lanes are done with snb
      if (lane < kNbr) snb[lane] = d.nb0[size_t(b) * kNbr + lane];
      __syncthreads();
      fine_step_cell<S>(cr<S>(d.c0[p0]), wr<S>(d.c0[p0 ^ 1]), nullptr, snb, R(d.om[0]), d.enc[0], d.dec[0], b,
                        (w.y & 1) * 256 + lane);
    } else {
      dense_step_cell<S, false, true>(cr<S>(d.c0[p0]), wr<S>(d.c0[p0 ^ 1]), d.g0, R(d.om[0]), d.enc[0], d.dec[0],
                                      uint32_t(w.y) * 256 + threadIdx.x);
    }
  }
}
// items: (member, 2048-cell chunk): 8 consecutive x-cells per lane, one vector load per direction (the
// single-context FP32 step's body)
template <class S>
__global__ __launch_bounds__(256) void kb_dense_step8(const MemberDesc* __restrict__ md, const int2* __restrict__ items,
                                                      int n, int p0) {
  using R = typename S::real;
  for (int it = blockIdx.x; it < n; it += gridDim.x) {
    const int2 w = items[it];
    const MemberDesc& d = md[w.x];
    dense_step_vec<S, 8, false, false>(cr<S>(d.c0[p0]), wr<S>(d.c0[p0 ^ 1]), d.g0, R(d.om[0]), d.enc[0], d.dec[0],
                                       uint32_t(w.y) * 256 + threadIdx.x);
  }
}
// items: (member, fine tile); p: parity of the parent level's current buffer
template <class S, int kLevel>
__global__ __launch_bounds__(256) JP_FULL_OCCUPANCY(S) void kb_fill(const MemberDesc* __restrict__ md, const int2* __restrict__ items, int n,
                                               int p) {
  using R = typename S::real;
  __shared__ FillLds<R> L;
  for (int it = blockIdx.x; it < n; it += gridDim.x) {
    const int2 w = items[it];
    const MemberDesc& d = md[w.x];
    const int lane = opaque_lane();
    if constexpr (kLevel == 1)
      fill_l1_tile<S>(cr<S>(d.c0[p]), static_cast<R*>(d.gh1), d.l1pcb, d.l1oct, d.nb1, d.slots, d.G, d.nqs01, d.dec[0], d.enc[1], w.y,
                      lane, L);
    else
      fill_l2_tile<S>(cr<S>(d.f1[p]), static_cast<R*>(d.gh2), d.l2pl1, d.l2oct, d.f1pos, d.nb1, d.nb2, d.slots, d.G,
                      d.nqs12, d.dec[1], d.enc[2], w.y, lane, L);
  }
}
// items: (member, fine tile)
template <class S, int kLevel>
__global__ __launch_bounds__(512) JP_FULL_OCCUPANCY(S) void kb_fine_step(const MemberDesc* __restrict__ md, const int2* __restrict__ items,
                                                    int n, int p) {
  using R = typename S::real;
  __shared__ int snb[kNbr];
  for (int it = blockIdx.x; it < n; it += gridDim.x) {
    const int2 w = items[it];
    const MemberDesc& d = md[w.x];
    const int* nb = kLevel == 1 ? d.nb1 : d.nb2;
    const int lane = opaque_lane();
    __syncthreads();  // the previous item's lanes are done with snb
    if (lane < kNbr) snb[lane] = nb[size_t(w.y) * kNbr + lane];
    __syncthreads();
    if constexpr (kLevel == 1)
      fine_step_cell<S>(cr<S>(d.f1[p]), wr<S>(d.f1[p ^ 1]), static_cast<const R*>(d.gh1) + size_t(w.y) * kGhostSlots, snb, R(d.om[1]), d.enc[1],
                        d.dec[1], w.y, lane);
    else
      fine_step_cell<S>(cr<S>(d.f2[p]), wr<S>(d.f2[p ^ 1]), static_cast<const R*>(d.gh2) + size_t(w.y) * kGhostSlots, snb, R(d.om[2]), d.enc[2],
                        d.dec[2], w.y, lane);
  }
}
// Second substep of a fine level with the restriction into its parent fused in (the single-context
// k_fine_step_r): items (member, slab unit of fine_step_restrict); pf: the fine level's parity, pc: the
// parent level's.
template <class S, int kLevel>
__global__ __launch_bounds__(512) JP_FULL_OCCUPANCY(S) void kb_fine_step_r(const MemberDesc* __restrict__ md, const int2* __restrict__ items,
                                                      int n, int pf, int pc) {
  using R = typename S::real;
  __shared__ RestrictLds<R> L;
  for (int it = blockIdx.x; it < n; it += gridDim.x) {
    const int2 w = items[it];
    const MemberDesc& d = md[w.x];
    const int lane = opaque_lane();
    __syncthreads();  // the previous item's lanes are done with the neighbour entries and the partial sums
    if constexpr (kLevel == 1)
      fine_step_restrict<S>(cr<S>(d.f1[pf]), wr<S>(d.f1[pf ^ 1]), static_cast<const R*>(d.gh1), d.nb1, d.l1pcb, w.y, lane,
                            R(d.om[1]), d.enc[1], d.dec[1], wr<S>(d.c0[pc]), d.nqi10, d.enc[0], L);
    else
      fine_step_restrict<S>(cr<S>(d.f2[pf]), wr<S>(d.f2[pf ^ 1]), static_cast<const R*>(d.gh2), d.nb2, d.l2pl1, w.y, lane,
                            R(d.om[2]), d.enc[2], d.dec[2], wr<S>(d.f1[pc]), d.nqi21, d.enc[1], L);
  }
}

#define JP_DISPATCH(prec, CALL)                   \
  switch (prec) {                                 \
    case 0: { using S = FormF32; CALL; } break;   \
    case 1: { using S = FormF16; CALL; } break;   \
    case 2: { using S = FormBF16; CALL; } break;  \
    case 3: { using S = FormF64; CALL; } break;   \
    default: return hipErrorInvalidValue;         \
  }

static inline dim3 grid_of(int n, int groups) { return dim3(unsigned(std::max(1, std::min(n, groups)))); }

hipError_t ens_launch(int prec, EnsKernel k, const MemberDesc* md, const int2* items, int n, int groups, int pa, int pb,
                      hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const dim3 g = grid_of(n, groups);
  switch (k) {
    case EK_DENSE: JP_DISPATCH(prec, hipLaunchKernelGGL((kb_dense_step<S>), g, dim3(256), 0, st, md, items, n, pa)); break;
    case EK_FILL1: JP_DISPATCH(prec, hipLaunchKernelGGL((kb_fill<S, 1>), g, dim3(256), 0, st, md, items, n, pa)); break;
    case EK_FILL2: JP_DISPATCH(prec, hipLaunchKernelGGL((kb_fill<S, 2>), g, dim3(256), 0, st, md, items, n, pa)); break;
    case EK_STEP1: JP_DISPATCH(prec, hipLaunchKernelGGL((kb_fine_step<S, 1>), g, dim3(512), 0, st, md, items, n, pa)); break;
    case EK_STEP2: JP_DISPATCH(prec, hipLaunchKernelGGL((kb_fine_step<S, 2>), g, dim3(512), 0, st, md, items, n, pa)); break;
    case EK_STEP1R:
      JP_DISPATCH(prec, hipLaunchKernelGGL((kb_fine_step_r<S, 1>), g, dim3(512), 0, st, md, items, n, pa, pb));
      break;
    case EK_STEP2R:
      JP_DISPATCH(prec, hipLaunchKernelGGL((kb_fine_step_r<S, 2>), g, dim3(512), 0, st, md, items, n, pa, pb));
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t ens_launch_dense8(int prec, const MemberDesc* md, const int2* items, int n, int groups, int pa,
                             hipStream_t st) {
  if (n <= 0) return hipSuccess;
  JP_DISPATCH(prec, hipLaunchKernelGGL((kb_dense_step8<S>), grid_of(n, groups), dim3(256), 0, st, md, items, n, pa));
  return hipGetLastError();
}

}  // namespace jp
