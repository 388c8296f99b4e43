// Batched ensemble: per-member device descriptor and the batched launch interface (ensemble.hip).
#pragma once
#include "../../include/jp.h"
#include "amr_internal.hpp"

namespace jp {

// Everything a batched kernel needs about one member. Buffers that ping-pong are stored as pairs; the
// launch passes the parity that selects the current one (all members of an ensemble swap in step).
struct MemberDesc {
  void* c0[2];
  void* f1[2];
  void* f2[2];
  void *gh1, *gh2;
  const int *ref0, *fof0, *rl0, *l1pcb, *l1oct, *f1pos, *l1grid, *ref1, *fof1, *rl1, *l2pl1, *l2oct;
  const int *nb0, *nb1, *nb2;  // nb0: an AMR member's level-0 neighbour table (null for a uniform member)
  const uint32_t* slots;
  Dense g0;  // a uniform member's lattice
  AmrGeo G;
  double om[3];
  float enc[3], dec[3];
  float nqs01, nqi10, nqs12, nqi21;
};

// EK_FILL1 / EK_FILL2: ghost records (prep of the parent level computed in the fill). EK_STEP2R /
// EK_STEP1R: the second substep of a fine level with the restriction into its parent fused in.
enum EnsKernel { EK_DENSE = 0, EK_FILL1, EK_FILL2, EK_STEP2, EK_STEP2R, EK_STEP1, EK_STEP1R, EK_COUNT };
static_assert(EK_COUNT == JP_ENS_KERNELS, "one tuned launch shape per batched kernel");

// One batched launch of kernel k over n items with at most `groups` workgroups. pa: parity of the
// buffer read, pb: the parent level's parity (the fused restriction's target; EK_STEP2R / EK_STEP1R only).
hipError_t ens_launch(int prec, EnsKernel k, const MemberDesc* md, const int2* items, int n, int groups, int pa, int pb,
                      hipStream_t st);
// The dense step in its second tile shape: items (member, 2048-cell chunk), 8 consecutive x-cells per lane
// (every member's cell count must be a multiple of 2048).
constexpr int kDense8Chunk = 2048;
hipError_t ens_launch_dense8(int prec, const MemberDesc* md, const int2* items, int n, int groups, int pa, hipStream_t st);

}  // namespace jp
