// Internal launch interface between the C-ABI layer (api.cpp) and the kernel translation units.
// Populations of a dense periodic lattice (the uniform lattice and AMR level 0) sit at
// f[q*qs + z*zs + y*N + x], x fastest, in one of two layouts: direction planes (qs = N^3 + pad, zs = N^2:
// AMR level 0, and the uniform lattice on request) or z-layers (qs = N^2 + pad, zs = 19 qs: the 19
// directions of one z-layer adjacent, the uniform default). Scattered AMR tiles (levels 1 and 2) are stored
// tile-major: f[(tile*19 + q)*512 + cell], cell = (k*8 + j)*8 + i.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "lattice.hpp"

namespace jp {

struct Dense {
  uint32_t N;      // edge
  uint64_t qs;     // elements between directions
  uint64_t zs;     // elements between z-layers
  uint64_t elems;  // elements of the buffer
  FastDiv divN;
  // element of direction q at z-layer z, in-layer offset r = y*N + x
  __host__ __device__ uint64_t at(int q, uint32_t z, uint32_t r) const { return uint64_t(q) * qs + uint64_t(z) * zs + r; }
};
// Direction planes of N^3 + pad elements.
inline Dense make_dense(int N, uint64_t pad) {
  Dense d;
  d.N = uint32_t(N);
  d.qs = uint64_t(N) * N * N + pad;
  d.zs = uint64_t(N) * N;
  d.elems = d.qs * Q;
  d.divN = make_fastdiv(uint32_t(N));
  return d;
}
// z-layers: the 19 directions of a layer adjacent, each followed by pad elements (a multiple of 32).
inline Dense make_dense_layers(int N, uint64_t pad = 0) {
  Dense d;
  d.N = uint32_t(N);
  d.qs = uint64_t(N) * N + pad;
  d.zs = d.qs * Q;
  d.elems = d.zs * uint64_t(N);
  d.divN = make_fastdiv(uint32_t(N));
  return d;
}
// Work item of workgroup w in a launch of n one-item workgroups. The dispatcher deals workgroups to the 8
// XCDs round-robin (w -> XCD w % 8), and each XCD has its own L2: giving every XCD a contiguous run of
// items keeps neighbouring items (the octants of one parent tile, the columns of one lattice band, whose
// faces and halos overlap) in one L2 instead of spreading them over all eight. A bijection of [0, n).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int w, int n) {
  const int x = w % kXcds, i = w / kXcds, q = n / kXcds, r = n % kXcds;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

template <class R> struct alignas(sizeof(R) * 4) V4 {
  R a, b, c, d;
};
inline size_t prec_bytes(int prec) { return prec == 3 ? 8 : (prec == 0 ? 4 : 2); }

// ---- dense lattice (uniform.hip) ----
hipError_t dense_init(int prec, void* f, const Dense& g, double u0, double kf, const double* phase, float enc,
                      hipStream_t st);
hipError_t dense_step(int prec, const void* fa, void* fb, const Dense& g, double omega, float enc, float dec,
                      int variant, hipStream_t st);
// Column shape {TX, TY, TZ, lanes} of the two-step fused pass for this form and edge: shape `pick` of the
// list, or with pick < 0 the first that divides the edge; false where none applies (the 16-bit forms,
// FP64, an edge no shape divides, or a pick that does not divide it): the lattice runs single steps.
int step_pair_shape_count();
bool step_pair_shape(int prec, int N, int pick, int* out4);
// Waves per CU the single step (pick < 0, launch variant `variant`) or the fused pair (column form `pick`)
// runs at: workgroups a CU holds at once (registers, LDS, workgroup size) x waves per workgroup; 0 if none.
int dense_step_waves_per_cu(int prec, const Dense& g, int variant, int pick);
// Two steps fa -> (t + 1) -> fb in one launch, bit-identical to two dense_step calls; needs a shape.
hipError_t dense_step_pair(int prec, const void* fa, void* fb, const Dense& g, double omega, float enc, float dec,
                           int pick, hipStream_t st);
// macro field (rho, u, v, w): float4 for FP32/FP16/BF16 storage, double4 for FP64 storage
hipError_t dense_macro(int prec, const void* f, const Dense& g, float dec, void* field, hipStream_t st);
// Q-criterion, speed and (when vel != nullptr) velocity of an N^3 field of (rho, u, v, w) (field.hip)
hipError_t field_qspeed(bool dbl_field, bool diff_in_field, const void* field, int N, float* Q, float* spd, float* vel,
                        hipStream_t st);
// E, enstrophy, mass sums of a macro field; diff_in_field_type selects the uniform driver's central
// differences (taken in the field's type) over the AMR reduction's (taken in double)
hipError_t field_diag(bool dbl_field, bool diff_in_field_type, const void* field, int N, double* partial,
                      int nblocks, double* out3, hipStream_t st);
constexpr int kDiagBlocks = 1024;
// E, enstrophy, mass of the dense lattice in one fused pass from the populations (no macro field): one
// partial per work item (3 doubles each, dense_diag_items(N) of them) at `partial`, the three means at out3
hipError_t dense_diag(int prec, const void* f, const Dense& g, float dec, double* partial, double* out3,
                      hipStream_t st);
uint32_t dense_diag_items(int N);
// The library's fixed per-device scratch array (no allocation): its device address on `device`, which must
// be current, and how many partial slots it holds (the 3 results follow them).
hipError_t noalloc_scratch(int device, double** ptr, size_t* slots);
hipError_t copy_probe(const void* src, void* dst, size_t bytes, hipStream_t st);
// mode 0 copy src -> dst, 1 read src, 2 write dst; bytes a multiple of 4096
hipError_t bandwidth_probe(int mode, const void* src, void* dst, size_t bytes, hipStream_t st);
// tile-major <-> dense plane conversion for the test hooks (raw elements, any precision)
hipError_t dense_to_tiles(int prec, const void* f, const Dense& g, void* tiles, hipStream_t st);
hipError_t tiles_to_dense(int prec, const void* tiles, const Dense& g, void* f, hipStream_t st);
// max |f - w_q| over the lattice (FP32 state), used to pick the shifted-store scale
hipError_t fill_i32(int* p, int v, size_t n, hipStream_t st);

}  // namespace jp
