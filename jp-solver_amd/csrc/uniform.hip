// Dense periodic D3Q19 lattice on gfx950: TGV initial state, fused pull stream-collide, macroscopic
// field, FP64 energy/enstrophy/mass reduction. Used for the uniform solver and for AMR level 0.
//
// Layout: 19 x-fastest planes, f[q*plane + (z*N + y)*N + x]. A wave of 64 lanes covers 64 consecutive
// cells of one x-row, so every population load is a contiguous 256 B (FP32) run, shifted by one element
// for directions with c_x != 0 (the partial line is shared with the neighbouring wave of the same block
// and served by the CU's L1). Each element of every plane is read exactly once per step, so the kernel
// moves the algorithmic 2*19 elements per cell update and nothing else.
#include <algorithm>

#include "dense_bodies.hpp"

namespace jp {

// TGV initial state, src/tgv3d_uniform.cu:65-77: FP64 trigonometry, FP32 (working type) equilibrium.
template <class S, int BS>
__global__ __launch_bounds__(BS) void k_dense_init(typename S::raw* __restrict__ f, Dense g, double u0, double kf,
                                                   double px, double py, double pz, float enc) {
  using R = typename S::real;
  const uint32_t cell = blockIdx.x * BS + threadIdx.x;
  const uint32_t row = fdiv(cell, g.divN), x = cell - row * g.N, z = fdiv(row, g.divN), y = row - z * g.N;
  const double X = (double(x) + 0.5) * kf + px, Y = (double(y) + 0.5) * kf + py, Z = (double(z) + 0.5) * kf + pz;
  const double ux = u0 * sin(X) * cos(Y) * cos(Z);
  const double uy = -u0 * cos(X) * sin(Y) * cos(Z);
  const double p = (u0 * u0 / 16.0) * (cos(2 * X) + cos(2 * Y)) * (cos(2 * Z) + 2.0);
  const double rho = 1.0 + 3.0 * p;
  const EqPre<R> e = eq_pre<R>(R(rho), R(ux), R(uy), R(0));
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    f[uint64_t(q) * g.plane + cell] = S::template store<q>(feq<q>(e), enc);
  });
}

// Fused pull stream-collide over the whole lattice, one cell per lane.
template <class S, int BS, bool kNtLoad, bool kNtStore>
__global__ __launch_bounds__(BS) void k_dense_step(const typename S::raw* __restrict__ fa,
                                                   typename S::raw* __restrict__ fb, Dense g,
                                                   typename S::real omega, float enc, float dec) {
  dense_step_cell<S, kNtLoad, kNtStore>(fa, fb, g, omega, enc, dec, blockIdx.x * BS + threadIdx.x);
}

// Macroscopic field, src/tgv3d_uniform.cu:80-90.
template <class S, int BS>
__global__ __launch_bounds__(BS) void k_dense_macro(const typename S::raw* __restrict__ f, Dense g, float dec,
                                                    V4<typename S::real>* __restrict__ F) {
  using R = typename S::real;
  const uint32_t cell = blockIdx.x * BS + threadIdx.x;
  R v[Q];
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    v[q] = S::template load<q>(f[uint64_t(q) * g.plane + cell], dec);
  });
  const Mom<R> m = moments(v);
  V4<R> o;
  o.a = m.rho;
  o.b = m.ux / m.rho;
  o.c = m.uy / m.rho;
  o.d = m.uz / m.rho;
  F[cell] = o;
}

// ---- deterministic FP64 block reduction ----
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
template <int BS> __device__ inline void block_sum3(double a, double b, double c, double* out) {
  __shared__ double sh[3][BS / 64];
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l == 0) {
    sh[0][w] = a;
    sh[1][w] = b;
    sh[2][w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int i = 0; i < BS / 64; ++i) {
      s0 += sh[0][i];
      s1 += sh[1][i];
      s2 += sh[2][i];
    }
    out[0] = s0;
    out[1] = s1;
    out[2] = s2;
  }
}

// Per-cell E, enstrophy, mass of a macro field: the uniform driver's host loop
// (src/tgv3d_uniform.cu:99-108, 174-183; differences in the field's type) or the AMR device reduction
// (src/tgv3d_amr.cu:538-549; differences in double). Fixed grid, fixed per-thread order: deterministic.
template <class R, bool kDiffInField, int BS>
__global__ __launch_bounds__(BS) void k_field_diag(const V4<R>* __restrict__ F, uint32_t N, FastDiv dN,
                                                   double* __restrict__ partial) {
  double E = 0, ens = 0, mass = 0;
  const uint32_t total = N * N * N;
  for (uint32_t cell = blockIdx.x * BS + threadIdx.x; cell < total; cell += gridDim.x * BS) {
    const uint32_t row = fdiv(cell, dN), x = cell - row * N, z = fdiv(row, dN), y = row - z * N;
    const Nbr n = neighbours(x, y, z, N);
    const V4<R> c = F[cell];
    const V4<R> xp = F[n.zs[1] + n.ys[1] + n.xs[0]], xm = F[n.zs[1] + n.ys[1] + n.xs[2]];
    const V4<R> yp = F[n.zs[1] + n.ys[0] + n.xs[1]], ym = F[n.zs[1] + n.ys[2] + n.xs[1]];
    const V4<R> zp = F[n.zs[0] + n.ys[1] + n.xs[1]], zm = F[n.zs[2] + n.ys[1] + n.xs[1]];
    auto dif = [](R p, R m) -> double {
      if constexpr (kDiffInField) return double(p - m) * 0.5;
      else return (double(p) - double(m)) * 0.5;
    };
    const double dudy = dif(yp.b, ym.b), dudz = dif(zp.b, zm.b);
    const double dvdx = dif(xp.c, xm.c), dvdz = dif(zp.c, zm.c);
    const double dwdx = dif(xp.d, xm.d), dwdy = dif(yp.d, ym.d);
    const double ox = dwdy - dvdz, oy = dudz - dwdx, oz = dvdx - dudy;
    const double u = c.b, v = c.c, w = c.d;
    E += 0.5 * (u * u + v * v + w * w);
    ens += 0.5 * (ox * ox + oy * oy + oz * oz);
    mass += double(c.a);
  }
  block_sum3<BS>(E, ens, mass, partial + 3 * blockIdx.x);
}
template <int BS>
__global__ __launch_bounds__(BS) void k_diag_final(const double* __restrict__ partial, int nb, double inv,
                                                   double* __restrict__ out) {
  double a = 0, b = 0, c = 0;
  for (int i = threadIdx.x; i < nb; i += BS) {
    a += partial[3 * i];
    b += partial[3 * i + 1];
    c += partial[3 * i + 2];
  }
  __shared__ double res[3];
  block_sum3<BS>(a, b, c, res);
  __syncthreads();
  if (threadIdx.x < 3) out[threadIdx.x] = res[threadIdx.x] * inv;
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ void k_copy16(const u32x4* __restrict__ s, u32x4* __restrict__ d, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    __builtin_nontemporal_store(__builtin_nontemporal_load(s + i), d + i);
}
// Plain bandwidth kernels, one 16 B element per lane and one pass over the buffer (the fastest simple
// forms on MI355X): mode 0 copies s to d, 1 reads s (the sum is stored only if it matches a value the
// zero-filled buffer cannot produce, so the loads stay live), 2 writes d.
template <int kMode> __global__ __launch_bounds__(256) void k_probe(const u32x4* __restrict__ s, u32x4* __restrict__ d) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if constexpr (kMode == 0) {
    d[i] = s[i];
  } else if constexpr (kMode == 1) {
    const u32x4 v = s[i];
    if ((v.x ^ v.y ^ v.z ^ v.w) == 0x9e3779b9u) d[0] = v;
  } else {
    const u32x4 v = {uint32_t(i), 1u, 2u, 3u};
    d[i] = v;
  }
}

template <class T, bool kToTiles>
__global__ void k_dense_tiles(const T* __restrict__ src, T* __restrict__ dst, Dense g) {
  // one block per (tile, q); 512 threads = the tile's cells
  const uint32_t b = blockIdx.x / Q, q = blockIdx.x % Q, cell = threadIdx.x;
  const uint32_t gx = g.N / MB, bx = b % gx, by = (b / gx) % gx, bz = b / (gx * gx);
  const uint32_t i = cell % MB, j = (cell / MB) % MB, k = cell / (MB * MB);
  const uint64_t d = uint64_t(q) * g.plane + ((uint64_t(bz * MB + k) * g.N + (by * MB + j)) * g.N + (bx * MB + i));
  const uint64_t t = (uint64_t(b) * Q + q) * MB3 + cell;
  if (kToTiles) dst[t] = src[d];
  else dst[d] = src[t];
}

__global__ void k_fill_i32(int* p, int v, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}

// ---- launchers ----
constexpr int kBS = 256;

template <class S> static hipError_t init_t(void* f, const Dense& g, double u0, double kf, const double* ph, float enc,
                                            hipStream_t st) {
  const uint64_t cells = uint64_t(g.N) * g.N * g.N;
  hipLaunchKernelGGL((k_dense_init<S, kBS>), dim3(uint32_t(cells / kBS)), dim3(kBS), 0, st,
                     static_cast<typename S::raw*>(f), g, u0, kf, ph[0], ph[1], ph[2], enc);
  return hipGetLastError();
}
hipError_t dense_init(int prec, void* f, const Dense& g, double u0, double kf, const double* phase, float enc,
                      hipStream_t st) {
  switch (prec) {
    case 0: return init_t<FormF32>(f, g, u0, kf, phase, enc, st);
    case 1: return init_t<FormF16>(f, g, u0, kf, phase, enc, st);
    case 2: return init_t<FormBF16>(f, g, u0, kf, phase, enc, st);
    case 3: return init_t<FormF64>(f, g, u0, kf, phase, enc, st);
  }
  return hipErrorInvalidValue;
}

// V consecutive x-cells per lane (V = 1: dense_step_cell)
template <class S, int BS, int V, bool kNtLoad, bool kNtStore>
__global__ __launch_bounds__(BS) void k_dense_step_vec(const typename S::raw* __restrict__ fa,
                                                       typename S::raw* __restrict__ fb, Dense g,
                                                       typename S::real omega, float enc, float dec) {
  dense_step_vec<S, V, kNtLoad, kNtStore>(fa, fb, g, omega, enc, dec, blockIdx.x * BS + threadIdx.x);
}

template <class S, int BS, int V, bool NL, bool NS>
static hipError_t step_v(const void* fa, void* fb, const Dense& g, double omega, float enc, float dec, hipStream_t st) {
  using R = typename S::real;
  const uint64_t cells = uint64_t(g.N) * g.N * g.N;
  if constexpr (V == 1)
    hipLaunchKernelGGL((k_dense_step<S, BS, NL, NS>), dim3(uint32_t(cells / BS)), dim3(BS), 0, st,
                       static_cast<const typename S::raw*>(fa), static_cast<typename S::raw*>(fb), g, R(omega), enc, dec);
  else
    hipLaunchKernelGGL((k_dense_step_vec<S, BS, V, NL, NS>), dim3(uint32_t(cells / (uint64_t(BS) * V))), dim3(BS), 0, st,
                       static_cast<const typename S::raw*>(fa), static_cast<typename S::raw*>(fb), g, R(omega), enc, dec);
  return hipGetLastError();
}
template <class S, int V>
static hipError_t step_bs(const void* fa, void* fb, const Dense& g, double omega, float enc, float dec, int v,
                          hipStream_t st) {
  const bool nl = !(v & 4), ns = !(v & 8);
#define JP_STEP_V(BS_)                                                                 \
  if (nl && ns) return step_v<S, BS_, V, true, true>(fa, fb, g, omega, enc, dec, st);   \
  if (nl) return step_v<S, BS_, V, true, false>(fa, fb, g, omega, enc, dec, st);        \
  if (ns) return step_v<S, BS_, V, false, true>(fa, fb, g, omega, enc, dec, st);        \
  return step_v<S, BS_, V, false, false>(fa, fb, g, omega, enc, dec, st);
  switch (v & 3) {
    case 1: { JP_STEP_V(128) }
    case 2: { JP_STEP_V(512) }
    case 3: { JP_STEP_V(1024) }
    default: { JP_STEP_V(256) }
  }
#undef JP_STEP_V
}
// variant: bits 0-1 workgroup size (256, 128, 512, 1024); bit 2 cached loads; bit 3 cached stores;
// bits 4-5 cells per lane (1, 2, 4, 8). Shapes that do not tile N^3 fall back to smaller ones.
template <class S> static hipError_t step_t(const void* fa, void* fb, const Dense& g, double omega, float enc,
                                            float dec, int v, hipStream_t st) {
  const uint64_t cells = uint64_t(g.N) * g.N * g.N;
  int bs = (v & 3) == 1 ? 128 : (v & 3) == 2 ? 512 : (v & 3) == 3 ? 1024 : 256;
  int V = 1 << ((v >> 4) & 3);
  if (g.N % V) V = 1;
  while (V > 1 && cells % (uint64_t(bs) * V)) V >>= 1;
  if (cells % bs) {
    v &= ~3;  // N^3 is a multiple of 512; 1024 needs an even tile count
    bs = 256;
  }
  switch (V) {
    case 2: return step_bs<S, 2>(fa, fb, g, omega, enc, dec, v, st);
    case 4: return step_bs<S, 4>(fa, fb, g, omega, enc, dec, v, st);
    case 8: return step_bs<S, 8>(fa, fb, g, omega, enc, dec, v, st);
    default: return step_bs<S, 1>(fa, fb, g, omega, enc, dec, v, st);
  }
}
hipError_t dense_step(int prec, const void* fa, void* fb, const Dense& g, double omega, float enc, float dec, int variant,
                      hipStream_t st) {
  switch (prec) {
    case 0: return step_t<FormF32>(fa, fb, g, omega, enc, dec, variant, st);
    case 1: return step_t<FormF16>(fa, fb, g, omega, enc, dec, variant, st);
    case 2: return step_t<FormBF16>(fa, fb, g, omega, enc, dec, variant, st);
    case 3: return step_t<FormF64>(fa, fb, g, omega, enc, dec, variant, st);
  }
  return hipErrorInvalidValue;
}

template <class S> static hipError_t macro_t(const void* f, const Dense& g, float dec, void* F, hipStream_t st) {
  const uint64_t cells = uint64_t(g.N) * g.N * g.N;
  hipLaunchKernelGGL((k_dense_macro<S, kBS>), dim3(uint32_t(cells / kBS)), dim3(kBS), 0, st,
                     static_cast<const typename S::raw*>(f), g, dec, static_cast<V4<typename S::real>*>(F));
  return hipGetLastError();
}
hipError_t dense_macro(int prec, const void* f, const Dense& g, float dec, void* field, hipStream_t st) {
  switch (prec) {
    case 0: return macro_t<FormF32>(f, g, dec, field, st);
    case 1: return macro_t<FormF16>(f, g, dec, field, st);
    case 2: return macro_t<FormBF16>(f, g, dec, field, st);
    case 3: return macro_t<FormF64>(f, g, dec, field, st);
  }
  return hipErrorInvalidValue;
}

hipError_t field_diag(bool dbl, bool diff_in_field, const void* field, int N, double* partial, int nblocks,
                      double* out3, hipStream_t st) {
  const FastDiv dN = make_fastdiv(uint32_t(N));
  if (dbl) {
    if (diff_in_field)
      hipLaunchKernelGGL((k_field_diag<double, true, kBS>), dim3(nblocks), dim3(kBS), 0, st,
                         static_cast<const V4<double>*>(field), uint32_t(N), dN, partial);
    else
      hipLaunchKernelGGL((k_field_diag<double, false, kBS>), dim3(nblocks), dim3(kBS), 0, st,
                         static_cast<const V4<double>*>(field), uint32_t(N), dN, partial);
  } else {
    if (diff_in_field)
      hipLaunchKernelGGL((k_field_diag<float, true, kBS>), dim3(nblocks), dim3(kBS), 0, st,
                         static_cast<const V4<float>*>(field), uint32_t(N), dN, partial);
    else
      hipLaunchKernelGGL((k_field_diag<float, false, kBS>), dim3(nblocks), dim3(kBS), 0, st,
                         static_cast<const V4<float>*>(field), uint32_t(N), dN, partial);
  }
  const double inv = 1.0 / (double(N) * N * N);
  hipLaunchKernelGGL((k_diag_final<kBS>), dim3(1), dim3(kBS), 0, st, partial, nblocks, inv, out3);
  return hipGetLastError();
}

hipError_t bandwidth_probe(int mode, const void* src, void* dst, size_t bytes, hipStream_t st) {
  const size_t n = bytes / 16;  // a multiple of 256
  const dim3 grid(uint32_t(n / 256));
  auto s = static_cast<const u32x4*>(src);
  auto d = static_cast<u32x4*>(dst);
  if (mode == 0) hipLaunchKernelGGL(k_probe<0>, grid, dim3(256), 0, st, s, d);
  else if (mode == 1) hipLaunchKernelGGL(k_probe<1>, grid, dim3(256), 0, st, s, d);
  else hipLaunchKernelGGL(k_probe<2>, grid, dim3(256), 0, st, s, d);
  return hipGetLastError();
}

hipError_t copy_probe(const void* src, void* dst, size_t bytes, hipStream_t st) {
  const size_t n = bytes / 16;
  hipLaunchKernelGGL(k_copy16, dim3(8192), dim3(256), 0, st, static_cast<const u32x4*>(src), static_cast<u32x4*>(dst), n);
  return hipGetLastError();
}

template <class T> static hipError_t conv_t(const void* s, const Dense& g, void* d, bool to_tiles, hipStream_t st) {
  const uint32_t gx = g.N / MB, nb = gx * gx * gx;
  if (to_tiles)
    hipLaunchKernelGGL((k_dense_tiles<T, true>), dim3(nb * Q), dim3(MB3), 0, st, static_cast<const T*>(s),
                       static_cast<T*>(d), g);
  else
    hipLaunchKernelGGL((k_dense_tiles<T, false>), dim3(nb * Q), dim3(MB3), 0, st, static_cast<const T*>(s),
                       static_cast<T*>(d), g);
  return hipGetLastError();
}
hipError_t dense_to_tiles(int prec, const void* f, const Dense& g, void* tiles, hipStream_t st) {
  switch (prec_bytes(prec)) {
    case 2: return conv_t<uint16_t>(f, g, tiles, true, st);
    case 4: return conv_t<uint32_t>(f, g, tiles, true, st);
    default: return conv_t<uint64_t>(f, g, tiles, true, st);
  }
}
hipError_t tiles_to_dense(int prec, const void* tiles, const Dense& g, void* f, hipStream_t st) {
  switch (prec_bytes(prec)) {
    case 2: return conv_t<uint16_t>(tiles, g, f, false, st);
    case 4: return conv_t<uint32_t>(tiles, g, f, false, st);
    default: return conv_t<uint64_t>(tiles, g, f, false, st);
  }
}

hipError_t fill_i32(int* p, int v, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const uint32_t blocks = uint32_t(std::min<size_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

}  // namespace jp
