// Dense periodic D3Q19 lattice on gfx950: TGV initial state, fused pull stream-collide, macroscopic
// field, FP64 energy/enstrophy/mass reduction. Used for the uniform solver and for AMR level 0.
//
// Layout: 19 x-fastest planes, f[q*plane + (z*N + y)*N + x]. A wave of 64 lanes covers 64 consecutive
// cells of one x-row, so every population load is a contiguous 256 B (FP32) run, shifted by one element
// for directions with c_x != 0 (the partial line is shared with the neighbouring wave of the same block
// and served by the CU's L1). Each element of every plane is read exactly once per step, so the single
// step moves the algorithmic 2*19 elements per cell update and nothing else. The FP32 uniform lattice
// runs two steps per launch instead (k_dense_step2), which keeps the intermediate step on the chip and
// moves fewer bytes than two single steps.
#include <algorithm>
#include <mutex>

#include "dense_bodies.hpp"

namespace jp {

// TGV initial state, src/tgv3d_uniform.cu:65-77: FP64 trigonometry, FP32 (working type) equilibrium.
template <class S, int BS>
__global__ __launch_bounds__(BS) void k_dense_init(typename S::raw* __restrict__ f, Dense g, double u0, double kf,
                                                   double px, double py, double pz, float enc) {
  using R = typename S::real;
  const uint32_t cell = blockIdx.x * BS + threadIdx.x;
  const uint32_t row = fdiv(cell, g.divN), x = cell - row * g.N, z = fdiv(row, g.divN), y = row - z * g.N;
  const EqPre<R> e = tgv_equilibrium<R>(x, y, z, u0, kf, px, py, pz);
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    f[g.at(q, z, y * g.N + x)] = S::template store<q>(feq<q>(e), enc);
  });
}

// Fused pull stream-collide over the whole lattice, one cell per lane.
template <class S, int BS, bool kNtLoad, bool kNtStore>
__global__ __launch_bounds__(BS) void k_dense_step(const typename S::raw* __restrict__ fa,
                                                   typename S::raw* __restrict__ fb, Dense g,
                                                   typename S::real omega, float enc, float dec) {
  dense_step_cell<S, kNtLoad, kNtStore>(fa, fb, g, omega, enc, dec, blockIdx.x * BS + threadIdx.x);
}

// The same when a workgroup lies inside one x-row (N a multiple of BS): y and z are the workgroup's, so
// every direction's source row starts at a workgroup-uniform address (scalar registers) and a lane adds
// only its 32-bit x offset, one of three per lane. The loads and stores then take the scalar-base address
// form and need no 64-bit lane arithmetic (about 95 of the 16-bit step's 425 vector instructions per cell
// in the form above). Each cell's arithmetic is dense_step_cell's.
template <class S, int BS, bool kNtLoad, bool kNtStore>
__global__ __launch_bounds__(BS) void k_dense_step_row(const typename S::raw* __restrict__ fa,
                                                       typename S::raw* __restrict__ fb, Dense g,
                                                       typename S::real omega, float enc, float dec) {
  using R = typename S::real;
  using raw = typename S::raw;
  const uint32_t N = g.N, per_row = N / BS;
  const uint32_t row = blockIdx.x / per_row, x = (blockIdx.x - row * per_row) * BS + threadIdx.x;
  const uint32_t z = fdiv(row, g.divN), y = row - z * N;
  // index c + 1 of a direction component c: the source lies at destination - c, periodic
  const uint32_t xb[3] = {uint32_t((x == N - 1 ? 0 : x + 1) * sizeof(raw)), uint32_t(x * sizeof(raw)),
                          uint32_t((x == 0 ? N - 1 : x - 1) * sizeof(raw))};
  const uint64_t zo[3] = {uint64_t(z == N - 1 ? 0 : z + 1) * g.zs, uint64_t(z) * g.zs, uint64_t(z == 0 ? N - 1 : z - 1) * g.zs};
  const uint32_t yo[3] = {(y == N - 1 ? 0 : y + 1) * N, y * N, (y == 0 ? N - 1 : y - 1) * N};
  R fin[Q];
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const char* base = reinterpret_cast<const char*>(fa + (uint64_t(q) * g.qs + zo[CZ(q) + 1] + yo[CY(q) + 1]));
    const raw* p = reinterpret_cast<const raw*>(base + xb[CX(q) + 1]);
    if constexpr (kNtLoad) fin[q] = S::template in<q>(__builtin_nontemporal_load(p), dec);
    else fin[q] = S::template in<q>(*p, dec);
  });
  const Collider<S> col(fin, dec, enc);
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const raw v = S::template out<q>(col.template post<q>(fin[q], omega), enc);
    char* base = reinterpret_cast<char*>(fb + (uint64_t(q) * g.qs + zo[1] + yo[1]));
    raw* p = reinterpret_cast<raw*>(base + xb[1]);
    if constexpr (kNtStore) __builtin_nontemporal_store(v, p);
    else *p = v;
  });
}

// Macroscopic field, src/tgv3d_uniform.cu:80-90.
template <class S, int BS>
__global__ __launch_bounds__(BS) void k_dense_macro(const typename S::raw* __restrict__ f, Dense g, float dec,
                                                    V4<typename S::real>* __restrict__ F) {
  using R = typename S::real;
  const uint32_t cell = blockIdx.x * BS + threadIdx.x;
  const uint32_t z = fdiv(fdiv(cell, g.divN), g.divN), r = cell - z * g.N * g.N;
  R v[Q];
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    v[q] = S::template load<q>(f[g.at(q, z, r)], dec);
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

// Energy, enstrophy and mass of the dense lattice in one pass, with no macroscopic field in HBM (the
// uniform driver's diagnostics, src/tgv3d_uniform.cu:80-90, 99-108, 170-183). A work item is a column of
// kDiagTX x kDiagTY cells over a segment of at most kDiagSeg z-planes. The workgroup computes rho and u of
// each plane of the column's (kDiagTX + 2) x (kDiagTY + 2) extended tile, one plane ahead, into a ring of
// three LDS planes; each lane accumulates its cell's terms in FP64 exactly as k_dense_macro +
// k_field_diag<R, true> do (the same values, the same per-cell operation order). One partial per item,
// stored at the item's index, so k_diag_final adds them in a fixed order: the sum is deterministic.
// Items are dealt to the XCDs in contiguous runs (xcd_tile), so neighbouring columns share halo rows in
// one L2.
constexpr int kDiagTX = 64, kDiagTY = 8, kDiagSeg = 128, kDiagBS = kDiagTX * kDiagTY;
template <class S>
__global__ __launch_bounds__(kDiagBS) void k_dense_diag(const typename S::raw* __restrict__ f, Dense g, float dec,
                                                       uint32_t ntx, uint32_t ncol, uint32_t nitems,
                                                       double* __restrict__ partial) {
  using R = typename S::real;
  constexpr int EX = kDiagTX + 2, EY = kDiagTY + 2, NE = EX * EY;
  __shared__ V4<R> ring[3][NE];
  const uint32_t N = g.N;
  const uint32_t item = uint32_t(xcd_tile(int(blockIdx.x), int(nitems)));
  const uint32_t col = item % ncol, seg = item / ncol;
  const uint32_t x0 = (col % ntx) * kDiagTX, y0 = (col / ntx) * kDiagTY;
  const uint32_t z0 = seg * kDiagSeg, z1 = min(N, z0 + kDiagSeg);
  // rho, u of the extended tile of plane z (x and y periodic)
  auto plane = [&](V4<R>* P, uint32_t z) {
    for (int e = threadIdx.x; e < NE; e += kDiagBS) {
      const int ex = e % EX, ey = e / EX;
      const uint32_t x = (x0 + ex + N - 1) % N, y = (y0 + ey + N - 1) % N, r = y * N + x;
      R v[Q];
      sfor<Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        v[q] = S::template load<q>(f[g.at(q, z, r)], dec);
      });
      const Mom<R> m = moments(v);
      V4<R> o;
      o.a = m.rho;
      o.b = m.ux / m.rho;
      o.c = m.uy / m.rho;
      o.d = m.uz / m.rho;
      P[e] = o;
    }
  };
  const int lx = threadIdx.x % kDiagTX, ly = threadIdx.x / kDiagTX;
  const bool live = x0 + lx < N && y0 + ly < N;
  const int c = (ly + 1) * EX + lx + 1;
  double E = 0, ens = 0, mass = 0;
  int sm = 0, sc = 1, sp = 2;  // ring slots of planes z - 1, z, z + 1
  plane(ring[sm], z0 == 0 ? N - 1 : z0 - 1);
  plane(ring[sc], z0);
  for (uint32_t z = z0; z < z1; ++z) {
    plane(ring[sp], z + 1 == N ? 0 : z + 1);
    __syncthreads();
    if (live) {
      const V4<R>* P = ring[sc];
      const V4<R> cc = P[c], xp = P[c + 1], xm = P[c - 1], yp = P[c + EX], ym = P[c - EX];
      const V4<R> zp = ring[sp][c], zm = ring[sm][c];
      auto dif = [](R p, R m) -> double { return double(p - m) * 0.5; };
      const double dudy = dif(yp.b, ym.b), dudz = dif(zp.b, zm.b);
      const double dvdx = dif(xp.c, xm.c), dvdz = dif(zp.c, zm.c);
      const double dwdx = dif(xp.d, xm.d), dwdy = dif(yp.d, ym.d);
      const double ox = dwdy - dvdz, oy = dudz - dwdx, oz = dvdx - dudy;
      const double u = cc.b, v = cc.c, w = cc.d;
      E += 0.5 * (u * u + v * v + w * w);
      ens += 0.5 * (ox * ox + oy * oy + oz * oz);
      mass += double(cc.a);
    }
    __syncthreads();
    const int t = sm;
    sm = sc;
    sc = sp;
    sp = t;
  }
  block_sum3<kDiagBS>(E, ens, mass, partial + 3 * item);
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
  const uint64_t d = g.at(int(q), bz * MB + k, (by * MB + j) * g.N + (bx * MB + i));
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

// ---- two fused steps per pass ----
// A workgroup owns a TX x TY column of the lattice over TZ consecutive z-planes and streams up the column:
// at iteration it it computes the intermediate step (t + 1) on the plane zo0 - 1 + it of an extended
// (TX + 2) x (TY + 2) tile (phase A, each cell pulling from the lattice at t, exactly as the single
// step), keeps it in LDS, then computes the step t + 2 of the column's plane zo0 + it - 2 from the three
// intermediate planes around it (phase B) and writes it. The intermediate state never leaves the chip:
// per two steps HBM sees the t-state of the extended tile read once (its x-y halo and the two extra
// z-planes re-read by the neighbouring columns) and the t + 2 state written once, instead of two full
// reads and writes. Every cell's arithmetic is the single step's (src/tgv3d_uniform.cu:41-61), the
// intermediate populations pass through the storage form as a stored population would, so the result
// is bit-identical to two single steps.
//
// LDS holds each direction's intermediate planes as a ring: the second step at plane o pulls direction q
// from plane o - c_z, and the newest intermediate plane is o + 1, so phase B of a plane reads the newest
// plane of a direction with c_z = -1, the second newest of one with c_z = 0 and the third newest of one with
// c_z = +1. Rings of 1, 2 and 3 planes (38 in all) need a second barrier after phase B, before the next
// phase A overwrites a slot phase B reads. The 16-bit forms keep one plane more per ring (2, 3 and 4, 57 in
// all, 58.4 KB): phase A of the next plane then writes slots no phase B of this plane reads, one barrier
// per plane (between A and B) orders the rings, and the waves with no phase-B cell (832 lanes, 512 output
// cells) go on to the next plane's phase A while the others finish phase B (512^3: FP16 85.0-87.4 ->
// 87.8-88.2 GLUPS, BF16 85.4-86.2 -> 90.0). The FP32 pair, which waits on memory more than on issue, ran
// 1-2 % slower so (63.4-63.5 against 64.4-64.7 GLUPS) and keeps two barriers.
// A ring plane of direction q holds only the TX x TY values the second step pulls, indexed by the output
// cell that pulls them: the extended cell (ex, ey) feeds direction q of output cell (ex - 1 + c_x,
// ey - 1 + c_y) and nothing else, so phase A stores each value at most once and phase B reads every
// direction at its own cell index.
__host__ __device__ constexpr int ring_depth(int q, int extra) { return 2 + extra + CZ(q); }
__host__ __device__ constexpr int ring_base(int q, int extra) {
  int s = 0;
  for (int i = 0; i < q; ++i) s += ring_depth(i, extra);
  return s;
}

// Workgroup barrier that orders LDS only: the population loads of the next plane stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <class S, int TX, int TY, int BS>
__global__ __launch_bounds__(BS) void k_dense_step2(const typename S::raw* __restrict__ fa,
                                                    typename S::raw* __restrict__ fb, Dense g, uint32_t TZ,
                                                    typename S::real omega, float enc, float dec) {
  using raw = typename S::raw;
  using R = typename S::real;
  constexpr int EX = TX + 2, EY = TY + 2, E = EX * EY, T = TX * TY;
  constexpr int PA = (E + BS - 1) / BS, PB = (T + BS - 1) / BS;
  constexpr int X = sizeof(raw) == 2 ? 1 : 0;  // extra ring plane per direction: one barrier per plane
  __shared__ raw ring[ring_base(Q, X) * T];
  const uint32_t N = g.N;
  const uint32_t ntx = N / TX, nty = N / TY, nt = ntx * nty * (N / TZ);
  // every XCD streams a contiguous run of columns, so halo rows shared by neighbouring columns are read
  // through one L2
  const uint32_t t = uint32_t(xcd_tile(int(blockIdx.x), int(nt)));
  const uint32_t bx = t % ntx, by = (t / ntx) % nty, zo0 = (t / (ntx * nty)) * TZ;
  const int tid = threadIdx.x;
  // Phase-A cells of this lane: extended-tile cell e = a * BS + tid. Spare lanes of the last pass take
  // the last cell's addresses (they load and discard), so every lane issues the same loads and the
  // compiler can count them: the next plane's loads then stay in flight while the current one computes.
  uint32_t colx[PA][3], coly[PA][3];
  int exa[PA], eya[PA];
  bool va[PA];
#pragma unroll
  for (int a = 0; a < PA; ++a) {
    const int e0 = a * BS + tid;
    va[a] = e0 < E;
    const int e = e0 < E ? e0 : E - 1;
    exa[a] = e % EX;
    eya[a] = e / EX;
    const int gx = int(bx * TX) + exa[a] - 1, gy = int(by * TY) + eya[a] - 1;
    const uint32_t x = gx < 0 ? N - 1 : (gx >= int(N) ? 0 : gx), y = gy < 0 ? N - 1 : (gy >= int(N) ? 0 : gy);
    // source column of c = +1, 0, -1 (index 1 - c): destination - c, periodic
    colx[a][0] = x == 0 ? N - 1 : x - 1;
    colx[a][1] = x;
    colx[a][2] = x == N - 1 ? 0 : x + 1;
    coly[a][0] = (y == 0 ? N - 1 : y - 1) * N;
    coly[a][1] = y * N;
    coly[a][2] = (y == N - 1 ? 0 : y + 1) * N;
  }
  auto load_plane = [&](raw(&r)[PA][Q], int it) {
    const int gp = int(zo0) - 1 + it;
    const uint32_t z = gp < 0 ? N - 1 : (gp >= int(N) ? 0 : gp);
    const uint64_t zs[3] = {uint64_t(z == 0 ? N - 1 : z - 1) * g.zs, uint64_t(z) * g.zs, uint64_t(z == N - 1 ? 0 : z + 1) * g.zs};
#pragma unroll
    for (int a = 0; a < PA; ++a)
      sfor<Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        // workgroup-uniform base of the source plane (scalar registers) plus the lane's 32-bit byte offset
        // within it: the load takes the scalar-base address form and needs no 64-bit lane arithmetic
        const char* base = reinterpret_cast<const char*>(fa + (uint64_t(q) * g.qs + zs[1 - CZ(q)]));
        r[a][q] = *reinterpret_cast<const raw*>(base + (coly[a][1 - CY(q)] + colx[a][1 - CX(q)]) * uint32_t(sizeof(raw)));
      });
  };
  auto body = [&](const raw(&cur)[PA][Q], int it) {
    const int m2 = it & 1, m3 = it % 3, m4 = it & 3;
    // phase A: intermediate plane it of the extended tile, into slot it mod depth
#pragma unroll
    for (int a = 0; a < PA; ++a) {
      if (va[a]) {
        R fin[Q];
        sfor<Q>([&](auto qc) { fin[decltype(qc)::value] = S::template in<decltype(qc)::value>(cur[a][decltype(qc)::value], dec); });
        Collider<S>(fin, dec, enc).post_all(fin, omega);
        sfor<Q>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          constexpr int d = ring_depth(q, X);
          const int slot = d == 1 ? 0 : (d == 2 ? m2 : (d == 3 ? m3 : m4));
          const int ox = exa[a] - 1 + CX(q), oy = eya[a] - 1 + CY(q);  // the output cell pulling q from here
          if (unsigned(ox) < unsigned(TX) && unsigned(oy) < unsigned(TY))
            ring[(ring_base(q, X) + slot) * T + oy * TX + ox] = S::template out<q>(fin[q], enc);
        });
      }
    }
    lds_barrier();
    // phase B: output plane o = zo0 + it - 2 from intermediate planes it - 2, it - 1, it
    if (it >= 2) {
      const uint32_t o = zo0 + it - 2;
#pragma unroll
      for (int b = 0; b < PB; ++b) {
        const int c = b * BS + tid;
        if (c < T) {
          const int ox = c % TX, oy = c / TX;
          R fin[Q];
          sfor<Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int d = ring_depth(q, X);
            // plane o - c_z: c_z = -1 the newest (that of it), 0 that of it - 1, +1 that of it - 2
            const int slot = X == 0 ? (d == 1 ? 0 : (d == 2 ? (m2 ^ 1) : (m3 == 2 ? 0 : m3 + 1)))
                                    : (d == 2 ? m2 : (d == 3 ? (m3 == 0 ? 2 : m3 - 1) : (m4 ^ 2)));
            fin[q] = S::template in<q>(ring[(ring_base(q, X) + slot) * T + c], dec);
          });
          Collider<S>(fin, dec, enc).post_all(fin, omega);
          const uint32_t r = (by * TY + oy) * N + bx * TX + ox;
          sfor<Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            fb[g.at(q, o, r)] = S::template out<q>(fin[q], enc);
          });
        }
      }
    }
    if constexpr (X == 0) lds_barrier();
  };
  // intermediate planes 0 .. TZ + 1 in pairs (TZ is even); the last prefetch repeats plane TZ + 1
  raw ra[PA][Q], rb[PA][Q];
  load_plane(ra, 0);
  for (int it = 0; it < int(TZ) + 2; it += 2) {
    load_plane(rb, it + 1);
    body(ra, it);
    load_plane(ra, it + 2 < int(TZ) + 2 ? it + 2 : int(TZ) + 1);
    body(rb, it + 1);
  }
}

// Column shapes of the fused pair, in order of preference (measured on MI355X at 512^3 FP32: 128 x 4
// with 832 lanes 58.3 GLUPS, 64 x 8 / 704 lanes 54.3, against 34.2 for single steps; taller columns need
// two phase-A passes per lane, whose doubled prefetch registers spill at 1024 lanes). FP32, FP16 and BF16
// fuse by default. The 16-bit forms move half the bytes, so their fused pass (2.5 collisions per cell where
// the single step has 1) is bound by arithmetic rather than bytes; since they collide on the stored scale
// (no decode or encode multiplies) their pair runs 83.0 / 82.8 GLUPS at 512^3 against 79.5 / 78.9 for
// their single steps (FP16 / BF16, 128 x 4 columns on the padded z-layers). FP64 would need
// twice the LDS per cell. Tried and dropped: buffer addressing (one descriptor per z-layer, 32-bit lane
// offsets) cut the pair's registers from 102 to 69 and allowed two workgroups per CU, but ran no faster
// (FP32 128 x 4: 63.3 against 64.4 GLUPS with global addresses); with the direction's offset in a scalar
// register it also cut 12-14 % of the VALU instructions, and still lost on the issue-bound 16-bit pairs
// (FP16 81.2 against 82.5, BF16 80.4 against 82.1, same process).
struct PairShape {
  int tx, ty, bs;
};
static const PairShape kPairShapes[] = {{128, 4, 832}, {64, 8, 704}, {32, 16, 640}};
constexpr int kShapeCount = int(sizeof(kPairShapes) / sizeof(kPairShapes[0]));
// Launch forms 0 .. kShapeCount - 1: each column shape with the longest z-segment; the next kShapeCount: the
// same with half that segment (at least 32 planes), for the autotuner to try where boxes differ.
int step_pair_shape_count() { return 2 * kShapeCount; }
bool step_pair_shape(int prec, int N, int pick, int* out4) {
  if (prec == 3) return false;
  // z-segment: the longest of 256, 128, 64, 32 planes that divides N (at 512^3 on padded z-layers 256 planes
  // ran 62.0 against 61.2 GLUPS for 128: two halo planes per segment are 0.8 % of the reads instead of 1.6 %)
  const int tz0 = N % 256 == 0 ? 256 : (N % 128 == 0 ? 128 : (N % 64 == 0 ? 64 : (N % 32 == 0 ? 32 : 0)));
  if (!tz0) return false;
  for (int i = 0; i < step_pair_shape_count(); ++i) {
    const PairShape& p = kPairShapes[i % kShapeCount];
    const int tz = i < kShapeCount ? tz0 : tz0 / 2;
    if ((pick >= 0 && i != pick) || N % p.tx || N % p.ty || tz < 32) continue;
    out4[0] = p.tx;
    out4[1] = p.ty;
    out4[2] = tz;
    out4[3] = p.bs;
    return true;
  }
  return false;
}

// occ != nullptr: instead of launching, *occ = workgroups of this kernel one CU holds at once
template <class S, int TX, int TY, int BS>
static hipError_t pair_launch(const void* fa, void* fb, const Dense& g, int tz, double omega, float enc, float dec,
                              hipStream_t st, int* occ = nullptr) {
  using raw = typename S::raw;
  if (occ) return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, k_dense_step2<S, TX, TY, BS>, BS, 0);
  const uint32_t nt = (g.N / TX) * (g.N / TY) * (g.N / uint32_t(tz));
  hipLaunchKernelGGL((k_dense_step2<S, TX, TY, BS>), dim3(nt), dim3(BS), 0, st, static_cast<const raw*>(fa),
                     static_cast<raw*>(fb), g, uint32_t(tz), typename S::real(omega), enc, dec);
  return hipGetLastError();
}
template <class S>
static hipError_t pair_shape(const void* fa, void* fb, const Dense& g, const int* s, double omega, float enc, float dec,
                             hipStream_t st, int* occ) {
  if (s[0] == 128) return pair_launch<S, 128, 4, 832>(fa, fb, g, s[2], omega, enc, dec, st, occ);
  if (s[0] == 64) return pair_launch<S, 64, 8, 704>(fa, fb, g, s[2], omega, enc, dec, st, occ);
  return pair_launch<S, 32, 16, 640>(fa, fb, g, s[2], omega, enc, dec, st, occ);
}
static hipError_t pair_dispatch(int prec, const void* fa, void* fb, const Dense& g, double omega, float enc, float dec,
                                int pick, hipStream_t st, int* occ) {
  int s[4];
  if (!step_pair_shape(prec, int(g.N), pick, s)) return hipErrorInvalidValue;
  switch (prec) {
    case 0: return pair_shape<FormF32>(fa, fb, g, s, omega, enc, dec, st, occ);
    case 1: return pair_shape<FormF16>(fa, fb, g, s, omega, enc, dec, st, occ);
    case 2: return pair_shape<FormBF16>(fa, fb, g, s, omega, enc, dec, st, occ);
  }
  return hipErrorInvalidValue;
}
hipError_t dense_step_pair(int prec, const void* fa, void* fb, const Dense& g, double omega, float enc, float dec,
                           int pick, hipStream_t st) {
  return pair_dispatch(prec, fa, fb, g, omega, enc, dec, pick, st, nullptr);
}

template <class S, int BS, int V, bool NL, bool NS>
static hipError_t step_v(const void* fa, void* fb, const Dense& g, double omega, float enc, float dec, hipStream_t st,
                         int* occ) {
  using R = typename S::real;
  const uint64_t cells = uint64_t(g.N) * g.N * g.N;
  const bool in_row = g.N % BS == 0;  // every workgroup inside one x-row
  if (occ) {  // workgroups one CU holds at once, instead of a launch
    if constexpr (V == 1) {
      if (in_row) return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, k_dense_step_row<S, BS, NL, NS>, BS, 0);
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, k_dense_step<S, BS, NL, NS>, BS, 0);
    } else {
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, k_dense_step_vec<S, BS, V, NL, NS>, BS, 0);
    }
  }
  if constexpr (V == 1) {
    if (in_row)
      hipLaunchKernelGGL((k_dense_step_row<S, BS, NL, NS>), dim3(uint32_t(cells / BS)), dim3(BS), 0, st,
                         static_cast<const typename S::raw*>(fa), static_cast<typename S::raw*>(fb), g, R(omega), enc, dec);
    else
      hipLaunchKernelGGL((k_dense_step<S, BS, NL, NS>), dim3(uint32_t(cells / BS)), dim3(BS), 0, st,
                         static_cast<const typename S::raw*>(fa), static_cast<typename S::raw*>(fb), g, R(omega), enc, dec);
  }
  else
    hipLaunchKernelGGL((k_dense_step_vec<S, BS, V, NL, NS>), dim3(uint32_t(cells / (uint64_t(BS) * V))), dim3(BS), 0, st,
                       static_cast<const typename S::raw*>(fa), static_cast<typename S::raw*>(fb), g, R(omega), enc, dec);
  return hipGetLastError();
}
template <class S, int V>
static hipError_t step_bs(const void* fa, void* fb, const Dense& g, double omega, float enc, float dec, int v,
                          hipStream_t st, int* occ) {
  const bool nl = !(v & 4), ns = !(v & 8);
#define JP_STEP_V(BS_)                                                                      \
  if (nl && ns) return step_v<S, BS_, V, true, true>(fa, fb, g, omega, enc, dec, st, occ);   \
  if (nl) return step_v<S, BS_, V, true, false>(fa, fb, g, omega, enc, dec, st, occ);        \
  if (ns) return step_v<S, BS_, V, false, true>(fa, fb, g, omega, enc, dec, st, occ);        \
  return step_v<S, BS_, V, false, false>(fa, fb, g, omega, enc, dec, st, occ);
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
                                            float dec, int v, hipStream_t st, int* occ, int* lanes) {
  const uint64_t cells = uint64_t(g.N) * g.N * g.N;
  int bs = (v & 3) == 1 ? 128 : (v & 3) == 2 ? 512 : (v & 3) == 3 ? 1024 : 256;
  int V = 1 << ((v >> 4) & 3);
  if (g.N % V) V = 1;
  while (V > 1 && cells % (uint64_t(bs) * V)) V >>= 1;
  if (cells % bs) {
    v &= ~3;  // N^3 is a multiple of 512; 1024 needs an even tile count
    bs = 256;
  }
  if (lanes) *lanes = bs;
  switch (V) {
    case 2: return step_bs<S, 2>(fa, fb, g, omega, enc, dec, v, st, occ);
    case 4: return step_bs<S, 4>(fa, fb, g, omega, enc, dec, v, st, occ);
    case 8: return step_bs<S, 8>(fa, fb, g, omega, enc, dec, v, st, occ);
    default: return step_bs<S, 1>(fa, fb, g, omega, enc, dec, v, st, occ);
  }
}
hipError_t dense_step(int prec, const void* fa, void* fb, const Dense& g, double omega, float enc, float dec, int variant,
                      hipStream_t st) {
  switch (prec) {
    case 0: return step_t<FormF32>(fa, fb, g, omega, enc, dec, variant, st, nullptr, nullptr);
    case 1: return step_t<FormF16>(fa, fb, g, omega, enc, dec, variant, st, nullptr, nullptr);
    case 2: return step_t<FormBF16>(fa, fb, g, omega, enc, dec, variant, st, nullptr, nullptr);
    case 3: return step_t<FormF64>(fa, fb, g, omega, enc, dec, variant, st, nullptr, nullptr);
  }
  return hipErrorInvalidValue;
}
int dense_step_waves_per_cu(int prec, const Dense& g, int variant, int pick) {
  int blocks = 0, lanes = 0, s[4];
  hipError_t e = hipErrorInvalidValue;
  if (pick >= 0) {
    if (step_pair_shape(prec, int(g.N), pick, s)) {
      e = pair_dispatch(prec, nullptr, nullptr, g, 0.0, 1.f, 1.f, pick, nullptr, &blocks);
      lanes = s[3];
    }
  } else {
    switch (prec) {
      case 0: e = step_t<FormF32>(nullptr, nullptr, g, 0.0, 1.f, 1.f, variant, nullptr, &blocks, &lanes); break;
      case 1: e = step_t<FormF16>(nullptr, nullptr, g, 0.0, 1.f, 1.f, variant, nullptr, &blocks, &lanes); break;
      case 2: e = step_t<FormBF16>(nullptr, nullptr, g, 0.0, 1.f, 1.f, variant, nullptr, &blocks, &lanes); break;
      case 3: e = step_t<FormF64>(nullptr, nullptr, g, 0.0, 1.f, 1.f, variant, nullptr, &blocks, &lanes); break;
    }
  }
  return e == hipSuccess ? blocks * ((lanes + 63) / 64) : 0;
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

static void diag_grid(uint32_t N, uint32_t* ntx, uint32_t* ncol, uint32_t* nitems) {
  *ntx = (N + kDiagTX - 1) / kDiagTX;
  *ncol = *ntx * ((N + kDiagTY - 1) / kDiagTY);
  *nitems = *ncol * ((N + kDiagSeg - 1) / kDiagSeg);
}
uint32_t dense_diag_items(int N) {
  uint32_t a, b, n;
  diag_grid(uint32_t(N), &a, &b, &n);
  return n;
}
template <class S>
static hipError_t ddiag_t(const void* f, const Dense& g, float dec, double* partial, double* out3, hipStream_t st) {
  uint32_t ntx, ncol, nitems;
  diag_grid(g.N, &ntx, &ncol, &nitems);
  hipLaunchKernelGGL((k_dense_diag<S>), dim3(nitems), dim3(kDiagBS), 0, st, static_cast<const typename S::raw*>(f),
                     g, dec, ntx, ncol, nitems, partial);
  const double inv = 1.0 / (double(g.N) * g.N * g.N);
  hipLaunchKernelGGL((k_diag_final<kBS>), dim3(1), dim3(kBS), 0, st, partial, int(nitems), inv, out3);
  return hipGetLastError();
}
hipError_t dense_diag(int prec, const void* f, const Dense& g, float dec, double* partial, double* out3,
                      hipStream_t st) {
  switch (prec) {
    case 0: return ddiag_t<FormF32>(f, g, dec, partial, out3, st);
    case 1: return ddiag_t<FormF16>(f, g, dec, partial, out3, st);
    case 2: return ddiag_t<FormBF16>(f, g, dec, partial, out3, st);
    case 3: return ddiag_t<FormF64>(f, g, dec, partial, out3, st);
  }
  return hipErrorInvalidValue;
}

// Fixed device array for the partial sums of dense_diag when the caller cannot allocate (jp_diag_noalloc):
// kNoAllocSlots partials of 3 doubles, then the 3 results (768 KiB). It holds the partials of every lattice
// the library creates (edge <= 1280: 32000 items). Each GPU the library runs on holds its own copy until the
// process ends; every call on one GPU shares it. Host code must not pass `scratch` itself to a kernel (in
// host code the name is a host-side stand-in): noalloc_scratch looks up each device's address once, with
// that device current, and every launch takes that address.
constexpr size_t kNoAllocSlots = 32768;
static __device__ __align__(256) double scratch[3 * kNoAllocSlots + 8];
constexpr int kMaxDevices = 64;
static std::mutex scratch_mu;
static double* scratch_dev[kMaxDevices] = {};
hipError_t noalloc_scratch(int device, double** ptr, size_t* slots) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // the caller made `device` current
    if (e != hipSuccess) return e;
    scratch_dev[device] = static_cast<double*>(p);
  }
  *ptr = scratch_dev[device];
  *slots = kNoAllocSlots;
  return hipSuccess;
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
