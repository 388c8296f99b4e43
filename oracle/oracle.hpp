// TEST INFRASTRUCTURE ONLY. CPU restatement of jp-solver's D3Q19 TGV time step (uniform and 3-level
// block AMR), used as the checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
// Nothing in the product library links, loads or calls this code.
//
// Every function follows the reference method in src/tgv3d_uniform.cu and src/tgv3d_amr.cu line for line
// in its arithmetic (same operation order, IEEE single precision with no FMA contraction), so that the
// GPU library, built with the same contraction rule, can be compared bit for bit. The population layout
// here is the reference's tile-major one: f[(b*19 + q)*512 + cell], cell = (k*8 + j)*8 + i,
// b = (bz*gx + by)*gx + bx (src/tgv3d_uniform.cu:42-53).
//
// Precision forms (not in the reference): the populations may be stored as FP16 or BF16 values of the
// shifted quantity (f - w_q) * scale, decoded to FP32 for the collision, or kept in FP64 throughout.
// The storage rule is restated here so that the GPU forms can be checked against it bit for bit.
//
// Parity pins (see tests/test_oracle_*.py): closed forms of the initial diagnostics, conservation and
// fixed-point identities of the collision, the omega = 0 pure-shift permutation, Hermite-projection
// identities, brute-force table logic, std::sort / std::exclusive_scan / stable std::copy_if semantics
// for the Thrust call sites, numpy's float16 rounding for the FP16 store. The coarse-fine interface
// (seeding, ghost fill, the fine steps' ghost reads, restriction, migration, sensor) is pinned by closed
// forms on affine fields carrying a traceless Hermite mode, a renormalised constant, and an omega = 0
// shift of coded values (tests/interface_pins.py); tests/test_oracle_mutants.py plants plausible mistakes
// in copies of these sources and checks that the pins catch each one.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace orc {

constexpr int Q = 19, MB = 8, MB3 = MB * MB * MB, EB = MB + 2, EB3 = EB * EB * EB, HALF = MB / 2;

// D3Q19 velocity set, q order of src/tgv3d_uniform.cu:28-30 and src/tgv3d_amr.cu:41-43.
static const int cvx[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int cvy[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int cvz[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};

template <class R> inline R wq(int q) {
  // src/tgv3d_uniform.cu:31-33: 1/3, 1/18 (x6), 1/36 (x12), each rounded in the working type.
  return q == 0 ? R(1) / R(3) : (q < 7 ? R(1) / R(18) : R(1) / R(36));
}

// ---- storage forms ------------------------------------------------------------------------------
// IEEE binary16 round-to-nearest-even from binary32, subnormals kept (what numpy.float16 does).
inline uint16_t f32_to_f16(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  uint32_t sign = (u >> 16) & 0x8000u;
  uint32_t ax = u & 0x7fffffffu;
  if (ax >= 0x7f800000u) return uint16_t(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
  if (ax >= 0x477ff000u) return uint16_t(sign | 0x7c00u);  // rounds to >= 65520: overflow to inf
  if (ax < 0x38800000u) {                                    // below the smallest normal half 2^-14
    if (ax < 0x33000000u) return uint16_t(sign);             // below half the smallest subnormal
    uint32_t e = ax >> 23, m = (ax & 0x7fffffu) | 0x800000u;
    int shift = 126 - int(e);  // value = m * 2^(e-150); half subnormal unit is 2^-24
    uint32_t r = m >> shift, rem = m & ((1u << shift) - 1u), halfway = 1u << (shift - 1);
    if (rem > halfway || (rem == halfway && (r & 1u))) ++r;
    return uint16_t(sign | r);
  }
  uint32_t e = (ax >> 23) - 127 + 15, m = ax & 0x7fffffu;
  uint32_t r = (e << 10) | (m >> 13), rem = m & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) ++r;
  return uint16_t(sign | r);
}
inline float f16_to_f32(uint16_t h) {
  uint32_t sign = uint32_t(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu, u;
  if (e == 0) {
    if (m == 0) u = sign;
    else {  // subnormal: m * 2^-24, exact in binary32
      float v = float(m) * 5.9604644775390625e-08f;
      std::memcpy(&u, &v, 4);
      u |= sign;
    }
  } else if (e == 31) u = sign | 0x7f800000u | (m << 13);
  else u = sign | ((e - 15 + 127) << 23) | (m << 13);
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
// bfloat16 round-to-nearest-even (NaN kept quiet).
inline uint16_t f32_to_bf16(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
inline float bf16_to_f32(uint16_t h) {
  uint32_t u = uint32_t(h) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// A storage form: raw element type, compute type, decode / encode of one population of direction q
// (load / store). The collision works in the form's own representation (in / out; from_full maps a
// population, e.g. a ghost value, to it): populations f for the plain forms, deviations d = f - w_q for
// the shifted FP16/BF16 forms (shift = true).
struct SF32 {
  using raw = float;
  using real = float;
  static constexpr bool shift = false;
  static float load(raw r, int, float) { return r; }
  static raw store(float f, int, float) { return f; }
  static float in(raw r, int, float) { return r; }
  static raw out(float f, int, float) { return f; }
  static float from_full(float f, int) { return f; }
};
struct SF64 {
  using raw = double;
  using real = double;
  static constexpr bool shift = false;
  static double load(raw r, int, float) { return r; }
  static raw store(double f, int, float) { return f; }
  static double in(raw r, int, float) { return r; }
  static raw out(double f, int, float) { return f; }
  static double from_full(double f, int) { return f; }
};
// FP16 of (f - w_q) * scale; scale is a power of two, inv its exact inverse.
struct SF16 {
  using raw = uint16_t;
  using real = float;
  static constexpr bool shift = true;
  static float load(raw r, int q, float inv) { return f16_to_f32(r) * inv + wq<float>(q); }
  static raw store(float f, int q, float scale) { return f32_to_f16((f - wq<float>(q)) * scale); }
  static float in(raw r, int, float inv) { return f16_to_f32(r) * inv; }
  static raw out(float d, int, float scale) { return f32_to_f16(d * scale); }
  static float from_full(float f, int q) { return f - wq<float>(q); }
};
struct SBF16 {
  using raw = uint16_t;
  using real = float;
  static constexpr bool shift = true;
  static float load(raw r, int q, float inv) { return bf16_to_f32(r) * inv + wq<float>(q); }
  static raw store(float f, int q, float scale) { return f32_to_bf16((f - wq<float>(q)) * scale); }
  static float in(raw r, int, float inv) { return bf16_to_f32(r) * inv; }
  static raw out(float d, int, float scale) { return f32_to_bf16(d * scale); }
  static float from_full(float f, int q) { return f - wq<float>(q); }
};

// Scaling of a stored level: enc multiplies before rounding, dec multiplies after widening.
struct Scale {
  float enc = 1.f, dec = 1.f;
};

// ---- lattice core (src/tgv3d_uniform.cu:35-38; src/tgv3d_amr.cu:52-67) -----------------------------
template <class R> inline R feq_q(int q, R rho, R ux, R uy, R uz) {
  R cu = R(3) * (R(cvx[q]) * ux + R(cvy[q]) * uy + R(cvz[q]) * uz);
  return wq<R>(q) * rho * (R(1) + cu + R(0.5) * cu * cu - R(1.5) * (ux * ux + uy * uy + uz * uz));
}
template <class R> inline void regularize(R* n) {
  R Pxx = 0, Pyy = 0, Pzz = 0, Pxy = 0, Pxz = 0, Pyz = 0;
  for (int q = 0; q < Q; ++q) {
    R v = n[q];
    Pxx += R(cvx[q] * cvx[q]) * v;
    Pyy += R(cvy[q] * cvy[q]) * v;
    Pzz += R(cvz[q] * cvz[q]) * v;
    Pxy += R(cvx[q] * cvy[q]) * v;
    Pxz += R(cvx[q] * cvz[q]) * v;
    Pyz += R(cvy[q] * cvz[q]) * v;
  }
  for (int q = 0; q < Q; ++q) {
    R Qxx = R(cvx[q] * cvx[q]) - R(1) / R(3), Qyy = R(cvy[q] * cvy[q]) - R(1) / R(3),
      Qzz = R(cvz[q] * cvz[q]) - R(1) / R(3);
    R Qxy = R(cvx[q] * cvy[q]), Qxz = R(cvx[q] * cvz[q]), Qyz = R(cvy[q] * cvz[q]);
    n[q] = R(4.5) * wq<R>(q) *
           (Qxx * Pxx + Qyy * Pyy + Qzz * Pzz + R(2) * (Qxy * Pxy + Qxz * Pxz + Qyz * Pyz));
  }
}
template <class R> inline void moments(const R* f, R& rho, R& ux, R& uy, R& uz) {
  rho = 0;
  ux = 0;
  uy = 0;
  uz = 0;
  for (int q = 0; q < Q; ++q) {
    rho += f[q];
    ux += R(cvx[q]) * f[q];
    uy += R(cvy[q]) * f[q];
    uz += R(cvz[q]) * f[q];
  }
}
// feq_q - w_q = w_q ((rho - 1) + rho (cu + cu^2/2 - 1.5 u^2)), the equilibrium of the shifted forms.
template <class R> inline R deq_q(int q, R drho, R rho, R ux, R uy, R uz) {
  R cu = R(3) * (R(cvx[q]) * ux + R(cvy[q]) * uy + R(cvz[q]) * uz);
  return wq<R>(q) * (drho + rho * (cu + R(0.5) * cu * cu - R(1.5) * (ux * ux + uy * uy + uz * uz)));
}
// BGK collision of one cell in the form's working representation, in place
// (src/tgv3d_uniform.cu:55-60; src/tgv3d_amr.cu:268-273): f_q - omega (f_q - feq_q), feq_q =
// w_q rho (1 + cu + cu^2/2 - 1.5 u^2), cu = 3 c_q.u, evaluated with fused multiply-adds at fixed points:
//   u.u = fma(uz, uz, fma(uy, uy, ux ux)), t = 1.5 u.u, s_q = fma(cu/2, cu, 1 - t),
//   d_q = fma(-w_q rho, s_q + cu, f_q) (= f_q - feq_q), f_q' = fma(-omega, d_q, f_q).
// The reference's compiler contracts these multiply-adds too (nvcc's default --fmad=true); writing them
// out fixes the rounding so that the GPU and this restatement agree bit for bit, and halves the
// collision's instruction count. std::fma and the GPU's fma both round once (IEEE fusedMultiplyAdd).
// The shifted forms collide the deviations d_q = f_q - w_q (rho - 1 = sum d_q, rho u = sum c_q d_q), with
// feq_q - w_q = w_q ((rho - 1) + rho (cu + cu^2/2 - 1.5 u^2)) as fma(-w_q, fma(rho, g_q + cu, rho - 1),
// d_q), g_q = fma(cu/2, cu, -t): the same collision in exact arithmetic, but in FP32 summing w_q + d_q
// would round rho at the fixed low bits of the w_q, a per-step bias of ~1e-7 that drifts the mass.
template <class S> inline void collide(typename S::real* fin, typename S::real om) {
  using R = typename S::real;
  R rho, ux, uy, uz, drho = R(0);
  moments(fin, rho, ux, uy, uz);
  if (S::shift) {
    drho = rho;
    rho = R(1) + drho;
  }
  ux /= rho;
  uy /= rho;
  uz /= rho;
  const R t15 = R(1.5) * std::fma(uz, uz, std::fma(uy, uy, ux * ux));
  const R omt = R(1) - t15;
  for (int q = 0; q < Q; ++q) {
    const R f = fin[q];
    R d;
    if (q == 0) {
      d = S::shift ? std::fma(-wq<R>(0), std::fma(rho, -t15, drho), f) : std::fma(-(wq<R>(0) * rho), omt, f);
    } else {
      const R cu = R(3) * (R(cvx[q]) * ux + R(cvy[q]) * uy + R(cvz[q]) * uz);
      const R hc = R(0.5) * cu;
      if (S::shift)
        d = std::fma(-wq<R>(q), std::fma(rho, std::fma(hc, cu, -t15) + cu, drho), f);
      else
        d = std::fma(-(wq<R>(q) * rho), std::fma(hc, cu, omt) + cu, f);
    }
    fin[q] = std::fma(-om, d, f);
  }
}
inline int pmod(int a, int n) { return ((a % n) + n) % n; }

// ---- uniform solver (src/tgv3d_uniform.cu) --------------------------------------------------------
// TGV initial state, src/tgv3d_uniform.cu:65-77 (and src/tgv3d_amr.cu:431-440). The phase shifts are the
// ensemble extension; with zero phase the arithmetic is the reference's exactly.
template <class S>
void init_tgv(typename S::raw* f, int N, double u0, double kf, const double* phase, Scale sc) {
  using R = typename S::real;
  const int gx = N / MB, nb = gx * gx * gx;
  const double px = phase ? phase[0] : 0.0, py = phase ? phase[1] : 0.0, pz = phase ? phase[2] : 0.0;
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b) {
    const int bx = b % gx, by = (b / gx) % gx, bz = b / (gx * gx);
    for (int cell = 0; cell < MB3; ++cell) {
      const int i = cell % MB, j = (cell / MB) % MB, k = cell / (MB * MB);
      double x = (bx * MB + i + 0.5) * kf, y = (by * MB + j + 0.5) * kf, z = (bz * MB + k + 0.5) * kf;
      if (phase) {
        x += px;
        y += py;
        z += pz;
      }
      double ux = u0 * std::sin(x) * std::cos(y) * std::cos(z);
      double uy = -u0 * std::cos(x) * std::sin(y) * std::cos(z);
      double p = (u0 * u0 / 16.0) * (std::cos(2 * x) + std::cos(2 * y)) * (std::cos(2 * z) + 2.0);
      double rho = 1.0 + 3.0 * p;
      for (int q = 0; q < Q; ++q)
        f[(size_t(b) * Q + q) * MB3 + cell] =
            S::store(feq_q<R>(q, R(rho), R(ux), R(uy), R(0)), q, sc.enc);
    }
  }
}

// Fused pull stream-collide, src/tgv3d_uniform.cu:41-61.
template <class S>
void step_uniform(const typename S::raw* fa, typename S::raw* fb, int N, typename S::real omega, Scale sc) {
  using R = typename S::real;
  const int gx = N / MB, nb = gx * gx * gx;
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b) {
    const int bx = b % gx, by = (b / gx) % gx, bz = b / (gx * gx);
    for (int cell = 0; cell < MB3; ++cell) {
      const int i = cell % MB, j = (cell / MB) % MB, k = cell / (MB * MB);
      const int gi = bx * MB + i, gj = by * MB + j, gk = bz * MB + k;
      R fin[Q];
      for (int q = 0; q < Q; ++q) {
        int sgi = (gi - cvx[q] + N) % N, sgj = (gj - cvy[q] + N) % N, sgk = (gk - cvz[q] + N) % N;
        int sbx = sgi / MB, sby = sgj / MB, sbz = sgk / MB;
        int nbk = (sbz * gx + sby) * gx + sbx;
        int scl = ((sgk - sbz * MB) * MB + (sgj - sby * MB)) * MB + (sgi - sbx * MB);
        fin[q] = S::in(fa[(size_t(nbk) * Q + q) * MB3 + scl], q, sc.dec);
      }
      collide<S>(fin, omega);
      for (int q = 0; q < Q; ++q) fb[(size_t(b) * Q + q) * MB3 + cell] = S::out(fin[q], q, sc.enc);
    }
  }
}

// Macroscopic field (rho, u, v, w) on the linear (gk*N+gj)*N+gi grid, src/tgv3d_uniform.cu:80-90.
template <class S> void macro_uniform(const typename S::raw* f, int N, typename S::real* out, Scale sc) {
  using R = typename S::real;
  const int gx = N / MB, nb = gx * gx * gx;
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b) {
    const int bx = b % gx, by = (b / gx) % gx, bz = b / (gx * gx);
    for (int cell = 0; cell < MB3; ++cell) {
      const int i = cell % MB, j = (cell / MB) % MB, k = cell / (MB * MB);
      R v[Q];
      for (int q = 0; q < Q; ++q) v[q] = S::load(f[(size_t(b) * Q + q) * MB3 + cell], q, sc.dec);
      R rho, ux, uy, uz;
      moments(v, rho, ux, uy, uz);
      size_t idx = ((size_t(bz * MB + k) * N + (by * MB + j)) * N + (bx * MB + i)) * 4;
      out[idx + 0] = rho;
      out[idx + 1] = ux / rho;
      out[idx + 2] = uy / rho;
      out[idx + 3] = uz / rho;
    }
  }
}

// Host diagnostics of the uniform driver, src/tgv3d_uniform.cu:99-108 and 174-183: the central
// differences are taken in the field's own type (float for the FP32 field) and widened to double.
// Returns the means <0.5|u|^2>, <0.5|omega|^2>, <rho>. Sums are per z-plane, then over planes in order.
template <class R> void diag_uniform(const R* F, int N, double out[3]) {
  std::vector<double> pe(N), pz(N), pm(N);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < N; ++k) {
    double E = 0, ens = 0, mass = 0;
    const int kp = (k + 1) % N, km = (k - 1 + N) % N;
    for (int j = 0; j < N; ++j) {
      const int jp = (j + 1) % N, jm = (j - 1 + N) % N;
      for (int i = 0; i < N; ++i) {
        const int ip = (i + 1) % N, im = (i - 1 + N) % N;
        auto at = [&](int a, int b, int c, int comp) { return F[((size_t(c) * N + b) * N + a) * 4 + 1 + comp]; };
        size_t p = (size_t(k) * N + j) * N + i;
        mass += F[p * 4 + 0];
        double u = F[p * 4 + 1], v = F[p * 4 + 2], w = F[p * 4 + 3];
        E += 0.5 * (u * u + v * v + w * w);
        double g[3][3];
        for (int c = 0; c < 3; ++c) {
          g[c][0] = (at(ip, j, k, c) - at(im, j, k, c)) * 0.5;
          g[c][1] = (at(i, jp, k, c) - at(i, jm, k, c)) * 0.5;
          g[c][2] = (at(i, j, kp, c) - at(i, j, km, c)) * 0.5;
        }
        double ox = g[2][1] - g[1][2], oy = g[0][2] - g[2][0], oz = g[1][0] - g[0][1];
        ens += 0.5 * (ox * ox + oy * oy + oz * oz);
      }
    }
    pe[k] = E;
    pz[k] = ens;
    pm[k] = mass;
  }
  double E = 0, ens = 0, mass = 0;
  for (int k = 0; k < N; ++k) {
    E += pe[k];
    ens += pz[k];
    mass += pm[k];
  }
  const double inv = 1.0 / (double(N) * N * N);
  out[0] = E * inv;
  out[1] = ens * inv;
  out[2] = mass * inv;
}

// Output field of a frame: Q-criterion, speed and velocity from an N^3 field (rho, u, v, w). The uniform
// host writer (src/tgv3d_uniform.cu:96-119) differences two float field values in float before widening
// (diff_in_field); the AMR device kernel k_qspeed (src/tgv3d_amr.cu:553-564) widens first.
template <class R>
void qspeed(const R* F, int N, bool diff_in_field, float* Qo, float* spd, float* vel) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < N; ++k)
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i) {
        const int ip = (i + 1) % N, im = (i - 1 + N) % N, jp = (j + 1) % N, jm = (j - 1 + N) % N, kp = (k + 1) % N,
                  km = (k - 1 + N) % N;
        auto at = [&](int a, int b, int c, int comp) -> R { return F[((size_t(c) * N + b) * N + a) * 4 + 1 + comp]; };
        auto d = [&](R a, R b) -> double { return diff_in_field ? double(R(a - b)) * 0.5 : (double(a) - double(b)) * 0.5; };
        double g[3][3];
        for (int c = 0; c < 3; ++c) {
          g[c][0] = d(at(ip, j, k, c), at(im, j, k, c));
          g[c][1] = d(at(i, jp, k, c), at(i, jm, k, c));
          g[c][2] = d(at(i, j, kp, c), at(i, j, km, c));
        }
        double Q = 0;
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) Q -= 0.5 * g[a][b] * g[b][a];
        const size_t p = (size_t(k) * N + j) * N + i;
        const double u = F[p * 4 + 1], v = F[p * 4 + 2], w = F[p * 4 + 3];
        Qo[p] = float(Q);
        spd[p] = float(std::sqrt(u * u + v * v + w * w));
        if (vel) {
          vel[p * 3 + 0] = float(u);
          vel[p * 3 + 1] = float(v);
          vel[p * 3 + 2] = float(w);
        }
      }
}

// AMR diagnostics reduction, src/tgv3d_amr.cu:538-549: here the differences are taken in double.
template <class R> void diag_composite(const R* F, int N, double out[3]) {
  std::vector<double> pe(N), pz(N), pm(N);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < N; ++k) {
    double E = 0, ens = 0, mass = 0;
    const int kp = (k + 1) % N, km = (k - 1 + N) % N;
    for (int j = 0; j < N; ++j) {
      const int jp = (j + 1) % N, jm = (j - 1 + N) % N;
      for (int i = 0; i < N; ++i) {
        const int ip = (i + 1) % N, im = (i - 1 + N) % N;
        auto V = [&](int a, int b, int c, int comp) -> double {
          return F[((size_t(c) * N + b) * N + a) * 4 + 1 + comp];
        };
        size_t p = (size_t(k) * N + j) * N + i;
        double u = F[p * 4 + 1], v = F[p * 4 + 2], w = F[p * 4 + 3], rho = F[p * 4 + 0];
        double dudy = (V(i, jp, k, 0) - V(i, jm, k, 0)) * 0.5, dudz = (V(i, j, kp, 0) - V(i, j, km, 0)) * 0.5;
        double dvdx = (V(ip, j, k, 1) - V(im, j, k, 1)) * 0.5, dvdz = (V(i, j, kp, 1) - V(i, j, km, 1)) * 0.5;
        double dwdx = (V(ip, j, k, 2) - V(im, j, k, 2)) * 0.5, dwdy = (V(i, jp, k, 2) - V(i, jm, k, 2)) * 0.5;
        double ox = dwdy - dvdz, oy = dudz - dwdx, oz = dvdx - dudy;
        E += 0.5 * (u * u + v * v + w * w);
        ens += 0.5 * (ox * ox + oy * oy + oz * oz);
        mass += rho;
      }
    }
    pe[k] = E;
    pz[k] = ens;
    pm[k] = mass;
  }
  double E = 0, ens = 0, mass = 0;
  for (int k = 0; k < N; ++k) {
    E += pe[k];
    ens += pz[k];
    mass += pm[k];
  }
  const double inv = 1.0 / (double(N) * N * N);
  out[0] = E * inv;
  out[1] = ens * inv;
  out[2] = mass * inv;
}

}  // namespace orc
