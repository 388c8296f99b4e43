// D3Q19 lattice core for gfx950 device code: velocity set, weights, equilibrium, Hermite
// regularisation and the population storage forms. Everything direction-dependent is resolved at
// compile time (each q is a template constant), so zero velocity components cost no instructions.
//
// Arithmetic contract: kernels are compiled with -ffp-contract=off and IEEE division/sqrt, and every
// expression keeps the operation order of the reference method (src/tgv3d_uniform.cu:35-38, 55-60;
// src/tgv3d_amr.cu:52-67). Dropped terms are only exact zeros (c = 0 components), which can change the
// sign of a zero intermediate but never a non-zero result.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

namespace jp {

constexpr int Q = 19, MB = 8, MB3 = 512, HALF = 4;

// q order {0; +x,-x,+y,-y,+z,-z; xy(4); xz(4); yz(4)} as in src/tgv3d_uniform.cu:28-30.
__host__ __device__ constexpr int CX(int q) {
  constexpr int t[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  return t[q];
}
__host__ __device__ constexpr int CY(int q) {
  constexpr int t[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
  return t[q];
}
__host__ __device__ constexpr int CZ(int q) {
  constexpr int t[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
  return t[q];
}
template <class R> __host__ __device__ constexpr R W(int q) {
  return q == 0 ? R(1) / R(3) : (q < 7 ? R(1) / R(18) : R(1) / R(36));
}

// Compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1.
template <class F, int... I> __host__ __device__ inline void sfor_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F> __host__ __device__ inline void sfor(F&& f) {
  sfor_impl(f, std::make_integer_sequence<int, N>{});
}

// c_q . u summed over the non-zero components in x, y, z order.
template <int q, class R> __device__ inline R cdot(R ux, R uy, R uz) {
  constexpr int cx = CX(q), cy = CY(q), cz = CZ(q);
  R acc = R(0);
  if constexpr (cx != 0) acc = cx > 0 ? ux : -ux;
  if constexpr (cy != 0) {
    R t = cy > 0 ? uy : -uy;
    if constexpr (cx != 0) acc = acc + t;
    else acc = t;
  }
  if constexpr (cz != 0) {
    R t = cz > 0 ? uz : -uz;
    if constexpr (cx != 0 || cy != 0) acc = acc + t;
    else acc = t;
  }
  return acc;
}

// Moments rho, rho*u accumulated in q order (src/tgv3d_uniform.cu:55-57).
template <class R> struct Mom {
  R rho, ux, uy, uz;
};
template <class R> __device__ inline Mom<R> moments(const R* f) {
  Mom<R> m;
  m.rho = f[0];
  m.ux = R(0);
  m.uy = R(0);
  m.uz = R(0);
  // the first non-zero component of each axis is q = 1 (+x), 3 (+y), 5 (+z)
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if constexpr (q > 0) m.rho = m.rho + f[q];
    if constexpr (CX(q) != 0) {
      if constexpr (q == 1) m.ux = f[q];
      else m.ux = CX(q) > 0 ? m.ux + f[q] : m.ux - f[q];
    }
    if constexpr (CY(q) != 0) {
      if constexpr (q == 3) m.uy = f[q];
      else m.uy = CY(q) > 0 ? m.uy + f[q] : m.uy - f[q];
    }
    if constexpr (CZ(q) != 0) {
      if constexpr (q == 5) m.uz = f[q];
      else m.uz = CZ(q) > 0 ? m.uz + f[q] : m.uz - f[q];
    }
  });
  return m;
}

// Equilibrium factors shared by all q of one cell: w_q*rho per weight class and 1.5|u|^2.
template <class R> struct EqPre {
  R wr[3], t15, ux, uy, uz;
};
template <class R> __device__ inline EqPre<R> eq_pre(R rho, R ux, R uy, R uz) {
  EqPre<R> e;
  e.wr[0] = W<R>(0) * rho;
  e.wr[1] = W<R>(1) * rho;
  e.wr[2] = W<R>(7) * rho;
  e.t15 = R(1.5) * (ux * ux + uy * uy + uz * uz);
  e.ux = ux;
  e.uy = uy;
  e.uz = uz;
  return e;
}
// w_q rho (1 + cu + cu^2/2 - 1.5 u^2), cu = 3 c.u (src/tgv3d_uniform.cu:35-38).
template <int q, class R> __device__ inline R feq(const EqPre<R>& e) {
  constexpr int cls = q == 0 ? 0 : (q < 7 ? 1 : 2);
  if constexpr (q == 0) {
    return e.wr[0] * (R(1) - e.t15);
  } else {
    R cu = R(3) * cdot<q>(e.ux, e.uy, e.uz);
    return e.wr[cls] * (R(1) + cu + R(0.5) * cu * cu - e.t15);
  }
}

// Equilibrium minus the weight, w_q ((rho - 1) + rho (cu + cu^2/2 - 1.5 u^2)), for the shifted forms.
template <int q, class R> __device__ inline R deq(const EqPre<R>& e, R drho, R rho) {
  if constexpr (q == 0) {
    return W<R>(0) * (drho + rho * (R(0) - e.t15));
  } else {
    R cu = R(3) * cdot<q>(e.ux, e.uy, e.uz);
    return W<R>(q) * (drho + rho * (cu + R(0.5) * cu * cu - e.t15));
  }
}

// Second-order Hermite projection of a non-equilibrium vector (src/tgv3d_amr.cu:56-67). Terms whose
// coefficient c_a c_b is zero are dropped at compile time: the accumulators start at +0 and can never
// become -0, so adding the reference's 0 * v terms leaves them unchanged; of the off-diagonal products
// each direction has at most one non-zero, and the others only add zeros, whose sign difference is lost
// as soon as the projected value is added to a (non-zero) equilibrium by every caller.
template <class R> __device__ inline void regularize(R* n) {
  R Pxx = R(0), Pyy = R(0), Pzz = R(0), Pxy = R(0), Pxz = R(0), Pyz = R(0);
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int cx = CX(q), cy = CY(q), cz = CZ(q);
    const R v = n[q];
    if constexpr (cx != 0) Pxx = Pxx + v;
    if constexpr (cy != 0) Pyy = Pyy + v;
    if constexpr (cz != 0) Pzz = Pzz + v;
    if constexpr (cx * cy != 0) Pxy = cx * cy > 0 ? Pxy + v : Pxy - v;
    if constexpr (cx * cz != 0) Pxz = cx * cz > 0 ? Pxz + v : Pxz - v;
    if constexpr (cy * cz != 0) Pyz = cy * cz > 0 ? Pyz + v : Pyz - v;
  });
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int cx = CX(q), cy = CY(q), cz = CZ(q);
    const R third = R(1) / R(3);
    const R Qxx = R(cx * cx) - third, Qyy = R(cy * cy) - third, Qzz = R(cz * cz) - third;
    R d = Qxx * Pxx + Qyy * Pyy + Qzz * Pzz;
    if constexpr (cx * cy != 0) d = d + R(2) * (cx * cy > 0 ? Pxy : -Pxy);
    else if constexpr (cx * cz != 0) d = d + R(2) * (cx * cz > 0 ? Pxz : -Pxz);
    else if constexpr (cy * cz != 0) d = d + R(2) * (cy * cz > 0 ? Pyz : -Pyz);
    n[q] = R(4.5) * W<R>(q) * d;
  });
}

// ---- storage forms ----
// A form maps the stored element of direction q to the working type and back: load / store convert
// populations f. FP16/BF16 hold the shifted value (f - w_q) * enc with enc a power of two (dec = 1/enc
// exactly), and collide in that shifted representation (kShift): in / out convert the collision's
// working value. The working value is f for the plain forms and the stored deviation itself,
// D = (f - w_q) * enc, for the shifted ones: the collision takes the scale into its constants (Collider),
// so neither the decode nor the encode multiplies. Ghost records hold working values too (ghost_value).
struct FormF32 {
  using raw = float;
  using real = float;
  static constexpr int id = 0;
  static constexpr bool kShift = false;
  template <int q> __device__ static inline real load(raw r, float) { return r; }
  template <int q> __device__ static inline raw store(real f, float) { return f; }
  template <int q> __device__ static inline real in(raw r, float) { return r; }
  template <int q> __device__ static inline raw out(real f, float) { return f; }
};
struct FormF64 {
  using raw = double;
  using real = double;
  static constexpr int id = 3;
  static constexpr bool kShift = false;
  template <int q> __device__ static inline real load(raw r, float) { return r; }
  template <int q> __device__ static inline raw store(real f, float) { return f; }
  template <int q> __device__ static inline real in(raw r, float) { return r; }
  template <int q> __device__ static inline raw out(real f, float) { return f; }
};
struct FormF16 {
  using raw = _Float16;
  using real = float;
  static constexpr int id = 1;
  static constexpr bool kShift = true;
  template <int q> __device__ static inline real load(raw r, float dec) { return float(r) * dec + W<float>(q); }
  template <int q> __device__ static inline raw store(real f, float enc) {
    return static_cast<_Float16>((f - W<float>(q)) * enc);
  }
  template <int q> __device__ static inline real in(raw r, float) { return float(r); }
  template <int q> __device__ static inline raw out(real d, float) {
    // the empty asm keeps the conversion a separate rounding: fused into the multiply-add that produced d
    // (v_fma_mixlo_f16) the value would be rounded once, from the exact result, instead of twice
    asm volatile("" : "+v"(d));
    return static_cast<_Float16>(d);
  }
};
// binary32 -> bfloat16, round to nearest even: one v_cvt_pk_bf16_f32 on gfx950 (the integer form, add
// 0x7fff plus the kept lsb and truncate, took five instructions per value); NaN stays NaN.
__device__ inline uint16_t bf16_rne(float x) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(x)); }
struct FormBF16 {
  using raw = uint16_t;
  using real = float;
  static constexpr int id = 2;
  static constexpr bool kShift = true;
  template <int q> __device__ static inline real load(raw r, float dec) {
    return __uint_as_float(uint32_t(r) << 16) * dec + W<float>(q);
  }
  template <int q> __device__ static inline raw store(real f, float enc) { return bf16_rne((f - W<float>(q)) * enc); }
  template <int q> __device__ static inline real in(raw r, float) { return __uint_as_float(uint32_t(r) << 16); }
  template <int q> __device__ static inline raw out(real d, float) { return bf16_rne(d); }
};

// Fused multiply-add in the working type (one rounding, IEEE fusedMultiplyAdd).
template <class R> __device__ __forceinline__ R fmx(R a, R b, R c) {
  if constexpr (sizeof(R) == 4) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}

// The three velocity quotients u = j / rho of a cell, each the correctly rounded quotient (IEEE division),
// from one reciprocal instead of three division sequences (10 instructions each, 2 of them serialised on
// VCC). Binary32: Y0 = v_rcp_f32(b) (relative error at most 2^-23). Each Newton step in binary64,
// Y <- fma(Y, fma(-b, Y, 1), Y), squares the relative error and adds one rounding (2^-53), so two steps
// from any start within 2^-13 leave |Y b - 1| <= 1.5 2^-52 (from v_rcp_f32's start, 2^-53 (1 + 2^-38)), and
// Q = RN64(a Y) lies within 2^-51 |a / b| of a / b. A float
// quotient a / b never lies within 2^-49 |a / b| of a binary32 rounding boundary (a midpoint m 2^e with m
// odd, 2^24 <= m < 2^25): with a = A 2^s, b = B 2^t, |A|, |B| < 2^24, the distance is
// |A 2^s - B m 2^(t+e)| / |b|, whose numerator is a non-zero multiple of 2^min(s, t + e) (equality would
// need odd(A) = odd(B) m >= 2^24), i.e. at least 2^e / |B| > |m 2^e| 2^-49 or |a / b| 2^-24. So RN32(Q) =
// RN32(a / b) bit for bit, zero signs included (a = +-0 gives +-0), for every finite a and normal b > 0
// whose quotient is a normal float (rho is near 1 and |j| < 1 here). Binary64 keeps IEEE division.
template <class R> __device__ __forceinline__ void quot3(R b, R& x, R& y, R& z) {
  if constexpr (sizeof(R) == 4) {
    const double B = double(b);
    double Y = double(__builtin_amdgcn_rcpf(b));
    double E = __builtin_fma(-B, Y, 1.0);
    Y = __builtin_fma(Y, E, Y);
    E = __builtin_fma(-B, Y, 1.0);
    Y = __builtin_fma(Y, E, Y);
    x = float(double(x) * Y);
    y = float(double(y) * Y);
    z = float(double(z) * Y);
  } else {
    x = x / b;
    y = y / b;
    z = z / b;
  }
}

// BGK collision of one cell in the form's working representation (src/tgv3d_uniform.cu:55-60,
// src/tgv3d_amr.cu:268-273): post<q>(x) = x - omega (x - feq_q), feq_q = w_q rho (1 + cu + cu^2/2 - 1.5 u^2),
// cu = 3 c_q.u, with the multiply-adds fused at fixed points (the reference's compiler contracts them too,
// nvcc's default --fmad=true; the oracle writes the same fused operations, so the two agree bit for bit):
//   u.u = fma(uz, uz, fma(uy, uy, ux ux)), t = 1.5 u.u, s_q = fma(cu/2, cu, 1 - t),
//   post<q>(x) = fma(-omega, fma(-w_q rho, s_q + cu, x), x).
// The shifted forms collide the deviations d_q = f_q - w_q: rho - 1 = sum d_q, rho u = sum c_q d_q, and
// feq_q - w_q = w_q ((rho - 1) + rho (cu + cu^2/2 - 1.5 u^2)) enters as fma(-w_q, fma(rho, g_q + cu,
// rho - 1), d_q) with g_q = fma(cu/2, cu, -t). In exact arithmetic this is the same collision; in FP32
// every operand is then of the order of the deviations, whereas summing w_q + d_q rounds rho at the fixed
// low bits of the w_q, a bias of ~1e-7 per cell and step that would show as a linear drift of the mass.
// They work on the stored scale, D_q = enc d_q: the sums of D_q times dec give rho - 1 and rho u, and the
// relaxation reads fma(-om, fma(-(w_q enc), fma(rho, g_q + cu, rho - 1), D_q), D_q). Multiplying by a power
// of two is exact and commutes with rounding while values stay in the normal range (stored deviations are
// at least 2^-24 dec and, like w_q enc, far from the range's ends), so every partial sum and both
// multiply-adds equal enc times those of the collision of the d_q, bit for bit, and the result is the
// value the encode of that collision would store.
template <class S> struct Collider {
  using R = typename S::real;
  R ux, uy, uz, t15, omt, wr[3], drho, rho;
  // fin in the form's working representation; dec / enc: the shifted forms' scale (unused otherwise)
  __device__ Collider(const R* fin, float dec, float enc) {
    const Mom<R> m = moments(fin);
    R jx = m.ux, jy = m.uy, jz = m.uz;
    if constexpr (S::kShift) {
      drho = m.rho * R(dec);
      rho = R(1) + drho;
      jx = jx * R(dec);
      jy = jy * R(dec);
      jz = jz * R(dec);
      wr[0] = -W<R>(0) * R(enc);
      wr[1] = -W<R>(1) * R(enc);
      wr[2] = -W<R>(7) * R(enc);
    } else {
      drho = R(0);
      rho = m.rho;
      wr[0] = W<R>(0) * rho;
      wr[1] = W<R>(1) * rho;
      wr[2] = W<R>(7) * rho;
    }
    ux = jx;
    uy = jy;
    uz = jz;
    quot3(rho, ux, uy, uz);
    t15 = R(1.5) * fmx<R>(uz, uz, fmx<R>(uy, uy, ux * ux));
    omt = R(1) - t15;
  }
  // the relaxed value of direction q from sc = s_q + cu (plain forms) or g_q + cu (shifted forms); wr holds
  // w rho (plain) or -w enc (shifted) per weight class
  template <int q> __device__ R relax(R x, R om, R sc) const {
    constexpr int cls = q == 0 ? 0 : (q < 7 ? 1 : 2);
    if constexpr (S::kShift) return fmx<R>(-om, fmx<R>(wr[cls], fmx<R>(rho, sc, drho), x), x);
    else return fmx<R>(-om, fmx<R>(-wr[cls], sc, x), x);
  }
  template <int q> __device__ R post(R x, R om) const {
    if constexpr (q == 0) {
      return relax<0>(x, om, S::kShift ? -t15 : omt);
    } else {
      const R cu = R(3) * cdot<q>(ux, uy, uz);
      const R hc = R(0.5) * cu;
      return relax<q>(x, om, fmx<R>(hc, cu, S::kShift ? -t15 : omt) + cu);
    }
  }
  // post<q> of all 19 directions in place. The opposite directions q and q + 1 (q odd) share cu/2 and
  // s_q (g_q): c_{q+1} = -c_q makes cu change sign exactly, (-a)(-b) = ab in IEEE arithmetic, and
  // s_q - cu = s_q + (-cu), so each value equals post<q>'s bit for bit, with two fewer operations per pair.
  __device__ void post_all(R* f, R om) const {
    f[0] = post<0>(f[0], om);
    sfor<9>([&](auto pc) {
      constexpr int q = 1 + 2 * decltype(pc)::value;
      const R cu = R(3) * cdot<q>(ux, uy, uz);
      const R hc = R(0.5) * cu;
      const R s = fmx<R>(hc, cu, S::kShift ? -t15 : omt);
      f[q] = relax<q>(f[q], om, s + cu);
      f[q + 1] = relax<q + 1>(f[q + 1], om, s - cu);
    });
  }
};

// A population f of direction q (run-time) in the form's working representation, for ghost records.
template <class S> __device__ inline typename S::real ghost_value(typename S::real f, int q, float enc) {
  if constexpr (S::kShift) return (f - (q == 0 ? W<float>(0) : (q < 7 ? W<float>(1) : W<float>(7)))) * enc;
  else return f;
}

// Runtime-q decode for kernels that gather one direction at a time.
template <class S> __device__ inline typename S::real load_rt(typename S::raw r, int q, float dec) {
  if constexpr (S::id == 1) return float(r) * dec + (q == 0 ? W<float>(0) : (q < 7 ? W<float>(1) : W<float>(7)));
  else if constexpr (S::id == 2)
    return __uint_as_float(uint32_t(r) << 16) * dec + (q == 0 ? W<float>(0) : (q < 7 ? W<float>(1) : W<float>(7)));
  else return r;
}

// Unsigned division by a run-time constant: q = (n * m) >> s, exact for n < 2^31 and d < 2^16.
struct FastDiv {
  uint64_t m;
  uint32_t s, d;
};
inline FastDiv make_fastdiv(uint32_t d) {
  uint32_t s = 0;
  while ((1u << s) < d) ++s;
  FastDiv f;
  f.d = d;
  f.s = 32 + s;
  f.m = ((uint64_t(1) << f.s) / d) + 1;
  return f;
}
__device__ inline uint32_t fdiv(uint32_t n, const FastDiv& f) { return uint32_t((uint64_t(n) * f.m) >> f.s); }

}  // namespace jp
