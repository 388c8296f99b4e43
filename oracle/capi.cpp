// TEST INFRASTRUCTURE ONLY (see oracle.hpp). C entry points of the CPU oracle for ctypes: a uniform
// context (src/tgv3d_uniform.cu:142-195) and an AMR context (src/tgv3d_amr.cu:643-837) over the four
// storage forms. Buffers crossing this boundary use the reference's tile-major layout.
#include <omp.h>

#include <chrono>
#include <cstdio>
#include <memory>

#include "amr.hpp"

using namespace orc;

namespace {

struct UniBase {
  virtual ~UniBase() = default;
  virtual void init(const double* phase) = 0;
  virtual void step(long n) = 0;
  virtual void diag(double* out) = 0;
  virtual void qspeed_out(float* Q, float* spd, float* vel) = 0;
  virtual size_t elem() const = 0;
  virtual void get_raw(void* dst) const = 0;
  virtual void set_raw(const void* src) = 0;
  virtual void get_f64(double* dst) const = 0;
  virtual void set_omega(double w) = 0;
  int N = 0;
  double re = 0, u0 = 0, nu = 0, tau = 0, kf = 0;
  Scale sc;
};

template <class S> struct Uni : UniBase {
  using raw = typename S::raw;
  using R = typename S::real;
  std::vector<raw> fa, fb;
  R omega;
  Uni(int n, double re_, double u0_, Scale s) {
    N = n;
    re = re_;
    u0 = u0_;
    sc = s;
    // src/tgv3d_uniform.cu:151-152
    nu = u0 * double(N) / (2.0 * M_PI * re);
    tau = 3.0 * nu + 0.5;
    omega = R(1.0 / tau);
    kf = 2.0 * M_PI / double(N);
    fa.assign(size_t(N) * N * N * Q, raw(0));
    fb = fa;
  }
  void init(const double* phase) override { init_tgv<S>(fa.data(), N, u0, kf, phase, sc); }
  void step(long n) override {
    for (long s = 0; s < n; ++s) {
      step_uniform<S>(fa.data(), fb.data(), N, omega, sc);
      std::swap(fa, fb);
    }
  }
  void diag(double* out) override {
    std::vector<R> F(size_t(N) * N * N * 4);
    macro_uniform<S>(fa.data(), N, F.data(), sc);
    diag_uniform(F.data(), N, out);
  }
  void qspeed_out(float* Q, float* spd, float* vel) override {  // src/tgv3d_uniform.cu:170, 110-119
    std::vector<R> F(size_t(N) * N * N * 4);
    macro_uniform<S>(fa.data(), N, F.data(), sc);
    qspeed(F.data(), N, true, Q, spd, vel);
  }
  void set_omega(double w) override { omega = R(w); }
  size_t elem() const override { return sizeof(raw); }
  void get_raw(void* dst) const override { std::memcpy(dst, fa.data(), fa.size() * sizeof(raw)); }
  void set_raw(const void* src) override { std::memcpy(fa.data(), src, fa.size() * sizeof(raw)); }
  void get_f64(double* dst) const override {
    const size_t nb = fa.size() / (size_t(Q) * MB3);
    for (size_t b = 0; b < nb; ++b)
      for (int q = 0; q < Q; ++q)
        for (int c = 0; c < MB3; ++c) {
          size_t i = (b * Q + q) * MB3 + c;
          dst[i] = double(S::load(fa[i], q, sc.dec));
        }
  }
};

struct AmrBase {
  virtual ~AmrBase() = default;
  virtual int init(const float* sen0, const float* sen1pos) = 0;
  virtual int regrid(const float* sen0, const float* sen1pos) = 0;
  virtual void step(long n) = 0;
  virtual void diag(double* out) = 0;
  virtual void qspeed_out(float* Q, float* spd) = 0;
  virtual void counts(long* out) const = 0;
  virtual int table(int which, void* dst, long cap) const = 0;
  virtual long level_elems(int level) const = 0;
  virtual size_t elem() const = 0;
  virtual void get_raw(int level, void* dst) const = 0;
  virtual void set_raw(int level, const void* src) = 0;
  virtual void get_f64(int level, double* dst) const = 0;
  virtual void params(double* out) const = 0;
  virtual void run(int op) = 0;
  virtual void set_omega(int level, double w) = 0;
  virtual long ghost_elems(int level) const = 0;
  virtual void get_ghost(int level, double* dst) const = 0;
  virtual void set_ghost(int level, const double* src) = 0;
};

template <class S> struct AmrO : AmrBase {
  Amr<S> a;
  explicit AmrO(const AmrParams& p) : a(p) {}
  int init(const float* s0, const float* s1) override {
    a.init_tgv_l0();
    a.initial_build(s0, s1);
    return a.overflow ? 3 : 0;
  }
  int regrid(const float* s0, const float* s1) override {
    a.regrid(s0, s1);
    return a.overflow ? 3 : 0;
  }
  void step(long n) override {
    for (long s = 0; s < n; ++s) a.base_step();
  }
  void diag(double* out) override { a.diag(out); }
  void qspeed_out(float* Q, float* spd) override {  // src/tgv3d_amr.cu:871, 890
    std::vector<typename S::real> F;
    a.composite(F);
    qspeed(F.data(), a.N2, false, Q, spd, nullptr);
  }
  void counts(long* o) const override {
    long v[] = {a.cnb0, a.nref0, a.nf1, a.nref1, a.nf2, a.n_skin0, a.n_skin1, a.cap_nref0, a.cap1,
                a.cap_nref1 == INT_MAX ? -1 : a.cap_nref1, a.cap2, a.hw_nf1, a.hw_nf2, a.overflow ? 1 : 0};
    for (int i = 0; i < 14; ++i) o[i] = v[i];
  }
  // table ids match include/jp.h (JP_TABLE_*)
  int table(int which, void* dst, long cap) const override {
    const std::vector<int>* t = nullptr;
    long n = 0;
    switch (which) {
      case 0: t = &a.ref0; n = a.cnb0; break;
      case 1: t = &a.fof0; n = a.cnb0; break;
      case 2: t = &a.rl0; n = a.nref0; break;
      case 3: t = &a.l1pcb; n = a.nf1; break;
      case 4: t = &a.l1oct; n = a.nf1; break;
      case 5: t = &a.f1pos; n = a.nf1; break;
      case 6: t = &a.l1grid; n = a.NG1; break;
      case 7: t = &a.ref1; n = a.nf1; break;
      case 8: t = &a.fof1; n = a.nf1; break;
      case 9: t = &a.rl1; n = a.nref1; break;
      case 10: t = &a.l2pl1; n = a.nf2; break;
      case 11: t = &a.l2oct; n = a.nf2; break;
      case 12: t = &a.skin0; n = a.n_skin0; break;
      case 13: t = &a.skin1; n = a.n_skin1; break;
      case 14: {
        if (cap < a.cnb0) return -1;
        std::memcpy(dst, a.sen0.data(), size_t(a.cnb0) * 4);
        return a.cnb0;
      }
      case 15: {
        if (cap < a.nf1) return -1;
        std::memcpy(dst, a.sen1.data(), size_t(a.nf1) * 4);
        return a.nf1;
      }
      default: return -1;
    }
    if (cap < n) return -1;
    std::memcpy(dst, t->data(), size_t(n) * 4);
    return int(n);
  }
  long level_elems(int level) const override {
    long nt = level == 0 ? a.cnb0 : (level == 1 ? a.nf1 : a.nf2);
    return nt * Q * MB3;
  }
  const std::vector<typename S::raw>& lv(int level) const { return level == 0 ? a.c0a : (level == 1 ? a.f1a : a.f2a); }
  std::vector<typename S::raw>& lv(int level) { return level == 0 ? a.c0a : (level == 1 ? a.f1a : a.f2a); }
  size_t elem() const override { return sizeof(typename S::raw); }
  void get_raw(int level, void* dst) const override {
    std::memcpy(dst, lv(level).data(), size_t(level_elems(level)) * sizeof(typename S::raw));
  }
  void set_raw(int level, const void* src) override {
    std::memcpy(lv(level).data(), src, size_t(level_elems(level)) * sizeof(typename S::raw));
  }
  void get_f64(int level, double* dst) const override {
    const auto& v = lv(level);
    const float dec = a.P.sc[level].dec;
    long n = level_elems(level);
    for (long i = 0; i < n; ++i) dst[i] = double(S::load(v[i], int((i / MB3) % Q), dec));
  }
  void params(double* o) const override {
    double v[] = {a.nu, a.tau0, a.tau1, a.tau2, double(a.om0), double(a.om1), double(a.om2), a.nqs01, a.nqi10,
                  a.nqs12, a.nqi21, a.kf};
    for (int i = 0; i < 12; ++i) o[i] = v[i];
  }
  // One operation of the subcycle or the regrid on the current state, as base_step / regrid run it
  // (src/tgv3d_amr.cu:825-836, 299-308, 382-393, 442-453). The interface pins drive each one alone.
  void run(int op) override {
    switch (op) {
      case 0:  // prep over the level-0 skin + level-1 ghost fill (amr.cu:825-826)
        a.prep_list(a.c0a, a.prep0, a.skin0, a.n_skin0, a.nqs01, a.P.sc[0].dec, true);
        a.fillcf_L1p();
        break;
      case 1:  // prep over the level-1 skin + level-2 ghost fill (amr.cu:830-831)
        a.prep_list(a.f1a, a.prep1, a.skin1, a.n_skin1, a.nqs12, a.P.sc[1].dec, false);
        a.fillcf_L2p();
        break;
      case 2: a.step_L1(); break;
      case 3: a.step_L2(); break;
      case 4: a.restrict_level(a.c0a, a.f1a, a.rl0, a.fof0, a.nref0, a.nqi10, a.P.sc[0], a.P.sc[1]); break;
      case 5: a.restrict_level(a.f1a, a.f2a, a.rl1, a.fof1, a.nref1, a.nqi21, a.P.sc[1], a.P.sc[2]); break;
      case 6: a.seed_L1(); break;
      case 7: a.seed_L2(); break;
      case 8:  // sensors of both levels into the sen0 / sen1 tables
        a.sensor(a.c0a, a.cnb0, a.sen0, a.P.sc[0].dec);
        a.sensor(a.f1a, a.nf1, a.sen1, a.P.sc[1].dec);
        break;
    }
  }
  void set_omega(int level, double w) override {
    using R = typename S::real;
    (level == 0 ? a.om0 : (level == 1 ? a.om1 : a.om2)) = R(w);
  }
  // Ghost boxes in the reference's layout gh[(tile*19 + q)*1000 + (ek*10 + ej)*10 + ei] (amr.cu:179-185).
  long ghost_elems(int level) const override { return long(level == 1 ? a.nf1 : a.nf2) * Q * EB3; }
  void get_ghost(int level, double* dst) const override {
    const auto& g = level == 1 ? a.gh1 : a.gh2;
    long n = ghost_elems(level);
    for (long i = 0; i < n; ++i) dst[i] = double(g[i]);
  }
  void set_ghost(int level, const double* src) override {
    using R = typename S::real;
    auto& g = level == 1 ? a.gh1 : a.gh2;
    long n = ghost_elems(level);
    for (long i = 0; i < n; ++i) g[i] = R(src[i]);
  }
};

Scale make_scale(float enc) {
  Scale s;
  s.enc = enc;
  s.dec = 1.f / enc;
  return s;
}

}  // namespace

extern "C" {

int orc_threads() { return omp_get_max_threads(); }
void orc_set_threads(int n) { omp_set_num_threads(n); }

// ---- binary16 / bfloat16 conversions (pinned against numpy in the tests) ----
void orc_f32_to_f16(const float* x, uint16_t* y, long n) {
  for (long i = 0; i < n; ++i) y[i] = f32_to_f16(x[i]);
}
void orc_f16_to_f32(const uint16_t* x, float* y, long n) {
  for (long i = 0; i < n; ++i) y[i] = f16_to_f32(x[i]);
}
void orc_f32_to_bf16(const float* x, uint16_t* y, long n) {
  for (long i = 0; i < n; ++i) y[i] = f32_to_bf16(x[i]);
}

// ---- lattice core probes (pinned by identities in the tests) ----
void orc_feq(int q, float rho, float ux, float uy, float uz, float* out) { *out = feq_q<float>(q, rho, ux, uy, uz); }
void orc_regularize(float* n) { regularize<float>(n); }
void orc_qspeed_field(const float* F, int N, int diff_in_field, float* Q, float* spd, float* vel) {
  qspeed<float>(F, N, diff_in_field != 0, Q, spd, vel);
}
void orc_regularize64(double* n) { regularize<double>(n); }

// ---- uniform context ----
void* orc_uni_create(int prec, int N, double re, double u0, float scale) {
  Scale s = make_scale(scale);
  switch (prec) {
    case 0: return new Uni<SF32>(N, re, u0, s);
    case 1: return new Uni<SF16>(N, re, u0, s);
    case 2: return new Uni<SBF16>(N, re, u0, s);
    case 3: return new Uni<SF64>(N, re, u0, s);
  }
  return nullptr;
}
void orc_uni_destroy(void* h) { delete static_cast<UniBase*>(h); }
void orc_uni_init(void* h, const double* phase) { static_cast<UniBase*>(h)->init(phase); }
void orc_uni_step(void* h, long n) { static_cast<UniBase*>(h)->step(n); }
void orc_uni_diag(void* h, double* out) { static_cast<UniBase*>(h)->diag(out); }
void orc_uni_qspeed(void* h, float* Q, float* spd, float* vel) { static_cast<UniBase*>(h)->qspeed_out(Q, spd, vel); }
long orc_uni_elem(void* h) { return long(static_cast<UniBase*>(h)->elem()); }
void orc_uni_get_raw(void* h, void* dst) { static_cast<UniBase*>(h)->get_raw(dst); }
void orc_uni_set_raw(void* h, const void* src) { static_cast<UniBase*>(h)->set_raw(src); }
void orc_uni_get_f64(void* h, double* dst) { static_cast<UniBase*>(h)->get_f64(dst); }
// Overrides the relaxation rate (omega = 0 turns the step into a pure periodic shift).
void orc_uni_set_omega(void* h, double w) { static_cast<UniBase*>(h)->set_omega(w); }
// Wall-clock seconds of n uniform steps (the CPU baseline), after the state is initialised.
double orc_uni_time_steps(void* h, long n) {
  auto t0 = std::chrono::steady_clock::now();
  static_cast<UniBase*>(h)->step(n);
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ---- AMR context ----
// ip: n, l1_mode, l2_mode, skipprep, skiprestr, prec, develop
// dp: re, u0, frac1, frac2, budget_bytes, phase[3], scale[3]
void* orc_amr_create(const int* ip, const double* dp) {
  AmrParams p;
  p.n = ip[0];
  p.l1_mode = ip[1];
  p.l2_mode = ip[2];
  p.skipprep = ip[3];
  p.skiprestr = ip[4];
  p.develop = ip[6];
  p.re = dp[0];
  p.u0 = dp[1];
  p.frac1 = dp[2];
  p.frac2 = dp[3];
  p.budget_bytes = dp[4];
  for (int i = 0; i < 3; ++i) p.phase[i] = dp[5 + i];
  for (int l = 0; l < 3; ++l) p.sc[l] = make_scale(float(dp[8 + l]));
  switch (ip[5]) {
    case 0: return new AmrO<SF32>(p);
    case 1: return new AmrO<SF16>(p);
    case 2: return new AmrO<SBF16>(p);
    case 3: return new AmrO<SF64>(p);
  }
  return nullptr;
}
void orc_amr_destroy(void* h) { delete static_cast<AmrBase*>(h); }
int orc_amr_init(void* h, const float* sen0, const float* sen1pos) { return static_cast<AmrBase*>(h)->init(sen0, sen1pos); }
int orc_amr_regrid(void* h, const float* sen0, const float* sen1pos) {
  return static_cast<AmrBase*>(h)->regrid(sen0, sen1pos);
}
void orc_amr_step(void* h, long n) { static_cast<AmrBase*>(h)->step(n); }
void orc_amr_diag(void* h, double* out) { static_cast<AmrBase*>(h)->diag(out); }
void orc_amr_qspeed(void* h, float* Q, float* spd) { static_cast<AmrBase*>(h)->qspeed_out(Q, spd); }
void orc_amr_counts(void* h, long* out) { static_cast<AmrBase*>(h)->counts(out); }
int orc_amr_table(void* h, int which, void* dst, long cap) { return static_cast<AmrBase*>(h)->table(which, dst, cap); }
long orc_amr_level_elems(void* h, int level) { return static_cast<AmrBase*>(h)->level_elems(level); }
long orc_amr_elem(void* h) { return long(static_cast<AmrBase*>(h)->elem()); }
void orc_amr_get_raw(void* h, int level, void* dst) { static_cast<AmrBase*>(h)->get_raw(level, dst); }
void orc_amr_set_raw(void* h, int level, const void* src) { static_cast<AmrBase*>(h)->set_raw(level, src); }
void orc_amr_get_f64(void* h, int level, double* dst) { static_cast<AmrBase*>(h)->get_f64(level, dst); }
void orc_amr_params(void* h, double* out) { static_cast<AmrBase*>(h)->params(out); }
// Interface test hooks: one subcycle/regrid operation alone, a level's relaxation rate, ghost boxes.
void orc_amr_run(void* h, int op) { static_cast<AmrBase*>(h)->run(op); }
void orc_amr_set_omega(void* h, int level, double w) { static_cast<AmrBase*>(h)->set_omega(level, w); }
long orc_amr_ghost_elems(void* h, int level) { return static_cast<AmrBase*>(h)->ghost_elems(level); }
void orc_amr_get_ghost(void* h, int level, double* dst) { static_cast<AmrBase*>(h)->get_ghost(level, dst); }
void orc_amr_set_ghost(void* h, int level, const double* src) { static_cast<AmrBase*>(h)->set_ghost(level, src); }
}
