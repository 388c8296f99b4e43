// TEST INFRASTRUCTURE ONLY (see oracle.hpp). CPU restatement of the 3-level block AMR of
// src/tgv3d_amr.cu: level steps, coarse-fine interface, restriction, seeding, migration, the device
// regrid pipeline (with std::sort, std::exclusive_scan and stable std::copy_if in place of the Thrust
// calls at src/tgv3d_amr.cu:730-812), pool sizing, and the composite diagnostics.
#pragma once
#include <climits>
#include <string>

#include "oracle.hpp"

namespace orc {

enum Mode { MODE_ALL = 0, MODE_SLAB = 1, MODE_SENSOR = 2 };

struct AmrParams {
  int n = 32;
  double re = 800, u0 = 0.05;
  double phase[3] = {0, 0, 0};
  int l1_mode = MODE_ALL, l2_mode = MODE_ALL;
  double frac1 = 0.35, frac2 = 0.35;
  double budget_bytes = 1e30;  // fine-pool VRAM budget (src/tgv3d_amr.cu:684-688)
  int skipprep = 0, skiprestr = 0;
  int develop = 300;  // hidden level-0 steps before a sensor-mode initial build (src/tgv3d_amr.cu:779)
  Scale sc[3];
};

template <class S> struct Amr {
  using raw = typename S::raw;
  using R = typename S::real;
  AmrParams P;
  int N, cgx, cnb0, l1gx, N1, N2, NG1;
  double nu, tau0, tau1, tau2, kf;
  R om0, om1, om2;
  float nqs01, nqi10, nqs12, nqi21;
  int cap_nref0, cap1, cap_nref1 = INT_MAX, cap2 = 0;
  long hw_nf1 = 0, hw_nf2 = 0;
  int nref0 = 0, nf1 = 0, nref1 = 0, nf2 = 0, n_skin0 = 0, n_skin1 = 0;
  bool overflow = false;
  std::vector<raw> c0a, c0b, f1a, f1b, f2a, f2b;
  std::vector<R> prep0, prep1, gh1, gh2;  // prep/ghost hold decoded values (the reference's are FP32)
  std::vector<int> ref0, fof0, rl0, l1pcb, l1oct, f1pos, l1grid, ref1, fof1, rl1, l2pl1, l2oct, skin0, skin1;
  std::vector<int> o_ref0, o_fof0, o_l1grid, o_ref1, o_fof1, rank, deep;
  std::vector<float> sen0, sen1;

  // Level parameters, src/tgv3d_amr.cu:658-664; pool sizing, 675-691.
  explicit Amr(const AmrParams& p) : P(p) {
    N = P.n;
    cgx = N / MB;
    cnb0 = cgx * cgx * cgx;
    l1gx = 2 * cgx;
    N1 = 2 * N;
    N2 = 4 * N;
    NG1 = l1gx * l1gx * l1gx;
    nu = P.u0 * double(N) / (2.0 * M_PI * P.re);
    tau0 = 3.0 * nu + 0.5;
    tau1 = 2.0 * tau0 - 0.5;
    tau2 = 2.0 * tau1 - 0.5;
    om0 = R(1.0 / tau0);
    om1 = R(1.0 / tau1);
    om2 = R(1.0 / tau2);
    nqs01 = float((tau1 - 1.0) / (2.0 * (tau0 - 1.0)));
    nqi10 = float((2.0 * (tau0 - 1.0)) / (tau1 - 1.0));
    nqs12 = float((tau2 - 1.0) / (2.0 * (tau1 - 1.0)));
    nqi21 = float((2.0 * (tau1 - 1.0)) / (tau2 - 1.0));
    kf = 2.0 * M_PI / double(N);
    cap_nref0 = (P.l1_mode == MODE_ALL) ? cnb0
                                         : std::min(cnb0, std::max(1, int(std::ceil(P.frac1 * cnb0 * 1.4))));
    while (cap_nref0 > 1 && !pool_fits(8L * cap_nref0, 0)) cap_nref0--;
    cap1 = 8 * cap_nref0;
    size_t nc = size_t(cnb0) * Q * MB3;
    c0a.assign(nc, raw(0));
    c0b.assign(nc, raw(0));
    prep0.assign(nc, R(0));
    f1a.assign(size_t(cap1) * Q * MB3, raw(0));
    l1_pool(false);
    const int n1 = 8 * cnb0;  // level-1 tables are sized for full refinement; only the pools grow
    ref0.assign(cnb0, 0);
    fof0.assign(cnb0, -1);
    o_ref0 = ref0;
    o_fof0 = fof0;
    l1pcb.assign(n1, 0);
    l1oct.assign(n1, 0);
    rl0.assign(n1, 0);
    f1pos.assign(n1, 0);
    l1grid.assign(NG1, -1);
    o_l1grid = l1grid;
    ref1.assign(n1, 0);
    fof1.assign(n1, -1);
    o_ref1 = ref1;
    o_fof1 = fof1;
    rl1.assign(n1, 0);
    sen0.assign(cnb0, 0.f);
    sen1.assign(n1, 0.f);
    skin0.assign(cnb0, 0);
    skin1.assign(n1, 0);
    rank.assign(n1, 0);
    deep.assign(n1, 0);
  }
  void l1_pool(bool with_f1a) {
    if (with_f1a) f1a.assign(size_t(cap1) * Q * MB3, raw(0));
    f1b.assign(size_t(cap1) * Q * MB3, raw(0));
    prep1.assign(size_t(cap1) * Q * MB3, R(0));
    gh1.assign(size_t(cap1) * Q * EB3, R(0));
  }
  // Ties at the level-0 threshold can exceed the level-1 pool (the reference writes out of bounds
  // there; a symmetric TGV ties every tile at small N0): the pool grows to max(nref0, 1.5 x) tiles
  // when the budget allows (never beyond full refinement); the target cap_nref0 stays. keep_f1a: a
  // regrid still migrates from the old level-1 data.
  bool grow_L1(bool keep_f1a) {
    int cn = std::min(cnb0, std::max(nref0, (cap1 / 8) * 3 / 2));
    while (cn > nref0 && !pool_fits(8L * cn, cap2)) cn--;
    if (!pool_fits(8L * cn, cap2)) return false;
    cap1 = 8 * cn;
    l1_pool(!keep_f1a);
    return true;
  }
  bool pool_fits(long c1, long c2) const {
    const double PB = double(Q) * 4 * (2 * MB3 + EB3);
    return double(c1 + c2) * PB * 6 / 5 < P.budget_bytes;
  }

  // ---- interpolation helpers, src/tgv3d_amr.cu:70-144 ----
  // index of coarse cell (ci,cj,ck) in the dense tile-major L0 arrays
  size_t l0idx(int ci, int cj, int ck, int q) const {
    int cbx = ci / MB, cby = cj / MB, cbz = ck / MB, cb = (cbz * cgx + cby) * cgx + cbx;
    int cc = ((ck - cbz * MB) * MB + (cj - cby * MB)) * MB + (ci - cbx * MB);
    return (size_t(cb) * Q + q) * MB3 + cc;
  }
  static void tri(int g, int& base, float& t) {
    double xc = (double(g) + 0.5) / 2.0 - 0.5;
    base = int(std::floor(xc));
    t = float(xc - base);
  }
  // prolong_L0, src/tgv3d_amr.cu:70-91: per corner moments + regularised non-equilibrium.
  void prolong_L0(const std::vector<raw>& cf, int cfi, int cfj, int cfk, float nqs, R* out) const {
    int ib, jb, kb;
    float tx, ty, tz;
    tri(cfi, ib, tx);
    tri(cfj, jb, ty);
    tri(cfk, kb, tz);
    for (int q = 0; q < Q; ++q) out[q] = 0;
    for (int dk = 0; dk < 2; ++dk)
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          float wgt = (di ? tx : 1 - tx) * (dj ? ty : 1 - ty) * (dk ? tz : 1 - tz);
          int ci = pmod(ib + di, N), cj = pmod(jb + dj, N), ck = pmod(kb + dk, N);
          R fq[Q], nq[Q], fr, fx, fy, fz;
          for (int q = 0; q < Q; ++q) fq[q] = S::load(cf[l0idx(ci, cj, ck, q)], q, P.sc[0].dec);
          moments(fq, fr, fx, fy, fz);
          R uxl = fx / fr, uyl = fy / fr, uzl = fz / fr;
          for (int q = 0; q < Q; ++q) nq[q] = fq[q] - feq_q<R>(q, fr, uxl, uyl, uzl);
          regularize(nq);
          for (int q = 0; q < Q; ++q) out[q] += R(wgt) * (feq_q<R>(q, fr, uxl, uyl, uzl) + R(nqs) * nq[q]);
        }
  }
  // prolong_L1, src/tgv3d_amr.cu:94-118: scattered L1 parent via l1grid, absent corners renormalised.
  void prolong_L1(const std::vector<raw>& f1, const std::vector<int>& grid, int gfi, int gfj, int gfk,
                  float nqs, R* out) const {
    int ib, jb, kb;
    float tx, ty, tz, wsum = 0;
    tri(gfi, ib, tx);
    tri(gfj, jb, ty);
    tri(gfk, kb, tz);
    for (int q = 0; q < Q; ++q) out[q] = 0;
    for (int dk = 0; dk < 2; ++dk)
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          float wgt = (di ? tx : 1 - tx) * (dj ? ty : 1 - ty) * (dk ? tz : 1 - tz);
          int ci = pmod(ib + di, N1), cj = pmod(jb + dj, N1), ck = pmod(kb + dk, N1);
          int lbx = ci / MB, lby = cj / MB, lbz = ck / MB, lb = grid[(lbz * l1gx + lby) * l1gx + lbx];
          if (lb < 0) continue;
          int cc = ((ck - lbz * MB) * MB + (cj - lby * MB)) * MB + (ci - lbx * MB);
          R fq[Q], nq[Q], fr, fx, fy, fz;
          for (int q = 0; q < Q; ++q) fq[q] = S::load(f1[(size_t(lb) * Q + q) * MB3 + cc], q, P.sc[1].dec);
          moments(fq, fr, fx, fy, fz);
          R uxl = fx / fr, uyl = fy / fr, uzl = fz / fr;
          for (int q = 0; q < Q; ++q) nq[q] = fq[q] - feq_q<R>(q, fr, uxl, uyl, uzl);
          regularize(nq);
          wsum += wgt;
          for (int q = 0; q < Q; ++q) out[q] += R(wgt) * (feq_q<R>(q, fr, uxl, uyl, uzl) + R(nqs) * nq[q]);
        }
    if (wsum > 0.f && wsum < 0.999f) {
      R inv = R(1) / R(wsum);
      for (int q = 0; q < Q; ++q) out[q] *= inv;
    }
  }
  // prolong_prep_L0 / prolong_prep_L1, src/tgv3d_amr.cu:122-144.
  void prolong_prep_L0(int cfi, int cfj, int cfk, R* out) const {
    int ib, jb, kb;
    float tx, ty, tz;
    tri(cfi, ib, tx);
    tri(cfj, jb, ty);
    tri(cfk, kb, tz);
    for (int q = 0; q < Q; ++q) out[q] = 0;
    for (int dk = 0; dk < 2; ++dk)
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          float wgt = (di ? tx : 1 - tx) * (dj ? ty : 1 - ty) * (dk ? tz : 1 - tz);
          int ci = pmod(ib + di, N), cj = pmod(jb + dj, N), ck = pmod(kb + dk, N);
          for (int q = 0; q < Q; ++q) out[q] += R(wgt) * prep0[l0idx(ci, cj, ck, q)];
        }
  }
  void prolong_prep_L1(int gfi, int gfj, int gfk, R* out) const {
    int ib, jb, kb;
    float tx, ty, tz, wsum = 0;
    tri(gfi, ib, tx);
    tri(gfj, jb, ty);
    tri(gfk, kb, tz);
    for (int q = 0; q < Q; ++q) out[q] = 0;
    for (int dk = 0; dk < 2; ++dk)
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          float wgt = (di ? tx : 1 - tx) * (dj ? ty : 1 - ty) * (dk ? tz : 1 - tz);
          int ci = pmod(ib + di, N1), cj = pmod(jb + dj, N1), ck = pmod(kb + dk, N1);
          int lbx = ci / MB, lby = cj / MB, lbz = ck / MB, lb = l1grid[(lbz * l1gx + lby) * l1gx + lbx];
          if (lb < 0) continue;
          int cc = ((ck - lbz * MB) * MB + (cj - lby * MB)) * MB + (ci - lbx * MB);
          wsum += wgt;
          for (int q = 0; q < Q; ++q) out[q] += R(wgt) * prep1[(size_t(lb) * Q + q) * MB3 + cc];
        }
    if (wsum > 0.f && wsum < 0.999f) {
      R inv = R(1) / R(wsum);
      for (int q = 0; q < Q; ++q) out[q] *= inv;
    }
  }

  // ---- kernels ----
  // prep = feq + nqs*regularize(f - feq) over the listed tiles, src/tgv3d_amr.cu:160-171.
  void prep_list(const std::vector<raw>& fin, std::vector<R>& pr, const std::vector<int>& list, int nlist,
                 float nqs, float dec, bool dense_l0) {
    (void)dense_l0;
#pragma omp parallel for schedule(static)
    for (int li = 0; li < nlist; ++li) {
      const int b = list[li];
      for (int cell = 0; cell < MB3; ++cell) {
        R f[Q], nq[Q], rho, ux, uy, uz;
        for (int q = 0; q < Q; ++q) f[q] = S::load(fin[(size_t(b) * Q + q) * MB3 + cell], q, dec);
        moments(f, rho, ux, uy, uz);
        ux /= rho;
        uy /= rho;
        uz /= rho;
        for (int q = 0; q < Q; ++q) nq[q] = f[q] - feq_q<R>(q, rho, ux, uy, uz);
        regularize(nq);
        for (int q = 0; q < Q; ++q) pr[(size_t(b) * Q + q) * MB3 + cell] = feq_q<R>(q, rho, ux, uy, uz) + R(nqs) * nq[q];
      }
    }
  }
  // L1 ghost fill from prep0, src/tgv3d_amr.cu:173-186.
  void fillcf_L1p() {
#pragma omp parallel for schedule(static)
    for (int fb = 0; fb < nf1; ++fb) {
      const int cb = l1pcb[fb], oct = l1oct[fb], qx = oct & 1, qy = (oct >> 1) & 1, qz = oct >> 2;
      const int cbx = cb % cgx, cby = (cb / cgx) % cgx, cbz = cb / (cgx * cgx);
      const int ox = cbx * 2 * MB + qx * MB, oy = cby * 2 * MB + qy * MB, oz = cbz * 2 * MB + qz * MB;
      for (int s = 0; s < EB3; ++s) {
        int ei = s % EB, ej = (s / EB) % EB, ek = s / (EB * EB), i = ei - 1, j = ej - 1, k = ek - 1;
        if (i >= 0 && i < MB && j >= 0 && j < MB && k >= 0 && k < MB) continue;
        int gfi = pmod(ox + i, N1), gfj = pmod(oy + j, N1), gfk = pmod(oz + k, N1);
        int scb = ((gfk / (2 * MB)) * cgx + (gfj / (2 * MB))) * cgx + (gfi / (2 * MB));
        if (ref0[scb]) continue;
        R out[Q];
        prolong_prep_L0(gfi, gfj, gfk, out);
        for (int q = 0; q < Q; ++q) gh1[(size_t(fb) * Q + q) * EB3 + s] = out[q];
      }
    }
  }
  // L2 ghost fill from prep1, src/tgv3d_amr.cu:187-202.
  void fillcf_L2p() {
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nf2; ++b) {
      const int pl1 = l2pl1[b], oct = l2oct[b], qx = oct & 1, qy = (oct >> 1) & 1, qz = oct >> 2;
      const int p = f1pos[pl1], px = p % l1gx, py = (p / l1gx) % l1gx, pz = p / (l1gx * l1gx);
      const int ox = px * 2 * MB + qx * MB, oy = py * 2 * MB + qy * MB, oz = pz * 2 * MB + qz * MB;
      for (int s = 0; s < EB3; ++s) {
        int ei = s % EB, ej = (s / EB) % EB, ek = s / (EB * EB), i = ei - 1, j = ej - 1, k = ek - 1;
        if (i >= 0 && i < MB && j >= 0 && j < MB && k >= 0 && k < MB) continue;
        int g2i = pmod(ox + i, N2), g2j = pmod(oy + j, N2), g2k = pmod(oz + k, N2);
        int lb = l1grid[((g2k / (2 * MB)) * l1gx + (g2j / (2 * MB))) * l1gx + (g2i / (2 * MB))];
        if (lb >= 0 && ref1[lb]) continue;
        R out[Q];
        prolong_prep_L1(g2i, g2j, g2k, out);
        for (int q = 0; q < Q; ++q) gh2[(size_t(b) * Q + q) * EB3 + s] = out[q];
      }
    }
  }
  // Dense coarse step, src/tgv3d_amr.cu:204-220 (identical to the uniform step).
  void step_L0() {
    step_uniform<S>(c0a.data(), c0b.data(), N, om0, P.sc[0]);
    std::swap(c0a, c0b);
  }
  // Fine L1 pull, src/tgv3d_amr.cu:241-274.
  void step_L1() {
    const Scale sc = P.sc[1];
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nf1; ++b) {
      const int cb = l1pcb[b], oct = l1oct[b], qx = oct & 1, qy = (oct >> 1) & 1, qz = oct >> 2;
      const int cbx = cb % cgx, cby = (cb / cgx) % cgx, cbz = cb / (cgx * cgx);
      int sref[27], sfof[27];
      for (int c = 0; c < 27; ++c) {
        int dx = c % 3 - 1, dy = (c / 3) % 3 - 1, dz = c / 9 - 1;
        int nx = (cbx + dx + cgx) % cgx, ny = (cby + dy + cgx) % cgx, nz = (cbz + dz + cgx) % cgx;
        int ncb = (nz * cgx + ny) * cgx + nx;
        sref[c] = ref0[ncb];
        sfof[c] = fof0[ncb];
      }
      const int ox = cbx * 2 * MB + qx * MB, oy = cby * 2 * MB + qy * MB, oz = cbz * 2 * MB + qz * MB;
      for (int cell = 0; cell < MB3; ++cell) {
        const int i = cell % MB, j = (cell / MB) % MB, k = cell / (MB * MB), gfi = ox + i, gfj = oy + j, gfk = oz + k;
        R fin[Q];
        for (int q = 0; q < Q; ++q) {
          int sfi = (gfi - cvx[q] + N1) % N1, sfj = (gfj - cvy[q] + N1) % N1, sfk = (gfk - cvz[q] + N1) % N1;
          int ddx = sfi / (2 * MB) - cbx, ddy = sfj / (2 * MB) - cby, ddz = sfk / (2 * MB) - cbz;
          if (ddx > 1) ddx -= cgx; else if (ddx < -1) ddx += cgx;
          if (ddy > 1) ddy -= cgx; else if (ddy < -1) ddy += cgx;
          if (ddz > 1) ddz -= cgx; else if (ddz < -1) ddz += cgx;
          int sidx = (ddz + 1) * 9 + (ddy + 1) * 3 + (ddx + 1);
          if (sref[sidx]) {
            int sqx = (sfi % (2 * MB)) / MB, sqy = (sfj % (2 * MB)) / MB, sqz = (sfk % (2 * MB)) / MB;
            int sfb = sfof[sidx] + (sqz * 4 + sqy * 2 + sqx), sc2 = ((sfk % MB) * MB + (sfj % MB)) * MB + (sfi % MB);
            fin[q] = S::in(f1a[(size_t(sfb) * Q + q) * MB3 + sc2], q, sc.dec);
          } else {
            int ei = i - cvx[q] + 1, ej = j - cvy[q] + 1, ek = k - cvz[q] + 1;
            fin[q] = S::from_full(gh1[(size_t(b) * Q + q) * EB3 + (ek * EB + ej) * EB + ei], q);
          }
        }
        collide_store(fin, om1, f1b, b, cell, sc);
      }
    }
    std::swap(f1a, f1b);
  }
  // Fine L2 pull, src/tgv3d_amr.cu:331-356.
  void step_L2() {
    const Scale sc = P.sc[2];
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nf2; ++b) {
      const int pl1 = l2pl1[b], oct = l2oct[b], qx = oct & 1, qy = (oct >> 1) & 1, qz = oct >> 2;
      const int p = f1pos[pl1], px = p % l1gx, py = (p / l1gx) % l1gx, pz = p / (l1gx * l1gx);
      const int ox = px * 2 * MB + qx * MB, oy = py * 2 * MB + qy * MB, oz = pz * 2 * MB + qz * MB;
      for (int cell = 0; cell < MB3; ++cell) {
        const int i = cell % MB, j = (cell / MB) % MB, k = cell / (MB * MB), g2i = ox + i, g2j = oy + j, g2k = oz + k;
        R fin[Q];
        for (int q = 0; q < Q; ++q) {
          int sfi = (g2i - cvx[q] + N2) % N2, sfj = (g2j - cvy[q] + N2) % N2, sfk = (g2k - cvz[q] + N2) % N2;
          int lb = l1grid[((sfk / (2 * MB)) * l1gx + (sfj / (2 * MB))) * l1gx + (sfi / (2 * MB))];
          if (lb >= 0 && ref1[lb]) {
            int sqx = (sfi % (2 * MB)) / MB, sqy = (sfj % (2 * MB)) / MB, sqz = (sfk % (2 * MB)) / MB;
            int sfb = fof1[lb] + (sqz * 4 + sqy * 2 + sqx), sc2 = ((sfk % MB) * MB + (sfj % MB)) * MB + (sfi % MB);
            fin[q] = S::in(f2a[(size_t(sfb) * Q + q) * MB3 + sc2], q, sc.dec);
          } else {
            int ei = i - cvx[q] + 1, ej = j - cvy[q] + 1, ek = k - cvz[q] + 1;
            fin[q] = S::from_full(gh2[(size_t(b) * Q + q) * EB3 + (ek * EB + ej) * EB + ei], q);
          }
        }
        collide_store(fin, om2, f2b, b, cell, sc);
      }
    }
    std::swap(f2a, f2b);
  }
  // fin: pulled values in the form's working representation (src/tgv3d_amr.cu:268-273)
  static void collide_store(R* fin, R om, std::vector<raw>& out, int b, int cell, Scale sc) {
    collide<S>(fin, om);
    for (int q = 0; q < Q; ++q) out[(size_t(b) * Q + q) * MB3 + cell] = S::out(fin[q], q, sc.enc);
  }
  // Restriction fine -> parent, src/tgv3d_amr.cu:275-298 (L1 -> L0) and 358-381 (L2 -> L1).
  void restrict_level(std::vector<raw>& cf, const std::vector<raw>& ff, const std::vector<int>& rl,
                      const std::vector<int>& fof, int n_ref, float nqi, Scale csc, Scale fsc) {
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n_ref; ++r) {
      const int cb = rl[r];
      for (int cell = 0; cell < MB3; ++cell) {
        const int lci = cell % MB, lcj = (cell / MB) % MB, lck = cell / (MB * MB);
        const int qx = lci / HALF, qy = lcj / HALF, qz = lck / HALF, oct = qz * 4 + qy * 2 + qx, fb = fof[cb] + oct;
        const int fi0 = 2 * (lci % HALF), fj0 = 2 * (lcj % HALF), fk0 = 2 * (lck % HALF);
        R avg[Q], nq[Q], rho, ux, uy, uz;
        for (int q = 0; q < Q; ++q) avg[q] = 0;
        for (int dk = 0; dk < 2; ++dk)
          for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di) {
              int sc2 = ((fk0 + dk) * MB + (fj0 + dj)) * MB + (fi0 + di);
              for (int q = 0; q < Q; ++q) avg[q] += R(0.125) * S::load(ff[(size_t(fb) * Q + q) * MB3 + sc2], q, fsc.dec);
            }
        moments(avg, rho, ux, uy, uz);
        ux /= rho;
        uy /= rho;
        uz /= rho;
        for (int q = 0; q < Q; ++q) nq[q] = avg[q] - feq_q<R>(q, rho, ux, uy, uz);
        regularize(nq);
        const int cc = (lck * MB + lcj) * MB + lci;
        for (int q = 0; q < Q; ++q)
          cf[(size_t(cb) * Q + q) * MB3 + cc] = S::store(feq_q<R>(q, rho, ux, uy, uz) + R(nqi) * nq[q], q, csc.enc);
      }
    }
  }
  // Seeding, src/tgv3d_amr.cu:299-308 and 382-393.
  void seed_L1() {
#pragma omp parallel for schedule(static)
    for (int fb = 0; fb < nf1; ++fb) {
      const int cb = l1pcb[fb], oct = l1oct[fb], qx = oct & 1, qy = (oct >> 1) & 1, qz = oct >> 2;
      const int cbx = cb % cgx, cby = (cb / cgx) % cgx, cbz = cb / (cgx * cgx);
      for (int cell = 0; cell < MB3; ++cell) {
        const int i = cell % MB, j = (cell / MB) % MB, k = cell / (MB * MB);
        R out[Q];
        prolong_L0(c0a, cbx * 2 * MB + qx * MB + i, cby * 2 * MB + qy * MB + j, cbz * 2 * MB + qz * MB + k, nqs01, out);
        for (int q = 0; q < Q; ++q) f1a[(size_t(fb) * Q + q) * MB3 + cell] = S::store(out[q], q, P.sc[1].enc);
      }
    }
  }
  void seed_L2() {
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nf2; ++b) {
      const int pl1 = l2pl1[b], oct = l2oct[b], qx = oct & 1, qy = (oct >> 1) & 1, qz = oct >> 2;
      const int p = f1pos[pl1], px = p % l1gx, py = (p / l1gx) % l1gx, pz = p / (l1gx * l1gx);
      for (int cell = 0; cell < MB3; ++cell) {
        const int i = cell % MB, j = (cell / MB) % MB, k = cell / (MB * MB);
        R out[Q];
        prolong_L1(f1a, l1grid, px * 2 * MB + qx * MB + i, py * 2 * MB + qy * MB + j, pz * 2 * MB + qz * MB + k, nqs12, out);
        for (int q = 0; q < Q; ++q) f2a[(size_t(b) * Q + q) * MB3 + cell] = S::store(out[q], q, P.sc[2].enc);
      }
    }
  }
  // Migration, src/tgv3d_amr.cu:397-430.
  void migrate_L1() {
#pragma omp parallel for schedule(static)
    for (int nfb = 0; nfb < nf1; ++nfb) {
      const int cb = l1pcb[nfb], oct = l1oct[nfb];
      for (int cell = 0; cell < MB3; ++cell) {
        if (o_ref0[cb]) {
          int ofb = o_fof0[cb] + oct;
          for (int q = 0; q < Q; ++q) f1b[(size_t(nfb) * Q + q) * MB3 + cell] = f1a[(size_t(ofb) * Q + q) * MB3 + cell];
        } else {
          int qx = oct & 1, qy = (oct >> 1) & 1, qz = oct >> 2, cbx = cb % cgx, cby = (cb / cgx) % cgx, cbz = cb / (cgx * cgx);
          int i = cell % MB, j = (cell / MB) % MB, k = cell / (MB * MB);
          R out[Q];
          prolong_L0(c0a, cbx * 2 * MB + qx * MB + i, cby * 2 * MB + qy * MB + j, cbz * 2 * MB + qz * MB + k, nqs01, out);
          for (int q = 0; q < Q; ++q) f1b[(size_t(nfb) * Q + q) * MB3 + cell] = S::store(out[q], q, P.sc[1].enc);
        }
      }
    }
    std::swap(f1a, f1b);
  }
  void migrate_L2() {
#pragma omp parallel for schedule(static)
    for (int nfb = 0; nfb < nf2; ++nfb) {
      const int pl1 = l2pl1[nfb], oct = l2oct[nfb], p = f1pos[pl1];
      const int old_l1b = o_l1grid[p];
      for (int cell = 0; cell < MB3; ++cell) {
        if (old_l1b >= 0 && o_ref1[old_l1b]) {
          int of2 = o_fof1[old_l1b] + oct;
          for (int q = 0; q < Q; ++q) f2b[(size_t(nfb) * Q + q) * MB3 + cell] = f2a[(size_t(of2) * Q + q) * MB3 + cell];
        } else {
          int px = p % l1gx, py = (p / l1gx) % l1gx, pz = p / (l1gx * l1gx), qx = oct & 1, qy = (oct >> 1) & 1, qz = oct >> 2;
          int i = cell % MB, j = (cell / MB) % MB, k = cell / (MB * MB);
          R out[Q];
          prolong_L1(f1a, l1grid, px * 2 * MB + qx * MB + i, py * 2 * MB + qy * MB + j, pz * 2 * MB + qz * MB + k, nqs12, out);
          for (int q = 0; q < Q; ++q) f2b[(size_t(nfb) * Q + q) * MB3 + cell] = S::store(out[q], q, P.sc[2].enc);
        }
      }
    }
    std::swap(f2a, f2b);
  }
  // Sensor: per tile max over cells of sqrt(sum_q (f-feq)^2)/rho, src/tgv3d_amr.cu:442-453.
  void sensor(const std::vector<raw>& ff, int nblk, std::vector<float>& d, float dec) const {
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nblk; ++b) {
      float mx = 0.f;
      bool first = true;
      for (int cell = 0; cell < MB3; ++cell) {
        R f[Q], rho, ux, uy, uz, s = 0;
        for (int q = 0; q < Q; ++q) f[q] = S::load(ff[(size_t(b) * Q + q) * MB3 + cell], q, dec);
        moments(f, rho, ux, uy, uz);
        ux /= rho;
        uy /= rho;
        uz /= rho;
        for (int q = 0; q < Q; ++q) {
          R d2 = f[q] - feq_q<R>(q, rho, ux, uy, uz);
          s += d2 * d2;
        }
        float v = float(std::sqrt(s) / rho);
        mx = first ? v : std::fmax(mx, v);
        first = false;
      }
      d[b] = mx;
    }
  }
  // 6-face-neighbour presence in l1grid, src/tgv3d_amr.cu:459-474.
  bool is_deep(int b) const {
    int p = f1pos[b], px = p % l1gx, py = (p / l1gx) % l1gx, pz = p / (l1gx * l1gx);
    static const int nb6[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (int d = 0; d < 6; ++d) {
      int nx = (px + nb6[d][0] + l1gx) % l1gx, ny = (py + nb6[d][1] + l1gx) % l1gx, nz = (pz + nb6[d][2] + l1gx) % l1gx;
      if (l1grid[(nz * l1gx + ny) * l1gx + nx] < 0) return false;
    }
    return true;
  }
  // exclusive scan of 0/1 flags, then the count nref = rank[n-1] + ref[n-1] (src/tgv3d_amr.cu:733-735).
  int scan_count(const std::vector<int>& flags, int n) {
    if (n <= 0) return 0;
    std::exclusive_scan(flags.begin(), flags.begin() + n, rank.begin(), 0);
    return rank[n - 1] + flags[n - 1];
  }
  // threshold = ascending-sorted[n - tgt] (src/tgv3d_amr.cu:728-731)
  static float threshold(std::vector<float> v, int tgt) {
    std::sort(v.begin(), v.end());
    return v[v.size() - tgt];
  }
  // Table scatter, src/tgv3d_amr.cu:476-494. Writes past the pool capacity are refused and flagged.
  void scatter_L1() {
    for (int cb = 0; cb < cnb0; ++cb) {
      if (!ref0[cb]) {
        fof0[cb] = -1;
        continue;
      }
      int r = rank[cb];
      fof0[cb] = 8 * r;
      if (8 * r + 8 > cap1) {
        overflow = true;
        continue;
      }
      rl0[r] = cb;
      int cbx = cb % cgx, cby = (cb / cgx) % cgx, cbz = cb / (cgx * cgx);
      for (int oct = 0; oct < 8; ++oct) {
        int qx = oct & 1, qy = (oct >> 1) & 1, qz = oct >> 2, fb = 8 * r + oct;
        int pos = ((cbz * 2 + qz) * l1gx + (cby * 2 + qy)) * l1gx + (cbx * 2 + qx);
        l1pcb[fb] = cb;
        l1oct[fb] = oct;
        f1pos[fb] = pos;
        l1grid[pos] = fb;
      }
    }
  }
  void scatter_L2(int cap2_) {
    for (int b = 0; b < nf1; ++b) {
      if (!ref1[b]) {
        fof1[b] = -1;
        continue;
      }
      int r = rank[b];
      fof1[b] = 8 * r;
      if (8 * r + 8 > cap2_) {
        overflow = true;
        continue;
      }
      rl1[r] = b;
      for (int oct = 0; oct < 8; ++oct) {
        l2pl1[8 * r + oct] = b;
        l2oct[8 * r + oct] = oct;
      }
    }
  }
  // Skin lists: refined or with a refined 26-neighbour, ascending (src/tgv3d_amr.cu:502-513, 758-762).
  void build_skins() {
    n_skin0 = 0;
    for (int cb = 0; cb < cnb0; ++cb) {
      int bx = cb % cgx, by = (cb / cgx) % cgx, bz = cb / (cgx * cgx), f = 0;
      for (int dz = -1; dz <= 1 && !f; ++dz)
        for (int dy = -1; dy <= 1 && !f; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            int nx = (bx + dx + cgx) % cgx, ny = (by + dy + cgx) % cgx, nz = (bz + dz + cgx) % cgx;
            if (ref0[(nz * cgx + ny) * cgx + nx]) {
              f = 1;
              break;
            }
          }
      if (f) skin0[n_skin0++] = cb;
    }
    n_skin1 = 0;
    for (int b = 0; b < nf1; ++b) {
      int p = f1pos[b], px = p % l1gx, py = (p / l1gx) % l1gx, pz = p / (l1gx * l1gx), f = 0;
      for (int dz = -1; dz <= 1 && !f; ++dz)
        for (int dy = -1; dy <= 1 && !f; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            int nx = (px + dx + l1gx) % l1gx, ny = (py + dy + l1gx) % l1gx, nz = (pz + dz + l1gx) % l1gx;
            int nb = l1grid[(nz * l1gx + ny) * l1gx + nx];
            if (nb >= 0 && ref1[nb]) {
              f = 1;
              break;
            }
          }
      if (f) skin1[n_skin1++] = b;
    }
  }
  void ensure_l2_pool() {
    f2a.assign(size_t(cap2) * Q * MB3, raw(0));
    f2b = f2a;
    gh2.assign(size_t(cap2) * Q * EB3, R(0));
    l2pl1.assign(cap2, 0);
    l2oct.assign(cap2, 0);
  }

  // ---- pipeline ----
  void init_tgv_l0() { init_tgv<S>(c0a.data(), N, P.u0, kf, P.phase, P.sc[0]); }
  void develop(int n) {
    for (int s = 0; s < n; ++s) step_L0();
  }
  // Initial build, src/tgv3d_amr.cu:776-815. sen0_in / sen1_by_pos (indexed by L1 grid position)
  // replace the computed sensors when given: the hook that feeds both sides the same sensor field.
  void initial_build(const float* sen0_in, const float* sen1_by_pos) {
    if (P.l1_mode == MODE_ALL) std::fill(ref0.begin(), ref0.end(), 1);
    else if (P.l1_mode == MODE_SLAB)
      for (int cb = 0; cb < cnb0; ++cb) ref0[cb] = (cb / (cgx * cgx) < cgx / 2) ? 1 : 0;
    else {
      develop(P.develop);
      sensor(c0a, cnb0, sen0, P.sc[0].dec);
      if (sen0_in) std::copy(sen0_in, sen0_in + cnb0, sen0.begin());
      int tgt0 = std::min(cap_nref0, std::max(1, int(P.frac1 * cnb0)));
      float thr = threshold(std::vector<float>(sen0.begin(), sen0.end()), tgt0);
      for (int i = 0; i < cnb0; ++i) ref0[i] = sen0[i] >= thr ? 1 : 0;
    }
    nref0 = scan_count(ref0, cnb0);
    nf1 = 8 * nref0;
    hw_nf1 = std::max(hw_nf1, long(nf1));
    if (nf1 > cap1 && !grow_L1(false)) {
      overflow = true;
      return;
    }
    std::fill(l1grid.begin(), l1grid.end(), -1);
    std::fill(fof0.begin(), fof0.end(), -1);
    if (nref0) scatter_L1();
    if (overflow) return;
    if (nf1) {
      std::fill(gh1.begin(), gh1.begin() + size_t(nf1) * Q * EB3, R(0));
      seed_L1();
    }
    nref1 = 0;
    nf2 = 0;
    if (nf1) {
      sensor(f1a, nf1, sen1, P.sc[1].dec);
      if (sen1_by_pos)
        for (int b = 0; b < nf1; ++b) sen1[b] = sen1_by_pos[f1pos[b]];
      for (int b = 0; b < nf1; ++b) deep[b] = is_deep(b) ? 1 : 0;
      std::fill(ref1.begin(), ref1.end(), 0);
      if (P.l2_mode == MODE_ALL)
        for (int b = 0; b < nf1; ++b) ref1[b] = deep[b];
      else {
        std::vector<float> c;  // stable copy_if of the deep sensors
        for (int b = 0; b < nf1; ++b)
          if (deep[b]) c.push_back(sen1[b]);
        int ndeep = int(c.size());
        if (ndeep > 0) {
          int tgt1 = std::max(1, int(P.frac2 * ndeep));
          float thr = threshold(c, tgt1);
          for (int b = 0; b < nf1; ++b) ref1[b] = (deep[b] && sen1[b] >= thr) ? 1 : 0;
        }
      }
      nref1 = scan_count(ref1, nf1);
      nf2 = 8 * nref1;
      hw_nf2 = std::max(hw_nf2, long(nf2));
      cap_nref1 = std::max(1, int(std::ceil(nref1 * 2.5)));
      if (P.l2_mode == MODE_ALL) cap_nref1 = std::max(cap_nref1, nf1);
      cap_nref1 = std::min(cap_nref1, cap1);
      while (cap_nref1 > 1 && !pool_fits(cap1, 8L * cap_nref1)) cap_nref1--;
      cap2 = 8 * cap_nref1;
    } else {
      cap_nref1 = 1;
      cap2 = 8;
    }
    ensure_l2_pool();
    if (nf2 > cap2) overflow = true;
    std::fill(fof1.begin(), fof1.end(), -1);
    if (nf2) {
      scatter_L2(cap2);
      if (overflow) return;
      seed_L2();
    }
    build_skins();
  }
  // Regrid, src/tgv3d_amr.cu:722-764.
  void regrid(const float* sen0_in, const float* sen1_by_pos) {
    o_ref0 = ref0;
    o_fof0 = fof0;
    o_l1grid = l1grid;
    o_ref1 = ref1;
    o_fof1 = fof1;
    sensor(c0a, cnb0, sen0, P.sc[0].dec);
    if (sen0_in) std::copy(sen0_in, sen0_in + cnb0, sen0.begin());
    int tgt0 = std::min(cap_nref0, std::max(1, int(P.frac1 * cnb0)));
    float thr = threshold(std::vector<float>(sen0.begin(), sen0.end()), tgt0);
    for (int i = 0; i < cnb0; ++i) ref0[i] = sen0[i] >= thr ? 1 : 0;
    nref0 = scan_count(ref0, cnb0);
    nf1 = 8 * nref0;
    hw_nf1 = std::max(hw_nf1, long(nf1));
    if (nf1 > cap1 && !grow_L1(true)) {
      overflow = true;
      return;
    }
    std::fill(l1grid.begin(), l1grid.end(), -1);
    std::fill(fof0.begin(), fof0.end(), -1);
    if (nref0) scatter_L1();
    if (overflow) return;
    if (nf1) migrate_L1();
    if (f1b.size() != f1a.size()) f1b.assign(f1a.size(), raw(0));
    nref1 = 0;
    nf2 = 0;
    if (nf1) {
      sensor(f1a, nf1, sen1, P.sc[1].dec);
      if (sen1_by_pos)
        for (int b = 0; b < nf1; ++b) sen1[b] = sen1_by_pos[f1pos[b]];
      for (int b = 0; b < nf1; ++b) deep[b] = is_deep(b) ? 1 : 0;
      std::vector<float> c;
      for (int b = 0; b < nf1; ++b)
        if (deep[b]) c.push_back(sen1[b]);
      int ndeep = int(c.size());
      std::fill(ref1.begin(), ref1.end(), 0);
      if (ndeep > 0) {
        int tgt1 = std::min(cap_nref1, std::max(1, int(P.frac2 * ndeep)));
        float thr1 = threshold(c, tgt1);
        for (int b = 0; b < nf1; ++b) ref1[b] = (deep[b] && sen1[b] >= thr1) ? 1 : 0;
      }
      nref1 = scan_count(ref1, nf1);
      nf2 = 8 * nref1;
      hw_nf2 = std::max(hw_nf2, long(nf2));
      // Ties at the threshold can exceed the level-2 pool (the reference writes out of bounds there):
      // the pool grows when the budget allows, the refinement target cap_nref1 stays as measured.
      std::vector<raw> oldf2;
      if (nf2 > cap2) {
        int cn = std::max(nref1, (cap2 / 8) * 3 / 2);
        while (cn > nref1 && !pool_fits(cap1, 8L * cn)) cn--;
        if (!pool_fits(cap1, 8L * cn)) {
          overflow = true;
          return;
        }
        cap2 = 8 * cn;
        oldf2 = f2a;
        ensure_l2_pool();
        f2a = oldf2;
      }
      std::fill(fof1.begin(), fof1.end(), -1);
      if (nref1) scatter_L2(cap2);
      if (overflow) return;
      if (nf2) migrate_L2();
      if (f2b.size() != f2a.size()) f2b.assign(f2a.size(), raw(0));
    }
    build_skins();
  }
  // One coarse step of the nested 1:2:4 subcycle, src/tgv3d_amr.cu:821-837.
  void base_step() {
    step_L0();
    if (!nf1) return;
    if (!P.skipprep) {
      prep_list(c0a, prep0, skin0, n_skin0, nqs01, P.sc[0].dec, true);
      fillcf_L1p();
    }
    for (int s1 = 0; s1 < 2; ++s1) {
      if (nf2) {
        if (!P.skipprep) {
          prep_list(f1a, prep1, skin1, n_skin1, nqs12, P.sc[1].dec, false);
          fillcf_L2p();
        }
        for (int s2 = 0; s2 < 2; ++s2) step_L2();
        if (!P.skiprestr) restrict_level(f1a, f2a, rl1, fof1, nref1, nqi21, P.sc[1], P.sc[2]);
      }
      step_L1();
    }
    if (!P.skiprestr) restrict_level(c0a, f1a, rl0, fof0, nref0, nqi10, P.sc[0], P.sc[1]);
  }
  // Composite macro field at N2 = 4 N0, src/tgv3d_amr.cu:515-536.
  void composite(std::vector<R>& F) const {
    F.assign(size_t(N2) * N2 * N2 * 4, R(0));
#pragma omp parallel for schedule(static)
    for (int gk = 0; gk < N2; ++gk)
      for (int gj = 0; gj < N2; ++gj)
        for (int gi = 0; gi < N2; ++gi) {
          size_t idx = (size_t(gk) * N2 + gj) * N2 + gi;
          int lb = l1grid[((gk / (2 * MB)) * l1gx + (gj / (2 * MB))) * l1gx + (gi / (2 * MB))];
          const std::vector<raw>* src;
          size_t blk;
          int sc2;
          float dec;
          if (lb >= 0 && ref1[lb]) {
            int qx = (gi % (2 * MB)) / MB, qy = (gj % (2 * MB)) / MB, qz = (gk % (2 * MB)) / MB;
            blk = size_t(fof1[lb] + (qz * 4 + qy * 2 + qx));
            sc2 = ((gk % MB) * MB + (gj % MB)) * MB + (gi % MB);
            src = &f2a;
            dec = P.sc[2].dec;
          } else if (lb >= 0) {
            int g1i = gi / 2, g1j = gj / 2, g1k = gk / 2;
            blk = size_t(lb);
            sc2 = ((g1k % MB) * MB + (g1j % MB)) * MB + (g1i % MB);
            src = &f1a;
            dec = P.sc[1].dec;
          } else {
            int c0i = gi / 4, c0j = gj / 4, c0k = gk / 4, cbx = c0i / MB, cby = c0j / MB, cbz = c0k / MB;
            blk = size_t((cbz * cgx + cby) * cgx + cbx);
            sc2 = ((c0k % MB) * MB + (c0j % MB)) * MB + (c0i % MB);
            src = &c0a;
            dec = P.sc[0].dec;
          }
          R v[Q], r, x, y, z;
          for (int q = 0; q < Q; ++q) v[q] = S::load((*src)[(blk * Q + q) * MB3 + sc2], q, dec);
          moments(v, r, x, y, z);
          F[idx * 4 + 0] = r;
          F[idx * 4 + 1] = x / r;
          F[idx * 4 + 2] = y / r;
          F[idx * 4 + 3] = z / r;
        }
  }
  void diag(double out[3]) const {
    std::vector<R> F;
    composite(F);
    diag_composite(F.data(), N2, out);
  }
};

}  // namespace orc
