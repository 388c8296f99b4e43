// Device regrid primitives (src/tgv3d_amr.cu:455-513 and the Thrust calls at 728-762, 779-812):
// exact order-statistic threshold, marking, 2:1 deep mask, exclusive scan, stable compaction, block-table
// scatter, dilation, and the 27-neighbour tables the fine step reads.
//
// The threshold is the (count - tgt)-th ascending sensor value, found by a 4-pass radix select over
// order-preserving 32-bit keys in one workgroup (LDS histograms), so the sensor is never sorted or
// copied to the host. The scan and the compaction use wavefront ballots: a 0/1 flag per lane, the
// prefix within the wave from mbcnt, wave totals combined through LDS. All results are exact and
// independent of scheduling.
#include <algorithm>
#include <climits>

#include "amr_internal.hpp"

namespace jp {

__device__ inline uint32_t fkey(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float fkey_inv(uint32_t k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k); }

constexpr int kSelBS = 1024;

// Item count of a launch: from the device when dn is given, else n. The thread-per-item kernels walk
// their items with a grid stride, so a launch recorded before a regrid covers any later count.
__device__ inline int item_count(int n, const int* dn) { return dn ? *dn : n; }
#define JP_ITEMS(i, n) for (int i = int(blockIdx.x) * 256 + int(threadIdx.x); i < (n); i += int(gridDim.x) * 256)

// One workgroup. count = #entries (mask != 0); tgt = min(cap, max(1, int(frac*count)));
// thr = value at ascending rank count - tgt. Four 8-bit radix passes over the orderable keys; the keys
// and their mask bits stay in LDS between passes when they fit. The digit holding the rank is found by
// one wave: each lane owns 4 bins, a wave prefix sum gives the
// cumulative counts and the lane whose range contains the rank resolves it (the first bin whose
// cumulative count exceeds the rank, as a sequential walk would).
constexpr int kSelLds = 12288;
__global__ __launch_bounds__(kSelBS) void k_select(const float* __restrict__ sen, const int* __restrict__ mask, int n_host,
                                                   const int* __restrict__ dn, double frac, int cap, float* __restrict__ thr,
                                                   int* __restrict__ cnt, int count_slot) {
  __shared__ unsigned hist[256];
  __shared__ unsigned s_prefix, s_rank;
  __shared__ int s_count;
  __shared__ uint32_t keys[kSelLds];
  __shared__ uint8_t valid[kSelLds];
  const int tid = threadIdx.x;
  const int n = item_count(n_host, dn);
  const bool in_lds = n <= kSelLds;
  if (in_lds)
    for (int i = tid; i < n; i += kSelBS) {
      valid[i] = (!mask || mask[i]) ? 1 : 0;
      keys[i] = fkey(sen[i]);
    }
  unsigned prefix = 0;
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 24 - 8 * pass;
    const unsigned himask = pass == 0 ? 0u : (0xffffffffu << (32 - 8 * pass));
    for (int b = tid; b < 256; b += kSelBS) hist[b] = 0;
    __syncthreads();
    for (int i = tid; i < n; i += kSelBS) {
      uint32_t k;
      if (in_lds) {
        if (!valid[i]) continue;
        k = keys[i];
      } else {
        if (mask && !mask[i]) continue;
        k = fkey(sen[i]);
      }
      if ((k & himask) == (prefix & himask)) atomicAdd(&hist[(k >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (tid < 64) {
      const unsigned h0 = hist[4 * tid], h1 = hist[4 * tid + 1], h2 = hist[4 * tid + 2], h3 = hist[4 * tid + 3];
      const unsigned local = h0 + h1 + h2 + h3;
      unsigned incl = local;
      for (int off = 1; off < 64; off <<= 1) {
        const unsigned v = __shfl_up(incl, off, 64);
        if (tid >= off) incl += v;
      }
      const unsigned tot = __shfl(incl, 63, 64);
      unsigned r;
      if (pass == 0) {
        int tgt = int(frac * double(tot));
        tgt = tgt < 1 ? 1 : tgt;
        tgt = tgt > cap ? cap : tgt;
        r = tot > 0 ? unsigned(int(tot) - tgt) : 0u;
        if (tid == 0) {
          s_count = int(tot);
          cnt[count_slot] = int(tot);
          cnt[CNT_TGT] = tgt;
        }
      } else {
        r = s_rank;
      }
      const unsigned excl = incl - local;
      if (excl <= r && r < incl) {  // exactly one lane when r < tot
        unsigned acc = excl;
        int d = 4 * tid;
        if (acc + h0 <= r) {
          acc += h0;
          ++d;
          if (acc + h1 <= r) {
            acc += h1;
            ++d;
            if (acc + h2 <= r) {
              acc += h2;
              ++d;
            }
          }
        }
        s_rank = r - acc;
        s_prefix = prefix | (unsigned(d) << shift);
      }
    }
    __syncthreads();
    prefix = s_prefix;
    if (s_count == 0) break;
  }
  if (tid == 0 && s_count > 0) *thr = fkey_inv(prefix);
}

__global__ void k_mark_thr(const float* __restrict__ sen, const float* __restrict__ thr, int* __restrict__ ref, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) ref[i] = (sen[i] >= *thr) ? 1 : 0;
}
// ref1 = deep && sen >= thr when the deep set is non-empty, else 0 (src/tgv3d_amr.cu:459-466, 745-749)
__global__ void k_mark_ref1(const float* __restrict__ sen, const float* __restrict__ thr, const int* __restrict__ deep,
                            const int* __restrict__ cnt, int* __restrict__ ref1, int n_host, const int* __restrict__ dn) {
  const int nf1 = item_count(n_host, dn);
  JP_ITEMS(b, nf1) ref1[b] = (cnt[CNT_NDEEP] > 0 && deep[b] && sen[b] >= *thr) ? 1 : 0;
}
// modes: 0 all (k_setall), 1 slab bz < cgx/2 (k_mark_slab), 2 copy of deep (k_mark_ref1_all); amr.cu:497-500
__global__ void k_set(int* __restrict__ ref, int n, int mode, int cgx, const int* __restrict__ deep) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  ref[i] = mode == 0 ? 1 : (mode == 1 ? ((i / (cgx * cgx) < cgx / 2) ? 1 : 0) : deep[i]);
}
// deep = all 6 face neighbours present in l1grid, periodic (src/tgv3d_amr.cu:468-474)
__global__ void k_deepmask(const int* __restrict__ f1pos, const int* __restrict__ l1grid, int l1gx, int n_host,
                           const int* __restrict__ dn, int* __restrict__ deep) {
  const int nf1 = item_count(n_host, dn);
  JP_ITEMS(b, nf1) {
    const int p = f1pos[b], px = p % l1gx, py = (p / l1gx) % l1gx, pz = p / (l1gx * l1gx);
    const int d6[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    int ok = 1;
    for (int d = 0; d < 6; ++d) {
      const int nx = (px + d6[d][0] + l1gx) % l1gx, ny = (py + d6[d][1] + l1gx) % l1gx, nz = (pz + d6[d][2] + l1gx) % l1gx;
      if (l1grid[(nz * l1gx + ny) * l1gx + nx] < 0) {
        ok = 0;
        break;
      }
    }
    deep[b] = ok;
  }
}

// Ballot prefix helpers (wave64).
__device__ inline unsigned lane_prefix(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi(unsigned(m >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(m), 0u));
}

// Exclusive scan (kCompact = false: rank[i] = #set flags before i) or stable compaction
// (kCompact = true: out[#set before i] = i for set i). Total to cnt[slot]. One workgroup.
template <bool kCompact>
__global__ __launch_bounds__(kSelBS) void k_ballot_scan(const int* __restrict__ flags, int n_host, const int* __restrict__ dn,
                                                        int* __restrict__ out, int* __restrict__ cnt, int slot, int tiles_slot,
                                                        int cap_tiles) {
  __shared__ unsigned wtot[kSelBS / 64];
  __shared__ unsigned s_carry;
  const int tid = threadIdx.x, w = tid >> 6;
  const int n = item_count(n_host, dn);
  if (tid == 0) s_carry = 0;
  __syncthreads();
  for (int base = 0; base < n; base += kSelBS) {
    const int i = base + tid;
    const int f = (i < n && flags[i] != 0) ? 1 : 0;
    const unsigned long long m = __ballot(f);
    const unsigned lp = lane_prefix(m);
    if ((tid & 63) == 0) wtot[w] = unsigned(__popcll(m));
    __syncthreads();
    unsigned off = s_carry;
    for (int v = 0; v < w; ++v) off += wtot[v];
    if (i < n) {
      if (kCompact) {
        if (f) out[off + lp] = i;
      } else {
        out[i] = int(off + lp);
      }
    }
    __syncthreads();
    if (tid == 0) {
      unsigned t = 0;
      for (int v = 0; v < kSelBS / 64; ++v) t += wtot[v];
      s_carry += t;
    }
    __syncthreads();
  }
  if (tid == 0) {
    cnt[slot] = int(s_carry);
    if (tiles_slot >= 0) {
      const int t = 8 * min(int(s_carry), cap_tiles / 8);
      cnt[tiles_slot] = t;
      cnt[tiles_slot + 2] = max(cnt[tiles_slot + 2], t);
    }
  }
}

// Block tables of level 1 from ref0 and its rank (src/tgv3d_amr.cu:476-486). Tiles beyond the pool
// capacity are not written; the overflow flag is raised instead. With clamp, a parent whose tiles do not
// fit is left unrefined, so that the tables describe a consistent (smaller) grid that later launches may
// read; without it, ref and fof0 keep the parent for a scatter repeated after the pool grows.
__global__ void k_scatter_l1(int* __restrict__ ref, const int* __restrict__ rank, int cnb0, int cgx, int l1gx, int cap1,
                             int clamp, int* __restrict__ fof0, int* __restrict__ rl0, int* __restrict__ l1pcb,
                             int* __restrict__ l1oct, int* __restrict__ l1grid, int* __restrict__ f1pos,
                             int* __restrict__ cnt) {
  const int cb = blockIdx.x * 256 + threadIdx.x;
  if (cb >= cnb0) return;
  if (!ref[cb]) {
    fof0[cb] = -1;
    return;
  }
  const int r = rank[cb];
  if (8 * r + 8 > cap1) {
    atomicOr(&cnt[CNT_OVERFLOW], 1);
    if (clamp) ref[cb] = 0;
    fof0[cb] = clamp ? -1 : 8 * r;
    return;
  }
  fof0[cb] = 8 * r;
  rl0[r] = cb;
  const int cbx = cb % cgx, cby = (cb / cgx) % cgx, cbz = cb / (cgx * cgx);
  for (int oct = 0; oct < 8; ++oct) {
    const int fb = 8 * r + oct;
    const int pos = ((cbz * 2 + (oct >> 2)) * l1gx + (cby * 2 + ((oct >> 1) & 1))) * l1gx + (cbx * 2 + (oct & 1));
    l1pcb[fb] = cb;
    l1oct[fb] = oct;
    f1pos[fb] = pos;
    l1grid[pos] = fb;
  }
}
// Level-2 tables (src/tgv3d_amr.cu:488-494); overflow and clamp as for level 1.
__global__ void k_scatter_l2(int* __restrict__ ref1, const int* __restrict__ rank, int n_host, const int* __restrict__ dn,
                             int cap2, int clamp, int* __restrict__ fof1, int* __restrict__ rl1, int* __restrict__ l2pl1,
                             int* __restrict__ l2oct, int* __restrict__ cnt) {
  const int nf1 = item_count(n_host, dn);
  JP_ITEMS(b, nf1) {
    if (!ref1[b]) {
      fof1[b] = -1;
      continue;
    }
    const int r = rank[b];
    if (8 * r + 8 > cap2) {
      atomicOr(&cnt[CNT_OVERFLOW], 1);
      if (clamp) ref1[b] = 0;
      fof1[b] = clamp ? -1 : 8 * r;
      continue;
    }
    fof1[b] = 8 * r;
    rl1[r] = b;
    for (int oct = 0; oct < 8; ++oct) {
      l2pl1[8 * r + oct] = b;
      l2oct[8 * r + oct] = oct;
    }
  }
}
// Skin flags: refined, or a refined 26-neighbour (src/tgv3d_amr.cu:502-513).
__global__ void k_dilate_l0(const int* __restrict__ ref0, int cnb0, int cgx, int* __restrict__ flags) {
  const int cb = blockIdx.x * 256 + threadIdx.x;
  if (cb >= cnb0) return;
  const int bx = cb % cgx, by = (cb / cgx) % cgx, bz = cb / (cgx * cgx);
  int f = 0;
  for (int d = 0; d < 27 && !f; ++d) {
    const int nx = (bx + d % 3 - 1 + cgx) % cgx, ny = (by + (d / 3) % 3 - 1 + cgx) % cgx, nz = (bz + d / 9 - 1 + cgx) % cgx;
    f = ref0[(nz * cgx + ny) * cgx + nx] ? 1 : 0;
  }
  flags[cb] = f;
}
__global__ void k_dilate_l1(const int* __restrict__ ref1, const int* __restrict__ f1pos, const int* __restrict__ l1grid,
                            int n_host, const int* __restrict__ dn, int l1gx, int* __restrict__ flags) {
  const int nf1 = item_count(n_host, dn);
  JP_ITEMS(b, nf1) {
    const int p = f1pos[b], px = p % l1gx, py = (p / l1gx) % l1gx, pz = p / (l1gx * l1gx);
    int f = 0;
    for (int d = 0; d < 27 && !f; ++d) {
      const int nx = (px + d % 3 - 1 + l1gx) % l1gx, ny = (py + (d / 3) % 3 - 1 + l1gx) % l1gx,
                nz = (pz + d / 9 - 1 + l1gx) % l1gx;
      const int nb = l1grid[(nz * l1gx + ny) * l1gx + nx];
      f = (nb >= 0 && ref1[nb]) ? 1 : 0;
    }
    flags[b] = f;
  }
}
// 27 same-level neighbours of each level-1 tile: the tile at the neighbouring level-1 grid position.
__global__ void k_nbr_l1(const int* __restrict__ f1pos, const int* __restrict__ l1grid, int l1gx, int n_host,
                         const int* __restrict__ dn, int* __restrict__ nb1) {
  const int nf1 = item_count(n_host, dn);
  JP_ITEMS(t, nf1 * kNbr) {
    const int b = t / kNbr, d = t % kNbr;
    const int p = f1pos[b], px = p % l1gx, py = (p / l1gx) % l1gx, pz = p / (l1gx * l1gx);
    const int nx = (px + d % 3 - 1 + l1gx) % l1gx, ny = (py + (d / 3) % 3 - 1 + l1gx) % l1gx,
              nz = (pz + d / 9 - 1 + l1gx) % l1gx;
    nb1[t] = l1grid[(nz * l1gx + ny) * l1gx + nx];
  }
}
// Level-2 neighbours: the level-2 grid position's parent level-1 tile must be present and refined.
__global__ void k_nbr_l2(const int* __restrict__ l2pl1, const int* __restrict__ l2oct, const int* __restrict__ f1pos,
                         const int* __restrict__ l1grid, const int* __restrict__ ref1, const int* __restrict__ fof1,
                         int l1gx, int n_host, const int* __restrict__ dn, int* __restrict__ nb2) {
  const int nf2 = item_count(n_host, dn);
  const int l2gx = 2 * l1gx;
  JP_ITEMS(t, nf2 * kNbr) {
    const int b = t / kNbr, d = t % kNbr;
    const int p = f1pos[l2pl1[b]], oct = l2oct[b];
    const int x = (p % l1gx) * 2 + (oct & 1), y = ((p / l1gx) % l1gx) * 2 + ((oct >> 1) & 1),
              z = (p / (l1gx * l1gx)) * 2 + (oct >> 2);
    const int nx = (x + d % 3 - 1 + l2gx) % l2gx, ny = (y + (d / 3) % 3 - 1 + l2gx) % l2gx, nz = (z + d / 9 - 1 + l2gx) % l2gx;
    const int lb = l1grid[((nz >> 1) * l1gx + (ny >> 1)) * l1gx + (nx >> 1)];
    nb2[t] = (lb >= 0 && ref1[lb]) ? fof1[lb] + (nz & 1) * 4 + (ny & 1) * 2 + (nx & 1) : -1;
  }
}
__global__ void k_gather_by_pos(const float* __restrict__ by_pos, const int* __restrict__ f1pos, int n_host,
                                const int* __restrict__ dn, float* __restrict__ out) {
  const int nf1 = item_count(n_host, dn);
  JP_ITEMS(b, nf1) out[b] = by_pos[f1pos[b]];
}

static inline dim3 g256(long n) { return dim3(unsigned((n + 255) / 256)); }
// Launch width of a thread-per-item launch over `per` threads per item; 0 = nothing to launch. A
// device count launches at least one workgroup, which loops over whatever count it reads.
static inline unsigned width256(Items it, long per = 1) {
  const long w = (long(std::max(it.n, 0)) * per + 255) / 256;
  return unsigned(it.dn ? std::max(w, 1L) : w);
}

hipError_t rg_select(const float* sen, const int* mask, Items n, double frac, int cap, float* thr, int* cnt, int slot,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_select, dim3(1), dim3(kSelBS), 0, st, sen, mask, n.n, n.dn, frac, cap, thr, cnt, slot);
  return hipGetLastError();
}
hipError_t rg_mark_thr(const float* sen, const float* thr, int* ref, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mark_thr, g256(n), dim3(256), 0, st, sen, thr, ref, n);
  return hipGetLastError();
}
hipError_t rg_mark_ref1(const float* sen, const float* thr, const int* deep, const int* cnt, int* ref1, Items nf1,
                        hipStream_t st) {
  if (const unsigned w = width256(nf1))
    hipLaunchKernelGGL(k_mark_ref1, dim3(w), dim3(256), 0, st, sen, thr, deep, cnt, ref1, nf1.n, nf1.dn);
  return hipGetLastError();
}
hipError_t rg_set(int* ref, int n, int mode, int cgx, const int* deep, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_set, g256(n), dim3(256), 0, st, ref, n, mode, cgx, deep);
  return hipGetLastError();
}
hipError_t rg_deepmask(const int* f1pos, const int* l1grid, int l1gx, Items nf1, int* deep, hipStream_t st) {
  if (const unsigned w = width256(nf1))
    hipLaunchKernelGGL(k_deepmask, dim3(w), dim3(256), 0, st, f1pos, l1grid, l1gx, nf1.n, nf1.dn, deep);
  return hipGetLastError();
}
hipError_t rg_scan(const int* flags, Items n, int* rank, int* cnt, int slot, int tiles_slot, int cap_tiles,
                   hipStream_t st) {
  hipLaunchKernelGGL((k_ballot_scan<false>), dim3(1), dim3(kSelBS), 0, st, flags, n.n, n.dn, rank, cnt, slot, tiles_slot,
                     cap_tiles);
  return hipGetLastError();
}
hipError_t rg_compact(const int* flags, Items n, int* out, int* cnt, int slot, hipStream_t st) {
  hipLaunchKernelGGL((k_ballot_scan<true>), dim3(1), dim3(kSelBS), 0, st, flags, n.n, n.dn, out, cnt, slot, -1, 0);
  return hipGetLastError();
}
hipError_t rg_scatter_l1(int* ref, const int* rank, int cnb0, int cgx, int l1gx, int cap1, bool clamp, int* fof0, int* rl0,
                         int* l1pcb, int* l1oct, int* l1grid, int* f1pos, int* cnt, hipStream_t st) {
  hipLaunchKernelGGL(k_scatter_l1, g256(cnb0), dim3(256), 0, st, ref, rank, cnb0, cgx, l1gx, cap1, clamp ? 1 : 0, fof0, rl0,
                     l1pcb, l1oct, l1grid, f1pos, cnt);
  return hipGetLastError();
}
hipError_t rg_scatter_l2(int* ref1, const int* rank, Items nf1, int cap2, bool clamp, int* fof1, int* rl1, int* l2pl1,
                         int* l2oct, int* cnt, hipStream_t st) {
  if (const unsigned w = width256(nf1))
    hipLaunchKernelGGL(k_scatter_l2, dim3(w), dim3(256), 0, st, ref1, rank, nf1.n, nf1.dn, cap2, clamp ? 1 : 0, fof1, rl1,
                       l2pl1, l2oct, cnt);
  return hipGetLastError();
}
hipError_t rg_dilate_l0(const int* ref0, int cnb0, int cgx, int* flags, hipStream_t st) {
  hipLaunchKernelGGL(k_dilate_l0, g256(cnb0), dim3(256), 0, st, ref0, cnb0, cgx, flags);
  return hipGetLastError();
}
hipError_t rg_dilate_l1(const int* ref1, const int* f1pos, const int* l1grid, Items nf1, int l1gx, int* flags,
                        hipStream_t st) {
  if (const unsigned w = width256(nf1))
    hipLaunchKernelGGL(k_dilate_l1, dim3(w), dim3(256), 0, st, ref1, f1pos, l1grid, nf1.n, nf1.dn, l1gx, flags);
  return hipGetLastError();
}
hipError_t rg_nbr_l1(const int* f1pos, const int* l1grid, int l1gx, Items nf1, int* nb1, hipStream_t st) {
  if (const unsigned w = width256(nf1, kNbr))
    hipLaunchKernelGGL(k_nbr_l1, dim3(w), dim3(256), 0, st, f1pos, l1grid, l1gx, nf1.n, nf1.dn, nb1);
  return hipGetLastError();
}
hipError_t rg_nbr_l2(const int* l2pl1, const int* l2oct, const int* f1pos, const int* l1grid, const int* ref1,
                     const int* fof1, int l1gx, Items nf2, int* nb2, hipStream_t st) {
  if (const unsigned w = width256(nf2, kNbr))
    hipLaunchKernelGGL(k_nbr_l2, dim3(w), dim3(256), 0, st, l2pl1, l2oct, f1pos, l1grid, ref1, fof1, l1gx, nf2.n, nf2.dn,
                       nb2);
  return hipGetLastError();
}
hipError_t rg_gather_by_pos(const float* by_pos, const int* f1pos, Items nf1, float* out, hipStream_t st) {
  if (const unsigned w = width256(nf1))
    hipLaunchKernelGGL(k_gather_by_pos, dim3(w), dim3(256), 0, st, by_pos, f1pos, nf1.n, nf1.dn, out);
  return hipGetLastError();
}
__global__ void k_differs(const int* __restrict__ a, const int* __restrict__ b, int n_host, const int* __restrict__ dn,
                          int* __restrict__ cnt, int slot) {
  const int n = item_count(n_host, dn);
  JP_ITEMS(i, n) if (a[i] != b[i]) atomicOr(&cnt[slot], 1);
}
hipError_t rg_differs(const int* a, const int* b, Items n, int* cnt, int slot, hipStream_t st) {
  if (hipError_t e = fill_i32(cnt + slot, 0, 1, st)) return e;
  if (const unsigned w = width256(n)) hipLaunchKernelGGL(k_differs, dim3(w), dim3(256), 0, st, a, b, n.n, n.dn, cnt, slot);
  return hipGetLastError();
}

struct CopyList {
  const int* src[5];
  int* dst[5];
  long n[5];
  int count;
};
__global__ void k_copy_i32_multi(CopyList L, long nmax) {
  const long i = long(blockIdx.x) * 256 + threadIdx.x;
  if (i >= nmax) return;
  for (int k = 0; k < L.count; ++k)
    if (i < L.n[k]) L.dst[k][i] = L.src[k][i];
}
hipError_t rg_copy_i32_multi(const int* const* src, int* const* dst, const long* n, int count, hipStream_t st) {
  if (count < 1 || count > 5) return hipErrorInvalidValue;
  CopyList L{};
  long nmax = 0;
  for (int k = 0; k < count; ++k) {
    L.src[k] = src[k];
    L.dst[k] = dst[k];
    L.n[k] = n[k];
    nmax = std::max(nmax, n[k]);
  }
  L.count = count;
  if (nmax <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_i32_multi, dim3(unsigned((nmax + 255) / 256)), dim3(256), 0, st, L, nmax);
  return hipGetLastError();
}

}  // namespace jp
