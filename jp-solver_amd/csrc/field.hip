// Output field of a frame: Q-criterion, speed and velocity from the macroscopic field (rho, u, v, w)
// (src/tgv3d_uniform.cu:96-119 on the host; src/tgv3d_amr.cu:553-564 on the device), computed on the
// device so only the output arrays cross to the host. Velocity gradients are periodic central differences
// in FP64; the uniform reference takes the difference of two float field values in float before widening
// (kDiffInField), the AMR reference widens first.
#include "internal.hpp"

namespace jp {

template <class R, bool kDiffInField>
__global__ __launch_bounds__(256) void k_qspeed(const V4<R>* __restrict__ F, uint32_t N, FastDiv dN, float* __restrict__ Qo,
                                                float* __restrict__ spd, float* __restrict__ vel) {
  const uint32_t cell = blockIdx.x * 256 + threadIdx.x;
  if (cell >= N * N * N) return;
  const uint32_t row = fdiv(cell, dN), i = cell - row * N, k = fdiv(row, dN), j = row - k * N;
  const uint32_t ip = i == N - 1 ? 0 : i + 1, im = i == 0 ? N - 1 : i - 1;
  const uint32_t jp = j == N - 1 ? 0 : j + 1, jm = j == 0 ? N - 1 : j - 1;
  const uint32_t kp = k == N - 1 ? 0 : k + 1, km = k == 0 ? N - 1 : k - 1;
  auto at = [&](uint32_t a, uint32_t b, uint32_t c) { return F[(c * N + b) * N + a]; };
  const V4<R> o = F[cell];
  const V4<R> xp = at(ip, j, k), xm = at(im, j, k), yp = at(i, jp, k), ym = at(i, jm, k), zp = at(i, j, kp), zm = at(i, j, km);
  auto d = [&](R a, R b) -> double {
    if constexpr (kDiffInField) return double(R(a - b)) * 0.5;
    else return (double(a) - double(b)) * 0.5;
  };
  // g[c][e] = d u_c / d x_e
  double g[3][3];
  g[0][0] = d(xp.b, xm.b); g[0][1] = d(yp.b, ym.b); g[0][2] = d(zp.b, zm.b);
  g[1][0] = d(xp.c, xm.c); g[1][1] = d(yp.c, ym.c); g[1][2] = d(zp.c, zm.c);
  g[2][0] = d(xp.d, xm.d); g[2][1] = d(yp.d, ym.d); g[2][2] = d(zp.d, zm.d);
  double Q = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) Q -= 0.5 * g[a][b] * g[b][a];
  const double u = o.b, v = o.c, w = o.d;
  Qo[cell] = float(Q);
  spd[cell] = float(sqrt(u * u + v * v + w * w));
  if (vel) {
    vel[size_t(cell) * 3 + 0] = float(u);
    vel[size_t(cell) * 3 + 1] = float(v);
    vel[size_t(cell) * 3 + 2] = float(w);
  }
}

hipError_t field_qspeed(bool dbl_field, bool diff_in_field, const void* field, int N, float* Q, float* spd, float* vel,
                        hipStream_t st) {
  const uint64_t n3 = uint64_t(N) * N * N;
  const dim3 grid(uint32_t((n3 + 255) / 256));
  const FastDiv dN = make_fastdiv(uint32_t(N));
  if (dbl_field)
    hipLaunchKernelGGL((k_qspeed<double, false>), grid, dim3(256), 0, st, static_cast<const V4<double>*>(field), uint32_t(N),
                       dN, Q, spd, vel);
  else if (diff_in_field)
    hipLaunchKernelGGL((k_qspeed<float, true>), grid, dim3(256), 0, st, static_cast<const V4<float>*>(field), uint32_t(N),
                       dN, Q, spd, vel);
  else
    hipLaunchKernelGGL((k_qspeed<float, false>), grid, dim3(256), 0, st, static_cast<const V4<float>*>(field), uint32_t(N),
                       dN, Q, spd, vel);
  return hipGetLastError();
}

}  // namespace jp
