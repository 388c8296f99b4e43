// Command-line lookup with the reference's semantics (src/tgv3d_uniform.cu:92-94): `--key value`
// pairs, the first match wins, unknown keys are ignored, a key in the last position has no value.
#pragma once
#include <cstdlib>
#include <cstring>

namespace jpcli {
inline int argi(int c, char** v, const char* k, int d) {
  for (int i = 1; i < c - 1; ++i)
    if (!std::strcmp(v[i], k)) return std::atoi(v[i + 1]);
  return d;
}
inline double argd(int c, char** v, const char* k, double d) {
  for (int i = 1; i < c - 1; ++i)
    if (!std::strcmp(v[i], k)) return std::atof(v[i + 1]);
  return d;
}
inline const char* args(int c, char** v, const char* k, const char* d) {
  for (int i = 1; i < c - 1; ++i)
    if (!std::strcmp(v[i], k)) return v[i + 1];
  return d;
}
// --prec fp32|fp16|bf16|fp64 (default fp32); -1 for an unknown name
inline int prec_of(const char* s) {
  if (!std::strcmp(s, "fp32")) return 0;
  if (!std::strcmp(s, "fp16")) return 1;
  if (!std::strcmp(s, "bf16")) return 2;
  if (!std::strcmp(s, "fp64")) return 3;
  return -1;
}
}  // namespace jpcli
