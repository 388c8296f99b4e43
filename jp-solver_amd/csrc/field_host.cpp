// Frame output of the drivers (the reference's --vtk path): the .vti of one output field and the .vtp of the
// AMR level boxes. The arrays are produced on the device by jp_field; this file only lays them out in the
// reference's file formats (src/tgv3d_uniform.cu:110-140; src/tgv3d_amr.cu:602-641), which its
// visualisation scripts read.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/jp.h"
#include "ctx.hpp"

namespace {

struct File {
  FILE* f;
  explicit File(const char* path) : f(std::fopen(path, "wb")) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

// ImageData with appended raw arrays, each preceded by its UInt64 byte count.
struct VtiArray {
  const char* name;
  int comps;
  const float* data;
  size_t count;
};
bool write_image(FILE* f, int n, const char* point_data_attrs, const std::vector<VtiArray>& arrays) {
  std::fprintf(f, "<?xml version=\"1.0\"?>\n");
  std::fprintf(f, "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n");
  std::fprintf(f, "  <ImageData WholeExtent=\"0 %d 0 %d 0 %d\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n", n - 1, n - 1, n - 1);
  std::fprintf(f, "    <Piece Extent=\"0 %d 0 %d 0 %d\">\n", n - 1, n - 1, n - 1);
  std::fprintf(f, "      <PointData %s>\n", point_data_attrs);
  uint64_t offset = 0;
  for (const VtiArray& a : arrays) {
    if (a.comps == 1)
      std::fprintf(f, "        <DataArray type=\"Float32\" Name=\"%s\" format=\"appended\" offset=\"%llu\"/>\n", a.name,
                   static_cast<unsigned long long>(offset));
    else
      std::fprintf(f,
                   "        <DataArray type=\"Float32\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"appended\" "
                   "offset=\"%llu\"/>\n",
                   a.name, a.comps, static_cast<unsigned long long>(offset));
    offset += 8 + uint64_t(a.count) * 4;
  }
  std::fprintf(f, "      </PointData>\n    </Piece>\n  </ImageData>\n");
  std::fprintf(f, "  <AppendedData encoding=\"raw\">\n_");
  for (const VtiArray& a : arrays) {
    const uint64_t nb = uint64_t(a.count) * 4;
    if (std::fwrite(&nb, 8, 1, f) != 1) return false;
    if (nb && std::fwrite(a.data, 1, nb, f) != nb) return false;
  }
  std::fprintf(f, "\n  </AppendedData>\n</VTKFile>\n");
  return !std::ferror(f);
}

}  // namespace

extern "C" {

int jp_write_vti(jp_ctx* c, const char* path) {
  if (!c || !path) return JP_ERR_ARG;
  const long n = jp_field_edge(c);
  if (n <= 0) return jp::fail(c, JP_ERR_STATE, "no output field for this context");
  const size_t cells = size_t(n) * n * n;
  const bool uniform = c->kind == 0;
  // the arrays come back through a pinned buffer kept by the context (several times the pageable rate)
  const size_t need = cells * (uniform ? 5 : 2) * sizeof(float);
  if (c->host_out_bytes < need) {
    if (c->host_out) (void)hipHostFree(c->host_out);
    c->host_out = nullptr;
    c->host_out_bytes = 0;
    if (hipHostMalloc(reinterpret_cast<void**>(&c->host_out), need, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return jp::fail(c, JP_ERR_NOMEM, "pinned host buffer for the frame");
    }
    c->host_out_bytes = need;
  }
  float *hq = c->host_out, *hs = hq + cells, *hv = hs + cells;
  if (int r = jp_field(c, hq, hs, uniform ? hv : nullptr, long(cells))) return r;
  File f(path);
  if (!f.f) return jp::fail(c, JP_ERR_ARG, std::string("cannot open ") + path);
  const VtiArray q{"Q", 1, hq, cells}, s{"speed", 1, hs, cells}, v{"velocity", 3, hv, 3 * cells};
  const bool ok = uniform ? write_image(f.f, int(n), "Scalars=\"Q\" Vectors=\"velocity\"", {q, s, v})
                          : write_image(f.f, int(n), "Scalars=\"Q\"", {q, s});
  return ok ? JP_OK : jp::fail(c, JP_ERR_ARG, std::string("short write to ") + path);
}

int jp_write_boxes(jp_ctx* c, const char* path) {
  if (!c || !path) return JP_ERR_ARG;
  if (c->kind != 1) return jp::fail(c, JP_ERR_STATE, "level boxes exist for AMR contexts only");
  if (int r = jp::amr_sync_counts(c)) return r;
  const jp::AmrState& a = c->amr;
  auto table = [&](int which, std::vector<int>& out, long len) -> int {
    out.resize(size_t(len));
    long got = 0;
    return len ? jp_get_table(c, which, out.data(), len, &got) : JP_OK;
  };
  std::vector<int> l1pcb, l1oct, l2pl1, l2oct, f1pos;
  if (int r = table(JP_TABLE_L1PCB, l1pcb, a.nf1)) return r;
  if (int r = table(JP_TABLE_L1OCT, l1oct, a.nf1)) return r;
  if (int r = table(JP_TABLE_F1POS, f1pos, a.nf1)) return r;
  if (int r = table(JP_TABLE_L2PL1, l2pl1, a.nf2)) return r;
  if (int r = table(JP_TABLE_L2OCT, l2oct, a.nf2)) return r;
  // box origin and edge in finest (4 N0) cells, and its level
  struct Box {
    double x, y, z, s;
    int level;
  };
  std::vector<Box> boxes;
  boxes.reserve(l1pcb.size() + l2pl1.size());
  for (size_t b = 0; b < l1pcb.size(); ++b) {  // a level-1 tile spans 16 finest cells
    const int cb = l1pcb[b], oct = l1oct[b];
    const int bx = cb % a.cgx, by = (cb / a.cgx) % a.cgx, bz = cb / (a.cgx * a.cgx);
    boxes.push_back({double(2 * (bx * 16 + (oct & 1) * 8)), double(2 * (by * 16 + ((oct >> 1) & 1) * 8)),
                     double(2 * (bz * 16 + (oct >> 2) * 8)), 16.0, 1});
  }
  for (size_t b = 0; b < l2pl1.size(); ++b) {  // a level-2 tile spans 8
    const int p = f1pos[size_t(l2pl1[b])], oct = l2oct[b];
    const int px = p % a.l1gx, py = (p / a.l1gx) % a.l1gx, pz = p / (a.l1gx * a.l1gx);
    boxes.push_back({double(px * 16 + (oct & 1) * 8), double(py * 16 + ((oct >> 1) & 1) * 8),
                     double(pz * 16 + (oct >> 2) * 8), 8.0, 2});
  }
  File f(path);
  if (!f.f) return jp::fail(c, JP_ERR_ARG, std::string("cannot open ") + path);
  const long nb = long(boxes.size());
  std::fprintf(f.f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"LittleEndian\">\n");
  std::fprintf(f.f, "  <PolyData>\n    <Piece NumberOfPoints=\"%ld\" NumberOfLines=\"%ld\">\n", nb * 8, nb * 12);
  std::fprintf(f.f, "      <PointData><DataArray type=\"Int32\" Name=\"level\" format=\"ascii\">\n");
  for (const Box& b : boxes)
    for (int v = 0; v < 8; ++v) std::fprintf(f.f, "%d\n", b.level);
  std::fprintf(f.f, "      </DataArray></PointData>\n      <Points>\n");
  std::fprintf(f.f, "        <DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (const Box& b : boxes)  // corner v: x + s (v & 1), y + s (v >> 1 & 1), z + s (v >> 2)
    for (int v = 0; v < 8; ++v)
      std::fprintf(f.f, "%g %g %g\n", b.x + ((v & 1) ? b.s : 0.0), b.y + ((v & 2) ? b.s : 0.0), b.z + ((v & 4) ? b.s : 0.0));
  std::fprintf(f.f, "        </DataArray>\n      </Points>\n      <Lines>\n");
  std::fprintf(f.f, "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n");
  static const int kEdge[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
  for (long b = 0; b < nb; ++b)
    for (const auto& e : kEdge) std::fprintf(f.f, "%ld %ld\n", b * 8 + e[0], b * 8 + e[1]);
  std::fprintf(f.f, "        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
  for (long e = 0; e < nb * 12; ++e) std::fprintf(f.f, "%ld\n", 2 * (e + 1));
  std::fprintf(f.f, "        </DataArray>\n      </Lines>\n    </Piece>\n  </PolyData>\n</VTKFile>\n");
  return std::ferror(f.f) ? jp::fail(c, JP_ERR_ARG, std::string("short write to ") + path) : JP_OK;
}

}  // extern "C"
