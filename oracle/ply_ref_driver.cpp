// Reference-side PLY writer / reader for tests: drives the reference's vendored
// tinyply (third_party/tinyply) with the same property requests as
// GaussianModel::savePly / loadPly (src/gaussian_model.cpp:860-1069).
// TEST INFRASTRUCTURE ONLY (built into oracle/_ref by oracle/Makefile).
//
//   ply_ref write <raw_in> <ply_out>   raw_in: int64 P, int32 K, then xyz [P,3],
//                                      f_dc [P,3] (channel order), f_rest [P,3K]
//                                      (channel-major), opacity [P], scale [P,3],
//                                      rot [P,4], all float32
//   ply_ref read <ply_in> <K> <raw_out>  the same raw layout back
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "tinyply.h"

namespace {

struct Arrays {
  int64_t P = 0;
  int32_t K = 0;
  std::vector<float> xyz, dc, rest, op, sc, rot;
};

bool read_raw(const char* path, Arrays& a) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  bool ok = std::fread(&a.P, 8, 1, f) == 1 && std::fread(&a.K, 4, 1, f) == 1;
  auto rd = [&](std::vector<float>& v, size_t n) {
    v.resize(n);
    ok = ok && (n == 0 || std::fread(v.data(), 4, n, f) == n);
  };
  rd(a.xyz, 3 * a.P);
  rd(a.dc, 3 * a.P);
  rd(a.rest, 3 * (size_t)a.K * a.P);
  rd(a.op, a.P);
  rd(a.sc, 3 * a.P);
  rd(a.rot, 4 * a.P);
  std::fclose(f);
  return ok;
}

bool write_raw(const char* path, const Arrays& a) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  bool ok = std::fwrite(&a.P, 8, 1, f) == 1 && std::fwrite(&a.K, 4, 1, f) == 1;
  for (const auto* v : {&a.xyz, &a.dc, &a.rest, &a.op, &a.sc, &a.rot})
    ok = ok && (v->empty() || std::fwrite(v->data(), 4, v->size(), f) == v->size());
  return (std::fclose(f) == 0) && ok;
}

std::vector<std::string> names(const char* stem, int n) {
  std::vector<std::string> v(n);
  for (int i = 0; i < n; ++i) v[i] = std::string(stem) + std::to_string(i);
  return v;
}

int do_write(const char* raw, const char* out) {
  Arrays a;
  if (!read_raw(raw, a)) return 2;
  std::vector<float> normals(3 * a.P, 0.f);
  std::filebuf fb;
  fb.open(out, std::ios::out | std::ios::binary);
  std::ostream os(&fb);
  tinyply::PlyFile ply;
  auto add = [&](const std::vector<std::string>& n, std::vector<float>& v) {
    ply.add_properties_to_element("vertex", n, tinyply::Type::FLOAT32, a.P, reinterpret_cast<uint8_t*>(v.data()),
                                  tinyply::Type::INVALID, 0);
  };
  add({"x", "y", "z"}, a.xyz);
  add({"nx", "ny", "nz"}, normals);
  add(names("f_dc_", 3), a.dc);
  add(names("f_rest_", 3 * a.K), a.rest);
  add({"opacity"}, a.op);
  add(names("scale_", 3), a.sc);
  add(names("rot_", 4), a.rot);
  ply.write(os, true);
  fb.close();
  return 0;
}

int do_read(const char* in, int K, const char* raw) {
  std::ifstream is(in, std::ios::binary);
  if (!is) return 2;
  tinyply::PlyFile ply;
  ply.parse_header(is);
  auto req = [&](const std::vector<std::string>& n) { return ply.request_properties_from_element("vertex", n); };
  auto xyz = req({"x", "y", "z"});
  auto dc = req(names("f_dc_", 3));
  std::shared_ptr<tinyply::PlyData> rest;
  if (K > 0) rest = req(names("f_rest_", 3 * K));
  auto op = req({"opacity"});
  auto sc = req(names("scale_", 3));
  auto rot = req(names("rot_", 4));
  ply.read(is);
  Arrays a;
  a.P = (int64_t)xyz->count;
  a.K = K;
  auto cp = [](std::shared_ptr<tinyply::PlyData>& d, std::vector<float>& v) {
    v.resize(d->buffer.size_bytes() / 4);
    std::memcpy(v.data(), d->buffer.get(), d->buffer.size_bytes());
  };
  cp(xyz, a.xyz);
  cp(dc, a.dc);
  if (K > 0) cp(rest, a.rest);
  cp(op, a.op);
  cp(sc, a.sc);
  cp(rot, a.rot);
  return write_raw(raw, a) ? 0 : 3;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 4 && std::string(argv[1]) == "write") return do_write(argv[2], argv[3]);
  if (argc == 5 && std::string(argv[1]) == "read") return do_read(argv[2], std::atoi(argv[3]), argv[4]);
  std::fprintf(stderr, "usage: ply_ref write <raw> <ply> | read <ply> <K> <raw>\n");
  return 1;
}
