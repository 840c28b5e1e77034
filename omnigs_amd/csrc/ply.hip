// 3DGS PLY reader / writer (host code): the vertex layout of the reference's
// GaussianModel::savePly / loadPly (src/gaussian_model.cpp:860-1069), which
// write and read it through tinyply:
//   x y z nx ny nz f_dc_0..2 f_rest_0..(3K-1) opacity scale_0..2 rot_0..3
// all float32; normals are zero; f_dc_c = features_dc[p][0][c] and
// f_rest_{c K + j} = features_rest[p][j][c] (channel-major, K = (D+1)^2 - 1);
// opacity, scale and rot are the raw (pre-activation) parameters.
//
// The writer emits tinyply's header text and interleaved little-endian rows.
// The reader accepts any property order and scalar type in binary (either
// byte order) or ASCII files, converting to float; a missing property is an
// error naming it.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/omnigs_lonlat.h"

namespace {

char g_ply_error[512];

int ply_fail(const std::string& msg) {
  std::snprintf(g_ply_error, sizeof(g_ply_error), "%s", msg.c_str());
  return OMNIGS_ERR_INVALID_ARGUMENT;
}

struct Prop {
  std::string name;
  int bytes;
  char kind;  // 'i' signed, 'u' unsigned, 'f' float
};

struct Header {
  enum Format { kAscii, kLittle, kBig } format = kLittle;
  long long vertices = -1;
  std::vector<Prop> props;  // of the vertex element
  long long data_offset = 0;
  bool vertex_first = true;  // the vertex element is the first one
};

bool type_of(const std::string& t, Prop& p) {
  struct T {
    const char* a;
    const char* b;
    int n;
    char k;
  };
  static const T tab[] = {{"char", "int8", 1, 'i'},    {"uchar", "uint8", 1, 'u'},  {"short", "int16", 2, 'i'},
                          {"ushort", "uint16", 2, 'u'}, {"int", "int32", 4, 'i'},     {"uint", "uint32", 4, 'u'},
                          {"float", "float32", 4, 'f'}, {"double", "float64", 8, 'f'}};
  for (const T& e : tab)
    if (t == e.a || t == e.b) {
      p.bytes = e.n;
      p.kind = e.k;
      return true;
    }
  return false;
}

int parse_header(FILE* f, Header& h) {
  char line[1024];
  if (!std::fgets(line, sizeof(line), f) || std::strncmp(line, "ply", 3) != 0) return ply_fail("not a PLY file");
  bool in_vertex = false;
  int element_index = 0;
  while (std::fgets(line, sizeof(line), f)) {
    std::string s(line);
    while (!s.empty() && (s.back() == '\n' || s.back() == '\r' || s.back() == ' ')) s.pop_back();
    if (s == "end_header") {
      h.data_offset = std::ftell(f);
      if (h.vertices < 0) return ply_fail("no vertex element");
      return OMNIGS_OK;
    }
    char a[256] = {0}, b[256] = {0}, c[256] = {0}, d[256] = {0};
    const int n = std::sscanf(s.c_str(), "%255s %255s %255s %255s", a, b, c, d);
    if (n <= 0) continue;
    const std::string k(a);
    if (k == "format") {
      const std::string fm(b);
      if (fm == "ascii") h.format = Header::kAscii;
      else if (fm == "binary_little_endian") h.format = Header::kLittle;
      else if (fm == "binary_big_endian") h.format = Header::kBig;
      else return ply_fail("unknown PLY format " + fm);
    } else if (k == "element") {
      in_vertex = std::string(b) == "vertex";
      if (in_vertex) {
        h.vertices = std::atoll(c);
        h.vertex_first = element_index == 0;
      }
      ++element_index;
    } else if (k == "property" && in_vertex) {
      if (std::string(b) == "list") return ply_fail("list properties in the vertex element are not supported");
      Prop p;
      if (!type_of(b, p)) return ply_fail(std::string("unknown property type ") + b);
      p.name = c;
      h.props.push_back(p);
    }
  }
  return ply_fail("PLY header has no end_header");
}

double to_double(const unsigned char* q, const Prop& p, bool swap) {
  unsigned char t[8];
  for (int i = 0; i < p.bytes; ++i) t[i] = swap ? q[p.bytes - 1 - i] : q[i];
  switch (p.kind) {
    case 'f':
      if (p.bytes == 4) {
        float v;
        std::memcpy(&v, t, 4);
        return v;
      } else {
        double v;
        std::memcpy(&v, t, 8);
        return v;
      }
    case 'i':
      if (p.bytes == 1) return (double)(int8_t)t[0];
      if (p.bytes == 2) { int16_t v; std::memcpy(&v, t, 2); return v; }
      { int32_t v; std::memcpy(&v, t, 4); return v; }
    default:
      if (p.bytes == 1) return (double)t[0];
      if (p.bytes == 2) { uint16_t v; std::memcpy(&v, t, 2); return v; }
      { uint32_t v; std::memcpy(&v, t, 4); return v; }
  }
}

bool host_little_endian() {
  const uint16_t x = 1;
  unsigned char c;
  std::memcpy(&c, &x, 1);
  return c == 1;
}

// Property names of the layout, in file order.
std::vector<std::string> layout_names(int K) {
  std::vector<std::string> n = {"x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"};
  for (int i = 0; i < 3 * K; ++i) n.push_back("f_rest_" + std::to_string(i));
  n.push_back("opacity");
  for (int i = 0; i < 3; ++i) n.push_back("scale_" + std::to_string(i));
  for (int i = 0; i < 4; ++i) n.push_back("rot_" + std::to_string(i));
  return n;
}

}  // namespace

extern "C" {

const char* omnigs_ply_last_error(void) { return g_ply_error; }

int omnigs_ply_write(const char* path, long long P, int sh_degree, const float* xyz, const float* features_dc,
                     const float* features_rest, const float* opacity, const float* scaling, const float* rotation) {
  if (!path || P < 0 || sh_degree < 0 || sh_degree > 3) return ply_fail("bad path, P or sh_degree");
  const int K = (sh_degree + 1) * (sh_degree + 1) - 1;
  if (P > 0 && (!xyz || !features_dc || (K > 0 && !features_rest) || !opacity || !scaling || !rotation))
    return ply_fail("missing array");
  if (!host_little_endian()) return ply_fail("big-endian hosts are not supported");
  FILE* f = std::fopen(path, "wb");
  if (!f) return ply_fail(std::string("cannot open ") + path + ": " + std::strerror(errno));
  std::string hdr = "ply\nformat binary_little_endian 1.0\nelement vertex " + std::to_string(P) + "\n";
  for (const std::string& n : layout_names(K)) hdr += "property float " + n + "\n";
  hdr += "end_header\n";
  bool ok = std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
  const int row = 9 + 3 * K + 8;
  const long long chunk = 1 << 14;
  std::vector<float> buf((size_t)chunk * row);
  for (long long p0 = 0; ok && p0 < P; p0 += chunk) {
    const long long n = P - p0 < chunk ? P - p0 : chunk;
    for (long long i = 0; i < n; ++i) {
      const long long p = p0 + i;
      float* r = &buf[(size_t)i * row];
      for (int c = 0; c < 3; ++c) r[c] = xyz[3 * p + c];
      r[3] = r[4] = r[5] = 0.f;
      for (int c = 0; c < 3; ++c) r[6 + c] = features_dc[3 * p + c];
      for (int c = 0; c < 3; ++c)
        for (int j = 0; j < K; ++j) r[9 + c * K + j] = features_rest[((size_t)p * K + j) * 3 + c];
      r[9 + 3 * K] = opacity[p];
      for (int c = 0; c < 3; ++c) r[10 + 3 * K + c] = scaling[3 * p + c];
      for (int c = 0; c < 4; ++c) r[13 + 3 * K + c] = rotation[4 * p + c];
    }
    ok = std::fwrite(buf.data(), sizeof(float) * row, (size_t)n, f) == (size_t)n;
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? OMNIGS_OK : ply_fail(std::string("write failed: ") + path);
}

int omnigs_ply_count(const char* path, long long* P) {
  if (!path || !P) return ply_fail("NULL path or P");
  FILE* f = std::fopen(path, "rb");
  if (!f) return ply_fail(std::string("cannot open ") + path + ": " + std::strerror(errno));
  Header h;
  const int st = parse_header(f, h);
  std::fclose(f);
  if (st == OMNIGS_OK) *P = h.vertices;
  return st;
}

int omnigs_ply_read(const char* path, long long P, int sh_degree, float* xyz, float* features_dc,
                    float* features_rest, float* opacity, float* scaling, float* rotation) {
  if (!path || sh_degree < 0 || sh_degree > 3) return ply_fail("bad path or sh_degree");
  const int K = (sh_degree + 1) * (sh_degree + 1) - 1;
  FILE* f = std::fopen(path, "rb");
  if (!f) return ply_fail(std::string("cannot open ") + path + ": " + std::strerror(errno));
  Header h;
  int st = parse_header(f, h);
  if (st != OMNIGS_OK) {
    std::fclose(f);
    return st;
  }
  if (h.vertices != P) {
    std::fclose(f);
    return ply_fail("vertex count " + std::to_string(h.vertices) + " != P " + std::to_string(P));
  }
  if (!h.vertex_first && h.format != Header::kAscii) {
    std::fclose(f);
    return ply_fail("the vertex element must come first");
  }
  // destination of every property of the layout: (array, stride, offset)
  const std::vector<std::string> names = layout_names(K);
  std::vector<int> where(h.props.size(), -1);
  for (size_t n = 0; n < names.size(); ++n) {
    if (names[n][0] == 'n' && names[n].size() == 2) continue;  // normals are ignored
    size_t i = 0;
    while (i < h.props.size() && h.props[i].name != names[n]) ++i;
    if (i == h.props.size()) {
      std::fclose(f);
      return ply_fail("property " + names[n] + " missing");
    }
    where[i] = (int)n;
  }
  auto store = [&](long long p, int n, float v) {
    if (n < 3) xyz[3 * p + n] = v;
    else if (n >= 6 && n < 9) features_dc[3 * p + (n - 6)] = v;
    else if (n >= 9 && n < 9 + 3 * K) {
      const int k = n - 9, c = k / K, j = k % K;
      features_rest[((size_t)p * K + j) * 3 + c] = v;
    } else if (n == 9 + 3 * K) opacity[p] = v;
    else if (n < 13 + 3 * K) scaling[3 * p + (n - 10 - 3 * K)] = v;
    else rotation[4 * p + (n - 13 - 3 * K)] = v;
  };
  if (h.format == Header::kAscii) {
    for (long long p = 0; p < P; ++p)
      for (size_t i = 0; i < h.props.size(); ++i) {
        double v;
        if (std::fscanf(f, "%lf", &v) != 1) {
          std::fclose(f);
          return ply_fail("short ASCII PLY body");
        }
        if (where[i] >= 0) store(p, where[i], (float)v);
      }
  } else {
    const bool swap = (h.format == Header::kBig) == host_little_endian();
    std::vector<size_t> off(h.props.size());
    size_t row = 0;
    for (size_t i = 0; i < h.props.size(); ++i) {
      off[i] = row;
      row += (size_t)h.props[i].bytes;
    }
    std::fseek(f, (long)h.data_offset, SEEK_SET);
    const long long chunk = 1 << 14;
    std::vector<unsigned char> buf((size_t)chunk * row);
    for (long long p0 = 0; p0 < P; p0 += chunk) {
      const long long n = P - p0 < chunk ? P - p0 : chunk;
      if (std::fread(buf.data(), row, (size_t)n, f) != (size_t)n) {
        std::fclose(f);
        return ply_fail("short binary PLY body");
      }
      for (long long r = 0; r < n; ++r) {
        const unsigned char* q = &buf[(size_t)r * row];
        for (size_t i = 0; i < h.props.size(); ++i)
          if (where[i] >= 0) store(p0 + r, where[i], (float)to_double(q + off[i], h.props[i], swap));
      }
    }
  }
  std::fclose(f);
  return OMNIGS_OK;
}

}  // extern "C"
