// CPU oracle for the equirectangular (camera_type = 3) Gaussian rasterizer.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product path (omnigs_amd/, include/)
// links, loads or calls this code; tests/, the smoke check and the CPU
// baseline leg of bench.py use it as the checker.
//
// A restatement of the reference algorithm in scalar C++, templated over the
// arithmetic type (float = the reference's FP32 arithmetic, double = the FP64
// error yardstick).  Every function cites the reference lines it follows.
// glm (absent here) is restated from its documented semantics: mat3 is
// column-major, m[c][r] is column c / row r, and mat3(a..i) fills column 0
// with (a,b,c).  The evaluation order of the reference's float expressions is
// kept where it decides rounding (tile rects, ndc2Pix in FP64, key bits).
//
// Pinning: every stage has a check in tests/test_oracle.py that a plausible
// mistake would fail -- closed forms (projection, Jacobian, single Gaussian
// pixel value), brute-force binning, bg=0/bg=1 identity, FP32 vs FP64, and
// central finite differences of the FP64 oracle for every backward term.
// Conventions that no check can pin (only the reference text fixes them) are
// listed in tests/test_oracle.py::SOURCE_ONLY_CONVENTIONS.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#endif

namespace omnigs_oracle {

constexpr int kChannels = 3;  // config.h:25 NUM_CHANNELS

// Running rounding-error bound (Wilkinson's running error analysis, as in
// Higham, "Accuracy and Stability of Numerical Algorithms", 2nd ed., Sec. 3.3):
// v is an operation's result evaluated in FP64 (the exact value to first
// order) and e a first-order bound on the error that carrying out the same
// operations in FP32 accumulates, in units of the FP32 unit roundoff
// u = 2^-24.  Every operation propagates its operands' bounds and adds one
// rounding of its own result:
//   z = x +- y:  e = ex + ey + |z|          z = x y:  e = |y| ex + |x| ey + |z|
//   z = x / y:   e = (ex + |z| ey) / |y| + |z|
//   z = sqrt(x): e = ex / (2 z) + |z|       z = exp(x): e = |z| (ex + |x| + 2)
// (exp also covers an exp2 evaluation whose argument x log2(e) is rounded).
// Inputs and constants carry e = 0: both sides of a comparison read the same
// FP32 values.  Decisions compare v.  Used to instantiate the backward stages
// and give each gradient entry its own rounding scale (e u).
struct Rerr {
  double v = 0.0, e = 0.0;
  constexpr Rerr() = default;
  constexpr Rerr(double x) : v(x), e(0.0) {}
  constexpr Rerr(double x, double err) : v(x), e(err) {}
  Rerr& operator+=(Rerr o);
  Rerr& operator-=(Rerr o);
  Rerr& operator*=(Rerr o);
};
inline Rerr operator+(Rerr a, Rerr b) {
  const double z = a.v + b.v;
  return {z, a.e + b.e + std::fabs(z)};
}
inline Rerr operator-(Rerr a, Rerr b) {
  const double z = a.v - b.v;
  return {z, a.e + b.e + std::fabs(z)};
}
inline Rerr operator-(Rerr a) { return {-a.v, a.e}; }
inline Rerr operator*(Rerr a, Rerr b) {
  const double z = a.v * b.v;
  return {z, std::fabs(b.v) * a.e + std::fabs(a.v) * b.e + std::fabs(z)};
}
inline Rerr operator/(Rerr a, Rerr b) {
  const double z = a.v / b.v;
  return {z, (a.e + std::fabs(z) * b.e) / std::fabs(b.v) + std::fabs(z)};
}
inline Rerr& Rerr::operator+=(Rerr o) { return *this = *this + o; }
inline Rerr& Rerr::operator-=(Rerr o) { return *this = *this - o; }
inline Rerr& Rerr::operator*=(Rerr o) { return *this = *this * o; }
inline bool operator<(Rerr a, Rerr b) { return a.v < b.v; }
inline bool operator>(Rerr a, Rerr b) { return a.v > b.v; }
inline bool operator==(Rerr a, Rerr b) { return a.v == b.v; }
inline bool operator!=(Rerr a, Rerr b) { return a.v != b.v; }

// Math functions of the backward stages: std:: for float / double (the same
// overloads as before), the error-propagating forms for Rerr.
template <class R> inline R osqrt(R x) { return std::sqrt(x); }
template <class R> inline R oexp(R x) { return std::exp(x); }
template <class R> inline R onearbyint(R x) { return std::nearbyint(x); }
inline Rerr osqrt(Rerr x) {
  const double z = std::sqrt(x.v);
  return {z, (z > 0.0 ? x.e / (2.0 * z) : 0.0) + z};
}
inline Rerr oexp(Rerr x) {
  const double z = std::exp(x.v);
  return {z, z * (x.e + std::fabs(x.v) + 2.0)};
}
inline Rerr onearbyint(Rerr x) { return {std::nearbyint(x.v), 0.0}; }
template <class R> struct is_rerr : std::false_type {};
template <> struct is_rerr<Rerr> : std::true_type {};

// Real SH constants, auxiliary.h:32-49 (textbook values of the real basis).
template <class R> struct SH {
  static constexpr R C0 = R(0.28209479177387814);
  static constexpr R C1 = R(0.4886025119029199);
  static constexpr R C2[5] = {R(1.0925484305920792), R(-1.0925484305920792), R(0.31539156525252005),
                              R(-1.0925484305920792), R(0.5462742152960396)};
  static constexpr R C3[7] = {R(-0.5900435899266435), R(2.890611442640554), R(-0.4570457994644658),
                              R(0.3731763325901154), R(-0.4570457994644658), R(1.445305721320277),
                              R(-0.5900435899266435)};
};

// M_1_PIf32 / M_2_PIf32 as float literals in FP32, exact in FP64 mode.
template <class R> inline R inv_pi() { return R(0.318309886183790671537767526745028724); }
template <class R> inline R two_inv_pi() { return R(0.636619772367581343075535053490057448); }

template <class R> struct V3 {
  R x, y, z;
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// glm::mat3 restated: c[col][row].
template <class R> struct M3 {
  R c[3][3];
  static M3 cols(R a0, R a1, R a2, R b0, R b1, R b2, R d0, R d1, R d2) {
    M3 m;
    m.c[0][0] = a0; m.c[0][1] = a1; m.c[0][2] = a2;
    m.c[1][0] = b0; m.c[1][1] = b1; m.c[1][2] = b2;
    m.c[2][0] = d0; m.c[2][1] = d1; m.c[2][2] = d2;
    return m;
  }
};
// glm operator*(mat3, mat3): column i of the result = A0*B[i][0] + A1*B[i][1] + A2*B[i][2].
template <class R> inline M3<R> mul(const M3<R>& A, const M3<R>& B) {
  M3<R> o;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o.c[i][j] = (A.c[0][j] * B.c[i][0] + A.c[1][j] * B.c[i][1]) + A.c[2][j] * B.c[i][2];
  return o;
}
template <class R> inline M3<R> transpose(const M3<R>& A) {
  M3<R> o;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o.c[i][j] = A.c[j][i];
  return o;
}

// transformPoint4x3 / transformVec4x3Transpose, auxiliary.h:85-93, :116-124.
template <class R> inline V3<R> xform_point(const V3<R>& p, const R* m) {
  return {m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12], m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13],
          m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14]};
}
template <class R> inline V3<R> xform_vec_T(const V3<R>& p, const R* m) {
  return {m[0] * p.x + m[1] * p.y + m[2] * p.z, m[4] * p.x + m[5] * p.y + m[6] * p.z,
          m[8] * p.x + m[9] * p.y + m[10] * p.z};
}

// ndc2Pix, auxiliary.h:51-54: the double literals promote the expression to FP64.
template <class R> inline R ndc2pix(R v, int S) { return R(((double(v) + 1.0) * S - 1.0) * 0.5); }

struct Rect {
  uint32_t x0, y0, x1, y1;
};
// getRect, auxiliary.h:56-66, with its exact float evaluation order:
// (int)((p - r) / 16) and (int)((((p + r) + 16) - 1) / 16), clamped to [0, grid].
inline Rect tile_rect(float px, float py, int radius, int gx, int gy, int bx, int by) {
  auto lo = [](float p, int r, int b, int g) {
    int v = (int)((p - (float)r) / (float)b);
    v = std::max(0, v);
    return (uint32_t)std::min(g, v);
  };
  auto hi = [](float p, int r, int b, int g) {
    int v = (int)((((p + (float)r) + (float)b) - 1.0f) / (float)b);
    v = std::max(0, v);
    return (uint32_t)std::min(g, v);
  };
  return {lo(px, radius, bx, gx), lo(py, radius, by, gy), hi(px, radius, bx, gx), hi(py, radius, by, gy)};
}
inline Rect tile_rect(double px, double py, int radius, int gx, int gy, int bx, int by) {
  return tile_rect((float)px, (float)py, radius, gx, gy, bx, by);
}

// Seam-wrap extension (no reference counterpart; the reference clamps): the
// column range uses floor, may leave [0, gx) and is read modulo gx; a range
// wider than the ring becomes the ring.  Columns are returned as signed ints.
struct WRect {
  int x0, y0, x1, y1;
};
inline WRect tile_rect_wrap(float px, float py, int radius, int gx, int gy, int bx, int by) {
  Rect c = tile_rect(px, py, radius, gx, gy, bx, by);
  int x0 = (int)std::floor((px - (float)radius) / (float)bx);
  int x1 = (int)std::floor((((px + (float)radius) + (float)bx) - 1.0f) / (float)bx);
  if (x1 - x0 >= gx) {
    x0 = 0;
    x1 = gx;
  }
  return {x0, (int)c.y0, x1, (int)c.y1};
}
inline uint32_t wrap_col(int x, int gx) { return (uint32_t)(((x % gx) + gx) % gx); }
template <class R> inline R seam_dx(R dx, int W) { return dx - R(W) * onearbyint(dx / R(W)); }

// getHigherMsb, rasterizer_impl.cu:47-62 (bit count of n: highest set bit + 1).
inline uint32_t higher_msb(uint32_t n) {
  uint32_t msb = sizeof(n) * 4, step = msb;
  while (step > 1) {
    step /= 2;
    if (n >> msb) msb += step; else msb -= step;
  }
  if (n >> msb) msb++;
  return msb;
}

template <class R> struct Inputs {
  int P = 0, D = 0, M = 0, W = 0, H = 0;
  int block_x = 16, block_y = 16;  // config.h:26-27
  int seam_wrap = 0;                // extension: 1 = wrap tile columns and dx around the seam
  const R* bg = nullptr;            // [3]
  const R* means = nullptr;         // [P,3]
  const R* shs = nullptr;           // [P,M,3] or null
  const R* colors = nullptr;        // [P,3] or null
  const R* opac = nullptr;          // [P]
  const R* scales = nullptr;        // [P,3] or null
  R mod = R(1);
  const R* rots = nullptr;          // [P,4] (w,x,y,z) or null
  const R* cov3d_pre = nullptr;     // [P,6] or null
  const R* view = nullptr;          // [16], Tcw^T column-major
  const R* campos = nullptr;        // [3]
  // > 1: the OpenMP CPU baseline -- Gaussians, pixel rows and tiles in
  // parallel, a parallel stable sort (same result), and the render backward
  // reduced per Gaussian in sorted-instance order (the same sums for any
  // thread count, rounded in another order than the serial oracle).  No
  // decision signature is recorded in this mode.
  int threads = 1;
};

template <class R> struct State {
  // geometry (rasterizer_impl.h:48-63)
  std::vector<int32_t> radii;
  std::vector<R> xy, depth, cov3d, rgb, conic;  // conic: (A,B,C,opacity)
  std::vector<uint8_t> clamped;
  std::vector<uint32_t> tiles_touched, offsets;
  // binning (rasterizer_impl.h:80-90)
  uint32_t R_ = 0;
  std::vector<uint64_t> keys;
  std::vector<uint32_t> point_list;
  // image (rasterizer_impl.h:68-75); ranges sized per tile
  std::vector<uint32_t> ranges;  // [tiles,2]
  std::vector<R> final_T, out_color;
  std::vector<uint32_t> n_contrib;
  // per pixel: smallest relative distance of any alpha >= 1/255 or T >= 1e-4
  // decision to its threshold (a decision closer than float rounding to its
  // threshold may legitimately differ between two correct FP32 evaluations)
  std::vector<R> margin;
  uint64_t decision_sig = 0;  // hash of every discrete decision (FD masking)
  // per-Gaussian and per-pixel blend signatures of the render: gsig[g] is
  // the XOR over the pixels that blend Gaussian g of a hash of the pixel;
  // psig[pixel] hashes the sequence of Gaussians the pixel blends.  Only
  // blended pairs enter: a pair skipped (power > 0, alpha < 1/255, after the
  // stop) or absent from the tile list contributes nothing to any output.
  // Two evaluations whose signatures agree blended the same pairs for that
  // Gaussian or pixel.
  std::vector<uint64_t> gsig, psig;
  // backward
  std::vector<R> dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors;  // [P,3], [P,4], [P], [P,3]
  std::vector<R> dL_dmeans3D, dL_dcov3D, dL_dsh, dL_dscales, dL_drot, dpx_dt, dpy_dt;
  // running-error mode (R = Rerr) only: Gaussians blended in a pixel where a
  // render-backward decision (power > 0, alpha >= 1/255) lies within rounding
  // of its threshold, so that another correct FP32 evaluation may decide it
  // the other way (for information: the error bounds already carry the
  // effect of such a flip -- the pair's own terms, and the change of T and of
  // the colour behind for every instance in front of it)
  std::vector<uint8_t> fragile;
  uint64_t fragile_pixels = 0;
};

inline void sig_mix(uint64_t& h, uint64_t v) {
  h ^= v + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
}
// splitmix64 finaliser (Steele, Lea, Flood 2014), for order-independent XOR signatures
inline uint64_t sig_hash(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

// computeCov3D, forward.cu:194-228: Sigma = (S R)^T (S R), q = (r,x,y,z) used unnormalised.
template <class R> void cov3d_forward(const R* s, R mod, const R* q, R* out) {
  M3<R> S = M3<R>::cols(1, 0, 0, 0, 1, 0, 0, 0, 1);
  S.c[0][0] = mod * s[0];
  S.c[1][1] = mod * s[1];
  S.c[2][2] = mod * s[2];
  R r = q[0], x = q[1], y = q[2], z = q[3];
  M3<R> Rm = M3<R>::cols(R(1) - R(2) * (y * y + z * z), R(2) * (x * y - r * z), R(2) * (x * z + r * y),
                         R(2) * (x * y + r * z), R(1) - R(2) * (x * x + z * z), R(2) * (y * z - r * x),
                         R(2) * (x * z - r * y), R(2) * (y * z + r * x), R(1) - R(2) * (x * x + y * y));
  M3<R> Mm = mul(S, Rm);
  M3<R> Sig = mul(transpose(Mm), Mm);
  out[0] = Sig.c[0][0]; out[1] = Sig.c[0][1]; out[2] = Sig.c[0][2];
  out[3] = Sig.c[1][1]; out[4] = Sig.c[1][2]; out[5] = Sig.c[2][2];
}

// Lon/lat Jacobian pieces, forward.cu:135-155 and backward.cu:339-360.
template <class R> struct LonlatJac {
  R dpx_dtx, dpx_dtz, dpy_dtx, dpy_dty, dpy_dtz;
  R W_div_2pi, H_div_pi;
};
template <class R> LonlatJac<R> lonlat_jacobian(V3<R> t, int W, int H) {
  const R eps = R(1e-7);
  R rxz2 = t.x * t.x + t.z * t.z;
  R rxz2_inv = R(1) / (rxz2 + eps);
  R rxz = osqrt(rxz2);
  R rxz_inv = R(1) / (rxz + eps);
  R r2 = rxz2 + t.y * t.y;
  R r2_inv = R(1) / (r2 + eps);
  LonlatJac<R> j;
  j.W_div_2pi = R(W) * R(0.5) * inv_pi<R>();
  j.H_div_pi = R(H) * inv_pi<R>();
  j.dpx_dtx = j.W_div_2pi * t.z * rxz2_inv;
  j.dpx_dtz = -j.W_div_2pi * t.x * rxz2_inv;
  j.dpy_dtx = -j.H_div_pi * t.x * t.y * rxz_inv * r2_inv;
  j.dpy_dty = j.H_div_pi * rxz * r2_inv;
  j.dpy_dtz = -j.H_div_pi * t.z * t.y * rxz_inv * r2_inv;
  return j;
}

// computeCov2DLonlat, forward.cu:130-189: cov = T^T Vrk^T T (+0.3 I), T = W J (glm).
template <class R> V3<R> cov2d_lonlat(V3<R> mean, int W, int H, const R* c3, const R* view) {
  V3<R> t = xform_point(mean, view);
  LonlatJac<R> j = lonlat_jacobian(t, W, H);
  M3<R> J = M3<R>::cols(j.dpx_dtx, 0, j.dpx_dtz, j.dpy_dtx, j.dpy_dty, j.dpy_dtz, 0, 0, 0);
  M3<R> Wm = M3<R>::cols(view[0], view[4], view[8], view[1], view[5], view[9], view[2], view[6], view[10]);
  M3<R> T = mul(Wm, J);
  M3<R> V = M3<R>::cols(c3[0], c3[1], c3[2], c3[1], c3[3], c3[4], c3[2], c3[4], c3[5]);
  M3<R> cov = mul(mul(transpose(T), transpose(V)), T);
  cov.c[0][0] += R(0.3);
  cov.c[1][1] += R(0.3);
  return {cov.c[0][0], cov.c[0][1], cov.c[1][1]};
}

// computeColorFromSH (forward), forward.cu:30-83.
template <class R> V3<R> sh_to_rgb(int deg, int M, const R* sh_g, V3<R> pos, V3<R> campos, uint8_t* clamped3) {
  using K = SH<R>;
  V3<R> dir = pos - campos;
  R len = std::sqrt(dot(dir, dir));
  dir = {dir.x / len, dir.y / len, dir.z / len};
  auto s = [&](int i) { return V3<R>{sh_g[3 * i], sh_g[3 * i + 1], sh_g[3 * i + 2]}; };
  V3<R> res = K::C0 * s(0);
  if (deg > 0) {
    R x = dir.x, y = dir.y, z = dir.z;
    res = ((res - (K::C1 * y) * s(1)) + (K::C1 * z) * s(2)) - (K::C1 * x) * s(3);
    if (deg > 1) {
      R xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
      res = res + ((((K::C2[0] * xy) * s(4) + (K::C2[1] * yz) * s(5)) +
                    (K::C2[2] * (R(2) * zz - xx - yy)) * s(6)) +
                   (K::C2[3] * xz) * s(7)) +
            (K::C2[4] * (xx - yy)) * s(8);
      if (deg > 2) {
        res = res + ((((((K::C3[0] * y * (R(3) * xx - yy)) * s(9) + (K::C3[1] * xy * z) * s(10)) +
                        (K::C3[2] * y * (R(4) * zz - xx - yy)) * s(11)) +
                       (K::C3[3] * z * (R(2) * zz - R(3) * xx - R(3) * yy)) * s(12)) +
                      (K::C3[4] * x * (R(4) * zz - xx - yy)) * s(13)) +
                     (K::C3[5] * z * (xx - yy)) * s(14)) +
              (K::C3[6] * x * (xx - R(3) * yy)) * s(15);
      }
    }
  }
  res = res + V3<R>{R(0.5), R(0.5), R(0.5)};
  clamped3[0] = res.x < 0;
  clamped3[1] = res.y < 0;
  clamped3[2] = res.z < 0;
  return {std::max(res.x, R(0)), std::max(res.y, R(0)), std::max(res.z, R(0))};
}

// preprocessLonlatCUDA, forward.cu:593-703 (too_close auxiliary.h:198-220,
// point3ToLonlatScreen auxiliary.h:236-248).
template <class R> void preprocess(const Inputs<R>& in, State<R>& st) {
  const int P = in.P;
  const int gx = (in.W + in.block_x - 1) / in.block_x, gy = (in.H + in.block_y - 1) / in.block_y;
  st.radii.assign(P, 0);
  st.xy.assign(2 * P, 0);
  st.depth.assign(P, 0);
  st.cov3d.assign(6 * P, 0);
  st.rgb.assign(3 * P, 0);
  st.conic.assign(4 * P, 0);
  st.clamped.assign(3 * P, 0);
  st.tiles_touched.assign(P, 0);
  V3<R> campos{in.campos[0], in.campos[1], in.campos[2]};
#pragma omp parallel for schedule(static) if (in.threads > 1) num_threads(in.threads)
  for (int i = 0; i < P; ++i) {
    V3<R> p{in.means[3 * i], in.means[3 * i + 1], in.means[3 * i + 2]};
    V3<R> t = xform_point(p, in.view);
    R rr = t.x * t.x + t.y * t.y + t.z * t.z;
    if (rr <= R(0.04)) continue;
    R r = std::sqrt(rr);
    R inv_r = R(1) / (r + R(1e-7));
    R lon = std::atan2(t.x, t.z);
    R lat = std::asin(t.y * inv_r);
    R sx = lon * inv_pi<R>(), sy = lat * two_inv_pi<R>();
    const R* c3;
    if (in.cov3d_pre) {
      c3 = in.cov3d_pre + 6 * i;
    } else {
      cov3d_forward(in.scales + 3 * i, in.mod, in.rots + 4 * i, &st.cov3d[6 * i]);
      c3 = &st.cov3d[6 * i];
    }
    V3<R> cov = cov2d_lonlat(p, in.W, in.H, c3, in.view);
    R det = cov.x * cov.z - cov.y * cov.y;
    if (det == R(0)) continue;
    R det_inv = R(1) / det;
    V3<R> conic{cov.z * det_inv, -cov.y * det_inv, cov.x * det_inv};
    R mid = R(0.5) * (cov.x + cov.z);
    R l1 = mid + std::sqrt(std::max(R(0.1), mid * mid - det));
    R l2 = mid - std::sqrt(std::max(R(0.1), mid * mid - det));
    R radius_f = std::ceil(R(3) * std::sqrt(std::max(l1, l2)));
    int radius = (int)radius_f;
    R px = ndc2pix(sx, in.W), py = ndc2pix(sy, in.H);
    Rect rc = tile_rect(px, py, radius, gx, gy, in.block_x, in.block_y);
    if (in.seam_wrap) {
      WRect w = tile_rect_wrap((float)px, (float)py, radius, gx, gy, in.block_x, in.block_y);
      rc = {0, (uint32_t)w.y0, (uint32_t)(w.x1 - w.x0), (uint32_t)w.y1};  // width only
    }
    if ((rc.x1 - rc.x0) * (rc.y1 - rc.y0) == 0) continue;
    if (!in.colors) {
      V3<R> c = sh_to_rgb(in.D, in.M, in.shs + (size_t)3 * in.M * i, p, campos, &st.clamped[3 * i]);
      st.rgb[3 * i] = c.x; st.rgb[3 * i + 1] = c.y; st.rgb[3 * i + 2] = c.z;
    }
    st.depth[i] = r;
    st.radii[i] = radius;
    st.xy[2 * i] = px; st.xy[2 * i + 1] = py;
    st.conic[4 * i] = conic.x; st.conic[4 * i + 1] = conic.y; st.conic[4 * i + 2] = conic.z;
    st.conic[4 * i + 3] = in.opac[i];
    st.tiles_touched[i] = (rc.y1 - rc.y0) * (rc.x1 - rc.x0);
  }
}

// Depth bits used in the sort key: the float bit pattern of the FP32 depth
// (duplicateWithKeys, rasterizer_impl.cu:133).  The FP64 oracle keys on the
// depth rounded to float, which keeps both precisions' orders comparable.
template <class R> inline uint32_t depth_bits(R d) {
  float f = (float)d;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// Inclusive scan + duplicateWithKeys + stable radix sort + identifyTileRanges:
// rasterizer_impl.cu:94-167, :622-679.  CUB's LSD radix sort is stable, so the
// sort is std::stable_sort on the 64-bit key over bits [0, 32 + getHigherMsb(tiles)).
template <class R> void binning(const Inputs<R>& in, State<R>& st) {
  const int P = in.P;
  const int gx = (in.W + in.block_x - 1) / in.block_x, gy = (in.H + in.block_y - 1) / in.block_y;
  const int tiles = gx * gy;
  st.offsets.resize(P);
  uint32_t acc = 0;
  for (int i = 0; i < P; ++i) st.offsets[i] = (acc += st.tiles_touched[i]);
  st.R_ = P ? st.offsets[P - 1] : 0;
  std::vector<uint64_t> keys(st.R_);
  std::vector<uint32_t> vals(st.R_);
#pragma omp parallel for schedule(dynamic, 1024) if (in.threads > 1) num_threads(in.threads)
  for (int i = 0; i < P; ++i) {
    if (st.radii[i] <= 0) continue;
    uint32_t off = i == 0 ? 0 : st.offsets[i - 1];
    WRect rc;
    if (in.seam_wrap) {
      rc = tile_rect_wrap((float)st.xy[2 * i], (float)st.xy[2 * i + 1], st.radii[i], gx, gy, in.block_x, in.block_y);
    } else {
      Rect c = tile_rect(st.xy[2 * i], st.xy[2 * i + 1], st.radii[i], gx, gy, in.block_x, in.block_y);
      rc = {(int)c.x0, (int)c.y0, (int)c.x1, (int)c.y1};
    }
    for (int y = rc.y0; y < rc.y1; ++y)
      for (int x = rc.x0; x < rc.x1; ++x) {
        const uint32_t col = in.seam_wrap ? wrap_col(x, gx) : (uint32_t)x;
        keys[off] = ((uint64_t)((uint32_t)y * gx + col) << 32) | depth_bits(st.depth[i]);
        vals[off] = (uint32_t)i;
        ++off;
      }
  }
  const uint32_t end_bit = 32 + higher_msb((uint32_t)tiles);
  const uint64_t mask = end_bit >= 64 ? ~0ULL : ((1ULL << end_bit) - 1);
  std::vector<uint32_t> perm(st.R_);
  std::iota(perm.begin(), perm.end(), 0u);
  auto by_key = [&](uint32_t a, uint32_t b) { return (keys[a] & mask) < (keys[b] & mask); };
#ifdef _OPENMP
  if (in.threads > 1) {
    omp_set_num_threads(in.threads);
    __gnu_parallel::stable_sort(perm.begin(), perm.end(), by_key);
  } else
#endif
    std::stable_sort(perm.begin(), perm.end(), by_key);
  st.keys.resize(st.R_);
  st.point_list.resize(st.R_);
  for (uint32_t k = 0; k < st.R_; ++k) {
    st.keys[k] = keys[perm[k]];
    st.point_list[k] = vals[perm[k]];
  }
  st.ranges.assign(2 * (size_t)tiles, 0);
  for (uint32_t k = 0; k < st.R_; ++k) {
    uint32_t t = (uint32_t)(st.keys[k] >> 32);
    if (k == 0) st.ranges[2 * t] = 0;
    else {
      uint32_t pt = (uint32_t)(st.keys[k - 1] >> 32);
      if (pt != t) {
        st.ranges[2 * pt + 1] = k;
        st.ranges[2 * t] = k;
      }
    }
    if (k == st.R_ - 1) st.ranges[2 * t + 1] = st.R_;
  }
}

// renderCUDA (forward), forward.cu:346-467: per pixel front-to-back blending
// over its tile's list; identical per-pixel semantics (the block-level vote only
// ends the loop once every pixel of the tile is done, which does not change
// any pixel's result).
template <class R> void render(const Inputs<R>& in, State<R>& st) {
  const int W = in.W, H = in.H, bx = in.block_x, by = in.block_y;
  const int gx = (W + bx - 1) / bx;
  const size_t N = (size_t)W * H;
  const R* feat = in.colors ? in.colors : st.rgb.data();
  st.final_T.assign(N, 0);
  st.n_contrib.assign(N, 0);
  st.out_color.assign(kChannels * N, 0);
  st.margin.assign(N, R(1));
  st.gsig.assign(in.P, 0);
  st.psig.assign(N, 0);
  uint64_t sig = 1469598103934665603ULL;
  const bool par = in.threads > 1;
  sig_mix(sig, st.R_);
#pragma omp parallel for schedule(dynamic, 1) if (par) num_threads(in.threads)
  for (int py = 0; py < H; ++py)
    for (int px = 0; px < W; ++px) {
      const size_t pix = (size_t)py * W + px;
      uint64_t ps = 1469598103934665603ULL;
      uint32_t cur = 0;  // the Gaussian of the pair being decided
      auto mix = [&](uint64_t v) {
        if (!par) sig_mix(sig, v);
        if (v >= 4) {  // blended (1: power > 0, 2: alpha < 1/255, 3: stop)
          __atomic_fetch_xor(&st.gsig[cur], sig_hash(pix), __ATOMIC_RELAXED);
          sig_mix(ps, cur);
        }
      };
      const int tile = (py / by) * gx + (px / bx);
      const uint32_t b = st.ranges[2 * tile], e = st.ranges[2 * tile + 1];
      const R fx = (R)px, fy = (R)py;
      R T = 1, C[kChannels] = {0, 0, 0};
      uint32_t contributor = 0, last = 0;
      R mg = 1;
      for (uint32_t k = b; k < e; ++k) {
        ++contributor;
        const uint32_t g = st.point_list[k];
        cur = g;
        R dx = st.xy[2 * g] - fx;
        const R dy = st.xy[2 * g + 1] - fy;
        if (in.seam_wrap) dx = seam_dx(dx, W);
        const R* co = &st.conic[4 * g];
        const R power = R(-0.5) * (co[0] * dx * dx + co[2] * dy * dy) - co[1] * dx * dy;
        if (power > R(0)) { mix(1); continue; }
        const R alpha = std::min(R(0.99), co[3] * std::exp(power));
        mg = std::min(mg, std::abs(alpha * R(255) - R(1)));
        if (alpha < R(1) / R(255)) { mix(2); continue; }
        const R test_T = T * (R(1) - alpha);
        mg = std::min(mg, std::abs(test_T * R(10000) - R(1)));
        if (test_T < R(0.0001)) { mix(3); break; }
        for (int ch = 0; ch < kChannels; ++ch) C[ch] += feat[kChannels * g + ch] * alpha * T;
        T = test_T;
        last = contributor;
        mix(4 + k);
      }
      const size_t pid = (size_t)py * W + px;
      st.final_T[pid] = T;
      st.n_contrib[pid] = last;
      st.margin[pid] = mg;
      st.psig[pid] = ps;
      for (int ch = 0; ch < kChannels; ++ch) st.out_color[ch * N + pid] = C[ch] + T * in.bg[ch];
    }
  for (int i = 0; i < in.P; ++i) sig_mix(sig, (uint64_t)st.radii[i] * 131 + st.tiles_touched[i]);
  st.decision_sig = par ? 0 : sig;
}

template <class R> void forward(const Inputs<R>& in, State<R>& st) {
  preprocess(in, st);
  binning(in, st);
  render(in, st);
}

// renderCUDA (backward), backward.cu:672-843: reverse traversal per pixel with
// T recovered by division; gradients accumulated per Gaussian (the reference's
// atomics summed in pixel order).  dL_dmean2D is in NDC units (x 0.5W / 0.5H,
// backward.cu:739-740) and the conic .y slot holds half the off-diagonal term.
template <class R> void render_backward(const Inputs<R>& in, State<R>& st, const R* dL_dpix) {
  const int W = in.W, H = in.H, bx = in.block_x, by = in.block_y, P = in.P;
  const int gx = (W + bx - 1) / bx;
  const size_t N = (size_t)W * H;
  const R* col = in.colors ? in.colors : st.rgb.data();
  st.dL_dmean2D.assign(3 * (size_t)P, 0);
  st.dL_dconic.assign(4 * (size_t)P, 0);
  st.dL_dopacity.assign(P, 0);
  st.dL_dcolors.assign(3 * (size_t)P, 0);
  st.fragile.assign(is_rerr<R>::value ? P : 0, 0);
  st.fragile_pixels = 0;
  const R ddelx_dx = R(0.5 * W), ddely_dy = R(0.5 * H);
  // One pixel's reverse traversal; sink(k, g, vals[9]) receives the
  // contributions of sorted instance k (Gaussian g): dcolor x3, dmean2D x2,
  // dconic (0, 1, 3) x3, dopacity.
  auto pixel = [&](int px, int py, auto&& sink) {
    const int tile = (py / by) * gx + (px / bx);
    const uint32_t b = st.ranges[2 * tile], e = st.ranges[2 * tile + 1];
    const size_t pid = (size_t)py * W + px;
    const R fx = (R)px, fy = (R)py;
    const R T_final = st.final_T[pid];
    R T = T_final;
    const uint32_t last = st.n_contrib[pid];
    R dpix[kChannels], accum[kChannels] = {0, 0, 0}, last_color[kChannels] = {0, 0, 0};
    for (int ch = 0; ch < kChannels; ++ch) dpix[ch] = dL_dpix[ch * N + pid];
    R last_alpha = 0;
    R bg_dot = 0;
    for (int ch = 0; ch < kChannels; ++ch) bg_dot += in.bg[ch] * dpix[ch];
    bool frag = false;
    double pend[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // error bounds of an ambiguous pair blended here
    bool have_pend = false;
    // contributor index k - b (0-based); only those before last are blended.
    for (uint32_t kk = e; kk > b; --kk) {
      const uint32_t k = kk - 1;
      if (k - b >= last) continue;
      const uint32_t g = st.point_list[k];
      R dx = st.xy[2 * g] - fx;
      const R dy = st.xy[2 * g + 1] - fy;
      if (in.seam_wrap) dx = seam_dx(dx, W);
      const R* co = &st.conic[4 * g];
      const R power = R(-0.5) * (co[0] * dx * dx + co[2] * dy * dy) - co[1] * dx * dy;
      bool ambiguous = false;
      if constexpr (is_rerr<R>::value) {
        // a decision within rounding of its threshold may go the other way in
        // another correct FP32 evaluation (power > 0 near the mean, alpha
        // near 1/255): its possible effect enters the error bounds below
        constexpr double u = 0x1p-24;
        if (std::fabs(power.v) <= 8.0 * power.e * u + 64.0 * u) ambiguous = true;
        const R oG = co[3] * oexp(power);
        const double thr = 1.0 / 255.0;
        if (std::fabs(oG.v - thr) <= 8.0 * oG.e * u + thr * u * (64.0 + 8.0 * std::fabs(power.v))) ambiguous = true;
        if (ambiguous) frag = true;
      }
      const bool skip = power > R(0) || std::min(R(0.99), co[3] * oexp(power)) < R(1) / R(255);
      if (skip && !ambiguous) continue;
      if constexpr (is_rerr<R>::value) {
        if (ambiguous) {
          // the pair as if blended (G at most 1, alpha at least the threshold)
          constexpr double u = 0x1p-24;
          const double Gv = std::min(1.0, std::exp(std::min(0.0, power.v)));
          const double av = std::max(1.0 / 255.0, std::min(0.99, co[3].v * Gv));
          const double Th = T.v / (1.0 - av), r = av / (1.0 - av);
          double dLa = 0.0;
          for (int ch = 0; ch < kChannels; ++ch) {
            const double c = col[kChannels * g + ch].v;
            const double acc_h = last_alpha.v * last_color[ch].v + (1.0 - last_alpha.v) * accum[ch].v;
            dLa += std::fabs((c - acc_h) * dpix[ch].v);
          }
          dLa = dLa * Th + std::fabs(T_final.v / (1.0 - av) * bg_dot.v);
          const double dG = co[3].v * dLa, gdx = Gv * std::fabs(dx.v), gdy = Gv * std::fabs(dy.v);
          R vh[9];
          for (int ch = 0; ch < kChannels; ++ch) vh[ch] = R(0.0, av * Th * std::fabs(dpix[ch].v) / u);
          vh[3] = R(0.0, dG * (gdx * std::fabs(co[0].v) + gdy * std::fabs(co[1].v)) * std::fabs(ddelx_dx.v) / u);
          vh[4] = R(0.0, dG * (gdy * std::fabs(co[2].v) + gdx * std::fabs(co[1].v)) * std::fabs(ddely_dy.v) / u);
          vh[5] = R(0.0, 0.5 * gdx * std::fabs(dx.v) * dG / u);
          vh[6] = R(0.0, 0.5 * gdx * std::fabs(dy.v) * dG / u);
          vh[7] = R(0.0, 0.5 * gdy * std::fabs(dy.v) * dG / u);
          vh[8] = R(0.0, Gv * dLa / u);
          // the instances in front see T and the colour behind them change by
          // up to alpha / (1 - alpha) and alpha (c - acc)
          T.e += std::fabs(T.v) * r / u;
          for (int ch = 0; ch < kChannels; ++ch)
            accum[ch].e += av * (std::fabs(col[kChannels * g + ch].v) + std::fabs(accum[ch].v)) / u;
          if (skip) {
            sink(k, g, vh);
            continue;
          }
          // blended here: its own terms may be absent on the other side
          for (int j = 0; j < 9; ++j) pend[j] = vh[j].e;
          have_pend = true;
        }
      }
      const R G = oexp(power);
      const R alpha = std::min(R(0.99), co[3] * G);
      T = T / (R(1) - alpha);
      const R dch_dcolor = alpha * T;
      R dL_dalpha = 0;
      R v[9];
      for (int ch = 0; ch < kChannels; ++ch) {
        const R c = col[kChannels * g + ch];
        accum[ch] = last_alpha * last_color[ch] + (R(1) - last_alpha) * accum[ch];
        last_color[ch] = c;
        dL_dalpha += (c - accum[ch]) * dpix[ch];
        v[ch] = dch_dcolor * dpix[ch];
      }
      dL_dalpha *= T;
      last_alpha = alpha;
      dL_dalpha += (-T_final / (R(1) - alpha)) * bg_dot;
      const R dL_dG = co[3] * dL_dalpha;
      const R gdx = G * dx, gdy = G * dy;
      const R dG_ddelx = -gdx * co[0] - gdy * co[1];
      const R dG_ddely = -gdy * co[2] - gdx * co[1];
      v[3] = dL_dG * dG_ddelx * ddelx_dx;
      v[4] = dL_dG * dG_ddely * ddely_dy;
      v[5] = R(-0.5) * gdx * dx * dL_dG;
      v[6] = R(-0.5) * gdx * dy * dL_dG;
      v[7] = R(-0.5) * gdy * dy * dL_dG;
      v[8] = G * dL_dalpha;
      if constexpr (is_rerr<R>::value) {
        if (have_pend)
          for (int j = 0; j < 9; ++j) v[j].e += pend[j];
        have_pend = false;
      }
      sink(k, g, v);
    }
    if constexpr (is_rerr<R>::value) {
      if (frag) {
        for (uint32_t k = b; k < e && k - b < last; ++k) __atomic_store_n(&st.fragile[st.point_list[k]], 1, __ATOMIC_RELAXED);
        __atomic_fetch_add(&st.fragile_pixels, 1, __ATOMIC_RELAXED);
      }
    }
  };
  auto add_to_gaussian = [&](uint32_t g, const R* v) {
    for (int ch = 0; ch < kChannels; ++ch) st.dL_dcolors[3 * g + ch] += v[ch];
    st.dL_dmean2D[3 * g] += v[3];
    st.dL_dmean2D[3 * g + 1] += v[4];
    st.dL_dconic[4 * g] += v[5];
    st.dL_dconic[4 * g + 1] += v[6];
    st.dL_dconic[4 * g + 3] += v[7];
    st.dL_dopacity[g] += v[8];
  };
  if (in.threads <= 1) {
    for (int py = 0; py < H; ++py)
      for (int px = 0; px < W; ++px) pixel(px, py, [&](uint32_t, uint32_t g, const R* v) { add_to_gaussian(g, v); });
    return;
  }
  // Tile-parallel: every sorted instance belongs to one tile, so per-instance
  // partial sums have a single writer; then each Gaussian adds its instances
  // in sorted order.
  const int gy = (H + by - 1) / by;
  std::vector<R> part((size_t)st.R_ * 9, R(0));
#pragma omp parallel for schedule(dynamic, 1) num_threads(in.threads)
  for (int t = 0; t < gx * gy; ++t) {
    const int tx = t % gx, ty = t / gx;
    for (int py = ty * by; py < std::min(H, (ty + 1) * by); ++py)
      for (int px = tx * bx; px < std::min(W, (tx + 1) * bx); ++px)
        pixel(px, py, [&](uint32_t k, uint32_t, const R* v) {
          for (int j = 0; j < 9; ++j) part[(size_t)k * 9 + j] += v[j];
        });
  }
  std::vector<uint32_t> first(P + 1, 0), idx(st.R_);
  for (uint32_t k = 0; k < st.R_; ++k) ++first[st.point_list[k] + 1];
  for (int g = 0; g < P; ++g) first[g + 1] += first[g];
  {
    std::vector<uint32_t> fill(first.begin(), first.end() - 1);
    for (uint32_t k = 0; k < st.R_; ++k) idx[fill[st.point_list[k]]++] = k;
  }
#pragma omp parallel for schedule(static) num_threads(in.threads)
  for (int g = 0; g < P; ++g)
    for (uint32_t j = first[g]; j < first[g + 1]; ++j) add_to_gaussian((uint32_t)g, &part[(size_t)idx[j] * 9]);
}

// computeCov2DLonLatCUDA, backward.cu:297-485: dL/dconic -> dL/dcov3D and the
// covariance part of dL/dmean (assigned), plus the dpx_dt / dpy_dt rows.
template <class R> void cov2d_backward(const Inputs<R>& in, State<R>& st) {
  const int P = in.P;
  st.dL_dmeans3D.assign(3 * (size_t)P, 0);
  st.dL_dcov3D.assign(6 * (size_t)P, 0);
  st.dpx_dt.assign(3 * (size_t)P, 0);
  st.dpy_dt.assign(3 * (size_t)P, 0);
  const R* view = in.view;
#pragma omp parallel for schedule(static) if (in.threads > 1) num_threads(in.threads)
  for (int i = 0; i < P; ++i) {
    if (!(st.radii[i] > 0)) continue;
    const R* c3 = in.cov3d_pre ? in.cov3d_pre + 6 * i : &st.cov3d[6 * i];
    V3<R> mean{in.means[3 * i], in.means[3 * i + 1], in.means[3 * i + 2]};
    V3<R> dcon{st.dL_dconic[4 * i], st.dL_dconic[4 * i + 1], st.dL_dconic[4 * i + 3]};
    V3<R> t = xform_point(mean, view);
    const R txtx = t.x * t.x, tyty = t.y * t.y, tztz = t.z * t.z, txtytz = t.x * t.y * t.z;
    const R rxz2 = txtx + tztz;
    const R rxz2_inv = R(1) / (rxz2 + R(1e-7));
    const R rxz4_inv = rxz2_inv * rxz2_inv;
    const R rxz = osqrt(rxz2);
    const R rxz_inv = R(1) / (rxz + R(1e-7));
    const R r2 = rxz2 + tyty;
    const R r2_inv = R(1) / (r2 + R(1e-7));
    const R r4_inv = r2_inv * r2_inv;
    const R rxz_r4_inv = rxz_inv * r4_inv;
    const R rxz3_r4_inv = rxz2_inv * rxz_r4_inv;
    const R yy_minus_rxz2 = tyty - rxz2;
    LonlatJac<R> j = lonlat_jacobian(t, in.W, in.H);
    st.dpx_dt[3 * i] = j.dpx_dtx; st.dpx_dt[3 * i + 1] = 0; st.dpx_dt[3 * i + 2] = j.dpx_dtz;
    st.dpy_dt[3 * i] = j.dpy_dtx; st.dpy_dt[3 * i + 1] = j.dpy_dty; st.dpy_dt[3 * i + 2] = j.dpy_dtz;
    M3<R> J = M3<R>::cols(j.dpx_dtx, 0, j.dpx_dtz, j.dpy_dtx, j.dpy_dty, j.dpy_dtz, 0, 0, 0);
    M3<R> Wm = M3<R>::cols(view[0], view[4], view[8], view[1], view[5], view[9], view[2], view[6], view[10]);
    M3<R> V = M3<R>::cols(c3[0], c3[1], c3[2], c3[1], c3[3], c3[4], c3[2], c3[4], c3[5]);
    M3<R> T = mul(Wm, J);
    M3<R> cov = mul(mul(transpose(T), transpose(V)), T);
    const R a = cov.c[0][0] += R(0.3);
    const R b = cov.c[0][1];
    const R c = cov.c[1][1] += R(0.3);
    const R denom = a * c - b * b;
    R dL_da = 0, dL_db = 0, dL_dc = 0;
    const R d2inv = R(1) / ((denom * denom) + R(1e-7));
    R* dcov = &st.dL_dcov3D[6 * i];
    auto Tm = [&](int col, int row) { return T.c[col][row]; };
    if (d2inv != R(0)) {
      dL_da = d2inv * (-c * c * dcon.x + R(2) * b * c * dcon.y + (denom - a * c) * dcon.z);
      dL_dc = d2inv * (-a * a * dcon.z + R(2) * a * b * dcon.y + (denom - a * c) * dcon.x);
      dL_db = d2inv * R(2) * (b * c * dcon.x - (denom + R(2) * b * b) * dcon.y + a * b * dcon.z);
      dcov[0] = Tm(0, 0) * Tm(0, 0) * dL_da + Tm(0, 0) * Tm(1, 0) * dL_db + Tm(1, 0) * Tm(1, 0) * dL_dc;
      dcov[3] = Tm(0, 1) * Tm(0, 1) * dL_da + Tm(0, 1) * Tm(1, 1) * dL_db + Tm(1, 1) * Tm(1, 1) * dL_dc;
      dcov[5] = Tm(0, 2) * Tm(0, 2) * dL_da + Tm(0, 2) * Tm(1, 2) * dL_db + Tm(1, 2) * Tm(1, 2) * dL_dc;
      dcov[1] = R(2) * Tm(0, 0) * Tm(0, 1) * dL_da + (Tm(0, 0) * Tm(1, 1) + Tm(0, 1) * Tm(1, 0)) * dL_db +
                R(2) * Tm(1, 0) * Tm(1, 1) * dL_dc;
      dcov[2] = R(2) * Tm(0, 0) * Tm(0, 2) * dL_da + (Tm(0, 0) * Tm(1, 2) + Tm(0, 2) * Tm(1, 0)) * dL_db +
                R(2) * Tm(1, 0) * Tm(1, 2) * dL_dc;
      dcov[4] = R(2) * Tm(0, 2) * Tm(0, 1) * dL_da + (Tm(0, 1) * Tm(1, 2) + Tm(0, 2) * Tm(1, 1)) * dL_db +
                R(2) * Tm(1, 1) * Tm(1, 2) * dL_dc;
    }
    auto Vm = [&](int col, int row) { return V.c[col][row]; };
    // dL/dT (upper 2x3), backward.cu:432-443
    R dT[2][3];
    for (int k = 0; k < 3; ++k) {
      const R r0 = Tm(0, 0) * Vm(k, 0) + Tm(0, 1) * Vm(k, 1) + Tm(0, 2) * Vm(k, 2);
      const R r1 = Tm(1, 0) * Vm(k, 0) + Tm(1, 1) * Vm(k, 1) + Tm(1, 2) * Vm(k, 2);
      dT[0][k] = R(2) * r0 * dL_da + r1 * dL_db;
      dT[1][k] = R(2) * r1 * dL_dc + r0 * dL_db;
    }
    auto Wv = [&](int col, int row) { return Wm.c[col][row]; };
    // dL/dJ, backward.cu:447-451
    const R dJ00 = Wv(0, 0) * dT[0][0] + Wv(0, 1) * dT[0][1] + Wv(0, 2) * dT[0][2];
    const R dJ02 = Wv(2, 0) * dT[0][0] + Wv(2, 1) * dT[0][1] + Wv(2, 2) * dT[0][2];
    const R dJ10 = Wv(0, 0) * dT[1][0] + Wv(0, 1) * dT[1][1] + Wv(0, 2) * dT[1][2];
    const R dJ11 = Wv(1, 0) * dT[1][0] + Wv(1, 1) * dT[1][1] + Wv(1, 2) * dT[1][2];
    const R dJ12 = Wv(2, 0) * dT[1][0] + Wv(2, 1) * dT[1][1] + Wv(2, 2) * dT[1][2];
    // second derivatives of the lon/lat projection, backward.cu:455-475 (supp. Eq. 43-55)
    const R Hp = j.H_div_pi, Wp = j.W_div_2pi;
    const R temp1 = Hp * yy_minus_rxz2 * rxz_r4_inv;
    const R temp2 = Hp * txtytz * (r2 + R(2) * rxz2) * rxz3_r4_inv;
    const R temp3 = Wp * (txtx - tztz) * rxz4_inv;
    const R temp4 = Wp * R(2) * t.x * t.z * rxz4_inv;
    const R temp5 = Hp * t.y * rxz3_r4_inv;
    const R dtx = -dJ00 * temp4 + dJ02 * temp3 + dJ10 * temp5 * (R(2) * txtx * rxz2 - tztz * r2) +
                  dJ11 * t.x * temp1 + dJ12 * temp2;
    const R dty = dJ10 * t.x * temp1 - dJ11 * Hp * R(2) * rxz * t.y * r4_inv + dJ12 * t.z * temp1;
    const R dtz = dJ00 * temp3 + dJ02 * temp4 + dJ10 * temp2 + dJ11 * t.z * temp1 +
                  dJ12 * temp5 * (R(2) * tztz * rxz2 - txtx * r2);
    V3<R> dm = xform_vec_T(V3<R>{dtx, dty, dtz}, view);
    st.dL_dmeans3D[3 * i] = dm.x; st.dL_dmeans3D[3 * i + 1] = dm.y; st.dL_dmeans3D[3 * i + 2] = dm.z;
  }
}

// computeColorFromSH (backward), backward.cu:30-151.
template <class R> void sh_backward(int deg, int M, const R* sh_g, V3<R> pos, V3<R> campos, const uint8_t* clamped3,
                                     const R* dL_dcolor3, R* dL_dmean3, R* dL_dsh_g) {
  using K = SH<R>;
  V3<R> d0 = pos - campos;
  R len = osqrt(dot(d0, d0));
  V3<R> dir{d0.x / len, d0.y / len, d0.z / len};
  auto s = [&](int i) { return V3<R>{sh_g[3 * i], sh_g[3 * i + 1], sh_g[3 * i + 2]}; };
  V3<R> g{dL_dcolor3[0] * (clamped3[0] ? R(0) : R(1)), dL_dcolor3[1] * (clamped3[1] ? R(0) : R(1)),
          dL_dcolor3[2] * (clamped3[2] ? R(0) : R(1))};
  auto put = [&](int i, R w) {
    dL_dsh_g[3 * i] = w * g.x; dL_dsh_g[3 * i + 1] = w * g.y; dL_dsh_g[3 * i + 2] = w * g.z;
  };
  V3<R> dx{0, 0, 0}, dy{0, 0, 0}, dz{0, 0, 0};
  const R x = dir.x, y = dir.y, z = dir.z;
  put(0, K::C0);
  if (deg > 0) {
    put(1, -K::C1 * y); put(2, K::C1 * z); put(3, -K::C1 * x);
    dx = R(-K::C1) * s(3);
    dy = R(-K::C1) * s(1);
    dz = K::C1 * s(2);
    if (deg > 1) {
      const R xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
      put(4, K::C2[0] * xy); put(5, K::C2[1] * yz); put(6, K::C2[2] * (R(2) * zz - xx - yy));
      put(7, K::C2[3] * xz); put(8, K::C2[4] * (xx - yy));
      dx = dx + (K::C2[0] * y) * s(4) + (K::C2[2] * R(2) * -x) * s(6) + (K::C2[3] * z) * s(7) + (K::C2[4] * R(2) * x) * s(8);
      dy = dy + (K::C2[0] * x) * s(4) + (K::C2[1] * z) * s(5) + (K::C2[2] * R(2) * -y) * s(6) + (K::C2[4] * R(2) * -y) * s(8);
      dz = dz + (K::C2[1] * y) * s(5) + (K::C2[2] * R(2) * R(2) * z) * s(6) + (K::C2[3] * x) * s(7);
      if (deg > 2) {
        put(9, K::C3[0] * y * (R(3) * xx - yy));
        put(10, K::C3[1] * xy * z);
        put(11, K::C3[2] * y * (R(4) * zz - xx - yy));
        put(12, K::C3[3] * z * (R(2) * zz - R(3) * xx - R(3) * yy));
        put(13, K::C3[4] * x * (R(4) * zz - xx - yy));
        put(14, K::C3[5] * z * (xx - yy));
        put(15, K::C3[6] * x * (xx - R(3) * yy));
        dx = dx + ((((((K::C3[0] * R(3) * R(2) * xy) * s(9) + (K::C3[1] * yz) * s(10)) + (K::C3[2] * R(-2) * xy) * s(11)) +
                     (K::C3[3] * R(-3) * R(2) * xz) * s(12)) + (K::C3[4] * (R(-3) * xx + R(4) * zz - yy)) * s(13)) +
                   (K::C3[5] * R(2) * xz) * s(14)) + (K::C3[6] * R(3) * (xx - yy)) * s(15);
        dy = dy + ((((((K::C3[0] * R(3) * (xx - yy)) * s(9) + (K::C3[1] * xz) * s(10)) +
                      (K::C3[2] * (R(-3) * yy + R(4) * zz - xx)) * s(11)) + (K::C3[3] * R(-3) * R(2) * yz) * s(12)) +
                    (K::C3[4] * R(-2) * xy) * s(13)) + (K::C3[5] * R(-2) * yz) * s(14)) +
             (K::C3[6] * R(-3) * R(2) * xy) * s(15);
        dz = dz + ((((K::C3[1] * xy) * s(10) + (K::C3[2] * R(4) * R(2) * yz) * s(11)) +
                    (K::C3[3] * R(3) * (R(2) * zz - xx - yy)) * s(12)) + (K::C3[4] * R(4) * R(2) * xz) * s(13)) +
             (K::C3[5] * (xx - yy)) * s(14);
      }
    }
  }
  const V3<R> ddir{dot(dx, g), dot(dy, g), dot(dz, g)};
  // dnormvdv, auxiliary.h:134-144
  const R sum2 = d0.x * d0.x + d0.y * d0.y + d0.z * d0.z;
  const R invsum32 = R(1) / osqrt(sum2 * sum2 * sum2);
  dL_dmean3[0] += ((sum2 - d0.x * d0.x) * ddir.x - d0.y * d0.x * ddir.y - d0.z * d0.x * ddir.z) * invsum32;
  dL_dmean3[1] += (-d0.x * d0.y * ddir.x + (sum2 - d0.y * d0.y) * ddir.y - d0.z * d0.y * ddir.z) * invsum32;
  dL_dmean3[2] += (-d0.x * d0.z * ddir.x - d0.y * d0.z * ddir.y + (sum2 - d0.z * d0.z) * ddir.z) * invsum32;
}

// computeCov3D (backward), backward.cu:489-552: dL/dscale (w.r.t. s, ignoring the
// modifier) and dL/dq w.r.t. the raw quaternion.
template <class R> void cov3d_backward(const R* sc, R mod, const R* q, const R* dc, R* dscale, R* drot) {
  const R r = q[0], x = q[1], y = q[2], z = q[3];
  M3<R> Rm = M3<R>::cols(R(1) - R(2) * (y * y + z * z), R(2) * (x * y - r * z), R(2) * (x * z + r * y),
                         R(2) * (x * y + r * z), R(1) - R(2) * (x * x + z * z), R(2) * (y * z - r * x),
                         R(2) * (x * z - r * y), R(2) * (y * z + r * x), R(1) - R(2) * (x * x + y * y));
  const R s[3] = {mod * sc[0], mod * sc[1], mod * sc[2]};
  M3<R> S = M3<R>::cols(s[0], 0, 0, 0, s[1], 0, 0, 0, s[2]);
  M3<R> Mm = mul(S, Rm);
  M3<R> dSig = M3<R>::cols(dc[0], R(0.5) * dc[1], R(0.5) * dc[2], R(0.5) * dc[1], dc[3], R(0.5) * dc[4],
                           R(0.5) * dc[2], R(0.5) * dc[4], dc[5]);
  M3<R> dM = mul(Mm, dSig);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) dM.c[i][j] = R(2) * dM.c[i][j];
  M3<R> Rt = transpose(Rm);
  M3<R> dMt = transpose(dM);
  for (int k = 0; k < 3; ++k)
    dscale[k] = Rt.c[k][0] * dMt.c[k][0] + Rt.c[k][1] * dMt.c[k][1] + Rt.c[k][2] * dMt.c[k][2];
  for (int k = 0; k < 3; ++k)
    for (int j = 0; j < 3; ++j) dMt.c[k][j] *= s[k];
  auto m = [&](int col, int row) { return dMt.c[col][row]; };
  drot[0] = R(2) * z * (m(0, 1) - m(1, 0)) + R(2) * y * (m(2, 0) - m(0, 2)) + R(2) * x * (m(1, 2) - m(2, 1));
  drot[1] = R(2) * y * (m(1, 0) + m(0, 1)) + R(2) * z * (m(2, 0) + m(0, 2)) + R(2) * r * (m(1, 2) - m(2, 1)) -
            R(4) * x * (m(2, 2) + m(1, 1));
  drot[2] = R(2) * x * (m(1, 0) + m(0, 1)) + R(2) * r * (m(2, 0) - m(0, 2)) + R(2) * z * (m(1, 2) + m(2, 1)) -
            R(4) * y * (m(2, 2) + m(0, 0));
  drot[3] = R(2) * r * (m(0, 1) - m(1, 0)) + R(2) * x * (m(2, 0) + m(0, 2)) + R(2) * y * (m(1, 2) + m(2, 1)) -
            R(4) * z * (m(1, 1) + m(0, 0));
}

// preprocessLonLatCUDA (backward), backward.cu:613-669.
template <class R> void preprocess_backward(const Inputs<R>& in, State<R>& st) {
  const int P = in.P;
  st.dL_dsh.assign((size_t)P * in.M * 3, 0);
  st.dL_dscales.assign(3 * (size_t)P, 0);
  st.dL_drot.assign(4 * (size_t)P, 0);
  V3<R> campos{in.campos[0], in.campos[1], in.campos[2]};
  const R dsx_dpx = R(2) / (R)in.W, dsy_dpy = R(2) / (R)in.H;
#pragma omp parallel for schedule(static) if (in.threads > 1) num_threads(in.threads)
  for (int i = 0; i < P; ++i) {
    if (!(st.radii[i] > 0)) continue;
    const R dL_dpx = st.dL_dmean2D[3 * i] * dsx_dpx;
    const R dL_dpy = st.dL_dmean2D[3 * i + 1] * dsy_dpy;
    const R* ax = &st.dpx_dt[3 * i];
    const R* ay = &st.dpy_dt[3 * i];
    V3<R> dt{dL_dpx * ax[0] + dL_dpy * ay[0], dL_dpx * ax[1] + dL_dpy * ay[1], dL_dpx * ax[2] + dL_dpy * ay[2]};
    V3<R> dm = xform_vec_T(dt, in.view);
    st.dL_dmeans3D[3 * i] += dm.x; st.dL_dmeans3D[3 * i + 1] += dm.y; st.dL_dmeans3D[3 * i + 2] += dm.z;
    V3<R> pos{in.means[3 * i], in.means[3 * i + 1], in.means[3 * i + 2]};
    if (in.shs)
      sh_backward(in.D, in.M, in.shs + (size_t)3 * in.M * i, pos, campos, &st.clamped[3 * i], &st.dL_dcolors[3 * i],
                  &st.dL_dmeans3D[3 * i], &st.dL_dsh[(size_t)3 * in.M * i]);
    if (in.scales)
      cov3d_backward(in.scales + 3 * i, in.mod, in.rots + 4 * i, &st.dL_dcov3D[6 * i], &st.dL_dscales[3 * i],
                     &st.dL_drot[4 * i]);
  }
}

template <class R> void backward(const Inputs<R>& in, State<R>& st, const R* dL_dpix) {
  render_backward(in, st, dL_dpix);
  cov2d_backward(in, st);
  preprocess_backward(in, st);
}

}  // namespace omnigs_oracle
