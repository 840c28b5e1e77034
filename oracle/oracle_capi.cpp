// extern "C" surface of the CPU oracle, for ctypes (tests/ and bench.py's CPU
// baseline leg only).  TEST INFRASTRUCTURE: never linked by the product.
//
// A handle holds one precision's State.  Stages run independently so a test
// can feed the GPU's outputs of stage k into the oracle's stage k+1
// ("identical preprocess outputs" parity) via oracle_set().
#include <chrono>
#include <functional>
#include <cstdio>
#include <cstring>
#include <string>

#include "omnigs_oracle.hpp"

using namespace omnigs_oracle;

namespace {

struct HandleBase {
  virtual ~HandleBase() = default;
  int precision = 32;
};

template <class R> struct Handle : HandleBase {
  // owned copies of every input (so the caller's buffers may go away)
  int P = 0, D = 0, M = 0, W = 0, H = 0, bx = 16, by = 16, seam_wrap = 0, threads = 1;
  std::vector<R> bg, means, shs, colors, opac, scales, rots, cov3d_pre, view, campos;
  bool has_shs = false, has_colors = false, has_scales = false, has_rots = false, has_cov = false;
  R mod = 1;
  State<R> st;
  Inputs<R> inputs() const {
    Inputs<R> in;
    in.P = P; in.D = D; in.M = M; in.W = W; in.H = H; in.block_x = bx; in.block_y = by; in.seam_wrap = seam_wrap;
    in.threads = threads;
    in.bg = bg.data(); in.means = means.data(); in.opac = opac.data();
    in.shs = has_shs ? shs.data() : nullptr;
    in.colors = has_colors ? colors.data() : nullptr;
    in.scales = has_scales ? scales.data() : nullptr;
    in.rots = has_rots ? rots.data() : nullptr;
    in.cov3d_pre = has_cov ? cov3d_pre.data() : nullptr;
    in.mod = mod; in.view = view.data(); in.campos = campos.data();
    return in;
  }
};

// Inputs arrive in the handle's precision; a running-error handle (Rerr) reads
// FP64 values and gives them e = 0.
template <class R> void copy_in(std::vector<R>& dst, const void* src, size_t n, bool& has) {
  has = src != nullptr;
  dst.assign(n, R(0));
  if (!src) return;
  if constexpr (is_rerr<R>::value) {
    const double* d = static_cast<const double*>(src);
    for (size_t i = 0; i < n; ++i) dst[i] = Rerr(d[i]);
  } else {
    std::memcpy(dst.data(), src, n * sizeof(R));
  }
}
template <class R> void copy_in(std::vector<R>& dst, const void* src, size_t n) {
  bool h;
  copy_in(dst, src, n, h);
}

// Named arrays of the state, for get/set.  elem = element size in bytes;
// real = the array holds the handle's real type (a running-error handle then
// exchanges FP64 values, or the error bounds of names prefixed "err:").
template <class R> bool lookup(State<R>& st, const std::string& name, void*& ptr, size_t& count, size_t& elem,
                               std::function<void(size_t)>& resize, bool* real = nullptr) {
#define OMNIGS_ARR(NM, VEC)                                                    \
  if (name == NM) {                                                            \
    ptr = (void*)st.VEC.data();                                                \
    count = st.VEC.size();                                                     \
    elem = sizeof(st.VEC[0]);                                                  \
    if (real) *real = std::is_same<std::decay_t<decltype(st.VEC[0])>, R>::value; \
    resize = [&st](size_t n) { st.VEC.resize(n); };                            \
    return true;                                                               \
  }
  OMNIGS_ARR("radii", radii)
  OMNIGS_ARR("xy", xy)
  OMNIGS_ARR("depth", depth)
  OMNIGS_ARR("cov3d", cov3d)
  OMNIGS_ARR("rgb", rgb)
  OMNIGS_ARR("conic", conic)
  OMNIGS_ARR("clamped", clamped)
  OMNIGS_ARR("tiles_touched", tiles_touched)
  OMNIGS_ARR("offsets", offsets)
  OMNIGS_ARR("keys", keys)
  OMNIGS_ARR("point_list", point_list)
  OMNIGS_ARR("ranges", ranges)
  OMNIGS_ARR("final_T", final_T)
  OMNIGS_ARR("n_contrib", n_contrib)
  OMNIGS_ARR("margin", margin)
  OMNIGS_ARR("out_color", out_color)
  OMNIGS_ARR("dL_dmean2D", dL_dmean2D)
  OMNIGS_ARR("dL_dconic", dL_dconic)
  OMNIGS_ARR("dL_dopacity", dL_dopacity)
  OMNIGS_ARR("dL_dcolors", dL_dcolors)
  OMNIGS_ARR("dL_dmeans3D", dL_dmeans3D)
  OMNIGS_ARR("dL_dcov3D", dL_dcov3D)
  OMNIGS_ARR("dL_dsh", dL_dsh)
  OMNIGS_ARR("dL_dscales", dL_dscales)
  OMNIGS_ARR("dL_drot", dL_drot)
  OMNIGS_ARR("dpx_dt", dpx_dt)
  OMNIGS_ARR("dpy_dt", dpy_dt)
  OMNIGS_ARR("fragile", fragile)
  OMNIGS_ARR("gsig", gsig)
  OMNIGS_ARR("psig", psig)
#undef OMNIGS_ARR
  return false;
}

template <class R> Handle<R>* as(void* h) { return static_cast<Handle<R>*>(static_cast<HandleBase*>(h)); }

// precision 32 / 64: FP32 / FP64 oracle; kRunningError: FP64 values with
// running FP32 rounding-error bounds (backward stages only).
constexpr int kRunningError = 65;

template <class F> auto dispatch(void* h, F&& f) {
  HandleBase* b = static_cast<HandleBase*>(h);
  if (b->precision == kRunningError) return f(*as<Rerr>(h));
  if (b->precision == 64) return f(*as<double>(h));
  return f(*as<float>(h));
}

}  // namespace

extern "C" {

// Inputs are in the handle's precision (float32 or float64 arrays); any of
// shs / colors / scales / rots / cov3d_pre may be NULL, as in the reference.
void* oracle_create(int precision, int P, int D, int M, int W, int H, int block_x, int block_y, const void* bg,
                    const void* means, const void* shs, const void* colors, const void* opac, const void* scales,
                    double mod, const void* rots, const void* cov3d_pre, const void* view, const void* campos) {
  auto fill = [&](auto* h) {
    using R = typename std::remove_reference<decltype(h->mod)>::type;
    h->P = P; h->D = D; h->M = M; h->W = W; h->H = H; h->bx = block_x; h->by = block_y;
    h->mod = (R)mod;
    copy_in<R>(h->bg, bg, 3);
    copy_in<R>(h->means, means, 3 * (size_t)P);
    copy_in<R>(h->shs, shs, (size_t)P * M * 3, h->has_shs);
    copy_in<R>(h->colors, colors, 3 * (size_t)P, h->has_colors);
    copy_in<R>(h->opac, opac, (size_t)P);
    copy_in<R>(h->scales, scales, 3 * (size_t)P, h->has_scales);
    copy_in<R>(h->rots, rots, 4 * (size_t)P, h->has_rots);
    copy_in<R>(h->cov3d_pre, cov3d_pre, 6 * (size_t)P, h->has_cov);
    copy_in<R>(h->view, view, 16);
    copy_in<R>(h->campos, campos, 3);
  };
  if (precision == kRunningError) {
    auto* h = new Handle<Rerr>();
    h->precision = kRunningError;
    fill(h);
    return static_cast<HandleBase*>(h);
  }
  if (precision == 64) {
    auto* h = new Handle<double>();
    h->precision = 64;
    fill(h);
    return static_cast<HandleBase*>(h);
  }
  auto* h = new Handle<float>();
  h->precision = 32;
  fill(h);
  return static_cast<HandleBase*>(h);
}

void oracle_destroy(void* h) { delete static_cast<HandleBase*>(h); }

// Extension switch: 1 = wrap tile columns and dx around the longitude seam.
void oracle_set_seam_wrap(void* h, int on) {
  dispatch(h, [&](auto& hd) { hd.seam_wrap = on ? 1 : 0; return 0; });
}

// Threads of the OpenMP CPU baseline (1 = the serial oracle).  Returns the
// count in effect (1 when built without OpenMP).
int oracle_set_threads(void* h, int n) {
#ifdef _OPENMP
  const int t = n < 1 ? 1 : n;
#else
  const int t = 1;
  (void)n;
#endif
  dispatch(h, [&](auto& hd) { hd.threads = t; return 0; });
  return t;
}

// stage: 0 preprocess, 1 binning, 2 render, 3 render_backward (needs dL_dpix),
// 4 cov2d_backward, 5 preprocess_backward, 6 cov3d (the 3D covariances from
// scales and rotations, as preprocess computes them).  A running-error handle
// runs stages 3-6 only, on state set from outside.  Returns 0 on success.
int oracle_run(void* h, int stage, const void* dL_dpix) {
  return dispatch(h, [&](auto& hd) -> int {
    using R = typename std::remove_reference<decltype(hd.mod)>::type;
    auto in = hd.inputs();
    std::vector<R> dpix;
    const R* dp = static_cast<const R*>(dL_dpix);
    if constexpr (is_rerr<R>::value) {
      if (stage < 3) return 2;
      if (dL_dpix) {
        const size_t n = (size_t)kChannels * in.W * in.H;
        dpix.resize(n);
        for (size_t i = 0; i < n; ++i) dpix[i] = Rerr(static_cast<const double*>(dL_dpix)[i]);
        dp = dpix.data();
      }
    } else {
      switch (stage) {
        case 0: preprocess(in, hd.st); return 0;
        case 1: binning(in, hd.st); return 0;
        case 2: render(in, hd.st); return 0;
      }
    }
    switch (stage) {
      case 3: render_backward(in, hd.st, dp); return 0;
      case 4: cov2d_backward(in, hd.st); return 0;
      case 5: preprocess_backward(in, hd.st); return 0;
      case 6:
        if (!in.scales || !in.rots) return 3;
        hd.st.cov3d.assign(6 * (size_t)in.P, R(0));
        for (int i = 0; i < in.P; ++i) cov3d_forward(in.scales + 3 * i, in.mod, in.rots + 4 * i, &hd.st.cov3d[6 * i]);
        return 0;
    }
    return 1;
  });
}

long long oracle_fragile_pixels(void* h) {
  return dispatch(h, [](auto& hd) { return (long long)hd.st.fragile_pixels; });
}

long long oracle_num_rendered(void* h) {
  return dispatch(h, [](auto& hd) { return (int)hd.st.R_; });
}

unsigned long long oracle_decision_sig(void* h) {
  return dispatch(h, [](auto& hd) { return (unsigned long long)hd.st.decision_sig; });
}

// A running-error handle exchanges its real arrays as FP64: the values, or
// with the prefix "err:" their error bounds (in units of 2^-24).
namespace {
bool err_name(const char*& name) {
  if (std::strncmp(name, "err:", 4) != 0) return false;
  name += 4;
  return true;
}
}  // namespace

// Size of a named array in elements (-1 if unknown) and its element size.
long long oracle_size(void* h, const char* name, int* elem_bytes) {
  return dispatch(h, [&](auto& hd) -> long long {
    using R = typename std::remove_reference<decltype(hd.mod)>::type;
    const char* nm = name;
    const bool err = err_name(nm);
    void* p; size_t n, e; std::function<void(size_t)> rs; bool real = false;
    if (!lookup(hd.st, nm, p, n, e, rs, &real)) return -1;
    if (err && !(is_rerr<R>::value && real)) return -1;
    if (is_rerr<R>::value && real) e = sizeof(double);
    if (elem_bytes) *elem_bytes = (int)e;
    return (long long)n;
  });
}

int oracle_get(void* h, const char* name, void* dst, long long count) {
  return dispatch(h, [&](auto& hd) -> int {
    using R = typename std::remove_reference<decltype(hd.mod)>::type;
    const char* nm = name;
    const bool err = err_name(nm);
    void* p; size_t n, e; std::function<void(size_t)> rs; bool real = false;
    if (!lookup(hd.st, nm, p, n, e, rs, &real)) return 1;
    if ((size_t)count != n) return 2;
    if constexpr (is_rerr<R>::value) {
      if (real) {
        const Rerr* src = static_cast<const Rerr*>(p);
        double* d = static_cast<double*>(dst);
        for (size_t i = 0; i < n; ++i) d[i] = err ? src[i].e : src[i].v;
        return 0;
      }
    }
    if (err) return 1;
    std::memcpy(dst, p, n * e);
    return 0;
  });
}

// On a running-error handle, "err:<name>" sets the error bounds of an array
// already set (input uncertainty, in units of 2^-24), leaving its values.
int oracle_set(void* h, const char* name, const void* src, long long count) {
  return dispatch(h, [&](auto& hd) -> int {
    using R = typename std::remove_reference<decltype(hd.mod)>::type;
    const char* nm = name;
    if (err_name(nm)) {
      void* p; size_t n, e; std::function<void(size_t)> rs; bool real = false;
      if (!lookup(hd.st, nm, p, n, e, rs, &real)) return 1;
      if constexpr (is_rerr<R>::value) {
        if (!real || (size_t)count != n) return 2;
        Rerr* d = static_cast<Rerr*>(p);
        const double* ev = static_cast<const double*>(src);
        for (long long i = 0; i < count; ++i) d[i].e = ev[i];
        return 0;
      }
      return 1;
    }
    void* p; size_t n, e; std::function<void(size_t)> rs; bool real = false;
    if (!lookup(hd.st, name, p, n, e, rs, &real)) return 1;
    rs((size_t)count);
    lookup(hd.st, name, p, n, e, rs);
    bool done = false;
    if constexpr (is_rerr<R>::value) {
      if (real) {
        Rerr* d = static_cast<Rerr*>(p);
        const double* sv = static_cast<const double*>(src);
        for (long long i = 0; i < count; ++i) d[i] = Rerr(sv[i]);
        done = true;
      }
    }
    if (!done) std::memcpy(p, src, (size_t)count * e);
    if (std::string(name) == "point_list") hd.st.R_ = (uint32_t)count;
    return 0;
  });
}

int oracle_higher_msb(unsigned n) { return (int)higher_msb(n); }

// Full forward + backward, timed with steady_clock (the CPU baseline leg).
// Returns the elapsed seconds of fwd+bwd; -1 on error.
double oracle_time_fwd_bwd(void* h, const void* dL_dpix) {
  return dispatch(h, [&](auto& hd) -> double {
    using R = typename std::remove_reference<decltype(hd.mod)>::type;
    if constexpr (is_rerr<R>::value) {
      return -1.0;
    } else {
      auto in = hd.inputs();
      auto t0 = std::chrono::steady_clock::now();
      forward(in, hd.st);
      backward(in, hd.st, static_cast<const R*>(dL_dpix));
      auto t1 = std::chrono::steady_clock::now();
      return std::chrono::duration<double>(t1 - t0).count();
    }
  });
}

}  // extern "C"
