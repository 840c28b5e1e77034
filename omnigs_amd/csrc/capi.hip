// C ABI entry points (include/omnigs_lonlat.h): host orchestration of the
// forward (preprocess -> depth sort -> scan -> expand -> tile sort -> ranges
// -> render) and the backward (render backward -> fused preprocess backward).
// Stage order and buffer contract follow the reference's
// LonlatRasterizer::forward / backward (rasterizer_impl.cu:540-697, :701-795).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/omnigs_lonlat.h"
#include "kernels.hpp"
#include "mlp.hpp"

namespace {

thread_local char g_last_error[512] = "";

// Seam wrap reads tile columns modulo the grid width and pixel offsets modulo
// the image width; the two rings agree only when the tiles tile the width
// exactly, so other widths are refused rather than dropping instances near
// the seam.
constexpr const char* kWrapWidthMsg = "seam_wrap needs an image width that is a multiple of the tile width";

int fail_hip(hipError_t e, const char* where) {
  std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s", where, hipGetErrorString(e));
  return OMNIGS_ERR_HIP;
}
int fail_arg(const char* what) {
  std::snprintf(g_last_error, sizeof(g_last_error), "invalid argument: %s", what);
  return OMNIGS_ERR_INVALID_ARGUMENT;
}

#define OMNIGS_TRY(expr, where)                 \
  do {                                          \
    hipError_t _e = (expr);                     \
    if (_e != hipSuccess) return fail_hip(_e, where); \
  } while (0)

// ---- stage timers (HIP events on the launch stream) ----
struct StageEvent {
  int stage;
  hipEvent_t start, stop;
};
std::mutex g_prof_mu;
int g_prof_mask = 0;  // bit s: time stage s
std::vector<StageEvent> g_prof_events;

// Events are kept in a pool and reused: creating one costs the calling thread
// microseconds, which would delay the launches between two timed stages.
std::vector<hipEvent_t> g_prof_pool;

bool take_event(hipEvent_t* e) {
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!g_prof_pool.empty()) {
      *e = g_prof_pool.back();
      g_prof_pool.pop_back();
      return true;
    }
  }
  return hipEventCreate(e) == hipSuccess;
}

struct StageTimer {
  int stage;
  hipStream_t s;
  StageEvent ev{};
  bool on;
  StageTimer(int st, hipStream_t stream) : stage(st), s(stream) {
    {
      std::lock_guard<std::mutex> lk(g_prof_mu);
      on = st >= 0 && st < OMNIGS_NUM_STAGES && ((g_prof_mask >> st) & 1);
    }
    if (on) {
      ev.stage = st;
      if (!take_event(&ev.start) || !take_event(&ev.stop) || hipEventRecord(ev.start, s) != hipSuccess)
        on = false;
    }
  }
  ~StageTimer() {
    if (!on) return;
    if (hipEventRecord(ev.stop, s) == hipSuccess) {
      std::lock_guard<std::mutex> lk(g_prof_mu);
      g_prof_events.push_back(ev);
    }
  }
};

char* grow(omnigs_buffer b, size_t bytes) {
  if (!b.resize) return nullptr;
  return b.resize(b.ctx, bytes);
}

}  // namespace

using namespace omnigs;

extern "C" {

size_t omnigs_lonlat_geometry_bytes(int P) { return carve_geom(nullptr, (size_t)(P < 0 ? 0 : P)).bytes; }
size_t omnigs_lonlat_binning_bytes(int R) { return carve_bin(nullptr, bin_capacity((uint32_t)(R < 0 ? 0 : R))).bytes; }
size_t omnigs_lonlat_image_bytes(int width, int height) {
  if (width <= 0 || height <= 0) return 0;
  return carve_img(nullptr, width, height).bytes;
}

const char* omnigs_status_string(int status) {
  switch (status) {
    case OMNIGS_OK: return "ok";
    case OMNIGS_ERR_INVALID_ARGUMENT: return "invalid argument";
    case OMNIGS_ERR_NEED_COLORS: return "For non-RGB, provide precomputed Gaussian colors!";
    case OMNIGS_ERR_ALLOC: return "buffer allocation failed";
    case OMNIGS_ERR_HIP: return "HIP error";
  }
  return "unknown status";
}

const char* omnigs_last_error(void) { return g_last_error; }

const char* omnigs_version(void) { return "omnigs-mi355x 0.1 (gfx950)"; }

int omnigs_profile_enable(int mask) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_mask = mask & OMNIGS_ALL_STAGES;
  return OMNIGS_OK;
}

int omnigs_profile_read(double* ms, int* count) {
  std::vector<StageEvent> evs;
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    evs.swap(g_prof_events);
  }
  int status = OMNIGS_OK;
  for (auto& e : evs) {
    float t = 0.f;
    hipError_t err = hipEventSynchronize(e.stop);
    if (err == hipSuccess) err = hipEventElapsedTime(&t, e.start, e.stop);
    if (err != hipSuccess) status = fail_hip(err, "profile_read");
    else if (e.stage >= 0 && e.stage < OMNIGS_NUM_STAGES) {
      if (ms) ms[e.stage] += t;
      if (count) count[e.stage] += 1;
    }
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_pool.push_back(e.start);
    g_prof_pool.push_back(e.stop);
  }
  return status;
}

const char* omnigs_stage_name(int stage) {
  static const char* names[OMNIGS_NUM_STAGES] = {"preprocess", "depth_sort", "scan", "duplicate", "tile_sort",
                                                 "ranges", "render", "gather_grads", "render_backward",
                                                 "preprocess_backward", "decoder_backward"};
  return (stage >= 0 && stage < OMNIGS_NUM_STAGES) ? names[stage] : "unknown";
}

int omnigs_lonlat_state(const char* geom_buffer, const char* binning_buffer, const char* image_buffer, int P, int R,
                        int width, int height, omnigs_lonlat_state_view* out, size_t binning_bytes) {
  if (!out || P < 0 || R < 0 || width <= 0 || height <= 0) return fail_arg("state view");
  std::memset(out, 0, sizeof(*out));
  if (geom_buffer) {
    const GeomState g = carve_geom(const_cast<char*>(geom_buffer), P);
    out->depth_key = g.depth_key;
    out->radii = g.radii;
    out->render_records = reinterpret_cast<const float*>(g.rec);
    out->clamped = g.clamped;
    out->tiles_touched = g.tiles_touched;
    out->depth_order = g.order;
    out->offsets = g.offsets;
  }
  if (binning_buffer && R > 0) {
    // the arrays viewed precede the gradient records, so a forward-only
    // layout (omnigs_lonlat_forward_checkpointed) is viewed the same way
    uint32_t C = bin_capacity((uint32_t)R);
    if (binning_bytes) {
      C = bin_capacity_of_bytes((uint32_t)R, binning_bytes, true);
      if (C == 0) C = bin_capacity_of_bytes((uint32_t)R, binning_bytes, false);
      if (C == 0) return fail_arg("binning_bytes is not the size of a forward's binning buffer for this R");
    }
    const BinState b = carve_bin(const_cast<char*>(binning_buffer), C);
    const bool alt = radix_passes(bit_width_u32((uint32_t)tile_grid(width, height).tiles())) % 2;
    out->point_list = alt ? b.gid_alt : b.gid;
    out->tile_keys = alt ? b.tile_key_alt : b.tile_key;
    out->blend = b.blend;
  }
  if (image_buffer) {
    const ImgState s = carve_img(const_cast<char*>(image_buffer), width, height);
    out->final_T = s.final_T;
    out->n_contrib = s.n_contrib;
    out->ranges = reinterpret_cast<const uint32_t*>(s.ranges);
  }
  return OMNIGS_OK;
}

size_t omnigs_loss_workspace_bytes(int channels, int height, int width, int skip_bottom) {
  if (channels <= 0 || height <= 0 || width <= 0 || skip_bottom < 0 || skip_bottom >= height) return 0;
  return loss_workspace_floats(channels, height, width, skip_bottom) * sizeof(float);
}

int omnigs_photometric_loss(const float* image, const float* gt_image, const float* mask, int channels, int height,
                            int width, int skip_bottom, float lambda_dssim, void* workspace, size_t workspace_bytes,
                            float* dL_dimage, float* loss_out, void* stream) {
  if (channels <= 0 || height <= 0 || width <= 0) return fail_arg("empty image");
  if (skip_bottom < 0 || skip_bottom >= height) return fail_arg("skip_bottom outside [0, height)");
  if (!(lambda_dssim >= 0.f && lambda_dssim <= 1.f)) return fail_arg("lambda_dssim outside [0, 1]");
  if (!image || !gt_image || !workspace || !dL_dimage || !loss_out) return fail_arg("missing required pointer");
  if (workspace_bytes < omnigs_loss_workspace_bytes(channels, height, width, skip_bottom))
    return fail_arg("workspace smaller than omnigs_loss_workspace_bytes()");
  OMNIGS_TRY(launch_loss(image, gt_image, mask, channels, height, width, skip_bottom, lambda_dssim,
                         static_cast<float*>(workspace), dL_dimage, loss_out, (hipStream_t)stream),
             "photometric loss");
  return OMNIGS_OK;
}

int omnigs_photometric_loss_noalloc(const float* image, const float* gt_image, const float* mask, int channels,
                                    int height, int width, int skip_bottom, float lambda_dssim, float* dL_dimage,
                                    float* loss_out, void* stream) {
  if (channels <= 0 || height <= 0 || width <= 0) return fail_arg("empty image");
  if (skip_bottom < 0 || skip_bottom >= height) return fail_arg("skip_bottom outside [0, height)");
  if (!(lambda_dssim >= 0.f && lambda_dssim <= 1.f)) return fail_arg("lambda_dssim outside [0, 1]");
  if (!image || !gt_image || !dL_dimage || !loss_out) return fail_arg("missing required pointer");
  if (loss_workspace_floats(channels, height, width, skip_bottom) > kLossScratchFloats)
    return fail_arg("image too large for the allocation-free loss (workspace over OMNIGS_LOSS_NOALLOC_MAX_BYTES)");
  float* ws = nullptr;
  OMNIGS_TRY(loss_scratch(&ws), "loss scratch address");
  OMNIGS_TRY(launch_loss(image, gt_image, mask, channels, height, width, skip_bottom, lambda_dssim, ws, dL_dimage,
                         loss_out, (hipStream_t)stream),
             "photometric loss (no-alloc)");
  return OMNIGS_OK;
}

int omnigs_lonlat_mark_visible(int P, uint8_t* present, void* stream) {
  if (P < 0) return fail_arg("P < 0");
  if (P > 0 && !present) return fail_arg("present is NULL");
  OMNIGS_TRY(launch_mark_visible(P, present, (hipStream_t)stream), "mark_visible");
  return OMNIGS_OK;
}

}  // extern "C"

namespace {

// Pinned host words per calling thread for the forward's read-back of the
// instance count and its overflow flag (allocated on first use, kept for the
// thread's lifetime).
uint32_t* pinned_word() {
  thread_local uint32_t* word = nullptr;
  thread_local bool tried = false;
  if (!tried) {
    tried = true;
    void* p = nullptr;
    if (hipHostMalloc(&p, 64, hipHostMallocDefault) == hipSuccess) word = static_cast<uint32_t*>(p);
  }
  return word;
}

// One event per calling thread for the forward's read-back of R (created on
// first use, kept for the thread's lifetime).
hipEvent_t readback_event() {
  thread_local hipEvent_t ev = nullptr;
  thread_local bool tried = false;
  if (!tried) {
    tried = true;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) ev = nullptr;
  }
  return ev;
}

// Expansion, tile sort, tile ranges and (render = true) the render forward,
// into a binning buffer carved for bin_capacity(n) instances.  R_dev NULL:
// n is the exact instance count R.  Otherwise n is a capacity guess and the
// kernels read the live count {R, overflow flag} at R_dev, stopping there
// (their grids cover the capacity).
int bin_and_render(const omnigs_lonlat_forward_args* a, omnigs_buffer binning_buffer, omnigs_buffer image_buffer,
                   const GeomState& g, ImgState& img, const TileGrid& grid, uint32_t n, const uint32_t* R_dev,
                   hipStream_t s, bool render, bool records) {
  const uint32_t C = bin_capacity(n);
  char* bbuf = grow(binning_buffer, carve_bin(nullptr, C, records).bytes);
  if (!bbuf) {
    std::snprintf(g_last_error, sizeof(g_last_error), "binning buffer allocation failed");
    return OMNIGS_ERR_ALLOC;
  }
  BinState b = carve_bin(bbuf, C, records);
  if (render && !records) {
    // checkpointed forward: the blend bytes go to the image buffer's tail,
    // which the caller keeps for the backward (the image state is carved
    // again: the buffer may have moved)
    const int W = a->width, H = a->height;
    char* ibuf = grow(image_buffer, img_blend_bytes(W, H, C));
    if (!ibuf) {
      std::snprintf(g_last_error, sizeof(g_last_error), "image buffer allocation failed");
      return OMNIGS_ERR_ALLOC;
    }
    img = carve_img(ibuf, W, H);
    b.blend_ok = reinterpret_cast<uint32_t*>(ibuf + img_blend_offset(W, H));
    b.blend = reinterpret_cast<uint8_t*>(ibuf + img_blend_offset(W, H) + kAlign);
  }
  const uint32_t grid_n = R_dev ? C : n;  // instances the grids cover
  const int tile_bits = bit_width_u32((uint32_t)grid.tiles());
  {
    // the expansion also counts the tile sort's first digit
    StageTimer tm(OMNIGS_STAGE_DUPLICATE, s);
    OMNIGS_TRY(launch_duplicate(a->P, grid_n, R_dev, g.order, g.offsets, g.rect_depth, grid, b.tile_key, b.gid,
                                reinterpret_cast<uint32_t*>(b.grad_opacity), b.sort_counts, tile_bits, img.ranges,
                                b.blend_ok, s),
               "duplicate");
  }
  SortResult ts{b.tile_key, b.gid};
  if (grid_n > 0) {
    StageTimer tm(OMNIGS_STAGE_TILE_SORT, s);
    OMNIGS_TRY(launch_radix_sort(b.tile_key, b.tile_key_alt, b.gid, b.gid, b.gid_alt, grid_n, tile_bits,
                                 b.sort_counts, b.scan_partials, s, &ts, R_dev, true),
               "tile sort");
  }
  {
    StageTimer tm(OMNIGS_STAGE_RANGES, s);
    OMNIGS_TRY(launch_tile_ranges(ts.keys, grid_n, img.ranges, grid.tiles(), s, R_dev, grid_n > 0), "ranges");
  }
  if (!render) return OMNIGS_OK;
  {
    StageTimer tm(OMNIGS_STAGE_RENDER, s);
    OMNIGS_TRY(launch_render_forward(grid, img.ranges, ts.vals, a->width, a->height, g.rec, img.final_T,
                                     img.n_contrib, a->background, a->out_color, b.blend, b.blend_ok, s),
               "render");
  }
  return OMNIGS_OK;
}

// Checks a forward's arguments (P > 0; out_color only when render) and
// returns its tile grid.
int check_forward_args(const omnigs_lonlat_forward_args* a, bool render, TileGrid* grid) {
  const int W = a->width, H = a->height;
  if (a->P > kMaxGaussians) return fail_arg("more than 89478485 Gaussians in one call");
  if (!a->means3D || !a->opacities || !a->viewmatrix || !a->campos || !a->background || (render && !a->out_color))
    return fail_arg("missing required pointer");
  const bool mlp = a->mlp_packed && a->latents;
  if (!mlp && (a->mlp_packed || a->latents)) return fail_arg("mlp_packed and latents must be given together");
  if (!mlp && !a->colors_precomp && !a->shs) return fail_arg("neither shs nor colors_precomp");
  if (!mlp && !a->colors_precomp && (a->D < 0 || a->D > 3 || a->M < (a->D + 1) * (a->D + 1)))
    return fail_arg("SH degree / coefficient count");
  if (!a->cov3D_precomp && (!a->scales || !a->rotations)) return fail_arg("neither scales+rotations nor cov3D");
  *grid = tile_grid(W, H, a->seam_wrap ? 1 : 0);
  if (grid->gx > kMaxGridCols || grid->gy > 65535) return fail_arg("tile grid over 32766 columns or 65535 rows");
  if (grid->tiles() > kMaxTiles) return fail_arg("more than 2^21 tiles in one image");
  if (a->seam_wrap && W % grid->bx != 0) return fail_arg(kWrapWidthMsg);
  return OMNIGS_OK;
}

// Preprocess, then the Gaussians by (depth bits, id) into g.order: a 32-bit
// stable radix sort with index values (an even pass count ends in g.order).
int preprocess_and_sort(const omnigs_lonlat_forward_args* a, const GeomState& g, int* radii, hipStream_t s) {
  const bool mlp = a->mlp_packed && a->latents;
  PreprocessArgs pa{a->P, a->D, a->M, a->width, a->height, a->means3D, a->shs, a->colors_precomp, a->opacities,
                    a->scales, a->scale_modifier, a->rotations, a->cov3D_precomp, a->viewmatrix, a->campos,
                    a->seam_wrap ? 1 : 0, mlp ? a->mlp_packed : nullptr, mlp ? a->latents : nullptr};
  {
    StageTimer tm(OMNIGS_STAGE_PREPROCESS, s);
    OMNIGS_TRY(launch_preprocess(pa, g, radii, s), "preprocess");
  }
  static_assert(32 % (2 * kRadixBits) == 0, "the depth sort must end in g.order");
  SortResult ds;
  {
    StageTimer tm(OMNIGS_STAGE_DEPTH_SORT, s);
    OMNIGS_TRY(launch_radix_sort(g.depth_key, g.key_alt, nullptr, g.order, g.order_alt, (size_t)a->P, 32,
                                 g.sort_counts, g.scan_partials, s, &ds),
               "depth sort");
  }
  return OMNIGS_OK;
}

// Shared by forward (render = true) and recompute (render = false, image
// buffer given and kept).
int forward_impl(const omnigs_lonlat_forward_args* a, omnigs_buffer geometry_buffer, omnigs_buffer binning_buffer,
                 omnigs_buffer image_buffer, char* kept_image, void* stream_, int* num_rendered, bool render,
                 bool records = true) {
  if (!a || !num_rendered) return fail_arg("NULL args");
  *num_rendered = 0;
  const int P = a->P, W = a->width, H = a->height;
  if (P < 0 || W <= 0 || H <= 0) return fail_arg("P < 0 or empty image");
  if (P == 0) return OMNIGS_OK;  // the reference boundary skips the rasterizer entirely
  if (!render && !kept_image) return fail_arg("recompute needs the forward's image buffer");
  TileGrid grid;
  {
    const int st = check_forward_args(a, render, &grid);
    if (st != OMNIGS_OK) return st;
  }
  hipStream_t s = (hipStream_t)stream_;

  char* gbuf = grow(geometry_buffer, carve_geom(nullptr, P).bytes);
  char* ibuf = render ? grow(image_buffer, carve_img(nullptr, W, H).bytes) : kept_image;
  if (!gbuf || !ibuf) {
    std::snprintf(g_last_error, sizeof(g_last_error), "geometry/image buffer allocation failed");
    return OMNIGS_ERR_ALLOC;
  }
  const GeomState g = carve_geom(gbuf, P);
  ImgState img = carve_img(ibuf, W, H);
  int* radii = a->radii ? a->radii : g.radii;
  {
    const int st = preprocess_and_sort(a, g, radii, s);
    if (st != OMNIGS_OK) return st;
  }
  // Instance offsets in depth order.
  {
    StageTimer tm(OMNIGS_STAGE_SCAN, s);
    // key_alt is free after the depth sort: it holds tiles_touched in depth order.
    // The same launches scan the preprocess blocks' slot counts into the first
    // gradient slot of every block (the backward's records).
    OMNIGS_TRY(launch_scan(g.tiles_touched, g.order, g.offsets, (size_t)P, false, g.scan_partials, s, g.key_alt,
                           g.block_slots, slot_blocks(P), g.slot_base, g.dup_rect, g.rect_depth),
               "scan");
  }

  // The instance count R sizes the binning buffer, and the host needs it to
  // return it.  Its read-back ({R, overflow flag} after the scan's block
  // partials) is queued first; the binning and the render are then queued
  // against the capacity bucket of the previous call's R with 1/16 headroom
  // when that call had the same Gaussian count and image size (the kernels
  // read the live count on the device and stop there), and only
  // then does the host wait for R -- so the GPU has the rest of the forward
  // to run meanwhile.  That binning stands whenever its capacity covers R
  // (and is under twice R): the backward recovers the capacity from the
  // binning buffer's size.  A call whose R outgrew it (and the first call on
  // a thread) bins again with bin_capacity(R).
  const uint32_t* total = g.scan_partials + scan_blocks((size_t)P);
  uint32_t* pw = pinned_word();
  hipEvent_t ev = readback_event();
  const bool async = pw && ev;
  if (async) {
    OMNIGS_TRY(hipMemcpyAsync(pw, total, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "read R");
    OMNIGS_TRY(hipEventRecord(ev, s), "read R event");
  }
  // The previous call on this thread, if it rendered the same number of
  // Gaussians at the same size (a training loop between densifications):
  // another scene or size bins exactly, rather than against a capacity from
  // an unrelated call.
  struct LastCall {
    uint32_t R;
    int P, W, H;
  };
  thread_local LastCall last{0u, 0, 0, 0};
  const bool same_shape = last.P == a->P && last.W == a->width && last.H == a->height;
  const uint32_t guess = async && same_shape && last.R ? bin_capacity(last.R + last.R / 16u) : 0u;
  if (guess > 0) {
    const int st =
        bin_and_render(a, binning_buffer, image_buffer, g, img, grid, guess, total, s, render, records);
    if (st != OMNIGS_OK) return st;
  }
  uint32_t Rw[2] = {0u, 0u};
  if (async) {
    hipError_t q;
    while ((q = hipEventQuery(ev)) == hipErrorNotReady) {
    }
    OMNIGS_TRY(q, "sync R");
    Rw[0] = pw[0];
    Rw[1] = pw[1];
  } else {
    OMNIGS_TRY(hipMemcpyAsync(Rw, total, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "read R");
    OMNIGS_TRY(hipStreamSynchronize(s), "sync R");
  }
  if (Rw[1] != 0u || Rw[0] > 0x7fffffffu) return fail_arg("instance count exceeds int range");
  const uint32_t R = Rw[0];
  const bool covered = guess > 0 && R > 0 && guess >= R && (uint64_t)guess <= 2ull * R;
  if (!covered) {
    const int st = bin_and_render(a, binning_buffer, image_buffer, g, img, grid, R, nullptr, s, render, records);
    if (st != OMNIGS_OK) return st;
  }
  last = LastCall{R, a->P, a->width, a->height};
  *num_rendered = (int)R;
  return OMNIGS_OK;
}

// Checks a backward's arguments; buffers: the forward's three buffers must
// be given (else only the image buffer).
int check_backward_args(const omnigs_lonlat_backward_args* a, bool buffers) {
  const int P = a->P, W = a->width, H = a->height, R = a->R;
  if (P < 0 || R < 0 || W <= 0 || H <= 0) return fail_arg("negative size or empty image");
  if (P == 0) return OMNIGS_OK;
  if (!a->means3D || !a->viewmatrix || !a->campos || !a->background || !a->dL_dpix || !a->image_buffer ||
      (buffers && (!a->geom_buffer || (R > 0 && !a->binning_buffer))))
    return fail_arg("missing required pointer");
  if (!a->dL_dmean2D || !a->dL_dconic || !a->dL_dopacity || !a->dL_dcolor || !a->dL_dmean3D || !a->dL_dcov3D)
    return fail_arg("missing gradient output");
  const bool mlp = a->mlp_packed && a->latents;
  if (!mlp && (a->mlp_packed || a->latents)) return fail_arg("mlp_packed and latents must be given together");
  if (mlp && (!a->dL_dlatent || !a->dL_dmlp || !a->mlp_workspace)) return fail_arg("decoder gradient output missing");
  if (mlp && a->mlp_workspace_bytes < omnigs_mlp_workspace_bytes(P)) return fail_arg("mlp_workspace too small");
  if (!mlp && a->shs && !a->dL_dsh) return fail_arg("dL_dsh is NULL while shs is given");
  if (!mlp && !a->colors_precomp && !a->shs) return fail_arg("neither shs nor colors_precomp");
  if (!a->cov3D_precomp && (!a->scales || !a->rotations)) return fail_arg("neither scales+rotations nor cov3D");
  if (a->scales && (!a->dL_dscale || !a->dL_drot)) return fail_arg("dL_dscale / dL_drot is NULL");
  const TileGrid grid = tile_grid(W, H, a->seam_wrap ? 1 : 0);
  if (a->seam_wrap && W % grid.bx != 0) return fail_arg(kWrapWidthMsg);
  return OMNIGS_OK;
}

// The preprocess backward (reading the per-Gaussian 2D gradients, or with
// `gather` summing them from the records itself) and the decoder's backward.
int backward_tail(const omnigs_lonlat_backward_args* a, const GeomState& g, const GatherSrc& gather, hipStream_t s) {
  const int P = a->P;
  const bool mlp = a->mlp_packed && a->latents;
  const int* radii = a->radii ? a->radii : g.radii;
  PreprocessBackwardArgs pb{P, a->D, a->M, a->width, a->height, a->means3D, mlp ? nullptr : a->shs, a->scales,
                            a->scale_modifier, a->rotations,
                            a->cov3D_precomp, a->viewmatrix, a->campos, radii, g.clamped, a->dL_dmean2D,
                            a->dL_dconic, a->dL_dcolor, a->dL_dmean3D, a->dL_dcov3D, a->dL_dsh,
                            a->dL_dscale, a->dL_drot, a->dpx_dt, a->dpy_dt, gather,
                            a->dL_dmean2D, a->dL_dconic, a->dL_dopacity, a->dL_dcolor};
  if (mlp) {
    // the decoder's backward runs on the Gaussians with a colour gradient
    // only, listed by the preprocess backward
    pb.act_list = mlp_act_list(static_cast<float*>(a->mlp_workspace), P);
    pb.act_count = mlp_act_count(static_cast<float*>(a->mlp_workspace), P);
    pb.dL_dlatent = a->dL_dlatent;
  }
  {
    StageTimer tm(OMNIGS_STAGE_PREPROCESS_BACKWARD, s);
    OMNIGS_TRY(launch_preprocess_backward(pb, s), "preprocess backward");
  }
  if (mlp) {
    StageTimer tm(OMNIGS_STAGE_DECODER_BACKWARD, s);
    OMNIGS_TRY(launch_mlp_backward(P, a->means3D, a->campos, radii, a->latents, a->mlp_packed, a->dL_dcolor,
                                   a->dL_dmean3D, a->dL_dlatent, a->dL_dmlp, static_cast<float*>(a->mlp_workspace), s,
                                   1, true),
               "decoder backward");
  }
  return OMNIGS_OK;
}

}  // namespace

extern "C" {

int omnigs_lonlat_forward(const omnigs_lonlat_forward_args* a, omnigs_buffer geometry_buffer,
                          omnigs_buffer binning_buffer, omnigs_buffer image_buffer, void* stream, int* num_rendered) {
  return forward_impl(a, geometry_buffer, binning_buffer, image_buffer, nullptr, stream, num_rendered, true);
}

int omnigs_lonlat_forward_checkpointed(const omnigs_lonlat_forward_args* a, omnigs_buffer geometry_buffer,
                                       omnigs_buffer binning_buffer, omnigs_buffer image_buffer, void* stream,
                                       int* num_rendered) {
  return forward_impl(a, geometry_buffer, binning_buffer, image_buffer, nullptr, stream, num_rendered, true, false);
}

int omnigs_lonlat_recompute(const omnigs_lonlat_forward_args* a, omnigs_buffer geometry_buffer,
                            omnigs_buffer binning_buffer, char* image_buffer, void* stream, int* num_rendered) {
  return forward_impl(a, geometry_buffer, binning_buffer, omnigs_buffer{nullptr, nullptr}, image_buffer, stream,
                      num_rendered, false);
}

int omnigs_lonlat_backward(const omnigs_lonlat_backward_args* a, void* stream_) {
  if (!a) return fail_arg("NULL args");
  {
    const int st = check_backward_args(a, true);
    if (st != OMNIGS_OK || a->P == 0) return st;
  }
  const int P = a->P, W = a->width, H = a->height, R = a->R;
  hipStream_t s = (hipStream_t)stream_;
  const TileGrid grid = tile_grid(W, H, a->seam_wrap ? 1 : 0);
  const GeomState g = carve_geom(a->geom_buffer, P);
  const ImgState img = carve_img(a->image_buffer, W, H);

  // The point list ends in gid (even pass count) or gid_alt (odd).
  const uint32_t* point_list = nullptr;
  float4* grad_rec = nullptr;
  float* grad_opacity = nullptr;
  const uint8_t* blend = nullptr;
  const uint32_t* blend_ok = nullptr;
  if (R > 0) {
    // the forward's layout, named by the size it asked for (required: a
    // forward may keep a layout larger than R's own)
    const uint32_t C = bin_capacity_of_bytes((uint32_t)R, a->binning_bytes);
    if (C == 0) return fail_arg("binning_bytes is not the size of a forward's binning buffer for this R");
    const BinState b = carve_bin(a->binning_buffer, C);
    const int passes = radix_passes(bit_width_u32((uint32_t)grid.tiles()));
    point_list = (passes % 2) ? b.gid_alt : b.gid;
    grad_rec = b.grad_rec;
    grad_opacity = b.grad_opacity;
    blend = b.blend;
    blend_ok = b.blend_ok;
  }

  if (R > 0) {
    StageTimer tm(OMNIGS_STAGE_RENDER_BACKWARD, s);
    OMNIGS_TRY(launch_render_backward(grid, img.ranges, point_list, W, H, a->background, g.rec, img.final_T,
                                      img.n_contrib, a->dL_dpix, g.slot_base, (uint32_t)R, grad_rec, grad_opacity,
                                      blend, blend_ok, s),
               "render backward");
  }
  // The per-Gaussian sums of the records (the four 2D gradient outputs, every
  // entry written, zeros where no tile touched the Gaussian) are gathered
  // inside the preprocess backward, which uses them from registers.
  return backward_tail(
      a, g, GatherSrc{P, (uint32_t)R, nullptr, g.tiles_touched, g.slot_local, g.slot_base, grad_rec, grad_opacity},
      s);
}

int omnigs_lonlat_backward_checkpointed(const omnigs_lonlat_forward_args* f, const omnigs_lonlat_backward_args* a,
                                        omnigs_buffer geometry_buffer, omnigs_buffer binning_buffer,
                                        unsigned int band_instances, void* stream_, int* num_bands) {
  if (!f || !a) return fail_arg("NULL args");
  if (num_bands) *num_bands = 0;
  {
    const int st = check_backward_args(a, false);
    if (st != OMNIGS_OK || a->P == 0) return st;
  }
  if (f->P != a->P || f->width != a->width || f->height != a->height || (f->seam_wrap != 0) != (a->seam_wrap != 0))
    return fail_arg("forward and backward arguments describe different calls");
  TileGrid grid;
  {
    const int st = check_forward_args(f, false, &grid);
    if (st != OMNIGS_OK) return st;
  }
  const int P = a->P, W = a->width, H = a->height;
  hipStream_t s = (hipStream_t)stream_;

  // The geometry buffer, then a tail: a band's compacted owner rects, the
  // per-scan-tile owner counts, the per-row instance counts' difference
  // array, and each Gaussian's first gradient slot in the band.
  const size_t geom_bytes = carve_geom(nullptr, P).bytes;
  const size_t rect_bytes = align_up((size_t)P * sizeof(uint2));
  const size_t part_bytes = align_up((scan_blocks((size_t)P) + 2) * sizeof(uint32_t));
  const size_t diff_bytes = align_up(((size_t)grid.gy + 1) * sizeof(uint32_t));
  const size_t slot_bytes = align_up((size_t)P * sizeof(uint32_t));
  char* gbuf = grow(geometry_buffer, geom_bytes + rect_bytes + part_bytes + diff_bytes + slot_bytes +
                                         align_up((size_t)grid.tiles() * sizeof(uint2)));
  if (!gbuf) {
    std::snprintf(g_last_error, sizeof(g_last_error), "geometry buffer allocation failed");
    return OMNIGS_ERR_ALLOC;
  }
  const GeomState g = carve_geom(gbuf, P);
  // the whole rects in depth order (once), and per band its owners' ids,
  // inclusive offsets and rects (free arrays after the depth sort, and the
  // tail), with the per-scan-tile partials of its instances and owners
  uint2* rect_d = g.rect_depth;
  uint32_t* order_c = g.order_alt;
  uint32_t* offsets_c = g.offsets;
  uint2* rect_c = reinterpret_cast<uint2*>(gbuf + geom_bytes);
  uint32_t* part_area = g.scan_partials;
  uint32_t* part_count = reinterpret_cast<uint32_t*>(gbuf + geom_bytes + rect_bytes);
  uint32_t* row_diff = reinterpret_cast<uint32_t*>(gbuf + geom_bytes + rect_bytes + part_bytes);
  uint32_t* slot_words = reinterpret_cast<uint32_t*>(gbuf + geom_bytes + rect_bytes + part_bytes + diff_bytes);
  // a band's tile ranges (the image buffer keeps the forward's, through
  // which the forward's blend bytes are read)
  uint2* band_ranges = reinterpret_cast<uint2*>(gbuf + geom_bytes + rect_bytes + part_bytes + diff_bytes + slot_bytes);
  const uint32_t* band_total = part_area + scan_blocks((size_t)P);    // {R_b, overflow flag}
  const uint32_t* band_owners = part_count + scan_blocks((size_t)P);  // owners with an instance
  const ImgState img = carve_img(a->image_buffer, W, H);
  // The forward's blend bytes, when the image buffer is a checkpointed
  // forward's (its size covers them): each band's render backward then
  // evaluates only the (wave, instance) pairs the forward blended, as the
  // stored backward does (a band's list of a tile is the forward's list, so
  // position p there is position p of the forward's).  Else the quadrant
  // masks pick the pairs.
  const bool fwd_blend = a->R > 0 && a->image_bytes >= img_blend_bytes(W, H, (uint32_t)a->R);
  const uint32_t* blend_ok = fwd_blend ? reinterpret_cast<const uint32_t*>(a->image_buffer + img_blend_offset(W, H))
                                       : nullptr;
  const uint8_t* blend = fwd_blend ? reinterpret_cast<const uint8_t*>(a->image_buffer + img_blend_offset(W, H) + kAlign)
                                   : nullptr;
  {
    const int st = preprocess_and_sort(f, g, f->radii ? f->radii : g.radii, s);
    if (st != OMNIGS_OK) return st;
  }
  std::vector<uint32_t> diff((size_t)grid.gy + 1);
  {
    StageTimer tm(OMNIGS_STAGE_SCAN, s);
    OMNIGS_TRY(launch_row_counts(P, g.dup_rect, grid.gy, row_diff, s), "row counts");
    OMNIGS_TRY(launch_gather_rects(g.dup_rect, g.order, P, rect_d, s), "rects in depth order");
  }
  OMNIGS_TRY(hipMemcpyAsync(diff.data(), row_diff, diff.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
             "read row counts");
  OMNIGS_TRY(hipStreamSynchronize(s), "sync row counts");

  // Bands of whole tile rows, each of at most `budget` instances (a single
  // row above it makes a band of its own).
  std::vector<uint64_t> rows((size_t)grid.gy);
  uint32_t run = 0;
  uint64_t R = 0;
  for (int y = 0; y < grid.gy; ++y) {
    run += diff[(size_t)y];  // mod 2^32; each row holds at most R < 2^31 instances
    rows[(size_t)y] = run;
    R += run;
  }
  if (R != (uint64_t)a->R) return fail_arg("recomputed instance count differs from the forward's R");
  const uint64_t budget = band_instances ? (uint64_t)band_instances
                                         : std::max<uint64_t>((uint64_t)1 << 20, (R + 3) / 4);
  struct Band {
    int y0, y1;
    uint32_t n;
  };
  std::vector<Band> bands;
  uint64_t biggest = 0;
  for (int y = 0; y < grid.gy;) {
    const int y0 = y;
    uint64_t n = rows[(size_t)y++];
    while (y < grid.gy && n + rows[(size_t)y] <= budget) n += rows[(size_t)y++];
    if (n > 0) bands.push_back({y0, y, (uint32_t)n});
    biggest = std::max(biggest, n);
  }
  BinState b{};
  if (biggest > 0) {
    const uint32_t C = bin_capacity((uint32_t)biggest);
    char* bbuf = grow(binning_buffer, carve_bin(nullptr, C).bytes);
    if (!bbuf) {
      std::snprintf(g_last_error, sizeof(g_last_error), "binning buffer allocation failed");
      return OMNIGS_ERR_ALLOC;
    }
    b = carve_bin(bbuf, C);
  }

  // The 2D gradient outputs: a Gaussian's first band writes its sums and
  // later bands add theirs; the first band also writes the zeros of the
  // Gaussians without instances.  No band: all zeros.
  if (bands.empty()) {
    StageTimer tm(OMNIGS_STAGE_PREPROCESS_BACKWARD, s);
    OMNIGS_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(a->dL_dmean2D), 3 * (size_t)P, 0u, s), "zero dmean2D");
    OMNIGS_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(a->dL_dconic), 4 * (size_t)P, 0u, s), "zero dconic");
    OMNIGS_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(a->dL_dopacity), (size_t)P, 0u, s), "zero dopacity");
    OMNIGS_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(a->dL_dcolor), 3 * (size_t)P, 0u, s), "zero dcolor");
  }
  // Per band: its owners (offsets, rects, and the first gradient slot of
  // each, in owner order), expansion, tile sort and ranges (the lists of its
  // tiles are the forward's: every Gaussian over a tile, in (depth, id)
  // order), the render backward over its tiles, and the records' sums added
  // to the 2D gradients.
  const int tile_bits = bit_width_u32((uint32_t)grid.tiles());
  for (const Band& band : bands) {
    {
      // the Gaussians outside the band (zero area, interleaved in depth
      // order) are dropped from the expansion's owner list
      StageTimer tm(OMNIGS_STAGE_SCAN, s);
      OMNIGS_TRY(launch_band_owners(rect_d, g.order, (size_t)P, (uint32_t)band.y0, (uint32_t)band.y1, part_area,
                                    part_count, order_c, offsets_c, rect_c, slot_words, s),
                 "band owners");
    }
    {
      // band.n (from the row counts) equals the owners' total; the kernel
      // also stops at the device total
      StageTimer tm(OMNIGS_STAGE_DUPLICATE, s);
      OMNIGS_TRY(launch_duplicate(P, band.n, band_total, order_c, offsets_c, rect_c, grid, b.tile_key, b.gid,
                                  reinterpret_cast<uint32_t*>(b.grad_opacity), b.sort_counts, tile_bits, band_ranges,
                                  b.blend_ok, s, band_owners),
                 "band duplicate");
    }
    SortResult ts;
    {
      StageTimer tm(OMNIGS_STAGE_TILE_SORT, s);
      OMNIGS_TRY(launch_radix_sort(b.tile_key, b.tile_key_alt, b.gid, b.gid, b.gid_alt, band.n, tile_bits,
                                   b.sort_counts, b.scan_partials, s, &ts, nullptr, true),
                 "band tile sort");
    }
    {
      StageTimer tm(OMNIGS_STAGE_RANGES, s);
      OMNIGS_TRY(launch_tile_ranges(ts.keys, band.n, band_ranges, grid.tiles(), s, nullptr, true), "band ranges");
    }
    {
      StageTimer tm(OMNIGS_STAGE_RENDER_BACKWARD, s);
      OMNIGS_TRY(launch_render_backward(grid, band_ranges, ts.vals, W, H, a->background, g.rec, img.final_T,
                                        img.n_contrib, a->dL_dpix, nullptr, band.n, b.grad_rec, b.grad_opacity,
                                        blend, blend_ok, s, band.y0, band.y1, slot_words,
                                        fwd_blend ? img.ranges : nullptr),
                 "band render backward");
    }
    {
      StageTimer tm(OMNIGS_STAGE_PREPROCESS_BACKWARD, s);
      OMNIGS_TRY(launch_band_gather(P, g.dup_rect, (uint32_t)band.y0, (uint32_t)band.y1, slot_words, band.n, b.grad_rec,
                                    b.grad_opacity, a->dL_dmean2D, a->dL_dconic, a->dL_dopacity, a->dL_dcolor,
                                    &band == &bands.front(), s),
                 "band gather");
    }
  }
  if (num_bands) *num_bands = (int)bands.size();
  return backward_tail(a, g, GatherSrc{P, 0u, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr}, s);
}

// ---------------------------------------------------------------------------
// Batched multi-view path.
namespace {

struct BatchState {
  GeomState g;  // per item (V P)
  BinState b;   // per instance (capacity)
  float* final_T;
  uint32_t* n_contrib;
  uint2* ranges;
  float* dmean2D;  // per item [3]
  float* dconic;   // per item [4]
  float* dopacity; // per item
  float* dcolor;   // per item [3]
  size_t bytes;
};

BatchState carve_batch(char* base, size_t items, size_t pixels, size_t tiles, size_t capacity) {
  BatchState st;
  size_t off = 0;
  auto at = [&](size_t n) { return base ? base + off : nullptr; };
  st.g = carve_geom(at(0), items);
  off += st.g.bytes;
  st.b = carve_bin(at(0), capacity);
  off += st.b.bytes;
  Carver<char> c(at(0));
  st.final_T = c.take<float>(pixels);
  st.n_contrib = c.take<uint32_t>(pixels);
  st.ranges = c.take<uint2>(tiles);
  st.dmean2D = c.take<float>(3 * items);
  st.dconic = c.take<float>(4 * items);
  st.dopacity = c.take<float>(items);
  st.dcolor = c.take<float>(3 * items);
  off += align_up(c.off) + kAlign;
  st.bytes = off;
  return st;
}

int make_view_table(int P, int V, const int* widths, const int* heights, int bx, int by, int wrap, ViewTable& vt) {
  if (V < 1 || V > kMaxViews) return fail_arg("V outside [1, 32]");
  if (!widths || !heights) return fail_arg("widths / heights missing");
  if (!tile_shape_supported(bx, by)) return fail_arg("tile shape must be 16x16, 32x8 or 8x32");
  vt = ViewTable{};
  vt.V = V;
  vt.P = P;
  vt.bx = bx;
  vt.by = by;
  vt.wrap = wrap ? 1 : 0;
  uint64_t tiles = 0, pix = 0;
  for (int v = 0; v < V; ++v) {
    if (widths[v] <= 0 || heights[v] <= 0) return fail_arg("empty view");
    const TileGrid g = tile_grid(widths[v], heights[v], wrap, bx, by);
    if (g.gx > kMaxGridCols || g.gy > 65535) return fail_arg("view tile grid over 32766 columns or 65535 rows");
    if (g.tiles() > kMaxTiles) return fail_arg("more than 2^21 tiles in one view");
    if (wrap && widths[v] % bx != 0) return fail_arg(kWrapWidthMsg);
    vt.W[v] = widths[v];
    vt.H[v] = heights[v];
    vt.gx[v] = g.gx;
    vt.gy[v] = g.gy;
    vt.tile_base[v] = (uint32_t)tiles;
    vt.pix_base[v] = (uint32_t)pix;
    tiles += (uint64_t)g.tiles();
    pix += (uint64_t)widths[v] * heights[v];
  }
  if (tiles > 0x7fffffffULL || pix > 0xffffffffULL) return fail_arg("batch too large");
  if ((uint64_t)V * (uint64_t)P > 0x7fffffffULL) return fail_arg("V * P exceeds int range");
  vt.tile_base[V] = (uint32_t)tiles;
  vt.pix_base[V] = (uint32_t)pix;
  // Tile keys: compact global ids, unless padding every view to 2^b keys
  // lets the tile sort order only the b low bits in fewer passes (a stable
  // pass over the V P items then groups the depth order by view).
  uint32_t widest = 0;
  for (int v = 0; v < V; ++v) widest = std::max(widest, vt.tile_base[v + 1] - vt.tile_base[v]);
  const int compact_bits = bit_width_u32((uint32_t)tiles);
  const int local_bits = std::max(1, bit_width_u32(widest - 1u));
  if (V > 1 && radix_passes(local_bits) < radix_passes(compact_bits)) {
    vt.key_shift = local_bits;
    vt.key_count = (uint32_t)V << local_bits;
    vt.sort_bits = local_bits;
  } else {
    vt.key_shift = 0;
    vt.key_count = (uint32_t)tiles;
    vt.sort_bits = compact_bits;
  }
  return OMNIGS_OK;
}

}  // namespace

size_t omnigs_lonlat_batch_workspace_bytes(int P, int V, const int* widths, const int* heights, int tile_x,
                                           int tile_y, unsigned int capacity) {
  ViewTable vt;
  if (P < 0 || make_view_table(P, V, widths, heights, tile_x, tile_y, 0, vt) != OMNIGS_OK) return 0;
  return carve_batch(nullptr, (size_t)V * P, vt.pix_base[V], vt.key_count, capacity).bytes;
}

int omnigs_lonlat_batch_forward(const omnigs_lonlat_batch_args* a, void* stream_) {
  if (!a) return fail_arg("NULL args");
  if (a->P < 0) return fail_arg("P < 0");
  if (a->P > kMaxGaussians) return fail_arg("more than 89478485 Gaussians in one call");
  ViewTable vt;
  const int st = make_view_table(a->P, a->V, a->widths, a->heights, a->tile_x, a->tile_y, a->seam_wrap, vt);
  if (st != OMNIGS_OK) return st;
  const int P = a->P;
  const bool mlp = a->mlp_packed && a->latents;
  if (!mlp && (a->mlp_packed || a->latents)) return fail_arg("mlp_packed and latents must be given together");
  if (!a->means3D || !a->opacities || !a->viewmatrices || !a->campos || !a->background || !a->out_color ||
      !a->status || !a->workspace)
    return fail_arg("missing required pointer");
  if (!mlp && !a->colors_precomp && !a->shs) return fail_arg("neither shs nor colors_precomp");
  if (!mlp && !a->colors_precomp && (a->D < 0 || a->D > 3 || a->M < (a->D + 1) * (a->D + 1)))
    return fail_arg("SH degree / coefficient count");
  if (!a->cov3D_precomp && (!a->scales || !a->rotations)) return fail_arg("neither scales+rotations nor cov3D");
  if (a->capacity == 0) return fail_arg("capacity 0");
  const size_t items = (size_t)a->V * P;
  const BatchState bs = carve_batch(static_cast<char*>(a->workspace), items, vt.pix_base[vt.V], vt.key_count,
                                    a->capacity);
  if (a->workspace_bytes < bs.bytes) return fail_arg("workspace smaller than omnigs_lonlat_batch_workspace_bytes()");
  hipStream_t s = (hipStream_t)stream_;
  if (P == 0) {
    OMNIGS_TRY(launch_fill_u32(a->status, 2, 0u, s), "status");
    return OMNIGS_OK;
  }
  const GeomState& g = bs.g;
  PreprocessArgs pa{P, a->D, a->M, 0, 0, a->means3D, a->shs, a->colors_precomp, a->opacities, a->scales,
                    a->scale_modifier, a->rotations, a->cov3D_precomp, nullptr, nullptr, a->seam_wrap ? 1 : 0,
                    mlp ? a->mlp_packed : nullptr, mlp ? a->latents : nullptr, a->tile_x, a->tile_y};
  {
    StageTimer tm(OMNIGS_STAGE_PREPROCESS, s);
    OMNIGS_TRY(launch_preprocess_views(pa, a->viewmatrices, a->campos, vt, g, g.radii, s), "batch preprocess");
  }
  SortResult ds;
  {
    StageTimer tm(OMNIGS_STAGE_DEPTH_SORT, s);
    OMNIGS_TRY(launch_radix_sort(g.depth_key, g.key_alt, nullptr, g.order, g.order_alt, items, 32, g.sort_counts,
                                 g.scan_partials, s, &ds),
               "batch depth sort");
    if (vt.key_shift) {
      // padded keys: the depth order regrouped view by view (a stable sort by
      // view; key_alt and offsets are free until the scan), into order_alt
      SortResult vs;
      OMNIGS_TRY(launch_view_keys(ds.vals, items, (uint32_t)P, g.key_alt, s), "batch view keys");
      OMNIGS_TRY(launch_radix_sort(g.key_alt, g.offsets, ds.vals, g.order_alt, g.order_alt, items,
                                   bit_width_u32((uint32_t)vt.V - 1u), g.sort_counts, g.scan_partials, s, &vs),
                 "batch view sort");
      ds.vals = vs.vals;
    }
  }
  {
    StageTimer tm(OMNIGS_STAGE_SCAN, s);
    OMNIGS_TRY(launch_scan(g.tiles_touched, ds.vals, g.offsets, items, false, g.scan_partials, s, g.key_alt,
                           g.block_slots, slot_blocks(P), g.slot_base, g.dup_rect, g.rect_depth),
               "batch scan");
  }
  // {instance count, 32-bit overflow flag} after the offset scan's block partials
  const uint32_t* R_dev = g.scan_partials + scan_blocks(items);
  const BinState& b = bs.b;
  const int tile_bits = vt.sort_bits;
  {
    StageTimer tm(OMNIGS_STAGE_DUPLICATE, s);
    OMNIGS_TRY(launch_duplicate_views((int)items, a->capacity, R_dev, a->status, ds.vals, g.offsets, g.rect_depth,
                                      vt, b.tile_key, b.gid, reinterpret_cast<uint32_t*>(b.grad_opacity),
                                      b.sort_counts, tile_bits, bs.ranges, b.blend_ok, s),
               "batch duplicate");
  }
  SortResult ts;
  {
    StageTimer tm(OMNIGS_STAGE_TILE_SORT, s);
    OMNIGS_TRY(launch_radix_sort(b.tile_key, b.tile_key_alt, b.gid, b.gid, b.gid_alt, a->capacity, tile_bits,
                                 b.sort_counts, b.scan_partials, s, &ts, R_dev, true),
               "batch tile sort");
  }
  {
    StageTimer tm(OMNIGS_STAGE_RANGES, s);
    OMNIGS_TRY(launch_tile_ranges(ts.keys, a->capacity, bs.ranges, (int)vt.key_count, s, R_dev,
                                  items > 0 && a->capacity > 0),
               "batch ranges");
  }
  {
    StageTimer tm(OMNIGS_STAGE_RENDER, s);
    OMNIGS_TRY(launch_render_forward_views(vt, bs.ranges, ts.vals, g.rec, bs.final_T, bs.n_contrib, a->background,
                                           a->out_color, b.blend, b.blend_ok, s),
               "batch render");
  }
  return OMNIGS_OK;
}

int omnigs_lonlat_batch_backward(const omnigs_lonlat_batch_args* a, const omnigs_lonlat_batch_grads* gr,
                                 void* stream_) {
  if (!a || !gr) return fail_arg("NULL args");
  ViewTable vt;
  const int st = make_view_table(a->P, a->V, a->widths, a->heights, a->tile_x, a->tile_y, a->seam_wrap, vt);
  if (st != OMNIGS_OK) return st;
  const int P = a->P;
  if (P == 0) return OMNIGS_OK;
  const bool mlp = a->mlp_packed && a->latents;
  if (!gr->dL_dpix || !gr->dL_dmean3D || !gr->dL_dopacity || !gr->dL_dcov3D || !a->workspace)
    return fail_arg("missing gradient pointer");
  if (!mlp && a->shs && !gr->dL_dsh) return fail_arg("dL_dsh is NULL while shs is given");
  if (a->scales && (!gr->dL_dscale || !gr->dL_drot)) return fail_arg("dL_dscale / dL_drot is NULL");
  if (mlp && (!gr->dL_dlatent || !gr->dL_dmlp || !gr->mlp_workspace ||
              gr->mlp_workspace_bytes < omnigs_mlp_workspace_bytes(P)))
    return fail_arg("decoder gradient outputs / workspace missing");
  const size_t items = (size_t)a->V * P;
  const BatchState bs = carve_batch(static_cast<char*>(a->workspace), items, vt.pix_base[vt.V], vt.key_count,
                                    a->capacity);
  if (a->workspace_bytes < bs.bytes) return fail_arg("workspace smaller than omnigs_lonlat_batch_workspace_bytes()");
  hipStream_t s = (hipStream_t)stream_;
  const GeomState& g = bs.g;
  const BinState& b = bs.b;
  const uint32_t* point_list = (radix_passes(vt.sort_bits) % 2) ? b.gid_alt : b.gid;
  {
    StageTimer tm(OMNIGS_STAGE_RENDER_BACKWARD, s);
    OMNIGS_TRY(launch_render_backward_views(vt, bs.ranges, point_list, a->background, g.rec, bs.final_T,
                                            bs.n_contrib, gr->dL_dpix, g.slot_base, a->capacity, b.grad_rec,
                                            b.grad_opacity, b.blend, b.blend_ok, s),
               "batch render backward");
  }
  {
    StageTimer tm(OMNIGS_STAGE_GATHER_GRADS, s);
    // the live instance count (0 after a 32-bit overflow), as the forward's kernels read it
    OMNIGS_TRY(launch_gather_grads((int)items, P, a->capacity, g.scan_partials + scan_blocks(items), g.tiles_touched,
                                   g.slot_local, g.slot_base, b.grad_rec, b.grad_opacity, bs.dmean2D, bs.dconic,
                                   bs.dopacity, bs.dcolor, s),
               "batch gather grads");
  }
  PreprocessBackwardArgs pb{P, a->D, a->M, 0, 0, a->means3D, mlp ? nullptr : a->shs, a->scales, a->scale_modifier,
                            a->rotations, a->cov3D_precomp, nullptr, nullptr, g.radii, g.clamped, bs.dmean2D,
                            bs.dconic, bs.dcolor, gr->dL_dmean3D, gr->dL_dcov3D, mlp ? nullptr : gr->dL_dsh,
                            gr->dL_dscale, gr->dL_drot, nullptr, nullptr};
  if (mlp) {
    // the decoder's backward runs on the Gaussians with a colour gradient in
    // some view, listed by the preprocess backward (all of their views)
    pb.act_list = mlp_act_list(static_cast<float*>(gr->mlp_workspace), P);
    pb.act_count = mlp_act_count(static_cast<float*>(gr->mlp_workspace), P);
    pb.dL_dlatent = gr->dL_dlatent;
  }
  {
    StageTimer tm(OMNIGS_STAGE_PREPROCESS_BACKWARD, s);
    OMNIGS_TRY(launch_preprocess_backward_views(pb, a->viewmatrices, a->campos, vt, bs.dopacity, gr->dL_dopacity, s),
               "batch preprocess backward");
  }
  if (mlp) {
    StageTimer tm(OMNIGS_STAGE_DECODER_BACKWARD, s);
    OMNIGS_TRY(launch_mlp_backward(P, a->means3D, a->campos, g.radii, a->latents, a->mlp_packed, bs.dcolor,
                                   gr->dL_dmean3D, gr->dL_dlatent, gr->dL_dmlp, static_cast<float*>(gr->mlp_workspace),
                                   s, a->V, true),
               "batch decoder backward");
  }
  return OMNIGS_OK;
}

int omnigs_lonlat_batch_state(const omnigs_lonlat_batch_args* a, omnigs_lonlat_batch_state_view* out) {
  if (!a || !out) return fail_arg("NULL args");
  std::memset(out, 0, sizeof(*out));
  if (a->P < 0 || !a->workspace) return fail_arg("P < 0 or no workspace");
  ViewTable vt;
  const int st = make_view_table(a->P, a->V, a->widths, a->heights, a->tile_x, a->tile_y, a->seam_wrap, vt);
  if (st != OMNIGS_OK) return st;
  const size_t items = (size_t)a->V * a->P;
  const BatchState bs =
      carve_batch(static_cast<char*>(a->workspace), items, vt.pix_base[vt.V], vt.key_count, a->capacity);
  if (a->workspace_bytes < bs.bytes) return fail_arg("workspace smaller than omnigs_lonlat_batch_workspace_bytes()");
  omnigs_lonlat_state_view& v = out->s;
  v.depth_key = bs.g.depth_key;
  v.radii = bs.g.radii;
  v.render_records = reinterpret_cast<const float*>(bs.g.rec);
  v.clamped = bs.g.clamped;
  v.tiles_touched = bs.g.tiles_touched;
  v.depth_order = bs.g.order;
  v.offsets = bs.g.offsets;
  const bool alt = radix_passes(vt.sort_bits) % 2;
  v.point_list = alt ? bs.b.gid_alt : bs.b.gid;
  v.tile_keys = alt ? bs.b.tile_key_alt : bs.b.tile_key;
  v.blend = bs.b.blend;
  v.final_T = bs.final_T;
  v.n_contrib = bs.n_contrib;
  v.ranges = reinterpret_cast<const uint32_t*>(bs.ranges);
  out->dL_dmean2D = bs.dmean2D;
  out->dL_dconic = bs.dconic;
  out->dL_dopacity = bs.dopacity;
  out->dL_dcolor = bs.dcolor;
  out->key_shift = (unsigned int)vt.key_shift;
  out->key_count = vt.key_count;
  return OMNIGS_OK;
}

int omnigs_mlp_param_count(void) { return kThetaSize; }
size_t omnigs_mlp_packed_bytes(void) { return kMlpPackedBytes; }
size_t omnigs_mlp_workspace_bytes(int P) { return P < 0 ? 0 : mlp_workspace_floats(P) * sizeof(float); }

int omnigs_mlp_pack(const float* theta, void* packed, void* stream) {
  if (!theta || !packed) return fail_arg("NULL theta or packed");
  OMNIGS_TRY(launch_mlp_pack(theta, packed, (hipStream_t)stream), "mlp pack");
  return OMNIGS_OK;
}

int omnigs_mlp_eval(int N, const uint16_t* latents, const float* dirs, const void* packed, float* rgb, void* stream) {
  if (N < 0) return fail_arg("N < 0");
  if (N > 0 && (!latents || !dirs || !packed || !rgb)) return fail_arg("missing pointer");
  OMNIGS_TRY(launch_mlp_eval(N, latents, dirs, packed, rgb, (hipStream_t)stream), "mlp eval");
  return OMNIGS_OK;
}

int omnigs_mlp_distill_step(const omnigs_mlp_distill_args* a, void* stream) {
  if (!a) return fail_arg("NULL args");
  if (a->P <= 0 || a->M < 16 || a->num_dirs < 1) return fail_arg("P <= 0, M < 16 or num_dirs < 1");
  if (!a->shs || !a->latents || !a->mlp_packed || !a->dL_dlatent || !a->dL_dmlp || !a->loss || !a->workspace)
    return fail_arg("missing pointer");
  if (a->workspace_bytes < omnigs_mlp_workspace_bytes(a->P)) return fail_arg("workspace too small");
  if (a->means3D)
    for (int i = 0; i < 3; ++i)
      if (!(a->view_box[i] <= a->view_box[3 + i])) return fail_arg("view_box: lo above hi");
  OMNIGS_TRY(launch_mlp_distill(a->P, a->M, a->shs, a->latents, a->mlp_packed, a->seed, a->step, a->num_dirs,
                                a->dL_dlatent, a->dL_dmlp, a->loss, static_cast<float*>(a->workspace),
                                (hipStream_t)stream, a->means3D, a->view_box),
             "mlp distill");
  return OMNIGS_OK;
}

void omnigs_mlp_seeded_view_centre(unsigned int seed, unsigned int g, unsigned int k, const float view_box[6],
                                   float out[3]) {
  const float3 c = seeded_view_centre(seed, g, k, make_float3(view_box[0], view_box[1], view_box[2]),
                                      make_float3(view_box[3], view_box[4], view_box[5]));
  out[0] = c.x;
  out[1] = c.y;
  out[2] = c.z;
}

void omnigs_mlp_seeded_direction(unsigned int seed, unsigned int g, unsigned int k, float out[3]) {
  const float3 d = seeded_direction(seed, g, k);
  out[0] = d.x;
  out[1] = d.y;
  out[2] = d.z;
}

int omnigs_adam_step(const omnigs_adam_group* groups, int n, float beta1, float beta2, void* stream) {
  if (n < 0 || n > kAdamMaxGroups || (n > 0 && !groups)) return fail_arg("group count outside [0, 8]");
  if (!(beta1 >= 0.f && beta1 < 1.f && beta2 >= 0.f && beta2 < 1.f)) return fail_arg("betas outside [0, 1)");
  AdamGroupTable t{};
  t.n = n;
  t.one_minus_beta1 = (float)(1.0 - (double)beta1);
  t.one_minus_beta2 = (float)(1.0 - (double)beta2);
  size_t off = 0;
  for (int i = 0; i < n; ++i) {
    const omnigs_adam_group& g = groups[i];
    if (g.numel > 0 && (!g.param || !g.grad || !g.exp_avg || !g.exp_avg_sq)) return fail_arg("missing group pointer");
    if (g.step < 1) return fail_arg("step < 1");
    // scalars as torch::optim::Adam forms them (double), used in FP32
    const double bc1 = 1.0 - std::pow((double)beta1, (double)g.step);
    const double bc2 = 1.0 - std::pow((double)beta2, (double)g.step);
    t.g[i] = AdamGroupDev{g.param, g.grad, g.exp_avg, g.exp_avg_sq, g.param_bf16, (float)((double)g.lr / bc1),
                          (float)std::sqrt(bc2), g.eps};
    t.offset[i] = off;
    off += g.numel;
  }
  t.offset[n] = off;
  OMNIGS_TRY(launch_adam(t, beta1, beta2, (hipStream_t)stream), "adam");
  return OMNIGS_OK;
}

int omnigs_densification_stats(int P, const int* radii, const float* dL_dmean2D, float* max_radii2D,
                               float* xyz_gradient_accum, float* denom, void* stream) {
  if (P < 0) return fail_arg("P < 0");
  if (P > 0 && (!radii || !dL_dmean2D || !max_radii2D || !xyz_gradient_accum || !denom))
    return fail_arg("missing pointer");
  OMNIGS_TRY(launch_densify_stats(P, radii, dL_dmean2D, max_radii2D, xyz_gradient_accum, denom, (hipStream_t)stream),
             "densification stats");
  return OMNIGS_OK;
}

float omnigs_expon_lr(int step, float lr_init, float lr_final, int lr_delay_steps, float lr_delay_mult,
                      int max_steps) {
  return expon_lr(step, lr_init, lr_final, lr_delay_steps, lr_delay_mult, max_steps);
}

int omnigs_f32_to_bf16(const float* in, uint16_t* out, size_t n, void* stream) {
  if (n > 0 && (!in || !out)) return fail_arg("missing pointer");
  OMNIGS_TRY(launch_to_bf16(in, out, n, (hipStream_t)stream), "to bf16");
  return OMNIGS_OK;
}

}  // extern "C"
