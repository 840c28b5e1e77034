// Host-side launchers of the gfx950 kernels (one translation unit per stage).
#pragma once

#include "common.hpp"

namespace omnigs {

struct PreprocessArgs {
  int P, D, M, W, H;
  const float* means3D;
  const float* shs;
  const float* colors_precomp;
  const float* opacities;
  const float* scales;
  float scale_modifier;
  const float* rotations;
  const float* cov3D_precomp;
  const float* viewmatrix;
  const float* campos;
  int seam_wrap;
  // colour decoder (colours from the decoder when both are set)
  const void* mlp_packed;
  const uint16_t* latents;  // BF16 [P,16]
  int tile_x = kTileX, tile_y = kTileY;
};

struct PreprocessBackwardArgs {
  int P, D, M, W, H;
  const float* means3D;
  const float* shs;
  const float* scales;
  float scale_modifier;
  const float* rotations;
  const float* cov3D_precomp;
  const float* viewmatrix;
  const float* campos;
  const int* radii;
  const uint8_t* clamped;
  const float* dL_dmean2D;  // [P,3]
  const float* dL_dconic;   // [P,4]
  const float* dL_dcolor;   // [P,3]
  float* dL_dmean3D;
  float* dL_dcov3D;
  float* dL_dsh;
  float* dL_dscale;
  float* dL_drot;
  float* dpx_dt;
  float* dpy_dt;
  // Fused gather (gather.tiles_touched set): the 2D gradients are summed
  // from the render backward's records in this kernel and written to these
  // outputs, instead of read from dL_dmean2D / dL_dconic / dL_dcolor.
  GatherSrc gather;
  float* out_dmean2D;
  float* out_dconic;
  float* out_dopacity;
  float* out_dcolor;
  // Colour decoder (act_list set): the Gaussians of each 64-Gaussian wave
  // whose dL/dcolor is not zero, in index order, at act_list[64 w ..], their
  // number at act_count[w]; dL/dlatent of the others is zeroed here (the
  // decoder's backward writes it for the listed ones only).
  uint32_t* act_list;
  uint32_t* act_count;
  float* dL_dlatent;
};

// preprocess.hip
hipError_t launch_preprocess(const PreprocessArgs& a, const GeomState& g, int* radii, hipStream_t s);
hipError_t launch_mark_visible(int P, uint8_t* present, hipStream_t s);
// Batched: items v P + g of the V views in vt (GeomState arrays sized V P).
hipError_t launch_preprocess_views(const PreprocessArgs& a, const float* viewmats, const float* campos,
                                   const ViewTable& vt, const GeomState& g, int* radii, hipStream_t s);

// sort.hip
// Inclusive (or exclusive) scan of n u32 values, optionally gathered through
// idx (out[i] = sum of in[idx[j]] for j <= i).  partials needs scan_blocks(n)+2 entries:
// the grand total lands in partials[nb] and a 32-bit overflow flag in partials[nb+1] (nb = scan_blocks(n)).
// side_in (optional): a short array (a few thousand values) whose exclusive
// scan goes to side_out in the same launches (the gradient slot bases).
// rect_in (optional, with idx and gather_buf): the values are the area words
// of these expansion records, gathered through idx (in is then unused), and
// the rest of each gathered record is written to rect_out in scan order.
hipError_t launch_scan(const uint32_t* in, const uint32_t* idx, uint32_t* out, size_t n, bool exclusive,
                       uint32_t* partials, hipStream_t s, uint32_t* gather_buf = nullptr,
                       const uint32_t* side_in = nullptr, size_t side_n = 0, uint32_t* side_out = nullptr,
                       const uint2* rect_in = nullptr, uint2* rect_out = nullptr);
// Per tile row y: (row_diff[0] + ... + row_diff[y]) mod 2^32 is the number of
// instances in row y (the rect widths of the Gaussians whose rects span it);
// row_diff has gy + 1 entries, cleared here.
hipError_t launch_row_counts(int P, const uint2* dup_rect, int gy, uint32_t* row_diff, hipStream_t s);
// rect_d[i] = dup_rect[order[i]] (the whole rects in depth order).
hipError_t launch_gather_rects(const uint2* dup_rect, const uint32_t* order, int P, uint2* rect_d, hipStream_t s);
// A band's expansion input: of the n depth-ordered rects rect_d, those with
// rows in [by0, by1), in order -- their Gaussians (from order), inclusive
// instance offsets and band rects {x0, y0 | w << 16} -> order_c / offsets_c /
// rect_c.  part_area / part_count: scan_blocks(n) + 2 entries each; the
// band's instance count and overflow flag land in part_area[scan_blocks(n)]
// and [+ 1], its owner count in part_count[scan_blocks(n)].  The band's
// gradient slots follow its owners: slot_words[g] = owner g's first slot (its
// exclusive offset), for the render backward and the gather (other entries
// are left alone).
hipError_t launch_band_owners(const uint2* rect_d, const uint32_t* order, size_t n, uint32_t by0, uint32_t by1,
                              uint32_t* part_area, uint32_t* part_count, uint32_t* order_c, uint32_t* offsets_c,
                              uint2* rect_c, uint32_t* slot_words, hipStream_t s);
// Stable LSD radix sort of (key, value) pairs over key bits [0, bits).  When
// vals_in is NULL the values are the element indices.  The result is in
// (keys_out, vals_out) if the pass count is even... callers use the returned
// pointers.  counts: sort_workspace_words(n) entries (partials is unused).
struct SortResult {
  uint32_t* keys;
  uint32_t* vals;
};
// n_dev (optional): the element count on the device, n then being the capacity.
// first_counted: the first pass's count table is already in counts (written
// by the expansion, whose blocks are the sort's tiles).
hipError_t launch_radix_sort(uint32_t* keys, uint32_t* keys_alt, const uint32_t* vals_in, uint32_t* vals,
                             uint32_t* vals_alt, size_t n, int bits, uint32_t* counts, uint32_t* partials,
                             hipStream_t s, SortResult* out, const uint32_t* n_dev = nullptr,
                             bool first_counted = false);
// Digit mask of the first pass of a `bits`-bit sort.
uint32_t radix_first_mask(int bits);
// no_record: the gradient slots' opacity words, set to kNoRecord (one per instance).
// rect_depth: the Gaussians' rects in depth order (written by launch_scan).
// counts (optional): receives the first-pass count table of a key_bits-bit
// tile sort of the expanded keys (launch_radix_sort's first_counted).
// ranges: cleared for launch_tile_ranges (its `zeroed`).  blend_ok: cleared
// (BinState::blend).  R_dev (optional): the live count {R, overflow flag} on
// the device; R is then the capacity the grid covers.  Every owner (entry of
// order / offsets / rect_depth) must hold at least one instance, save for
// owners past the last instance.  P_dev (optional): the owner count on the
// device, P then being its bound.
hipError_t launch_duplicate(int P, uint32_t R, const uint32_t* R_dev, const uint32_t* order, const uint32_t* offsets,
                            const uint2* rect_depth, TileGrid grid, uint32_t* tile_key, uint32_t* gid,
                            uint32_t* no_record, uint32_t* counts, int key_bits, uint2* ranges, uint32_t* blend_ok,
                            hipStream_t s, const uint32_t* P_dev = nullptr);
// zeroed: ranges were cleared by the expansion (else they are cleared here).
hipError_t launch_tile_ranges(const uint32_t* sorted_tile_keys, size_t R, uint2* ranges, int tiles,
                              hipStream_t s, const uint32_t* R_dev = nullptr, bool zeroed = false);
// Batched expansion: items = V P, capacity instances at most; R_dev is the
// device-side instance count (last inclusive offset); status [2] receives
// {count needed, 1 if it exceeds the capacity}.
hipError_t launch_duplicate_views(int items, uint32_t capacity, const uint32_t* R_dev, uint32_t* status,
                                  const uint32_t* order, const uint32_t* offsets, const uint2* rect_depth,
                                  const ViewTable& vt, uint32_t* tile_key, uint32_t* gid, uint32_t* no_record,
                                  uint32_t* counts, int key_bits, uint2* ranges, uint32_t* blend_ok, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
// keys[i] = order[i] / P: the view of each batch item in the given order.
hipError_t launch_view_keys(const uint32_t* order, size_t n, uint32_t P, uint32_t* keys, hipStream_t s);

// render.hip
// blend (optional, R bytes): bit w of instance i set when wave w of its tile
// blended it; *blend_ok is set to 1 (BinState::blend).
hipError_t launch_render_forward(TileGrid grid, const uint2* ranges, const uint32_t* point_list, int W, int H,
                                 const float4* rec, float* final_T, uint32_t* n_contrib, const float* bg,
                                 float* out_color, uint8_t* blend, uint32_t* blend_ok, hipStream_t s);
// The backward writes each staged (tile, Gaussian) instance's nine gradient
// terms to its slot -- grad_rec [nslots][2] float4 + grad_opacity [nslots] --
// at slot_base[Gaussian / kSlotBlock] + the record's slot word + the tile's
// position in the Gaussian's rect; slots it does not write must hold
// kNoRecord in grad_opacity.  gather_grads sums each item's slots into the
// reference's per-Gaussian arrays (every entry written).  P: Gaussians per
// view (items = V P); R_dev (optional): the live slot count.  blend (with
// *blend_ok == 1): the forward's per-wave blend bits, which select the
// (wave, instance) pairs to evaluate; else the quadrant masks do.
// [band_y0, band_y1) (band_y1 < 0: to the last row): the tile rows rendered;
// the lists and slots then hold the rect rows from band_y0 on.  slot_words
// (optional): each Gaussian's first slot, in place of slot_base plus the
// record's slot word.  blend_ranges (optional): the tile ranges of the list
// the blend bytes were recorded against (the forward's, when `ranges` are a
// band's of the same per-tile lists); else `ranges`.
hipError_t launch_render_backward(TileGrid grid, const uint2* ranges, const uint32_t* point_list, int W, int H,
                                  const float* bg, const float4* rec, const float* final_T, const uint32_t* n_contrib,
                                  const float* dL_dpix, const uint32_t* slot_base, uint32_t nslots, float4* grad_rec,
                                  float* grad_opacity, const uint8_t* blend, const uint32_t* blend_ok, hipStream_t s,
                                  int band_y0 = 0, int band_y1 = -1, const uint32_t* slot_words = nullptr,
                                  const uint2* blend_ranges = nullptr);
hipError_t launch_gather_grads(int items, int P, uint32_t nslots, const uint32_t* R_dev,
                               const uint32_t* tiles_touched, const uint32_t* slot_local, const uint32_t* slot_base,
                               const float4* grad_rec, const float* grad_opacity, float* dL_dmean2D,
                               float* dL_dconic, float* dL_dopacity, float* dL_dcolors, hipStream_t s);
// One band's records added to the per-Gaussian sums (the four arrays in the
// layouts launch_gather_grads writes): Gaussian i owns the slots from
// slot_words[i] on (launch_band_owners writes it), as many as its rect
// (dup_rect) has tiles in the rows [by0, by1).  The band holding a rect's
// first row writes the sums (bands run top to bottom); zero_empty (the first
// band) also writes the zeros of Gaussians with an empty rect.
hipError_t launch_band_gather(int P, const uint2* dup_rect, uint32_t by0, uint32_t by1, const uint32_t* slot_words,
                              uint32_t nslots, const float4* grad_rec, const float* grad_opacity, float* dL_dmean2D,
                              float* dL_dconic, float* dL_dopacity, float* dL_dcolors, bool zero_empty, hipStream_t s);

// Batched render over every tile of every view of vt; pixel arrays are the
// views' slices laid out one after the other (colours [3, H_v, W_v] per view).
hipError_t launch_render_forward_views(const ViewTable& vt, const uint2* ranges, const uint32_t* point_list,
                                       const float4* rec, float* final_T, uint32_t* n_contrib, const float* bg,
                                       float* out_color, uint8_t* blend, uint32_t* blend_ok, hipStream_t s);
hipError_t launch_render_backward_views(const ViewTable& vt, const uint2* ranges, const uint32_t* point_list,
                                        const float* bg, const float4* rec, const float* final_T,
                                        const uint32_t* n_contrib, const float* dL_dpix, const uint32_t* slot_base,
                                        uint32_t nslots, float4* grad_rec, float* grad_opacity, const uint8_t* blend,
                                        const uint32_t* blend_ok, hipStream_t s);
bool tile_shape_supported(int bx, int by);

// preprocess_backward.hip
hipError_t launch_preprocess_backward(const PreprocessBackwardArgs& a, hipStream_t s);
// Batched: the per-item arrays of a (radii, clamped, dL_dmean2D, dL_dconic,
// dL_dcolor) are [V P, ...]; parameter gradients are summed over views.
hipError_t launch_preprocess_backward_views(const PreprocessBackwardArgs& a, const float* viewmats,
                                            const float* campos, const ViewTable& vt, const float* dopacity_items,
                                            float* dL_dopacity, hipStream_t s);

// mlp.hip
constexpr int kMlpTrainBlocks = 256;  // persistent grid of the decoder training pass (one block per CU)
int mlp_train_blocks(int P);
size_t mlp_workspace_floats(int P);
hipError_t launch_mlp_pack(const float* theta, void* packed, hipStream_t s);
hipError_t launch_mlp_eval(int N, const uint16_t* latents, const float* dirs, const void* packed, float* rgb,
                           hipStream_t s);
hipError_t launch_mlp_distill(int P, int M, const float* shs, const float* latents, const void* packed, uint32_t seed,
                              int step, int num_dirs, float* dL_dlatent, float* dL_dtheta, float* loss, float* ws,
                              hipStream_t s, const float* means3D = nullptr, const float* view_box = nullptr);
hipError_t launch_mlp_backward(int P, const float* means3D, const float* campos, const int* radii,
                               const uint16_t* latents, const void* packed, const float* dL_dcolor, float* dL_dmean3D,
                               float* dL_dlatent, float* dL_dtheta, float* ws, hipStream_t s, int V = 1,
                               bool listed = false);
// The active-Gaussian list inside the decoder workspace (see
// PreprocessBackwardArgs::act_list): list [P rounded up to 256], then the
// per-wave counts.
uint32_t* mlp_act_list(float* ws, int P);
uint32_t* mlp_act_count(float* ws, int P);
struct AdamGroupDev {
  float* param;
  const float* grad;
  float* exp_avg;
  float* exp_avg_sq;
  uint16_t* param_bf16;  // optional BF16 copy of the updated parameter
  float step_size;       // lr / (1 - beta1^t)
  float bc2_sqrt;        // sqrt(1 - beta2^t)
  float eps;
};
constexpr int kAdamMaxGroups = 8;
struct AdamGroupTable {
  AdamGroupDev g[kAdamMaxGroups];
  size_t offset[kAdamMaxGroups + 1];  // element offsets; offset[n] = total
  int n;
  float one_minus_beta1, one_minus_beta2;
};
hipError_t launch_adam(const AdamGroupTable& tab, float beta1, float beta2, hipStream_t s);
hipError_t launch_to_bf16(const float* in, uint16_t* out, size_t n, hipStream_t s);

// train.hip
hipError_t launch_densify_stats(int P, const int* radii, const float* dmean2D, float* max_radii2D, float* accum,
                                float* denom, hipStream_t s);
float expon_lr(int step, float lr_init, float lr_final, int lr_delay_steps, float lr_delay_mult, int max_steps);

// loss.hip
// Floats of workspace the fused loss needs for a [C,H,W] image with `skip`
// bottom rows dropped (0 when nothing is kept).
size_t loss_workspace_floats(int C, int H, int W, int skip);
// The allocation-free loss's workspace: a file-scope device array of
// kLossScratchFloats floats (80 MiB) per GPU; loss_scratch gives its device
// address on the current device (looked up once per device).
constexpr size_t kLossScratchFloats = (size_t)80 << 18;
hipError_t loss_scratch(float** out);
hipError_t launch_loss(const float* x, const float* y, const float* mask, int C, int H, int W, int skip, float lambda,
                       float* workspace, float* dL_dx, float* out, hipStream_t s);

}  // namespace omnigs
