/*
 * C ABI of the MI355X (gfx950) equirectangular Gaussian rasterizer.
 *
 * Plain pointers and sizes only.  Every float pointer is DEVICE memory (FP32,
 * C-contiguous), exactly as the reference passes tensor data_ptr()s into
 * CudaRasterizer::LonlatRasterizer.  NULL marks an absent optional input,
 * as an empty tensor's data_ptr() does in the reference.
 *
 * The entry points replace, one for one:
 *   omnigs_lonlat_forward      <- LonlatRasterizer::forward
 *                                 (cuda_rasterizer/rasterizer.h:103-122,
 *                                  rasterizer_impl.cu:540-697)
 *   omnigs_lonlat_backward     <- LonlatRasterizer::backward
 *                                 (cuda_rasterizer/rasterizer.h:124-154,
 *                                  rasterizer_impl.cu:701-795)
 *   omnigs_lonlat_mark_visible <- LonlatRasterizer::markVisible
 *                                 (cuda_rasterizer/rasterizer.h:99-101,
 *                                  rasterizer_impl.cu:185-192)
 * The three scratch buffers are grown through caller callbacks, the C form of
 * the reference's std::function<char*(size_t)> resize functors
 * (src/rasterize_points.cu:41-47).  Their contents are private to this
 * library; the backward must receive the same three buffers and the R the
 * forward returned.
 *
 * All work is enqueued on `stream` (a hipStream_t; NULL = the default
 * stream).  The forward synchronises `stream` once, to read the number of
 * rendered instances R back to the host, as the reference does
 * (rasterizer_impl.cu:627-628).  The backward does not synchronise.
 */
#ifndef OMNIGS_LONLAT_H
#define OMNIGS_LONLAT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The LibTorch boundary maps them to the reference's
 * exceptions (src/rasterize_points.cu:72-75, :158-161; rasterizer_impl.cu:585-588). */
enum omnigs_status {
  OMNIGS_OK = 0,
  OMNIGS_ERR_INVALID_ARGUMENT = 1, /* bad sizes, missing required pointer, unsupported tile shape */
  OMNIGS_ERR_NEED_COLORS = 2,      /* NUM_CHANNELS != 3 without precomputed colours (kept for parity) */
  OMNIGS_ERR_ALLOC = 3,            /* a buffer callback returned NULL */
  OMNIGS_ERR_HIP = 4               /* a HIP runtime call or launch failed */
};

/* Grows a scratch buffer to at least `nbytes` and returns its device address
 * (NULL on failure).  `ctx` is passed through unchanged. */
typedef char* (*omnigs_resize_fn)(void* ctx, size_t nbytes);

typedef struct omnigs_buffer {
  omnigs_resize_fn resize;
  void* ctx;
} omnigs_buffer;

typedef struct omnigs_lonlat_forward_args {
  int P;                       /* number of Gaussians (at most 89478485 per call) */
  int D;                       /* active SH degree (0..3) */
  int M;                       /* SH coefficients per Gaussian (shs is [P,M,3]) */
  int width, height;           /* equirectangular image size in pixels (at most 32766 tile
                                  columns, 65535 tile rows and 2^21 - 1 tiles, else
                                  OMNIGS_ERR_INVALID_ARGUMENT) */
  const float* background;     /* [3] */
  const float* means3D;        /* [P,3] world positions */
  const float* shs;            /* [P,M,3] or NULL (then colors_precomp is used) */
  const float* colors_precomp; /* [P,3] or NULL */
  const float* opacities;      /* [P] (activated) */
  const float* scales;         /* [P,3] (activated) or NULL */
  float scale_modifier;
  const float* rotations;      /* [P,4] (w,x,y,z), used as given */
  const float* cov3D_precomp;  /* [P,6] upper triangle or NULL */
  const float* viewmatrix;     /* [16] Tcw^T row-major (= Tcw column-major) */
  const float* campos;         /* [3] camera centre in world coordinates */
  int prefiltered;             /* accepted and ignored, as in the reference's lonlat path */
  float* out_color;            /* [3,H,W] output image (planar CHW) */
  int* radii;                  /* [P] output radii in pixels, or NULL (kept internally) */
  int seam_wrap;               /* 0: clamp tile columns at the image edge (the reference's behaviour);
                                  1: wrap them around the longitude seam (extension, no reference
                                  counterpart; the width must be a multiple of 16, the tile width) */
  /* Colour decoder (extension, no reference counterpart): when both are set
   * the colours come from the decoder (omnigs_mlp_*) evaluated inside
   * preprocess, and shs / colors_precomp are ignored. */
  const void* mlp_packed;      /* omnigs_mlp_packed_bytes() bytes from omnigs_mlp_pack */
  const uint16_t* latents;     /* [P,16] BF16 per-Gaussian latents */
} omnigs_lonlat_forward_args;

typedef struct omnigs_lonlat_backward_args {
  int P, D, M, R;              /* R = value returned by the forward */
  int width, height;
  const float* background;     /* [3] */
  const float* means3D;        /* [P,3] */
  const float* shs;            /* [P,M,3] or NULL */
  const float* colors_precomp; /* [P,3] or NULL */
  const float* scales;         /* [P,3] or NULL */
  float scale_modifier;
  const float* rotations;      /* [P,4] or NULL */
  const float* cov3D_precomp;  /* [P,6] or NULL */
  const float* viewmatrix;     /* [16] */
  const float* campos;         /* [3] */
  const int* radii;            /* [P] radii from the forward, or NULL (internal copy) */
  char* geom_buffer;           /* the forward's three scratch buffers; the backward also keeps its
                                  per-instance gradient records in the binning buffer and a per-tile
                                  word in the image buffer, so two backward calls on one forward's
                                  buffers must not run concurrently */
  char* binning_buffer;
  char* image_buffer;
  const float* dL_dpix;        /* [3,H,W] upstream gradient of out_color */
  /* Outputs, every entry written (zero for Gaussians that touch no tile or
   * have radius 0).  The first four are the per-Gaussian sums of the render
   * backward's per-(tile, Gaussian) terms (gathered per Gaussian, without
   * device-scope atomics; within a tile the waves combine their sums with
   * LDS atomics, so the last bits can vary between runs, as with the
   * reference's atomics). */
  float* dL_dmean2D;           /* [P,3] NDC units (x 0.5W, 0.5H), z unused */
  float* dL_dconic;            /* [P,4] (2x2 row-major; slots 0,1,3 used, .y holds half the off-diagonal) */
  float* dL_dopacity;          /* [P] */
  float* dL_dcolor;            /* [P,3] */
  float* dL_dmean3D;           /* [P,3] */
  float* dL_dcov3D;            /* [P,6] */
  float* dL_dsh;               /* [P,M,3] or NULL when shs is NULL */
  float* dL_dscale;            /* [P,3] or NULL when scales is NULL */
  float* dL_drot;              /* [P,4] or NULL when scales is NULL */
  float* dpx_dt;               /* [P,3] optional: d(px)/d(t) rows of the lon/lat Jacobian */
  float* dpy_dt;               /* [P,3] optional */
  int seam_wrap;               /* must equal the forward's */
  /* Colour decoder (as in the forward).  When set, dL_dsh is unused and the
   * decoder's backward adds the view-direction term to dL_dmean3D. */
  const void* mlp_packed;
  const uint16_t* latents;     /* [P,16] BF16 */
  float* dL_dlatent;           /* [P,16] output */
  float* dL_dmlp;              /* [omnigs_mlp_param_count()] output: gradient of the FP32 parameter vector */
  void* mlp_workspace;         /* device scratch of omnigs_mlp_workspace_bytes(P) bytes */
  size_t mlp_workspace_bytes;
  /* The size in bytes the forward's last call of the binning buffer's
   * resize callback asked for (the reference's resize functor makes it the
   * tensor's size, resize_({N}); a caller whose allocation only grows must
   * remember the request, not its allocation).  The forward bins against the
   * previous call's capacity with headroom and keeps that binning whenever it
   * covers R, so the buffer's layout is that of a capacity this size
   * identifies.  Required when R > 0: a size that matches no layout for R
   * (0 included) is OMNIGS_ERR_INVALID_ARGUMENT. */
  size_t binning_bytes;
  /* The size in bytes the forward's last call of the image buffer's resize
   * callback asked for.  Read by omnigs_lonlat_backward_checkpointed only:
   * omnigs_lonlat_forward_checkpointed keeps the render's per-instance blend
   * bytes at the image buffer's end, and a size that covers them lets the
   * banded backward evaluate only the pairs the forward blended.  0 (or the
   * plain forward's size) selects the pairs by each ellipse's reach instead;
   * the gradients agree either way. */
  size_t image_bytes;
} omnigs_lonlat_backward_args;

/* Forward: preprocess, binning, render.  On success *num_rendered = R. */
int omnigs_lonlat_forward(const omnigs_lonlat_forward_args* args, omnigs_buffer geometry_buffer,
                          omnigs_buffer binning_buffer, omnigs_buffer image_buffer, void* stream,
                          int* num_rendered);

/* Forward for the checkpointed backward: as omnigs_lonlat_forward (same
 * image, radii and R), but the binning buffer is carved without the
 * backward's per-instance gradient records -- 17 of its 53 bytes per
 * instance are allocated -- so the forward's peak HBM drops with the
 * backward's.  The image buffer grows by one byte per instance of the
 * binning capacity (plus 256), which holds the render's record of which
 * waves blended each instance.  Its geometry and binning buffers serve no
 * stored backward: pass the image buffer (and its size, image_bytes) to
 * omnigs_lonlat_backward_checkpointed.  Has no
 * counterpart in the reference, whose autograd wrapper keeps all three
 * buffers for its backward (src/gaussian_rasterizer.cpp:85-98); it serves
 * memory-bound training callers. */
int omnigs_lonlat_forward_checkpointed(const omnigs_lonlat_forward_args* args, omnigs_buffer geometry_buffer,
                                       omnigs_buffer binning_buffer, omnigs_buffer image_buffer, void* stream,
                                       int* num_rendered);

/* Checkpointed form: re-runs preprocess and binning of a forward whose
 * geometry and binning buffers were released, into fresh buffers, against
 * the forward's kept image buffer (its per-pixel final_T and n_contrib are
 * left as the forward wrote them; the per-tile ranges are rewritten with the
 * same values).  `args` must be the forward's arguments (out_color unused).
 * All kernels are deterministic, so *num_rendered equals the forward's R.
 * Has no counterpart in the reference, whose autograd wrapper keeps all three
 * buffers (src/gaussian_rasterizer.cpp:85-98); it serves memory-bound
 * training callers. */
int omnigs_lonlat_recompute(const omnigs_lonlat_forward_args* args, omnigs_buffer geometry_buffer,
                            omnigs_buffer binning_buffer, char* image_buffer, void* stream, int* num_rendered);

/* Backward: render backward, then preprocess backward (covariance, 2D mean,
 * SH and scale/rotation terms in one kernel). */
int omnigs_lonlat_backward(const omnigs_lonlat_backward_args* args, void* stream);

/* Checkpointed backward in tile-row bands, for a forward whose geometry and
 * binning buffers were released: re-runs preprocess and the depth sort into
 * a fresh geometry buffer, then bins and runs the render backward one band of
 * whole tile rows at a time -- at most `band_instances` instances per band (0:
 * max(2^20, R / 4); a row holding more is a band of its own) -- adding each
 * band's per-Gaussian sums to the 2D gradient outputs, and ends with the
 * preprocess backward.  The binning buffer is grown to the largest band only,
 * so the gradient records (36 of the 53 bytes per instance) never exist for
 * all R instances at once.
 *   fwd: the forward's arguments (out_color unused; radii, if given, is
 *        rewritten with the forward's values).
 *   bwd: as for omnigs_lonlat_backward with R = the forward's R and
 *        image_buffer = the forward's kept image buffer (read only: its
 *        per-pixel state, per-tile ranges and, when image_bytes covers them,
 *        the checkpointed forward's blend bytes); geom_buffer and
 *        binning_buffer are unused (the callbacks grow fresh ones).
 * The outputs equal omnigs_lonlat_backward's up to the order in which a
 * Gaussian's per-tile terms are summed across bands.  The recomputed R must
 * equal bwd->R (else OMNIGS_ERR_INVALID_ARGUMENT).  Synchronises `stream`
 * once, to read the per-row instance counts.  *num_bands (optional): bands
 * run.  Has no counterpart in the reference, whose autograd wrapper keeps
 * all three buffers from the forward to the backward
 * (src/gaussian_rasterizer.cpp:85-98); it serves memory-bound training
 * callers, for whom the peak HBM of the backward is what bounds the image
 * size. */
int omnigs_lonlat_backward_checkpointed(const omnigs_lonlat_forward_args* fwd,
                                        const omnigs_lonlat_backward_args* bwd, omnigs_buffer geometry_buffer,
                                        omnigs_buffer binning_buffer, unsigned int band_instances, void* stream,
                                        int* num_bands);

/* markVisible for camera_type 3: every Gaussian is visible (present[i] = 1). */
int omnigs_lonlat_mark_visible(int P, uint8_t* present, void* stream);

/* Fused training loss and its image gradient, the mapper's per-iteration
 * loss (reference src/gaussian_mapper.cpp:384-407 with loss_utils.h:31-129):
 *   masked = image * mask                      (mask [H,W], NULL = ones)
 *   loss   = (1 - lambda) mean|masked - gt| + lambda (1 - mean SSIM(masked, gt))
 * over the top height - skip_bottom rows (skip_bottom = round(H *
 * skip_bottom_ratio) in the reference; 0 keeps every row).  SSIM is
 * loss_utils::ssim: 11x11 Gaussian window with sigma 1.5, zero padding of 5,
 * C1 = 0.01^2, C2 = 0.03^2.  lambda = 0 computes the L1 term alone.
 * image, gt_image, dL_dimage: [channels,height,width] FP32 device arrays;
 * dL_dimage is d loss / d image over every row (zero in the dropped rows).
 * loss_out: 3 device floats {loss, L1, SSIM (0 when lambda = 0)}.
 * workspace: device memory of omnigs_loss_workspace_bytes() bytes.  Replaces
 * the torch ops plus loss.backward() the reference runs for this. */
size_t omnigs_loss_workspace_bytes(int channels, int height, int width, int skip_bottom);
int omnigs_photometric_loss(const float* image, const float* gt_image, const float* mask, int channels, int height,
                            int width, int skip_bottom, float lambda_dssim, void* workspace, size_t workspace_bytes,
                            float* dL_dimage, float* loss_out, void* stream);

/* omnigs_photometric_loss without a workspace argument and without any
 * allocation: no counterpart in the reference; for callers that cannot
 * allocate (code under HIP graph capture, tight training or serving loops,
 * harnesses that forbid hipMalloc).  The workspace is a file-scope device
 * array of OMNIGS_LOSS_NOALLOC_MAX_BYTES that every GPU the library runs on
 * holds until the process ends; its device address is looked up on the first
 * call on each device.  Images whose omnigs_loss_workspace_bytes() exceeds it
 * (beyond 3 x 1024 x 2048 with lambda_dssim != 0) are refused with
 * OMNIGS_ERR_INVALID_ARGUMENT.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (two streams, two threads) race.  Same
 * arguments and results as omnigs_photometric_loss. */
#define OMNIGS_LOSS_NOALLOC_MAX_BYTES (80u << 20)
int omnigs_photometric_loss_noalloc(const float* image, const float* gt_image, const float* mask, int channels,
                                    int height, int width, int skip_bottom, float lambda_dssim, float* dL_dimage,
                                    float* loss_out, void* stream);

/* Batched multi-view forward / backward (extension; the reference renders one
 * view per RasterizeGaussiansCUDA call, rasterize_points.cu:49-164).  V <= 32
 * equirectangular views of one Gaussian set, each with its own size and pose,
 * in one launch per stage: preprocess reads each Gaussian once for all views,
 * binning runs over (view, Gaussian) items with view-major tile ids, the
 * render kernels cover every tile of every view, and the backward sums each
 * Gaussian's gradients over the views in registers.  Per-view images match
 * V single-view forwards.  Tile shape 16x16 (the reference's), 32x8 or 8x32.
 * No host synchronisation and no allocation happen inside (the instance
 * count stays on the device, bounded by the workspace's capacity), so a
 * caller may capture forward + backward into one HIP graph.  status (device,
 * 2 uints) receives {instances needed, 1 if that exceeded capacity}, with
 * 0xFFFFFFFF needed when the count does not fit 32 bits; after
 * an overflow the images are incomplete and the caller re-runs with a larger
 * capacity.  Pixel arrays (out_color, dL_dpix) hold view v's [3, H_v, W_v]
 * block at float offset 3 * sum_{u<v} W_u H_u. */
#define OMNIGS_MAX_VIEWS 32
typedef struct omnigs_lonlat_batch_args {
  int P, D, M, V;
  int tile_x, tile_y;
  int seam_wrap;
  const int* widths;           /* host [V] */
  const int* heights;          /* host [V] */
  const float* viewmatrices;   /* device [V,16] */
  const float* campos;         /* device [V,3] */
  const float* background;     /* device [3] */
  const float* means3D;        /* [P,3] */
  const float* shs;            /* [P,M,3] or NULL */
  const float* colors_precomp; /* [P,3] or NULL */
  const float* opacities;      /* [P] */
  const float* scales;         /* [P,3] or NULL */
  float scale_modifier;
  const float* rotations;      /* [P,4] or NULL */
  const float* cov3D_precomp;  /* [P,6] or NULL */
  const void* mlp_packed;      /* colour decoder (with latents) or NULL */
  const uint16_t* latents;
  unsigned int capacity;       /* instance capacity of the workspace, all views together */
  void* workspace;             /* omnigs_lonlat_batch_workspace_bytes() bytes, kept between forward and backward */
  size_t workspace_bytes;
  float* out_color;            /* device, all views */
  unsigned int* status;        /* device [2] */
} omnigs_lonlat_batch_args;

typedef struct omnigs_lonlat_batch_grads {
  const float* dL_dpix;        /* device, layout of out_color */
  float* dL_dmean3D;           /* [P,3] */
  float* dL_dopacity;          /* [P] */
  float* dL_dcov3D;            /* [P,6] */
  float* dL_dsh;               /* [P,M,3] when shs is given */
  float* dL_dscale;            /* [P,3] when scales is given */
  float* dL_drot;              /* [P,4] when scales is given */
  float* dL_dlatent;           /* [P,16] with the decoder */
  float* dL_dmlp;              /* [omnigs_mlp_param_count()] with the decoder */
  void* mlp_workspace;         /* omnigs_mlp_workspace_bytes(P) with the decoder */
  size_t mlp_workspace_bytes;
} omnigs_lonlat_batch_grads;

size_t omnigs_lonlat_batch_workspace_bytes(int P, int V, const int* widths, const int* heights, int tile_x,
                                           int tile_y, unsigned int capacity);
int omnigs_lonlat_batch_forward(const omnigs_lonlat_batch_args* args, void* stream);
int omnigs_lonlat_batch_backward(const omnigs_lonlat_batch_args* args, const omnigs_lonlat_batch_grads* grads,
                                 void* stream);

/* Colour decoder (extension; no counterpart in the reference, whose colours
 * come from computeColorFromSH, cuda_rasterizer/forward.cu:30-83).
 * rgb = sigmoid(W3 relu(W2 relu(W1 [latent(16), Y2(dir)(9)] + b1) + b2) + b3)
 * with Y2 the degree-2 real SH basis of the view direction, BF16 operands and
 * FP32 accumulation on the MFMA unit.  The FP32 parameter vector (theta,
 * omnigs_mlp_param_count() floats) holds, in order: W1 [64][25], b1 [64],
 * W2 [64][64], b2 [64], W3 [3][64], b3 [3] (weights row-major [out][in]).
 * omnigs_mlp_pack converts it into the kernels' BF16 operand image. */
int omnigs_mlp_param_count(void);
size_t omnigs_mlp_packed_bytes(void);
int omnigs_mlp_pack(const float* theta, void* packed, void* stream);
/* rgb[i] = decoder(latents[i], dirs[i]) for N unit directions [N,3]. */
int omnigs_mlp_eval(int N, const uint16_t* latents, const float* dirs, const void* packed, float* rgb, void* stream);
/* Device scratch of one decoder training pass over P Gaussians. */
size_t omnigs_mlp_workspace_bytes(int P);

/* One distillation pass: every Gaussian takes one of its num_dirs seeded
 * view directions (chosen by seed and step), the teacher colour is its SH
 * colour there (degree 3, + 0.5, clamped at 0, as computeColorFromSH), and
 * the loss is the mean over P x 3 of (decoder - teacher)^2.  Writes the
 * gradients of the FP32 latents and of theta, and the loss.
 * Directions: with means3D NULL, uniform on the sphere
 * (omnigs_mlp_seeded_direction); with means3D [P,3] given, direction k of
 * Gaussian g is the view direction unit(mean_g - c) from the seeded view
 * centre c = omnigs_mlp_seeded_view_centre(seed, g, k, view_box), uniform in
 * the box view_box = {lo x, lo y, lo z, hi x, hi y, hi z} where the cameras
 * that will render the scene stand -- the directions the colours are
 * actually seen from, as the reference's SH are fitted on its training
 * views. */
typedef struct omnigs_mlp_distill_args {
  int P, M;                /* shs is [P,M,3], M >= 16 */
  const float* shs;
  const float* latents;    /* [P,16] FP32 master latents (rounded to BF16 for the MFMA) */
  const void* mlp_packed;
  unsigned int seed;
  int step;
  int num_dirs;            /* directions per Gaussian (>= 1) */
  float* dL_dlatent;       /* [P,16] */
  float* dL_dmlp;          /* [param_count] */
  float* loss;             /* 1 float */
  void* workspace;         /* omnigs_mlp_workspace_bytes(P) */
  size_t workspace_bytes;
  const float* means3D;    /* [P,3] or NULL (directions on the sphere) */
  float view_box[6];       /* camera-centre box with means3D: lo xyz, hi xyz */
} omnigs_mlp_distill_args;
int omnigs_mlp_distill_step(const omnigs_mlp_distill_args* args, void* stream);

/* Seeded view centre k of Gaussian g in view_box (lo xyz, hi xyz), used by
 * the distillation when means3D is given. */
void omnigs_mlp_seeded_view_centre(unsigned int seed, unsigned int g, unsigned int k, const float view_box[6],
                                   float out[3]);
/* Seeded direction k of Gaussian g used by the distillation (unit vector). */
void omnigs_mlp_seeded_direction(unsigned int seed, unsigned int g, unsigned int k, float out[3]);

/* Fused Adam step over up to 8 parameter groups in one launch, the update of
 * torch::optim::Adam as the reference's GaussianModel uses it (betas (0.9,
 * 0.999) by default, eps 1e-15, no weight decay; gaussian_model.cpp:485-518):
 *   m = b1 m + (1 - b1) g;  v = b2 v + (1 - b2) g^2;
 *   p -= lr / (1 - b1^step) * m / (sqrt(v) / sqrt(1 - b2^step) + eps).
 * step is the 1-based step count of this update (torch increments it first).
 * param_bf16, when set, receives a BF16 copy of the updated parameter. */
typedef struct omnigs_adam_group {
  float* param;
  const float* grad;
  float* exp_avg;
  float* exp_avg_sq;
  uint16_t* param_bf16;
  size_t numel;
  float lr;
  float eps;
  long long step;
} omnigs_adam_group;
int omnigs_adam_step(const omnigs_adam_group* groups, int n_groups, float beta1, float beta2, void* stream);

/* Densification statistics of one training step in one pass (the reference's
 * max_radii2D update, gaussian_mapper.cpp:419-428, and
 * GaussianModel::addDensificationStats, gaussian_model.cpp:839-853): for
 * radii[i] > 0, max_radii2D[i] = max(max_radii2D[i], radii[i]),
 * xyz_gradient_accum[i] += |dL_dmean2D[i].xy|, denom[i] += 1.  dL_dmean2D is
 * the backward's [P,3] screen-space gradient; the other arrays are [P]. */
int omnigs_densification_stats(int P, const int* radii, const float* dL_dmean2D, float* max_radii2D,
                               float* xyz_gradient_accum, float* denom, void* stream);

/* Position learning rate of GaussianModel::exponLrFunc
 * (gaussian_model.cpp:1140-1153): log-linear from lr_init at step 0 to
 * lr_final at max_steps, with the optional sine warm-up from lr_delay_mult
 * over lr_delay_steps; 0 for step < 0 or lr_init = lr_final = 0. */
float omnigs_expon_lr(int step, float lr_init, float lr_final, int lr_delay_steps, float lr_delay_mult,
                      int max_steps);

/* FP32 -> BF16 (round to nearest even). */
int omnigs_f32_to_bf16(const float* in, uint16_t* out, size_t n, void* stream);

/* 3DGS PLY files (host memory), the vertex layout the reference's
 * GaussianModel::savePly / loadPly write and read through tinyply
 * (src/gaussian_model.cpp:860-1069): x y z nx ny nz f_dc_0..2
 * f_rest_0..(3K-1) opacity scale_0..2 rot_0..3, float32, normals zero,
 * f_rest channel-major (f_rest_{c K + j} = features_rest[p][j][c],
 * K = (sh_degree + 1)^2 - 1); raw (pre-activation) opacity, scale, rotation.
 * features_dc is [P,1,3], features_rest [P,K,3], opacity [P], scaling [P,3],
 * rotation [P,4].  The writer emits binary little-endian files with
 * tinyply's header text; the reader takes any property order, scalar type
 * and format (binary either byte order, or ASCII). */
int omnigs_ply_write(const char* path, long long P, int sh_degree, const float* xyz, const float* features_dc,
                     const float* features_rest, const float* opacity, const float* scaling, const float* rotation);
int omnigs_ply_count(const char* path, long long* P);
int omnigs_ply_read(const char* path, long long P, int sh_degree, float* xyz, float* features_dc,
                    float* features_rest, float* opacity, float* scaling, float* rotation);
const char* omnigs_ply_last_error(void);

/* Byte sizes the forward requests from each buffer callback. */
size_t omnigs_lonlat_geometry_bytes(int P);
size_t omnigs_lonlat_binning_bytes(int R);
size_t omnigs_lonlat_image_bytes(int width, int height);

/* Device addresses of the arrays inside the three private buffers, for
 * inspection (parity tests compare them with a CPU restatement stage by
 * stage).  Has no counterpart in the reference, whose buffer carving
 * (rasterizer_impl.h:48-102) is not exported.  Any of the three buffer
 * pointers may be NULL; the matching fields are then NULL. */
typedef struct omnigs_lonlat_state_view {
  /* geometry buffer, per Gaussian */
  const uint32_t* depth_key;     /* the depth-sort keys in depth order, aligned with depth_order:
                                    depth_key[i] = float bits of |t| of Gaussian depth_order[i]
                                    (0xFFFFFFFF when culled) */
  const int32_t* radii;          /* internal radii (written when the forward got radii == NULL) */
  const float* render_records;   /* [P,12]: x, y, conic A, B, C, opacity, r, g, b, then three 32-bit
                                    integers stored in the float slots: radius, the Gaussian's first
                                    gradient slot within its 256-Gaussian block, and 0 */
  const uint8_t* clamped;        /* [P] bit c = channel c clamped */
  const uint32_t* tiles_touched; /* [P] */
  const uint32_t* depth_order;   /* [P] Gaussian ids by (depth bits, id) */
  const uint32_t* offsets;       /* [P] inclusive instance offsets in depth order */
  /* binning buffer, per instance */
  const uint32_t* point_list;    /* [R] Gaussian id per sorted instance */
  const uint32_t* tile_keys;     /* [R] tile id per sorted instance */
  /* image buffer */
  const float* final_T;          /* [H,W] */
  const uint32_t* n_contrib;     /* [H,W] */
  const uint32_t* ranges;        /* [tiles,2] */
  /* binning buffer, per sorted instance: bit w set when wave w of its tile (8x8 quadrant w, row-major
     over the tile's quadrants) blended it in the render forward; valid once omnigs_lonlat_forward ran
     (omnigs_lonlat_forward_checkpointed writes these bytes at its image buffer's end instead, after
     omnigs_lonlat_image_bytes(width, height) + 256 bytes) */
  const uint8_t* blend;
} omnigs_lonlat_state_view;

int omnigs_lonlat_state(const char* geom_buffer, const char* binning_buffer, const char* image_buffer, int P, int R,
                        int width, int height, omnigs_lonlat_state_view* out,
                        size_t binning_bytes /* as omnigs_lonlat_backward_args::binning_bytes */);

/* The same view of a batched workspace (omnigs_lonlat_batch_forward /
 * _backward), for stage-by-stage parity tests.  Per-item arrays hold V P
 * items, item v P + g being Gaussian g in view v; point_list / tile_keys hold
 * the batch's sorted instances (their count is status[0] of the forward),
 * Gaussian ids being item ids; final_T / n_contrib hold every view's pixels in
 * the layout of out_color (one plane each).  Tile t of view v has the key
 * key_base(v) + t: with key_shift 0 (compact keys) key_base(v) is the sum of
 * the earlier views' tile counts and the instances run view by view; with
 * key_shift > 0 (padded keys, chosen when they save the tile sort a pass)
 * key_base(v) = v << key_shift and the instances are ordered by the key's low
 * key_shift bits, each (view, tile) group contiguous.  Either way each group
 * is in depth order and ranges (key_count entries) holds its [start, end) in
 * point_list per key.  depth_key / depth_order are the depth sort's (by
 * depth, item id).  The 2D gradients are the backward's per-item sums (the
 * single-view backward's dL_dmean2D, dL_dconic [4], dL_dopacity and
 * dL_dcolor, per item).  No counterpart in the reference. */
typedef struct omnigs_lonlat_batch_state_view {
  omnigs_lonlat_state_view s;
  const float* dL_dmean2D;  /* [V P, 3] */
  const float* dL_dconic;   /* [V P, 4] */
  const float* dL_dopacity; /* [V P] */
  const float* dL_dcolor;   /* [V P, 3] */
  unsigned int key_shift;
  unsigned int key_count;
} omnigs_lonlat_batch_state_view;
int omnigs_lonlat_batch_state(const omnigs_lonlat_batch_args* args, omnigs_lonlat_batch_state_view* out);

/* Stage timers: when enabled, the selected stages of the forward and backward
 * are bracketed by HIP events recorded on the launch stream.  read() waits for
 * the recorded events, adds each stage's elapsed milliseconds into
 * ms[stage] and count[stage] (arrays of OMNIGS_NUM_STAGES), and clears them.
 * Measurement support only; no counterpart in the reference. */
enum omnigs_stage {
  OMNIGS_STAGE_PREPROCESS = 0,
  OMNIGS_STAGE_DEPTH_SORT = 1,
  OMNIGS_STAGE_SCAN = 2,
  OMNIGS_STAGE_DUPLICATE = 3,
  OMNIGS_STAGE_TILE_SORT = 4,
  OMNIGS_STAGE_RANGES = 5,
  OMNIGS_STAGE_RENDER = 6,
  OMNIGS_STAGE_GATHER_GRADS = 7, /* per-Gaussian sum of the per-instance records (batched calls; the
                                    single-view backward sums them inside the preprocess backward) */
  OMNIGS_STAGE_RENDER_BACKWARD = 8,
  OMNIGS_STAGE_PREPROCESS_BACKWARD = 9,
  OMNIGS_STAGE_DECODER_BACKWARD = 10,
  OMNIGS_NUM_STAGES = 11
};
/* mask: 0 = off, otherwise bit s (1 << s) times stage s; OMNIGS_ALL_STAGES
 * times every stage.  Timing one stage costs its launch stream two event
 * records per call. */
#define OMNIGS_ALL_STAGES ((1 << OMNIGS_NUM_STAGES) - 1)
int omnigs_profile_enable(int mask);
int omnigs_profile_read(double* ms, int* count);
const char* omnigs_stage_name(int stage);

/* Human-readable text of a status code, and of the last HIP error seen. */
const char* omnigs_status_string(int status);
const char* omnigs_last_error(void);

/* Library version string. */
const char* omnigs_version(void);

#ifdef __cplusplus
}
#endif

#endif /* OMNIGS_LONLAT_H */
