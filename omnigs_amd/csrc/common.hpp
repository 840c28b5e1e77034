// Shared device helpers and buffer layouts of the gfx950 equirectangular
// rasterizer.  Written for CDNA4: 64-lane wavefronts, 160 KB LDS per CU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace omnigs {

constexpr int kWave = 64;
constexpr int kTileX = 16;  // tile width in pixels (reference config.h:26)
constexpr int kTileY = 16;  // tile height in pixels (reference config.h:27)
constexpr int kTilePixels = kTileX * kTileY;
constexpr int kChannels = 3;

// Real SH constants (textbook values; the reference lists them at
// cuda_rasterizer/auxiliary.h:32-49).
constexpr float kSH0 = 0.28209479177387814f;
constexpr float kSH1 = 0.4886025119029199f;
constexpr float kSH2_0 = 1.0925484305920792f;
constexpr float kSH2_1 = -1.0925484305920792f;
constexpr float kSH2_2 = 0.31539156525252005f;
constexpr float kSH2_3 = -1.0925484305920792f;
constexpr float kSH2_4 = 0.5462742152960396f;
constexpr float kSH3_0 = -0.5900435899266435f;
constexpr float kSH3_1 = 2.890611442640554f;
constexpr float kSH3_2 = -0.4570457994644658f;
constexpr float kSH3_3 = 0.3731763325901154f;
constexpr float kSH3_4 = -0.4570457994644658f;
constexpr float kSH3_5 = 1.445305721320277f;
constexpr float kSH3_6 = -0.5900435899266435f;

constexpr float kInvPi = 0.318309886183790671537767526745028724f;
constexpr float kTwoInvPi = 0.636619772367581343075535053490057448f;

struct TileGrid {
  int gx, gy;
  int wrap;    // 1: tile columns wrap around the longitude seam (extension; 0 = the reference's clamp)
  int bx, by;  // tile size in pixels (16 x 16 is the reference's; 32 x 8 and 8 x 32 are extensions)
  __host__ __device__ int tiles() const { return gx * gy; }
};
inline TileGrid tile_grid(int W, int H, int wrap = 0, int bx = kTileX, int by = kTileY) {
  return {(W + bx - 1) / bx, (H + by - 1) / by, wrap, bx, by};
}

// Views of a batched launch (an extension: the reference renders one view per
// call).  Item i of the batch is Gaussian i % P in view i / P; global tile ids
// and pixel offsets run view after view.  Passed to kernels by value.
constexpr int kMaxViews = 32;
struct ViewTable {
  int V, P;
  int bx, by, wrap;
  int W[kMaxViews], H[kMaxViews], gx[kMaxViews], gy[kMaxViews];
  uint32_t tile_base[kMaxViews + 1];  // first global tile of view v; [V] = total tiles
  uint32_t pix_base[kMaxViews + 1];   // first pixel of view v; [V] = total pixels
  // Tile keys of the expansion, tile sort and ranges: tile t of view v has
  // key key_base(v) + t.  Compact (key_shift 0): the global tile ids.  Padded
  // (key_shift > 0): v << key_shift | t, every view's tiles padded to
  // 2^key_shift keys, used when a sort over those low bits alone takes fewer
  // passes than over the compact ids; the instances are then expanded view by
  // view (the depth order re-sorted by view) and the tile sort orders them by
  // the low bits only, leaving each (view, tile) group contiguous and in depth
  // order.
  int key_shift;
  uint32_t key_count;  // keys (range entries): tile_base[V], or V << key_shift
  int sort_bits;       // key bits the tile sort orders by
  __host__ __device__ uint32_t key_base(int v) const {
    return key_shift ? (uint32_t)v << key_shift : tile_base[v];
  }
  __host__ __device__ TileGrid grid(int v) const { return {gx[v], gy[v], wrap, bx, by}; }
  // view owning global tile t (V is small: a scan over constant kernel arguments)
  __device__ int view_of_tile(uint32_t t) const {
    int v = 0;
    while (v + 1 < V && t >= tile_base[v + 1]) ++v;
    return v;
  }
};

// Tile rectangle [x0,x1) x [y0,y1) of a Gaussian at pixel (px,py) with integer
// radius r.  The float evaluation order is the reference's getRect
// (auxiliary.h:56-66) so that tile sets agree bit for bit:
// lo = (int)((p - r) / 16), hi = (int)((((p + r) + 16) - 1) / 16), clamped.
struct Rect {
  int x0, y0, x1, y1;
};
__device__ __forceinline__ int rect_lo(float p, int r, int tile, int g) {
  int v = (int)((p - (float)r) / (float)tile);
  return min(g, max(0, v));
}
__device__ __forceinline__ int rect_hi(float p, int r, int tile, int g) {
  int v = (int)((((p + (float)r) + (float)tile) - 1.0f) / (float)tile);
  return min(g, max(0, v));
}
// Seam-wrap mode (an extension; the reference clamps at the image edge,
// auxiliary.h:56-66): the column range is floored instead of truncated, may
// leave [0, gx) and is read modulo gx; a range wider than the ring is the ring.
__device__ __forceinline__ Rect tile_rect(float2 p, int r, TileGrid g) {
  if (g.wrap) {
    int x0 = (int)floorf((p.x - (float)r) / (float)g.bx);
    int x1 = (int)floorf((((p.x + (float)r) + (float)g.bx) - 1.0f) / (float)g.bx);
    if (x1 - x0 >= g.gx) {
      x0 = 0;
      x1 = g.gx;
    }
    return {x0, rect_lo(p.y, r, g.by, g.gy), x1, rect_hi(p.y, r, g.by, g.gy)};
  }
  return {rect_lo(p.x, r, g.bx, g.gx), rect_lo(p.y, r, g.by, g.gy), rect_hi(p.x, r, g.bx, g.gx),
          rect_hi(p.y, r, g.by, g.gy)};
}
// A Gaussian's expansion record (see GeomState::dup_rect), 8 bytes:
// {x0 (low 16 bits, signed) | h << 16, y0 | w << 16}, the area being w h; a
// culled one has h = 0 and a one-column rect.  Fits while the tile grid is
// below kMaxGridCols columns (|x0| <= gx + 1 in seam-wrap mode) and 65536
// rows.
constexpr int kMaxGridCols = 32766;
__device__ __forceinline__ uint2 pack_dup_rect(bool live, int area, const Rect& rc) {
  if (!live || area <= 0) return make_uint2(0u, 1u << 16);
  const uint32_t w = (uint32_t)(rc.x1 - rc.x0), h = (uint32_t)(rc.y1 - rc.y0);  // area = w h > 0
  return make_uint2(((uint32_t)rc.x0 & 0xFFFFu) | (h << 16), (uint32_t)rc.y0 | (w << 16));
}
// Its area and the rect {x0, y0 | w << 16} that the expansion reads.
__device__ __forceinline__ uint32_t dup_rect_area(uint2 r) { return (r.x >> 16) * (r.y >> 16); }
__device__ __forceinline__ uint2 dup_rect_xy(uint2 r) {
  return make_uint2((uint32_t)(int32_t)(int16_t)(r.x & 0xFFFFu), r.y);
}
// The record of the rect's rows inside the tile-row band [by0, by1) (same
// columns), or the empty record when it has none there.  The band [0, 65535)
// holds every rect whole (rows fit 16 bits).
__device__ __forceinline__ uint2 band_dup_rect(uint2 r, uint32_t by0, uint32_t by1) {
  const uint32_t y0 = r.y & 0xFFFFu, y1 = y0 + (r.x >> 16);
  const uint32_t b0 = max(y0, by0), b1 = min(y1, by1);
  if (b1 <= b0) return make_uint2(0u, 1u << 16);
  return make_uint2((r.x & 0xFFFFu) | ((b1 - b0) << 16), b0 | (r.y & 0xFFFF0000u));
}
// Tile column of rect column x (identity in clamp mode).
__device__ __forceinline__ int wrap_col(int x, TileGrid g) { return g.wrap ? ((x % g.gx) + g.gx) % g.gx : x; }

// Position of tile (tx, ty) in a Gaussian's rect enumeration -- rows
// [y0, y1), columns x0 .. x0 + w - 1 read modulo gx in seam-wrap mode -- which
// is the order the duplicate pass lays out the Gaussian's instances in.
__device__ __forceinline__ uint32_t rect_local(const Rect& rc, int w, int tx, int ty, TileGrid g) {
  int c = tx - rc.x0;
  if (g.wrap) c = ((c % g.gx) + g.gx) % g.gx;
  return (uint32_t)((ty - rc.y0) * w + c);
}

// ---------------------------------------------------------------------------
// Buffer layouts.  The three byte buffers are private to this library; the
// backward re-derives every array's offset from (P), (R, tiles) and (W, H),
// like the reference's GeometryState / BinningState / ImageState carving
// (rasterizer_impl.h:48-102), but with this library's own contents.
// ---------------------------------------------------------------------------
constexpr size_t kAlign = 256;
__host__ __device__ inline size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// Aligns absolute addresses (like the reference's obtain(), rasterizer_impl.h:37-43);
// sizes computed from a NULL base carry kAlign bytes of slack for that.
template <class T> struct Carver {
  size_t off = 0;
  char* base;
  __host__ __device__ explicit Carver(char* b) : base(b) {}
  template <class U> U* take(size_t count) {
    const uintptr_t a = align_up(reinterpret_cast<uintptr_t>(base) + off);
    off = a - reinterpret_cast<uintptr_t>(base);
    U* p = reinterpret_cast<U*>(a);
    off += count * sizeof(U);
    return p;
  }
};

// Radix sort geometry: 256 threads x 16 keys per block, 8-bit digits.
constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
inline size_t sort_blocks(size_t n) { return (n + kSortTile - 1) / kSortTile; }
// Radix sort workspace (uint32 words): per-block digit counts [kRadix][blocks]
// (scanned in place along each digit's row) and the kRadix digit totals.
inline size_t sort_workspace_words(size_t n) { return (size_t)kRadix * (sort_blocks(n) + 1); }

// Scan geometry: 256 threads x 16 items per block.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;
inline size_t scan_blocks(size_t n) { return (n + kScanTile - 1) / kScanTile; }

// ---------------------------------------------------------------------------
// Wave / block scans (64-lane wavefronts, 4 waves per 256-thread block).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v, int lane) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t n = __shfl_up(v, o, kWave);
    if (lane >= o) v += n;
  }
  return v;
}

// Exclusive scan of one value per thread across a 256-thread block.
// Returns the exclusive prefix; *total receives the block sum.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[kScanThreads / kWave];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const uint32_t inc = wave_inclusive_scan(v, lane);
  if (lane == kWave - 1) wsum[w] = inc;
  __syncthreads();
  uint32_t woff = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kScanThreads / kWave; ++i) {
    const uint32_t s = wsum[i];
    woff += i < w ? s : 0u;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return woff + inc - v;
}

// Gradient slots are numbered per block of kSlotBlock Gaussians (the
// preprocess's thread blocks): block b's slots start at slot_base[b].
constexpr int kSlotBlock = 256;
// The render kernels hold a tile-list position times 48 (a staging byte
// offset) in 32 bits, so a tile's list -- at most P entries -- is bounded.
constexpr int kMaxGaussians = 89478485;  // floor((2^32 - 1) / 48)
// Tiles per image (per view in a batch).  A Gaussian touches at most this
// many, so a scan block of kScanTile tile counts sums below 2^32 and only the
// scan's grand total can wrap (scan_partials_kernel flags it).
constexpr int kMaxTiles = (1 << 21) - 1;
inline size_t slot_blocks(size_t n) { return (n + kSlotBlock - 1) / kSlotBlock; }
// Opacity word of a gradient slot that holds no record (a NaN pattern the
// hardware never produces: its NaNs are the canonical 0x7FC00000).
constexpr uint32_t kNoRecord = 0xFFFFFFFFu;

// Per-Gaussian state written by the forward preprocess.
struct GeomState {
  uint32_t* depth_key;     // float bits of the camera distance (sort key); ~0 if culled
  int32_t* radii;          // internal radii (used when the caller passes none)
  // Render record, 48 B per Gaussian so one gather fetches everything the
  // tile kernels need: {x, y, A, B}, {C, opacity, r, g}, {b, radius, slot, 0}
  // (pixel-space mean, conic, opacity, colour -- SH-derived or precomputed --,
  // the radius as int bits, and `slot`, the offset of the Gaussian's first
  // gradient slot inside its 256-Gaussian block: the block's slots start at
  // slot_base[block]).
  float4* rec;
  uint8_t* clamped;        // bit c set when channel c was clamped at 0
  uint32_t* tiles_touched; // rect area
  // Expansion input per Gaussian, in index order (pack_dup_rect, 8 bytes);
  // the offset scan gathers it once through the depth order into rect_depth
  // {x0, y0 | w << 16}, which the expansion then reads contiguously.
  uint2* dup_rect;
  uint2* rect_depth;
  uint32_t* slot_local;    // copy of the record's `slot` (dense, for the gradient gather)
  uint32_t* block_slots;   // per 256-Gaussian block: its Gaussians' total rect area
  uint32_t* slot_base;     // exclusive scan of block_slots (written by the forward scan)
  uint32_t* order;         // Gaussians sorted by (depth bits, index)
  uint32_t* order_alt;     // ping-pong for the depth sort
  uint32_t* key_alt;       // ping-pong keys for the depth sort
  uint32_t* offsets;       // inclusive scan of tiles_touched in depth order
  uint32_t* scan_partials; // scan block sums
  uint32_t* sort_counts;   // radix sort workspace (sort_workspace_words)
  size_t bytes;
};

inline GeomState carve_geom(char* base, size_t P) {
  GeomState g;
  Carver<char> c(base);
  g.depth_key = c.take<uint32_t>(P);
  g.radii = c.take<int32_t>(P);
  g.rec = c.take<float4>(3 * P);
  g.clamped = c.take<uint8_t>(P);
  g.tiles_touched = c.take<uint32_t>(P);
  g.dup_rect = c.take<uint2>(P);
  g.rect_depth = c.take<uint2>(P);
  g.slot_local = c.take<uint32_t>(P);
  g.block_slots = c.take<uint32_t>(slot_blocks(P) + 1);
  g.slot_base = c.take<uint32_t>(slot_blocks(P) + 1);
  g.order = c.take<uint32_t>(P);
  g.order_alt = c.take<uint32_t>(P);
  g.key_alt = c.take<uint32_t>(P);
  g.offsets = c.take<uint32_t>(P);
  g.scan_partials = c.take<uint32_t>(scan_blocks(P) + 2);  // + total, overflow flag
  g.sort_counts = c.take<uint32_t>(sort_workspace_words(P));
  g.bytes = align_up(c.off) + kAlign;
  return g;
}

// Where the per-Gaussian gradient sums come from: the backward's records and
// the slot layout (see gather_sum).  tiles_touched == nullptr: not used.
struct GatherSrc {
  int P;
  uint32_t nslots;
  const uint32_t* R_dev;
  const uint32_t* tiles_touched;
  const uint32_t* slot_local;
  const uint32_t* slot_base;
  const float4* grad_rec;
  const float* grad_opacity;
};

// Per-instance state: tile keys and Gaussian ids, ping-pong for the tile sort;
// the backward's per-(tile, Gaussian) gradient terms.  Gaussian g's instances
// own the slots [slot_base[g / 256] + slot(g), ... + tiles_touched[g]), one per
// tile of its rect in row-major order, so that consecutive Gaussians' slots
// are adjacent; the opacity word of a slot the backward did not write holds
// kNoRecord.
constexpr int kGradTerms = 9;  // dcolor x3, dmean2D x2, dconic x3, dopacity
struct BinState {
  uint32_t* tile_key;
  uint32_t* tile_key_alt;
  uint32_t* gid;
  uint32_t* gid_alt;
  float4* grad_rec;       // [R][2]: dcolor x3, dmean2D x2, dconic x3 (32 B, one sector)
  float* grad_opacity;    // [R]: dopacity, or kNoRecord
  uint32_t* sort_counts;  // radix sort workspace (sort_workspace_words)
  uint32_t* scan_partials;
  // Per instance (in the tile lists' order): bit w set when some lane of the
  // tile's wave w blended it in the render forward.  Valid only when
  // *blend_ok == 1 (set by the render forward, cleared by the expansion, so
  // a binning recomputed without a render forward falls back to the masks).
  uint8_t* blend;
  uint32_t* blend_ok;
  size_t bytes;
};

struct RenderRec {
  float2 xy;
  float4 co;  // A, B, C, opacity
  float3 rgb;
};
__device__ __forceinline__ RenderRec load_rec(const float4* __restrict__ rec, uint32_t g) {
  const float4 a = rec[3 * g], b = rec[3 * g + 1], c = rec[3 * g + 2];
  return {make_float2(a.x, a.y), make_float4(a.z, a.w, b.x, b.y), make_float3(b.z, b.w, c.x)};
}
__device__ __forceinline__ void store_rec(float4* __restrict__ rec, uint32_t g, float2 xy, float4 co, float3 rgb,
                                          int radius, uint32_t slot) {
  rec[3 * g] = make_float4(xy.x, xy.y, co.x, co.y);
  rec[3 * g + 1] = make_float4(co.z, co.w, rgb.x, rgb.y);
  rec[3 * g + 2] = make_float4(rgb.z, __int_as_float(radius), __uint_as_float(slot), 0.f);
}

// Instance capacity of a single-view binning buffer: the instance count R
// rounded up to a multiple of 2^(bit_width(R) - 5) (at least 1024), a bucket
// of at most about 3%.  A function of R alone, so that the backward carves
// the forward's layout from the R the forward returned, while the forward can
// carve and queue its binning against the previous call's bucket before it
// reads R back.
inline uint32_t bin_capacity(uint32_t R) {
  if (R == 0) return 0;
  int bw = 0;
  for (uint32_t v = R; v; v >>= 1) ++bw;
  const int sh = bw - 5 > 10 ? bw - 5 : 10;
  const uint64_t g = 1ull << sh;
  const uint64_t c = ((uint64_t)R + g - 1) / g * g;
  return c > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)c;
}

// records = false: the forward-only layout, the same arrays up to the
// backward's gradient records (36 of the 53 bytes per instance), which it
// leaves out -- a prefix of the full layout, for a forward whose backward
// bins again (the checkpointed form).
inline BinState carve_bin(char* base, size_t R, bool records = true) {
  BinState b;
  Carver<char> c(base);
  b.tile_key = c.take<uint32_t>(R);
  b.tile_key_alt = c.take<uint32_t>(R);
  b.gid = c.take<uint32_t>(R);
  b.gid_alt = c.take<uint32_t>(R);
  b.sort_counts = c.take<uint32_t>(sort_workspace_words(R));
  b.scan_partials = c.take<uint32_t>(1);
  b.blend = c.take<uint8_t>(R);
  b.blend_ok = c.take<uint32_t>(1);
  b.grad_rec = records ? c.take<float4>(2 * (size_t)R) : nullptr;
  b.grad_opacity = records ? c.take<float>(R) : nullptr;
  b.bytes = align_up(c.off) + kAlign;
  return b;
}

// The capacity whose binning layout occupies `bytes` bytes: a forward may
// bin against a larger capacity than bin_capacity(R) (the previous call's
// bucket with headroom), so the backward recovers it from the buffer's size.
// The capacities a forward uses are bin_capacity values (multiples of at
// least 1024), on which the layout's size strictly increases.  0 when no
// such capacity at or above R has that size.
inline uint32_t bin_capacity_of_bytes(uint32_t R, size_t bytes, bool records = true) {
  if (R == 0) return 0;
  for (uint32_t C = bin_capacity(R);;) {
    const size_t b = carve_bin(nullptr, C, records).bytes;
    if (b == bytes) return C;
    if (b > bytes) return 0;
    const uint32_t next = bin_capacity(C + 1u);
    if (next <= C) return 0;
    C = next;
  }
}

// Per-pixel / per-tile state.
struct ImgState {
  float* final_T;
  uint32_t* n_contrib;
  uint2* ranges;
  size_t bytes;
};

inline ImgState carve_img(char* base, int W, int H) {
  ImgState s;
  Carver<char> c(base);
  const size_t N = (size_t)W * H;
  s.final_T = c.take<float>(N);
  s.n_contrib = c.take<uint32_t>(N);
  s.ranges = c.take<uint2>((size_t)tile_grid(W, H).tiles());
  s.bytes = align_up(c.off) + kAlign;
  return s;
}

// The checkpointed forward keeps the render's per-instance blend bytes in
// the image buffer, behind the per-pixel and per-tile state, since its
// backward has no forward binning buffer: a kAlign-byte header holding the
// blend_ok word, then one byte per instance of the binning capacity C.  The
// backward reads them through the forward's tile ranges, which stay in the
// image buffer.
inline size_t img_blend_offset(int W, int H) { return carve_img(nullptr, W, H).bytes; }
inline size_t img_blend_bytes(int W, int H, uint32_t C) { return img_blend_offset(W, H) + kAlign + align_up(C); }

// Number of 8-bit passes for a key range of `bits` bits, and whether the
// sorted result ends in the *_alt arrays (odd pass count).
inline int radix_passes(int bits) { return (bits + kRadixBits - 1) / kRadixBits; }

// bit count of n (highest set bit + 1); the reference's getHigherMsb
// (rasterizer_impl.cu:47-62) computes the same quantity.
inline int bit_width_u32(uint32_t n) {
  int b = 0;
  while (b < 32 && (n >> b)) ++b;
  return b;
}

}  // namespace omnigs
