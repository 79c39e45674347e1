// ft_kernels.h — the gfx950 kernels (compiled by hipcc into libft_hip.so, and by hiprtc
// together with scene-specialised code emitted by ft_emit.cpp).
//
//   ft_k_depth          depth preview, one thread per pixel looping over the AA grid
//                       (render_depth_tile renderer.rs:3931-3971);
//   ft_k_whitted_N      one thread per (pixel, AA sample): Whitted trace with an N-entry ray
//                       stack, writes the sample radiance;
//   ft_k_resolve        per pixel: sums the samples in AA order, scales by 1/aa and tone maps
//                       (render_ray_tile renderer.rs:3874-3929, renderer.rs:5221-5239).
#pragma once

#include "ft_shade.h"

namespace ftd {

struct CamArgs {
  float ox, oy, oz, fx, fy, fz, rx, ry, rz, ux, uy, uz, tan_fov, aspect;
  int width, height, aa, grid;
  int tile_w;  // pixel tile of one wave: tile_w x (64 / tile_w), tile_w in {8, 16, 32}; 0: sample-major
  int y0, rows;  // the launch covers image rows [y0, y0 + rows) (a band of a progressive render)
};

__device__ __forceinline__ void primary_ray(const CamArgs& c, int x, int y, int s, float3& o, float3& d) {
  // pixel_sample_offsets (renderer.rs:4838-4854): sample s sits at grid cell (s % g, s / g)
  const float g = (float)c.grid;
  const float sx = ((float)(s % c.grid) + 0.5f) / g, sy = ((float)(s / c.grid) + 0.5f) / g;
  const float px = (((float)x + sx) / (float)c.width) * 2.0f - 1.0f;
  const float py = 1.0f - (((float)y + sy) / (float)c.height) * 2.0f;
  const float ndx = px * c.aspect, ndy = py;
  float3 dir = add(add(mk(c.fx, c.fy, c.fz), mul(mk(c.rx, c.ry, c.rz), ndx * c.tan_fov)),
                   mul(mk(c.ux, c.uy, c.uz), ndy * c.tan_fov));
  d = normalize(normalize(dir));
  o = mk(c.ox, c.oy, c.oz);
}

__device__ __forceinline__ float depth_preview_value(float t, float max_dist) {  // renderer.rs:5215-5219
  float dd = clampf(t / fmaxf(max_dist, 1.0e-6f), 0.0f, 1.0f);
  float s = 1.0f - dd;
  return s * s;
}

// Pixels are visited in tiles of 64 per wavefront (8x8, 16x4 or 32x2, chosen per launch;
// blockDim 64 x waves per block) so that the lanes of a wave march adjacent rays.
// tile_w == 0 (Whitted, aa in {4, 16, 64}): a wave holds every AA sample of a square block
// of 64 / aa pixels (sample = lane % aa), so its rays start from neighbouring points of
// the same few pixels and tend to hit the same surfaces.
__device__ __forceinline__ int sample_block_w(int aa) { return aa == 4 ? 4 : (aa == 16 ? 2 : 1); }
__device__ __forceinline__ bool tile_xy_of(const CamArgs& c, int tile, int lane, int& x, int& y) {
  if (c.tile_w == 0) {
    const int bw = sample_block_w(c.aa), pi = lane / c.aa;
    const int blocks_x = (c.width + bw - 1) / bw;
    x = (tile % blocks_x) * bw + pi % bw;
    y = c.y0 + (tile / blocks_x) * bw + pi / bw;
    return x < c.width && y < c.y0 + c.rows;
  }
  const int tw = c.tile_w, th = 64 / tw;
  const int tiles_x = (c.width + tw - 1) / tw;
  const int tx = tile % tiles_x, ty = tile / tiles_x;
  x = tx * tw + lane % tw;
  y = c.y0 + ty * th + lane / tw;
  return x < c.width && y < c.y0 + c.rows;
}
__device__ __forceinline__ bool tile_xy(const CamArgs& c, int tile, int& x, int& y) {
  return tile_xy_of(c, tile, (int)threadIdx.x, x, y);
}
// 8x8 tiles (the differentiable kernels)
__device__ __forceinline__ bool tile_xy(int width, int height, int tile, int& x, int& y) {
  const int lane = threadIdx.x;
  const int tiles_x = (width + 7) / 8;
  const int tx = tile % tiles_x, ty = tile / tiles_x;
  x = tx * 8 + (lane & 7);
  y = ty * 8 + (lane >> 3);
  return x < width && y < height;
}
__device__ __forceinline__ int wave_index() { return blockIdx.x * blockDim.y + threadIdx.y; }

__device__ __forceinline__ void depth_tile(int slot, const CamArgs& cam, const TraceParams& tp, int tile,
                                           unsigned char* __restrict__ rgb, float* __restrict__ t_out) {
  int x, y;
  if (!tile_xy(cam, tile, x, y)) return;
  float sum = 0.0f;
  float first_t = -1.0f;
  for (int s = 0; s < cam.aa; ++s) {
    float3 o, d;
    primary_ray(cam, x, y, s, o, d);
    Hit hit;
    bool h = raymarch(slot, o, d, tp.max_steps, tp.eps, tp.step_scale, 0.0f, tp.max_dist, hit);
    sum = sum + (h ? depth_preview_value(hit.t, tp.max_dist) : 0.0f);
    if (s == 0) first_t = h ? hit.t : -1.0f;
  }
  const unsigned char v = to_u8(sum * (1.0f / (float)cam.aa));
  const size_t pix = (size_t)y * cam.width + x;
  rgb[pix * 3] = v;
  rgb[pix * 3 + 1] = v;
  rgb[pix * 3 + 2] = v;
  if (t_out) t_out[pix] = first_t;
}

template <int kRayStack>
__device__ __forceinline__ void whitted_tile(int slot, const CamArgs& cam, const TraceParams& tp, int aov, int tile,
                                             int s, float* __restrict__ samples, int* __restrict__ overflow,
                                             const float* __restrict__ t_cache = nullptr) {
  int x, y;
  if (!tile_xy(cam, tile, x, y)) return;
  float3 o, d;
  primary_ray(cam, x, y, s, o, d);
  const float t0 = t_cache && s == 0 && cam.aa == 1 ? t_cache[(size_t)y * cam.width + x] : -2.0f;
  float3 c = aov >= 0 ? trace_aov(slot, tp, o, d, aov) : trace_whitted<kRayStack>(slot, tp, o, d, overflow, t0);
  const size_t idx = ((size_t)s * cam.height + y) * cam.width + x;
  samples[idx * 3] = c.x;
  samples[idx * 3 + 1] = c.y;
  samples[idx * 3 + 2] = c.z;
}

__device__ __forceinline__ void depth_kernel_body(int slot, CamArgs cam, TraceParams tp, unsigned char* __restrict__ rgb,
                                                  float* __restrict__ t_out) {
  depth_tile(slot, cam, tp, wave_index(), rgb, t_out);
}

template <int kRayStack>
__device__ __forceinline__ void whitted_kernel_body(int slot, CamArgs cam, TraceParams tp, int aov,
                                                    float* __restrict__ samples, int* __restrict__ overflow) {
  const int s = cam.tile_w == 0 ? (int)threadIdx.x % cam.aa : (int)blockIdx.y;
  whitted_tile<kRayStack>(slot, cam, tp, aov, wave_index(), s, samples, overflow);
}

// Batches: many (scene, view) images in one launch.  Wave w renders tile w % tiles_per_view
// of the view ranked w / tiles_per_view in `order` (views sorted by expected cost, most
// expensive first, so the long tiles start early and the tail of the launch is short).
// Every view has the batch's image size and sample count; its camera is cams[view].
struct BatchArgs {
  const int* order;     // rank -> view
  const int* slots;     // view -> scene slot
  const CamArgs* cams;  // view -> camera
  int n_views, tiles_per_view;
};

// XCD-aware placement of the persistent depth pass.  The command processor hands
// consecutive workgroups to the eight XCDs in turn (blockIdx.x % kXcds), and each XCD has
// its own L2, from which a view's node program is re-read on every scene evaluation.
// View ranks are dealt to XCDs (rank % kXcds); a wave takes the chunks of its own XCD's
// views first.  The mapping only decides who renders what: the images do not depend on it.
constexpr int kXcds = 8;
__device__ __forceinline__ int views_on_xcd(int n_views, int x) { return (n_views - x + kXcds - 1) / kXcds; }

// Tiled launches keep the linear order: dealing views to XCDs without work stealing left
// some XCDs with the expensive views (the tiled Whitted pass of 32 scenes ran 7% slower).
__device__ __forceinline__ bool batch_item(const BatchArgs& b, int& view, int& tile, int& slot, CamArgs& cam) {
  const int w = __builtin_amdgcn_readfirstlane(wave_index());
  const int rank = w / b.tiles_per_view;
  if (rank >= b.n_views) return false;
  view = __builtin_amdgcn_readfirstlane(uniform_ptr(b.order)[rank]);
  tile = w - rank * b.tiles_per_view;
  slot = __builtin_amdgcn_readfirstlane(uniform_ptr(b.slots)[view]);
  cam = load_uniform(uniform_ptr(b.cams) + view);
  return true;
}

__device__ __forceinline__ void batch_depth_body(BatchArgs b, TraceParams tp, unsigned char* __restrict__ rgb) {
  int view, tile, slot;
  CamArgs cam;
  if (!batch_item(b, view, tile, slot, cam)) return;
  depth_tile(slot, cam, tp, tile, rgb + (size_t)view * cam.width * cam.height * 3, nullptr);
}

// The batch depth pass as a persistent kernel with per-lane refill (one AA sample per
// pixel).  A wave takes chunks of kBatchChunkTiles tiles of one view from a global counter;
// inside a chunk every lane whose ray has ended takes the chunk's next pixel at once, so the
// wave keeps 64 rays in flight until the chunk runs dry instead of waiting for its slowest
// ray per tile.  Each ray's march is MarchLane's step sequence, the one raymarch runs, so
// depths and distances are those of ft_k_batch_depth bit for bit.  `queue` starts at zero;
// every wave leaves once the counter passes the last chunk.  t_cache (optional) receives
// each pixel's hit distance (-1: miss).
// The counters are per XCD (queue[x] counts the chunks taken from XCD x's views, ranks
// x, x + kXcds, ...): a wave works through its own XCD's chunks first, then takes what is
// left on the others, so the load stays balanced while most views stay in one L2.
constexpr int kBatchChunkTiles = 4;
__device__ __forceinline__ void batch_depth_persistent_body(BatchArgs b, TraceParams tp, unsigned char* __restrict__ rgb,
                                                           float* __restrict__ t_cache, unsigned* __restrict__ queue) {
  const int chunks_per_view = (b.tiles_per_view + kBatchChunkTiles - 1) / kBatchChunkTiles;
  const int home = (int)blockIdx.x % kXcds;
  int probe = 0;  // queues found empty so far (this wave tries home, home + 1, ... in turn)
  for (;;) {
    const int x = (home + probe) % kXcds;
    const unsigned n_chunks = (unsigned)views_on_xcd(b.n_views, x) * (unsigned)chunks_per_view;
    unsigned c = 0u;
    if (threadIdx.x == 0) c = atomicAdd(queue + x, 1u);
    c = __builtin_amdgcn_readlane(c, 0);
    if (c >= n_chunks) {
      if (++probe == kXcds) return;
      continue;
    }
    const int rank = x + kXcds * (int)(c / (unsigned)chunks_per_view);
    const int first = (int)(c % (unsigned)chunks_per_view) * kBatchChunkTiles;
    const int n_items = min(kBatchChunkTiles, b.tiles_per_view - first) * 64;
    const int view = __builtin_amdgcn_readfirstlane(uniform_ptr(b.order)[rank]);
    const int slot = __builtin_amdgcn_readfirstlane(uniform_ptr(b.slots)[view]);
    const CamArgs cam = load_uniform(uniform_ptr(b.cams) + view);
    const DevScene& S = scene_of(slot);
    const SceneEval ev{slot};
    const AtlasCoarse coarse{slot, S.atlas_on != 0, S.atlas_switch};
    const size_t npix = (size_t)cam.width * cam.height;
    unsigned char* const out = rgb + (size_t)view * npix * 3;
    float* const tout = t_cache ? t_cache + (size_t)view * npix : nullptr;
    // depth_tile's arithmetic for one sample
    auto write = [&](int p, bool h, float t) {
      float sum = 0.0f;
      sum = sum + (h ? depth_preview_value(t, tp.max_dist) : 0.0f);
      const unsigned char v = to_u8(sum * (1.0f / (float)cam.aa));
      out[(size_t)p * 3] = v;
      out[(size_t)p * 3 + 1] = v;
      out[(size_t)p * 3 + 2] = v;
      if (tout) tout[p] = h ? t : -1.0f;
    };
    MarchLane m;
    int next = 0;  // items of the chunk handed out so far (wave-uniform)
    int pix = -1;  // this lane's pixel, -1: idle
    for (;;) {
      const unsigned long long idle = __builtin_amdgcn_ballot_w64(pix < 0);
      if (idle != 0ull && next < n_items) {
        if (pix < 0) {
          const int it = next + lanes_below(idle);
          int x, y;
          if (it < n_items && tile_xy_of(cam, first + it / 64, it % 64, x, y)) {
            float3 o, d;
            primary_ray(cam, x, y, 0, o, d);
            const int p = y * cam.width + x;
            if (m.init(S.bmin, S.bmax, o, d, tp.max_steps, tp.eps, tp.step_scale, 0.0f, tp.max_dist, coarse)) pix = p;
            else write(p, false, 0.0f);
          }
        }
        next = min(next + (int)__popcll(idle), n_items);
      }
      if (__builtin_amdgcn_ballot_w64(pix >= 0) == 0ull) {
        if (next >= n_items) break;
        continue;
      }
      if (pix >= 0) {
        const int r = m.step(ev, coarse);
        if (r != 0) {
          write(pix, r == 2, m.hi);
          pix = -1;
        }
      }
    }
  }
}

template <int kRayStack>
__device__ __forceinline__ void batch_whitted_body(BatchArgs b, TraceParams tp, int aov, float* __restrict__ samples,
                                                   int* __restrict__ overflow, const float* __restrict__ t_cache) {
  int view, tile, slot;
  CamArgs cam;
  if (!batch_item(b, view, tile, slot, cam)) return;
  const size_t npix = (size_t)cam.width * cam.height;
  const size_t per_view = (size_t)cam.aa * npix * 3;
  whitted_tile<kRayStack>(slot, cam, tp, aov, tile, blockIdx.y, samples + (size_t)view * per_view, overflow,
                          t_cache ? t_cache + (size_t)view * npix : nullptr);
}

// Tile order of a width x height image: bands of 8 rows top to bottom, each band's 8-column
// tiles left to right, each tile's pixels row by row (the last band and the last tile of a
// band may be narrower).  A bijection of [0, width * height) for any image size.
__device__ __forceinline__ int tile_order_pixel(int q, int width, int height) {
  const int band = q / (8 * width), r = q - band * 8 * width;
  const int hb = min(8, height - 8 * band);
  const int tx = r / (8 * hb), r2 = r - tx * 8 * hb;
  const int wt = min(8, width - 8 * tx);
  const int ly = r2 / wt, lx = r2 - ly * wt;
  return (8 * band + ly) * width + 8 * tx + lx;
}

// Hit-first pixel order for the batch Whitted pass.  After the depth pass, one 256-thread
// block per view writes the view's pixel indices, those whose primary ray hit (t_cache >= 0)
// first, then the misses, each group in tile order (a stable partition: the order is fixed
// by the image, not by timing).  Shading waves then hold 64 hits, instead of the few hits of
// a tile on a silhouette next to lanes that finished with the sky, and the 64 hits of a wave
// lie in one or two neighbouring 8x8 tiles, so their bounce and shadow rays stay coherent.
__device__ __forceinline__ void batch_partition_body(int width, int height, const float* __restrict__ t_cache,
                                                    unsigned* __restrict__ list) {
  __shared__ int cnt[256];
  const int npix = width * height;
  const int view = (int)blockIdx.x, t = (int)threadIdx.x;
  const float* tv = t_cache + (size_t)view * npix;
  unsigned* lv = list + (size_t)view * npix;
  const int seg = (npix + 255) / 256;
  const int q0 = min(t * seg, npix), q1 = min(q0 + seg, npix);
  int hits = 0;
  for (int q = q0; q < q1; ++q) hits += tv[tile_order_pixel(q, width, height)] >= 0.0f ? 1 : 0;
  cnt[t] = hits;
  __syncthreads();
  // exclusive scan of the per-thread hit counts (256 entries, read by one thread each)
  int before = 0, total = 0;
  for (int i = 0; i < 256; ++i) {
    const int c = cnt[i];
    if (i < t) before += c;
    total += c;
  }
  int h = before, m = total + (q0 - before);  // misses before q0 = q0 - hits before q0
  for (int q = q0; q < q1; ++q) {
    const int p = tile_order_pixel(q, width, height);
    if (tv[p] >= 0.0f) lv[h++] = (unsigned)p;
    else lv[m++] = (unsigned)p;
  }
}

// The batch Whitted pass over the hit-first order: wave k of a view traces list entries
// 64k .. 64k + 63, each from the depth pass's primary hit (one sample per pixel).
template <int kRayStack>
__device__ __forceinline__ void batch_whitted_list_body(BatchArgs b, TraceParams tp, float* __restrict__ samples,
                                                       int* __restrict__ overflow, const float* __restrict__ t_cache,
                                                       const unsigned* __restrict__ list) {
  const int w = __builtin_amdgcn_readfirstlane(wave_index());
  const int rank = w / b.tiles_per_view;
  if (rank >= b.n_views) return;
  const int view = __builtin_amdgcn_readfirstlane(uniform_ptr(b.order)[rank]);
  const int k = w - rank * b.tiles_per_view;
  const int slot = __builtin_amdgcn_readfirstlane(uniform_ptr(b.slots)[view]);
  const CamArgs cam = load_uniform(uniform_ptr(b.cams) + view);
  const int npix = cam.width * cam.height;
  const int e = k * 64 + (int)threadIdx.x;
  if (e >= npix) return;
  const int p = (int)list[(size_t)view * npix + e];
  const int x = p % cam.width, y = p / cam.width;
  float3 o, d;
  primary_ray(cam, x, y, 0, o, d);
  const float3 c = trace_whitted<kRayStack>(slot, tp, o, d, overflow, t_cache[(size_t)view * npix + p]);
  float* const out = samples + ((size_t)view * npix + p) * 3;
  out[0] = c.x;
  out[1] = c.y;
  out[2] = c.z;
}

// Distance atlas construction (exact scene distances sampled on a grid).
struct AtlasGrid {
  float ox, oy, oz, voxel;
  int nx, ny, nz;
  float band;      // cells whose centre distance is within this get a brick
  int max_bricks;
};

__device__ __forceinline__ long linear_thread() {
  return ((long)blockIdx.x * blockDim.y + threadIdx.y) * blockDim.x + threadIdx.x;
}

// Pass 1, one thread per cell: exact distance at the centre; near-surface cells take a
// brick slot (a vector atomic on a global counter), or -2 when the slots ran out.
__device__ __forceinline__ void atlas_cells_body(int slot, AtlasGrid g, float* __restrict__ cell_d,
                                                 int* __restrict__ cells, int* __restrict__ brick_cell,
                                                 unsigned* __restrict__ counter) {
  const long li = linear_thread();
  const int n = g.nx * g.ny * g.nz;
  const int i = li < n ? (int)li : n;
  const int ci = i < n ? i : n - 1;
  const int cx = ci % g.nx, cy = (ci / g.nx) % g.ny, cz = ci / (g.nx * g.ny);
  const float cell = 8.0f * g.voxel;
  const float3 p = mk(g.ox + ((float)cx + 0.5f) * cell, g.oy + ((float)cy + 0.5f) * cell, g.oz + ((float)cz + 0.5f) * cell);
  const float d = scene_dist(slot, p);  // every lane evaluates (wave-uniform programs)
  if (i >= n) return;
  cell_d[i] = d;
  int b = -1;
  if (fabsf(d) <= g.band) {
    const unsigned k = atomicAdd(counter, 1u);
    if ((int)k < g.max_bricks) {
      b = (int)k;
      brick_cell[k] = i;
    } else {
      b = -2;
    }
  }
  cells[i] = b;
}

// Pass 2, one thread per brick sample: exact distance at the voxel corner.
__device__ __forceinline__ void atlas_bricks_body(int slot, AtlasGrid g, const int* __restrict__ brick_cell,
                                                  int n_bricks, int fp16, void* __restrict__ bricks) {
  const long i = linear_thread();
  const long n = (long)n_bricks * kBrickSamples;
  const long si = i < n ? i : n - 1;
  const int b = (int)(si / kBrickSamples), s = (int)(si % kBrickSamples);
  const int c = brick_cell[b];
  const int cx = c % g.nx, cy = (c / g.nx) % g.ny, cz = c / (g.nx * g.ny);
  const int sx = s % kBrickSide, sy = (s / kBrickSide) % kBrickSide, sz = s / (kBrickSide * kBrickSide);
  const float cell = 8.0f * g.voxel;
  const float3 p = mk(g.ox + (float)cx * cell + (float)sx * g.voxel, g.oy + (float)cy * cell + (float)sy * g.voxel,
                      g.oz + (float)cz * cell + (float)sz * g.voxel);
  const float d = scene_dist(slot, p);
  if (i >= n) return;
  if (fp16) ((_Float16*)bricks)[i] = (_Float16)d;
  else ((float*)bricks)[i] = d;
}

// One thread per pixel; blockIdx.y selects the image of a batch (samples and outputs of
// image k start at k times the per-image size).
__device__ __forceinline__ void resolve_kernel_body(CamArgs cam, int aov, const float* __restrict__ samples,
                                                 unsigned char* __restrict__ rgb, float* __restrict__ hdr) {
  const int npix = cam.width * cam.height;
  const int pix = cam.y0 * cam.width + blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= (cam.y0 + cam.rows) * cam.width) return;
  samples += (size_t)blockIdx.y * cam.aa * npix * 3;
  rgb += (size_t)blockIdx.y * npix * 3;
  if (hdr) hdr += (size_t)blockIdx.y * npix * 3;
  float3 sum = mk(0, 0, 0);
  for (int s = 0; s < cam.aa; ++s) {
    const float* c = samples + ((size_t)s * npix + pix) * 3;
    sum = add(sum, mk(c[0], c[1], c[2]));
  }
  const float3 avg = mul(sum, 1.0f / (float)cam.aa);
  unsigned char* o = rgb + (size_t)pix * 3;
  if (aov >= 0) { o[0] = to_u8(avg.x); o[1] = to_u8(avg.y); o[2] = to_u8(avg.z); }
  else { o[0] = to_u8_reinhard(avg.x); o[1] = to_u8_reinhard(avg.y); o[2] = to_u8_reinhard(avg.z); }
  if (hdr) { hdr[(size_t)pix * 3] = avg.x; hdr[(size_t)pix * 3 + 1] = avg.y; hdr[(size_t)pix * 3 + 2] = avg.z; }
}

// Clears the overflow flag with a kernel (a memset node would misbehave under graph capture).
__device__ __forceinline__ void clear_flag_body(int* flag) {
  if (threadIdx.x == 0) flag[0] = 0;
}

__device__ __forceinline__ void clear_u64_body(unsigned long long* v, int n) {
  if ((int)threadIdx.x < n) v[threadIdx.x] = 0ull;
}

}  // namespace ftd

// FT_WAVES_PER_EU: the occupancy the march kernels are compiled for (their VGPR budget,
// 512 / waves); the launch-shape autotuner builds scene-specialised variants of it.
#ifndef FT_WAVES_PER_EU
#define FT_WAVES_PER_EU 4
#endif
#define FT_KERNEL extern "C" __global__ __launch_bounds__(256)
#define FT_MARCH_OCCUPANCY __attribute__((amdgpu_waves_per_eu(FT_WAVES_PER_EU)))

// FT_RTC_KERNEL selects the one kernel a runtime-compiled (scene-specialised) code object
// holds: 0 depth, 1/2/3 Whitted with a 16/48/128-entry ray stack.  The precompiled library
// holds them all.  FT_COUNTED names the kernel *_counted: the FLOP-counting launches of
// instrumented runs then show up apart from the timed ones in a kernel trace.
#if defined(FT_COUNTED)
#define FT_KNAME(n) n##_counted
#else
#define FT_KNAME(n) n
#endif
#if !defined(FT_RTC_KERNEL) || FT_RTC_KERNEL == 0
FT_KERNEL FT_MARCH_OCCUPANCY void FT_KNAME(ft_k_depth)(int slot, ftd::CamArgs cam, ftd::TraceParams tp, unsigned char* __restrict__ rgb,
                                             float* __restrict__ t_out) {
  ftd::depth_kernel_body(slot, cam, tp, rgb, t_out);
}
#endif
#if !defined(FT_RTC_KERNEL) || FT_RTC_KERNEL == 1
FT_KERNEL FT_MARCH_OCCUPANCY void FT_KNAME(ft_k_whitted_16)(int slot, ftd::CamArgs cam, ftd::TraceParams tp, int aov,
                                                  float* __restrict__ samples, int* __restrict__ overflow) {
  ftd::whitted_kernel_body<16>(slot, cam, tp, aov, samples, overflow);
}
#endif
#if !defined(FT_RTC_KERNEL) || FT_RTC_KERNEL == 2
FT_KERNEL FT_MARCH_OCCUPANCY void FT_KNAME(ft_k_whitted_48)(int slot, ftd::CamArgs cam, ftd::TraceParams tp, int aov,
                                                  float* __restrict__ samples, int* __restrict__ overflow) {
  ftd::whitted_kernel_body<48>(slot, cam, tp, aov, samples, overflow);
}
#endif
#if !defined(FT_RTC_KERNEL) || FT_RTC_KERNEL == 3
FT_KERNEL FT_MARCH_OCCUPANCY void FT_KNAME(ft_k_whitted_128)(int slot, ftd::CamArgs cam, ftd::TraceParams tp, int aov,
                                                   float* __restrict__ samples, int* __restrict__ overflow) {
  ftd::whitted_kernel_body<128>(slot, cam, tp, aov, samples, overflow);
}
#endif
#if !defined(FT_RTC_KERNEL)
FT_KERNEL FT_MARCH_OCCUPANCY void ft_k_batch_depth(ftd::BatchArgs b, ftd::TraceParams tp, unsigned char* __restrict__ rgb) {
  ftd::batch_depth_body(b, tp, rgb);
}
FT_KERNEL FT_MARCH_OCCUPANCY void ft_k_batch_depth_persistent(ftd::BatchArgs b, ftd::TraceParams tp,
                                                              unsigned char* __restrict__ rgb,
                                                              float* __restrict__ t_cache, unsigned* __restrict__ queue) {
  ftd::batch_depth_persistent_body(b, tp, rgb, t_cache, queue);
}
FT_KERNEL void ft_k_batch_partition(int width, int height, const float* __restrict__ t_cache,
                                    unsigned* __restrict__ list) {
  ftd::batch_partition_body(width, height, t_cache, list);
}
FT_KERNEL FT_MARCH_OCCUPANCY void ft_k_batch_whitted_list_16(ftd::BatchArgs b, ftd::TraceParams tp,
                                                             float* __restrict__ samples, int* __restrict__ overflow,
                                                             const float* __restrict__ t_cache,
                                                             const unsigned* __restrict__ list) {
  ftd::batch_whitted_list_body<16>(b, tp, samples, overflow, t_cache, list);
}
FT_KERNEL FT_MARCH_OCCUPANCY void ft_k_batch_whitted_list_48(ftd::BatchArgs b, ftd::TraceParams tp,
                                                             float* __restrict__ samples, int* __restrict__ overflow,
                                                             const float* __restrict__ t_cache,
                                                             const unsigned* __restrict__ list) {
  ftd::batch_whitted_list_body<48>(b, tp, samples, overflow, t_cache, list);
}
FT_KERNEL FT_MARCH_OCCUPANCY void ft_k_batch_whitted_list_128(ftd::BatchArgs b, ftd::TraceParams tp,
                                                              float* __restrict__ samples, int* __restrict__ overflow,
                                                              const float* __restrict__ t_cache,
                                                              const unsigned* __restrict__ list) {
  ftd::batch_whitted_list_body<128>(b, tp, samples, overflow, t_cache, list);
}
FT_KERNEL FT_MARCH_OCCUPANCY void ft_k_batch_whitted_16(ftd::BatchArgs b, ftd::TraceParams tp, int aov,
                                                        float* __restrict__ samples, int* __restrict__ overflow,
                                                        const float* __restrict__ t_cache) {
  ftd::batch_whitted_body<16>(b, tp, aov, samples, overflow, t_cache);
}
FT_KERNEL FT_MARCH_OCCUPANCY void ft_k_batch_whitted_48(ftd::BatchArgs b, ftd::TraceParams tp, int aov,
                                                        float* __restrict__ samples, int* __restrict__ overflow,
                                                        const float* __restrict__ t_cache) {
  ftd::batch_whitted_body<48>(b, tp, aov, samples, overflow, t_cache);
}
FT_KERNEL FT_MARCH_OCCUPANCY void ft_k_batch_whitted_128(ftd::BatchArgs b, ftd::TraceParams tp, int aov,
                                                        float* __restrict__ samples, int* __restrict__ overflow,
                                                        const float* __restrict__ t_cache) {
  ftd::batch_whitted_body<128>(b, tp, aov, samples, overflow, t_cache);
}
FT_KERNEL void ft_k_resolve(ftd::CamArgs cam, int aov, const float* __restrict__ samples, unsigned char* __restrict__ rgb,
                            float* __restrict__ hdr) {
  ftd::resolve_kernel_body(cam, aov, samples, rgb, hdr);
}
FT_KERNEL FT_MARCH_OCCUPANCY void ft_k_atlas_cells(int slot, ftd::AtlasGrid g, float* __restrict__ cell_d, int* __restrict__ cells,
                                int* __restrict__ brick_cell, unsigned* __restrict__ counter) {
  ftd::atlas_cells_body(slot, g, cell_d, cells, brick_cell, counter);
}
FT_KERNEL FT_MARCH_OCCUPANCY void ft_k_atlas_bricks(int slot, ftd::AtlasGrid g, const int* __restrict__ brick_cell, int n_bricks, int fp16,
                                 void* __restrict__ bricks) {
  ftd::atlas_bricks_body(slot, g, brick_cell, n_bricks, fp16, bricks);
}
FT_KERNEL void ft_k_clear_flag(int* flag) { ftd::clear_flag_body(flag); }
FT_KERNEL void ft_k_clear_u64(unsigned long long* v, int n) { ftd::clear_u64_body(v, n); }
#endif
