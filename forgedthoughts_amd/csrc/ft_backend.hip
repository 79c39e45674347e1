// ft_backend.hip — the C ABI of libft_hip.so (include/ft_abi.h) and kernel dispatch.
//
// The kernels themselves are in ft_kernels.h.  Each march kernel exists twice:
//   * precompiled into this library, evaluating the scene through the node-program
//     interpreter (any scene, no start-up cost);
//   * runtime-compiled per scene with hiprtc from the source ft_emit.cpp writes, where the
//     scene distance and the Forge functions are straight-line code with every parameter
//     a literal — the role Cranelift plays for the reference (jit.rs).
// ft_scene_create decides which (FT_EMIT: "auto" default, "1" always, "0" never); the
// specialised code object for a kernel is built on its first launch and cached.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "ft_diff.h"
#include "ft_emit.h"
#include "ft_kernels.h"
#include "ft_rtc.h"
#include "ft_scene_build.h"

using namespace ftd;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define FT_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(e_ == hipErrorOutOfMemory ? FT_ERR_OUT_OF_MEMORY : FT_ERR_LAUNCH_FAILED,  \
                  std::string(#call) + ": " + hipGetErrorString(e_));                       \
  } while (0)

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

template <class T>
int upload(DevBuf& b, const std::vector<T>& v) {
  size_t bytes = v.size() * sizeof(T);
  if (bytes == 0) return FT_OK;
  FT_HIP(hipMalloc(&b.ptr, bytes));
  b.bytes = bytes;
  FT_HIP(hipMemcpy(b.ptr, v.data(), bytes, hipMemcpyHostToDevice));
  return FT_OK;
}

int ensure(DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes) return FT_OK;
  if (b.ptr) FT_HIP(hipFree(b.ptr));
  b.ptr = nullptr;
  b.bytes = 0;
  FT_HIP(hipMalloc(&b.ptr, bytes));
  b.bytes = bytes;
  return FT_OK;
}

}  // namespace

// Launch shape of a march kernel (ft_launch_config): block_waves x 64 threads per block,
// tile_w x (64 / tile_w) pixels per wave (Whitted tile_w 0: every AA sample of a square
// pixel block per wave), the occupancy the code is compiled for and the march steps per
// loop iteration.  The default code is the precompiled kernels'.
struct LaunchCfg {
  uint32_t block_waves = 4, tile_w = 8, wpe = 4, unroll = 1, shade_inline = 0;
  bool operator==(const LaunchCfg& o) const {
    return block_waves == o.block_waves && tile_w == o.tile_w && wpe == o.wpe && unroll == o.unroll &&
           shade_inline == o.shade_inline;
  }
  bool default_code() const { return wpe == 4 && unroll == 1 && shade_inline == 0; }
  bool valid(uint32_t kind) const {
    return (block_waves == 1 || block_waves == 2 || block_waves == 4) &&
           (tile_w == 8 || tile_w == 16 || tile_w == 32 || (tile_w == 0 && kind == FT_KIND_WHITTED)) &&
           wpe >= 1 && wpe <= 8 && (unroll == 1 || unroll == 2 || unroll == 4) && shade_inline <= 1;
  }
};

// Default shapes: one-wave blocks, since a block holds its CU slot and LDS until its last
// wave ends and the waves of a frame differ widely in cost (depth previews 4-21% faster
// than with four-wave blocks).  Whitted: sample-major waves; all samples of a pixel start
// at the same pixel, so the wave's rays take near-identical paths through the scene and its
// BVH (1080p at 16 spp: 59 ms against 103 ms for four-wave blocks of 8x8 pixel tiles, the
// same image byte for byte).
LaunchCfg default_cfg(uint32_t kind) {
  LaunchCfg c;
  c.block_waves = 1;
  if (kind == FT_KIND_WHITTED) c.tile_w = 0;
  return c;
}

struct ft_scene {
  fth::HostScene host;
  DevScene dev;
  int slot = -1;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevBuf nodes, prog, ops, lbprog, mats, hooks, objects, lights, fns, exprs, bvh, leaves;
  DevBuf rgb, tbuf, samples, hdr, overflow, flops;
  DevBuf atlas_cell_d, atlas_cells, atlas_brick_cell, atlas_bricks, atlas_counter;
  ft_render_stats stats{};
  bool timed = false;
  int* last_overflow = nullptr;  // overflow flag of the last Whitted render, read by finish()
  // scene specialisation
  fth::EmitPlan plan;
  std::string source;                              // emitted translation unit (empty: interpreter)
  // runtime-compiled code objects, one per (kernel, counted, occupancy, unroll), built on
  // first use: the scene-specialised kernels, the *_counted variants (FLOP counting on) and
  // launch-shape variants of the interpreter kernels
  std::map<uint32_t, std::pair<hipModule_t, hipFunction_t>> mods;
  LaunchCfg cfg[2] = {default_cfg(FT_KIND_DEPTH), default_cfg(FT_KIND_WHITTED)};  // depth [0], Whitted [1]
  uint64_t ir_hash = 0;
  double compile_ms = 0.0;
};

namespace {

std::mutex g_slot_mu;
bool g_slot_used[kMaxScenes] = {};

int acquire_slot() {
  std::lock_guard<std::mutex> lk(g_slot_mu);
  for (int k = 0; k < kMaxScenes; ++k)
    if (!g_slot_used[k]) { g_slot_used[k] = true; return k; }
  return -1;
}
void release_slot(int k) {
  if (k < 0) return;
  std::lock_guard<std::mutex> lk(g_slot_mu);
  g_slot_used[k] = false;
}

// Publishes the scene descriptor to its constant-memory slot (call with no render in flight).
int publish_module(ft_scene* s, hipModule_t m) {
  hipDeviceptr_t table = nullptr;
  size_t bytes = 0;
  FT_HIP(hipModuleGetGlobal(&table, &bytes, m, "ft_c_scenes"));
  if (bytes < (size_t)(s->slot + 1) * sizeof(DevScene)) return fail(FT_ERR_LAUNCH_FAILED, "scene table size mismatch");
  FT_HIP(hipMemcpyHtoD((hipDeviceptr_t)((char*)table + (size_t)s->slot * sizeof(DevScene)), &s->dev, sizeof(DevScene)));
  return FT_OK;
}

int publish(ft_scene* s) {
  FT_HIP(hipMemcpyToSymbol(HIP_SYMBOL(ft_c_scenes), &s->dev, sizeof(DevScene), (size_t)s->slot * sizeof(DevScene),
                           hipMemcpyHostToDevice));
  for (auto& kv : s->mods) {
    int rc = publish_module(s, kv.second.first);
    if (rc) return rc;
  }
  return FT_OK;
}

const char* const kKernelNames[fth::kRtcKernelCount] = {"ft_k_depth", "ft_k_whitted_16", "ft_k_whitted_48",
                                                        "ft_k_whitted_128"};
const void* const kKernelStubs[fth::kRtcKernelCount] = {(const void*)ft_k_depth, (const void*)ft_k_whitted_16,
                                                        (const void*)ft_k_whitted_48, (const void*)ft_k_whitted_128};

const LaunchCfg& cfg_of(const ft_scene* s, int k) { return s->cfg[k == fth::kRtcDepth ? 0 : 1]; }
uint32_t mod_key(int k, bool counted, const LaunchCfg& c) {
  return (uint32_t)k | (counted ? 1u << 4 : 0u) | (c.wpe << 8) | (c.unroll << 12) | (c.shade_inline << 16);
}
// Kernel k of shape c runs from a runtime-compiled code object (else the precompiled one).
bool needs_module(const ft_scene* s, bool counted, const LaunchCfg& c) {
  return counted || !s->source.empty() || !c.default_code() || fth::strict_math();
}
hipFunction_t module_function(const ft_scene* s, int k, bool counted, const LaunchCfg& c) {
  auto it = s->mods.find(mod_key(k, counted, c));
  return it == s->mods.end() ? nullptr : it->second.second;
}

// Loads a compiled code object for kernel k and publishes the scene table into it.
int load_module(ft_scene* s, int k, bool counted, const LaunchCfg& c, const std::vector<char>& code) {
  FT_HIP(hipStreamSynchronize(s->stream));
  hipModule_t m = nullptr;
  FT_HIP(hipModuleLoadData(&m, code.data()));
  const std::string name = std::string(kKernelNames[k]) + (counted ? "_counted" : "");
  hipFunction_t f = nullptr;
  FT_HIP(hipModuleGetFunction(&f, m, name.c_str()));
  s->mods[mod_key(k, counted, c)] = {m, f};
  return publish_module(s, m);
}

std::string module_source(const ft_scene* s) {
  return s->source.empty() ? std::string("#include \"ft_kernels.h\"\n") : s->source;
}

// Builds and loads the code object kernel `k` needs under shape `c` (no-op when it runs
// precompiled or is loaded).  With FLOP counting on, the *_counted variant is used; it is
// built for every scene (from the interpreter alone when the scene is not specialised).
int prepare(ft_scene* s, int k, const LaunchCfg& c) {
  const bool counted = s->dev.flops != nullptr;
  if (!needs_module(s, counted, c) || module_function(s, k, counted, c)) return FT_OK;
  std::vector<char> code;
  std::string err;
  double ms = 0.0;
  if (!fth::rtc_compile(module_source(s), k, counted, (int)c.wpe, (int)c.unroll, (int)c.shade_inline, code, ms, err))
    return fail(FT_ERR_COMPILE_FAILED, err);
  s->compile_ms += ms;
  return load_module(s, k, counted, c, code);
}
int prepare(ft_scene* s, int k) { return prepare(s, k, cfg_of(s, k)); }

// Launches march kernel `k` in the scene's launch shape for it: the FLOP-counting variant
// when counting is on, else the runtime-compiled object when one is needed, else the
// precompiled kernel.  `args` points at each argument.
int launch(ft_scene* s, int k, dim3 grid, size_t lds, void** args) {
  const LaunchCfg& c = cfg_of(s, k);
  int rc = prepare(s, k, c);
  if (rc) return rc;
  const bool counted = s->dev.flops != nullptr;
  const dim3 block(64, c.block_waves);
  if (needs_module(s, counted, c)) {
    hipFunction_t f = module_function(s, k, counted, c);
    FT_HIP(hipModuleLaunchKernel(f, grid.x, grid.y, grid.z, block.x, block.y, block.z, (unsigned)lds, s->stream, args,
                                 nullptr));
    return FT_OK;
  }
  FT_HIP(hipLaunchKernel(kKernelStubs[k], grid, block, args, lds, s->stream));
  return FT_OK;
}

int check_options(const ft_render_options* o) {
  if (!o) return fail(FT_ERR_INVALID_ARGUMENT, "render options are null");
  if (o->width == 0 || o->height == 0) return fail(FT_ERR_INVALID_ARGUMENT, "image size must be non-zero");
  if ((uint64_t)o->width * o->height > (1ull << 30)) return fail(FT_ERR_INVALID_ARGUMENT, "image too large");
  if (!(o->epsilon > 0.0f) || !(o->max_dist > 0.0f)) return fail(FT_ERR_INVALID_ARGUMENT, "epsilon and max_dist must be positive");
  return FT_OK;
}

CamArgs make_cam(const ft_scene* s, const ft_render_options& o, uint32_t aa, uint32_t tile_w = 8) {
  fth::CamBasis cb = fth::camera_basis(s->host, o);
  CamArgs c;
  c.ox = cb.origin[0]; c.oy = cb.origin[1]; c.oz = cb.origin[2];
  c.fx = cb.forward[0]; c.fy = cb.forward[1]; c.fz = cb.forward[2];
  c.rx = cb.right[0]; c.ry = cb.right[1]; c.rz = cb.right[2];
  c.ux = cb.up[0]; c.uy = cb.up[1]; c.uz = cb.up[2];
  c.tan_fov = cb.tan_fov;
  c.aspect = (float)o.width / (float)o.height;
  c.width = (int)o.width;
  c.height = (int)o.height;
  c.aa = (int)(aa < 1 ? 1 : aa);
  c.grid = 1;
  while (c.grid * c.grid < c.aa) ++c.grid;  // ceil(sqrt(aa))
  c.tile_w = (int)tile_w;
  c.y0 = 0;
  c.rows = c.height;
  return c;
}

TraceParams make_tp(const ft_render_options& o, uint32_t max_depth) {
  TraceParams tp;
  tp.max_steps = (int)o.max_steps;
  tp.max_depth = (int)(max_depth < 1 ? 1 : max_depth);
  tp.eps = o.epsilon;
  tp.step_scale = o.step_scale;
  tp.max_dist = o.max_dist;
  return tp;
}

// dynamic LDS of a block of `waves` wavefronts: three value-stack planes
size_t stack_lds(const ft_scene* s, uint32_t waves = 4) {
  return (size_t)3 * s->dev.stack_depth * 64 * waves * sizeof(float);
}

// one wave per pixel tile of the camera's tile shape over its rows, `waves` waves per block
// tile_w == 0: Whitted waves of all AA samples of a square pixel block (sample_block_w)
int sample_block_w(int aa) { return aa == 4 ? 4 : (aa == 16 ? 2 : 1); }
bool sample_major_ok(int aa) { return aa == 4 || aa == 16 || aa == 64; }

dim3 tile_grid(const CamArgs& c, uint32_t waves = 4) {
  int tiles;
  if (c.tile_w == 0) {
    const int bw = sample_block_w(c.aa);
    tiles = ((c.width + bw - 1) / bw) * ((c.rows + bw - 1) / bw);
  } else {
    const int th = 64 / c.tile_w;
    tiles = ((c.width + c.tile_w - 1) / c.tile_w) * ((c.rows + th - 1) / th);
  }
  return dim3((unsigned)((tiles + (int)waves - 1) / (int)waves), 1, 1);
}


int clear_flops(ft_scene* s) {
  s->stats.sdf_flops = 0;
  if (!s->dev.flops) return FT_OK;
  hipLaunchKernelGGL(ft_k_clear_u64, dim3(1), dim3(64), 0, s->stream, s->dev.flops, 2);
  FT_HIP(hipGetLastError());
  return FT_OK;
}

// Enqueues the depth preview of the camera's rows (the kernel must be prepared).
int depth_enqueue(ft_scene* s, const CamArgs& cam_in, const TraceParams& tp_in, unsigned char* rgb_dev, float* t_dev) {
  CamArgs cam = cam_in;
  TraceParams tp = tp_in;
  const LaunchCfg& lc = s->cfg[0];
  int slot = s->slot, rc;
  void* args[] = {&slot, &cam, &tp, &rgb_dev, &t_dev};
  FT_HIP(hipEventRecord(s->ev0, s->stream));
  if ((rc = launch(s, fth::kRtcDepth, tile_grid(cam, lc.block_waves), stack_lds(s, lc.block_waves), args))) return rc;
  FT_HIP(hipEventRecord(s->ev1, s->stream));
  s->timed = true;
  s->stats.primary_rays = (uint64_t)cam.width * cam.rows * cam.aa;
  s->stats.kernel_launches = 1;
  return FT_OK;
}

int depth_impl(ft_scene* s, const ft_render_options* o, uint32_t aa, unsigned char* rgb_dev, float* t_dev) {
  int rc = check_options(o);
  if (rc) return rc;
  if ((rc = clear_flops(s))) return rc;
  CamArgs cam = make_cam(s, *o, aa, s->cfg[0].tile_w);
  TraceParams tp = make_tp(*o, 1);
  FT_HIP(hipSetDevice(s->device));
  if ((rc = prepare(s, fth::kRtcDepth))) return rc;  // compile outside the timed region
  return depth_enqueue(s, cam, tp, rgb_dev, t_dev);
}

// Checks a Whitted request and picks its kernel: the ray-stack capacity is exact (every hit
// pops one entry and pushes at most max_children, none at max_depth).
int whitted_check(ft_scene* s, const ft_render_options* o, const ft_ray_settings* rs, CamArgs& cam, TraceParams& tp,
                  int& kernel) {
  int rc = check_options(o);
  if (rc) return rc;
  if (!rs) return fail(FT_ERR_INVALID_ARGUMENT, "ray settings are null");
  if (rs->debug_aov < -1 || rs->debug_aov > FT_AOV_HIT_T) return fail(FT_ERR_INVALID_ARGUMENT, "unknown debug AOV");
  if (s->host.dynamic_material >= 0)
    return fail(FT_ERR_UNSUPPORTED_OBJECT, "material " + std::to_string(s->host.dynamic_material) +
                                               " has per-hit hook functions, which this backend does not evaluate");
  // a sample-major shape needs aa in {4, 16, 64}; other sample counts use 8x8 pixel tiles
  const uint32_t tw = s->cfg[1].tile_w == 0 && !sample_major_ok((int)rs->aa_samples) ? 8u : s->cfg[1].tile_w;
  cam = make_cam(s, *o, rs->aa_samples, tw);
  if (cam.aa > 65535) return fail(FT_ERR_INVALID_ARGUMENT, "too many AA samples");
  tp = make_tp(*o, rs->max_depth);
  const long need = 1 + (long)(tp.max_depth - 1) * (long)(s->host.max_children - 1);
  if (need <= 16) kernel = fth::kRtcWhitted16;
  else if (need <= 48) kernel = fth::kRtcWhitted48;
  else if (need <= 128) kernel = fth::kRtcWhitted128;
  else return fail(FT_ERR_INVALID_ARGUMENT, "max_depth too large for the ray stack with this scene's materials");
  return FT_OK;
}

// Enqueues a checked Whitted render with the per-sample radiance in `samples`
// (aa * W * H * 3 floats) and the overflow flag at `ovf`; the kernel must be prepared.
int whitted_enqueue(ft_scene* s, const CamArgs& cam_in, const TraceParams& tp_in, int kernel, int aov,
                    unsigned char* rgb_dev, float* hdr_dev, float* samples, int* ovf) {
  CamArgs cam = cam_in;
  TraceParams tp = tp_in;
  int rc;
  FT_HIP(hipSetDevice(s->device));
  hipLaunchKernelGGL(ft_k_clear_flag, dim3(1), dim3(64), 0, s->stream, ovf);
  FT_HIP(hipGetLastError());
  const LaunchCfg& lc = s->cfg[1];
  dim3 grid = tile_grid(cam, lc.block_waves);
  grid.y = cam.tile_w == 0 ? 1u : (unsigned)cam.aa;  // sample-major waves hold every sample
  int slot = s->slot;
  void* args[] = {&slot, &cam, &tp, &aov, &samples, &ovf};
  FT_HIP(hipEventRecord(s->ev0, s->stream));
  if ((rc = launch(s, kernel, grid, stack_lds(s, lc.block_waves), args))) return rc;
  const int npix = cam.width * cam.rows;  // pixels of the launch's rows
  hipLaunchKernelGGL(ft_k_resolve, dim3((npix + 255) / 256), dim3(256), 0, s->stream, cam, aov, samples, rgb_dev,
                     hdr_dev);
  FT_HIP(hipGetLastError());
  FT_HIP(hipEventRecord(s->ev1, s->stream));
  s->last_overflow = ovf;
  s->timed = true;
  s->stats.primary_rays = (uint64_t)npix * cam.aa;
  s->stats.kernel_launches = 2;
  s->stats.ray_stack = kernel == fth::kRtcWhitted16 ? 16u : (kernel == fth::kRtcWhitted48 ? 48u : 128u);
  return FT_OK;
}

int whitted_impl(ft_scene* s, const ft_render_options* o, const ft_ray_settings* rs, unsigned char* rgb_dev,
                 float* hdr_dev) {
  CamArgs cam;
  TraceParams tp;
  int kernel;
  int rc = whitted_check(s, o, rs, cam, tp, kernel);
  if (rc) return rc;
  if ((rc = clear_flops(s))) return rc;
  rc = ensure(s->samples, (size_t)cam.aa * cam.width * cam.height * 3 * sizeof(float));
  if (rc) return rc;
  rc = ensure(s->overflow, sizeof(int));
  if (rc) return rc;
  FT_HIP(hipSetDevice(s->device));
  if ((rc = prepare(s, kernel))) return rc;  // compile outside the timed region
  return whitted_enqueue(s, cam, tp, kernel, rs->debug_aov, rgb_dev, hdr_dev, (float*)s->samples.ptr,
                         (int*)s->overflow.ptr);
}

int finish(ft_scene* s) {
  FT_HIP(hipStreamSynchronize(s->stream));
  if (s->dev.flops) {
    unsigned long long f[2] = {0, 0};
    FT_HIP(hipMemcpy(f, s->dev.flops, sizeof(f), hipMemcpyDeviceToHost));
    s->stats.sdf_flops = f[0];
    s->stats.atlas_samples = f[1];
  }
  if (s->timed) {
    float ms = 0.0f;
    FT_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->stats.last_kernel_ms = ms;
    s->timed = false;
  }
  if (s->last_overflow) {
    int ovf = 0;
    FT_HIP(hipMemcpy(&ovf, s->last_overflow, sizeof(int), hipMemcpyDeviceToHost));
    s->last_overflow = nullptr;
    if (ovf) return fail(FT_ERR_LAUNCH_FAILED, "ray stack overflow");
  }
  return FT_OK;
}

}  // namespace

// Device memory of ft_render_whitted_device_noalloc: the overflow flag in the first 256
// bytes, the per-sample radiance after it.  One copy per GPU, held until the process ends.
constexpr size_t kNoallocScratchBytes = (size_t)64 << 20;
constexpr size_t kNoallocSamplesOffset = 256;
static_assert((kNoallocScratchBytes - kNoallocSamplesOffset) / (3 * sizeof(float)) == FT_NOALLOC_MAX_SAMPLES,
              "FT_NOALLOC_MAX_SAMPLES must match the scratch array");
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (the array's name on the host is only a
// stand-in), looked up once per device with that device current.
int noalloc_scratch(int device, unsigned char** out) {
  if (device < 0 || device >= kMaxDevices) return fail(FT_ERR_INVALID_ARGUMENT, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    FT_HIP(hipSetDevice(device));
    void* p = nullptr;
    FT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = p;
  }
  *out = (unsigned char*)g_scratch_ptr[device];
  return FT_OK;
}
}  // namespace

namespace {

uint64_t fnv_bytes(uint64_t h, const void* p, size_t n) {
  const unsigned char* c = (const unsigned char*)p;
  for (size_t i = 0; i < n; ++i) {
    h ^= c[i];
    h *= 1099511628211ull;
  }
  return h;
}

// A hash of everything in the IR that decides the image (the tuning cache key's scene part).
uint64_t ir_hash(const ft_scene_ir* ir, uint32_t accel) {
  uint64_t h = fnv_bytes(1469598103934665603ull, &accel, sizeof(accel));
  h = fnv_bytes(h, &ir->root, sizeof(ir->root));
  h = fnv_bytes(h, ir->nodes, sizeof(ft_node) * ir->n_nodes);
  h = fnv_bytes(h, ir->materials, sizeof(ft_material) * ir->n_materials);
  if (ir->object_transforms) h = fnv_bytes(h, ir->object_transforms, sizeof(float) * 6 * ir->n_objects);
  if (ir->lights) h = fnv_bytes(h, ir->lights, sizeof(ft_light) * ir->n_lights);
  if (ir->functions) h = fnv_bytes(h, ir->functions, sizeof(ft_function) * ir->n_functions);
  if (ir->exprs) h = fnv_bytes(h, ir->exprs, sizeof(ft_expr) * ir->n_exprs);
  if (ir->material_hooks) h = fnv_bytes(h, ir->material_hooks, sizeof(ft_material_hooks) * ir->n_material_hooks);
  h = fnv_bytes(h, &ir->env_fn, sizeof(ir->env_fn));
  h = fnv_bytes(h, &ir->camera, sizeof(ir->camera));
  return h;
}

}  // namespace

// ------------------------------------------------------------------ progressive renders
namespace {

// render_preview_tiled / render_ray_with_accel_progressive (renderer.rs:3635-3724,
// 3751-3834): the image in rows of tiles of tile_size pixels; after each row of tiles the
// rows are copied back and the callback runs on the calling thread once per finished tile,
// in the reference's tile order (tile_jobs renderer.rs:3849-3862), with the image so far
// (unfinished tiles black).  A non-zero return stops the render.
int progressive_impl(ft_scene* s, bool whitted, const ft_render_options* o, const ft_ray_settings* rs, uint32_t aa,
                     uint32_t tile_size, ft_progress_fn cb, void* user, uint8_t* rgb_out) {
  const auto t0 = std::chrono::steady_clock::now();
  int rc = check_options(o);
  if (rc) return rc;
  if ((rc = clear_flops(s))) return rc;
  CamArgs cam;
  TraceParams tp;
  int kernel = fth::kRtcDepth;
  if (whitted) {
    if ((rc = whitted_check(s, o, rs, cam, tp, kernel))) return rc;
    if ((rc = ensure(s->samples, (size_t)cam.aa * cam.width * cam.height * 3 * sizeof(float)))) return rc;
    if ((rc = ensure(s->overflow, sizeof(int)))) return rc;
  } else {
    cam = make_cam(s, *o, aa, s->cfg[0].tile_w);
    tp = make_tp(*o, 1);
  }
  const size_t W = o->width, H = o->height;
  if ((rc = ensure(s->rgb, W * H * 3))) return rc;
  FT_HIP(hipSetDevice(s->device));
  if ((rc = prepare(s, kernel))) return rc;
  const uint32_t ts = std::max<uint32_t>(tile_size, 1u);
  const uint32_t tiles_x = (uint32_t)((W + ts - 1) / ts), tiles_y = (uint32_t)((H + ts - 1) / ts);
  std::vector<uint8_t> image(W * H * 3, 0);
  ft_progress pg{};
  pg.tiles_total = tiles_x * tiles_y;
  unsigned char* rgb_dev = (unsigned char*)s->rgb.ptr;
  for (uint32_t by = 0; by < tiles_y; ++by) {
    cam.y0 = (int)(by * ts);
    cam.rows = (int)std::min<size_t>(ts, H - cam.y0);
    rc = whitted ? whitted_enqueue(s, cam, tp, kernel, rs->debug_aov, rgb_dev, nullptr, (float*)s->samples.ptr,
                                   (int*)s->overflow.ptr)
                 : depth_enqueue(s, cam, tp, rgb_dev, nullptr);
    if (rc) return rc;
    if ((rc = finish(s))) return rc;
    const size_t off = (size_t)cam.y0 * W * 3, n = (size_t)cam.rows * W * 3;
    FT_HIP(hipMemcpy(image.data() + off, rgb_dev + off, n, hipMemcpyDeviceToHost));
    if (!cb) continue;
    for (uint32_t bx = 0; bx < tiles_x; ++bx) {
      ++pg.tiles_done;
      pg.elapsed_ms = (uint64_t)std::chrono::duration_cast<std::chrono::milliseconds>(
                          std::chrono::steady_clock::now() - t0).count();
      pg.tile_x = bx * ts;
      pg.tile_y = (uint32_t)cam.y0;
      pg.tile_w = (uint32_t)std::min<size_t>(ts, W - pg.tile_x);
      pg.tile_h = (uint32_t)cam.rows;
      pg.width = (uint32_t)W;
      pg.height = (uint32_t)H;
      pg.image = image.data();
      if (cb(&pg, user) != 0) return fail(FT_ERR_CALLBACK_ABORT, "render stopped by the progress callback");
    }
  }
  std::memcpy(rgb_out, image.data(), image.size());
  s->stats.primary_rays = (uint64_t)W * H * cam.aa;
  s->stats.kernel_launches = (uint64_t)tiles_y * (whitted ? 2 : 1);
  return FT_OK;
}

}  // namespace

// ------------------------------------------------------------------ launch-shape autotuner
namespace {

std::mutex g_tune_mu;
std::map<uint64_t, LaunchCfg> g_tune_cache;  // (scene, kind, request) -> best shape, per process

uint64_t tune_key(const ft_scene* s, uint32_t kind, const ft_render_options& o, uint32_t aa, uint32_t max_depth) {
  uint64_t h = fnv_bytes(s->ir_hash, &kind, sizeof(kind));
  h = fnv_bytes(h, &o, sizeof(o));
  h = fnv_bytes(h, &aa, sizeof(aa));
  h = fnv_bytes(h, &max_depth, sizeof(max_depth));
  const int emitted = s->source.empty() ? 0 : 1;
  return fnv_bytes(h, &emitted, sizeof(emitted));
}

bool tune_cache_get(uint64_t key, uint32_t kind, LaunchCfg& c) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune_cache.find(key);
  if (it != g_tune_cache.end()) {
    c = it->second;
    return true;
  }
  char name[48];
  std::snprintf(name, sizeof(name), "tune5_%016llx.txt", (unsigned long long)key);
  const std::string path = fth::cache_file(name);
  if (path.empty()) return false;
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  unsigned v[5];
  const bool ok = std::fscanf(f, "%u %u %u %u %u", &v[0], &v[1], &v[2], &v[3], &v[4]) == 5;
  std::fclose(f);
  if (!ok) return false;
  const LaunchCfg got{v[0], v[1], v[2], v[3], v[4]};
  if (!got.valid(kind)) return false;
  c = got;
  g_tune_cache[key] = c;
  return true;
}

void tune_cache_put(uint64_t key, const LaunchCfg& c) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[key] = c;
  char name[48];
  std::snprintf(name, sizeof(name), "tune5_%016llx.txt", (unsigned long long)key);
  const std::string path = fth::cache_file(name);
  if (path.empty()) return;
  const std::string tmp = path + ".tmp";
  if (FILE* f = std::fopen(tmp.c_str(), "w")) {  // best effort: the directory may not exist
    std::fprintf(f, "%u %u %u %u %u\n", c.block_waves, c.tile_w, c.wpe, c.unroll, c.shade_inline);
    std::fclose(f);
    std::rename(tmp.c_str(), path.c_str());
  }
}

// Kernel time (best of `reps` after one warm-up render) of the request under shape c.
int time_config(ft_scene* s, uint32_t kind, const LaunchCfg& c, const ft_render_options* o,
                const ft_ray_settings* rs, int reps, double& ms_out) {
  s->cfg[kind] = c;
  double best = 1e30;
  for (int r = 0; r <= reps; ++r) {
    int rc = kind == FT_KIND_DEPTH ? depth_impl(s, o, rs ? rs->aa_samples : 1, (unsigned char*)s->rgb.ptr, nullptr)
                                   : whitted_impl(s, o, rs, (unsigned char*)s->rgb.ptr, nullptr);
    if (rc) return rc;
    if ((rc = finish(s))) return rc;
    if (r > 0) best = std::min(best, (double)s->stats.last_kernel_ms);
  }
  ms_out = best;
  return FT_OK;
}

// Builds the code objects of `variants` for kernel k in parallel host threads and loads them.
int build_variants(ft_scene* s, int k, const std::vector<LaunchCfg>& variants) {
  std::vector<std::vector<char>> codes(variants.size());
  std::vector<std::string> errs(variants.size());
  std::vector<double> ms(variants.size(), 0.0);
  std::vector<char> ok(variants.size(), 1), need(variants.size(), 0);
  const std::string src = module_source(s);
  std::vector<std::thread> th;
  for (size_t i = 0; i < variants.size(); ++i) {
    if (!needs_module(s, false, variants[i]) || module_function(s, k, false, variants[i])) continue;
    need[i] = 1;
    th.emplace_back([&, i]() {
      ok[i] = fth::rtc_compile(src, k, false, (int)variants[i].wpe, (int)variants[i].unroll,
                               (int)variants[i].shade_inline, codes[i], ms[i], errs[i]);
    });
  }
  for (auto& t : th) t.join();
  for (size_t i = 0; i < variants.size(); ++i) {
    if (!need[i]) continue;
    if (!ok[i]) return fail(FT_ERR_COMPILE_FAILED, errs[i]);
    s->compile_ms += ms[i];
    int rc = load_module(s, k, false, variants[i], codes[i]);
    if (rc) return rc;
  }
  return FT_OK;
}

int autotune_impl(ft_scene* s, uint32_t kind, const ft_render_options* o, const ft_ray_settings* rs_in,
                  uint32_t flags, ft_tune_report* rep) {
  const auto t0 = std::chrono::steady_clock::now();
  std::memset(rep, 0, sizeof(*rep));
  if (s->dev.flops) return fail(FT_ERR_INVALID_ARGUMENT, "FLOP counting is on");
  int rc = check_options(o);
  if (rc) return rc;
  if (kind == FT_KIND_WHITTED && !rs_in) return fail(FT_ERR_INVALID_ARGUMENT, "ray settings are null");
  ft_ray_settings rs{};
  if (rs_in) rs = *rs_in;
  else rs.aa_samples = 1;
  // the tuning render: the request itself, or (Whitted) the same frame at <= 16 samples.  A
  // sample-major wave covers a pixel block whose size follows the sample count (4x4 pixels at
  // 4 spp, 2x2 at 16), so a render at fewer samples than the request can rank the shapes
  // differently (at 4 spp two-wave workgroups won, at 16 spp they took 68.7 against 59.5 ms).
  if (kind == FT_KIND_WHITTED && !(flags & FT_TUNE_FULL) && rs.aa_samples > 16) rs.aa_samples = 16;
  if ((rc = ensure(s->rgb, (size_t)o->width * o->height * 3))) return rc;
  int k = fth::kRtcDepth;
  if (kind == FT_KIND_WHITTED) {
    CamArgs cam;
    TraceParams tp;
    if ((rc = whitted_check(s, o, &rs, cam, tp, k))) return rc;
  }
  const uint64_t key = tune_key(s, kind, *o, rs.aa_samples, rs.max_depth);
  const LaunchCfg dflt = default_cfg(kind);
  LaunchCfg best = dflt;
  double best_ms = 0.0, dflt_ms = 0.0;
  uint32_t timed = 0;
  if (!(flags & FT_TUNE_NO_CACHE) && tune_cache_get(key, kind, best)) {
    rep->cached = 1;
    if ((rc = prepare(s, k, best))) return rc;
    if ((rc = time_config(s, kind, best, o, &rs, 1, best_ms))) return rc;
    if ((rc = time_config(s, kind, dflt, o, &rs, 1, dflt_ms))) return rc;
    timed = 2;
  } else {
    // stage 1: code variants (occupancy target, march unroll) in the default launch shape
    std::vector<LaunchCfg> code = {dflt};
    if (!(flags & FT_TUNE_SHAPE_ONLY)) {
      // (waves per EU, march unroll, shade inline) of each code variant
      static const uint32_t variants[7][3] = {{5, 1, 0}, {3, 1, 0}, {2, 1, 0}, {4, 2, 0}, {2, 2, 0}, {4, 1, 1},
                                              {2, 1, 1}};
      for (const auto& v : variants) {
        LaunchCfg c = dflt;
        c.wpe = v[0];
        c.unroll = v[1];
        c.shade_inline = v[2];
        if (kind == FT_KIND_WHITTED || c.shade_inline == 0) code.push_back(c);
      }
    }
    if ((rc = build_variants(s, k, code))) return rc;
    best_ms = 1e30;
    for (const LaunchCfg& c : code) {
      double ms = 0.0;
      if ((rc = time_config(s, kind, c, o, &rs, 2, ms))) return rc;
      ++timed;
      if (c == dflt) dflt_ms = ms;
      if (ms < best_ms) { best_ms = ms; best = c; }
    }
    // stage 2: launch shapes (waves per block, pixel tile) of the best code
    const LaunchCfg code_best = best;
    const bool sample_major = kind == FT_KIND_WHITTED && sample_major_ok((int)rs.aa_samples);
    for (uint32_t bw : {1u, 2u, 4u})
      for (uint32_t tw : {8u, 16u, 32u, 0u}) {
        if (tw == 0 && !sample_major) continue;
        LaunchCfg c = code_best;
        c.block_waves = bw;
        c.tile_w = tw;
        if (c == code_best) continue;
        double ms = 0.0;
        if ((rc = time_config(s, kind, c, o, &rs, 2, ms))) return rc;
        ++timed;
        if (ms < best_ms) { best_ms = ms; best = c; }
      }
    tune_cache_put(key, best);
  }
  s->cfg[kind] = best;
  rep->best.block_waves = best.block_waves;
  rep->best.tile_width = best.tile_w;
  rep->best.waves_per_eu = best.wpe;
  rep->best.march_unroll = best.unroll;
  rep->best.shade_inline = best.shade_inline;
  rep->best_ms = best_ms;
  rep->default_ms = dflt_ms;
  rep->candidates = timed;
  rep->tune_aa_samples = rs.aa_samples;
  rep->tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return FT_OK;
}

}  // namespace

static int create_scene(const ft_scene_ir* ir, uint32_t accel_mode, bool batch_member, ft_scene** out);

extern "C" {

const char* ft_last_error(void) { return g_err.c_str(); }

int ft_device_count(int* count) {
  if (!count) return fail(FT_ERR_INVALID_ARGUMENT, "count is null");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  *count = e == hipSuccess ? n : 0;
  return FT_OK;
}

int ft_scene_create(const ft_scene_ir* ir, uint32_t accel_mode, ft_scene** out) {
  return create_scene(ir, accel_mode, false, out);
}

}  // extern "C"

// Builds and uploads one scene.  Members of a batch share the batch's stream and always
// run on the interpreter.
static int create_scene(const ft_scene_ir* ir, uint32_t accel_mode, bool batch_member, ft_scene** out) {
  if (!out) return fail(FT_ERR_INVALID_ARGUMENT, "output handle is null");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(FT_ERR_GPU_UNAVAILABLE, "no HIP device available");
  ft_scene* s = new ft_scene();
  std::string err;
  if (!fth::build_scene(ir, accel_mode, s->host, err)) {
    delete s;
    return fail(err.find("root") != std::string::npos ? FT_ERR_MISSING_SCENE_ROOT : FT_ERR_INVALID_ARGUMENT, err);
  }
  int rc = FT_OK;
  auto cleanup = [&]() { ft_scene_destroy(s); };
  if (hipGetDevice(&s->device) != hipSuccess) { cleanup(); return fail(FT_ERR_GPU_UNAVAILABLE, "hipGetDevice failed"); }
  if (!batch_member && (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
                        hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)) {
    cleanup();
    return fail(FT_ERR_GPU_UNAVAILABLE, "stream/event creation failed");
  }
  fth::HostScene& h = s->host;
  if ((rc = upload(s->nodes, h.nodes)) || (rc = upload(s->prog, h.prog)) || (rc = upload(s->ops, h.ops)) ||
      (rc = upload(s->lbprog, h.lbprog)) ||
      (rc = upload(s->mats, h.mats)) || (rc = upload(s->hooks, h.hooks)) || (rc = upload(s->objects, h.objects)) || (rc = upload(s->lights, h.lights)) ||
      (rc = upload(s->fns, h.fns)) || (rc = upload(s->exprs, h.exprs)) || (rc = upload(s->bvh, h.bvh)) ||
      (rc = upload(s->leaves, h.leaves))) {
    std::string e = g_err;
    cleanup();
    return fail(rc, e);
  }
  DevScene& d = s->dev;
  std::memset(&d, 0, sizeof(d));
  d.nodes = (const DNode*)s->nodes.ptr;
  d.prog = (const unsigned*)s->prog.ptr;
  d.ops = (const DOp*)s->ops.ptr;
  d.lbprog = (const unsigned*)s->lbprog.ptr;
  d.mats = (const DParams*)s->mats.ptr;
  d.hooks = (const DHooks*)s->hooks.ptr;
  d.objects = (const float*)s->objects.ptr;
  d.lights = (const DLight*)s->lights.ptr;
  d.fns = (const ft_function*)s->fns.ptr;
  d.exprs = (const ft_expr*)s->exprs.ptr;
  d.bvh = (const DBvh*)s->bvh.ptr;
  d.leaves = (const DLeaf*)s->leaves.ptr;
  d.n_nodes = (int)h.nodes.size();
  d.root = h.root;
  d.n_prog = (int)h.prog.size();
  d.n_lbprog = (int)h.lbprog.size();
  d.n_mats = (int)h.mats.size();
  d.n_objects = (int)(h.objects.size() / 9);
  d.n_lights = (int)h.lights.size();
  d.env_fn = h.env_fn;
  d.accel = h.accel;
  d.simple = h.simple ? 1 : 0;
  d.root_walks = h.root_walks ? 1 : 0;
  d.n_leaves = (int)h.leaves.size();
  d.n_bvh = (int)h.bvh.size();
  d.stack_depth = std::max(1, h.max_stack);
  for (int k = 0; k < 3; ++k) { d.bmin[k] = h.bmin[k]; d.bmax[k] = h.bmax[k]; d.env_rad[k] = h.env_rad[k]; }
  // scene specialisation (FT_EMIT: unset/"auto", "1" always, "0" never)
  {
    const char* mode = std::getenv("FT_EMIT");
    const std::string m = mode ? mode : "auto";
    s->plan = fth::plan_emission(h, ir);
    bool emit = false;
    if (m == "1" || m == "always") emit = s->plan.any();
    else if (m != "0" && m != "never") emit = s->plan.any() && fth::emission_pays(h, ir, s->plan);
    if (emit && !batch_member) s->source = fth::emit_scene_source(h, ir, s->plan);
  }
  s->ir_hash = ir_hash(ir, accel_mode);
  s->slot = acquire_slot();
  if (s->slot < 0) {
    cleanup();
    return fail(FT_ERR_OUT_OF_MEMORY, "too many live scenes (limit " + std::to_string(kMaxScenes) + ")");
  }
  if ((rc = publish(s))) {
    std::string e = g_err;
    cleanup();
    return fail(rc, e);
  }
  *out = s;
  return FT_OK;
}

extern "C" {

int ft_scene_destroy(ft_scene* s) {
  if (!s) return FT_OK;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (DevBuf* b : {&s->nodes, &s->prog, &s->ops, &s->lbprog, &s->mats, &s->hooks, &s->objects, &s->lights, &s->fns, &s->exprs, &s->bvh,
                    &s->leaves, &s->rgb, &s->tbuf, &s->samples, &s->hdr, &s->overflow, &s->flops, &s->atlas_cell_d,
                    &s->atlas_cells, &s->atlas_brick_cell, &s->atlas_bricks, &s->atlas_counter})
    if (b->ptr) (void)hipFree(b->ptr);
  for (auto& kv : s->mods) (void)hipModuleUnload(kv.second.first);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  release_slot(s->slot);
  delete s;
  return FT_OK;
}

int ft_render_depth_device(ft_scene* s, const ft_render_options* o, uint32_t aa, void* rgb_dev, void* t_dev) {
  if (!s || !rgb_dev) return fail(FT_ERR_INVALID_ARGUMENT, "scene or output is null");
  return depth_impl(s, o, aa, (unsigned char*)rgb_dev, (float*)t_dev);
}

int ft_render_whitted_device(ft_scene* s, const ft_render_options* o, const ft_ray_settings* rs, void* rgb_dev,
                             void* hdr_dev) {
  if (!s || !rgb_dev) return fail(FT_ERR_INVALID_ARGUMENT, "scene or output is null");
  return whitted_impl(s, o, rs, (unsigned char*)rgb_dev, (float*)hdr_dev);
}

int ft_render_whitted_device_noalloc(ft_scene* s, const ft_render_options* o, const ft_ray_settings* rs,
                                     void* rgb_dev, void* hdr_dev) {
  if (!s || !rgb_dev) return fail(FT_ERR_INVALID_ARGUMENT, "scene or output is null");
  CamArgs cam;
  TraceParams tp;
  int kernel;
  int rc = whitted_check(s, o, rs, cam, tp, kernel);
  if (rc) return rc;
  const uint64_t n_samples = (uint64_t)cam.aa * cam.width * cam.height;
  if (n_samples > FT_NOALLOC_MAX_SAMPLES)
    return fail(FT_ERR_INVALID_ARGUMENT, "aa * width * height = " + std::to_string(n_samples) +
                                             " exceeds the allocation-free limit of " +
                                             std::to_string(FT_NOALLOC_MAX_SAMPLES) + " samples");
  if (s->dev.flops) return fail(FT_ERR_INVALID_ARGUMENT, "FLOP counting is on (its kernels are built on first use)");
  if (needs_module(s, false, s->cfg[1]) && !module_function(s, kernel, false, s->cfg[1]))
    return fail(FT_ERR_INVALID_ARGUMENT, "the scene's specialised Whitted kernel is not built yet: render once "
                                         "with ft_render_whitted or ft_render_whitted_device first");
  unsigned char* base = nullptr;
  if ((rc = noalloc_scratch(s->device, &base))) return rc;
  s->stats.sdf_flops = 0;
  return whitted_enqueue(s, cam, tp, kernel, rs->debug_aov, (unsigned char*)rgb_dev, (float*)hdr_dev,
                         (float*)(base + kNoallocSamplesOffset), (int*)base);
}

int ft_scene_synchronize(ft_scene* s) {
  if (!s) return fail(FT_ERR_INVALID_ARGUMENT, "scene is null");
  return finish(s);
}

int ft_render_depth(ft_scene* s, const ft_render_options* o, uint32_t aa, uint8_t* rgb_out, float* t_out) {
  if (!s || !rgb_out) return fail(FT_ERR_INVALID_ARGUMENT, "scene or output is null");
  int rc = check_options(o);
  if (rc) return rc;
  size_t npix = (size_t)o->width * o->height;
  if ((rc = ensure(s->rgb, npix * 3))) return rc;
  if (t_out && (rc = ensure(s->tbuf, npix * sizeof(float)))) return rc;
  if ((rc = depth_impl(s, o, aa, (unsigned char*)s->rgb.ptr, t_out ? (float*)s->tbuf.ptr : nullptr))) return rc;
  if ((rc = finish(s))) return rc;
  FT_HIP(hipMemcpy(rgb_out, s->rgb.ptr, npix * 3, hipMemcpyDeviceToHost));
  if (t_out) FT_HIP(hipMemcpy(t_out, s->tbuf.ptr, npix * sizeof(float), hipMemcpyDeviceToHost));
  return FT_OK;
}

int ft_render_whitted(ft_scene* s, const ft_render_options* o, const ft_ray_settings* rs, uint8_t* rgb_out,
                      float* hdr_out) {
  if (!s || !rgb_out) return fail(FT_ERR_INVALID_ARGUMENT, "scene or output is null");
  int rc = check_options(o);
  if (rc) return rc;
  size_t npix = (size_t)o->width * o->height;
  if ((rc = ensure(s->rgb, npix * 3))) return rc;
  if (hdr_out && (rc = ensure(s->hdr, npix * 3 * sizeof(float)))) return rc;
  if ((rc = whitted_impl(s, o, rs, (unsigned char*)s->rgb.ptr, hdr_out ? (float*)s->hdr.ptr : nullptr))) return rc;
  if ((rc = finish(s))) return rc;
  FT_HIP(hipMemcpy(rgb_out, s->rgb.ptr, npix * 3, hipMemcpyDeviceToHost));
  if (hdr_out) FT_HIP(hipMemcpy(hdr_out, s->hdr.ptr, npix * 3 * sizeof(float), hipMemcpyDeviceToHost));
  return FT_OK;
}

int ft_render_depth_progressive(ft_scene* s, const ft_render_options* o, uint32_t aa, uint32_t tile_size,
                                ft_progress_fn cb, void* user, uint8_t* rgb_out) {
  if (!s || !rgb_out) return fail(FT_ERR_INVALID_ARGUMENT, "scene or output is null");
  return progressive_impl(s, false, o, nullptr, aa, tile_size, cb, user, rgb_out);
}

int ft_render_whitted_progressive(ft_scene* s, const ft_render_options* o, const ft_ray_settings* rs,
                                  ft_progress_fn cb, void* user, uint8_t* rgb_out) {
  if (!s || !rgb_out) return fail(FT_ERR_INVALID_ARGUMENT, "scene or output is null");
  if (!rs) return fail(FT_ERR_INVALID_ARGUMENT, "ray settings are null");
  return progressive_impl(s, true, o, rs, rs->aa_samples, rs->tile_size, cb, user, rgb_out);
}

int ft_scene_stream(ft_scene* s, void** stream_out) {
  if (!s || !stream_out) return fail(FT_ERR_INVALID_ARGUMENT, "scene or output is null");
  *stream_out = (void*)s->stream;
  return FT_OK;
}

int ft_scene_stats(ft_scene* s, ft_render_stats* out) {
  if (!s || !out) return fail(FT_ERR_INVALID_ARGUMENT, "scene or output is null");
  *out = s->stats;
  out->flags = 0;
  if (!s->source.empty()) {
    if (s->plan.scene) out->flags |= FT_STATS_EMITTED_SCENE;
    if (s->plan.functions) out->flags |= FT_STATS_EMITTED_FUNCTIONS;
  }
  out->compile_ms = s->compile_ms;
  return FT_OK;
}

int ft_scene_build_atlas(ft_scene* s, float voxel, uint32_t bits, uint32_t max_bricks) {
  if (!s) return fail(FT_ERR_INVALID_ARGUMENT, "scene is null");
  FT_HIP(hipSetDevice(s->device));
  FT_HIP(hipStreamSynchronize(s->stream));
  if (bits == 0) {
    s->dev.atlas_on = 0;
    s->stats.atlas_bytes = 0;
    s->stats.atlas_bricks = 0;
    return publish(s);
  }
  if (bits != 16 && bits != 32) return fail(FT_ERR_INVALID_ARGUMENT, "atlas bits must be 16 or 32");
  if (!(voxel > 0.0f)) return fail(FT_ERR_INVALID_ARGUMENT, "voxel size must be positive");
  if (max_bricks == 0) max_bricks = 1u << 20;
  AtlasGrid g;
  const float cell = 8.0f * voxel;
  g.ox = s->host.bmin[0] - voxel;
  g.oy = s->host.bmin[1] - voxel;
  g.oz = s->host.bmin[2] - voxel;
  g.voxel = voxel;
  g.nx = (int)std::ceil((s->host.bmax[0] - s->host.bmin[0] + 2.0f * voxel) / cell);
  g.ny = (int)std::ceil((s->host.bmax[1] - s->host.bmin[1] + 2.0f * voxel) / cell);
  g.nz = (int)std::ceil((s->host.bmax[2] - s->host.bmin[2] + 2.0f * voxel) / cell);
  if (!(g.nx > 0 && g.ny > 0 && g.nz > 0) || (double)g.nx * g.ny * g.nz > (double)(1 << 26))
    return fail(FT_ERR_INVALID_ARGUMENT, "atlas grid too large for this voxel size");
  g.band = 1.5f * 0.8660254f * cell + 2.0f * voxel;
  // sample indices of the brick pass stay below 2^31
  g.max_bricks = (int)std::min<uint32_t>(max_bricks, (uint32_t)(((1u << 31) - 1) / kBrickSamples));
  const size_t n_cells = (size_t)g.nx * g.ny * g.nz;
  int rc;
  if ((rc = ensure(s->atlas_cell_d, n_cells * sizeof(float))) || (rc = ensure(s->atlas_cells, n_cells * sizeof(int))) ||
      (rc = ensure(s->atlas_brick_cell, (size_t)g.max_bricks * sizeof(int))) ||
      (rc = ensure(s->atlas_counter, sizeof(unsigned))))
    return rc;
  // the build evaluates the exact scene: no coarse steps, no counting
  s->dev.atlas_on = 0;
  unsigned long long* saved_flops = s->dev.flops;
  unsigned long long* saved_samples = s->dev.atlas_samples;
  s->dev.flops = nullptr;
  s->dev.atlas_samples = nullptr;
  if ((rc = publish(s))) return rc;
  hipLaunchKernelGGL(ft_k_clear_flag, dim3(1), dim3(64), 0, s->stream, (int*)s->atlas_counter.ptr);
  const unsigned cell_blocks = (unsigned)((n_cells + 255) / 256);
  hipLaunchKernelGGL(ft_k_atlas_cells, dim3(cell_blocks), dim3(64, 4), stack_lds(s), s->stream, s->slot, g,
                     (float*)s->atlas_cell_d.ptr, (int*)s->atlas_cells.ptr, (int*)s->atlas_brick_cell.ptr,
                     (unsigned*)s->atlas_counter.ptr);
  FT_HIP(hipGetLastError());
  unsigned count = 0;
  FT_HIP(hipStreamSynchronize(s->stream));
  FT_HIP(hipMemcpy(&count, s->atlas_counter.ptr, sizeof(count), hipMemcpyDeviceToHost));
  const int n_bricks = (int)std::min<unsigned>(count, (unsigned)g.max_bricks);
  const size_t sample_bytes = bits == 16 ? 2 : 4;
  if (n_bricks > 0) {
    if ((rc = ensure(s->atlas_bricks, (size_t)n_bricks * kBrickSamples * sample_bytes))) return rc;
    const size_t n_samples = (size_t)n_bricks * kBrickSamples;
    hipLaunchKernelGGL(ft_k_atlas_bricks, dim3((unsigned)((n_samples + 255) / 256)), dim3(64, 4), stack_lds(s),
                       s->stream, s->slot, g, (const int*)s->atlas_brick_cell.ptr, n_bricks, bits == 16 ? 1 : 0,
                       s->atlas_bricks.ptr);
    FT_HIP(hipGetLastError());
    FT_HIP(hipStreamSynchronize(s->stream));
  }
  DevScene& d = s->dev;
  d.atlas_cells = (const int*)s->atlas_cells.ptr;
  d.atlas_cell_d = (const float*)s->atlas_cell_d.ptr;
  d.atlas_bricks = s->atlas_bricks.ptr;
  d.atlas_origin[0] = g.ox;
  d.atlas_origin[1] = g.oy;
  d.atlas_origin[2] = g.oz;
  d.atlas_voxel = voxel;
  d.atlas_dims[0] = g.nx;
  d.atlas_dims[1] = g.ny;
  d.atlas_dims[2] = g.nz;
  d.atlas_fp16 = bits == 16 ? 1 : 0;
  d.atlas_switch = 2.0f * voxel;
  d.atlas_on = 1;
  d.flops = saved_flops;
  d.atlas_samples = saved_samples;
  s->stats.atlas_bricks = (uint64_t)n_bricks;
  s->stats.atlas_bytes = (uint64_t)n_cells * (sizeof(float) + sizeof(int)) + (uint64_t)n_bricks * kBrickSamples * sample_bytes;
  return publish(s);
}

int ft_scene_set_launch_config(ft_scene* s, uint32_t kind, const ft_launch_config* c) {
  if (!s || !c) return fail(FT_ERR_INVALID_ARGUMENT, "scene or config is null");
  if (kind > FT_KIND_WHITTED) return fail(FT_ERR_INVALID_ARGUMENT, "unknown kernel kind");
  const LaunchCfg l{c->block_waves, c->tile_width, c->waves_per_eu, c->march_unroll, c->shade_inline};
  if (!l.valid(kind)) return fail(FT_ERR_INVALID_ARGUMENT, "launch config out of range");
  s->cfg[kind] = l;
  return FT_OK;
}

int ft_scene_get_launch_config(ft_scene* s, uint32_t kind, ft_launch_config* c) {
  if (!s || !c) return fail(FT_ERR_INVALID_ARGUMENT, "scene or config is null");
  if (kind > FT_KIND_WHITTED) return fail(FT_ERR_INVALID_ARGUMENT, "unknown kernel kind");
  const LaunchCfg& l = s->cfg[kind];
  c->block_waves = l.block_waves;
  c->tile_width = l.tile_w;
  c->waves_per_eu = l.wpe;
  c->march_unroll = l.unroll;
  c->shade_inline = l.shade_inline;
  return FT_OK;
}

int ft_scene_autotune(ft_scene* s, uint32_t kind, const ft_render_options* o, const ft_ray_settings* rs,
                      uint32_t flags, ft_tune_report* report) {
  if (!s || !report) return fail(FT_ERR_INVALID_ARGUMENT, "scene or report is null");
  if (kind > FT_KIND_WHITTED) return fail(FT_ERR_INVALID_ARGUMENT, "unknown kernel kind");
  FT_HIP(hipSetDevice(s->device));
  return autotune_impl(s, kind, o, rs, flags, report);
}

int ft_scene_set_flop_counting(ft_scene* s, int enable) {
  if (!s) return fail(FT_ERR_INVALID_ARGUMENT, "scene is null");
  FT_HIP(hipStreamSynchronize(s->stream));
  if (enable) {
    int rc = ensure(s->flops, 2 * sizeof(unsigned long long));
    if (rc) return rc;
    s->dev.flops = (unsigned long long*)s->flops.ptr;
    s->dev.atlas_samples = s->dev.flops + 1;
  } else {
    s->dev.flops = nullptr;
    s->dev.atlas_samples = nullptr;
  }
  return publish(s);
}

int ft_emit_scene_source(const ft_scene_ir* ir, uint32_t accel_mode, char* buf, uint64_t cap, uint64_t* len_out) {
  fth::HostScene h;
  std::string err;
  if (!fth::build_scene(ir, accel_mode, h, err))
    return fail(err.find("root") != std::string::npos ? FT_ERR_MISSING_SCENE_ROOT : FT_ERR_INVALID_ARGUMENT, err);
  const std::string src = fth::emit_scene_source(h, ir, fth::plan_emission(h, ir));
  if (len_out) *len_out = src.size();
  if (buf && cap > 0) {
    const size_t n = std::min<size_t>(src.size(), (size_t)cap - 1);
    std::memcpy(buf, src.data(), n);
    buf[n] = '\0';
  }
  return FT_OK;
}

int ft_emit_compile(const ft_scene_ir* ir, uint32_t accel_mode, int kernel, double* compile_ms_out) {
  if (kernel < 0 || kernel >= fth::kRtcKernelCount) return fail(FT_ERR_INVALID_ARGUMENT, "unknown kernel");
  fth::HostScene h;
  std::string err;
  if (!fth::build_scene(ir, accel_mode, h, err))
    return fail(err.find("root") != std::string::npos ? FT_ERR_MISSING_SCENE_ROOT : FT_ERR_INVALID_ARGUMENT, err);
  const std::string src = fth::emit_scene_source(h, ir, fth::plan_emission(h, ir));
  std::vector<char> code;
  double ms = 0.0;
  if (!fth::rtc_compile(src, kernel, false, 4, 1, 0, code, ms, err)) return fail(FT_ERR_COMPILE_FAILED, err);
  if (compile_ms_out) *compile_ms_out = ms;
  return FT_OK;
}

// sizeof() of every ABI struct, so bindings can verify their layout.
int ft_abi_struct_size(const char* name) {
  if (!name) return -1;
  std::string n(name);
  if (n == "ft_node") return (int)sizeof(ft_node);
  if (n == "ft_expr") return (int)sizeof(ft_expr);
  if (n == "ft_function") return (int)sizeof(ft_function);
  if (n == "ft_material") return (int)sizeof(ft_material);
  if (n == "ft_material_hooks") return (int)sizeof(ft_material_hooks);
  if (n == "ft_light") return (int)sizeof(ft_light);
  if (n == "ft_camera") return (int)sizeof(ft_camera);
  if (n == "ft_scene_ir") return (int)sizeof(ft_scene_ir);
  if (n == "ft_render_options") return (int)sizeof(ft_render_options);
  if (n == "ft_ray_settings") return (int)sizeof(ft_ray_settings);
  if (n == "ft_render_stats") return (int)sizeof(ft_render_stats);
  if (n == "ft_view") return (int)sizeof(ft_view);
  if (n == "ft_launch_config") return (int)sizeof(ft_launch_config);
  if (n == "ft_tune_report") return (int)sizeof(ft_tune_report);
  if (n == "ft_progress") return (int)sizeof(ft_progress);
  return -1;
}

}  // extern "C"

// ------------------------------------------------------------------------- batches
struct ft_batch {
  std::vector<ft_scene*> scenes;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevBuf order, slots, cams, samples, overflow, flops, queue, tcache, plist;
  int cus = 0;  // compute units of the device (persistent launch size)
  bool counting = false;
  std::vector<int> h_order, h_slots;
  std::vector<CamArgs> h_cams;
  int stack_depth = 1;
  long max_children = 2;
  ft_render_stats stats{};
  bool timed = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t graph_key = 0;
};

namespace {

uint64_t mix_bytes(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = (const unsigned char*)p;
  for (size_t k = 0; k < n; ++k) {
    h ^= b[k];
    h *= 1099511628211ull;
  }
  return h;
}

void drop_graph(ft_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_key = 0;
}

// The batch depth pass keeps its primary hits for the Whitted pass of the same call when
// both run at one sample per pixel and the depth pass is the persistent kernel.
bool primary_cache(uint32_t flags, const CamArgs& c0) {
  return (flags & FT_BATCH_DEPTH) && (flags & FT_BATCH_WHITTED) && !(flags & FT_BATCH_TILED) && c0.aa == 1;
}

int ray_stack_kernel(long need) {
  if (need <= 16) return 16;
  if (need <= 48) return 48;
  if (need <= 128) return 128;
  return 0;
}

// Enqueues the launches of one batch render on the batch's stream (captured or not).
int batch_enqueue(ft_batch* b, uint32_t n_views, const TraceParams& tp, int aov, const CamArgs& c0, uint32_t flags,
                  unsigned char* depth_rgb, unsigned char* whitted_rgb, int& launches) {
  const int tpv = ((c0.width + 7) / 8) * ((c0.height + 7) / 8);
  const size_t npix = (size_t)c0.width * c0.height;
  const size_t lds = (size_t)3 * b->stack_depth * 256 * sizeof(float);
  float* samples = (float*)b->samples.ptr;
  int* ovf = (int*)b->overflow.ptr;
  launches = 0;
  if (flags & FT_BATCH_WHITTED) {
    hipLaunchKernelGGL(ft_k_clear_flag, dim3(1), dim3(64), 0, b->stream, ovf);
    ++launches;
  }
  if (flags & FT_BATCH_SEPARATE) {
    // the single-scene path, once per view
    for (uint32_t r = 0; r < n_views; ++r) {
      const int v = b->h_order[r];
      ft_scene* s = b->scenes[b->h_slots[v]];
      const CamArgs cam = b->h_cams[v];
      const size_t slds = stack_lds(s);
      if (flags & FT_BATCH_DEPTH) {
        hipLaunchKernelGGL(ft_k_depth, tile_grid(cam), dim3(64, 4), slds, b->stream, s->slot, cam, tp,
                           depth_rgb + (size_t)v * npix * 3, (float*)nullptr);
        ++launches;
      }
      if (flags & FT_BATCH_WHITTED) {
        const long need = 1 + (long)(tp.max_depth - 1) * (long)(s->host.max_children - 1);
        dim3 grid = tile_grid(cam);
        grid.y = (unsigned)cam.aa;
        float* vs = samples + (size_t)v * cam.aa * npix * 3;
        switch (ray_stack_kernel(need)) {
          case 16: hipLaunchKernelGGL(ft_k_whitted_16, grid, dim3(64, 4), slds, b->stream, s->slot, cam, tp, aov, vs, ovf); break;
          case 48: hipLaunchKernelGGL(ft_k_whitted_48, grid, dim3(64, 4), slds, b->stream, s->slot, cam, tp, aov, vs, ovf); break;
          default: hipLaunchKernelGGL(ft_k_whitted_128, grid, dim3(64, 4), slds, b->stream, s->slot, cam, tp, aov, vs, ovf); break;
        }
        hipLaunchKernelGGL(ft_k_resolve, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, b->stream, cam, aov, vs,
                           whitted_rgb + (size_t)v * npix * 3, (float*)nullptr);
        launches += 2;
      }
    }
    FT_HIP(hipGetLastError());
    return FT_OK;
  }
  BatchArgs ba;
  ba.order = (const int*)b->order.ptr;
  ba.slots = (const int*)b->slots.ptr;
  ba.cams = (const CamArgs*)b->cams.ptr;
  ba.n_views = (int)n_views;
  ba.tiles_per_view = tpv;
  const unsigned waves = (unsigned)((size_t)n_views * tpv);
  const dim3 grid((waves + 3) / 4);
  float* const t_cache = primary_cache(flags, c0) ? (float*)b->tcache.ptr : nullptr;
  if (flags & FT_BATCH_DEPTH) {
    if (c0.aa == 1 && !(flags & FT_BATCH_TILED)) {
      // persistent: four 4-wave blocks per CU, each wave drawing chunks from the counter;
      // with a Whitted pass in the same call its primary hits are kept for it (t_cache)
      unsigned* queue = (unsigned*)b->queue.ptr;  // one chunk counter per XCD
      const unsigned blocks = std::min((unsigned)std::max(b->cus, 1) * 4u, (waves + 3) / 4);
      hipLaunchKernelGGL(ft_k_clear_u64, dim3(1), dim3(64), 0, b->stream, (unsigned long long*)queue,
                         (int)(kXcds / 2));
      hipLaunchKernelGGL(ft_k_batch_depth_persistent, dim3(blocks), dim3(64, 4), lds, b->stream, ba, tp, depth_rgb,
                         t_cache, queue);
      launches += 2;
    } else {
      hipLaunchKernelGGL(ft_k_batch_depth, grid, dim3(64, 4), lds, b->stream, ba, tp, depth_rgb);
      ++launches;
    }
  }
  if ((flags & FT_BATCH_WHITTED) && t_cache && aov < 0) {
    // hit-first pixel order per view, then the Whitted pass over it from the primary hits
    unsigned* plist = (unsigned*)b->plist.ptr;
    hipLaunchKernelGGL(ft_k_batch_partition, dim3(n_views), dim3(256), 0, b->stream, c0.width, c0.height,
                       (const float*)t_cache, plist);
    switch (ray_stack_kernel(1 + (long)(tp.max_depth - 1) * (b->max_children - 1))) {
      case 16: hipLaunchKernelGGL(ft_k_batch_whitted_list_16, grid, dim3(64, 4), lds, b->stream, ba, tp, samples, ovf,
                                  (const float*)t_cache, (const unsigned*)plist); break;
      case 48: hipLaunchKernelGGL(ft_k_batch_whitted_list_48, grid, dim3(64, 4), lds, b->stream, ba, tp, samples, ovf,
                                  (const float*)t_cache, (const unsigned*)plist); break;
      default: hipLaunchKernelGGL(ft_k_batch_whitted_list_128, grid, dim3(64, 4), lds, b->stream, ba, tp, samples, ovf,
                                  (const float*)t_cache, (const unsigned*)plist); break;
    }
    hipLaunchKernelGGL(ft_k_resolve, dim3((unsigned)((npix + 255) / 256), n_views), dim3(256), 0, b->stream, c0, aov,
                       samples, whitted_rgb, (float*)nullptr);
    launches += 3;
  } else if (flags & FT_BATCH_WHITTED) {
    dim3 wg = grid;
    wg.y = (unsigned)c0.aa;
    switch (ray_stack_kernel(1 + (long)(tp.max_depth - 1) * (b->max_children - 1))) {
      case 16: hipLaunchKernelGGL(ft_k_batch_whitted_16, wg, dim3(64, 4), lds, b->stream, ba, tp, aov, samples, ovf,
                                  (const float*)t_cache); break;
      case 48: hipLaunchKernelGGL(ft_k_batch_whitted_48, wg, dim3(64, 4), lds, b->stream, ba, tp, aov, samples, ovf,
                                  (const float*)t_cache); break;
      default: hipLaunchKernelGGL(ft_k_batch_whitted_128, wg, dim3(64, 4), lds, b->stream, ba, tp, aov, samples, ovf,
                                  (const float*)t_cache); break;
    }
    hipLaunchKernelGGL(ft_k_resolve, dim3((unsigned)((npix + 255) / 256), n_views), dim3(256), 0, b->stream, c0, aov,
                       samples, whitted_rgb, (float*)nullptr);
    launches += 2;
  }
  FT_HIP(hipGetLastError());
  return FT_OK;
}

int batch_render_impl(ft_batch* b, const ft_view* views, uint32_t n_views, const ft_render_options* o,
                      const ft_ray_settings* rs, uint32_t flags, unsigned char* depth_rgb, unsigned char* whitted_rgb) {
  int rc = check_options(o);
  if (rc) return rc;
  if (!views || n_views == 0) return fail(FT_ERR_INVALID_ARGUMENT, "no views");
  if (!(flags & (FT_BATCH_DEPTH | FT_BATCH_WHITTED))) return fail(FT_ERR_INVALID_ARGUMENT, "nothing to render");
  if ((flags & FT_BATCH_DEPTH) && !depth_rgb) return fail(FT_ERR_INVALID_ARGUMENT, "depth output is null");
  if ((flags & FT_BATCH_WHITTED) && !whitted_rgb) return fail(FT_ERR_INVALID_ARGUMENT, "whitted output is null");
  if ((flags & FT_BATCH_WHITTED) && !rs) return fail(FT_ERR_INVALID_ARGUMENT, "ray settings are null");
  const int aov = rs ? rs->debug_aov : -1;
  if (aov < -1 || aov > FT_AOV_HIT_T) return fail(FT_ERR_INVALID_ARGUMENT, "unknown debug AOV");
  const uint32_t aa = (flags & FT_BATCH_WHITTED) ? rs->aa_samples : 1;
  const int tpv = (((int)o->width + 7) / 8) * (((int)o->height + 7) / 8);
  if ((uint64_t)n_views * tpv > (1ull << 31) / 64) return fail(FT_ERR_INVALID_ARGUMENT, "batch too large");
  TraceParams tp = make_tp(*o, rs ? rs->max_depth : 1);
  if ((flags & FT_BATCH_WHITTED) && !ray_stack_kernel(1 + (long)(tp.max_depth - 1) * (b->max_children - 1)))
    return fail(FT_ERR_INVALID_ARGUMENT, "max_depth too large for the ray stack with these scenes' materials");
  // per-view cameras, slots and the cost order
  std::vector<CamArgs> cams(n_views);
  std::vector<int> slots(n_views), order(n_views);
  std::vector<float> cost(n_views);
  for (uint32_t v = 0; v < n_views; ++v) {
    if (views[v].scene >= b->scenes.size()) return fail(FT_ERR_INVALID_ARGUMENT, "view names an unknown scene");
    ft_scene* s = b->scenes[views[v].scene];
    if ((flags & FT_BATCH_WHITTED) && s->host.dynamic_material >= 0)
      return fail(FT_ERR_UNSUPPORTED_OBJECT, "a scene has materials with per-hit hook functions");
    fth::CamBasis cb = fth::camera_basis(views[v].origin, views[v].target, views[v].fov_y_degrees);
    CamArgs& c = cams[v];
    c.ox = cb.origin[0]; c.oy = cb.origin[1]; c.oz = cb.origin[2];
    c.fx = cb.forward[0]; c.fy = cb.forward[1]; c.fz = cb.forward[2];
    c.rx = cb.right[0]; c.ry = cb.right[1]; c.rz = cb.right[2];
    c.ux = cb.up[0]; c.uy = cb.up[1]; c.uz = cb.up[2];
    c.tan_fov = cb.tan_fov;
    c.aspect = (float)o->width / (float)o->height;
    c.width = (int)o->width;
    c.height = (int)o->height;
    c.aa = (int)(aa < 1 ? 1 : aa);
    c.grid = 1;
    while (c.grid * c.grid < c.aa) ++c.grid;
    c.tile_w = 8;  // the batch kernels' 8x8 tiles (tiles_per_view)
    c.y0 = 0;
    c.rows = c.height;
    slots[v] = (int)views[v].scene;  // host side: scene index; device side: constant slot
    order[v] = (int)v;
    cost[v] = s->host.nodes[s->host.root].cost;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
  FT_HIP(hipSetDevice(b->device));
  const size_t npix = (size_t)o->width * o->height;
  if (flags & FT_BATCH_WHITTED) {
    if ((rc = ensure(b->samples, (size_t)n_views * cams[0].aa * npix * 3 * sizeof(float)))) return rc;
    if ((rc = ensure(b->overflow, sizeof(int)))) return rc;
  }
  if ((rc = ensure(b->queue, kXcds * sizeof(unsigned)))) return rc;
  if (primary_cache(flags, cams[0]) && ((rc = ensure(b->tcache, (size_t)n_views * npix * sizeof(float))) ||
                                        (rc = ensure(b->plist, (size_t)n_views * npix * sizeof(unsigned)))))
    return rc;
  if (b->cus == 0) FT_HIP(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device));
  // device tables (uploaded when they change; a captured graph reads them in place)
  std::vector<int> dev_slots(n_views);
  for (uint32_t v = 0; v < n_views; ++v) dev_slots[v] = b->scenes[slots[v]]->slot;
  const bool tables_changed = order != b->h_order || slots != b->h_slots || cams.size() != b->h_cams.size() ||
                              std::memcmp(cams.data(), b->h_cams.data(), cams.size() * sizeof(CamArgs)) != 0;
  if (tables_changed) {
    FT_HIP(hipStreamSynchronize(b->stream));
    if ((rc = ensure(b->order, n_views * sizeof(int))) || (rc = ensure(b->slots, n_views * sizeof(int))) ||
        (rc = ensure(b->cams, n_views * sizeof(CamArgs))))
      return rc;
    FT_HIP(hipMemcpy(b->order.ptr, order.data(), n_views * sizeof(int), hipMemcpyHostToDevice));
    FT_HIP(hipMemcpy(b->slots.ptr, dev_slots.data(), n_views * sizeof(int), hipMemcpyHostToDevice));
    FT_HIP(hipMemcpy(b->cams.ptr, cams.data(), n_views * sizeof(CamArgs), hipMemcpyHostToDevice));
    b->h_order = order;
    b->h_slots = slots;
    b->h_cams = cams;
  }
  int launches = 0;
  if (b->counting) {
    hipLaunchKernelGGL(ft_k_clear_u64, dim3(1), dim3(64), 0, b->stream, (unsigned long long*)b->flops.ptr, 2);
    FT_HIP(hipGetLastError());
  }
  FT_HIP(hipEventRecord(b->ev0, b->stream));
  if (flags & FT_BATCH_GRAPH) {
    uint64_t key = 1469598103934665603ull;
    key = mix_bytes(key, views, n_views * sizeof(ft_view));
    key = mix_bytes(key, o, sizeof(*o));
    if (rs) key = mix_bytes(key, rs, sizeof(*rs));
    key = mix_bytes(key, &flags, sizeof(flags));
    key = mix_bytes(key, &depth_rgb, sizeof(depth_rgb));
    key = mix_bytes(key, &whitted_rgb, sizeof(whitted_rgb));
    key = mix_bytes(key, &b->samples.ptr, sizeof(void*));
    key = mix_bytes(key, &b->order.ptr, sizeof(void*));
    key = mix_bytes(key, &b->queue.ptr, sizeof(void*));
    key = mix_bytes(key, &b->tcache.ptr, sizeof(void*));
    key = mix_bytes(key, &b->plist.ptr, sizeof(void*));
    if (!b->exec || key != b->graph_key || tables_changed) {
      drop_graph(b);
      FT_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      rc = batch_enqueue(b, n_views, tp, aov, cams[0], flags & ~FT_BATCH_GRAPH, depth_rgb, whitted_rgb, launches);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      FT_HIP(e);
      b->graph = g;
      FT_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
      b->graph_key = key;
    } else {
      launches = (int)b->stats.kernel_launches;  // same sequence as the captured call
    }
    FT_HIP(hipGraphLaunch(b->exec, b->stream));
  } else {
    if ((rc = batch_enqueue(b, n_views, tp, aov, cams[0], flags, depth_rgb, whitted_rgb, launches))) return rc;
  }
  FT_HIP(hipEventRecord(b->ev1, b->stream));
  b->timed = true;
  b->stats.primary_rays = (uint64_t)n_views * npix *
                          (((flags & FT_BATCH_DEPTH) ? 1u : 0u) + ((flags & FT_BATCH_WHITTED) ? (uint64_t)cams[0].aa : 0u));
  b->stats.kernel_launches = (uint64_t)launches;
  return FT_OK;
}

int batch_finish(ft_batch* b) {
  FT_HIP(hipStreamSynchronize(b->stream));
  if (b->counting) {
    unsigned long long f = 0;
    FT_HIP(hipMemcpy(&f, b->flops.ptr, sizeof(f), hipMemcpyDeviceToHost));
    b->stats.sdf_flops = f;
  }
  if (b->timed) {
    float ms = 0.0f;
    FT_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->stats.last_kernel_ms = ms;
    b->timed = false;
  }
  if (b->overflow.ptr) {
    int ovf = 0;
    FT_HIP(hipMemcpy(&ovf, b->overflow.ptr, sizeof(int), hipMemcpyDeviceToHost));
    if (ovf) return fail(FT_ERR_LAUNCH_FAILED, "ray stack overflow");
  }
  return FT_OK;
}

}  // namespace

extern "C" {

int ft_batch_create(const ft_scene_ir* const* scenes, uint32_t n_scenes, uint32_t accel_mode, ft_batch** out) {
  if (!out) return fail(FT_ERR_INVALID_ARGUMENT, "output handle is null");
  *out = nullptr;
  if (!scenes || n_scenes == 0) return fail(FT_ERR_INVALID_ARGUMENT, "no scenes");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(FT_ERR_GPU_UNAVAILABLE, "no HIP device available");
  ft_batch* b = new ft_batch();
  if (hipGetDevice(&b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    ft_batch_destroy(b);
    return fail(FT_ERR_GPU_UNAVAILABLE, "stream/event creation failed");
  }
  for (uint32_t k = 0; k < n_scenes; ++k) {
    ft_scene* s = nullptr;
    int rc = create_scene(scenes[k], accel_mode, true, &s);
    if (rc) {
      std::string e = "scene " + std::to_string(k) + ": " + g_err;
      ft_batch_destroy(b);
      return fail(rc, e);
    }
    b->scenes.push_back(s);
    b->stack_depth = std::max(b->stack_depth, s->dev.stack_depth);
    b->max_children = std::max<long>(b->max_children, s->host.max_children);
  }
  *out = b;
  return FT_OK;
}

int ft_batch_destroy(ft_batch* b) {
  if (!b) return FT_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (ft_scene* s : b->scenes) ft_scene_destroy(s);
  for (ft_scene* s : b->scenes) s->dev.flops = nullptr;  // the counter belongs to the batch
  for (DevBuf* d : {&b->order, &b->slots, &b->cams, &b->samples, &b->overflow, &b->flops, &b->queue, &b->tcache,
                     &b->plist})
    if (d->ptr) (void)hipFree(d->ptr);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return FT_OK;
}

int ft_batch_render_device(ft_batch* b, const ft_view* views, uint32_t n_views, const ft_render_options* o,
                           const ft_ray_settings* rs, uint32_t flags, void* depth_rgb_dev, void* whitted_rgb_dev) {
  if (!b) return fail(FT_ERR_INVALID_ARGUMENT, "batch is null");
  return batch_render_impl(b, views, n_views, o, rs, flags, (unsigned char*)depth_rgb_dev, (unsigned char*)whitted_rgb_dev);
}

int ft_batch_render(ft_batch* b, const ft_view* views, uint32_t n_views, const ft_render_options* o,
                    const ft_ray_settings* rs, uint32_t flags, uint8_t* depth_rgb_out, uint8_t* whitted_rgb_out) {
  if (!b || !o) return fail(FT_ERR_INVALID_ARGUMENT, "batch or options are null");
  const size_t bytes = (size_t)n_views * o->width * o->height * 3;
  void* dd = nullptr;
  void* dw = nullptr;
  if ((flags & FT_BATCH_DEPTH) && depth_rgb_out) FT_HIP(hipMalloc(&dd, bytes));
  if ((flags & FT_BATCH_WHITTED) && whitted_rgb_out) {
    hipError_t e = hipMalloc(&dw, bytes);
    if (e != hipSuccess) {
      if (dd) (void)hipFree(dd);
      return fail(FT_ERR_OUT_OF_MEMORY, "hipMalloc failed");
    }
  }
  int rc = batch_render_impl(b, views, n_views, o, rs, flags & ~FT_BATCH_GRAPH, (unsigned char*)dd, (unsigned char*)dw);
  if (!rc) rc = batch_finish(b);
  if (!rc && dd && hipMemcpy(depth_rgb_out, dd, bytes, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(FT_ERR_LAUNCH_FAILED, "copy of the depth images failed");
  if (!rc && dw && hipMemcpy(whitted_rgb_out, dw, bytes, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(FT_ERR_LAUNCH_FAILED, "copy of the Whitted images failed");
  if (dd) (void)hipFree(dd);
  if (dw) (void)hipFree(dw);
  return rc;
}

int ft_batch_synchronize(ft_batch* b) {
  if (!b) return fail(FT_ERR_INVALID_ARGUMENT, "batch is null");
  return batch_finish(b);
}

int ft_batch_stats(ft_batch* b, ft_render_stats* out) {
  if (!b || !out) return fail(FT_ERR_INVALID_ARGUMENT, "batch or output is null");
  *out = b->stats;
  return FT_OK;
}

int ft_batch_set_flop_counting(ft_batch* b, int enable) {
  if (!b) return fail(FT_ERR_INVALID_ARGUMENT, "batch is null");
  FT_HIP(hipStreamSynchronize(b->stream));
  drop_graph(b);  // a captured graph would not see the new scene descriptors' counter
  int rc = FT_OK;
  if (enable && (rc = ensure(b->flops, 2 * sizeof(unsigned long long)))) return rc;
  for (ft_scene* s : b->scenes) {
    s->dev.flops = enable ? (unsigned long long*)b->flops.ptr : nullptr;
    s->dev.atlas_samples = enable ? (unsigned long long*)b->flops.ptr + 1 : nullptr;
    if ((rc = publish(s))) return rc;
  }
  b->counting = enable != 0;
  return FT_OK;
}

}  // extern "C"

// ------------------------------------------------------------ differentiable rendering
struct ft_diff {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint32_t n_prims = 0, n_params = 0, n_views = 0, n_nodes = 0;
  ft_render_options opts{};
  float light[3] = {0, 1, 0};
  float normal_step = 1.0e-2f;
  std::vector<int> h_nodes;  // 4 ints per node
  DevBuf nodes, th, bounds, cams, t, s, tt, st, gt, gs, partials, red_a, red_b, grad, m, v, lpart, losses;
  bool have_target = false, have_forward = false;
  int adam_step = 0;
  ft_render_stats stats{};
  DevBuf flops;  // FLOP counter of the counted kernels (counting on), else unallocated
  bool counting = false;
};

namespace {

void diff_tree(int lo, int hi, std::vector<int>& nodes, int& joins) {
  if (hi - lo == 1) {
    nodes.insert(nodes.end(), {lo, -1, -1, -1});
    return;
  }
  const int mid = lo + (hi - lo) / 2;
  diff_tree(lo, mid, nodes, joins);
  const int l = (int)nodes.size() / 4 - 1;
  diff_tree(mid, hi, nodes, joins);
  const int r = (int)nodes.size() / 4 - 1;
  nodes.insert(nodes.end(), {-1, joins++, l, r});
}

void diff_bounds_host(const float* th, uint32_t n, uint32_t n_params, float out[6]) {
  float kmax = 0.0f;
  for (uint32_t j = 4 * n; j < n_params; ++j) kmax = std::fmax(kmax, std::fabs(th[j]));
  const float mg = kmax + 1.0e-3f;
  for (int k = 0; k < 3; ++k) { out[k] = __builtin_huge_valf(); out[3 + k] = -__builtin_huge_valf(); }
  for (uint32_t i = 0; i < n; ++i) {
    const float r = std::fabs(th[4 * i + 3]) + mg;
    for (int k = 0; k < 3; ++k) {
      out[k] = std::fmin(out[k], th[4 * i + k] - r);
      out[3 + k] = std::fmax(out[3 + k], th[4 * i + k] + r);
    }
  }
}

const int kDiffThreads = 128;  // 64 x 2 per block

DiffArgs diff_args(ft_diff* d) {
  DiffArgs a;
  a.nodes = (const int4*)d->nodes.ptr;
  a.th = (const float*)d->th.ptr;
  a.bounds = (const float*)d->bounds.ptr;
  a.cams = (const CamArgs*)d->cams.ptr;
  a.n_nodes = (int)d->n_nodes;
  a.n_prims = (int)d->n_prims;
  a.n_params = (int)d->n_params;
  a.n_views = (int)d->n_views;
  a.tiles_per_view = (((int)d->opts.width + 7) / 8) * (((int)d->opts.height + 7) / 8);
  a.max_steps = (int)d->opts.max_steps;
  a.eps = d->opts.epsilon;
  a.step_scale = d->opts.step_scale;
  a.max_dist = d->opts.max_dist;
  a.normal_step = d->normal_step;
  a.lx = d->light[0];
  a.ly = d->light[1];
  a.lz = d->light[2];
  return a;
}

size_t diff_lds(const ft_diff* d) {
  return (size_t)2 * ((d->n_params + 1) & ~1u) * sizeof(float) + (size_t)d->n_nodes * kDiffThreads * sizeof(float);
}
unsigned diff_blocks(const ft_diff* d) {
  const size_t waves = (size_t)d->n_views * ((((int)d->opts.width + 7) / 8) * (((int)d->opts.height + 7) / 8));
  return (unsigned)((waves + 1) / 2);
}
size_t diff_pixels(const ft_diff* d) { return (size_t)d->n_views * d->opts.width * d->opts.height; }

int diff_forward_enqueue(ft_diff* d) {
  if (d->counting)
    hipLaunchKernelGGL(ft_k_diff_forward_counted, dim3(diff_blocks(d)), dim3(64, 2), diff_lds(d), d->stream,
                       diff_args(d), (float*)d->t.ptr, (float*)d->s.ptr, (unsigned long long*)d->flops.ptr);
  else
    hipLaunchKernelGGL(ft_k_diff_forward, dim3(diff_blocks(d)), dim3(64, 2), diff_lds(d), d->stream, diff_args(d),
                       (float*)d->t.ptr, (float*)d->s.ptr);
  FT_HIP(hipGetLastError());
  d->have_forward = true;
  return FT_OK;
}

// Sums `rows` x cols doubles at `in` over rows into d->grad-like `out` (deterministic order).
int diff_reduce(ft_diff* d, const double* in, int rows, int cols, double* out) {
  const int group = 256;
  const double* src = in;
  double* bufs[2] = {(double*)d->red_a.ptr, (double*)d->red_b.ptr};
  int which = 0;
  while (rows > 1) {
    const int groups = (rows + group - 1) / group;
    double* dst = groups == 1 ? out : bufs[which];
    hipLaunchKernelGGL(ft_k_diff_reduce, dim3(groups, (cols + 255) / 256), dim3(256), 0, d->stream, src, rows, cols,
                       group, dst);
    FT_HIP(hipGetLastError());
    src = dst;
    which ^= 1;
    rows = groups;
  }
  if (src != out) FT_HIP(hipMemcpyAsync(out, src, cols * sizeof(double), hipMemcpyDeviceToDevice, d->stream));
  return FT_OK;
}

int diff_backward_enqueue(ft_diff* d, const float* gt, const float* gs) {
  const unsigned blocks = diff_blocks(d);
  if (d->counting)
    hipLaunchKernelGGL(ft_k_diff_adjoint_counted, dim3(blocks), dim3(64, 2), diff_lds(d), d->stream, diff_args(d),
                       (const float*)d->t.ptr, gt, gs, (double*)d->partials.ptr, (unsigned long long*)d->flops.ptr);
  else
    hipLaunchKernelGGL(ft_k_diff_adjoint, dim3(blocks), dim3(64, 2), diff_lds(d), d->stream, diff_args(d),
                       (const float*)d->t.ptr, gt, gs, (double*)d->partials.ptr);
  FT_HIP(hipGetLastError());
  return diff_reduce(d, (const double*)d->partials.ptr, (int)blocks, (int)d->n_params, (double*)d->grad.ptr);
}

int diff_loss_enqueue(ft_diff* d, float wt, float ws, double* loss_dst) {
  const size_t npix = diff_pixels(d);
  const unsigned blocks = (unsigned)((npix + 255) / 256);
  hipLaunchKernelGGL(ft_k_diff_loss, dim3(blocks), dim3(256), 0, d->stream, (int)d->opts.width, (int)d->opts.height,
                     (int)d->n_views, (const float*)d->t.ptr, (const float*)d->s.ptr, (const float*)d->tt.ptr,
                     (const float*)d->st.ptr, wt, ws, (float*)d->gt.ptr, (float*)d->gs.ptr, (double*)d->lpart.ptr);
  FT_HIP(hipGetLastError());
  return diff_reduce(d, (const double*)d->lpart.ptr, (int)blocks, 1, loss_dst);
}

}  // namespace

extern "C" {

int ft_diff_create(uint32_t n_prims, const ft_view* views, uint32_t n_views, const ft_render_options* o,
                   const float light_dir[3], float normal_step, ft_diff** out) {
  if (!out) return fail(FT_ERR_INVALID_ARGUMENT, "output handle is null");
  *out = nullptr;
  int rc = check_options(o);
  if (rc) return rc;
  if (n_prims < 1 || n_prims > 48) return fail(FT_ERR_INVALID_ARGUMENT, "n_prims must be in [1, 48]");
  if (!views || n_views == 0) return fail(FT_ERR_INVALID_ARGUMENT, "no views");
  if (!light_dir) return fail(FT_ERR_INVALID_ARGUMENT, "light direction is null");
  if (!(normal_step > 0.0f)) return fail(FT_ERR_INVALID_ARGUMENT, "normal_step must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(FT_ERR_GPU_UNAVAILABLE, "no HIP device available");
  ft_diff* d = new ft_diff();
  d->n_prims = n_prims;
  d->n_params = 5 * n_prims - 1;
  d->n_views = n_views;
  d->opts = *o;
  d->normal_step = normal_step;
  const float ll = std::sqrt(light_dir[0] * light_dir[0] + light_dir[1] * light_dir[1] + light_dir[2] * light_dir[2]);
  for (int k = 0; k < 3; ++k) d->light[k] = ll > 0.0f ? light_dir[k] / ll : (k == 1 ? 1.0f : 0.0f);
  int joins = 0;
  diff_tree(0, (int)n_prims, d->h_nodes, joins);
  d->n_nodes = (uint32_t)(d->h_nodes.size() / 4);
  auto bail = [&](int code) {
    std::string e = g_err;
    ft_diff_destroy(d);
    return fail(code, e);
  };
  if (hipGetDevice(&d->device) != hipSuccess || hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&d->ev0) != hipSuccess || hipEventCreate(&d->ev1) != hipSuccess) {
    g_err = "stream/event creation failed";
    return bail(FT_ERR_GPU_UNAVAILABLE);
  }
  std::vector<CamArgs> cams(n_views);
  for (uint32_t v = 0; v < n_views; ++v) {
    fth::CamBasis cb = fth::camera_basis(views[v].origin, views[v].target, views[v].fov_y_degrees);
    CamArgs& c = cams[v];
    c.ox = cb.origin[0]; c.oy = cb.origin[1]; c.oz = cb.origin[2];
    c.fx = cb.forward[0]; c.fy = cb.forward[1]; c.fz = cb.forward[2];
    c.rx = cb.right[0]; c.ry = cb.right[1]; c.rz = cb.right[2];
    c.ux = cb.up[0]; c.uy = cb.up[1]; c.uz = cb.up[2];
    c.tan_fov = cb.tan_fov;
    c.aspect = (float)o->width / (float)o->height;
    c.width = (int)o->width;
    c.height = (int)o->height;
    c.aa = 1;
    c.grid = 1;
    c.tile_w = 8;
    c.y0 = 0;
    c.rows = c.height;
  }
  const size_t npix = diff_pixels(d);
  const unsigned blocks = diff_blocks(d);
  const size_t lblocks = (npix + 255) / 256;
  const size_t red_rows = std::max<size_t>((blocks + 255) / 256, (lblocks + 255) / 256) + 1;
  std::vector<float> zeros(d->n_params, 0.0f);
  if ((rc = upload(d->nodes, d->h_nodes)) || (rc = upload(d->cams, cams)) ||
      (rc = ensure(d->th, d->n_params * sizeof(float))) || (rc = ensure(d->bounds, 6 * sizeof(float))) ||
      (rc = ensure(d->t, npix * sizeof(float))) || (rc = ensure(d->s, npix * sizeof(float))) ||
      (rc = ensure(d->tt, npix * sizeof(float))) || (rc = ensure(d->st, npix * sizeof(float))) ||
      (rc = ensure(d->gt, npix * sizeof(float))) || (rc = ensure(d->gs, npix * sizeof(float))) ||
      (rc = ensure(d->partials, (size_t)blocks * d->n_params * sizeof(double))) ||
      (rc = ensure(d->red_a, red_rows * std::max<size_t>(d->n_params, 1) * sizeof(double))) ||
      (rc = ensure(d->red_b, red_rows * std::max<size_t>(d->n_params, 1) * sizeof(double))) ||
      (rc = ensure(d->grad, d->n_params * sizeof(double))) || (rc = ensure(d->m, d->n_params * sizeof(double))) ||
      (rc = ensure(d->v, d->n_params * sizeof(double))) || (rc = ensure(d->lpart, lblocks * sizeof(double))) ||
      (rc = ensure(d->losses, sizeof(double))))
    return bail(rc);
  if (diff_lds(d) > 64 * 1024) {
    g_err = "scene too large for the adjoint's LDS budget";
    return bail(FT_ERR_INVALID_ARGUMENT);
  }
  *out = d;
  return FT_OK;
}

int ft_diff_destroy(ft_diff* d) {
  if (!d) return FT_OK;
  (void)hipSetDevice(d->device);
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  for (DevBuf* b : {&d->nodes, &d->th, &d->bounds, &d->cams, &d->t, &d->s, &d->tt, &d->st, &d->gt, &d->gs,
                    &d->partials, &d->red_a, &d->red_b, &d->grad, &d->m, &d->v, &d->lpart, &d->losses, &d->flops})
    if (b->ptr) (void)hipFree(b->ptr);
  if (d->ev0) (void)hipEventDestroy(d->ev0);
  if (d->ev1) (void)hipEventDestroy(d->ev1);
  if (d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
  return FT_OK;
}

int ft_diff_param_count(ft_diff* d, uint32_t* n) {
  if (!d || !n) return fail(FT_ERR_INVALID_ARGUMENT, "diff or output is null");
  *n = d->n_params;
  return FT_OK;
}

int ft_diff_set_params(ft_diff* d, const float* params) {
  if (!d || !params) return fail(FT_ERR_INVALID_ARGUMENT, "diff or params are null");
  float b[6];
  diff_bounds_host(params, d->n_prims, d->n_params, b);
  FT_HIP(hipStreamSynchronize(d->stream));
  FT_HIP(hipMemcpy(d->th.ptr, params, d->n_params * sizeof(float), hipMemcpyHostToDevice));
  FT_HIP(hipMemcpy(d->bounds.ptr, b, sizeof(b), hipMemcpyHostToDevice));
  d->have_forward = false;
  return FT_OK;
}

int ft_diff_get_params(ft_diff* d, float* params) {
  if (!d || !params) return fail(FT_ERR_INVALID_ARGUMENT, "diff or params are null");
  FT_HIP(hipStreamSynchronize(d->stream));
  FT_HIP(hipMemcpy(params, d->th.ptr, d->n_params * sizeof(float), hipMemcpyDeviceToHost));
  return FT_OK;
}

int ft_diff_forward_device(ft_diff* d, void* t_dev, void* s_dev) {
  if (!d) return fail(FT_ERR_INVALID_ARGUMENT, "diff is null");
  FT_HIP(hipSetDevice(d->device));
  FT_HIP(hipEventRecord(d->ev0, d->stream));
  int rc = diff_forward_enqueue(d);
  if (rc) return rc;
  FT_HIP(hipEventRecord(d->ev1, d->stream));
  const size_t bytes = diff_pixels(d) * sizeof(float);
  if (t_dev) FT_HIP(hipMemcpyAsync(t_dev, d->t.ptr, bytes, hipMemcpyDeviceToDevice, d->stream));
  if (s_dev) FT_HIP(hipMemcpyAsync(s_dev, d->s.ptr, bytes, hipMemcpyDeviceToDevice, d->stream));
  d->stats.primary_rays = diff_pixels(d);
  d->stats.kernel_launches = 1;
  return FT_OK;
}

int ft_diff_synchronize(ft_diff* d) {
  if (!d) return fail(FT_ERR_INVALID_ARGUMENT, "diff is null");
  FT_HIP(hipStreamSynchronize(d->stream));
  return FT_OK;
}

int ft_diff_forward(ft_diff* d, float* t_out, float* s_out) {
  if (!d) return fail(FT_ERR_INVALID_ARGUMENT, "diff is null");
  int rc = ft_diff_forward_device(d, nullptr, nullptr);
  if (rc) return rc;
  FT_HIP(hipStreamSynchronize(d->stream));
  float ms = 0.0f;
  FT_HIP(hipEventElapsedTime(&ms, d->ev0, d->ev1));
  d->stats.last_kernel_ms = ms;
  const size_t bytes = diff_pixels(d) * sizeof(float);
  if (t_out) FT_HIP(hipMemcpy(t_out, d->t.ptr, bytes, hipMemcpyDeviceToHost));
  if (s_out) FT_HIP(hipMemcpy(s_out, d->s.ptr, bytes, hipMemcpyDeviceToHost));
  return FT_OK;
}

int ft_diff_backward_device(ft_diff* d, const void* gt, const void* gs, double* grad_out) {
  if (!d || !grad_out) return fail(FT_ERR_INVALID_ARGUMENT, "diff or output is null");
  if (!d->have_forward) return fail(FT_ERR_INVALID_ARGUMENT, "backward needs a forward at the current parameters");
  FT_HIP(hipSetDevice(d->device));
  FT_HIP(hipEventRecord(d->ev0, d->stream));
  int rc = diff_backward_enqueue(d, (const float*)gt, (const float*)gs);
  if (rc) return rc;
  FT_HIP(hipEventRecord(d->ev1, d->stream));
  FT_HIP(hipStreamSynchronize(d->stream));
  float ms = 0.0f;
  FT_HIP(hipEventElapsedTime(&ms, d->ev0, d->ev1));
  d->stats.last_kernel_ms = ms;
  FT_HIP(hipMemcpy(grad_out, d->grad.ptr, d->n_params * sizeof(double), hipMemcpyDeviceToHost));
  return FT_OK;
}

int ft_diff_set_target(ft_diff* d, const float* t_target, const float* s_target) {
  if (!d || !t_target || !s_target) return fail(FT_ERR_INVALID_ARGUMENT, "diff or targets are null");
  const size_t bytes = diff_pixels(d) * sizeof(float);
  FT_HIP(hipStreamSynchronize(d->stream));
  FT_HIP(hipMemcpy(d->tt.ptr, t_target, bytes, hipMemcpyHostToDevice));
  FT_HIP(hipMemcpy(d->st.ptr, s_target, bytes, hipMemcpyHostToDevice));
  d->have_target = true;
  return FT_OK;
}

int ft_diff_loss_grad(ft_diff* d, float wt, float ws, double* loss_out, double* grad_out) {
  if (!d) return fail(FT_ERR_INVALID_ARGUMENT, "diff is null");
  if (!d->have_target) return fail(FT_ERR_INVALID_ARGUMENT, "no target images");
  FT_HIP(hipSetDevice(d->device));
  int rc;
  if ((rc = diff_forward_enqueue(d)) || (rc = diff_loss_enqueue(d, wt, ws, (double*)d->losses.ptr)) ||
      (rc = diff_backward_enqueue(d, (const float*)d->gt.ptr, (const float*)d->gs.ptr)))
    return rc;
  FT_HIP(hipStreamSynchronize(d->stream));
  if (loss_out) FT_HIP(hipMemcpy(loss_out, d->losses.ptr, sizeof(double), hipMemcpyDeviceToHost));
  if (grad_out) FT_HIP(hipMemcpy(grad_out, d->grad.ptr, d->n_params * sizeof(double), hipMemcpyDeviceToHost));
  return FT_OK;
}

int ft_diff_fit(ft_diff* d, uint32_t steps, float lr, float wt, float ws, double* losses_out) {
  if (!d) return fail(FT_ERR_INVALID_ARGUMENT, "diff is null");
  if (!d->have_target) return fail(FT_ERR_INVALID_ARGUMENT, "no target images");
  if (steps == 0) return FT_OK;
  FT_HIP(hipSetDevice(d->device));
  int rc = ensure(d->losses, steps * sizeof(double));
  if (rc) return rc;
  FT_HIP(hipStreamSynchronize(d->stream));
  if (d->adam_step == 0) {
    FT_HIP(hipMemset(d->m.ptr, 0, d->n_params * sizeof(double)));
    FT_HIP(hipMemset(d->v.ptr, 0, d->n_params * sizeof(double)));
  }
  FT_HIP(hipEventRecord(d->ev0, d->stream));
  for (uint32_t k = 0; k < steps; ++k) {
    if ((rc = diff_forward_enqueue(d)) || (rc = diff_loss_enqueue(d, wt, ws, (double*)d->losses.ptr + k)) ||
        (rc = diff_backward_enqueue(d, (const float*)d->gt.ptr, (const float*)d->gs.ptr)))
      return rc;
    ++d->adam_step;
    hipLaunchKernelGGL(ft_k_diff_adam, dim3(1), dim3(256), 0, d->stream, (float*)d->th.ptr, (const double*)d->grad.ptr,
                       (double*)d->m.ptr, (double*)d->v.ptr, (int)d->n_params, (int)d->n_prims, d->adam_step, lr, 0.9f,
                       0.999f, 1.0e-8f, (float*)d->bounds.ptr);
    FT_HIP(hipGetLastError());
  }
  FT_HIP(hipEventRecord(d->ev1, d->stream));
  FT_HIP(hipStreamSynchronize(d->stream));
  d->have_forward = false;
  float ms = 0.0f;
  FT_HIP(hipEventElapsedTime(&ms, d->ev0, d->ev1));
  d->stats.last_kernel_ms = ms;
  d->stats.kernel_launches = (uint64_t)steps * 7;
  d->stats.primary_rays = (uint64_t)steps * diff_pixels(d);
  if (losses_out) FT_HIP(hipMemcpy(losses_out, d->losses.ptr, steps * sizeof(double), hipMemcpyDeviceToHost));
  return FT_OK;
}

int ft_diff_stats(ft_diff* d, ft_render_stats* out) {
  if (!d || !out) return fail(FT_ERR_INVALID_ARGUMENT, "diff or output is null");
  *out = d->stats;
  out->sdf_flops = 0;
  if (d->counting) {
    FT_HIP(hipSetDevice(d->device));
    FT_HIP(hipStreamSynchronize(d->stream));
    unsigned long long f[2] = {0, 0};
    FT_HIP(hipMemcpy(f, d->flops.ptr, sizeof(f), hipMemcpyDeviceToHost));
    out->sdf_flops = f[0];
  }
  return FT_OK;
}

int ft_diff_set_flop_counting(ft_diff* d, int enable) {
  if (!d) return fail(FT_ERR_INVALID_ARGUMENT, "diff is null");
  FT_HIP(hipSetDevice(d->device));
  FT_HIP(hipStreamSynchronize(d->stream));
  if (enable) {
    int rc = ensure(d->flops, 2 * sizeof(unsigned long long));
    if (rc) return rc;
    hipLaunchKernelGGL(ft_k_clear_u64, dim3(1), dim3(64), 0, d->stream, (unsigned long long*)d->flops.ptr, 2);
    FT_HIP(hipGetLastError());
  }
  d->counting = enable != 0;
  return FT_OK;
}

}  // extern "C"
