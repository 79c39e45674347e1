/*
 * ft_abi.h — C ABI of the MI355X sphere-tracing backend (libft_hip.so).
 *
 * The reference renderer (ForgedThoughts, Rust) selects its backend through
 * `AccelMode` and dispatches to three public entry points:
 *   render_depth_png_with_accel            crates/forgedthoughts/src/renderer.rs:1139-1156
 *   render_preview_progressive_with_accel  crates/forgedthoughts/src/renderer.rs:1238-1277
 *   render_ray_progressive_with_accel      crates/forgedthoughts/src/renderer.rs:1201-1236
 * Those take an `EvalState`, compile it into a `CompiledScene` (renderer.rs:480-488,
 * `compile_scene` renderer.rs:1384-1401) and build an `Accelerator` (renderer.rs:609-623).
 * The per-point `Accelerator` trait is far too fine-grained to cross an FFI, so this
 * boundary sits one level up: the caller hands over the flattened compiled scene
 * (`ft_scene_ir`, the data of `CompiledScene` + `RenderSetup`, renderer.rs:480-501),
 * the library builds its device-side acceleration data once (`ft_scene_create`), and
 * renders whole images (`ft_render_depth`, `ft_render_whitted`).
 *
 * Conventions
 *   - every entry point returns an int status (FT_OK == 0); on failure a message is
 *     available from ft_last_error() (thread-local), mirroring RenderError
 *     (renderer.rs:158-168) plus device-side failures;
 *   - host output buffers are caller-owned; device memory is library-owned;
 *   - each scene handle owns one HIP stream; a handle is not thread-safe, distinct
 *     handles are;
 *   - `ft_render_options` / `ft_ray_settings` mirror RenderOptions (renderer.rs:31-41)
 *     and RaySettings (renderer.rs:128-134) field for field.
 */
#ifndef FT_ABI_H
#define FT_ABI_H

#if defined(__HIPCC_RTC__)
/* runtime-compiled device code: hiprtc provides the fixed-width integer types */
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::uint64_t uint64_t;
typedef __hip_internal::uint8_t uint8_t;
#else
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define FT_ABI_VERSION 6u

/* ---------------------------------------------------------------- status */
enum {
  FT_OK = 0,
  FT_ERR_MISSING_SCENE_ROOT = 1,   /* RenderError::MissingSceneRoot */
  FT_ERR_UNSUPPORTED_OBJECT = 2,   /* RenderError::UnsupportedObjectType */
  FT_ERR_EXPECTED_OBJECT = 3,      /* RenderError::ExpectedObject */
  FT_ERR_IMAGE = 4,                /* RenderError::Image */
  FT_ERR_INVALID_ARGUMENT = 5,
  FT_ERR_GPU_UNAVAILABLE = 6,
  FT_ERR_COMPILE_FAILED = 7,
  FT_ERR_OUT_OF_MEMORY = 8,
  FT_ERR_LAUNCH_FAILED = 9,
  FT_ERR_CALLBACK_ABORT = 10
};

/* --------------------------------------------------------- accel modes */
/* AccelMode (renderer.rs:58-63). */
enum { FT_ACCEL_NAIVE = 0, FT_ACCEL_BVH = 1, FT_ACCEL_BRICKS = 2 };

/* ----------------------------------------------------------- scene IR  */
/* SdfNode variants (renderer.rs:170-355).  Nodes are stored children-first
 * (post-order); `lhs`/`rhs` index into the node array. */
enum {
  FT_NODE_SPHERE = 0,
  FT_NODE_BOX = 1,
  FT_NODE_CYLINDER = 2,
  FT_NODE_TORUS = 3,
  FT_NODE_EXTRUDE_POLYGON = 4,
  FT_NODE_CUSTOM = 5,
  FT_NODE_DOMAIN_MODIFIER = 6,
  FT_NODE_DISTANCE_POST_MODIFIER = 7,
  FT_NODE_NOISE = 8,
  FT_NODE_UNION = 9,
  FT_NODE_INTERSECT = 10,
  FT_NODE_SUBTRACT = 11,
  FT_NODE_UNION_ROUND = 12,
  FT_NODE_UNION_CHAMFER = 13,
  FT_NODE_UNION_COLUMNS = 14,
  FT_NODE_UNION_STAIRS = 15,
  FT_NODE_UNION_SOFT = 16,
  FT_NODE_INTERSECT_ROUND = 17,
  FT_NODE_INTERSECT_CHAMFER = 18,
  FT_NODE_INTERSECT_COLUMNS = 19,
  FT_NODE_INTERSECT_STAIRS = 20,
  FT_NODE_DIFF_ROUND = 21,
  FT_NODE_DIFF_CHAMFER = 22,
  FT_NODE_DIFF_COLUMNS = 23,
  FT_NODE_DIFF_STAIRS = 24,
  FT_NODE_PIPE = 25,
  FT_NODE_ENGRAVE = 26,
  FT_NODE_GROOVE = 27,
  FT_NODE_TONGUE = 28,
  FT_NODE_SLICE = 29,
  FT_NODE_SMOOTH = 30,
  FT_NODE_KIND_COUNT = 31
};

/*
 * One SdfNode.  `param` layout per kind (values exactly as compile_sdf stores them,
 * renderer.rs:1414-1742, e.g. Box stores the HALF size):
 *   SPHERE            radius, shell
 *   BOX               half_x, half_y, half_z, round, shell
 *   CYLINDER          radius, half_height, round, shell
 *   TORUS             major_radius, minor_radius
 *   EXTRUDE_POLYGON   sides, radius, half_height, round, shell
 *   CUSTOM            bounds_half_x, bounds_half_y, bounds_half_z
 *   NOISE             octaves, scale, lacunarity
 *   *_ROUND/CHAMFER/SOFT, PIPE, ENGRAVE      r
 *   *_COLUMNS/STAIRS  r, n
 *   GROOVE, TONGUE    ra, rb
 *   SLICE             axis (0/1/2), min, max
 *   SMOOTH            k
 * DOMAIN_MODIFIER / DISTANCE_POST_MODIFIER / NOISE carry the bounds computed at scene
 * compile time (renderer.rs:1810-1948) in bounds_min/bounds_max; other kinds ignore them.
 * CUSTOM uses fn_domain / fn_distance / fn_post (the sdf definition's `domain`,
 * `distance`, `distance_post`); DOMAIN_MODIFIER uses fn_domain; DISTANCE_POST_MODIFIER
 * uses fn_post.  Unused function slots are -1.
 */
typedef struct ft_node {
  uint32_t kind;
  int32_t lhs;
  int32_t rhs;
  uint32_t object_id;
  uint32_t material_id;
  int32_t fn_domain;
  int32_t fn_distance;
  int32_t fn_post;
  float center[3];
  float rot_deg[3];
  float param[6];
  float bounds_min[3];
  float bounds_max[3];
} ft_node;

/* ------------------------------------------------------- expression IR */
/*
 * Forge function bodies (ast.rs:126-171 subset; no control flow exists in the
 * language) lowered to a typed expression DAG.  Helper calls are inlined and captured
 * `let` values are folded into constants by the front end.  Each function owns the
 * contiguous range [first, first+count) of `exprs`; operands a/b/c index into that
 * range (relative to `first`) and always precede the node; the last node is the
 * result.  `type` is FT_TYPE_SCALAR or FT_TYPE_VEC3.
 */
enum { FT_TYPE_SCALAR = 0, FT_TYPE_VEC3 = 1 };

enum {
  FT_EXPR_CONST = 0,      /* value */
  FT_EXPR_PARAM = 1,      /* a = parameter index */
  FT_EXPR_VEC3 = 2,       /* vec3(a, b, c) */
  FT_EXPR_GET_X = 3,      /* a.x */
  FT_EXPR_GET_Y = 4,
  FT_EXPR_GET_Z = 5,
  FT_EXPR_NEG = 6,
  FT_EXPR_ADD = 7,        /* binary ops broadcast scalar over vec3 (eval.rs:1228-1256) */
  FT_EXPR_SUB = 8,
  FT_EXPR_MUL = 9,
  FT_EXPR_DIV = 10,
  FT_EXPR_ABS = 11,       /* unary maps apply per component (eval.rs:4230-4242) */
  FT_EXPR_FLOOR = 12,
  FT_EXPR_CEIL = 13,
  FT_EXPR_FRACT = 14,
  FT_EXPR_SQRT = 15,
  FT_EXPR_SIN = 16,
  FT_EXPR_COS = 17,
  FT_EXPR_SATURATE = 18,
  FT_EXPR_MIN = 19,
  FT_EXPR_MAX = 20,
  FT_EXPR_POW = 21,
  FT_EXPR_CLAMP = 22,     /* clamp(x, a, b) with swapped bounds allowed */
  FT_EXPR_MIX = 23,
  FT_EXPR_STEP = 24,      /* step(edge, x) */
  FT_EXPR_SMOOTHSTEP = 25,
  FT_EXPR_DOT = 26,
  FT_EXPR_LENGTH = 27,
  FT_EXPR_NORMALIZE = 28,
  FT_EXPR_ROTATE_X = 29,  /* rotate_x(v, degrees) */
  FT_EXPR_ROTATE_Y = 30,
  FT_EXPR_ROTATE_Z = 31,
  FT_EXPR_SPHERE_DIST = 32,   /* Sphere.distance(p, r)      eval.rs:1899-1909 */
  FT_EXPR_BOX_DIST = 33,      /* Box.distance(p, half)      eval.rs:1881-1898 */
  FT_EXPR_CYLINDER_DIST = 34, /* Cylinder.distance(p, r, h) eval.rs:1910-1929 */
  FT_EXPR_TORUS_DIST = 35,    /* Torus.distance(p, R, r)    eval.rs:1930-1945 */
  FT_EXPR_VALUE_NOISE = 36,   /* value_noise_3d(p)          eval.rs:4292-4317 */
  FT_EXPR_FBM = 37,           /* fbm_3d(p, octaves, scale, lacunarity)  eval.rs:4319-4336 */
  FT_EXPR_OP_COUNT = 38
};

typedef struct ft_expr {
  uint32_t op;
  uint32_t type;
  int32_t a;
  int32_t b;
  int32_t c;
  int32_t d;
  float value;
  float pad;
} ft_expr;

typedef struct ft_function {
  uint32_t first;
  uint32_t count;
  uint32_t n_params;
  uint32_t result_type;
} ft_function;

/* ---------------------------------------------------------- materials */
/* Material / MaterialParams (materials/mod.rs:67-87, 214-221). */
enum { FT_MODEL_STANDARD = 0, FT_MODEL_LAMBERT = 1, FT_MODEL_METAL = 2, FT_MODEL_DIELECTRIC = 3 };
enum { FT_PATTERN_NONE = 0, FT_PATTERN_CHECKER3D = 1 };

typedef struct ft_material {
  uint32_t model;
  uint32_t thin_walled;
  uint32_t has_medium;
  uint32_t pattern;
  float color[3];
  float roughness;
  float metallic;
  float specular;
  float specular_weight;
  float ior;
  float specular_color[3];
  float clearcoat;
  float clearcoat_roughness;
  float transmission;
  float emission_strength;
  float medium_ior;
  float emission_color[3];
  float medium_density;
  float medium_absorption[3];
  float pattern_scale;
  float pattern_color_a[3];
  int32_t dynamic_material_id;   /* index into the sorted material-definition names when the
                                    definition has per-hit hook functions, else -1; Whitted
                                    renders fail with FT_ERR_UNSUPPORTED_OBJECT unless the
                                    hooks are listed in ft_scene_ir.material_hooks */
  float pattern_color_b[3];
  int32_t dynamic_override_id;   /* interned override object of a dynamic material, or -1 */
} ft_material;

/* Per-hit material hooks (resolve_leaf_material renderer.rs:7164-7536) of one material from
 * a definition: Forge functions of the shading context, -1 where the definition has none
 * (or, for a number hook, returns a vector).  The Metal arm reads color, roughness and the
 * emission hooks; the Dielectric and Lambert arms also ior and thin_walled.  Standard /
 * Layered definitions (renderer.rs:7259-7470) read every entry: their coat block replaces
 * the ior, specular colour, clearcoat and clearcoat roughness, so for them `ior` holds the
 * coat_ior hook, `specular_color` coat_color, `clearcoat_roughness` coat_roughness and
 * `clearcoat` coat_weight, the clearcoat being clamp(weight * coat_mask) with the static
 * weight where `clearcoat` is -1.  bump and normal hooks of any model reshape the normal the
 * direct lighting uses.  Every hook function takes FT_HOOK_PARAMS parameters, the shading
 * context (make_shading_context renderer.rs:7875-7911) in the order position,
 * local_position, normal, view_dir, front_face, object_id, material_id, then wo, wi,
 * current_ior, u1, u2, u3 (zero for these hooks); scalars are splatted. */
#define FT_HOOK_PARAMS 13u
typedef struct ft_material_hooks {
  uint32_t material;          /* index into ft_scene_ir.materials */
  int32_t color;
  int32_t roughness;
  int32_t ior;
  int32_t thin_walled;
  int32_t emission_color;
  int32_t emission_strength;
  int32_t metallic;           /* Standard / Layered definitions only, from here on */
  int32_t specular;
  int32_t specular_weight;
  int32_t transmission;
  int32_t specular_color;
  int32_t clearcoat_roughness;
  int32_t clearcoat;
  int32_t coat_mask;
  int32_t bump;               /* any model: the shading normal of build_bsdf_context */
  int32_t normal;             /* (renderer.rs:4397-4449, 7749-7827), a number and a vector */
  int32_t medium_ior;         /* the fields of the Medium object a medium hook returns */
  int32_t medium_absorption;  /* (medium_from_value renderer.rs:6834-6845): all three or none */
  int32_t medium_density;
  int32_t eval;               /* the BSDF value of shade_color (eval_bsdf renderer.rs:4905-4920):
                                 context normal = shading normal, view_dir = wo, plus wo, wi,
                                 current_ior; read while the material keeps its dynamic id */
  int32_t pad;
} ft_material_hooks;

/* -------------------------------------------------------------- lights */
/* PointLight / EnvLight / SphereLight (render_api.rs:171-289), radiance already
 * multiplied by intensity (renderer.rs:6786-6800). */
enum { FT_LIGHT_POINT = 0, FT_LIGHT_ENV = 1, FT_LIGHT_SPHERE = 2 };

typedef struct ft_light {
  uint32_t kind;
  uint32_t samples;
  float radius;
  float pad;
  float position[3];
  float pad1;
  float radiance[3];
  float pad2;
} ft_light;

/* `camera` binding as read by parse_camera (renderer.rs:6263-6286); absent fields fall
 * back to the scene centre and RenderOptions at render time. */
typedef struct ft_camera {
  uint32_t present;
  uint32_t has_origin;
  uint32_t has_target;
  uint32_t has_fov;
  float origin[3];
  float fov_y_degrees;
  float target[3];
  float pad;
} ft_camera;

typedef struct ft_scene_ir {
  uint32_t abi_version;
  uint32_t root;              /* index of the root node (normally n_nodes - 1) */
  uint32_t n_nodes;
  uint32_t n_materials;
  uint32_t n_objects;         /* object_transforms has 6 floats per object id */
  uint32_t n_lights;
  uint32_t n_functions;
  uint32_t n_exprs;
  int32_t env_fn;             /* environment `color(dir)` function or -1 */
  uint32_t pad;
  const ft_node* nodes;
  const ft_material* materials;
  const float* object_transforms;
  const ft_light* lights;
  const ft_function* functions;
  const ft_expr* exprs;
  ft_camera camera;
  uint32_t n_material_hooks;
  uint32_t pad2;
  const ft_material_hooks* material_hooks;  /* may be null when n_material_hooks is 0 */
} ft_scene_ir;

/* --------------------------------------------------------- render args */
typedef struct ft_render_options {
  uint32_t width;
  uint32_t height;
  uint32_t max_steps;
  float max_dist;
  float epsilon;
  float step_scale;
  float camera_z;
  float fov_y_degrees;
} ft_render_options;

/* RayDebugAov (renderer.rs:146-155); -1 = none. */
enum {
  FT_AOV_NONE = -1,
  FT_AOV_DEPTH = 0,
  FT_AOV_NORMAL = 1,
  FT_AOV_MATERIAL_ID = 2,
  FT_AOV_IOR = 3,
  FT_AOV_TRANSMISSION = 4,
  FT_AOV_FRESNEL = 5,
  FT_AOV_HIT_T = 6
};

typedef struct ft_ray_settings {
  uint32_t max_depth;
  uint32_t tile_size;
  uint32_t aa_samples;
  int32_t debug_aov;
} ft_ray_settings;

/* ------------------------------------------------------------ entry points */
typedef struct ft_scene ft_scene;

/* Returns the thread-local message of the last failed call ("" if none). */
const char* ft_last_error(void);

/* Number of usable GPUs (0 when no device is visible). */
int ft_device_count(int* count);

/* Builds device-side scene data for `accel_mode` (AccelMode::from_scene,
 * renderer.rs:1024-1129).  The IR is copied; the caller may free it afterwards. */
int ft_scene_create(const ft_scene_ir* ir, uint32_t accel_mode, ft_scene** out);
int ft_scene_destroy(ft_scene* scene);

/* Depth preview (render_depth_tile renderer.rs:3931-3971 via render_preview_tiled).
 * rgb_out: width*height*3 bytes, row-major, host memory.  t_out (nullable): width*height
 * floats, refined hit distance of the first AA sample (or -1 on a miss). */
int ft_render_depth(ft_scene* scene, const ft_render_options* options, uint32_t aa_samples,
                    uint8_t* rgb_out, float* t_out);

/* Whitted trace (render_ray_tile renderer.rs:3874-3929 + renderer/ray.rs).
 * hdr_out (nullable): width*height*3 floats, AA-averaged radiance before tone mapping. */
int ft_render_whitted(ft_scene* scene, const ft_render_options* options,
                      const ft_ray_settings* settings, uint8_t* rgb_out, float* hdr_out);

/* Progressive renders with a progress callback (RayProgress and the on_tile callback of
 * render_ray_progressive_with_accel renderer.rs:1201-1236, 3635-3724, and of
 * render_preview_progressive_with_accel renderer.rs:1238-1277, 3751-3834).  The image is
 * rendered in rows of tile_size x tile_size tiles (the reference's tile_jobs order,
 * renderer.rs:3849-3862); the callback runs on the calling thread once per finished tile
 * with the image so far (unfinished tiles black).  A non-zero return stops the render:
 * the call returns FT_ERR_CALLBACK_ABORT (the reference returns the callback's error).
 * rgb_out receives the finished image.  Synchronous. */
typedef struct ft_progress {
  uint32_t tiles_done;
  uint32_t tiles_total;
  uint64_t elapsed_ms;
  uint32_t tile_x, tile_y, tile_w, tile_h;  /* the tile just finished */
  uint32_t width, height;
  const uint8_t* image;                     /* width * height * 3 bytes */
} ft_progress;
typedef int (*ft_progress_fn)(const ft_progress* progress, void* user);
int ft_render_depth_progressive(ft_scene* scene, const ft_render_options* options, uint32_t aa_samples,
                                uint32_t tile_size, ft_progress_fn on_tile, void* user, uint8_t* rgb_out);
/* tile_size and aa_samples from the settings (RaySettings.tile_size / aa_samples). */
int ft_render_whitted_progressive(ft_scene* scene, const ft_render_options* options,
                                  const ft_ray_settings* settings, ft_progress_fn on_tile, void* user,
                                  uint8_t* rgb_out);

/* Device-resident variants: outputs are device pointers, the call enqueues on the
 * scene's stream and returns without synchronising (ft_scene_synchronize waits). */
int ft_render_depth_device(ft_scene* scene, const ft_render_options* options,
                           uint32_t aa_samples, void* rgb_dev, void* t_dev);
int ft_render_whitted_device(ft_scene* scene, const ft_render_options* options,
                             const ft_ray_settings* settings, void* rgb_dev, void* hdr_dev);
/* ft_render_whitted_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The per-sample radiance and the
 * ray-stack overflow flag live in a fixed 64 MiB device array of the library (one per GPU,
 * held until the process ends), so aa_samples * width * height must not exceed
 * FT_NOALLOC_MAX_SAMPLES (e.g. 1920x1080 at 2 spp, 512x512 at 16 spp); larger requests fail
 * with FT_ERR_INVALID_ARGUMENT.  The scene's Whitted kernel must already be built (render
 * once with ft_render_whitted or ft_render_whitted_device first: building a specialised
 * kernel loads a code object) and FLOP counting must be off.  Every call on one GPU uses
 * that GPU's array: calls that overlap on one GPU (renders of any scenes in flight at the
 * same time) race and must be serialised by the caller.  Enqueues and returns;
 * ft_scene_synchronize waits and reports a ray-stack overflow. */
#define FT_NOALLOC_MAX_SAMPLES 5592384u
int ft_render_whitted_device_noalloc(ft_scene* scene, const ft_render_options* options,
                                     const ft_ray_settings* settings, void* rgb_dev, void* hdr_dev);
int ft_scene_synchronize(ft_scene* scene);

/* The HIP stream (hipStream_t) the scene launches on, for callers that time it. */
int ft_scene_stream(ft_scene* scene, void** stream_out);

/* Per-kernel statistics of the last render on this scene. */
typedef struct ft_render_stats {
  uint64_t primary_rays;
  uint64_t kernel_launches;
  float last_kernel_ms;   /* hipEvent time of the last render's kernels */
  uint32_t flags;         /* FT_STATS_* */
  uint64_t sdf_flops;     /* FLOPs of SDF evaluation counted by the last render (counting on) */
  double compile_ms;      /* time spent compiling scene-specialised kernels so far */
  uint64_t atlas_samples; /* distance-atlas samples taken by the last render (counting on) */
  uint64_t atlas_bytes;   /* device bytes of the scene's distance atlas (0: none) */
  uint64_t atlas_bricks;  /* bricks in the atlas */
  uint32_t ray_stack;     /* ray-stack entries of the last Whitted kernel (ft_k_whitted_N), 0: none yet */
} ft_render_stats;
#define FT_STATS_EMITTED_SCENE 1u     /* scene distance runs as specialised straight-line code */
#define FT_STATS_EMITTED_FUNCTIONS 2u /* Forge functions run as specialised code */
int ft_scene_stats(ft_scene* scene, ft_render_stats* out);

/* ------------------------------------------------------------------ batches
 * Many scenes, each seen from several views, rendered together.  No counterpart in the
 * reference (it renders one scene per call, renderer.rs:1139-1277); a batch is what a
 * caller of those entry points would otherwise issue as one call per (scene, view).
 * Every view of a call has the options' image size and the settings' sample count; its
 * camera replaces the scene's `camera` binding (parse_camera renderer.rs:6263-6286).
 * Scenes of a batch run on the node-program interpreter (no per-scene compilation). */
typedef struct ft_batch ft_batch;

typedef struct ft_view {
  uint32_t scene;            /* index into the batch's scenes */
  float fov_y_degrees;
  float origin[3];
  float target[3];
} ft_view;

#define FT_BATCH_DEPTH 1u     /* depth preview per view (render_preview_progressive_with_accel) */
#define FT_BATCH_WHITTED 2u   /* Whitted trace per view (render_ray_progressive_with_accel) */
#define FT_BATCH_SEPARATE 4u  /* issue the per-view launches of the single-scene path instead of
                                 the batched launch (for comparison) */
#define FT_BATCH_GRAPH 8u     /* capture the launch sequence in a HIP graph; a later call with the
                                 same views, options, flags and outputs replays it */
#define FT_BATCH_TILED 16u    /* depth pass as one wave per 8x8 tile instead of the persistent
                                 kernel with per-lane refill (same images; for comparison) */

int ft_batch_create(const ft_scene_ir* const* scenes, uint32_t n_scenes, uint32_t accel_mode, ft_batch** out);
int ft_batch_destroy(ft_batch* batch);
/* Outputs are device pointers: n_views images of width*height*3 bytes each, view order;
 * either may be null when its flag is not set.  Enqueues and returns. */
int ft_batch_render_device(ft_batch* batch, const ft_view* views, uint32_t n_views,
                           const ft_render_options* options, const ft_ray_settings* settings,
                           uint32_t flags, void* depth_rgb_dev, void* whitted_rgb_dev);
/* The same into host buffers (synchronous). */
int ft_batch_render(ft_batch* batch, const ft_view* views, uint32_t n_views,
                    const ft_render_options* options, const ft_ray_settings* settings,
                    uint32_t flags, uint8_t* depth_rgb_out, uint8_t* whitted_rgb_out);
int ft_batch_synchronize(ft_batch* batch);
int ft_batch_stats(ft_batch* batch, ft_render_stats* out);
/* FLOP counting over every scene of the batch (see ft_scene_set_flop_counting). */
int ft_batch_set_flop_counting(ft_batch* batch, int enable);

/* --------------------------------------------------------- differentiable rendering
 * Gradients of depth and Lambert shade with respect to scene parameters.  No counterpart
 * in the reference (it renders only); the forward is its march (raymarch_hit
 * renderer.rs:4091-4139) on a parametric scene.
 * Scene: n_prims (1..48) spheres joined by a balanced union_round tree (split [lo, mid) + [mid, hi),
 * mid = lo + (hi - lo) / 2).  Parameters, n_params = 5 * n_prims - 1:
 *   [cx, cy, cz, r] per sphere, then the blend radius k of each join in post order.
 * Per view (1 spp, pixel centres) the forward gives t* (the march's hit distance refined
 * by one Newton step, -1 on a miss) and s = max(0, n . light) with n the central-difference
 * normal of step normal_step at o + t* d.  Gradients use pixels whose 3x3 neighbourhood
 * hits (interior pixels) and are summed in FP64. */
typedef struct ft_diff ft_diff;

int ft_diff_create(uint32_t n_prims, const ft_view* views, uint32_t n_views, const ft_render_options* options,
                   const float light_dir[3], float normal_step, ft_diff** out);
int ft_diff_destroy(ft_diff* diff);
int ft_diff_param_count(ft_diff* diff, uint32_t* n_params);
int ft_diff_set_params(ft_diff* diff, const float* params);
int ft_diff_get_params(ft_diff* diff, float* params);
/* Forward into host images (n_views * width * height floats each), synchronous. */
int ft_diff_forward(ft_diff* diff, float* t_out, float* s_out);
/* Forward into device images (either may be null: the result is still kept for backward). */
int ft_diff_forward_device(ft_diff* diff, void* t_dev, void* s_dev);
int ft_diff_synchronize(ft_diff* diff);
/* dL/dparams (FP64, host) for upstream per-pixel gradients dL/dt*, dL/ds (device images;
 * either may be null) at the parameters and images of the last forward.  Synchronous. */
int ft_diff_backward_device(ft_diff* diff, const void* grad_t_dev, const void* grad_s_dev, double* grad_out);
/* Target images (host) for the built-in L2 objective
 *   L = sum over pixels interior to both hit masks of w_t (t* - T)^2 + w_s (s - S)^2. */
int ft_diff_set_target(ft_diff* diff, const float* t_target, const float* s_target);
/* L and dL/dparams at the current parameters (FP64, host). */
int ft_diff_loss_grad(ft_diff* diff, float w_t, float w_s, double* loss_out, double* grad_out);
/* `steps` Adam steps on the device (beta1 0.9, beta2 0.999, eps 1e-8); losses_out
 * (nullable) receives L before each step.  Synchronous. */
int ft_diff_fit(ft_diff* diff, uint32_t steps, float lr, float w_t, float w_s, double* losses_out);
/* sdf_flops: FLOPs of the forward and adjoint tree passes since counting was turned on
 * (FLOP counting on), else 0. */
int ft_diff_stats(ft_diff* diff, ft_render_stats* out);
/* Turns FLOP counting on (1, clears the counter) or off (0): later forwards and adjoints run
 * the *_counted kernels, which add their per-lane tree-pass FLOPs with one atomic per wave. */
int ft_diff_set_flop_counting(ft_diff* diff, int enable);

/* Scene specialisation diagnostics (no GPU needed).  ft_emit_scene_source writes the
 * translation unit the emitter produces for `ir` under `accel_mode` (always emitted, the
 * FT_EMIT policy aside) into buf (cap bytes, NUL-terminated, *len_out = full length);
 * ft_emit_compile additionally compiles it with hiprtc for one kernel (0 depth, 1/2/3
 * Whitted with a 16/48/128-entry ray stack) and reports the compile time.  No counterpart
 * in the reference: its Cranelift JIT (crates/forgedthoughts/src/jit.rs) has no such hook;
 * these exist so a host without a GPU can check the emitted code. */
int ft_emit_scene_source(const ft_scene_ir* ir, uint32_t accel_mode, char* buf, uint64_t cap, uint64_t* len_out);
int ft_emit_compile(const ft_scene_ir* ir, uint32_t accel_mode, int kernel, double* compile_ms_out);

/* Distance atlas for coarse march steps (opt-in; no counterpart in the reference, whose
 * Bricks accelerator is a grid of leaf lists, renderer.rs:824-954).  Builds, on the GPU,
 * a grid of cells of 8^3 voxels of size `voxel` over the scene bounds; cells whose centre
 * lies within 1.5 half-diagonals + 2 voxels of the surface hold a 9^3 brick of exact
 * distances stored in `bits` = 16 (FP16) or 32 (FP32); the others keep their centre
 * distance.  Later renders on the scene step on the atlas while it reads more than two
 * voxels and evaluate the exact scene from there on (refinement, normal and shading are
 * always exact).  Hits are the reference's surface points but the step sequence differs,
 * so t agrees to about 2 * epsilon rather than bit for bit.  bits = 0 removes the atlas.
 * max_bricks bounds the memory (0: 1 << 20). */
int ft_scene_build_atlas(ft_scene* scene, float voxel, uint32_t bits, uint32_t max_bricks);

/* ----------------------------------------------------------- launch shapes
 * How the march kernels are launched and compiled for one scene.  No counterpart in the
 * reference, whose renderer is CPU loops over tiles and rows (renderer.rs:3659-3717). */
typedef struct ft_launch_config {
  uint32_t block_waves;   /* wavefronts per workgroup: 1, 2 or 4 */
  uint32_t tile_width;    /* pixel tile of one wavefront: 8 (8x8), 16 (16x4) or 32 (32x2); 0 (Whitted
                             with 4, 16 or 64 samples): every sample of a 4x4, 2x2 or 1x1 pixel
                             block (other sample counts use 8x8) */
  uint32_t waves_per_eu;  /* occupancy the kernel is compiled for, 1..8 (VGPR budget 512 / n) */
  uint32_t march_unroll;  /* march steps per loop iteration: 1, 2 or 4 */
  uint32_t shade_inline;  /* 1: the Whitted shading functions are inlined into the trace loop */
} ft_launch_config;
#define FT_KIND_DEPTH 0u
#define FT_KIND_WHITTED 1u
/* The defaults are {1, 8, 4, 1, 0} (depth) and {1, 0, 4, 1, 0} (Whitted), both with the code of
 * the precompiled kernels; any other occupancy, unroll or inlining runs from a runtime-compiled
 * code object built on first use. */
int ft_scene_set_launch_config(ft_scene* scene, uint32_t kind, const ft_launch_config* config);
int ft_scene_get_launch_config(ft_scene* scene, uint32_t kind, ft_launch_config* config);

typedef struct ft_tune_report {
  ft_launch_config best;
  double best_ms;            /* kernel time of the chosen shape on the tuning render */
  double default_ms;         /* kernel time of the default shape on the same render */
  double tune_ms;            /* wall time of the search, compilation included */
  uint32_t candidates;       /* shapes timed */
  uint32_t cached;           /* 1: the choice came from the tuning cache */
  uint32_t tune_aa_samples;  /* samples per pixel of the tuning render */
  uint32_t pad;
} ft_tune_report;
#define FT_TUNE_FULL 1u        /* time the request itself (default: Whitted at <= 16 samples) */
#define FT_TUNE_SHAPE_ONLY 2u  /* search block and tile shapes only (no code variants to build) */
#define FT_TUNE_NO_CACHE 4u    /* search even when the cache holds a choice */
/* Autotuner: times the scene's `kind` render of this request under candidate shapes,
 * first the code variants (occupancy 5/4/3/2, march unroll 1/2, shading calls or inlined) in
 * the default launch shape,
 * then block and tile shapes of the best code, and keeps the fastest as the scene's shape
 * for that kind.  Choices are cached per (scene IR hash, kind, request) in the process and,
 * when a cache directory is writable, on disk next to the code objects.  Synchronous. */
int ft_scene_autotune(ft_scene* scene, uint32_t kind, const ft_render_options* options,
                      const ft_ray_settings* settings, uint32_t flags, ft_tune_report* report);

/* Turns FLOP counting on (1) or off (0) for later renders on this scene.  Counting adds
 * atomics to every SDF evaluation: use it for work accounting, never for timing. */
int ft_scene_set_flop_counting(ft_scene* scene, int enable);

/* sizeof() of the named ABI struct ("ft_node", "ft_scene_ir", ...) or -1, so bindings
 * can verify their layout against the library they load. */
int ft_abi_struct_size(const char* name);

#ifdef __cplusplus
}
#endif

#endif /* FT_ABI_H */
