// ft_shade.h — per-hit resolution, BSDFs, lights and the Whitted integrator (device).
//
// The reference resolves the shading normal and material of a hit by walking the scene
// tree (renderer.rs:4397-4570, 6926-7651) and traces reflection/refraction recursively
// (renderer/ray.rs:3-435).  Here:
//   * the walk starts at the accel leaf that produced the hit distance: above the leaves
//     the tree is hard Unions only, and the reference walk there follows the side with
//     the smaller distance, i.e. exactly the leaf that wins the min;
//   * the walk below is iterative with explicit frames, evaluating child subtrees on
//     demand as the reference does;
//   * recursion is replaced by a per-lane stack of pending rays carrying a throughput:
//     every term of trace_hit_with_material is linear in the child radiance, so the
//     pixel value is the sum over the ray tree of path weight x local term.
#pragma once

#include "ft_device.h"

// FT_SHADE_INLINE=1 inlines the shading functions (hit_material, shade_color) into the
// Whitted loop (a code variant of the launch-shape autotuner); by default they are calls.
#if defined(FT_SHADE_INLINE) && FT_SHADE_INLINE
#define FT_SHADE_FN __forceinline__
#else
#define FT_SHADE_FN __noinline__
#endif

namespace ftd {

enum { TAG_STANDARD = 0, TAG_LAMBERT = 1, TAG_METAL = 2, TAG_DIELECTRIC = 3 };

struct Mat {  // Material (materials/mod.rs:214-221); a is used when !blend
  int blend, a_model, b_model, pad;
  float t;
  DParams a, b;
};

__device__ __forceinline__ float3 v3(const float* f) { return mk(f[0], f[1], f[2]); }
__device__ __forceinline__ void set3(float* f, float3 v) { f[0] = v.x; f[1] = v.y; f[2] = v.z; }

// lerp_spectrum (renderer.rs:5206-5213, materials/mod.rs:434-441)
__device__ __forceinline__ float3 lerp_spec(float3 a, float3 b, float t) {
  float tt = clampf(t, 0.0f, 1.0f);
  return mk(a.x * (1.0f - tt) + b.x * tt, a.y * (1.0f - tt) + b.y * tt, a.z * (1.0f - tt) + b.z * tt);
}
__device__ __forceinline__ float smoothstepf(float x) {  // smoothstepf(0, 1, x) (renderer.rs:7538-7541)
  float t = clampf(x / 1.0f, 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

// blend_params (renderer.rs:7602-7627)
__device__ inline void blend_params(DParams& a, const DParams& b, float t) {
  set3(a.color, lerp_spec(v3(a.color), v3(b.color), t));
  a.roughness = a.roughness + (b.roughness - a.roughness) * t;
  a.metallic = a.metallic + (b.metallic - a.metallic) * t;
  a.specular = a.specular + (b.specular - a.specular) * t;
  a.specular_weight = a.specular_weight + (b.specular_weight - a.specular_weight) * t;
  set3(a.specular_color, lerp_spec(v3(a.specular_color), v3(b.specular_color), t));
  a.ior = a.ior + (b.ior - a.ior) * t;
  a.clearcoat = a.clearcoat + (b.clearcoat - a.clearcoat) * t;
  a.clearcoat_roughness = a.clearcoat_roughness + (b.clearcoat_roughness - a.clearcoat_roughness) * t;
  a.transmission = a.transmission + (b.transmission - a.transmission) * t;
  a.thin_walled = t < 0.5f ? a.thin_walled : b.thin_walled;
  a.hook = -1.0f;  // a.dynamic_material_id = None (renderer.rs:7624): no eval hook after a lerp
  set3(a.emission_color, lerp_spec(v3(a.emission_color), v3(b.emission_color), t));
  a.emission_strength = a.emission_strength + (b.emission_strength - a.emission_strength) * t;
  if (a.has_medium != 0.0f && b.has_medium != 0.0f) {  // blend_medium (renderer.rs:7543-7566)
    a.medium_ior = a.medium_ior + (b.medium_ior - a.medium_ior) * t;
    set3(a.medium_absorption, lerp_spec(v3(a.medium_absorption), v3(b.medium_absorption), t));
    a.medium_density = a.medium_density + (b.medium_density - a.medium_density) * t;
  } else if (a.has_medium != 0.0f) {
    a.has_medium = t < 0.5f ? 1.0f : 0.0f;
  } else if (b.has_medium != 0.0f && t >= 0.5f) {
    a.has_medium = 1.0f;
    a.medium_ior = b.medium_ior;
    set3(a.medium_absorption, v3(b.medium_absorption));
    a.medium_density = b.medium_density;
  }
  a.pattern = 0.0f;
}

__device__ __forceinline__ int dom_model(const Mat& m) { return !m.blend ? m.a_model : (m.t < 0.5f ? m.a_model : m.b_model); }
__device__ __forceinline__ const DParams& dom_params(const Mat& m) { return (!m.blend || m.t < 0.5f) ? m.a : m.b; }

// blend_materials (renderer.rs:7629-7651)
__device__ inline void blend_materials(Mat& left, const Mat& right, float t) {
  if (!left.blend && !right.blend && left.a_model == right.a_model) {
    blend_params(left.a, right.a, t);
    return;
  }
  Mat out;
  out.blend = 1;
  out.a_model = dom_model(left);
  out.a = dom_params(left);
  out.b_model = dom_model(right);
  out.b = dom_params(right);
  out.t = clampf(t, 0.0f, 1.0f);
  out.pad = 0;
  left = out;
}

// The shading context of one hit as material hooks see it (make_shading_context
// renderer.rs:7875-7911): hit position, front-facing unit normal, unit view direction.
struct HookCtx {
  float3 pos, normal, view;
  float front;
};

// the clamps resolve_leaf_material applies to every leaf set (renderer.rs:7472-7530)
__device__ __forceinline__ void clamp_leaf(int model, DParams& a) {
  if (model != TAG_LAMBERT) a.roughness = clampf(a.roughness, 0.0f, 1.0f);
  if (model == TAG_STANDARD || model == TAG_DIELECTRIC) a.ior = clampf(a.ior, 1.0f, 3.0f);
  a.emission_strength = fmaxf(a.emission_strength, 0.0f);
}

// resolve_leaf_material (renderer.rs:7164-7536): the parameter set of a leaf material at
// `pos`; returns its model.  With a shading context, the material's per-hit hooks replace
// the static colour, roughness, ior, thin_walled and emission before the clamps (the arms
// of renderer.rs:7472-7530); a Standard / Layered definition also has its metallic,
// specular, specular weight, transmission and coat hooks (renderer.rs:7286-7470), the coat
// hooks listed under the parameter each replaces, the clearcoat being
// clamp(weight * coat_mask).
__device__ inline int leaf_params(int slot, unsigned mat_id, unsigned obj, float3 pos, DParams& a,
                                  const HookCtx* hc = nullptr) {
  const DevScene& S = scene_of(slot);
  unsigned id = mat_id < (unsigned)S.n_mats ? mat_id : 0u;
  a = S.mats[id];
  const int model = (int)a.model;
  const int hk = (int)a.hook;
  const bool hooked = hk >= 0 && hc != nullptr;
  const bool checker = a.pattern == (float)FT_PATTERN_CHECKER3D;
  if (checker || hooked) {
    const float3 lp = object_local(slot, obj, pos);
    if (hooked) {
      const DHooks H = S.hooks[hk];
      float3 args[FT_HOOK_PARAMS];
      args[0] = hc->pos;
      args[1] = lp;
      args[2] = hc->normal;
      args[3] = hc->view;
      args[4] = splat3(hc->front);
      args[5] = splat3((float)obj);
      args[6] = splat3((float)mat_id);
      for (int k = 7; k < (int)FT_HOOK_PARAMS; ++k) args[k] = mk(0.0f, 0.0f, 0.0f);
      if (H.color >= 0) {
        set3(a.color, eval_fn(slot, H.color, args));
      } else if (checker) {  // resolve_pattern_color (renderer.rs:7855-7873)
        const int sx = (int)floorf(lp.x * a.pattern_scale), sy = (int)floorf(lp.y * a.pattern_scale),
                  sz = (int)floorf(lp.z * a.pattern_scale);
        set3(a.color, ((sx + sy + sz) & 1) == 0 ? v3(a.pattern_a) : v3(a.pattern_b));
      }
      if (H.roughness >= 0) a.roughness = eval_fn(slot, H.roughness, args).x;
      if (H.ior >= 0) a.ior = eval_fn(slot, H.ior, args).x;
      if (H.thin_walled >= 0) a.thin_walled = eval_fn(slot, H.thin_walled, args).x >= 0.5f ? 1.0f : 0.0f;
      if (H.emission_color >= 0) set3(a.emission_color, eval_fn(slot, H.emission_color, args));
      if (H.emission_strength >= 0) a.emission_strength = eval_fn(slot, H.emission_strength, args).x;
      if (H.metallic >= 0) a.metallic = clampf(eval_fn(slot, H.metallic, args).x, 0.0f, 1.0f);
      if (H.specular >= 0) a.specular = clampf(eval_fn(slot, H.specular, args).x, 0.0f, 1.0f);
      if (H.specular_weight >= 0) a.specular_weight = clampf(eval_fn(slot, H.specular_weight, args).x, 0.0f, 1.0f);
      if (H.transmission >= 0) a.transmission = clampf(eval_fn(slot, H.transmission, args).x, 0.0f, 1.0f);
      if (H.specular_color >= 0) set3(a.specular_color, eval_fn(slot, H.specular_color, args));
      if (H.clearcoat_roughness >= 0)
        a.clearcoat_roughness = clampf(eval_fn(slot, H.clearcoat_roughness, args).x, 0.0f, 1.0f);
      if (H.clearcoat >= 0) a.clearcoat = clampf(eval_fn(slot, H.clearcoat, args).x, 0.0f, 1.0f);
      if (H.coat_mask >= 0)
        a.clearcoat = clampf(a.clearcoat * clampf(eval_fn(slot, H.coat_mask, args).x, 0.0f, 1.0f), 0.0f, 1.0f);
      if (H.medium_ior >= 0) {  // medium_from_value (renderer.rs:6834-6845) of the hook's Medium
        a.has_medium = 1.0f;
        a.medium_ior = clampf(eval_fn(slot, H.medium_ior, args).x, 1.0f, 3.0f);
        set3(a.medium_absorption, eval_fn(slot, H.medium_absorption, args));
        a.medium_density = fmaxf(eval_fn(slot, H.medium_density, args).x, 0.0f);
      }
    } else {  // resolve_pattern_color (renderer.rs:7855-7873)
      const int sx = (int)floorf(lp.x * a.pattern_scale), sy = (int)floorf(lp.y * a.pattern_scale),
                sz = (int)floorf(lp.z * a.pattern_scale);
      set3(a.color, ((sx + sy + sz) & 1) == 0 ? v3(a.pattern_a) : v3(a.pattern_b));
    }
  }
  clamp_leaf(model, a);
  return model;
}
// resolve_dynamic_bump / resolve_dynamic_normal / orthonormal_tangent (renderer.rs:7749-7827):
// the material of the hit's id reshapes the shading normal `gn` of build_bsdf_context
// (renderer.rs:4397-4449) through its bump hook (finite differences at 0.01 along a tangent
// frame) and then its normal hook, each result flipped to the side of the normal it
// started from.  Context: make_shading_context_with_normal (renderer.rs:7884-7911), the
// normal being the one reshaped and view_dir = -dir as build_bsdf_context receives it.
__device__ __noinline__ float3 apply_normal_hooks(int slot, float3 pos, bool front, unsigned obj, unsigned mat_id,
                                                  float3 view, float3 gn) {
  const DevScene& S = scene_of(slot);
  const unsigned id = mat_id < (unsigned)S.n_mats ? mat_id : 0u;
  const int hk = (int)S.mats[id].hook;
  if (hk < 0) return gn;
  const DHooks H = S.hooks[hk];
  if (H.bump < 0 && H.normal < 0) return gn;
  const float3 lp = object_local(slot, obj, pos);
  float3 args[FT_HOOK_PARAMS];
  args[0] = pos;
  args[1] = lp;
  args[2] = gn;
  args[3] = view;
  args[4] = splat3(front ? 1.0f : 0.0f);
  args[5] = splat3((float)obj);
  args[6] = splat3((float)mat_id);
  for (int k = 7; k < (int)FT_HOOK_PARAMS; ++k) args[k] = mk(0.0f, 0.0f, 0.0f);
  float3 n = gn;
  if (H.bump >= 0) {
    const float center = eval_fn(slot, H.bump, args).x;
    const float3 up = fabsf(gn.y) < 0.999f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
    const float3 tangent = normalize(cross(gn, up));
    const float3 bitangent = normalize(cross(gn, tangent));
    const float eps = 0.01f;
    args[1] = add(lp, mul(tangent, eps));
    const float dx = eval_fn(slot, H.bump, args).x - center;
    args[1] = add(lp, mul(bitangent, eps));
    const float dy = eval_fn(slot, H.bump, args).x - center;
    args[1] = lp;
    const float3 nb = normalize(sub(sub(gn, mul(tangent, dx / eps)), mul(bitangent, dy / eps)));
    if (len(nb) > 1.0e-6f) n = dot(nb, gn) < 0.0f ? mul(nb, -1.0f) : nb;
  }
  if (H.normal >= 0) {
    args[2] = n;
    const float3 nn = normalize(eval_fn(slot, H.normal, args));
    if (len(nn) > 1.0e-6f) n = dot(nn, n) < 0.0f ? mul(nn, -1.0f) : nn;
  }
  return n;
}

__device__ inline void leaf_material(int slot, unsigned mat_id, unsigned obj, float3 pos, Mat& m,
                                     const HookCtx* hc = nullptr) {
  m.blend = 0;
  m.a_model = leaf_params(slot, mat_id, obj, pos, m.a, hc);
  m.b_model = 0;
  m.t = 0.0f;
  m.pad = 0;
}
__device__ __forceinline__ int leaf_model(const DevScene& S, unsigned mat_id) {
  unsigned id = mat_id < (unsigned)S.n_mats ? mat_id : 0u;
  return (int)S.mats[id].model;
}

__device__ __forceinline__ int blend_class(int k) {
  switch (k) {
    case FT_NODE_UNION_ROUND: case FT_NODE_UNION_CHAMFER: case FT_NODE_UNION_SOFT:
    case FT_NODE_INTERSECT_ROUND: case FT_NODE_INTERSECT_CHAMFER: case FT_NODE_UNION_COLUMNS:
    case FT_NODE_UNION_STAIRS: case FT_NODE_INTERSECT_COLUMNS: case FT_NODE_INTERSECT_STAIRS: return 1;
    case FT_NODE_DIFF_ROUND: case FT_NODE_DIFF_CHAMFER: case FT_NODE_PIPE: case FT_NODE_ENGRAVE:
    case FT_NODE_DIFF_COLUMNS: case FT_NODE_DIFF_STAIRS: case FT_NODE_GROOVE: case FT_NODE_TONGUE: return 2;
    default: return 0;
  }
}
__device__ __forceinline__ float blend_radius(const DNode& n) {
  return (n.kind == FT_NODE_GROOVE || n.kind == FT_NODE_TONGUE) ? fmaxf(n.p[0], n.p[1]) : n.p[0];
}
// blend weight t of a smooth node from its children's distances
__device__ __forceinline__ float blend_t(const DNode& n, float dl, float dr) {
  float k = fmaxf(fabsf(blend_radius(n)), 1.0e-6f);
  if (blend_class(n.kind) == 2) return smoothstepf(0.5f + 0.5f * (((-dr) - dl) / k));
  return smoothstepf(0.5f + 0.5f * ((dl - dr) / k));
}
// hard boolean side choice (renderer.rs:4473-4498, 7000-7025)
__device__ __forceinline__ bool hard_left(int kind, float dl, float dr) {
  if (kind == FT_NODE_UNION) return dl <= dr;
  if (kind == FT_NODE_INTERSECT) return dl >= dr;
  return dl >= -dr;
}
__device__ __forceinline__ bool is_hard(int k) { return k == FT_NODE_UNION || k == FT_NODE_INTERSECT || k == FT_NODE_SUBTRACT; }

// estimate_node_normal (renderer.rs:4554-4570)
__device__ inline float3 node_normal(int slot, int node, float3 p, float eps) {
  const DevScene& S = scene_of(slot);
  float e = fmaxf(eps, 1.0e-5f);
  float d1 = node_dist(slot, node, mk(p.x + e, p.y - e, p.z - e));
  float d2 = node_dist(slot, node, mk(p.x - e, p.y - e, p.z + e));
  float d3 = node_dist(slot, node, mk(p.x - e, p.y + e, p.z - e));
  float d4 = node_dist(slot, node, mk(p.x + e, p.y + e, p.z + e));
  float3 n = add(add(add(mk(d1, -d1, -d1), mk(-d2, -d2, d2)), mk(-d3, d3, -d3)), mk(d4, d4, d4));
  if (len(n) > 1.0e-6f) return normalize(n);
  return mk(0.0f, 1.0f, 0.0f);
}

struct Frame { int node, phase; };

// resolve_surface_normal_from_node (renderer.rs:4462-4552), from `start`
__device__ __noinline__ float3 walk_normal(int slot, int start, float3 p, float eps) {
  const DevScene& S = scene_of(slot);
  Frame fr[kMaxWalk];
  float3 res[kMaxWalk];
  int fsp = 0, rsp = 0;
  fr[fsp++] = {start, 0};
  while (fsp > 0) {
    Frame f = fr[--fsp];
    const DNode& n = S.nodes[f.node];
    if (n.kind <= FT_NODE_NOISE) { res[rsp++] = node_normal(slot, f.node, p, eps); continue; }
    if (n.kind == FT_NODE_SLICE || n.kind == FT_NODE_SMOOTH) { fr[fsp++] = {n.lhs, 0}; continue; }
    if (is_hard(n.kind)) {
      if (f.phase == 1) { res[rsp - 1] = mul(res[rsp - 1], -1.0f); continue; }
      float dl = node_dist(slot, n.lhs, p), dr = node_dist(slot, n.rhs, p);
      if (hard_left(n.kind, dl, dr)) fr[fsp++] = {n.lhs, 0};
      else {
        if (n.kind == FT_NODE_SUBTRACT) fr[fsp++] = {f.node, 1};
        fr[fsp++] = {n.rhs, 0};
      }
      continue;
    }
    if (f.phase == 0) { fr[fsp++] = {f.node, 1}; fr[fsp++] = {n.lhs, 0}; continue; }
    if (f.phase == 1) { fr[fsp++] = {f.node, 2}; fr[fsp++] = {n.rhs, 0}; continue; }
    float3 right = res[--rsp], left = res[--rsp];
    float dl = node_dist(slot, n.lhs, p), dr = node_dist(slot, n.rhs, p);
    if (blend_class(n.kind) == 2) right = mul(right, -1.0f);
    float t = blend_t(n, dl, dr);
    float3 nn = add(mul(left, 1.0f - t), mul(right, t));
    res[rsp++] = len(nn) > 1.0e-6f ? normalize(nn) : (t < 0.5f ? left : right);
  }
  return res[0];
}

// resolve_material_from_node (renderer.rs:6945-7059), from `start`
__device__ __noinline__ void walk_material(int slot, int start, float3 p, Mat& out, int base = 0,
                                          const HookCtx* hc = nullptr) {
  const DevScene& S = scene_of(slot);
  Frame fr[kMaxWalk];
  Mat res[kMaxWalkMats];
  int fsp = 0, rsp = 0;
  fr[fsp++] = {start, 0};
  while (fsp > 0) {
    Frame f = fr[--fsp];
    const DNode& n = S.nodes[f.node];
    if (n.kind <= FT_NODE_CUSTOM) { leaf_material(slot, n.mat, n.obj, p, res[rsp++], hc); continue; }
    if (n.kind <= FT_NODE_NOISE || n.kind == FT_NODE_SLICE || n.kind == FT_NODE_SMOOTH) {
      fr[fsp++] = {n.lhs, 0};
      continue;
    }
    if (is_hard(n.kind)) {
      float dl = node_dist(slot, n.lhs, p, base), dr = node_dist(slot, n.rhs, p, base);
      fr[fsp++] = {hard_left(n.kind, dl, dr) ? n.lhs : n.rhs, 0};
      continue;
    }
    if (f.phase == 0) { fr[fsp++] = {f.node, 1}; fr[fsp++] = {n.lhs, 0}; continue; }
    if (f.phase == 1) { fr[fsp++] = {f.node, 2}; fr[fsp++] = {n.rhs, 0}; continue; }
    --rsp;
    float dl = node_dist(slot, n.lhs, p, base), dr = node_dist(slot, n.rhs, p, base);
    blend_materials(res[rsp - 1], res[rsp], blend_t(n, dl, dr));
  }
  out = res[0];
}

// resolve_split_material_from_node (renderer.rs:7061-7120)
__device__ __noinline__ bool split_material(int slot, int start, float3 p, Mat& a, Mat& b, float& t,
                                           const HookCtx* hc = nullptr) {
  const DevScene& S = scene_of(slot);
  int node = start;
  for (;;) {
    const DNode& n = S.nodes[node];
    if (is_hard(n.kind)) {
      float dl = node_dist(slot, n.lhs, p), dr = node_dist(slot, n.rhs, p);
      node = hard_left(n.kind, dl, dr) ? n.lhs : n.rhs;
      continue;
    }
    if (n.kind == FT_NODE_SMOOTH || n.kind == FT_NODE_NOISE) { node = n.lhs; continue; }
    if (blend_class(n.kind) == 0) return false;
    float dl = node_dist(slot, n.lhs, p), dr = node_dist(slot, n.rhs, p);
    walk_material(slot, n.lhs, p, a, 0, hc);
    walk_material(slot, n.rhs, p, b, 0, hc);
    t = blend_t(n, dl, dr);
    return true;
  }
}

// ----------------------------------------------------------- shading sweep
// walk_normal, walk_material and split_material above descend the tree from the hit leaf
// and, at every smooth boolean, evaluate both child subtrees again (O(n log n) subtree
// evaluations per hit, as the reference does).  The sweep below produces the same three
// results in one post-order pass over the node program: each node's distance and lower
// bound (exactly as eval_uniform computes them), its walk normal and its walk material are
// built from its children's.  Hard booleans above the first non-hard node are resolved
// first per lane (the walks' own side choice); modifier and noise subtrees, terminals of
// the walks, use the per-node functions.

// tetrahedral normal of a terminal node (estimate_node_normal renderer.rs:4554-4570)
template <class F>
__device__ __forceinline__ float3 tetra_normal(F&& f, float3 p, float eps) {
  const float e = fmaxf(eps, 1.0e-5f);
  const float d1 = f(mk(p.x + e, p.y - e, p.z - e));
  const float d2 = f(mk(p.x - e, p.y - e, p.z + e));
  const float d3 = f(mk(p.x - e, p.y + e, p.z - e));
  const float d4 = f(mk(p.x + e, p.y + e, p.z + e));
  const float3 n = add(add(add(mk(d1, -d1, -d1), mk(-d2, -d2, d2)), mk(-d3, d3, -d3)), mk(d4, d4, d4));
  if (len(n) > 1.0e-6f) return normalize(n);
  return mk(0.0f, 1.0f, 0.0f);
}

// Materials of the sweep by reference.  A material value is a single parameter set or a
// blend of two (blend_materials renderer.rs:7629-7651).  A sweep stack entry keeps the
// models, the blend weight and two handles: h >= 0 names the primitive or custom node whose
// leaf material it is (resolved at p when needed), h < 0 the pool entry -1-h holding a
// parameter set that a blend produced.  A smooth node over two single sets of one model
// lerps them field by field (blend_params); at t == 0 that is the left set itself and at
// t == 1 the right one up to rounding (a + (b - a) for b), so only 0 < t < 1 computes
// parameters.  A two-model blend takes each side's dominant set by handle.  Every pool
// entry has at most one live reference: a lerp overwrites the entry of an operand, and an
// entry whose reference is dropped returns to the pool.
struct MRef {
  int blend, a_model, b_model;
  float t;
  int ha, hb;
};
constexpr int kMatPool = 8;
struct MatPool {
  DParams e[kMatPool];
  unsigned used;  // bit i: entry i is referenced
  int overflow;
};
// A leaf handle with this bit names a leaf set that a smooth blend at t == 0 or 1 passed
// through: its parameters are the leaf's, its dynamic id (hook entry) is dropped.
constexpr int kLerped = 1 << 30;
__device__ __forceinline__ void mark_lerped(MatPool& pool, int& h) {
  if (h < 0) pool.e[-1 - h].hook = -1.0f;
  else h |= kLerped;
}
__device__ __forceinline__ void ref_params(int slot, const MatPool& pool, int h, float3 p, DParams& out,
                                           const HookCtx* hc) {
  if (h < 0) {
    out = pool.e[-1 - h];
  } else {
    const DNode& n = scene_of(slot).nodes[h & ~kLerped];
    leaf_params(slot, n.mat, n.obj, p, out, hc);
    if (h & kLerped) out.hook = -1.0f;
  }
}
__device__ inline void ref_to_mat(int slot, const MatPool& pool, const MRef& r, float3 p, Mat& m,
                                  const HookCtx* hc) {
  m.blend = r.blend;
  m.a_model = r.a_model;
  m.b_model = r.b_model;
  m.t = r.t;
  m.pad = 0;
  ref_params(slot, pool, r.ha, p, m.a, hc);
  if (r.blend) ref_params(slot, pool, r.hb, p, m.b, hc);
}
__device__ __forceinline__ int pool_take(MatPool& pool) {
  const unsigned free_ = ~pool.used & ((1u << kMatPool) - 1u);
  if (free_ == 0u) { pool.overflow = 1; return -1; }
  const int i = __builtin_ctz(free_);
  pool.used |= 1u << i;
  return i;
}
__device__ __forceinline__ void pool_drop(MatPool& pool, int h) {
  if (h < 0) pool.used &= ~(1u << (-1 - h));
}
__device__ __forceinline__ void drop_ref(MatPool& pool, const MRef& r) {
  pool_drop(pool, r.ha);
  if (r.blend) pool_drop(pool, r.hb);
}
// a walk material (modifier terminal) into the pool
__device__ inline MRef ref_of_mat(MatPool& pool, const Mat& m) {
  MRef r{m.blend, m.a_model, m.blend ? m.b_model : 0, m.blend ? m.t : 0.0f, 0, 0};
  const int ia = pool_take(pool);
  if (ia < 0) return r;
  pool.e[ia] = m.a;
  r.ha = -1 - ia;
  if (m.blend) {
    const int ib = pool_take(pool);
    if (ib < 0) return r;
    pool.e[ib] = m.b;
    r.hb = -1 - ib;
  }
  return r;
}
// the lerp of two single parameter sets of one model, 0 < t < 1, into a pool entry
// The handles go in and out by value: a reference would give the caller's MRef variables
// an address, and the emitted sweep's hundreds of them would live in scratch memory.
__device__ __noinline__ int blend_refs_lerp(int slot, MatPool& pool, int lha, int rha, float t, float3 p,
                                           const HookCtx* hc) {
  const int dst = lha < 0 ? -1 - lha : (rha < 0 ? -1 - rha : pool_take(pool));
  if (dst < 0) return lha;
  DParams a, b;
  ref_params(slot, pool, lha, p, a, hc);
  ref_params(slot, pool, rha, p, b, hc);
  blend_params(a, b, t);
  pool.e[dst] = a;
  if (lha < 0 && rha < 0) pool_drop(pool, rha);  // dst is l's entry
  return -1 - dst;
}
// blend_materials on references; r is consumed
__device__ __forceinline__ void blend_refs(int slot, MatPool& pool, MRef& l, const MRef& r, float t, float3 p,
                                           const HookCtx* hc) {
  if (!l.blend && !r.blend && l.a_model == r.a_model) {
    // t == 0 / 1: the left / right set itself (blend_params' lerp up to rounding), marked as
    // a lerp result, which has no dynamic id
    if (t == 0.0f) { pool_drop(pool, r.ha); mark_lerped(pool, l.ha); return; }
    if (t == 1.0f) { pool_drop(pool, l.ha); l = r; mark_lerped(pool, l.ha); return; }
    l.ha = blend_refs_lerp(slot, pool, l.ha, r.ha, t, p, hc);
    return;
  }
  const bool la = !l.blend || l.t < 0.5f, ra = !r.blend || r.t < 0.5f;
  if (l.blend) pool_drop(pool, la ? l.hb : l.ha);
  if (r.blend) pool_drop(pool, ra ? r.hb : r.ha);
  MRef out;
  out.blend = 1;
  out.a_model = la ? l.a_model : l.b_model;
  out.ha = la ? l.ha : l.hb;
  out.b_model = ra ? r.a_model : r.b_model;
  out.hb = ra ? r.ha : r.hb;
  out.t = clampf(t, 0.0f, 1.0f);
  l = out;
}

// The result of one hit's walks: shading normal and material references.  Parameter sets
// stay where they are (the scene's material table, or the pool for lerped and hooked sets)
// and are read when a BSDF needs them, so no per-hit copy of the parameters is made.
struct ShadeWalk {
  float3 normal;
  MRef a;        // the walk material, or (split) the split node's lhs material
  MRef b;        // (split) the split node's rhs material
  float split_t;
  bool split;
  bool mat_overflow;  // the material pool ran out: materials come from the per-node walks
  MatPool pool;
};

// The parameter set of a handle after finish_refs: a pool entry, or a leaf material without
// hooks or pattern read from the scene's table (with resolve_leaf_material's clamps).
__device__ __forceinline__ void load_ref(int slot, const MatPool& pool, int h, DParams& out) {
  if (h < 0) {
    out = pool.e[-1 - h];
    return;
  }
  const DevScene& S = scene_of(slot);
  const unsigned mat_id = S.nodes[h & ~kLerped].mat;
  out = S.mats[mat_id < (unsigned)S.n_mats ? mat_id : 0u];
  clamp_leaf((int)out.model, out);
  if (h & kLerped) out.hook = -1.0f;
}
// Leaf handles whose set depends on the hit (per-hit hooks, a pattern) are resolved once,
// into pool entries; false when the pool ran out.
__device__ __forceinline__ bool finish_handle(int slot, MatPool& pool, int& h, float3 p, const HookCtx* hc) {
  if (h < 0) return true;
  const DevScene& S = scene_of(slot);
  const unsigned mat_id = S.nodes[h & ~kLerped].mat;
  const DParams& m = S.mats[mat_id < (unsigned)S.n_mats ? mat_id : 0u];
  if (m.hook < 0.0f && m.pattern == 0.0f) return true;
  const int i = pool_take(pool);
  if (i < 0) return false;
  ref_params(slot, pool, h, p, pool.e[i], hc);
  h = -1 - i;
  return true;
}
__device__ __forceinline__ bool finish_ref(int slot, MatPool& pool, MRef& r, float3 p, const HookCtx* hc) {
  return finish_handle(slot, pool, r.ha, p, hc) && (!r.blend || finish_handle(slot, pool, r.hb, p, hc));
}
// the dominant side of a reference (dominant_material_model / _params, renderer.rs:7653-7683)
__device__ __forceinline__ int dom_model(const MRef& m) { return (!m.blend || m.t < 0.5f) ? m.a_model : m.b_model; }
__device__ __forceinline__ int dom_handle(const MRef& m) { return (!m.blend || m.t < 0.5f) ? m.ha : m.hb; }

// The sweep of node N (wave-uniform).  Stack entries: distance / lower bound in the LDS
// value planes, normal and material per lane in registers / private memory.
__device__ __forceinline__ void shade_sweep_uniform(const DevScene& S, int slot, int N, float3 p, float eps,
                                                    ShadeWalk& w, const HookCtx* hc) {
  const auto nodes = uniform_ptr(S.nodes);
  const auto prog = uniform_ptr(S.prog);
  const int pf = __builtin_amdgcn_readfirstlane(S.nodes[N].prog_first);
  const int pl = __builtin_amdgcn_readfirstlane(S.nodes[N].prog_last);
  const int split_node = __builtin_amdgcn_readfirstlane(S.nodes[N].walk_split);
  const int nth = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  float* __restrict__ sd = ft_lds + tid;
  float* __restrict__ slb = ft_lds + S.stack_depth * nth + tid;
  MRef ms[kMaxWalkMats];
  float3 ns[kMaxWalkMats];
  MatPool& pool = w.pool;
  pool.used = 0u;
  pool.overflow = 0;
  int sp = 0;
  w.split = false;
  for (int k = pf; k <= pl; ++k) {
    const unsigned ins = prog[k];
    const int ni = (int)(ins & ~kBeginFlag);
    const DNode n = load_uniform(nodes + ni);
    if (ins & kBeginFlag) {
      // modifier / noise subtree: one terminal of the walks
      const Info v = eval_uniform(S, slot, ni, p, sp);
      sd[sp * nth] = v.d;
      slb[sp * nth] = v.lb;
      ns[sp] = tetra_normal([&](float3 q) { return eval_uniform(S, slot, ni, q, sp).d; }, p, eps);
      Mat tm;
      walk_material(slot, ni, p, tm, sp, hc);
      ms[sp] = ref_of_mat(pool, tm);
      ++sp;
      k = n.prog_last;
      continue;
    }
    const int kind = n.kind;
    if (is_prim(kind)) {
      if (S.flops) atomicAdd(S.flops, 5ull * (unsigned long long)n.cost);  // value + 4 normal samples
      const float d = prim_distance(n, p);
      sd[sp * nth] = d;
      slb[sp * nth] = d;
      ns[sp] = tetra_normal([&](float3 q) { return prim_distance(n, q); }, p, eps);
      ms[sp] = MRef{0, leaf_model(S, n.mat), 0, 0.0f, ni, 0};
      ++sp;
      continue;
    }
    if (kind == FT_NODE_CUSTOM) {
      const float3 q = node_local(n, p);
      const float mn[3] = {-n.p[0], -n.p[1], -n.p[2]}, mx[3] = {n.p[0], n.p[1], n.p[2]};
      sd[sp * nth] = custom_distance(slot, n, q);
      slb[sp * nth] = aabb_lb(q, mn, mx);
      ns[sp] = tetra_normal([&](float3 r) { return custom_distance(slot, n, node_local(n, r)); }, p, eps);
      ms[sp] = MRef{0, leaf_model(S, n.mat), 0, 0.0f, ni, 0};
      ++sp;
      continue;
    }
    const int t = (sp - 1) * nth;
    if (kind == FT_NODE_SLICE) {
      const float c = comp(p, (int)n.p[0]);
      sd[t] = fmaxf(sd[t], fmaxf(n.p[1] - c, c - n.p[2]));
      if (n.flags & 2) slb[t] = aabb_lb(p, n.bmin, n.bmax);
      continue;
    }
    if (kind == FT_NODE_SMOOTH) {
      sd[t] -= n.p[0] * 0.1f;
      slb[t] -= n.p[0] * 0.1f;
      continue;
    }
    // binary: lhs entry l, rhs entry r
    --sp;
    const int l = t - nth, r = t;
    const int il = sp - 1, ir = sp;
    const float a = sd[l], b = sd[r];
    if (kind == FT_NODE_UNION || kind == FT_NODE_INTERSECT || kind == FT_NODE_SUBTRACT) {
      if (kind == FT_NODE_UNION) {
        const float llb = slb[l], rlb = slb[r];
        const bool take_l = llb <= rlb ? (a <= rlb || a <= b) : (!(b <= llb) && a <= b);
        if (!take_l) sd[l] = b;
        slb[l] = fminf(llb, rlb);
      } else if (kind == FT_NODE_INTERSECT) {
        if (!(a >= b)) sd[l] = b;
        slb[l] = fmaxf(slb[l], slb[r]);
      } else {
        if (!(a >= -b)) sd[l] = -b;
      }
      // the walks' side choice (renderer.rs:4473-4498, 7000-7025)
      if (!hard_left(kind, a, b)) {
        ns[il] = kind == FT_NODE_SUBTRACT ? mul(ns[ir], -1.0f) : ns[ir];
        drop_ref(pool, ms[il]);
        ms[il] = ms[ir];
      } else {
        drop_ref(pool, ms[ir]);
      }
      continue;
    }
    const float rr = n.p[0], nn = n.p[1];
    float d;
    switch (kind) {
      case FT_NODE_UNION_ROUND: d = op_union_round(a, b, rr); break;
      case FT_NODE_UNION_CHAMFER: d = op_union_chamfer(a, b, rr); break;
      case FT_NODE_UNION_COLUMNS: d = op_union_columns(a, b, rr, nn); break;
      case FT_NODE_UNION_STAIRS: d = op_union_stairs(a, b, rr, nn); break;
      case FT_NODE_UNION_SOFT: d = op_union_soft(a, b, rr); break;
      case FT_NODE_INTERSECT_ROUND: d = op_intersect_round(a, b, rr); break;
      case FT_NODE_INTERSECT_CHAMFER: d = op_intersect_chamfer(a, b, rr); break;
      case FT_NODE_INTERSECT_COLUMNS: d = -op_union_columns(-a, -b, rr, nn); break;
      case FT_NODE_INTERSECT_STAIRS: d = -op_union_stairs(-a, -b, rr, nn); break;
      case FT_NODE_DIFF_ROUND: d = op_intersect_round(a, -b, rr); break;
      case FT_NODE_DIFF_CHAMFER: d = op_intersect_chamfer(a, -b, rr); break;
      case FT_NODE_DIFF_COLUMNS: d = -op_union_columns(-a, b, rr, nn); break;
      case FT_NODE_DIFF_STAIRS: d = -op_union_stairs(-a, b, rr, nn); break;
      case FT_NODE_PIPE: d = fsqrt(a * a + b * b) - fabsf(rr); break;
      case FT_NODE_ENGRAVE: d = fmaxf(a, (a + fabsf(rr) - fabsf(b)) * kInvSqrt2); break;
      case FT_NODE_GROOVE: d = fmaxf(a, fminf(a + fabsf(rr), fabsf(nn) - fabsf(b))); break;
      default: d = fminf(a, fmaxf(a - fabsf(rr), fabsf(b) - fabsf(nn))); break;  // TONGUE
    }
    sd[l] = d;
    if (n.flags & 2) slb[l] = aabb_lb(p, n.bmin, n.bmax);
    // blend_surface_normals (renderer.rs:4525-4552)
    const float tt = blend_t(n, a, b);
    const float3 left = ns[il];
    const float3 right = blend_class(kind) == 2 ? mul(ns[ir], -1.0f) : ns[ir];
    const float3 bn = add(mul(left, 1.0f - tt), mul(right, tt));
    ns[il] = len(bn) > 1.0e-6f ? normalize(bn) : (tt < 0.5f ? left : right);
    if (ni == split_node) {
      // resolve_split_material_from_node (renderer.rs:7061-7120): the split node's children
      // keep their references (only unary nodes lie above it, so no later blend touches
      // their pool entries, and the walk material above it is not used)
      w.a = ms[il];
      w.b = ms[ir];
      w.split_t = tt;
      w.split = true;
      continue;
    }
    blend_refs(slot, pool, ms[il], ms[ir], tt, p, hc);
  }
  w.normal = ns[0];
  if (!w.split) w.a = ms[0];
  w.mat_overflow = pool.overflow != 0 || !finish_ref(slot, pool, w.a, p, hc) ||
                   (w.split && !finish_ref(slot, pool, w.b, p, hc));
}

#if defined(FT_EMITTED) && FT_EMITTED_SWEEP
// Provided by the scene emitter: shade_sweep_uniform of the root as straight-line code.
__device__ __forceinline__ void ftx_sweep(const DevScene& S, int slot, float3 p, float eps, ShadeWalk& w,
                                          const HookCtx* hc);
#endif

// The node where the reference's root walks for the hit at p begin to differ per leaf: the
// leaf they reach over a hard-union spine, or the root where the scene has no accel leaves
// or its leaves cannot be trusted to hold the walk (root_walks).
__device__ __forceinline__ int walk_start(int slot, const Hit& hit) {
  const DevScene& S = scene_of(slot);
  if (hit.leaf < 0 || S.root_walks) return S.root;
  if (S.accel == FT_ACCEL_NAIVE || S.n_bvh == 0 || S.n_leaves == 1) return S.leaves[hit.leaf].node;
  const int li = bvh_walk_leaf(slot, hit.pos);
  return li >= 0 ? S.leaves[li].node : S.root;
}

// split_material / walk_material from `start` (the reference's per-node walks), their
// parameter sets stored in the emptied pool (at most four entries).
__device__ __noinline__ void walks_to_refs(int slot, int start, float3 p, ShadeWalk& w, const HookCtx* hc) {
  Mat ma, mb;
  w.split = split_material(slot, start, p, ma, mb, w.split_t, hc);
  if (!w.split) walk_material(slot, start, p, ma, 0, hc);
  w.pool.used = 0u;
  w.pool.overflow = 0;
  w.a = ref_of_mat(w.pool, ma);
  if (w.split) w.b = ref_of_mat(w.pool, mb);
  w.mat_overflow = false;
}

// The three walks from `start` for one hit: hard booleans first (per lane, the walks'
// side choice), then the sweep of the node reached, lanes grouped by that node.
__device__ __noinline__ void shade_walks(int slot, int start, float3 p, float eps, ShadeWalk& w,
                                        const HookCtx* hc) {
  const DevScene& S = scene_of(slot);
  int node = start;
  float sign = 1.0f;
  for (;;) {
    const DNode& n = S.nodes[node];
    if (!is_hard(n.kind)) break;
    const float dl = node_dist(slot, n.lhs, p), dr = node_dist(slot, n.rhs, p);
    if (hard_left(n.kind, dl, dr)) {
      node = n.lhs;
    } else {
      if (n.kind == FT_NODE_SUBTRACT) sign = -sign;
      node = n.rhs;
    }
  }
  if (S.nodes[node].walk_split == -2 || !S.nodes[node].walk_ok) {
    w.normal = walk_normal(slot, start, p, eps);
    walks_to_refs(slot, start, p, w, hc);
    return;
  }
#if defined(FT_EMITTED) && FT_EMITTED_SWEEP
  if (node == S.root) {
    ftx_sweep(S, slot, p, eps, w, hc);
  } else
#endif
  {
    for_each_uniform(node, [&](int lead) { shade_sweep_uniform(S, slot, lead, p, eps, w, hc); });
  }
  w.normal = mul(w.normal, sign);
  if (w.mat_overflow) walks_to_refs(slot, start, p, w, hc);
}

// ---------------------------------------------------------------- BSDFs
constexpr float kPi = 3.14159265358979323846f;
__device__ __forceinline__ float ggx_d(float ndoth, float alpha) {
  float a2 = alpha * alpha, n2 = ndoth * ndoth;
  float denom = fmaxf(n2 * (a2 - 1.0f) + 1.0f, 1.0e-5f);
  return a2 / (kPi * denom * denom);
}
__device__ __forceinline__ float smith_g1(float x, float alpha) {
  float k = (alpha * alpha) * 0.5f;
  return x / fmaxf(x * (1.0f - k) + k, 1.0e-5f);
}
// metal.rs:7-21
__device__ inline float3 metal_eval(float3 color, float roughness, float3 n, float3 wi, float3 wo) {
  float ndotl = fmaxf(dot(n, wi), 0.0f), ndotv = fmaxf(dot(n, wo), 0.0f);
  if (ndotl <= 0.0f || ndotv <= 0.0f) return mk(0, 0, 0);
  float3 h = normalize(add(wi, wo));
  float ndoth = fmaxf(dot(n, h), 0.0f), vdoth = fmaxf(dot(wo, h), 0.0f);
  float r = clampf(roughness, 0.02f, 1.0f);
  float alpha = r * r;
  float d = ggx_d(ndoth, alpha);
  float g = smith_g1(ndotl, alpha) * smith_g1(ndotv, alpha);
  float m = powi5(1.0f - clampf(vdoth, 0.0f, 1.0f));
  float3 fres = add(add(color, mk(m, m, m)), mul(color, -m));
  return mul(fres, (d * g) / fmaxf(4.0f * ndotl * ndotv, 1.0e-5f));
}
__device__ __forceinline__ float f0_from_ior(float ior) {
  float i = fmaxf(ior, 1.0e-3f);
  float r = (i - 1.0f) / (i + 1.0f);
  return r * r;
}
// standard.rs:7-45
__device__ inline float3 standard_eval(const DParams& m, float3 n, float3 wi, float3 wo) {
  float ndotl = fmaxf(dot(n, wi), 0.0f), ndotv = fmaxf(dot(n, wo), 0.0f);
  if (ndotl <= 0.0f || ndotv <= 0.0f) return mk(0, 0, 0);
  float metallic = clampf(m.metallic, 0, 1), transmission = clampf(m.transmission, 0, 1);
  float clearcoat = clampf(m.clearcoat, 0, 1);
  float mw = clampf(1.0f - clampf(ndotv, 0, 1), 0, 1);
  float m2 = mw * mw;
  float coat_cover = clampf(clearcoat * (0.2f + 0.8f * (m2 * m2 * mw)) * 0.65f, 0.0f, 0.95f);
  float base_kd = clampf((1.0f - metallic) * (1.0f - transmission), 0, 1) * (1.0f - coat_cover);
  float3 diffuse = mul(v3(m.color), base_kd / kPi);
  float f0d = clampf(f0_from_ior(m.ior), 0, 1);
  float3 base_f0 = cmul(mul(mk(f0d, f0d, f0d), clampf(m.specular, 0, 1) * clampf(m.specular_weight, 0, 1)),
                        v3(m.specular_color));
  float3 spec_f0 = add(mul(base_f0, 1.0f - metallic), mul(v3(m.color), metallic));
  float3 spec = mul(metal_eval(spec_f0, m.roughness, n, wi, wo), 1.0f - coat_cover * 0.85f);
  float3 cc = mk(0, 0, 0);
  if (clearcoat > 1.0e-5f)
    cc = mul(metal_eval(mk(1, 1, 1), clampf(m.clearcoat_roughness, 0.02f, 1.0f), n, wi, wo), clearcoat * 0.18f);
  return add(add(diffuse, spec), cc);
}
__device__ inline float3 model_eval(int model, const DParams& p, float3 n, float3 wi, float3 wo) {
  switch (model) {
    case TAG_STANDARD: return standard_eval(p, n, wi, wo);
    case TAG_LAMBERT: return dot(n, wi) <= 0.0f ? mk(0, 0, 0) : mul(v3(p.color), 1.0f / kPi);
    case TAG_METAL: return metal_eval(v3(p.color), p.roughness, n, wi, wo);
    default: return mk(0, 0, 0);
  }
}
__device__ inline float3 mat_eval(const Mat& m, float3 n, float3 wi, float3 wo) {
  if (!m.blend) return model_eval(m.a_model, m.a, n, wi, wo);
  return lerp_spec(model_eval(m.a_model, m.a, n, wi, wo), model_eval(m.b_model, m.b, n, wi, wo), m.t);
}
__device__ __forceinline__ float3 emission_of(const DParams& p) { return mul(v3(p.emission_color), fmaxf(p.emission_strength, 0.0f)); }
__device__ inline float3 mat_emission(const Mat& m) {
  if (!m.blend) return emission_of(m.a);
  return lerp_spec(emission_of(m.a), emission_of(m.b), m.t);
}
// the same on a material reference (materials/mod.rs:291-305), parameter sets read on use
__device__ inline float3 ref_eval(int slot, const MatPool& pool, const MRef& m, float3 n, float3 wi, float3 wo) {
  DParams a;
  load_ref(slot, pool, m.ha, a);
  const float3 fa = model_eval(m.a_model, a, n, wi, wo);
  if (!m.blend) return fa;
  load_ref(slot, pool, m.hb, a);
  return lerp_spec(fa, model_eval(m.b_model, a, n, wi, wo), m.t);
}
__device__ inline float3 ref_emission(int slot, const MatPool& pool, const MRef& m) {
  DParams a;
  load_ref(slot, pool, m.ha, a);
  const float3 ea = emission_of(a);
  if (!m.blend) return ea;
  load_ref(slot, pool, m.hb, a);
  return lerp_spec(ea, emission_of(a), m.t);
}

// --------------------------------------------------------------- lights
// offset_ray_origin (renderer.rs:5067-5078)
__device__ __forceinline__ float3 offset_origin(float3 pos, float3 gn, float3 dir, float eps) {
  float3 n = len(gn) > 1.0e-8f ? normalize(gn) : normalize(dir);
  float sign = dot(dir, n) >= 0.0f ? 1.0f : -1.0f;
  float3 no = mul(n, sign * fmaxf(eps * 16.0f, 2.0e-4f));
  return add(add(pos, no), mul(normalize(dir), fmaxf(eps * 8.0f, 1.0e-4f)));
}

// shadow_visibility (renderer.rs:4706-4733)
__device__ __noinline__ float shadow_visibility(int slot, float3 o, float3 d, float max_t, float eps) {
  const DevScene& S = scene_of(slot);
  float t = eps * 2.0f, vis = 1.0f;
  const float lb_gate = fmaxf(eps * 8.0f, 0.02f);
  for (int k = 0; k < 80; ++k) {
    if (t >= max_t) return clampf(vis, 0.0f, 1.0f);
    float3 p = add(o, mul(d, t));
    float lower = fabsf(scene_lower_bound(slot, p, lb_gate));
    float h = lower > lb_gate ? lower : fabsf(scene_dist(slot, p));
    if (h < eps * 4.0f) return 0.0f;
    vis = fminf(vis, clampf(10.0f * h / t, 0.0f, 1.0f));
    t += fmaxf(h, eps * 2.0f);
  }
  return clampf(vis, 0.0f, 1.0f);
}

// render_api.rs:171-289
__device__ inline void sample_light(const DLight& L, float3 point, unsigned idx, unsigned count, float3& wi,
                                    float3& rad, float& dist) {
  if (L.kind == FT_LIGHT_ENV) {
    wi = mk(0.0f, 1.0f, 0.0f);
    rad = v3(L.rad);
    dist = __builtin_huge_valf();
    return;
  }
  float3 target = v3(L.pos);
  if (L.kind == FT_LIGHT_SPHERE && L.radius > 1.0e-5f) {
    float3 normal = normalize(sub(point, target));
    float3 up = fabsf(normal.y) < 0.99f ? mk(0, 1, 0) : mk(1, 0, 0);
    float3 tangent = normalize(cross(normal, up));
    float3 bitangent = normalize(cross(tangent, normal));
    float u1 = ((float)idx + 0.5f) / (float)(count < 1 ? 1 : count);
    unsigned bits = (idx >> 16) | (idx << 16);
    bits = ((bits & 0x55555555u) << 1) | ((bits & 0xAAAAAAAAu) >> 1);
    bits = ((bits & 0x33333333u) << 2) | ((bits & 0xCCCCCCCCu) >> 2);
    bits = ((bits & 0x0F0F0F0Fu) << 4) | ((bits & 0xF0F0F0F0u) >> 4);
    bits = ((bits & 0x00FF00FFu) << 8) | ((bits & 0xFF00FF00u) >> 8);
    float u2 = (float)bits * 2.3283064e-10f;
    float sx = 2.0f * u1 - 1.0f, sy = 2.0f * u2 - 1.0f;
    float dx = 0.0f, dy = 0.0f;
    if (!(fabsf(sx) <= 1.1920929e-7f && fabsf(sy) <= 1.1920929e-7f)) {
      float r, th;
      if (fabsf(sx) > fabsf(sy)) { r = sx; th = 0.78539816339744830962f * (sy / sx); }
      else { r = sy; th = 1.57079632679489661923f - 0.78539816339744830962f * (sx / sy); }
      dx = r * cosf(th);
      dy = r * sinf(th);
    }
    target = add(add(target, mul(tangent, dx * L.radius)), mul(bitangent, dy * L.radius));
  }
  float3 to = sub(target, point);
  dist = fmaxf(len(to), 1.0e-4f);
  wi = mul(to, 1.0f / dist);
  rad = mul(v3(L.rad), 1.0f / (dist * dist));
}

// shade_color (renderer.rs:4217-4276); `hk` is the hook entry of the material's dominant
// set (-1: none)
__device__ FT_SHADE_FN float3 shade_color(int slot, const MRef& mat, const MatPool& pool, int hk, float3 hit_pos,
                                          float3 hit_n, bool front, float3 nrm, float3 wo, float eps,
                                          float max_dist, unsigned obj, unsigned mat_id, float cur_ior) {
  const DevScene& S = scene_of(slot);
  if (S.n_lights == 0) return mk(1, 1, 1);
  // eval_bsdf (renderer.rs:4905-4920): while the material keeps its dynamic id, its eval hook
  // replaces the BSDF, with the context of bsdf_context_value (renderer.rs:5003-5030)
  const int ev = hk >= 0 ? S.hooks[hk].eval : -1;
  float3 args[FT_HOOK_PARAMS];
  if (ev >= 0) {
    args[0] = hit_pos;
    args[1] = object_local(slot, obj, hit_pos);
    args[2] = nrm;
    args[3] = wo;
    args[4] = splat3(front ? 1.0f : 0.0f);
    args[5] = splat3((float)obj);
    args[6] = splat3((float)mat_id);
    args[7] = wo;
    args[9] = splat3(cur_ior);
    for (int k = 10; k < (int)FT_HOOK_PARAMS; ++k) args[k] = mk(0.0f, 0.0f, 0.0f);
  }
  float3 n = normalize(nrm);
  float3 gn = front ? normalize(hit_n) : normalize(mul(hit_n, -1.0f));
  float3 color = mk(0, 0, 0);
  for (int li = 0; li < S.n_lights; ++li) {
    const DLight& L = S.lights[li];
    unsigned ns = L.kind == FT_LIGHT_SPHERE ? (L.samples < 1 ? 1u : L.samples) : 1u;
    float inv = 1.0f / (float)ns;
    for (unsigned si = 0; si < ns; ++si) {
      float3 wi, rad;
      float dist;
      sample_light(L, hit_pos, si, ns, wi, rad, dist);
      wi = normalize(wi);
      float ndotl = n.x * wi.x + n.y * wi.y + n.z * wi.z;
      if (ndotl <= 0.0f) continue;
      float3 so = offset_origin(hit_pos, gn, wi, eps);
      float max_t = isfinite(dist) ? fmaxf(dist - eps * 8.0f, 0.0f) : max_dist;
      if (max_t <= 0.0f) continue;
      float sh = shadow_visibility(slot, so, wi, max_t, eps);
      if (sh <= 0.0f) continue;
      float3 f;
      if (ev >= 0) {
        args[8] = wi;
        f = eval_fn(slot, ev, args);
      } else {
        f = ref_eval(slot, pool, mat, nrm, wi, wo);
      }
      color = add(color, mul(cmul(f, rad), ndotl * sh * inv));
    }
  }
  return color;
}

// --------------------------------------------------------------- Whitted
struct Medium {
  float ior, ax, ay, az, density;
};
__device__ __forceinline__ float3 transmittance(const Medium& m, float dist) {  // renderer.rs:5109-5125
  if (m.density <= 0.0f || dist <= 0.0f) return mk(1, 1, 1);
  float ex = m.density * fmaxf(dist, 0.0f);
  return mk(powf(clampf(m.ax, 0, 1), ex), powf(clampf(m.ay, 0, 1), ex), powf(clampf(m.az, 0, 1), ex));
}
__device__ __forceinline__ float fresnel_dielectric(float cos_i, float eta_i, float eta_t) {  // renderer.rs:5186-5204
  float ei = fmaxf(eta_i, 1.0e-4f), et = fmaxf(eta_t, 1.0e-4f);
  if (fabsf(ei - et) < 1.0e-6f) return 0.0f;
  float ci = clampf(cos_i, 0.0f, 1.0f);
  float eta = ei / et;
  float sin2t = eta * eta * (1.0f - ci * ci);
  if (sin2t >= 1.0f) return 1.0f;
  float ct = sqrtf(1.0f - sin2t);
  float rs = ((ei * ci) - (et * ct)) / fmaxf((ei * ci) + (et * ct), 1.0e-6f);
  float rp = ((et * ci) - (ei * ct)) / fmaxf((et * ci) + (ei * ct), 1.0e-6f);
  return 0.5f * (rs * rs + rp * rp);
}
__device__ __forceinline__ float3 reflect(float3 v, float3 n) { return sub(v, mul(n, 2.0f * dot(v, n))); }
__device__ __forceinline__ bool refract(float3 incident, float3 normal, float eta, float3& out) {  // renderer.rs:5080-5095
  float3 i = normalize(incident), n = normalize(normal);
  float cosi = clampf(dot(i, n), -1.0f, 1.0f);
  if (cosi > 0.0f) n = mul(n, -1.0f); else cosi = -cosi;
  float sin2t = eta * eta * (1.0f - cosi * cosi);
  if (sin2t > 1.0f) return false;
  float cost = sqrtf(1.0f - sin2t);
  out = normalize(add(mul(i, eta), mul(n, eta * cosi - cost)));
  return true;
}

struct RayEntry {
  float ox, oy, oz, dx, dy, dz;
  float tr, tg, tb;
  float m_ior, m_ax, m_ay, m_az, m_density;
  int depth;
  int owner;  // lane of the wave whose sample this ray contributes to
};
static_assert(sizeof(RayEntry) == 64, "RayEntry layout");

struct TraceParams {
  int max_steps, max_depth;
  float eps, step_scale, max_dist;
};

template <int kRayStack>
struct RayStack {
  RayEntry e[kRayStack];
  int sp;
  int overflow;
  int owner;      // owner of the rays pushed while the current ray is processed
  int max_depth;  // rays at this depth return black (ray.rs:12-14) and are never stored
  __device__ void push(float3 o, float3 d, float3 thr, const Medium& m, int depth) {
    if (depth >= max_depth) return;
    if (sp >= kRayStack) { overflow = 1; return; }
    RayEntry& r = e[sp++];
    r.ox = o.x; r.oy = o.y; r.oz = o.z; r.dx = d.x; r.dy = d.y; r.dz = d.z;
    r.tr = thr.x; r.tg = thr.y; r.tb = thr.z;
    r.m_ior = m.ior; r.m_ax = m.ax; r.m_ay = m.ay; r.m_az = m.az; r.m_density = m.density;
    r.depth = depth;
    r.owner = owner;
  }
};

__device__ __forceinline__ float3 environment(int slot, float3 dir) {
  const DevScene& S = scene_of(slot);
  if (S.env_fn >= 0) return eval_fn(slot, S.env_fn, &dir);
  return v3(S.env_rad);
}

// rough_lobe_directions (ray.rs:383-417)
template <int kRayStack>
__device__ inline void push_lobes(RayStack<kRayStack>& st, float3 hp, float3 gn, float3 ideal_dir, float rough,
                                  int count, float3 thr, const Medium& med, int depth, float eps) {
  float3 ideal = normalize(ideal_dir);
  float3 seed = fabsf(ideal.y) < 0.99f ? mk(0, 1, 0) : mk(1, 0, 0);
  float3 tangent = normalize(cross(ideal, seed));
  float3 bitangent = normalize(cross(tangent, ideal));
  float spread = clampf(rough, 0.0f, 0.5f) * 0.35f;
  int n = (count <= 1 || spread <= 1.0e-4f) ? 1 : (count > 7 ? 7 : count);
  const float ox[6] = {1.0f, -1.0f, 0.0f, 0.0f, 0.707f, -0.707f};
  const float oy[6] = {0.0f, 0.0f, 1.0f, -1.0f, 0.707f, 0.707f};
  float inv = 1.0f / (float)n;
  float3 w = mul(thr, inv);
  for (int k = 0; k < n; ++k) {
    float3 d = k == 0 ? ideal : normalize(add(add(ideal, mul(tangent, ox[k - 1] * spread)), mul(bitangent, oy[k - 1] * spread)));
    st.push(offset_origin(hp, gn, d, eps * (8.0f + rough * 12.0f)), d, w, med, depth + 1);
  }
}

// trace_hit_with_material (ray.rs:41-222) for one material of a hit; returns the local
// term (already weighted by `thr`) and pushes the children.
template <int kRayStack>
__device__ FT_SHADE_FN float3 hit_material(int slot, const TraceParams& tp, const Hit& hit, float3 dir,
                                           float3 nrm, const Medium& med, int depth, const MRef& mat,
                                           const MatPool& pool, float3 thr, RayStack<kRayStack>& st) {
  const float3 hp = hit.pos;
  const float3 normal = hit.front ? normalize(hit.n) : normalize(mul(hit.n, -1.0f));
  const float3 wo = normalize(mul(dir, -1.0f));
  DParams pr;
  load_ref(slot, pool, dom_handle(mat), pr);
  const float rough = clampf(pr.roughness, 0.0f, 1.0f);
  float transmission, ior, metallic, clearcoat, spec_strength;
  bool thin;
  float3 spec_color, base_color = v3(pr.color);
  switch (dom_model(mat)) {
    case TAG_STANDARD:
      transmission = clampf(pr.transmission, 0, 1); ior = clampf(pr.ior, 1, 3); thin = pr.thin_walled != 0.0f;
      spec_color = v3(pr.specular_color); metallic = clampf(pr.metallic, 0, 1); clearcoat = clampf(pr.clearcoat, 0, 1);
      spec_strength = clampf(pr.specular * pr.specular_weight, 0, 1);
      break;
    case TAG_LAMBERT:
      transmission = 0; ior = 1; thin = false; spec_color = mk(0, 0, 0); metallic = 0; clearcoat = 0; spec_strength = 0;
      break;
    case TAG_METAL:
      transmission = 0; ior = 1; thin = false; spec_color = base_color; metallic = 1; clearcoat = 0; spec_strength = 1;
      break;
    default:
      transmission = 1; ior = clampf(pr.ior, 1, 3); thin = pr.thin_walled != 0.0f; spec_color = mk(1, 1, 1);
      metallic = 0; clearcoat = 0; spec_strength = 1;
      break;
  }
  float eta_i, eta_t, next_ior;
  if (thin) { eta_i = 1.0f; eta_t = 1.0f; next_ior = med.ior; }
  else if (hit.front) { eta_i = med.ior; eta_t = ior; next_ior = ior; }
  else { eta_i = med.ior; eta_t = 1.0f; next_ior = 1.0f; }
  const float cosv = fabsf(dot(normal, wo));
  float fresnel;
  if (transmission > 1.0e-4f) fresnel = fresnel_dielectric(cosv, eta_i, eta_t);
  else {
    float f0 = clampf(f0_from_ior(ior), 0.0f, 1.0f);
    fresnel = clampf(f0 + (1.0f - f0) * powi5(1.0f - clampf(cosv, 0.0f, 1.0f)), 0.0f, 1.0f);
  }
  float reflect_weight;
  if (transmission > 1.0e-4f) reflect_weight = clampf(fresnel, 0.0f, 1.0f);
  else {
    float base_reflect = metallic > 0.5f ? 1.0f : clampf(spec_strength, 0.0f, 1.0f);
    reflect_weight = clampf(base_reflect * fresnel + metallic * 0.8f + clearcoat * 0.2f, 0.0f, 1.0f);
  }
  const float local_weight = clampf((1.0f - transmission) * (1.0f - reflect_weight), 0.0f, 1.0f);
  float3 color = mk(0, 0, 0);
  if (local_weight != 0.0f)
    color = mul(shade_color(slot, mat, pool, (int)pr.hook, hp, hit.n, hit.front, nrm, wo, tp.eps, tp.max_dist,
                            hit.obj, hit.mat, med.ior),
                local_weight);
  color = add(color, ref_emission(slot, pool, mat));

  if (reflect_weight > 1.0e-4f) {
    float3 tint = lerp_spec(spec_color, base_color, clampf(metallic, 0.0f, 1.0f));
    float3 w = mul(cmul(thr, tint), reflect_weight);
    if (transmission > 1.0e-4f && rough > 0.06f) {
      int n = rough < 0.12f ? 1 : (rough < 0.28f ? 2 : 3);
      push_lobes(st, hp, normal, normalize(reflect(dir, normal)), rough, n, w, med, depth, tp.eps);
    } else {
      float3 rd = normalize(reflect(dir, normal));
      st.push(offset_origin(hp, normal, rd, tp.eps * (1.0f + rough * 6.0f)), rd, w, med, depth + 1);
    }
  }
  if (transmission > 1.0e-4f) {
    Medium next = med;
    // transition_medium (renderer.rs:5153-5175): a Standard material changes the medium only
    // above transmission 0.01 (its own test, not the 1e-4 of the refraction branch)
    const int model = dom_model(mat);
    const bool changes_medium = model == TAG_DIELECTRIC || (model == TAG_STANDARD && pr.transmission > 0.01f);
    if (!thin && changes_medium) {
      if (!hit.front) next = Medium{1.0f, 1.0f, 1.0f, 1.0f, 0.0f};
      else if (pr.has_medium != 0.0f)
        next = Medium{clampf(pr.medium_ior, 1, 3), pr.medium_absorption[0], pr.medium_absorption[1],
                      pr.medium_absorption[2], fmaxf(pr.medium_density, 0.0f)};
      else next = Medium{clampf(pr.ior, 1, 3), 1.0f, 1.0f, 1.0f, 0.0f};
    }
    next.ior = clampf(next_ior, 1.0f, 3.0f);
    float3 tt = mk(1.0f - (1.0f - clampf(base_color.x, 0, 1)) * transmission,
                   1.0f - (1.0f - clampf(base_color.y, 0, 1)) * transmission,
                   1.0f - (1.0f - clampf(base_color.z, 0, 1)) * transmission);
    float3 w = mul(cmul(thr, tt), transmission * (1.0f - fresnel));
    float3 rdir;
    if (refract(dir, normal, eta_i / eta_t, rdir)) {
      if (rough > 0.06f) {
        int n = rough < 0.1f ? 2 : (rough < 0.24f ? 3 : 4);
        push_lobes(st, hp, normal, normalize(rdir), rough, n, w, next, depth, tp.eps);
      } else {
        st.push(offset_origin(hp, normal, rdir, tp.eps * (8.0f + rough * 12.0f)), normalize(rdir), w, next,
                depth + 1);
      }
    }
  }
  return cmul(color, thr);
}

// Index of the k-th (0-based) set bit of m; m has more than k set bits.
__device__ __forceinline__ int kth_set_bit(unsigned long long m, int k) {
  int pos = 0;
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    const unsigned long long low = m & ((1ull << w) - 1ull);
    const int c = __popcll(low);
    if (k >= c) {
      k -= c;
      m >>= w;
      pos += w;
    } else {
      m = low;
    }
  }
  return pos;
}
__device__ __forceinline__ int lanes_below(unsigned long long m) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}
__device__ __forceinline__ int lane_read(int v, int lane) { return __builtin_amdgcn_ds_bpermute(lane << 2, v); }
__device__ __forceinline__ float lane_read(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(lane << 2, __float_as_int(v)));
}

// Per-wave radiance accumulators: one RGB triple per lane's sample (blocks of 64 x 4).
__shared__ float ft_whitted_acc[4][64 * 3];

// trace_ray_recursive (ray.rs:3-38) with explicit stacks, for the 64 samples of one wave.
// Every term of the recursion is linear in the child radiance, so the sample radiance is
// the sum over its ray tree of path throughput x local term, in any order.  Each lane keeps
// a stack of pending rays tagged with the lane that owns their sample; before every round
// lanes whose stack ran empty take the top ray of lanes holding two or more (a cross-lane
// exchange), so the wave's rays are traced by all of its lanes instead of by the lanes of
// the deepest trees alone.  Contributions go to the owner's accumulator in LDS.  Returns
// the radiance of this lane's sample.  All lanes of the wave that are active on entry
// must call it together.
template <int kRayStack>
// t_cached >= -1: the primary ray's march is known to end there (-1: a miss), from the batch
// depth pass; the trace then takes raymarch_cached's hit instead of marching it again.
__device__ __noinline__ float3 trace_whitted(int slot, const TraceParams& tp, float3 o, float3 d, int* overflow,
                                             float t_cached = -2.0f) {
  const DevScene& S = scene_of(slot);
  const int me = (int)threadIdx.x;
  float* const acc = ft_whitted_acc[threadIdx.y];
  acc[me * 3] = 0.0f;
  acc[me * 3 + 1] = 0.0f;
  acc[me * 3 + 2] = 0.0f;
  RayStack<kRayStack> st;
  st.sp = 0;
  st.overflow = 0;
  st.owner = me;
  st.max_depth = tp.max_depth;
  st.push(o, d, mk(1, 1, 1), Medium{1.0f, 1.0f, 1.0f, 1.0f, 0.0f}, 0);
  const unsigned long long active = __builtin_amdgcn_ballot_w64(true);
  for (;;) {
    const unsigned long long idle = __builtin_amdgcn_ballot_w64(st.sp == 0);
    if (idle == active) break;
    const unsigned long long rich = __builtin_amdgcn_ballot_w64(st.sp >= 2);
    if (idle != 0ull && rich != 0ull) {
      const int n = min(__popcll(idle), __popcll(rich));
      const int ri = lanes_below(idle), rr = lanes_below(rich);
      const bool take = st.sp == 0 && ri < n, give = st.sp >= 2 && rr < n;
      const int src = take ? kth_set_bit(rich, ri) : me;
      const RayEntry top = st.e[st.sp > 0 ? st.sp - 1 : 0];
      RayEntry got;
      got.ox = lane_read(top.ox, src); got.oy = lane_read(top.oy, src); got.oz = lane_read(top.oz, src);
      got.dx = lane_read(top.dx, src); got.dy = lane_read(top.dy, src); got.dz = lane_read(top.dz, src);
      got.tr = lane_read(top.tr, src); got.tg = lane_read(top.tg, src); got.tb = lane_read(top.tb, src);
      got.m_ior = lane_read(top.m_ior, src); got.m_ax = lane_read(top.m_ax, src);
      got.m_ay = lane_read(top.m_ay, src); got.m_az = lane_read(top.m_az, src);
      got.m_density = lane_read(top.m_density, src);
      got.depth = lane_read(top.depth, src); got.owner = lane_read(top.owner, src);
      if (give) --st.sp;
      if (take) st.e[st.sp++] = got;
    }
    if (st.sp == 0) continue;
    const RayEntry r = st.e[--st.sp];
    if (r.depth >= tp.max_depth) continue;
    st.owner = r.owner;
    const float3 ro = mk(r.ox, r.oy, r.oz), rd = mk(r.dx, r.dy, r.dz);
    const float3 thr = mk(r.tr, r.tg, r.tb);
    const Medium med{r.m_ior, r.m_ax, r.m_ay, r.m_az, r.m_density};
    const float min_t = r.depth == 0 ? 0.0f : tp.eps * 2.0f;
    float3 c;
    Hit hit;
    const bool h = r.depth == 0 && r.owner == me && t_cached >= -1.0f
                       ? raymarch_cached(slot, ro, rd, tp.eps, t_cached, hit)
                       : raymarch(slot, ro, rd, tp.max_steps, tp.eps, tp.step_scale, min_t, tp.max_dist, hit);
    if (!h) {
      c = cmul(cmul(environment(slot, rd), transmittance(med, tp.max_dist)), thr);
    } else {
      const int start = walk_start(slot, hit);
      // build_bsdf_context: tree-walk shading normal at eps 1e-4 (renderer.rs:4397-4449), and
      // the hit's material or split materials, from one shading sweep
      ShadeWalk w;
      const HookCtx hc{hit.pos, hit.front ? normalize(hit.n) : normalize(mul(hit.n, -1.0f)),
                       normalize(mul(rd, -1.0f)), hit.front ? 1.0f : 0.0f};
      shade_walks(slot, start, hit.pos, 1.0e-4f, w, &hc);
      const float3 sn = apply_normal_hooks(slot, hit.pos, hit.front, hit.obj, hit.mat, mul(rd, -1.0f),
                                           normalize(hit.front ? w.normal : mul(w.normal, -1.0f)));
      const float3 att = cmul(thr, transmittance(med, hit.t));
      c = mk(0, 0, 0);
      // the split's two materials, lerped by t (ray.rs:31-35; a side of weight 0 adds
      // nothing), or the walk material
      const float tt = w.split ? clampf(w.split_t, 0.0f, 1.0f) : 0.0f;
      for (int side = 0; side < (w.split ? 2 : 1); ++side) {
        const float wgt = !w.split ? 1.0f : (side == 0 ? 1.0f - tt : tt);
        if (wgt == 0.0f) continue;
        c = add(c, hit_material(slot, tp, hit, rd, sn, med, r.depth, side == 0 ? w.a : w.b, w.pool,
                                w.split ? mul(att, wgt) : att, st));
      }
    }
    atomicAdd(&acc[r.owner * 3], c.x);
    atomicAdd(&acc[r.owner * 3 + 1], c.y);
    atomicAdd(&acc[r.owner * 3 + 2], c.z);
  }
  if (st.overflow) *overflow = 1;
  return mk(acc[me * 3], acc[me * 3 + 1], acc[me * 3 + 2]);
}

// trace_ray_debug_aov (ray.rs:224-285)
__device__ __noinline__ float3 trace_aov(int slot, const TraceParams& tp, float3 o, float3 d, int aov) {
  const DevScene& S = scene_of(slot);
  Hit hit;
  if (!raymarch(slot, o, d, tp.max_steps, tp.eps, tp.step_scale, 0.0f, tp.max_dist, hit)) return mk(0, 0, 0);
  const int start = walk_start(slot, hit);
  Mat mat;
  const HookCtx hc{hit.pos, hit.front ? normalize(hit.n) : normalize(mul(hit.n, -1.0f)),
                   normalize(mul(d, -1.0f)), hit.front ? 1.0f : 0.0f};
  walk_material(slot, start, hit.pos, mat, 0, &hc);
  float3 normal = hit.front ? normalize(hit.n) : normalize(mul(hit.n, -1.0f));
  int model = dom_model(mat);
  float ior = (model == TAG_DIELECTRIC || model == TAG_STANDARD) ? clampf(dom_params(mat).ior, 1, 3) : 1.0f;
  switch (aov) {
    case FT_AOV_DEPTH: { float v = clampf(1.0f - hit.t / tp.max_dist, 0, 1); return mk(v, v, v); }
    case FT_AOV_NORMAL: return mk(normal.x * 0.5f + 0.5f, normal.y * 0.5f + 0.5f, normal.z * 0.5f + 0.5f);
    case FT_AOV_MATERIAL_ID: {
      unsigned x = hit.mat * 0x9E3779B9u + 0x7F4A7C15u;
      x ^= x >> 16;
      x *= 0x85EBCA6Bu;
      x ^= x >> 13;
      return mk((float)(x & 0xFF) / 255.0f, (float)((x >> 8) & 0xFF) / 255.0f, (float)((x >> 16) & 0xFF) / 255.0f);
    }
    case FT_AOV_IOR: { float v = clampf((ior - 1.0f) / 2.0f, 0, 1); return mk(v, v, v); }
    case FT_AOV_TRANSMISSION: { float v = (model == TAG_DIELECTRIC || model == TAG_STANDARD) ? 1.0f : 0.0f; return mk(v, v, v); }
    case FT_AOV_FRESNEL: {
      float3 wo = normalize(mul(d, -1.0f));
      float f = fresnel_dielectric(fabsf(dot(normal, wo)), 1.0f, ior);
      return mk(f, f, f);
    }
    default: { float v = clampf(hit.t / tp.max_dist, 0, 1); return mk(v, v, v); }
  }
}

// spectrum_to_rgb8 / spectrum_to_rgb8_reinhard (renderer.rs:5221-5239)
__device__ __forceinline__ unsigned char to_u8(float v) {
  float mapped = fminf(powf(fmaxf(v, 0.0f), 1.0f / 2.2f), 1.0f);
  float s = mapped * 255.0f;
  if (!(s > 0.0f)) return 0;
  if (s >= 255.0f) return 255;
  return (unsigned char)s;
}
__device__ __forceinline__ unsigned char to_u8_reinhard(float v) {
  float x = fmaxf(v, 0.0f);
  return to_u8(x / (1.0f + x));
}

}  // namespace ftd
