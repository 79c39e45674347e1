// ft_device.h — device-side scene layout and evaluation for the gfx950 backend.
//
// Scene data lives in HBM as flat arrays built once per scene by ft_scene_build.cpp:
//   DNode     one per SdfNode (reference renderer.rs:170-355), with the per-primitive
//             rotation pre-resolved to the sin/cos pairs `to_local` (renderer.rs:6216-6240)
//             would compute on every call (same angles, same libm), and sdf_bounds
//             (renderer.rs:5241-5386) per node;
//   prog      the node tree flattened to a post-order instruction list.  Every lane of a
//             wavefront walks the same list, so control flow over nodes is wave-uniform
//             and the value stack pointer is uniform (the stack lives in VGPRs);
//             modifier nodes (domain / distance_post / noise) get a BEGIN instruction
//             ahead of their subtree that moves the evaluation point;
//   BVH       accel leaves (maximal non-Union subtrees, renderer.rs:659-670) under a
//             median-split BVH (renderer.rs:693-732), traversed near-first with the
//             reference's pruning rule (renderer.rs:743-822).
// One ray per lane; all arithmetic f32.
#pragma once

#if !defined(__HIPCC_RTC__)
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "ft_abi.h"
#include "ft_layout.h"

namespace ftd {

struct DevScene {
  const DNode* __restrict__ nodes;
  const unsigned* __restrict__ prog;
  const DOp* __restrict__ ops;           // prog as records (plus one padding record)
  const DParams* __restrict__ mats;
  const DHooks* __restrict__ hooks;      // per-hit material hooks (DParams::hook)
  const float* __restrict__ objects;     // 9 floats per object: centre, sin/cos z,y,x
  const DLight* __restrict__ lights;
  const ft_function* __restrict__ fns;
  const ft_expr* __restrict__ exprs;
  const DBvh* __restrict__ bvh;
  const DLeaf* __restrict__ leaves;
  const unsigned* __restrict__ lbprog;   // lower-bound program (post-order)
  int n_nodes, root, n_prog, n_lbprog;
  int n_mats, n_objects, n_lights, env_fn;
  int accel, n_leaves, n_bvh, stack_depth;   // stack_depth: LDS value-stack entries per thread
  int simple;                            // no custom SDF, modifier or noise node
  unsigned long long* flops;             // non-null: count evaluated FLOPs (instrumented runs)
  float bmin[3], bmax[3];                // scene bounds
  float env_rad[3];                      // sum of EnvLight radiance
  // sampled distance atlas (optional): a grid of cells of 8^3 voxels; cells near the
  // surface own a brick of 9^3 distance samples (corners of the voxels), the others keep
  // the exact distance at their centre
  const int* __restrict__ atlas_cells;   // per cell: brick index, -1 (far: use atlas_cell_d), -2 (exact)
  const float* __restrict__ atlas_cell_d;
  const void* __restrict__ atlas_bricks; // 729 samples per brick, FP16 or FP32
  unsigned long long* atlas_samples;     // non-null: count coarse samples (instrumented runs)
  float atlas_origin[3], atlas_voxel;
  int atlas_dims[3], atlas_fp16;
  int atlas_on;
  float atlas_switch;                    // exact evaluation below this distance (2 voxels)
  // the per-hit walks start at the root instead of the hit's accelerator leaf: the scene has
  // a node whose distance can fall below the bounds the accelerators prune with
  int root_walks;
};

// Scene descriptors live in constant memory, one slot per scene handle.  Device functions
// receive the slot index and re-read it with readfirstlane, so every scene field is a
// scalar (SGPR) load and node programs run with wave-uniform control flow.
constexpr int kMaxScenes = 1024;
}  // namespace ftd
extern "C" {
__constant__ ftd::DevScene ft_c_scenes[ftd::kMaxScenes];
}
namespace ftd {

__device__ __forceinline__ const DevScene& scene_of(int slot) {
  return ft_c_scenes[__builtin_amdgcn_readfirstlane(slot)];
}

// A wave-uniform pointer re-typed into the constant address space, so loads through it
// with uniform indices are scalar (s_load) instead of per-lane vector loads.
#if defined(__HIP_DEVICE_COMPILE__)
#define FT_CONST_AS __attribute__((address_space(4)))
#else
#define FT_CONST_AS
#endif
template <class T>
__device__ __forceinline__ const FT_CONST_AS T* uniform_ptr(const T* p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (const FT_CONST_AS T*)(((unsigned long long)hi << 32) | lo);
}
// Copies a record out of constant-address-space memory field by field (the implicit copy
// constructor only binds generic references).
template <class T>
__device__ __forceinline__ T load_uniform(const FT_CONST_AS T* p) {
  T out;
  const FT_CONST_AS unsigned* src = (const FT_CONST_AS unsigned*)p;
  unsigned* dst = (unsigned*)&out;
#pragma unroll
  for (unsigned k = 0; k < sizeof(T) / 4; ++k) dst[k] = src[k];
  return out;
}

}  // namespace ftd
#include "ft_math.h"
namespace ftd {


// to_local (renderer.rs:6216-6222)
__device__ __forceinline__ float3 node_local(const DNode& n, float3 p) {
  float3 q = mk(p.x - n.cx, p.y - n.cy, p.z - n.cz);
  if (n.flags & 1) {
    q = rot_z(q, n.szs, n.szc);
    q = rot_y(q, n.sys, n.syc);
    q = rot_x(q, n.sxs, n.sxc);
  }
  return q;
}
__device__ __forceinline__ float3 object_local(int slot, unsigned obj, float3 p) {
  const DevScene& S = scene_of(slot);
  if (obj >= (unsigned)S.n_objects) return p;
  const float* o = S.objects + 9 * obj;
  float3 q = mk(p.x - o[0], p.y - o[1], p.z - o[2]);
  q = rot_z(q, o[3], o[4]);
  q = rot_y(q, o[5], o[6]);
  q = rot_x(q, o[7], o[8]);
  return q;
}

#if defined(FT_EMITTED)
__device__ float3 ftx_fn(int fn, const float3* params);
#endif

__device__ __noinline__ float3 eval_fn(int slot, int fn, const float3* params) {
#if defined(FT_EMITTED) && FT_EMITTED_FUNCTIONS
  return ftx_fn(fn, params);
#endif
  const DevScene& S = scene_of(slot);
  const ft_function F = S.fns[fn];
  float3 v[kMaxExpr];
  const int count = (int)F.count;
  for (int k = 0; k < count; ++k) {
    const ft_expr e = S.exprs[F.first + k];
    float3 a = e.a >= 0 ? v[e.a] : mk(0, 0, 0);
    float3 b = e.b >= 0 ? v[e.b] : mk(0, 0, 0);
    float3 c = e.c >= 0 ? v[e.c] : mk(0, 0, 0);
    float3 r;
    switch (e.op) {
      case FT_EXPR_CONST: r = mk(e.value, e.value, e.value); break;
      case FT_EXPR_PARAM: r = params[e.a]; break;
      case FT_EXPR_VEC3: r = mk(a.x, b.x, c.x); break;
      case FT_EXPR_GET_X: r = mk(a.x, a.x, a.x); break;
      case FT_EXPR_GET_Y: r = mk(a.y, a.y, a.y); break;
      case FT_EXPR_GET_Z: r = mk(a.z, a.z, a.z); break;
      case FT_EXPR_NEG: r = mk(-a.x, -a.y, -a.z); break;
      case FT_EXPR_ADD: r = add(a, b); break;
      case FT_EXPR_SUB: r = sub(a, b); break;
      case FT_EXPR_MUL: r = cmul(a, b); break;
      case FT_EXPR_DIV: r = mk(a.x / b.x, a.y / b.y, a.z / b.z); break;
      case FT_EXPR_ABS: r = mk(fabsf(a.x), fabsf(a.y), fabsf(a.z)); break;
      case FT_EXPR_FLOOR: r = mk(floorf(a.x), floorf(a.y), floorf(a.z)); break;
      case FT_EXPR_CEIL: r = mk(ceilf(a.x), ceilf(a.y), ceilf(a.z)); break;
      case FT_EXPR_FRACT: r = mk(fractf_(a.x), fractf_(a.y), fractf_(a.z)); break;
      case FT_EXPR_SQRT: r = mk(sqrtf(a.x), sqrtf(a.y), sqrtf(a.z)); break;
      case FT_EXPR_SIN: r = mk(sinf(a.x), sinf(a.y), sinf(a.z)); break;
      case FT_EXPR_COS: r = mk(cosf(a.x), cosf(a.y), cosf(a.z)); break;
      case FT_EXPR_SATURATE: r = mk(saturatef_(a.x), saturatef_(a.y), saturatef_(a.z)); break;
      case FT_EXPR_MIN: r = mk(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); break;
      case FT_EXPR_MAX: r = mk(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); break;
      case FT_EXPR_POW: r = mk(powf(a.x, b.x), powf(a.y, b.y), powf(a.z, b.z)); break;
      case FT_EXPR_CLAMP: r = mk(fclamp_any(a.x, b.x, c.x), fclamp_any(a.y, b.y, c.y), fclamp_any(a.z, b.z, c.z)); break;
      case FT_EXPR_MIX: r = mk(lerpf(a.x, b.x, c.x), lerpf(a.y, b.y, c.y), lerpf(a.z, b.z, c.z)); break;
      case FT_EXPR_STEP: { float s = b.x < a.x ? 0.0f : 1.0f; r = mk(s, s, s); break; }
      case FT_EXPR_SMOOTHSTEP: r = splat3(smoothstep_builtin(a.x, b.x, c.x)); break;
      case FT_EXPR_DOT: { float s = dot(a, b); r = mk(s, s, s); break; }
      case FT_EXPR_LENGTH: { float s = len(a); r = mk(s, s, s); break; }
      case FT_EXPR_NORMALIZE: r = normalize_div(a); break;
      case FT_EXPR_ROTATE_X: r = rotate_deg_x(a, b.x); break;
      case FT_EXPR_ROTATE_Y: r = rotate_deg_y(a, b.x); break;
      case FT_EXPR_ROTATE_Z: r = rotate_deg_z(a, b.x); break;
      case FT_EXPR_SPHERE_DIST: { float s = len(a) - b.x; r = mk(s, s, s); break; }
      case FT_EXPR_BOX_DIST: r = splat3(box_dist_builtin(a, b)); break;
      case FT_EXPR_CYLINDER_DIST: r = splat3(cylinder_dist_builtin(a, b.x, c.x)); break;
      case FT_EXPR_TORUS_DIST: r = splat3(torus_dist_builtin(a, b.x, c.x)); break;
      case FT_EXPR_VALUE_NOISE: { float s = value_noise_3d(a); r = mk(s, s, s); break; }
      case FT_EXPR_FBM: {
        float3 d = e.d >= 0 ? v[e.d] : mk(0, 0, 0);
        float s = fbm_3d(a, b.x > 0.0f ? (int)b.x : 0, c.x, d.x);
        r = mk(s, s, s);
        break;
      }
      default: r = mk(0, 0, 0); break;
    }
    v[(int)e.pad] = r;  // the node's value slot (assigned at scene build)
  }
  return v[(int)S.exprs[F.first + count - 1].pad];
}

// eval_custom_sdf_distance (renderer.rs:6139-6179)
__device__ inline float custom_distance(int slot, const DNode& n, float3 p) {
  const DevScene& S = scene_of(slot);
  float3 dp = p;
  if (n.fn_domain >= 0) dp = eval_fn(slot, n.fn_domain, &p);
  float d = n.fn_distance >= 0 ? eval_fn(slot, n.fn_distance, &dp).x : 1.0e6f;
  if (n.fn_post >= 0) {
    float3 args[2] = {mk(d, d, d), dp};
    d = eval_fn(slot, n.fn_post, args).x;
  }
  return d;
}

__device__ __forceinline__ float prim_distance(const DNode& n, float3 p) {
  float3 q = node_local(n, p);
  switch (n.kind) {
    case FT_NODE_SPHERE: return sdf_sphere(q, n.p[0], n.p[1]);
    case FT_NODE_BOX: return sdf_box(q, n.p[0], n.p[1], n.p[2], n.p[3], n.p[4]);
    case FT_NODE_CYLINDER: return sdf_cylinder(q, n.p[0], n.p[1], n.p[2], n.p[3]);
    case FT_NODE_TORUS: return sdf_torus(q, n.p[0], n.p[1]);
    default: return sdf_ngon(q, (int)n.p[0], n.p[1], n.p[2], n.p[3], n.p[4]);
  }
}

// ------------------------------------------------------ node program
struct Info {
  float d, lb;
  unsigned obj, mat;
  int leaf;
};

__device__ __forceinline__ bool is_prim(int k) { return k <= FT_NODE_EXTRUDE_POLYGON; }

// primitive distance from a program record (prim_distance on the record's copy of the node)
__device__ __forceinline__ float op_prim_distance(const DOp& o, int kind, float3 p) {
  float3 q = mk(p.x - o.c[0], p.y - o.c[1], p.z - o.c[2]);
  if (o.code & kOpRotated) {
    q = rot_z(q, o.r[0], o.r[1]);
    q = rot_y(q, o.r[2], o.r[3]);
    q = rot_x(q, o.r[4], o.r[5]);
  }
  switch (kind) {
    case FT_NODE_SPHERE: return sdf_sphere(q, o.p[0], o.p[1]);
    case FT_NODE_BOX: return sdf_box(q, o.p[0], o.p[1], o.p[2], o.p[3], o.p[4]);
    case FT_NODE_CYLINDER: return sdf_cylinder(q, o.p[0], o.p[1], o.p[2], o.p[3]);
    case FT_NODE_TORUS: return sdf_torus(q, o.p[0], o.p[1]);
    default: return sdf_ngon(q, (int)o.p[0], o.p[1], o.p[2], o.p[3], o.p[4]);
  }
}
// the record's AABB lower bound (booleans, slice)
__device__ __forceinline__ float op_aabb_lb(const DOp& o, float3 p) {
  return aabb_lb_lit(p, o.c[0], o.c[1], o.c[2], o.r[0], o.r[1], o.r[2]);
}
// the blend booleans of sdf_distance_info (renderer.rs:5733-5953); keep_l: the ids come
// from the left operand
__device__ __forceinline__ float smooth_op(int kind, float a, float b, float rr, float nn, bool& keep_l) {
  switch (kind) {
    case FT_NODE_UNION_ROUND: keep_l = a <= b; return op_union_round(a, b, rr);
    case FT_NODE_UNION_CHAMFER: keep_l = a <= b; return op_union_chamfer(a, b, rr);
    case FT_NODE_UNION_COLUMNS: keep_l = a <= b; return op_union_columns(a, b, rr, nn);
    case FT_NODE_UNION_STAIRS: keep_l = a <= b; return op_union_stairs(a, b, rr, nn);
    case FT_NODE_UNION_SOFT: keep_l = a <= b; return op_union_soft(a, b, rr);
    case FT_NODE_INTERSECT_ROUND: keep_l = a >= b; return op_intersect_round(a, b, rr);
    case FT_NODE_INTERSECT_CHAMFER: keep_l = a >= b; return op_intersect_chamfer(a, b, rr);
    case FT_NODE_INTERSECT_COLUMNS: keep_l = a >= b; return -op_union_columns(-a, -b, rr, nn);
    case FT_NODE_INTERSECT_STAIRS: keep_l = a >= b; return -op_union_stairs(-a, -b, rr, nn);
    case FT_NODE_DIFF_ROUND: keep_l = a >= -b; return op_intersect_round(a, -b, rr);
    case FT_NODE_DIFF_CHAMFER: keep_l = a >= -b; return op_intersect_chamfer(a, -b, rr);
    case FT_NODE_DIFF_COLUMNS: keep_l = a >= -b; return -op_union_columns(-a, b, rr, nn);
    case FT_NODE_DIFF_STAIRS: keep_l = a >= -b; return -op_union_stairs(-a, b, rr, nn);
    case FT_NODE_PIPE: keep_l = true; return fsqrt(a * a + b * b) - fabsf(rr);
    case FT_NODE_ENGRAVE: keep_l = true; return fmaxf(a, (a + fabsf(rr) - fabsf(b)) * kInvSqrt2);
    case FT_NODE_GROOVE: keep_l = true; return fmaxf(a, fminf(a + fabsf(rr), fabsf(nn) - fabsf(b)));
    default: keep_l = true; return fminf(a, fmaxf(a - fabsf(rr), fabsf(b) - fabsf(nn)));  // TONGUE
  }
}

// Value stack of the node program: three LDS planes (distance, lower bound, id of the
// primitive node that produced the value), one column per thread, so a push/pop is one
// bank-conflict-free ds access even when lanes evaluate different programs.
extern __shared__ float ft_lds[];

// Evaluates the program range [pf, pl] at p: sdf_distance_info + sdf_lower_bound of the
// subtree root (renderer.rs:5413-5955).
// `base`: first value-stack entry this evaluation may use (entries below belong to a
// caller that is itself walking a node program).
// kSimple: the scene has no custom SDF, modifier or noise node, so the loop calls no
// function and keeps its state in registers.
template <bool kSimple = false>
__device__ __forceinline__ Info eval_uniform(const DevScene& S, int slot, int node, float3 p, int base = 0) {
  // the bounds and every record are wave-uniform; readfirstlane makes that visible to the
  // compiler so record fetches are scalar loads.  Records are streamed in program order and
  // the next one is requested before the current one is evaluated.
  const auto nodes = uniform_ptr(S.nodes);
  const auto ops = uniform_ptr(S.ops);
  const int pf = __builtin_amdgcn_readfirstlane(S.nodes[node].prog_first);
  const int pl = __builtin_amdgcn_readfirstlane(S.nodes[node].prog_last);
  const int nth = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  float* __restrict__ sd = ft_lds + base * nth + tid;
  float* __restrict__ slb = ft_lds + (S.stack_depth + base) * nth + tid;
  int* __restrict__ sid = reinterpret_cast<int*>(ft_lds + (2 * S.stack_depth + base) * nth) + tid;
  float3 pst[kMaxPointStack];
  int sp = 0, pp = 0;
  DOp o = load_uniform(ops + pf);
  for (int k = pf; k <= pl; ++k) {
    // read the current record first, then request the next one: the wait for the current
    // record's load then no longer covers the next, whose latency hides behind this node
    const int code = o.code;
    const int kind = code & kOpKindMask;
    asm volatile("" ::"s"(code));
    __builtin_amdgcn_sched_barrier(0);
    const DOp next = load_uniform(ops + k + 1);
    if (!kSimple && (code & (int)kBeginFlag)) {
      const DNode n = load_uniform(nodes + o.node);
      pst[pp++] = p;
      float3 q = node_local(n, p);
      if (n.kind == FT_NODE_DOMAIN_MODIFIER && n.fn_domain >= 0) q = eval_fn(slot, n.fn_domain, &q);
      p = q;
    } else if (is_prim(kind)) {
      const float d = op_prim_distance(o, kind, p);
      sd[sp * nth] = d;
      slb[sp * nth] = d;
      sid[sp * nth] = o.node;
      ++sp;
    } else if (!kSimple && kind == FT_NODE_CUSTOM) {
      const DNode n = load_uniform(nodes + o.node);
      float3 q = node_local(n, p);
      float mn[3] = {-n.p[0], -n.p[1], -n.p[2]}, mx[3] = {n.p[0], n.p[1], n.p[2]};
      sd[sp * nth] = custom_distance(slot, n, q);
      slb[sp * nth] = aabb_lb(q, mn, mx);
      sid[sp * nth] = o.node;
      ++sp;
    } else if (!kSimple &&
               (kind == FT_NODE_DOMAIN_MODIFIER || kind == FT_NODE_DISTANCE_POST_MODIFIER || kind == FT_NODE_NOISE)) {
      const int t = (sp - 1) * nth;
      const DNode n = load_uniform(nodes + o.node);
      float3 q = p;
      p = pst[--pp];
      float d = sd[t];
      if (kind == FT_NODE_DISTANCE_POST_MODIFIER && n.fn_post >= 0) {
        float3 args[2] = {mk(d, d, d), q};
        d = eval_fn(slot, n.fn_post, args).x;
      } else if (kind == FT_NODE_NOISE) {
        d = apply_noise(q, d, (int)n.p[0], n.p[1], n.p[2]);
      }
      sd[t] = d;
      if (n.flags & 2) slb[t] = aabb_lb(p, n.bmin, n.bmax);
    } else if (kind == FT_NODE_SLICE) {
      const int t = (sp - 1) * nth;
      float c = comp(p, (int)o.p[0]);
      sd[t] = fmaxf(sd[t], fmaxf(o.p[1] - c, c - o.p[2]));
      if (code & kOpNeedsLb) slb[t] = op_aabb_lb(o, p);
    } else if (kind == FT_NODE_SMOOTH) {
      const int t = (sp - 1) * nth;
      sd[t] -= o.p[0] * 0.1f;
      slb[t] -= o.p[0] * 0.1f;
    } else {
      // binary: lhs at l, rhs at r
      const int r = (sp - 1) * nth, l = r - nth;
      --sp;
      const float a = sd[l], b = sd[r];
      if (kind == FT_NODE_UNION) {
        // lower-bound short circuit (renderer.rs:5711-5731) evaluated eagerly
        const float llb = slb[l], rlb = slb[r];
        const bool take_l = llb <= rlb ? (a <= rlb || a <= b) : (!(b <= llb) && a <= b);
        if (!take_l) { sd[l] = b; sid[l] = sid[r]; }
        slb[l] = fminf(llb, rlb);
      } else if (kind == FT_NODE_INTERSECT) {
        if (!(a >= b)) { sd[l] = b; sid[l] = sid[r]; }
        slb[l] = fmaxf(slb[l], slb[r]);
      } else if (kind == FT_NODE_SUBTRACT) {
        if (!(a >= -b)) { sd[l] = -b; sid[l] = sid[r]; }
      } else {
        bool keep_l;
        const float d = smooth_op(kind, a, b, o.p[0], o.p[1], keep_l);
        if (!keep_l) sid[l] = sid[r];
        sd[l] = d;
        if (code & kOpNeedsLb) slb[l] = op_aabb_lb(o, p);
      }
    }
    o = next;
  }
  if (S.flops) atomicAdd(S.flops, (unsigned long long)S.nodes[S.prog[pl] & ~kBeginFlag].cost);
  Info out;
  out.d = sd[0];
  out.lb = slb[0];
  const DNode& w = S.nodes[sid[0]];
  out.obj = w.obj;
  out.mat = w.mat;
  out.leaf = w.leaf;
  return out;
}

// The same program walk for scenes without custom, modifier or noise nodes whose value
// stack is at most kRegStack deep, with the stack in registers: the stack pointer is
// wave-uniform, so the arrays are indexed through M0 (s_set_gpr_idx) and stay in VGPRs.
// Without LDS accesses in the loop the scalar loads of the next record are the only
// LGKM-counted operations, so their wait no longer waits for a stack read as well.
// kInfo = false leaves out the plane of producing primitives (distance-only evaluations).
constexpr int kRegStack = 16;
template <bool kInfo>
__device__ __forceinline__ Info eval_uniform_reg(const DevScene& S, int node, float3 p) {
  const auto ops = uniform_ptr(S.ops);
  const int pf = __builtin_amdgcn_readfirstlane(S.nodes[node].prog_first);
  const int pl = __builtin_amdgcn_readfirstlane(S.nodes[node].prog_last);
  float sd[kRegStack], slb[kRegStack];
  int sid[kRegStack];
  for (int i = 0; i < kRegStack; ++i) {
    sd[i] = 0.0f;
    slb[i] = 0.0f;
    if (kInfo) sid[i] = 0;
  }
  int sp = 0;
  DOp o = load_uniform(ops + pf);
  for (int k = pf; k <= pl; ++k) {
    const int code = o.code;
    const int kind = code & kOpKindMask;
    asm volatile("" ::"s"(code));
    __builtin_amdgcn_sched_barrier(0);
    const DOp next = load_uniform(ops + k + 1);
    sp = __builtin_amdgcn_readfirstlane(sp);
    // Every node reads the top two entries and writes one entry (a push, or the new top):
    // one indexed read per operand and one indexed write per plane, after the kind's
    // branch, so the planes are never copied to merge the branches.
    const int i1 = (sp - 1) & (kRegStack - 1), i0 = (sp - 2) & (kRegStack - 1);
    const float d1 = sd[i1], d0 = sd[i0], lb1 = slb[i1], lb0 = slb[i0];
    const int id1 = kInfo ? sid[i1] : 0, id0 = kInfo ? sid[i0] : 0;
    int wt;
    float vd, vlb;
    int vid;
    if (is_prim(kind)) {
      wt = sp & (kRegStack - 1);
      vd = op_prim_distance(o, kind, p);
      vlb = vd;
      vid = o.node;
      ++sp;
    } else if (kind >= FT_NODE_UNION_ROUND && kind <= FT_NODE_TONGUE) {
      // the blend joins first: with the primitives they are nearly every node of a program
      wt = i0;
      --sp;
      bool keep_l;
      vd = smooth_op(kind, d0, d1, o.p[0], o.p[1], keep_l);
      vid = keep_l ? id0 : id1;
      vlb = (code & kOpNeedsLb) ? op_aabb_lb(o, p) : lb0;
    } else if (kind == FT_NODE_SLICE) {
      wt = i1;
      const float c = comp(p, (int)o.p[0]);
      vd = fmaxf(d1, fmaxf(o.p[1] - c, c - o.p[2]));
      vlb = (code & kOpNeedsLb) ? op_aabb_lb(o, p) : lb1;
      vid = id1;
    } else if (kind == FT_NODE_SMOOTH) {
      wt = i1;
      vd = d1 - o.p[0] * 0.1f;
      vlb = lb1 - o.p[0] * 0.1f;
      vid = id1;
    } else {
      // hard booleans: lhs at i0, rhs at i1
      wt = i0;
      --sp;
      const float a = d0, b = d1;
      if (kind == FT_NODE_UNION) {
        // lower-bound short circuit (renderer.rs:5711-5731) evaluated eagerly
        const bool take_l = lb0 <= lb1 ? (a <= lb1 || a <= b) : (!(b <= lb0) && a <= b);
        vd = take_l ? a : b;
        vid = take_l ? id0 : id1;
        vlb = fminf(lb0, lb1);
      } else if (kind == FT_NODE_INTERSECT) {
        const bool keep_l = a >= b;
        vd = keep_l ? a : b;
        vid = keep_l ? id0 : id1;
        vlb = fmaxf(lb0, lb1);
      } else {  // FT_NODE_SUBTRACT
        const bool keep_l = a >= -b;
        vd = keep_l ? a : -b;
        vid = keep_l ? id0 : id1;
        vlb = lb0;
      }
    }
    sd[wt] = vd;
    slb[wt] = vlb;
    if (kInfo) sid[wt] = vid;
    o = next;
  }
  if (S.flops) atomicAdd(S.flops, (unsigned long long)S.nodes[S.prog[pl] & ~kBeginFlag].cost);
  Info out;
  out.d = sd[0];
  out.lb = slb[0];
  if (kInfo) {
    const DNode& w = S.nodes[sid[0]];
    out.obj = w.obj;
    out.mat = w.mat;
    out.leaf = w.leaf;
  } else {
    out.obj = 0u;
    out.mat = 0u;
    out.leaf = -1;
  }
  return out;
}

// True, for every active lane, when at least half of the active lanes hold v.
__device__ __forceinline__ bool wave_majority(bool v) {
  const unsigned long long yes = __builtin_amdgcn_ballot_w64(v), all = __builtin_amdgcn_ballot_w64(true);
  return __builtin_amdgcn_readfirstlane(2 * __popcll(yes) >= __popcll(all) ? 1 : 0) != 0;
}

// Runs f(key) once per distinct value of `key` among the active lanes, each time for the
// lanes holding that value (a waterfall).  The leader's value is read with a lane index
// from the ballot, so the compiler cannot mistake `key` for a uniform value.
template <class F>
__device__ __forceinline__ void for_each_uniform(int key, F&& f) {
  unsigned long long pending = __builtin_amdgcn_ballot_w64(true);
  while (pending) {
    const int leader = (int)__builtin_ctzll(pending);
    const int lead = __builtin_amdgcn_readlane(key, leader);
    const bool mine = key == lead;
    pending &= ~__builtin_amdgcn_ballot_w64(mine);
    if (mine) f(lead);
  }
}

// Evaluates the subtree of `node` at p for every active lane.  Lanes may ask for different
// subtrees (BVH leaves, tree walks): the loop serves one distinct subtree per iteration, so
// each program run is wave-uniform.
__device__ __noinline__ Info eval_node(int slot, int node, float3 p, int base = 0) {
  const DevScene& S = scene_of(slot);
  const int b = __builtin_amdgcn_readfirstlane(base);
  Info out;
  if (S.simple && S.stack_depth <= kRegStack)
    for_each_uniform(node, [&](int lead) { out = eval_uniform_reg<true>(S, lead, p); });
  else if (S.simple)
    for_each_uniform(node, [&](int lead) { out = eval_uniform<true>(S, slot, lead, p, b); });
  else
    for_each_uniform(node, [&](int lead) { out = eval_uniform<false>(S, slot, lead, p, b); });
  return out;
}
// The distance (and lower bound) of `node`'s subtree alone: Info.obj / mat / leaf are not set.
// Inlined into its callers (the march loop, shadow rays): a call per distance evaluation
// would save and restore the callee-saved VGPRs it uses through scratch memory every step.
__device__ __forceinline__ Info eval_node_dist(int slot, int node, float3 p, int base = 0) {
  const DevScene& S = scene_of(slot);
  if (!(S.simple && S.stack_depth <= kRegStack)) return eval_node(slot, node, p, base);
  Info out;
  for_each_uniform(node, [&](int lead) { out = eval_uniform_reg<false>(S, lead, p); });
  return out;
}

// sdf_lower_bound of the root (renderer.rs:5413-5543): post-order over the nodes reachable
// through Union/Intersect/Subtract/Smooth; other nodes are terminals.
// The same as a call, for the shading walks, which evaluate subtrees a few times per hit:
// inlining the interpreter there would only raise the register pressure of the sweep.
__device__ __noinline__ float node_dist(int slot, int node, float3 p, int base = 0) {
  return eval_node_dist(slot, node, p, base).d;
}

__device__ __noinline__ float lower_bound_root(int slot, float3 p) {
  const DevScene& S = scene_of(slot);
  const int nth = blockDim.x * blockDim.y;
  float* __restrict__ st0 = ft_lds + threadIdx.y * blockDim.x + threadIdx.x;  // value-stack plane 0
  auto st = [&](int i) -> float& { return st0[i * nth]; };
  int sp = 0;
  for (int k = 0; k < S.n_lbprog; ++k) {
    const DNode& n = S.nodes[S.lbprog[k]];
    switch (n.kind) {
      case FT_NODE_UNION: --sp; st(sp - 1) = fminf(st(sp - 1), st(sp)); break;
      case FT_NODE_INTERSECT: --sp; st(sp - 1) = fmaxf(st(sp - 1), st(sp)); break;
      case FT_NODE_SUBTRACT: --sp; break;
      case FT_NODE_SMOOTH: st(sp - 1) -= n.p[0] * 0.1f; break;
      case FT_NODE_CUSTOM: {
        float3 q = node_local(n, p);
        float mn[3] = {-n.p[0], -n.p[1], -n.p[2]}, mx[3] = {n.p[0], n.p[1], n.p[2]};
        st(sp++) = aabb_lb(q, mn, mx);
        break;
      }
      default:
        st(sp++) = is_prim(n.kind) ? prim_distance(n, p) : aabb_lb(p, n.bmin, n.bmax);
        break;
    }
  }
  return st(0);
}

// accel_leaf_distance_info (renderer.rs:629-657) of leaf li, the same leaf for every lane
// (li is wave-uniform): record and node fetches are scalar loads.  far_shortcut = false
// evaluates a custom leaf exactly wherever p is (sdf_distance_info of the leaf's subtree).
template <bool kInfo = true>
__device__ __forceinline__ Info leaf_info_uniform(const DevScene& S, int slot, int li, float3 p,
                                                  bool far_shortcut = true) {
  const DLeaf L = load_uniform(uniform_ptr(S.leaves) + li);
  Info o;
  if (L.is_custom && far_shortcut) {
    const float lb = aabb_lb(p, L.bmin, L.bmax);
    const float ex = L.bmax[0] - L.bmin[0], ey = L.bmax[1] - L.bmin[1], ez = L.bmax[2] - L.bmin[2];
    const float far_t = clampf(sqrtf(ex * ex + ey * ey + ez * ez) * 0.1f, 0.05f, 0.5f);
    if (lb > far_t) {
      const auto& n = uniform_ptr(S.nodes)[L.node];
      o.d = lb; o.lb = lb; o.obj = n.obj; o.mat = n.mat; o.leaf = li;
      return o;
    }
  }
  if (L.op >= 0) {
    // a single-primitive leaf: one record, no program walk
    const DOp r = load_uniform(uniform_ptr(S.ops) + L.op);
    const auto& n = uniform_ptr(S.nodes)[L.node];
    o.d = op_prim_distance(r, r.code & kOpKindMask, p);
    o.lb = o.d;
    o.obj = n.obj;
    o.mat = n.mat;
    o.leaf = li;
    if (S.flops) atomicAdd(S.flops, (unsigned long long)n.cost);
    return o;
  }
  o = kInfo ? eval_node(slot, L.node, p) : eval_node_dist(slot, L.node, p);
  o.leaf = li;
  return o;
}

// Per-wave traversal stacks (blocks of 64 x 4 threads).
__shared__ int ft_bvh_stack[4][kMaxBvhStack];

// bvh_distance_info (renderer.rs:743-786) for the wave's 64 points at once.  The wave walks
// one stack of BVH nodes: a node is entered when its lower bound does not exceed some lane's
// best distance, each lane testing the reference's `node_lb > best` against its own best and
// evaluating only the leaves that pass it; of two children the one most lanes find nearer is
// visited first.  For a lane whose best is >= 0 the answer is the minimum over the leaves
// its own pruning leaves open, whatever the visiting order (which can only change which of
// two equally distant leaves names the hit).
//
// A negative best is order-dependent: every lower bound is >= 0, so the reference's
// depth-first walk (nearer child first, lhs on ties) prunes all that follows the first leaf
// with a negative distance, and returns that leaf.  Such a leaf's bounds contain p, and
// between two children whose bounds both contain p the reference goes left, so its answer
// is the leftmost such leaf of the tree: the least pre-order node index here (the tree is
// stored parent first, lhs subtree before rhs).  The bricks grid's shell search (renderer.rs:
// 1024-1129) visits p's cell first, its leaves in ascending leaf index, so there the answer is
// the least leaf index.  A lane with a negative best therefore keeps open the nodes whose
// bounds contain p and that can still hold an earlier leaf, and takes a negative leaf by
// rank.  Nodes and leaves are fetched with scalar loads; no per-lane stack.  All lanes of
// the wave that are active on entry must call it together.
// kInfo = false: the distance alone (Info.obj / mat are not set); that form is inlined
// into the march loop (see eval_node_dist), the full one is a call (once per ray).
template <bool kInfo>
__device__ __forceinline__ Info bvh_distance_body(int slot, float3 p) {
  const DevScene& S = scene_of(slot);
  const auto bvh = uniform_ptr(S.bvh);
  int* const stack = ft_bvh_stack[threadIdx.y];
  const bool grid_order = S.accel == FT_ACCEL_BRICKS;
  int sp = 0;
  float best = __builtin_huge_valf();
  int best_rank = 0x7fffffff;  // rank of the leaf holding a negative best
  Info bi;
  bool have = false;
  stack[sp++] = 0;
  while (sp > 0) {
    const int ni = __builtin_amdgcn_readfirstlane(stack[--sp]);
    const DBvh nd = load_uniform(bvh + ni);
    if (S.flops) atomicAdd(S.flops, 3ull * kAabbFlops);
    const float lb = aabb_lb(p, nd.bmin, nd.bmax);
    const bool need = !(lb > best) || (best < 0.0f && lb == 0.0f && (grid_order || ni < best_rank));
    if (__builtin_amdgcn_ballot_w64(need) == 0ull) continue;
    if (nd.left < 0) {
      if (need) {
        const int li = -nd.left - 1;
        const Info info = leaf_info_uniform<kInfo>(S, slot, li, p);
        const int rank = grid_order ? li : ni;
        const bool take = best < 0.0f ? (info.d < 0.0f && rank < best_rank) : info.d < best;
        if (take) {
          best = info.d;
          best_rank = rank;
          bi = info;
          have = true;
        } else if (!have) {
          bi = info;
          have = true;
        }
      }
      continue;
    }
    const DBvh L = load_uniform(bvh + nd.left);
    const DBvh R = load_uniform(bvh + nd.right);
    const float llb = aabb_lb(p, L.bmin, L.bmax), rlb = aabb_lb(p, R.bmin, R.bmax);
    const bool left_first = wave_majority(!need || llb <= rlb);
    stack[sp++] = left_first ? nd.right : nd.left;
    stack[sp++] = left_first ? nd.left : nd.right;
  }
  if (!have) bi = kInfo ? eval_node(slot, S.root, p) : eval_node_dist(slot, S.root, p);
  return bi;
}
__device__ __noinline__ Info bvh_distance_info(int slot, float3 p) { return bvh_distance_body<true>(slot, p); }
// The accel leaf that the reference's tree walks from the root reach at p
// (resolve_material_from_node renderer.rs:6945-7059, resolve_surface_normal_from_node
// renderer.rs:4462-4570).  Over a spine of hard unions each walk descends to the child of
// smaller exact distance, lhs on ties, so it ends at the leaf of least distance, the least
// leaf index among equals (leaves are numbered left to right).  That is not always the leaf
// bvh_distance_info names: inside several primitives the accelerators stop at the first
// negative leaf.  Here a node is entered while its bounds lie within max(best, 0), so every
// leaf whose bounds contain p is evaluated once the best is negative; custom leaves are
// evaluated exactly (no far shortcut).  Returns -1 only for an empty tree.  All lanes of the
// wave that are active on entry must call it together.
__device__ __noinline__ int bvh_walk_leaf(int slot, float3 p) {
  const DevScene& S = scene_of(slot);
  const auto bvh = uniform_ptr(S.bvh);
  int* const stack = ft_bvh_stack[threadIdx.y];
  int sp = 0;
  float best = __builtin_huge_valf();
  int leaf = -1;
  stack[sp++] = 0;
  while (sp > 0) {
    const int ni = __builtin_amdgcn_readfirstlane(stack[--sp]);
    const DBvh nd = load_uniform(bvh + ni);
    const bool need = aabb_lb(p, nd.bmin, nd.bmax) <= fmaxf(best, 0.0f);
    if (__builtin_amdgcn_ballot_w64(need) == 0ull) continue;
    if (nd.left < 0) {
      if (need) {
        const int li = -nd.left - 1;
        const float d = leaf_info_uniform(S, slot, li, p, false).d;
        if (leaf < 0 || d < best || (d == best && li < leaf)) {
          best = d;
          leaf = li;
        }
      }
      continue;
    }
    const DBvh L = load_uniform(bvh + nd.left);
    const DBvh R = load_uniform(bvh + nd.right);
    const float llb = aabb_lb(p, L.bmin, L.bmax), rlb = aabb_lb(p, R.bmin, R.bmax);
    const bool left_first = wave_majority(!need || llb <= rlb);
    stack[sp++] = left_first ? nd.right : nd.left;
    stack[sp++] = left_first ? nd.left : nd.right;
  }
  return leaf;
}

// bvh_lower_bound (renderer.rs:788-822) for the wave's points, one node stack per wave:
// the least AABB lower bound over the accel leaves.  A node is skipped when its bound
// exceeds the lane's best, which no leaf below it can beat, so the minimum does not depend
// on the visiting order, and the child most lanes find nearer goes first (the best falls
// early and prunes more).  Its one caller, the shadow march, takes the exact distance
// wherever the bound is at or below `stop`, so a lane whose best reaches `stop` is done:
// the result is the exact minimum where that exceeds `stop`, else some value <= `stop`.
// Inlined into the shadow march (a call per step would save and restore callee-saved VGPRs
// through scratch memory).
__device__ __forceinline__ float bvh_lower_bound(int slot, float3 p, float stop) {
  const DevScene& S = scene_of(slot);
  const auto bvh = uniform_ptr(S.bvh);
  int* const stack = ft_bvh_stack[threadIdx.y];
  int sp = 0;
  float best = __builtin_huge_valf();
  stack[sp++] = 0;
  while (sp > 0) {
    const DBvh nd = load_uniform(bvh + __builtin_amdgcn_readfirstlane(stack[--sp]));
    const float lb = aabb_lb(p, nd.bmin, nd.bmax);
    const bool need = !(lb > best) && best > stop;
    if (__builtin_amdgcn_ballot_w64(need) == 0ull) {
      if (__builtin_amdgcn_ballot_w64(best > stop) == 0ull) break;
      continue;
    }
    if (nd.left < 0) {
      if (need) best = fminf(best, lb);
      continue;
    }
    const DBvh L = load_uniform(bvh + nd.left);
    const DBvh R = load_uniform(bvh + nd.right);
    const float llb = aabb_lb(p, L.bmin, L.bmax), rlb = aabb_lb(p, R.bmin, R.bmax);
    const bool left_first = wave_majority(!need || llb <= rlb);
    stack[sp++] = left_first ? nd.right : nd.left;
    stack[sp++] = left_first ? nd.left : nd.right;
  }
  return best;
}

// Accelerator::distance_info / lower_bound (renderer.rs:609-623, 1024-1129).  Bricks uses
// the BVH traversal: both return the minimum over accel leaves (the grid's shell search
// visits a superset of the leaves that can beat the running best).
#if defined(FT_EMITTED)
// Provided by the scene emitter (runtime-compiled scene-specialised code).
__device__ __forceinline__ Info ftx_scene_info(const DevScene& S, float3 p);
__device__ __forceinline__ float ftx_scene_dist(const DevScene& S, float3 p);
__device__ __forceinline__ float ftx_lower_bound(const DevScene& S, float3 p);
#endif

__device__ __forceinline__ Info scene_info(int slot, float3 p) {
#if defined(FT_EMITTED) && FT_EMITTED_SCENE
  return ftx_scene_info(scene_of(slot), p);
#endif
  const DevScene& S = scene_of(slot);
  if (S.accel != FT_ACCEL_NAIVE && S.n_bvh > 0) return bvh_distance_info(slot, p);
  return eval_node(slot, S.root, p);
}
// The distance of scene_info alone (the evaluations skip the winning-primitive selects).
__device__ __forceinline__ float scene_dist(int slot, float3 p) {
#if defined(FT_EMITTED) && FT_EMITTED_SCENE
  return ftx_scene_dist(scene_of(slot), p);
#endif
  const DevScene& S = scene_of(slot);
  if (S.accel != FT_ACCEL_NAIVE && S.n_bvh > 0) return bvh_distance_body<false>(slot, p).d;
  return eval_node_dist(slot, S.root, p).d;
}
// `stop`: the caller only needs the bound where it exceeds stop (bvh_lower_bound).
__device__ __forceinline__ float scene_lower_bound(int slot, float3 p, float stop) {
#if defined(FT_EMITTED) && FT_EMITTED_SCENE
  return ftx_lower_bound(scene_of(slot), p);
#endif
  const DevScene& S = scene_of(slot);
  if (S.accel != FT_ACCEL_NAIVE && S.n_bvh > 0) return bvh_lower_bound(slot, p, stop);
  return lower_bound_root(slot, p);
}

// ---------------------------------------------------------------- march
// The march itself is ft_march.h; the scene evaluation it samples is scene_info.
struct SceneEval {
  int slot;
  __device__ __forceinline__ Info operator()(float3 p) const { return scene_info(slot, p); }
  __device__ __forceinline__ float dist(float3 p) const { return scene_dist(slot, p); }
};

}  // namespace ftd
#include "ft_march.h"
namespace ftd {

constexpr int kBrickSide = 9;                 // samples per brick edge (8 voxels)
constexpr int kBrickSamples = kBrickSide * kBrickSide * kBrickSide;

// A lower bound of |distance| from the atlas, or -1 where the lane must evaluate exactly.
// The scene distance is 1-Lipschitz, so inside a voxel of size h the trilinear value of the
// corner samples exceeds the true |distance| by at most the weighted corner distance,
// sqrt(3) h / 2; FP16 corners add a relative rounding of 2^-11 each.  Subtracting both
// makes every coarse step conservative: it never passes a surface the exact march would
// stop at.
__device__ __forceinline__ float atlas_distance(const DevScene& S, float3 p) {
  const float cell = 8.0f * S.atlas_voxel;
  const float qx = (p.x - S.atlas_origin[0]) / cell, qy = (p.y - S.atlas_origin[1]) / cell,
              qz = (p.z - S.atlas_origin[2]) / cell;
  const int cx = (int)floorf(qx), cy = (int)floorf(qy), cz = (int)floorf(qz);
  if (cx < 0 || cy < 0 || cz < 0 || cx >= S.atlas_dims[0] || cy >= S.atlas_dims[1] || cz >= S.atlas_dims[2])
    return -1.0f;
  const size_t idx = ((size_t)cz * S.atlas_dims[1] + cy) * S.atlas_dims[0] + cx;
  if (S.atlas_samples) atomicAdd(S.atlas_samples, 1ull);
  const int b = S.atlas_cells[idx];
  if (b < 0) return b == -1 ? fabsf(S.atlas_cell_d[idx]) - 0.8660254f * cell : -1.0f;
  // trilinear interpolation inside the brick
  const float lx = (qx - (float)cx) * 8.0f, ly = (qy - (float)cy) * 8.0f, lz = (qz - (float)cz) * 8.0f;
  const int ix = min(max((int)lx, 0), 7), iy = min(max((int)ly, 0), 7), iz = min(max((int)lz, 0), 7);
  const float fx = lx - (float)ix, fy = ly - (float)iy, fz = lz - (float)iz;
  const size_t base = (size_t)b * kBrickSamples + ((size_t)iz * kBrickSide + iy) * kBrickSide + ix;
  float c[8];
  if (S.atlas_fp16) {
    const _Float16* h = (const _Float16*)S.atlas_bricks + base;
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = (float)h[((k >> 2) & 1) * 81 + ((k >> 1) & 1) * 9 + (k & 1)];
  } else {
    const float* f = (const float*)S.atlas_bricks + base;
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = f[((k >> 2) & 1) * 81 + ((k >> 1) & 1) * 9 + (k & 1)];
  }
  const float x00 = c[0] + (c[1] - c[0]) * fx, x10 = c[2] + (c[3] - c[2]) * fx;
  const float x01 = c[4] + (c[5] - c[4]) * fx, x11 = c[6] + (c[7] - c[6]) * fx;
  const float y0 = x00 + (x10 - x00) * fy, y1 = x01 + (x11 - x01) * fy;
  const float v = fabsf(y0 + (y1 - y0) * fz);
  const float rel = S.atlas_fp16 ? 1.0f / 1024.0f : 1.0f / 4194304.0f;
  return fmaxf(v - 0.8660254f * S.atlas_voxel - v * rel, 0.0f);
}

struct AtlasCoarse {
  int slot;
  bool on;
  float switch_dist;
  __device__ __forceinline__ bool enabled() const { return on; }
  __device__ __forceinline__ float operator()(float3 p) const { return atlas_distance(scene_of(slot), p); }
};

__device__ __noinline__ bool raymarch(int slot, float3 o, float3 d, int max_steps, float eps,
                                      float step_scale, float min_t, float max_t, Hit& hit) {
  const DevScene& S = scene_of(slot);
  const SceneEval ev{slot};
  const AtlasCoarse coarse{slot, S.atlas_on != 0, S.atlas_switch};
  return raymarch_core(ev, S.bmin, S.bmax, o, d, max_steps, eps, step_scale, min_t, max_t, hit, coarse);
}
// raymarch's result for a primary ray whose march ended at t_cached (-1: missed), taken
// from the batch depth pass.
__device__ __noinline__ bool raymarch_cached(int slot, float3 o, float3 d, float eps, float t_cached, Hit& hit) {
  const SceneEval ev{slot};
  return hit_from_t(ev, o, d, eps, t_cached, hit);
}

}  // namespace ftd
