// ft_scene_build.cpp — builds the device scene layout from the C-ABI scene IR.
//
// Runs once per ft_scene_create (the reference does the equivalent work in
// Accelerator::from_scene, renderer.rs:1024-1129).  Produces:
//   * DNode records with sdf_bounds (renderer.rs:5241-5386) and the sin/cos pairs of the
//     per-primitive rotation, computed with the host libm from the same angles the
//     reference converts on every `to_local` call (renderer.rs:6216-6240);
//   * the post-order node program and its lower-bound variant;
//   * accel leaves (collect_accel_leaves, renderer.rs:659-670) and, for the BVH and
//     bricks modes, a median-split BVH over them (build_bvh, renderer.rs:693-732).
#include "ft_scene_build.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <functional>

namespace fth {
namespace {

using ftd::DNode;

struct B3 { float mn[3], mx[3]; };

inline float deg2rad(float d) { return d * (3.14159265358979323846f / 180.0f); }
inline float fmaxr(float a, float b) { return std::fmax(a, b); }
inline float fminr(float a, float b) { return std::fmin(a, b); }

B3 cube(const float c[3], float r) {
  B3 b;
  for (int k = 0; k < 3; ++k) { b.mn[k] = c[k] - r; b.mx[k] = c[k] + r; }
  return b;
}
B3 bunion(const B3& a, const B3& b) {
  B3 o;
  for (int k = 0; k < 3; ++k) { o.mn[k] = fminr(a.mn[k], b.mn[k]); o.mx[k] = fmaxr(a.mx[k], b.mx[k]); }
  return o;
}

// rotate_x/y/z with a degree angle (renderer.rs:6224-6240)
void rot(float v[3], int axis, float deg) {
  float r = deg2rad(deg), s = std::sin(r), c = std::cos(r);
  float x = v[0], y = v[1], z = v[2];
  if (axis == 0) { v[1] = c * y - s * z; v[2] = s * y + c * z; }
  else if (axis == 1) { v[0] = c * x + s * z; v[2] = -s * x + c * z; }
  else { v[0] = c * x - s * y; v[1] = s * x + c * y; }
}

bool is_binary(uint32_t k) { return k >= FT_NODE_UNION && k <= FT_NODE_TONGUE; }
bool is_unary(uint32_t k) {
  return k == FT_NODE_DOMAIN_MODIFIER || k == FT_NODE_DISTANCE_POST_MODIFIER || k == FT_NODE_NOISE ||
         k == FT_NODE_SLICE || k == FT_NODE_SMOOTH;
}
bool is_modifier(uint32_t k) {
  return k == FT_NODE_DOMAIN_MODIFIER || k == FT_NODE_DISTANCE_POST_MODIFIER || k == FT_NODE_NOISE;
}
bool is_blend(uint32_t k) {
  return is_binary(k) && k != FT_NODE_UNION && k != FT_NODE_INTERSECT && k != FT_NODE_SUBTRACT;
}

bool validate(const ft_scene_ir* ir, std::string& err) {
  if (!ir) { err = "scene IR is null"; return false; }
  if (ir->abi_version != FT_ABI_VERSION) { err = "scene IR ABI version mismatch"; return false; }
  if (ir->n_material_hooks > 0 && !ir->material_hooks) { err = "material hook table is null"; return false; }
  for (uint32_t h = 0; h < ir->n_material_hooks; ++h)
    if (ir->material_hooks[h].material >= ir->n_materials) { err = "material hook names no material"; return false; }
  if (ir->n_nodes == 0 || !ir->nodes || ir->root >= ir->n_nodes) {
    err = "no suitable scene root found (expected binding like scene/result/b/s)";
    return false;
  }
  if (ir->n_materials == 0 || !ir->materials) { err = "scene IR has no materials"; return false; }
  for (uint32_t i = 0; i < ir->n_nodes; ++i) {
    const ft_node& n = ir->nodes[i];
    if (n.kind >= FT_NODE_KIND_COUNT) { err = "unsupported node kind " + std::to_string(n.kind); return false; }
    if ((is_binary(n.kind) || is_unary(n.kind)) && (n.lhs < 0 || (uint32_t)n.lhs >= i)) {
      err = "node " + std::to_string(i) + ": child index must precede the node";
      return false;
    }
    if (is_binary(n.kind) && (n.rhs < 0 || (uint32_t)n.rhs >= i)) {
      err = "node " + std::to_string(i) + ": child index must precede the node";
      return false;
    }
    for (int32_t f : {n.fn_domain, n.fn_distance, n.fn_post})
      if (f >= (int32_t)ir->n_functions) { err = "node function index out of range"; return false; }
    if (n.kind == FT_NODE_SLICE && (n.param[0] < 0.0f || n.param[0] > 2.0f)) { err = "slice axis out of range"; return false; }
    // arity of the hooks as the kernels call them: domain(p), distance(p), distance_post(d, p)
    auto arity_ok = [&](int32_t f, uint32_t want) {
      return f < 0 || (f < (int32_t)ir->n_functions && ir->functions[f].n_params == want);
    };
    if (!arity_ok(n.fn_domain, 1) || !arity_ok(n.fn_distance, 1) || !arity_ok(n.fn_post, 2)) {
      err = "node " + std::to_string(i) + ": hook function has the wrong number of parameters";
      return false;
    }
  }
  if (ir->env_fn >= 0 && ir->env_fn < (int32_t)ir->n_functions && ir->functions[ir->env_fn].n_params != 1) {
    err = "environment color(dir) must take one parameter";
    return false;
  }
  for (uint32_t f = 0; f < ir->n_functions; ++f) {
    const ft_function& F = ir->functions[f];
    if (F.count == 0 || F.first + F.count > ir->n_exprs) {
      err = "function " + std::to_string(f) + " has an invalid or too long expression range";
      return false;
    }
    for (uint32_t k = 0; k < F.count; ++k) {
      const ft_expr& e = ir->exprs[F.first + k];
      if (e.op >= FT_EXPR_OP_COUNT) { err = "unsupported expression op"; return false; }
      if (e.op == FT_EXPR_PARAM) {
        if (e.a < 0 || (uint32_t)e.a >= F.n_params || (F.n_params > 2 && F.n_params != FT_HOOK_PARAMS)) {
          err = "expression parameter out of range";
          return false;
        }
        continue;
      }
      for (int32_t o : {e.a, e.b, e.c, e.d})
        if (o >= (int32_t)k) { err = "expression operand must precede its use"; return false; }
    }
  }
  if (ir->env_fn >= (int32_t)ir->n_functions) { err = "environment function index out of range"; return false; }
  if (ir->n_objects > 0 && !ir->object_transforms) { err = "object transforms missing"; return false; }
  if (ir->n_lights > 0 && !ir->lights) { err = "lights missing"; return false; }
  return true;
}

}  // namespace

bool build_scene(const ft_scene_ir* ir, uint32_t accel, HostScene& S, std::string& err) {
  if (!validate(ir, err)) return false;
  if (accel > FT_ACCEL_BRICKS) { err = "unknown accel mode"; return false; }
  const uint32_t N = ir->n_nodes;
  S.accel = (int)accel;
  S.root = (int)ir->root;
  S.env_fn = ir->env_fn;
  S.camera = ir->camera;
  S.nodes.assign(N, DNode{});

  // nodes: transforms and bounds (children precede parents, so one forward pass)
  std::vector<B3> bounds(N);
  for (uint32_t i = 0; i < N; ++i) {
    const ft_node& n = ir->nodes[i];
    DNode& d = S.nodes[i];
    std::memset(&d, 0, sizeof(d));
    d.kind = (int)n.kind;
    d.lhs = n.lhs;
    d.rhs = n.rhs;
    d.obj = n.object_id;
    d.mat = n.material_id;
    d.leaf = -1;
    d.fn_domain = n.fn_domain;
    d.fn_distance = n.fn_distance;
    d.fn_post = n.fn_post;
    d.cx = n.center[0]; d.cy = n.center[1]; d.cz = n.center[2];
    for (int k = 0; k < 6; ++k) d.p[k] = n.param[k];
    if (n.rot_deg[0] != 0.0f || n.rot_deg[1] != 0.0f || n.rot_deg[2] != 0.0f) {
      d.flags |= 1;
      float rz = deg2rad(-n.rot_deg[2]), ry = deg2rad(-n.rot_deg[1]), rx = deg2rad(-n.rot_deg[0]);
      d.szs = std::sin(rz); d.szc = std::cos(rz);
      d.sys = std::sin(ry); d.syc = std::cos(ry);
      d.sxs = std::sin(rx); d.sxc = std::cos(rx);
    } else {
      d.szc = d.syc = d.sxc = 1.0f;
    }
    const float* c = n.center;
    B3 b;
    switch (n.kind) {
      case FT_NODE_SPHERE: b = cube(c, n.param[0]); break;
      case FT_NODE_BOX: {
        float h0 = n.param[0], h1 = n.param[1], h2 = n.param[2];
        b = cube(c, std::sqrt(h0 * h0 + h1 * h1 + h2 * h2));
        break;
      }
      case FT_NODE_CYLINDER: b = cube(c, std::sqrt(n.param[0] * n.param[0] + n.param[1] * n.param[1])); break;
      case FT_NODE_TORUS: b = cube(c, n.param[0] + n.param[1]); break;
      case FT_NODE_EXTRUDE_POLYGON: b = cube(c, std::sqrt(n.param[1] * n.param[1] + n.param[2] * n.param[2])); break;
      case FT_NODE_CUSTOM: {
        for (int k = 0; k < 3; ++k) { b.mn[k] = INFINITY; b.mx[k] = -INFINITY; }
        for (int m = 0; m < 8; ++m) {
          float v[3] = {(m & 4) ? n.param[0] : -n.param[0], (m & 2) ? n.param[1] : -n.param[1],
                        (m & 1) ? n.param[2] : -n.param[2]};
          rot(v, 0, n.rot_deg[0]);
          rot(v, 1, n.rot_deg[1]);
          rot(v, 2, n.rot_deg[2]);
          for (int k = 0; k < 3; ++k) {
            float w = c[k] + v[k];
            b.mn[k] = fminr(b.mn[k], w);
            b.mx[k] = fmaxr(b.mx[k], w);
          }
        }
        break;
      }
      case FT_NODE_DOMAIN_MODIFIER:
      case FT_NODE_DISTANCE_POST_MODIFIER:
      case FT_NODE_NOISE:
        for (int k = 0; k < 3; ++k) { b.mn[k] = n.bounds_min[k]; b.mx[k] = n.bounds_max[k]; }
        if (std::isnan(b.mn[0])) b = bounds[n.lhs];
        break;
      case FT_NODE_SUBTRACT: b = bounds[n.lhs]; break;
      case FT_NODE_SLICE: {
        b = bounds[n.lhs];
        int ax = (int)n.param[0];
        b.mn[ax] = fmaxr(b.mn[ax], n.param[1]);
        b.mx[ax] = fminr(b.mx[ax], n.param[2]);
        if (b.mx[ax] < b.mn[ax]) b.mx[ax] = b.mn[ax];
        break;
      }
      case FT_NODE_SMOOTH: {
        b = bounds[n.lhs];
        float e = n.param[0] * 0.1f;
        for (int k = 0; k < 3; ++k) { b.mn[k] -= e; b.mx[k] += e; }
        break;
      }
      default: b = bunion(bounds[n.lhs], bounds[n.rhs]); break;
    }
    bounds[i] = b;
    for (int k = 0; k < 3; ++k) { d.bmin[k] = b.mn[k]; d.bmax[k] = b.mx[k]; }
  }
  for (int k = 0; k < 3; ++k) { S.bmin[k] = bounds[S.root].mn[k]; S.bmax[k] = bounds[S.root].mx[k]; }

  // sdf_center (renderer.rs:4047-4089)
  std::function<void(int, float*)> center = [&](int i, float* out) {
    const ft_node& n = ir->nodes[i];
    if (n.kind <= FT_NODE_NOISE) { for (int k = 0; k < 3; ++k) out[k] = n.center[k]; return; }
    switch (n.kind) {
      case FT_NODE_SUBTRACT: case FT_NODE_DIFF_ROUND: case FT_NODE_DIFF_CHAMFER: case FT_NODE_DIFF_COLUMNS:
      case FT_NODE_DIFF_STAIRS: case FT_NODE_PIPE: case FT_NODE_ENGRAVE: case FT_NODE_GROOVE:
      case FT_NODE_TONGUE: case FT_NODE_SLICE: case FT_NODE_SMOOTH:
        center(n.lhs, out);
        return;
      default: {
        float l[3], r[3];
        center(n.lhs, l);
        center(n.rhs, r);
        for (int k = 0; k < 3; ++k) out[k] = (l[k] + r[k]) * 0.5f;
      }
    }
  };
  center(S.root, S.center);

  // program (post-order, BEGIN before modifier subtrees) and its stack requirement
  std::vector<int> depth_at(N, 0);
  int sp = 0, pp = 0;
  std::function<void(int)> emit = [&](int i) {
    const ft_node& n = ir->nodes[i];
    S.nodes[i].prog_first = (int)S.prog.size();
    if (is_modifier(n.kind)) {
      S.prog.push_back((unsigned)i | ftd::kBeginFlag);
      ++pp;
      S.max_points = std::max(S.max_points, pp);
      emit(n.lhs);
      --pp;
    } else if (is_binary(n.kind)) {
      emit(n.lhs);
      emit(n.rhs);
      --sp;
    } else if (is_unary(n.kind)) {
      emit(n.lhs);
    } else {
      ++sp;
      S.max_stack = std::max(S.max_stack, sp);
    }
    S.nodes[i].prog_last = (int)S.prog.size();
    S.prog.push_back((unsigned)i);
  };
  emit(S.root);
  if (S.max_stack > ftd::kMaxStack) { err = "scene tree needs a deeper evaluation stack than supported"; return false; }
  if (S.max_points > ftd::kMaxPointStack) { err = "modifier nesting deeper than supported"; return false; }

  // nodes whose lower bound a parent reads (Union / Intersect children); a Smooth or
  // Subtract node's bound is its lhs's, so the need passes down to that lhs.  Parents follow
  // their children in the IR, so one backward pass reaches every level.
  for (uint32_t i = N; i-- > 0;) {
    const ft_node& n = ir->nodes[i];
    if (n.kind == FT_NODE_UNION || n.kind == FT_NODE_INTERSECT) {
      S.nodes[n.lhs].flags |= 2;
      S.nodes[n.rhs].flags |= 2;
    } else if ((n.kind == FT_NODE_SMOOTH || n.kind == FT_NODE_SUBTRACT) && (S.nodes[i].flags & 2)) {
      S.nodes[n.lhs].flags |= 2;
    }
  }
  for (uint32_t i = 0; i < N; ++i) {
    const uint32_t k = ir->nodes[i].kind;
    if (k == FT_NODE_CUSTOM || is_modifier(k)) S.simple = false;
    // hg_sdf's columns formulas (op_union_columns renderer.rs:6072-6099) run on negated
    // inputs for the intersection and difference, so far from the shape they return values
    // far below the distance to its bounds: the accelerators prune such a leaf where the
    // reference's root walk (resolve_material_from_node from setup.root) still enters it
    if (k == FT_NODE_UNION_COLUMNS || k == FT_NODE_INTERSECT_COLUMNS || k == FT_NODE_DIFF_COLUMNS)
      S.root_walks = true;
  }
  // the program as self-contained records
  S.ops.assign(S.prog.size() + 1, ftd::DOp{});
  for (size_t k = 0; k < S.prog.size(); ++k) {
    const unsigned w = S.prog[k];
    const int i = (int)(w & ~ftd::kBeginFlag);
    const ftd::DNode& d = S.nodes[i];
    ftd::DOp& o = S.ops[k];
    std::memset(&o, 0, sizeof(o));
    o.node = i;
    if (w & ftd::kBeginFlag) {
      o.code = (int)ftd::kBeginFlag;
      continue;
    }
    o.code = d.kind | ((d.flags & 1) ? ftd::kOpRotated : 0) | ((d.flags & 2) ? ftd::kOpNeedsLb : 0);
    for (int m = 0; m < 5; ++m) o.p[m] = d.p[m];
    if (d.kind <= FT_NODE_EXTRUDE_POLYGON) {
      o.c[0] = d.cx; o.c[1] = d.cy; o.c[2] = d.cz;
      o.r[0] = d.szs; o.r[1] = d.szc; o.r[2] = d.sys; o.r[3] = d.syc; o.r[4] = d.sxs; o.r[5] = d.sxc;
    } else {
      for (int m = 0; m < 3; ++m) { o.c[m] = d.bmin[m]; o.r[m] = d.bmax[m]; }
    }
  }
  S.ops.back().code = FT_NODE_UNION;  // padding: read ahead, never executed

  // FLOPs of one evaluation of each subtree (work accounting of instrumented runs):
  // FMA = 2, add/mul/min/max/abs/compare = 1, sqrt/sin/cos = 1.
  auto fn_cost = [&](int32_t f) -> float {
    if (f < 0) return 0.0f;
    const ft_function& F = ir->functions[f];
    float c = 0.0f;
    for (uint32_t k = 0; k < F.count; ++k) {
      const ft_expr& e = ir->exprs[F.first + k];
      float w = e.type == FT_TYPE_VEC3 ? 3.0f : 1.0f;
      switch (e.op) {
        case FT_EXPR_CONST: case FT_EXPR_PARAM: case FT_EXPR_VEC3: case FT_EXPR_GET_X: case FT_EXPR_GET_Y:
        case FT_EXPR_GET_Z: break;
        case FT_EXPR_DOT: c += 5; break;
        case FT_EXPR_LENGTH: c += 6; break;
        case FT_EXPR_NORMALIZE: c += 10; break;
        case FT_EXPR_ROTATE_X: case FT_EXPR_ROTATE_Y: case FT_EXPR_ROTATE_Z: c += 9; break;
        case FT_EXPR_SPHERE_DIST: c += 7; break;
        case FT_EXPR_BOX_DIST: c += 20; break;
        case FT_EXPR_CYLINDER_DIST: c += 16; break;
        case FT_EXPR_TORUS_DIST: c += 12; break;
        case FT_EXPR_VALUE_NOISE: c += 150; break;
        case FT_EXPR_FBM: c += 600; break;
        case FT_EXPR_MIX: case FT_EXPR_CLAMP: case FT_EXPR_SMOOTHSTEP: c += 4 * w; break;
        default: c += w; break;
      }
    }
    return c;
  };
  for (uint32_t i = 0; i < N; ++i) {
    const ft_node& n = ir->nodes[i];
    DNode& d = S.nodes[i];
    float xf = (d.flags & 1) ? 21.0f : 3.0f;
    float lb = (d.flags & 2) ? (float)ftd::kAabbFlopsHost : 0.0f;
    float own;
    switch (n.kind) {
      case FT_NODE_SPHERE: own = xf + 7; break;
      case FT_NODE_BOX: own = xf + 23; break;
      case FT_NODE_CYLINDER: own = xf + 20; break;
      case FT_NODE_TORUS: own = xf + 12; break;
      case FT_NODE_EXTRUDE_POLYGON: own = xf + 30; break;
      case FT_NODE_CUSTOM: own = xf + 12 + fn_cost(n.fn_domain) + fn_cost(n.fn_distance) + fn_cost(n.fn_post); break;
      case FT_NODE_DOMAIN_MODIFIER: own = xf + fn_cost(n.fn_domain) + lb; break;
      case FT_NODE_DISTANCE_POST_MODIFIER: own = xf + fn_cost(n.fn_post) + lb; break;
      case FT_NODE_NOISE: own = xf + 130.0f * std::max(1.0f, n.param[0]) + lb; break;
      case FT_NODE_UNION: own = 4; break;
      case FT_NODE_INTERSECT: case FT_NODE_SUBTRACT: own = 2; break;
      case FT_NODE_UNION_CHAMFER: case FT_NODE_INTERSECT_CHAMFER: case FT_NODE_DIFF_CHAMFER: own = 7 + lb; break;
      case FT_NODE_UNION_COLUMNS: case FT_NODE_INTERSECT_COLUMNS: case FT_NODE_DIFF_COLUMNS: own = 30 + lb; break;
      case FT_NODE_UNION_STAIRS: case FT_NODE_INTERSECT_STAIRS: case FT_NODE_DIFF_STAIRS: own = 14 + lb; break;
      case FT_NODE_UNION_SOFT: own = 9 + lb; break;
      case FT_NODE_PIPE: case FT_NODE_ENGRAVE: case FT_NODE_GROOVE: case FT_NODE_TONGUE: own = 6 + lb; break;
      case FT_NODE_SLICE: own = 3 + lb; break;
      case FT_NODE_SMOOTH: own = 2; break;
      default: own = 13 + lb; break;  // *_ROUND
    }
    float sub = 0.0f;
    if (n.lhs >= 0 && (is_binary(n.kind) || is_unary(n.kind))) sub += S.nodes[n.lhs].cost;
    if (n.rhs >= 0 && is_binary(n.kind)) sub += S.nodes[n.rhs].cost;
    d.cost = own + sub;
  }

  // lower-bound program (sdf_lower_bound recursion, renderer.rs:5413-5543)
  int lsp = 0, lmax = 0;
  std::function<void(int)> emit_lb = [&](int i) {
    const ft_node& n = ir->nodes[i];
    if (n.kind == FT_NODE_UNION || n.kind == FT_NODE_INTERSECT || n.kind == FT_NODE_SUBTRACT) {
      emit_lb(n.lhs);
      emit_lb(n.rhs);
      --lsp;
    } else if (n.kind == FT_NODE_SMOOTH) {
      emit_lb(n.lhs);
    } else {
      ++lsp;
      lmax = std::max(lmax, lsp);
    }
    S.lbprog.push_back((unsigned)i);
  };
  emit_lb(S.root);
  if (lmax > ftd::kMaxStack) { err = "scene tree needs a deeper lower-bound stack than supported"; return false; }
  S.max_stack = std::max(S.max_stack, lmax);

  // accel leaves (collect_accel_leaves) and per-node leaf ids
  std::vector<int> leaf_nodes;
  std::function<void(int)> collect = [&](int i) {
    const ft_node& n = ir->nodes[i];
    if (n.kind == FT_NODE_UNION) { collect(n.lhs); collect(n.rhs); return; }
    leaf_nodes.push_back(i);
  };
  collect(S.root);
  for (size_t li = 0; li < leaf_nodes.size(); ++li) {
    int ln = leaf_nodes[li];
    ftd::DLeaf L;
    for (int k = 0; k < 3; ++k) { L.bmin[k] = bounds[ln].mn[k]; L.bmax[k] = bounds[ln].mx[k]; }
    L.node = ln;
    L.is_custom = ir->nodes[ln].kind == FT_NODE_CUSTOM ? 1 : 0;
    L.op = ir->nodes[ln].kind <= FT_NODE_EXTRUDE_POLYGON ? S.nodes[ln].prog_first : -1;
    L.pad[0] = L.pad[1] = L.pad[2] = 0;
    S.leaves.push_back(L);
    // every node of the leaf's subtree occupies prog[prog_first .. prog_last]
    for (int k = S.nodes[ln].prog_first; k <= S.nodes[ln].prog_last; ++k) S.nodes[S.prog[k] & ~ftd::kBeginFlag].leaf = (int)li;
  }

  // walk requirements inside leaves: frames and pending material results
  std::function<void(int, int, int)> walk_req = [&](int i, int frames, int mats) {
    const ft_node& n = ir->nodes[i];
    S.max_walk_frames = std::max(S.max_walk_frames, frames + 2);
    S.max_walk_mats = std::max(S.max_walk_mats, mats + 1);
    if (n.kind <= FT_NODE_NOISE) return;
    if (n.kind == FT_NODE_SLICE || n.kind == FT_NODE_SMOOTH) { walk_req(n.lhs, frames, mats); return; }
    if (is_blend(n.kind)) {
      walk_req(n.lhs, frames + 1, mats);
      walk_req(n.rhs, frames + 1, mats + 1);
      return;
    }
    walk_req(n.lhs, frames + 1, mats);
    walk_req(n.rhs, frames + 1, mats);
  };
  for (int ln : leaf_nodes) walk_req(ln, 0, 0);

  // shading sweep (one post-order pass that yields the tree-walk normal and materials):
  // stack depth with modifier / noise subtrees as single terminals, and the split node
  std::vector<int> sweep_depth(N, 1);
  for (uint32_t i = 0; i < N; ++i) {  // children precede parents in the IR
    const ft_node& n = ir->nodes[i];
    if (n.kind <= FT_NODE_NOISE) sweep_depth[i] = 1;
    else if (n.kind == FT_NODE_SLICE || n.kind == FT_NODE_SMOOTH) sweep_depth[i] = sweep_depth[n.lhs];
    else sweep_depth[i] = std::max(sweep_depth[n.lhs], sweep_depth[n.rhs] + 1);
  }
  for (uint32_t i = 0; i < N; ++i) {
    int split = -1, j = (int)i;
    for (;;) {
      const uint32_t k = ir->nodes[j].kind;
      if (is_blend(k)) { split = j; break; }
      if (k == FT_NODE_SMOOTH) { j = ir->nodes[j].lhs; continue; }
      if (k == FT_NODE_NOISE || k == FT_NODE_UNION || k == FT_NODE_INTERSECT || k == FT_NODE_SUBTRACT)
        split = (j == (int)i && k != FT_NODE_NOISE) ? -1 : -2;  // hard nodes below Smooth / Noise: per-node walks
      break;
    }
    S.nodes[i].walk_split = split;
    S.nodes[i].walk_ok = (sweep_depth[i] <= ftd::kMaxWalkMats && sweep_depth[i] <= S.max_stack) ? 1 : 0;
  }
  if (S.max_walk_frames > ftd::kMaxWalk || S.max_walk_mats > ftd::kMaxWalkMats) {
    err = "smooth-boolean nesting inside one accel leaf is deeper than supported";
    return false;
  }

  // BVH over leaves (BVH and bricks modes)
  if (accel != FT_ACCEL_NAIVE && !S.leaves.empty()) {
    std::function<int(std::vector<int>, int)> build = [&](std::vector<int> idx, int depth) -> int {
      S.max_bvh_depth = std::max(S.max_bvh_depth, depth);
      int me = (int)S.bvh.size();
      S.bvh.push_back(ftd::DBvh{});
      B3 b;
      for (int k = 0; k < 3; ++k) { b.mn[k] = S.leaves[idx[0]].bmin[k]; b.mx[k] = S.leaves[idx[0]].bmax[k]; }
      if (idx.size() == 1) {
        ftd::DBvh& nd = S.bvh[me];
        for (int k = 0; k < 3; ++k) { nd.bmin[k] = b.mn[k]; nd.bmax[k] = b.mx[k]; }
        nd.left = -(idx[0] + 1);
        nd.right = -1;
        return me;
      }
      for (size_t m = 1; m < idx.size(); ++m) {
        B3 o;
        for (int k = 0; k < 3; ++k) { o.mn[k] = S.leaves[idx[m]].bmin[k]; o.mx[k] = S.leaves[idx[m]].bmax[k]; }
        b = bunion(b, o);
      }
      float ex = b.mx[0] - b.mn[0], ey = b.mx[1] - b.mn[1], ez = b.mx[2] - b.mn[2];
      int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
      auto cen = [&](int l) { return (S.leaves[l].bmin[axis] + S.leaves[l].bmax[axis]) * 0.5f; };
      std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) { return cen(a) < cen(c); });
      size_t mid = idx.size() / 2;
      int l = build(std::vector<int>(idx.begin(), idx.begin() + mid), depth + 1);
      int r = build(std::vector<int>(idx.begin() + mid, idx.end()), depth + 1);
      ftd::DBvh& nd = S.bvh[me];
      for (int k = 0; k < 3; ++k) { nd.bmin[k] = b.mn[k]; nd.bmax[k] = b.mx[k]; }
      nd.left = l;
      nd.right = r;
      return me;
    };
    std::vector<int> all(S.leaves.size());
    for (size_t k = 0; k < all.size(); ++k) all[k] = (int)k;
    build(all, 1);
    if (S.max_bvh_depth + 2 > ftd::kMaxBvhStack) { err = "BVH deeper than the traversal stack"; return false; }
  }

  // materials
  bool dielectric_any = false, dielectric_rough = false, standard_transmissive = false, standard_rough = false;
  for (uint32_t m = 0; m < ir->n_materials; ++m) {
    const ft_material& a = ir->materials[m];
    ftd::DParams p;
    std::memset(&p, 0, sizeof(p));
    for (int k = 0; k < 3; ++k) {
      p.color[k] = a.color[k];
      p.specular_color[k] = a.specular_color[k];
      p.emission_color[k] = a.emission_color[k];
      p.medium_absorption[k] = a.medium_absorption[k];
      p.pattern_a[k] = a.pattern_color_a[k];
      p.pattern_b[k] = a.pattern_color_b[k];
    }
    p.roughness = a.roughness; p.metallic = a.metallic; p.specular = a.specular;
    p.specular_weight = a.specular_weight; p.ior = a.ior; p.clearcoat = a.clearcoat;
    p.clearcoat_roughness = a.clearcoat_roughness; p.transmission = a.transmission;
    p.emission_strength = a.emission_strength; p.thin_walled = a.thin_walled ? 1.0f : 0.0f;
    p.has_medium = a.has_medium ? 1.0f : 0.0f; p.medium_ior = a.medium_ior;
    p.medium_density = a.medium_density; p.pattern = (float)a.pattern; p.pattern_scale = a.pattern_scale;
    p.model = (float)a.model;
    p.hook = -1.0f;
    for (uint32_t h = 0; h < ir->n_material_hooks; ++h) {
      const ft_material_hooks& H = ir->material_hooks[h];
      if (H.material != m) continue;
      const int32_t ids[20] = {H.color, H.roughness, H.ior, H.thin_walled, H.emission_color, H.emission_strength,
                               H.metallic, H.specular, H.specular_weight, H.transmission, H.specular_color,
                               H.clearcoat_roughness, H.clearcoat, H.coat_mask, H.bump, H.normal,
                               H.medium_ior, H.medium_absorption, H.medium_density, H.eval};
      for (int32_t f : ids) {
        if (f >= (int32_t)ir->n_functions || (f >= 0 && ir->functions[f].n_params != FT_HOOK_PARAMS)) {
          err = "material hook is not a function of the shading context";
          return false;
        }
      }
      ftd::DHooks d{H.color, H.roughness, H.ior, H.thin_walled, H.emission_color, H.emission_strength,
                    H.metallic, H.specular, H.specular_weight, H.transmission, H.specular_color,
                    H.clearcoat_roughness, H.clearcoat, H.coat_mask, H.bump, H.normal,
                    H.medium_ior, H.medium_absorption, H.medium_density, H.eval};
      p.hook = (float)S.hooks.size();
      S.hooks.push_back(d);
      break;
    }
    S.mats.push_back(p);
    if (a.dynamic_material_id >= 0 && p.hook < 0.0f && S.dynamic_material < 0) S.dynamic_material = (int)m;
    // per-hit roughness / transmission hooks and lerps between two parameter sets of one
    // model can reach any value, so a model counts as (rough) transmissive when any of its
    // materials can be
    const ftd::DHooks* hk = p.hook >= 0.0f ? &S.hooks[(size_t)p.hook] : nullptr;
    const bool rough = a.roughness > 0.06f || (hk && hk->roughness >= 0);
    if (a.model == FT_MODEL_DIELECTRIC) {
      dielectric_any = true;
      dielectric_rough = dielectric_rough || rough;
    } else if (a.model == FT_MODEL_STANDARD) {
      standard_transmissive = standard_transmissive || a.transmission > 0.0f || (hk && hk->transmission >= 0);
      standard_rough = standard_rough || rough;
    }
  }
  const bool any_transmissive = dielectric_any || standard_transmissive;
  const bool any_rough_transmissive = (dielectric_any && dielectric_rough) || (standard_transmissive && standard_rough);
  // children of one hit: reflection (<= 3 lobes) + refraction (<= 4 lobes), per split material
  int per_mat = any_rough_transmissive ? 7 : (any_transmissive ? 2 : 1);
  bool any_blend = false;
  for (uint32_t i = 0; i < N; ++i) any_blend = any_blend || is_blend(ir->nodes[i].kind);
  S.max_children = per_mat * (any_blend ? 2 : 1);

  // objects
  for (uint32_t o = 0; o < ir->n_objects; ++o) {
    const float* t = ir->object_transforms + 6 * o;
    float rz = deg2rad(-t[5]), ry = deg2rad(-t[4]), rx = deg2rad(-t[3]);
    float v[9] = {t[0], t[1], t[2], std::sin(rz), std::cos(rz), std::sin(ry), std::cos(ry), std::sin(rx), std::cos(rx)};
    S.objects.insert(S.objects.end(), v, v + 9);
  }
  // lights
  S.env_rad[0] = S.env_rad[1] = S.env_rad[2] = 0.0f;
  for (uint32_t l = 0; l < ir->n_lights; ++l) {
    const ft_light& a = ir->lights[l];
    ftd::DLight L;
    std::memset(&L, 0, sizeof(L));
    L.kind = (int)a.kind;
    L.samples = a.samples;
    L.radius = a.radius;
    for (int k = 0; k < 3; ++k) { L.pos[k] = a.position[k]; L.rad[k] = a.radiance[k]; }
    S.lights.push_back(L);
    if (a.kind == FT_LIGHT_ENV)
      for (int k = 0; k < 3; ++k) S.env_rad[k] = S.env_rad[k] + a.radiance[k];
  }
  // functions
  if (ir->n_functions) S.fns.assign(ir->functions, ir->functions + ir->n_functions);
  if (ir->n_exprs) S.exprs.assign(ir->exprs, ir->exprs + ir->n_exprs);
  // The interpreter holds live values only.  Each node's result goes to a value slot that a
  // value past its last use has freed, so a function may have any number of nodes while at
  // most kMaxExpr values are live at once.  The device copy of the expressions names slots:
  // operands are rewritten to the slots of the nodes they read, and the slot a node writes
  // is kept in its `pad` field (a small integer, exact as a float).
  for (uint32_t f = 0; f < (uint32_t)S.fns.size(); ++f) {
    const ft_function& F = S.fns[f];
    const int n = (int)F.count;
    std::vector<int> last(n, -1), slot(n, -1), free_slots;
    auto refs = [](const ft_expr& e) {
      std::array<int32_t, 4> r{e.a, e.b, e.c, e.d};
      if (e.op == FT_EXPR_PARAM) r[0] = -1;  // a parameter index, not a node
      return r;
    };
    for (int k = 0; k < n; ++k)
      for (int32_t o : refs(S.exprs[F.first + k]))
        if (o >= 0) last[o] = k;
    last[n - 1] = n;  // the result
    int top = 0;
    for (int k = 0; k < n; ++k) {
      ft_expr& e = S.exprs[F.first + k];
      const std::array<int32_t, 4> r = refs(e);
      int32_t* fields[4] = {&e.a, &e.b, &e.c, &e.d};
      for (int q = 0; q < 4; ++q)
        if (r[q] >= 0) *fields[q] = slot[r[q]];
      for (int q = 0; q < 4; ++q) {  // operands read for the last time free their slots
        if (r[q] < 0 || last[r[q]] != k) continue;
        bool seen = false;
        for (int p = 0; p < q; ++p) seen = seen || r[p] == r[q];
        if (!seen) free_slots.push_back(slot[r[q]]);
      }
      if (free_slots.empty()) {
        slot[k] = top++;
      } else {
        slot[k] = free_slots.back();
        free_slots.pop_back();
      }
      if (top > ftd::kMaxExpr) {
        err = "function " + std::to_string(f) + " keeps more than " + std::to_string(ftd::kMaxExpr) +
              " values live";
        return false;
      }
      e.pad = (float)slot[k];
      if (last[k] < 0) free_slots.push_back(slot[k]);  // never read
    }
    S.max_fn_exprs = std::max(S.max_fn_exprs, top);
  }
  return true;
}

CamBasis camera_basis(const HostScene& s, const ft_render_options& o) {
  float origin[3], target[3], fov;
  const ft_camera& c = s.camera;
  for (int k = 0; k < 3; ++k) {
    float dflt_origin = s.center[k] + (k == 2 ? o.camera_z : 0.0f);
    origin[k] = (c.present && c.has_origin) ? c.origin[k] : dflt_origin;
    target[k] = (c.present && c.has_target) ? c.target[k] : s.center[k];
  }
  fov = (c.present && c.has_fov) ? c.fov_y_degrees : o.fov_y_degrees;
  return camera_basis(origin, target, fov);
}

CamBasis camera_basis(const float origin[3], const float target[3], float fov) {
  auto normalize = [](float v[3]) {
    float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l <= 1.1920929e-7f) return;
    float inv = 1.0f / l;
    for (int k = 0; k < 3; ++k) v[k] = v[k] * inv;
  };
  auto cross = [](const float a[3], const float b[3], float out[3]) {
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
  };
  CamBasis cb;
  float up0[3] = {0.0f, 1.0f, 0.0f};
  for (int k = 0; k < 3; ++k) { cb.origin[k] = origin[k]; cb.forward[k] = target[k] - origin[k]; }
  normalize(cb.forward);
  cross(cb.forward, up0, cb.right);
  normalize(cb.right);
  cross(cb.right, cb.forward, cb.up);
  normalize(cb.up);
  cb.tan_fov = std::tan(0.5f * deg2rad(fov));
  return cb;
}

}  // namespace fth
