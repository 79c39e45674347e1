// ft_emit.cpp — scene emitter: writes scene-specialised HIP device code for hiprtc.
//
// This is the counterpart of the reference's Cranelift JIT (jit.rs:396-747, 1166-1914):
// where the reference compiles each Forge `sdf` / `fn` body to host code, the emitter
// lowers the whole compiled scene to gfx950 device code:
//   * every Forge function of the IR becomes a straight-line `ftx_fn_<k>` (captures are
//     literals already, helper calls are inlined by the front end);
//   * the scene tree (or each non-trivial accel leaf) becomes straight-line SSA code with
//     all primitive parameters, transforms and blend radii as literals, following
//     sdf_distance_info / sdf_lower_bound node for node (renderer.rs:5413-5955);
//   * `ftx_lower_bound` restates sdf_lower_bound of the root.
// The output is compiled together with ft_kernels.h, whose hooks (FT_EMITTED_*) route the
// march, shadow and hook evaluation through the emitted functions.
#include "ft_emit.h"

#include <cmath>
#include <cstdio>
#include <functional>
#include <sstream>

namespace fth {
namespace {

std::string lit(float v) {
  if (std::isnan(v)) return "__builtin_nanf(\"\")";
  if (std::isinf(v)) return v > 0 ? "__builtin_huge_valf()" : "(-__builtin_huge_valf())";
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%af", (double)v);
  return std::string("(") + buf + ")";
}

bool is_binary(uint32_t k) { return k >= FT_NODE_UNION && k <= FT_NODE_TONGUE; }

struct Val {
  std::string d, lb, id;
};

class Emitter {
 public:
  Emitter(const HostScene& S, const ft_scene_ir* ir) : S_(S), ir_(ir) {}

  std::string functions() {
    std::ostringstream o;
    for (uint32_t f = 0; f < ir_->n_functions; ++f) o << function(f);
    o << "__device__ float3 ftx_fn(int fn, const float3* a) {\n  switch (fn) {\n";
    for (uint32_t f = 0; f < ir_->n_functions; ++f) {
      const ft_function& F = ir_->functions[f];
      o << "    case " << f << ": return ftx_fn_" << f << "(";
      for (uint32_t a = 0; a < F.n_params; ++a) o << (a ? ", " : "") << "a[" << a << "]";
      o << ");\n";
    }
    o << "    default: return mk(0.0f, 0.0f, 0.0f);\n  }\n}\n";
    return o.str();
  }

  // Straight-line sdf_distance_info of the subtree at `node`, as the body of a function
  // with parameter `p`; the winning primitive's node index is returned in `id`.
  std::string subtree_info(int node, const std::string& fname) {
    body_.str("");
    tmp_ = 0;
    Val v = emit(node, "p", false);
    std::ostringstream o;
    o << "__device__ __forceinline__ void " << fname << "(float3 p, float& d_out, float& lb_out, int& id_out) {\n"
      << body_.str() << "  d_out = " << v.d << ";\n  lb_out = " << (v.lb.empty() ? v.d : v.lb)
      << ";\n  id_out = " << v.id << ";\n}\n";
    return o.str();
  }

  // The same distance without the winning primitive (the march's inner steps).
  std::string subtree_dist(int node, const std::string& fname) {
    body_.str("");
    tmp_ = 0;
    ids_ = false;
    Val v = emit(node, "p", false);
    ids_ = true;
    std::ostringstream o;
    o << "__device__ __forceinline__ float " << fname << "(float3 p) {\n" << body_.str() << "  return " << v.d
      << ";\n}\n";
    return o.str();
  }

  std::string lower_bound(int node) {
    body_.str("");
    tmp_ = 0;
    std::string v = emit_lb(node, "p");
    std::ostringstream o;
    o << "__device__ __forceinline__ float ftx_lower_bound(const DevScene& S, float3 p) {\n" << body_.str()
      << "  return " << v << ";\n}\n";
    return o.str();
  }

 private:
  const HostScene& S_;
  const ft_scene_ir* ir_;
  std::ostringstream body_;
  int tmp_ = 0;
  bool ids_ = true;  // false: distances only (the id selects are left out)

  std::string fresh(const char* pfx) { return std::string(pfx) + std::to_string(tmp_++); }

  // ------------------------------------------------------------ functions
  std::string function(uint32_t f) {
    const ft_function& F = ir_->functions[f];
    std::ostringstream o;
    o << "__device__ __forceinline__ float3 ftx_fn_" << f << "(";
    for (uint32_t a = 0; a < F.n_params; ++a) o << (a ? ", " : "") << "float3 a" << a;
    o << ") {\n";
    auto v = [&](int k) { return "v" + std::to_string(k); };
    for (uint32_t k = 0; k < F.count; ++k) {
      const ft_expr& e = ir_->exprs[F.first + k];
      std::string a = e.a >= 0 ? v(e.a) : "", b = e.b >= 0 ? v(e.b) : "", c = e.c >= 0 ? v(e.c) : "",
                  d = e.d >= 0 ? v(e.d) : "";
      std::string r;
      auto map1 = [&](const char* fn) {
        return std::string("mk(") + fn + "(" + a + ".x), " + fn + "(" + a + ".y), " + fn + "(" + a + ".z))";
      };
      auto map2 = [&](const char* fn) {
        return std::string("mk(") + fn + "(" + a + ".x, " + b + ".x), " + fn + "(" + a + ".y, " + b + ".y), " + fn +
               "(" + a + ".z, " + b + ".z))";
      };
      auto splat = [&](const std::string& x) { return "splat3(" + x + ")"; };
      switch (e.op) {
        case FT_EXPR_CONST: r = splat(lit(e.value)); break;
        case FT_EXPR_PARAM: r = "a" + std::to_string(e.a); break;
        case FT_EXPR_VEC3: r = "mk(" + a + ".x, " + b + ".x, " + c + ".x)"; break;
        case FT_EXPR_GET_X: r = splat(a + ".x"); break;
        case FT_EXPR_GET_Y: r = splat(a + ".y"); break;
        case FT_EXPR_GET_Z: r = splat(a + ".z"); break;
        case FT_EXPR_NEG: r = "mk(-" + a + ".x, -" + a + ".y, -" + a + ".z)"; break;
        case FT_EXPR_ADD: r = "add(" + a + ", " + b + ")"; break;
        case FT_EXPR_SUB: r = "sub(" + a + ", " + b + ")"; break;
        case FT_EXPR_MUL: r = "cmul(" + a + ", " + b + ")"; break;
        case FT_EXPR_DIV: r = "mk(" + a + ".x / " + b + ".x, " + a + ".y / " + b + ".y, " + a + ".z / " + b + ".z)"; break;
        case FT_EXPR_ABS: r = map1("fabsf"); break;
        case FT_EXPR_FLOOR: r = map1("floorf"); break;
        case FT_EXPR_CEIL: r = map1("ceilf"); break;
        case FT_EXPR_FRACT: r = map1("fractf_"); break;
        case FT_EXPR_SQRT: r = map1("sqrtf"); break;
        case FT_EXPR_SIN: r = map1("sinf"); break;
        case FT_EXPR_COS: r = map1("cosf"); break;
        case FT_EXPR_SATURATE: r = map1("saturatef_"); break;
        case FT_EXPR_MIN: r = map2("fminf"); break;
        case FT_EXPR_MAX: r = map2("fmaxf"); break;
        case FT_EXPR_POW: r = map2("powf"); break;
        case FT_EXPR_CLAMP:
          r = "mk(fclamp_any(" + a + ".x, " + b + ".x, " + c + ".x), fclamp_any(" + a + ".y, " + b + ".y, " + c +
              ".y), fclamp_any(" + a + ".z, " + b + ".z, " + c + ".z))";
          break;
        case FT_EXPR_MIX:
          r = "mk(lerpf(" + a + ".x, " + b + ".x, " + c + ".x), lerpf(" + a + ".y, " + b + ".y, " + c + ".y), lerpf(" +
              a + ".z, " + b + ".z, " + c + ".z))";
          break;
        case FT_EXPR_STEP: r = splat("(" + b + ".x < " + a + ".x ? 0.0f : 1.0f)"); break;
        case FT_EXPR_SMOOTHSTEP: r = splat("smoothstep_builtin(" + a + ".x, " + b + ".x, " + c + ".x)"); break;
        case FT_EXPR_DOT: r = splat("dot(" + a + ", " + b + ")"); break;
        case FT_EXPR_LENGTH: r = splat("len(" + a + ")"); break;
        case FT_EXPR_NORMALIZE: r = "normalize_div(" + a + ")"; break;
        case FT_EXPR_ROTATE_X: r = "rotate_deg_x(" + a + ", " + b + ".x)"; break;
        case FT_EXPR_ROTATE_Y: r = "rotate_deg_y(" + a + ", " + b + ".x)"; break;
        case FT_EXPR_ROTATE_Z: r = "rotate_deg_z(" + a + ", " + b + ".x)"; break;
        case FT_EXPR_SPHERE_DIST: r = splat("(len(" + a + ") - " + b + ".x)"); break;
        case FT_EXPR_BOX_DIST: r = splat("box_dist_builtin(" + a + ", " + b + ")"); break;
        case FT_EXPR_CYLINDER_DIST: r = splat("cylinder_dist_builtin(" + a + ", " + b + ".x, " + c + ".x)"); break;
        case FT_EXPR_TORUS_DIST: r = splat("torus_dist_builtin(" + a + ", " + b + ".x, " + c + ".x)"); break;
        case FT_EXPR_VALUE_NOISE: r = splat("value_noise_3d(" + a + ")"); break;
        case FT_EXPR_FBM:
          r = splat("fbm_3d(" + a + ", " + b + ".x > 0.0f ? (int)" + b + ".x : 0, " + c + ".x, " + d + ".x)");
          break;
        default: r = "mk(0.0f, 0.0f, 0.0f)"; break;
      }
      o << "  const float3 " << v((int)k) << " = " << r << ";\n";
    }
    o << "  return " << v((int)F.count - 1) << ";\n}\n";
    return o.str();
  }

  // --------------------------------------------------------------- nodes
  std::string local(const ftd::DNode& n, const std::string& p) {
    std::string q = fresh("q");
    body_ << "  float3 " << q << " = mk(" << p << ".x - " << lit(n.cx) << ", " << p << ".y - " << lit(n.cy) << ", "
          << p << ".z - " << lit(n.cz) << ");\n";
    if (n.flags & 1) {
      body_ << "  " << q << " = rot_z(" << q << ", " << lit(n.szs) << ", " << lit(n.szc) << ");\n";
      body_ << "  " << q << " = rot_y(" << q << ", " << lit(n.sys) << ", " << lit(n.syc) << ");\n";
      body_ << "  " << q << " = rot_x(" << q << ", " << lit(n.sxs) << ", " << lit(n.sxc) << ");\n";
    }
    return q;
  }

  std::string aabb(const ftd::DNode& n, const std::string& p) {
    return "aabb_lb_lit(" + p + ", " + lit(n.bmin[0]) + ", " + lit(n.bmin[1]) + ", " + lit(n.bmin[2]) + ", " +
           lit(n.bmax[0]) + ", " + lit(n.bmax[1]) + ", " + lit(n.bmax[2]) + ")";
  }

  std::string prim(const ftd::DNode& n, const std::string& q) {
    switch (n.kind) {
      case FT_NODE_SPHERE: return "sdf_sphere(" + q + ", " + lit(n.p[0]) + ", " + lit(n.p[1]) + ")";
      case FT_NODE_BOX:
        return "sdf_box(" + q + ", " + lit(n.p[0]) + ", " + lit(n.p[1]) + ", " + lit(n.p[2]) + ", " + lit(n.p[3]) + ", " +
               lit(n.p[4]) + ")";
      case FT_NODE_CYLINDER:
        return "sdf_cylinder(" + q + ", " + lit(n.p[0]) + ", " + lit(n.p[1]) + ", " + lit(n.p[2]) + ", " + lit(n.p[3]) + ")";
      case FT_NODE_TORUS: return "sdf_torus(" + q + ", " + lit(n.p[0]) + ", " + lit(n.p[1]) + ")";
      default:
        return "sdf_ngon(" + q + ", " + std::to_string((int)n.p[0]) + ", " + lit(n.p[1]) + ", " + lit(n.p[2]) + ", " +
               lit(n.p[3]) + ", " + lit(n.p[4]) + ")";
    }
  }

  std::string custom(const ftd::DNode& n, const std::string& q) {
    std::string dp = q;
    if (n.fn_domain >= 0) {
      dp = fresh("dp");
      body_ << "  const float3 " << dp << " = ftx_fn_" << n.fn_domain << "(" << q << ");\n";
    }
    std::string d = fresh("cd");
    if (n.fn_distance >= 0) body_ << "  float " << d << " = ftx_fn_" << n.fn_distance << "(" << dp << ").x;\n";
    else body_ << "  float " << d << " = 1.0e6f;\n";
    if (n.fn_post >= 0) body_ << "  " << d << " = ftx_fn_" << n.fn_post << "(splat3(" << d << "), " << dp << ").x;\n";
    return d;
  }

  // need_lb: the caller reads the lower bound of this subtree
  Val emit(int i, const std::string& p, bool need_lb) {
    const ftd::DNode& n = S_.nodes[i];
    const int k = n.kind;
    Val out;
    if (k <= FT_NODE_EXTRUDE_POLYGON) {
      std::string q = local(n, p);
      out.d = fresh("d");
      body_ << "  const float " << out.d << " = " << prim(n, q) << ";\n";
      out.lb = out.d;
      out.id = std::to_string(i);
      return out;
    }
    if (k == FT_NODE_CUSTOM) {
      std::string q = local(n, p);
      out.d = custom(n, q);
      if (need_lb) {
        out.lb = fresh("lb");
        body_ << "  const float " << out.lb << " = aabb_lb_lit(" << q << ", " << lit(-n.p[0]) << ", " << lit(-n.p[1])
              << ", " << lit(-n.p[2]) << ", " << lit(n.p[0]) << ", " << lit(n.p[1]) << ", " << lit(n.p[2]) << ");\n";
      }
      out.id = std::to_string(i);
      return out;
    }
    if (k == FT_NODE_DOMAIN_MODIFIER || k == FT_NODE_DISTANCE_POST_MODIFIER || k == FT_NODE_NOISE) {
      std::string q = local(n, p);
      std::string cp = q;
      if (k == FT_NODE_DOMAIN_MODIFIER && n.fn_domain >= 0) {
        cp = fresh("dq");
        body_ << "  const float3 " << cp << " = ftx_fn_" << n.fn_domain << "(" << q << ");\n";
      }
      Val c = emit(n.lhs, cp, false);
      out.id = c.id;
      out.d = fresh("d");
      if (k == FT_NODE_DISTANCE_POST_MODIFIER && n.fn_post >= 0)
        body_ << "  const float " << out.d << " = ftx_fn_" << n.fn_post << "(splat3(" << c.d << "), " << q << ").x;\n";
      else if (k == FT_NODE_NOISE)
        body_ << "  const float " << out.d << " = apply_noise(" << q << ", " << c.d << ", " << (int)n.p[0] << ", "
              << lit(n.p[1]) << ", " << lit(n.p[2]) << ");\n";
      else
        body_ << "  const float " << out.d << " = " << c.d << ";\n";
      if (need_lb) {
        out.lb = fresh("lb");
        body_ << "  const float " << out.lb << " = " << aabb(n, p) << ";\n";
      }
      return out;
    }
    if (k == FT_NODE_SLICE) {
      Val c = emit(n.lhs, p, false);
      out.id = c.id;
      out.d = fresh("d");
      const char* ax = (int)n.p[0] == 0 ? ".x" : ((int)n.p[0] == 1 ? ".y" : ".z");
      body_ << "  const float " << out.d << " = fmaxf(" << c.d << ", fmaxf(" << lit(n.p[1]) << " - " << p << ax << ", "
            << p << ax << " - " << lit(n.p[2]) << "));\n";
      if (need_lb) {
        out.lb = fresh("lb");
        body_ << "  const float " << out.lb << " = " << aabb(n, p) << ";\n";
      }
      return out;
    }
    if (k == FT_NODE_SMOOTH) {
      Val c = emit(n.lhs, p, need_lb);
      out.id = c.id;
      out.d = fresh("d");
      body_ << "  const float " << out.d << " = " << c.d << " - " << lit(n.p[0] * 0.1f) << ";\n";
      if (need_lb) {
        out.lb = fresh("lb");
        body_ << "  const float " << out.lb << " = " << c.lb << " - " << lit(n.p[0] * 0.1f) << ";\n";
      }
      return out;
    }
    // binary
    const bool child_lb = k == FT_NODE_UNION || k == FT_NODE_INTERSECT;
    Val a = emit(n.lhs, p, child_lb || (need_lb && k == FT_NODE_SUBTRACT));
    Val b = emit(n.rhs, p, child_lb);
    out.d = fresh("d");
    out.id = ids_ ? fresh("id") : std::string("0");
    if (k == FT_NODE_UNION) {
      std::string tl = fresh("tl");
      body_ << "  const bool " << tl << " = " << a.lb << " <= " << b.lb << " ? (" << a.d << " <= " << b.lb << " || " << a.d
            << " <= " << b.d << ") : (!(" << b.d << " <= " << a.lb << ") && " << a.d << " <= " << b.d << ");\n";
      body_ << "  const float " << out.d << " = " << tl << " ? " << a.d << " : " << b.d << ";\n";
      if (ids_) body_ << "  const int " << out.id << " = " << tl << " ? " << a.id << " : " << b.id << ";\n";
      if (need_lb) {
        out.lb = fresh("lb");
        body_ << "  const float " << out.lb << " = fminf(" << a.lb << ", " << b.lb << ");\n";
      }
      return out;
    }
    if (k == FT_NODE_INTERSECT) {
      std::string tl = fresh("tl");
      body_ << "  const bool " << tl << " = " << a.d << " >= " << b.d << ";\n";
      body_ << "  const float " << out.d << " = " << tl << " ? " << a.d << " : " << b.d << ";\n";
      if (ids_) body_ << "  const int " << out.id << " = " << tl << " ? " << a.id << " : " << b.id << ";\n";
      if (need_lb) {
        out.lb = fresh("lb");
        body_ << "  const float " << out.lb << " = fmaxf(" << a.lb << ", " << b.lb << ");\n";
      }
      return out;
    }
    if (k == FT_NODE_SUBTRACT) {
      std::string tl = fresh("tl");
      body_ << "  const bool " << tl << " = " << a.d << " >= -" << b.d << ";\n";
      body_ << "  const float " << out.d << " = " << tl << " ? " << a.d << " : -" << b.d << ";\n";
      if (ids_) body_ << "  const int " << out.id << " = " << tl << " ? " << a.id << " : " << b.id << ";\n";
      if (need_lb) out.lb = a.lb;
      return out;
    }
    const std::string r = lit(n.p[0]), nn = lit(n.p[1]);
    std::string expr, keep;
    switch (k) {
      case FT_NODE_UNION_ROUND: expr = "op_union_round(" + a.d + ", " + b.d + ", " + r + ")"; keep = a.d + " <= " + b.d; break;
      case FT_NODE_UNION_CHAMFER: expr = "op_union_chamfer(" + a.d + ", " + b.d + ", " + r + ")"; keep = a.d + " <= " + b.d; break;
      case FT_NODE_UNION_COLUMNS: expr = "op_union_columns(" + a.d + ", " + b.d + ", " + r + ", " + nn + ")"; keep = a.d + " <= " + b.d; break;
      case FT_NODE_UNION_STAIRS: expr = "op_union_stairs(" + a.d + ", " + b.d + ", " + r + ", " + nn + ")"; keep = a.d + " <= " + b.d; break;
      case FT_NODE_UNION_SOFT: expr = "op_union_soft(" + a.d + ", " + b.d + ", " + r + ")"; keep = a.d + " <= " + b.d; break;
      case FT_NODE_INTERSECT_ROUND: expr = "op_intersect_round(" + a.d + ", " + b.d + ", " + r + ")"; keep = a.d + " >= " + b.d; break;
      case FT_NODE_INTERSECT_CHAMFER: expr = "op_intersect_chamfer(" + a.d + ", " + b.d + ", " + r + ")"; keep = a.d + " >= " + b.d; break;
      case FT_NODE_INTERSECT_COLUMNS: expr = "-op_union_columns(-" + a.d + ", -" + b.d + ", " + r + ", " + nn + ")"; keep = a.d + " >= " + b.d; break;
      case FT_NODE_INTERSECT_STAIRS: expr = "-op_union_stairs(-" + a.d + ", -" + b.d + ", " + r + ", " + nn + ")"; keep = a.d + " >= " + b.d; break;
      case FT_NODE_DIFF_ROUND: expr = "op_intersect_round(" + a.d + ", -" + b.d + ", " + r + ")"; keep = a.d + " >= -" + b.d; break;
      case FT_NODE_DIFF_CHAMFER: expr = "op_intersect_chamfer(" + a.d + ", -" + b.d + ", " + r + ")"; keep = a.d + " >= -" + b.d; break;
      case FT_NODE_DIFF_COLUMNS: expr = "-op_union_columns(-" + a.d + ", " + b.d + ", " + r + ", " + nn + ")"; keep = a.d + " >= -" + b.d; break;
      case FT_NODE_DIFF_STAIRS: expr = "-op_union_stairs(-" + a.d + ", " + b.d + ", " + r + ", " + nn + ")"; keep = a.d + " >= -" + b.d; break;
      case FT_NODE_PIPE: expr = "(fsqrt(" + a.d + " * " + a.d + " + " + b.d + " * " + b.d + ") - fabsf(" + r + "))"; keep = "true"; break;
      case FT_NODE_ENGRAVE: expr = "fmaxf(" + a.d + ", (" + a.d + " + fabsf(" + r + ") - fabsf(" + b.d + ")) * kInvSqrt2)"; keep = "true"; break;
      case FT_NODE_GROOVE: expr = "fmaxf(" + a.d + ", fminf(" + a.d + " + fabsf(" + r + "), fabsf(" + nn + ") - fabsf(" + b.d + ")))"; keep = "true"; break;
      default: expr = "fminf(" + a.d + ", fmaxf(" + a.d + " - fabsf(" + r + "), fabsf(" + b.d + ") - fabsf(" + nn + ")))"; keep = "true"; break;
    }
    body_ << "  const float " << out.d << " = " << expr << ";\n";
    if (ids_) body_ << "  const int " << out.id << " = (" << keep << ") ? " << a.id << " : " << b.id << ";\n";
    if (need_lb) {
      out.lb = fresh("lb");
      body_ << "  const float " << out.lb << " = " << aabb(n, p) << ";\n";
    }
    return out;
  }

 public:
  // ---------------------------------------------------------- shading sweep
  // shade_sweep_uniform of node `root` as straight-line code: every entry of the sweep
  // (distance, lower bound, walk normal, material reference) is an SSA value.
  std::string sweep(int root) {
    body_.str("");
    tmp_ = 0;
    prim_cost_ = 0.0;
    split_ = S_.nodes[root].walk_split;
    SVal v = sweep_node(root, false);
    std::ostringstream o;
    o << "__device__ __forceinline__ void ftx_sweep(const DevScene& S, int slot, float3 p, float eps, ShadeWalk& w,\n"
      << "                                          const HookCtx* hc) {\n"
      << "  MatPool& pool = w.pool;\n  pool.used = 0u;\n  pool.overflow = 0;\n  w.split = false;\n" << body_.str()
      << "  if (S.flops) atomicAdd(S.flops, " << (unsigned long long)(prim_cost_ + 0.5) << "ull);\n"
      << "  w.normal = " << v.n << ";\n  if (!w.split) w.a = " << v.m << ";\n"
      << "  w.mat_overflow = pool.overflow != 0 || !finish_ref(slot, pool, w.a, p, hc) ||\n"
      << "                   (w.split && !finish_ref(slot, pool, w.b, p, hc));\n}\n";
    return o.str();
  }

 private:
  struct SVal {
    std::string d, lb, n, m;
  };
  double prim_cost_ = 0.0;
  int split_ = -1;

  std::string local_expr(const ftd::DNode& n, const std::string& v) {
    std::string q = "mk(" + v + ".x - " + lit(n.cx) + ", " + v + ".y - " + lit(n.cy) + ", " + v + ".z - " + lit(n.cz) + ")";
    if (n.flags & 1)
      q = "rot_x(rot_y(rot_z(" + q + ", " + lit(n.szs) + ", " + lit(n.szc) + "), " + lit(n.sys) + ", " + lit(n.syc) +
          "), " + lit(n.sxs) + ", " + lit(n.sxc) + ")";
    return q;
  }

  SVal sweep_node(int i, bool need_lb) {
    const ftd::DNode& n = S_.nodes[i];
    const int k = n.kind;
    SVal out;
    if (k <= FT_NODE_EXTRUDE_POLYGON) {
      out.d = fresh("sd");
      out.n = fresh("sn");
      out.m = fresh("sm");
      body_ << "  const float " << out.d << " = " << prim(n, local_expr(n, "p")) << ";\n";
      body_ << "  const float3 " << out.n << " = tetra_normal([&](float3 r_) { return " << prim(n, local_expr(n, "r_"))
            << "; }, p, eps);\n";
      const unsigned id = n.mat < S_.mats.size() ? n.mat : 0u;
      const int model = S_.mats.empty() ? 0 : (int)S_.mats[id].model;
      body_ << "  MRef " << out.m << "{0, " << model << ", 0, 0.0f, " << i << ", 0};\n";
      out.lb = out.d;
      prim_cost_ += 5.0 * n.cost;
      return out;
    }
    if (k == FT_NODE_SLICE || k == FT_NODE_SMOOTH) {
      SVal c = sweep_node(n.lhs, k == FT_NODE_SMOOTH && need_lb);
      out.n = c.n;
      out.m = c.m;
      out.d = fresh("sd");
      if (k == FT_NODE_SLICE) {
        const char* ax = (int)n.p[0] == 0 ? ".x" : ((int)n.p[0] == 1 ? ".y" : ".z");
        body_ << "  const float " << out.d << " = fmaxf(" << c.d << ", fmaxf(" << lit(n.p[1]) << " - p" << ax << ", p" << ax
              << " - " << lit(n.p[2]) << "));\n";
        if (need_lb) {
          out.lb = fresh("sl");
          body_ << "  const float " << out.lb << " = " << aabb(n, "p") << ";\n";
        }
      } else {
        body_ << "  const float " << out.d << " = " << c.d << " - " << lit(n.p[0] * 0.1f) << ";\n";
        if (need_lb) {
          out.lb = fresh("sl");
          body_ << "  const float " << out.lb << " = " << c.lb << " - " << lit(n.p[0] * 0.1f) << ";\n";
        }
      }
      return out;
    }
    const bool hard = k == FT_NODE_UNION || k == FT_NODE_INTERSECT || k == FT_NODE_SUBTRACT;
    if (hard) {
      const bool child_lb = k != FT_NODE_SUBTRACT;
      SVal a = sweep_node(n.lhs, child_lb || need_lb), b = sweep_node(n.rhs, child_lb);
      out.d = fresh("sd");
      out.n = fresh("sn");
      out.m = fresh("sm");
      const std::string hl = fresh("hl");
      if (k == FT_NODE_UNION) {
        const std::string tl = fresh("tl");
        body_ << "  const bool " << tl << " = " << a.lb << " <= " << b.lb << " ? (" << a.d << " <= " << b.lb << " || " << a.d
              << " <= " << b.d << ") : (!(" << b.d << " <= " << a.lb << ") && " << a.d << " <= " << b.d << ");\n";
        body_ << "  const float " << out.d << " = " << tl << " ? " << a.d << " : " << b.d << ";\n";
        if (need_lb) {
          out.lb = fresh("sl");
          body_ << "  const float " << out.lb << " = fminf(" << a.lb << ", " << b.lb << ");\n";
        }
        body_ << "  const bool " << hl << " = " << a.d << " <= " << b.d << ";\n";
      } else if (k == FT_NODE_INTERSECT) {
        body_ << "  const float " << out.d << " = " << a.d << " >= " << b.d << " ? " << a.d << " : " << b.d << ";\n";
        if (need_lb) {
          out.lb = fresh("sl");
          body_ << "  const float " << out.lb << " = fmaxf(" << a.lb << ", " << b.lb << ");\n";
        }
        body_ << "  const bool " << hl << " = " << a.d << " >= " << b.d << ";\n";
      } else {
        body_ << "  const float " << out.d << " = " << a.d << " >= -" << b.d << " ? " << a.d << " : -" << b.d << ";\n";
        if (need_lb) out.lb = a.lb;
        body_ << "  const bool " << hl << " = " << a.d << " >= -" << b.d << ";\n";
      }
      body_ << "  const float3 " << out.n << " = " << hl << " ? " << a.n << " : "
            << (k == FT_NODE_SUBTRACT ? "mul(" + b.n + ", -1.0f)" : b.n) << ";\n";
      body_ << "  MRef " << out.m << ";\n  if (" << hl << ") { drop_ref(pool, " << b.m << "); " << out.m << " = " << a.m
            << "; } else { drop_ref(pool, " << a.m << "); " << out.m << " = " << b.m << "; }\n";
      return out;
    }
    // smooth booleans: value as in emit(), then blend_surface_normals and blend_materials
    SVal a = sweep_node(n.lhs, false), b = sweep_node(n.rhs, false);
    const std::string r = lit(n.p[0]), nn = lit(n.p[1]);
    const std::string& A = a.d;
    const std::string& B = b.d;
    std::string expr;
    switch (k) {
      case FT_NODE_UNION_ROUND: expr = "op_union_round(" + A + ", " + B + ", " + r + ")"; break;
      case FT_NODE_UNION_CHAMFER: expr = "op_union_chamfer(" + A + ", " + B + ", " + r + ")"; break;
      case FT_NODE_UNION_COLUMNS: expr = "op_union_columns(" + A + ", " + B + ", " + r + ", " + nn + ")"; break;
      case FT_NODE_UNION_STAIRS: expr = "op_union_stairs(" + A + ", " + B + ", " + r + ", " + nn + ")"; break;
      case FT_NODE_UNION_SOFT: expr = "op_union_soft(" + A + ", " + B + ", " + r + ")"; break;
      case FT_NODE_INTERSECT_ROUND: expr = "op_intersect_round(" + A + ", " + B + ", " + r + ")"; break;
      case FT_NODE_INTERSECT_CHAMFER: expr = "op_intersect_chamfer(" + A + ", " + B + ", " + r + ")"; break;
      case FT_NODE_INTERSECT_COLUMNS: expr = "-op_union_columns(-" + A + ", -" + B + ", " + r + ", " + nn + ")"; break;
      case FT_NODE_INTERSECT_STAIRS: expr = "-op_union_stairs(-" + A + ", -" + B + ", " + r + ", " + nn + ")"; break;
      case FT_NODE_DIFF_ROUND: expr = "op_intersect_round(" + A + ", -" + B + ", " + r + ")"; break;
      case FT_NODE_DIFF_CHAMFER: expr = "op_intersect_chamfer(" + A + ", -" + B + ", " + r + ")"; break;
      case FT_NODE_DIFF_COLUMNS: expr = "-op_union_columns(-" + A + ", " + B + ", " + r + ", " + nn + ")"; break;
      case FT_NODE_DIFF_STAIRS: expr = "-op_union_stairs(-" + A + ", " + B + ", " + r + ", " + nn + ")"; break;
      case FT_NODE_PIPE: expr = "(fsqrt(" + A + " * " + A + " + " + B + " * " + B + ") - fabsf(" + r + "))"; break;
      case FT_NODE_ENGRAVE: expr = "fmaxf(" + A + ", (" + A + " + fabsf(" + r + ") - fabsf(" + B + ")) * kInvSqrt2)"; break;
      case FT_NODE_GROOVE: expr = "fmaxf(" + A + ", fminf(" + A + " + fabsf(" + r + "), fabsf(" + nn + ") - fabsf(" + B + ")))"; break;
      default: expr = "fminf(" + A + ", fmaxf(" + A + " - fabsf(" + r + "), fabsf(" + B + ") - fabsf(" + nn + ")))"; break;
    }
    out.d = fresh("sd");
    body_ << "  const float " << out.d << " = " << expr << ";\n";
    if (need_lb) {
      out.lb = fresh("sl");
      body_ << "  const float " << out.lb << " = " << aabb(n, "p") << ";\n";
    }
    // blend_t: smoothstep of the children's distance difference over max(|radius|, 1e-6)
    const bool grooved = k == FT_NODE_GROOVE || k == FT_NODE_TONGUE;
    float rad = grooved ? std::fmax(n.p[0], n.p[1]) : n.p[0];
    rad = std::fabs(rad);
    if (!(rad >= 1.0e-6f)) rad = 1.0e-6f;
    const bool diff_class = k == FT_NODE_DIFF_ROUND || k == FT_NODE_DIFF_CHAMFER || k == FT_NODE_PIPE ||
                            k == FT_NODE_ENGRAVE || k == FT_NODE_DIFF_COLUMNS || k == FT_NODE_DIFF_STAIRS ||
                            k == FT_NODE_GROOVE || k == FT_NODE_TONGUE;
    const std::string t = fresh("st"), rn = fresh("rn"), bn = fresh("bn");
    if (diff_class)
      body_ << "  const float " << t << " = smoothstepf(0.5f + 0.5f * (((-" << B << ") - " << A << ") / " << lit(rad) << "));\n";
    else
      body_ << "  const float " << t << " = smoothstepf(0.5f + 0.5f * ((" << A << " - " << B << ") / " << lit(rad) << "));\n";
    body_ << "  const float3 " << rn << " = " << (diff_class ? "mul(" + b.n + ", -1.0f)" : b.n) << ";\n";
    body_ << "  const float3 " << bn << " = add(mul(" << a.n << ", 1.0f - " << t << "), mul(" << rn << ", " << t << "));\n";
    out.n = fresh("sn");
    body_ << "  const float3 " << out.n << " = len(" << bn << ") > 1.0e-6f ? normalize(" << bn << ") : (" << t
          << " < 0.5f ? " << a.n << " : " << rn << ");\n";
    out.m = fresh("sm");
    if (i == split_) {
      // the split node's children keep their references (only unary nodes lie above it)
      body_ << "  w.a = " << a.m << ";\n  w.b = " << b.m << ";\n  w.split_t = " << t << ";\n  w.split = true;\n"
            << "  const MRef " << out.m << " = " << a.m << ";\n";
      return out;
    }
    body_ << "  MRef " << out.m << " = " << a.m << ";\n  blend_refs(slot, pool, " << out.m << ", " << b.m << ", " << t
          << ", p, hc);\n";
    return out;
  }

  // sdf_lower_bound (renderer.rs:5413-5543)
  std::string emit_lb(int i, const std::string& p) {
    const ftd::DNode& n = S_.nodes[i];
    const int k = n.kind;
    std::string v = fresh("l");
    if (k == FT_NODE_UNION || k == FT_NODE_INTERSECT) {
      std::string a = emit_lb(n.lhs, p), b = emit_lb(n.rhs, p);
      body_ << "  const float " << v << " = " << (k == FT_NODE_UNION ? "fminf(" : "fmaxf(") << a << ", " << b << ");\n";
    } else if (k == FT_NODE_SUBTRACT) {
      return emit_lb(n.lhs, p);
    } else if (k == FT_NODE_SMOOTH) {
      std::string a = emit_lb(n.lhs, p);
      body_ << "  const float " << v << " = " << a << " - " << lit(n.p[0] * 0.1f) << ";\n";
    } else if (k <= FT_NODE_EXTRUDE_POLYGON) {
      std::string q = local(n, p);
      body_ << "  const float " << v << " = " << prim(n, q) << ";\n";
    } else if (k == FT_NODE_CUSTOM) {
      std::string q = local(n, p);
      body_ << "  const float " << v << " = aabb_lb_lit(" << q << ", " << lit(-n.p[0]) << ", " << lit(-n.p[1]) << ", "
            << lit(-n.p[2]) << ", " << lit(n.p[0]) << ", " << lit(n.p[1]) << ", " << lit(n.p[2]) << ");\n";
    } else {
      body_ << "  const float " << v << " = " << aabb(n, p) << ";\n";
    }
    return v;
  }
};

}  // namespace

EmitPlan plan_emission(const HostScene& S, const ft_scene_ir* ir) {
  EmitPlan plan;
  plan.functions = ir->n_functions > 0;
  const int root_nodes = S.nodes[S.root].prog_last - S.nodes[S.root].prog_first + 1;
  const bool single = S.accel == FT_ACCEL_NAIVE || S.leaves.size() <= 1;
  plan.scene = single && root_nodes <= kMaxEmittedNodes && root_nodes >= kMinEmittedNodes;
  const ftd::DNode& r = S.nodes[S.root];
  const bool hard_root = r.kind == FT_NODE_UNION || r.kind == FT_NODE_INTERSECT || r.kind == FT_NODE_SUBTRACT;
  plan.sweep = plan.scene && S.simple && !hard_root && r.walk_ok && r.walk_split != -2;
  return plan;
}

bool emission_pays(const HostScene& S, const ft_scene_ir* ir, const EmitPlan& plan) {
  if (plan.functions) return true;
  const int root_nodes = S.nodes[S.root].prog_last - S.nodes[S.root].prog_first + 1;
  return plan.scene && root_nodes >= kAutoEmitNodes;
}

std::string emit_scene_source(const HostScene& S, const ft_scene_ir* ir, const EmitPlan& plan) {
  Emitter E(S, ir);
  std::ostringstream o;
  o << "// generated by ft_emit.cpp\n#define FT_EMITTED 1\n#define FT_EMITTED_SCENE " << (plan.scene ? 1 : 0)
    << "\n#define FT_EMITTED_FUNCTIONS " << (plan.functions ? 1 : 0) << "\n#define FT_EMITTED_SWEEP "
    << (plan.sweep ? 1 : 0) << "\n#include \"ft_kernels.h\"\nnamespace ftd {\n";
  o << E.functions();
  if (plan.scene) {
    o << E.subtree_info(S.root, "ftx_root");
    o << E.subtree_dist(S.root, "ftx_root_d");
    // Accelerator::distance_info over a one-leaf BVH is that leaf's accel_leaf_distance_info
    const bool bvh_path = S.accel != FT_ACCEL_NAIVE && !S.bvh.empty();
    const bool custom_leaf_shortcut = bvh_path && S.leaves.size() == 1 && S.leaves[0].is_custom;
    o << "__device__ __forceinline__ Info ftx_scene_info(const DevScene& S, float3 p) {\n";
    if (bvh_path) o << "  if (S.flops) atomicAdd(S.flops, 3ull * kAabbFlops);\n";
    if (custom_leaf_shortcut) {
      const ftd::DLeaf& L = S.leaves[0];
      // accel_leaf_distance_info's far shortcut for a custom leaf (renderer.rs:629-657)
      float ex = L.bmax[0] - L.bmin[0], ey = L.bmax[1] - L.bmin[1], ez = L.bmax[2] - L.bmin[2];
      float far_t = std::sqrt(ex * ex + ey * ey + ez * ez) * 0.1f;
      far_t = far_t < 0.05f ? 0.05f : (far_t > 0.5f ? 0.5f : far_t);
      o << "  const float flb = aabb_lb_lit(p, " << lit(L.bmin[0]) << ", " << lit(L.bmin[1]) << ", " << lit(L.bmin[2])
        << ", " << lit(L.bmax[0]) << ", " << lit(L.bmax[1]) << ", " << lit(L.bmax[2]) << ");\n"
        << "  if (flb > " << lit(far_t) << ") { Info o; o.d = flb; o.lb = flb; o.obj = S.nodes[" << L.node
        << "].obj; o.mat = S.nodes[" << L.node << "].mat; o.leaf = 0; return o; }\n";
    }
    o << "  if (S.flops) atomicAdd(S.flops, (unsigned long long)S.nodes[" << S.root << "].cost);\n";
    o << "  float d, lb;\n  int id;\n  ftx_root(p, d, lb, id);\n  Info o;\n  o.d = d;\n  o.lb = lb;\n"
      << "  const DNode& w = S.nodes[id];\n  o.obj = w.obj;\n  o.mat = w.mat;\n  o.leaf = "
      << (bvh_path ? "0" : "w.leaf") << ";\n  return o;\n}\n";
    // the distance alone, as ftx_scene_info computes it
    o << "__device__ __forceinline__ float ftx_scene_dist(const DevScene& S, float3 p) {\n";
    if (bvh_path) o << "  if (S.flops) atomicAdd(S.flops, 3ull * kAabbFlops);\n";
    if (custom_leaf_shortcut) {
      const ftd::DLeaf& L = S.leaves[0];
      float ex = L.bmax[0] - L.bmin[0], ey = L.bmax[1] - L.bmin[1], ez = L.bmax[2] - L.bmin[2];
      float far_t = std::sqrt(ex * ex + ey * ey + ez * ez) * 0.1f;
      far_t = far_t < 0.05f ? 0.05f : (far_t > 0.5f ? 0.5f : far_t);
      o << "  const float flb = aabb_lb_lit(p, " << lit(L.bmin[0]) << ", " << lit(L.bmin[1]) << ", " << lit(L.bmin[2])
        << ", " << lit(L.bmax[0]) << ", " << lit(L.bmax[1]) << ", " << lit(L.bmax[2]) << ");\n"
        << "  if (flb > " << lit(far_t) << ") return flb;\n";
    }
    o << "  if (S.flops) atomicAdd(S.flops, (unsigned long long)S.nodes[" << S.root << "].cost);\n";
    o << "  return ftx_root_d(p);\n}\n";
    if (!bvh_path) {
      o << E.lower_bound(S.root);
    } else {
      // bvh_lower_bound over a one-node BVH: that node's AABB
      const ftd::DBvh& L = S.bvh[0];
      o << "__device__ __forceinline__ float ftx_lower_bound(const DevScene& S, float3 p) {\n  return aabb_lb_lit(p, "
        << lit(L.bmin[0]) << ", " << lit(L.bmin[1]) << ", " << lit(L.bmin[2]) << ", " << lit(L.bmax[0]) << ", "
        << lit(L.bmax[1]) << ", " << lit(L.bmax[2]) << ");\n}\n";
    }
    if (plan.sweep) o << E.sweep(S.root);
  } else {
    o << "__device__ __forceinline__ Info ftx_scene_info(const DevScene& S, float3 p) { return Info(); }\n"
      << "__device__ __forceinline__ float ftx_scene_dist(const DevScene& S, float3 p) { return 0.0f; }\n"
      << "__device__ __forceinline__ float ftx_lower_bound(const DevScene& S, float3 p) { return 0.0f; }\n";
  }
  o << "}  // namespace ftd\n";
  return o.str();
}

}  // namespace fth
