// ft_diff.h — differentiable sphere tree: forward depth + Lambert shade and the adjoint
// (implicit-differentiation) gradient with respect to every sphere's centre and radius and
// every join's blend radius.  No counterpart in the reference renderer; the forward
// follows its march (raymarch_hit renderer.rs:4091-4139, refine_hit_distance
// renderer.rs:4180-4200) and op_union_round (renderer.rs:6029-6034).
//
// Scene: n spheres joined by a balanced union_round tree (post-order node table).
// Parameters th: [cx, cy, cz, r] per sphere, then k per join (post order).
// Forward per pixel (1 spp, pixel centre): t_hit from the march, one Newton step
//   t* = t_hit - f(x)/(grad f(x) . d),  x = o + t_hit d
// and the shade s = max(0, n . l) with n the central-difference normal of step h at
// x* = o + t* d.
// Adjoint per pixel, given upstream gradients gt = dL/dt*, gs = dL/ds:
//   dt*/dth = -(df/dth)(x*) / (grad f(x*) . d)                     (implicit function)
//   ds/dth  = v . dG/dth,  v = [n.l > 0] (l - (n.l) n) / |G|,  G_k = f(x*+h e_k) - f(x*-h e_k)
//   dG_k/dth = df/dth(x*+h e_k) - df/dth(x*-h e_k) + (grad f(x*+h e_k) - grad f(x*-h e_k)) . d dt*/dth
// so dL/dth = sum over 7 points of a seed times df/dth at that point, each obtained by one
// reverse pass through the tree.  The march is recomputed, never stored.
#pragma once

#include "ft_kernels.h"

namespace ftd {

// FLOPs per lane of one tree pass (FMA = 2; add, mul, min, max, abs, compare, sqrt = 1):
// a sphere leaf translate 3 + length 6 + radius 1, a union_round join 13; a reverse pass
// adds per sphere 19 (normalised offset, gradient) and per join 22 (partials, child
// adjoints); one that accumulates parameter gradients adds per sphere 29 (seeded products
// and four 64-lane sums of 6 adds) and per join 10.
constexpr unsigned long long kDiffEvalSphere = 10, kDiffEvalJoin = 13;
constexpr unsigned long long kDiffRevSphere = 19, kDiffRevJoin = 22;
constexpr unsigned long long kDiffAccSphere = 29, kDiffAccJoin = 10;

struct DiffArgs {
  const int4* nodes;      // per node: leaf sphere (or -1), join (or -1), lhs node, rhs node
  const float* th;        // parameters
  const float* bounds;    // scene AABB (6 floats), derived from th on the device
  const CamArgs* cams;    // per view
  int n_nodes, n_prims, n_params, n_views, tiles_per_view;
  int max_steps;
  float eps, step_scale, max_dist, normal_step;
  float lx, ly, lz;       // unit light direction
};

// Per-block layout of the dynamic LDS: [n_params doubles: gradient accumulator]
// [n_nodes planes x threads floats: node values / adjoints].
__device__ __forceinline__ float* diff_vals(const DiffArgs& a) {
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  return ft_lds + 2 * ((a.n_params + 1) & ~1) + tid;
}
__device__ __forceinline__ double* diff_acc() { return reinterpret_cast<double*>(ft_lds); }

// f(p): every node value lands in its LDS plane (children are read from there).
__device__ __forceinline__ float diff_eval(const DiffArgs& a, float* __restrict__ vals, float3 p) {
  const int nth = blockDim.x * blockDim.y;
  const auto nodes = uniform_ptr(a.nodes);
  const auto th = uniform_ptr(a.th);
  const int nn = __builtin_amdgcn_readfirstlane(a.n_nodes);
  const int np = __builtin_amdgcn_readfirstlane(a.n_prims);
  float f = 0.0f;
  for (int k = 0; k < nn; ++k) {
    const int4 d = nodes[k];
    if (d.x >= 0) {
      const float3 q = mk(p.x - th[4 * d.x], p.y - th[4 * d.x + 1], p.z - th[4 * d.x + 2]);
      f = len(q) - th[4 * d.x + 3];
    } else {
      f = op_union_round(vals[d.z * nth], vals[d.w * nth], th[4 * np + d.y]);
    }
    vals[k * nth] = f;
  }
  return f;
}

struct DiffInfo {
  float d;
  unsigned obj, mat;
  int leaf;
};
struct DiffEval {
  const DiffArgs* a;
  float* vals;
  unsigned* evals;  // per-lane count of tree evaluations (work accounting)
  __device__ __forceinline__ DiffInfo operator()(float3 p) const {
    DiffInfo i;
    ++*evals;
    i.d = diff_eval(*a, vals, p);
    i.obj = 0;
    i.mat = 0;
    i.leaf = 0;
    return i;
  }
  __device__ __forceinline__ float dist(float3 p) const {
    ++*evals;
    return diff_eval(*a, vals, p);
  }
};

// Adds the FLOPs of `evals` plain passes, `revs` reverse passes and `accs` accumulating
// reverse passes of this lane to *flops (one atomic per wave).
__device__ __forceinline__ void diff_count(const DiffArgs& a, unsigned evals, unsigned revs, unsigned accs,
                                          unsigned long long* flops) {
  const unsigned long long np = (unsigned long long)a.n_prims, nj = np - 1;
  const unsigned long long ev = kDiffEvalSphere * np + kDiffEvalJoin * nj;
  const unsigned long long rv = ev + kDiffRevSphere * np + kDiffRevJoin * nj;
  const unsigned long long ac = rv + kDiffAccSphere * np + kDiffAccJoin * nj;
  unsigned long long f = ev * evals + rv * revs + ac * accs;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) f += __shfl_xor(f, m, 64);
  if (threadIdx.x == 0) atomicAdd(flops, f);
}

__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m, 64);
  return x;
}

// One reverse pass at p: values, then adjoints in place (seed 1).  Returns f(p) and sets
// grad = grad_p f.  With `acc`, adds seed * df/dth of every parameter, summed over the
// wave, to the block accumulator (lane 0 of each wave adds).
__device__ __forceinline__ float diff_reverse(const DiffArgs& a, float* __restrict__ vals, float3 p, float seed,
                                              bool acc, float3& grad) {
  const int nth = blockDim.x * blockDim.y;
  const auto nodes = uniform_ptr(a.nodes);
  const auto th = uniform_ptr(a.th);
  const int nn = __builtin_amdgcn_readfirstlane(a.n_nodes);
  const int np = __builtin_amdgcn_readfirstlane(a.n_prims);
  double* __restrict__ blk = diff_acc();
  const float f = diff_eval(a, vals, p);
  vals[(nn - 1) * nth] = 1.0f;  // adjoint of the root
  grad = mk(0.0f, 0.0f, 0.0f);
  const bool lane0 = threadIdx.x == 0;
  for (int k = nn - 1; k >= 0; --k) {
    const int4 d = nodes[k];
    const float adj = vals[k * nth];
    if (d.x >= 0) {
      const float3 q = mk(p.x - th[4 * d.x], p.y - th[4 * d.x + 1], p.z - th[4 * d.x + 2]);
      const float l = len(q);
      const float3 u = l > 0.0f ? mul(q, 1.0f / l) : mk(0.0f, 0.0f, 0.0f);
      grad = add(grad, mul(u, adj));
      if (acc) {
        const float w = seed * adj;
        const float sx = wave_sum(-w * u.x), sy = wave_sum(-w * u.y), sz = wave_sum(-w * u.z), sr = wave_sum(-w);
        if (lane0) {
          atomicAdd(&blk[4 * d.x], (double)sx);
          atomicAdd(&blk[4 * d.x + 1], (double)sy);
          atomicAdd(&blk[4 * d.x + 2], (double)sz);
          atomicAdd(&blk[4 * d.x + 3], (double)sr);
        }
      }
      continue;
    }
    const float av = vals[d.z * nth], bv = vals[d.w * nth];
    const float kk = th[4 * np + d.y];
    const float r = fmaxf(fabsf(kk), 1.0e-6f);
    const float ux = fmaxf(r - av, 0.0f), uy = fmaxf(r - bv, 0.0f);
    const float m = fminf(av, bv);
    const float L = fsqrt(ux * ux + uy * uy);
    float da = (m > r && av <= bv) ? 1.0f : 0.0f, db = (m > r && !(av <= bv)) ? 1.0f : 0.0f;
    float dr = m > r ? 0.0f : 1.0f;
    if (L > 0.0f) {
      da += ux / L;
      db += uy / L;
      dr -= (ux + uy) / L;
    }
    vals[d.z * nth] = adj * da;  // children are visited after their parent
    vals[d.w * nth] = adj * db;
    if (acc) {
      const float dk = fabsf(kk) > 1.0e-6f ? (kk > 0.0f ? dr : -dr) : 0.0f;
      const float s = wave_sum(seed * adj * dk);
      if (lane0) atomicAdd(&blk[4 * np + d.y], (double)s);
    }
  }
  return f;
}

__device__ __forceinline__ bool diff_item(const DiffArgs& a, int& view, int& tile, CamArgs& cam) {
  const int w = __builtin_amdgcn_readfirstlane(wave_index());
  view = w / a.tiles_per_view;
  if (view >= a.n_views) return false;
  tile = w - view * a.tiles_per_view;
  cam = load_uniform(uniform_ptr(a.cams) + view);
  return true;
}

// Forward: t*, s per pixel (t* = -1, s = 0 on a miss).  kCounted adds the work of the
// image's lanes to *flops.
template <bool kCounted>
__device__ __forceinline__ void diff_forward_body(DiffArgs a, float* __restrict__ t_out, float* __restrict__ s_out,
                                                  unsigned long long* flops) {
  int view, tile;
  CamArgs cam;
  if (!diff_item(a, view, tile, cam)) return;
  int x, y;
  const bool inside = tile_xy(cam.width, cam.height, tile, x, y);
  float* vals = diff_vals(a);
  float3 o, d;
  primary_ray(cam, inside ? x : 0, inside ? y : 0, 0, o, d);
  unsigned n_evals = 0;
  const DiffEval ev{&a, vals, &n_evals};
  Hit hit;
  const float bmin[3] = {a.bounds[0], a.bounds[1], a.bounds[2]};
  const float bmax[3] = {a.bounds[3], a.bounds[4], a.bounds[5]};
  const bool h = raymarch_core(ev, bmin, bmax, o, d, a.max_steps, a.eps, a.step_scale, 0.0f, a.max_dist, hit);
  // every lane runs the remaining evaluations (the tree walk is wave-uniform)
  float3 g;
  const float3 xh = add(o, mul(d, h ? hit.t : 0.0f));
  const float f = diff_reverse(a, vals, xh, 0.0f, false, g);
  const float gd = dot(g, d);
  const float ts = h ? (fabsf(gd) > 1.0e-6f ? hit.t - f / gd : hit.t) : -1.0f;
  const float3 xs = add(o, mul(d, h ? ts : 0.0f));
  const float e = a.normal_step;
  const float gx = diff_eval(a, vals, mk(xs.x + e, xs.y, xs.z)) - diff_eval(a, vals, mk(xs.x - e, xs.y, xs.z));
  const float gy = diff_eval(a, vals, mk(xs.x, xs.y + e, xs.z)) - diff_eval(a, vals, mk(xs.x, xs.y - e, xs.z));
  const float gz = diff_eval(a, vals, mk(xs.x, xs.y, xs.z + e)) - diff_eval(a, vals, mk(xs.x, xs.y, xs.z - e));
  const float3 G = mk(gx, gy, gz);
  const float gl = len(G);
  const float3 n = gl > 1.0e-12f ? mul(G, 1.0f / gl) : mk(0.0f, 1.0f, 0.0f);
  if (kCounted) diff_count(a, inside ? n_evals + 6 : 0, inside ? 1 : 0, 0, flops);
  if (!inside) return;
  const size_t idx = (size_t)view * cam.width * cam.height + (size_t)y * cam.width + x;
  t_out[idx] = ts;
  s_out[idx] = h ? fmaxf(dot(n, mk(a.lx, a.ly, a.lz)), 0.0f) : 0.0f;
}

// A pixel takes part in the gradient when it and its 8 neighbours hit (silhouette pixels
// are not differentiable by the implicit function theorem).
__device__ __forceinline__ bool interior_hit(const float* __restrict__ t, int w, int h, int x, int y) {
  if (x < 1 || y < 1 || x >= w - 1 || y >= h - 1) return false;
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx)
      if (!(t[(size_t)(y + dy) * w + (x + dx)] >= 0.0f)) return false;
  return true;
}

// Adjoint: accumulates sum over pixels of dL/dth into per-block partial sums (FP64).
// kCounted adds the work of the lanes with a non-zero upstream gradient to *flops.
template <bool kCounted>
__device__ __forceinline__ void diff_adjoint_body(DiffArgs a, const float* __restrict__ t_img,
                                                  const float* __restrict__ gt_img, const float* __restrict__ gs_img,
                                                  double* __restrict__ partials, unsigned long long* flops) {
  const int nth = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  double* blk = diff_acc();
  for (int j = tid; j < a.n_params; j += nth) blk[j] = 0.0;
  __syncthreads();
  int view, tile;
  CamArgs cam;
  if (diff_item(a, view, tile, cam)) {
    int x, y;
    const bool inside = tile_xy(cam.width, cam.height, tile, x, y);
    const size_t base = (size_t)view * cam.width * cam.height;
    float gt = 0.0f, gs = 0.0f, ts = 0.0f;
    if (inside && interior_hit(t_img + base, cam.width, cam.height, x, y)) {
      const size_t idx = base + (size_t)y * cam.width + x;
      ts = t_img[idx];
      gt = gt_img ? gt_img[idx] : 0.0f;
      gs = gs_img ? gs_img[idx] : 0.0f;
    }
    const bool active = gt != 0.0f || gs != 0.0f;
    if (__any(active)) {
      float* vals = diff_vals(a);
      float3 o, d;
      primary_ray(cam, inside ? x : 0, inside ? y : 0, 0, o, d);
      const float3 xs = add(o, mul(d, ts));
      const float e = a.normal_step;
      const float3 l = mk(a.lx, a.ly, a.lz);
      float fv[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const float off = (k & 1) ? -e : e;
        const int axis = k >> 1;
        fv[k] = diff_eval(a, vals, mk(xs.x + (axis == 0 ? off : 0.0f), xs.y + (axis == 1 ? off : 0.0f),
                                      xs.z + (axis == 2 ? off : 0.0f)));
      }
      const float3 G = mk(fv[0] - fv[1], fv[2] - fv[3], fv[4] - fv[5]);
      const float gl = len(G);
      float3 v = mk(0.0f, 0.0f, 0.0f);
      if (gl > 1.0e-12f) {
        const float3 n = mul(G, 1.0f / gl);
        const float nl = dot(n, l);
        if (nl > 0.0f) v = mul(sub(l, mul(n, nl)), gs / gl);
      }
      float c = 0.0f;
      float3 g;
#pragma unroll 1
      for (int k = 0; k < 6; ++k) {
        const float off = (k & 1) ? -e : e;
        const int axis = k >> 1;
        const float vk = axis == 0 ? v.x : (axis == 1 ? v.y : v.z);
        const float seed = (k & 1) ? -vk : vk;
        diff_reverse(a, vals,
                     mk(xs.x + (axis == 0 ? off : 0.0f), xs.y + (axis == 1 ? off : 0.0f), xs.z + (axis == 2 ? off : 0.0f)),
                     active ? seed : 0.0f, true, g);
        c += seed * dot(g, d);
      }
      diff_reverse(a, vals, xs, 0.0f, false, g);
      const float gd = dot(g, d);
      const float s0 = (active && fabsf(gd) > 1.0e-6f) ? -(gt + c) / gd : 0.0f;
      diff_reverse(a, vals, xs, s0, true, g);
      if (kCounted) diff_count(a, active ? 6 : 0, active ? 1 : 0, active ? 7 : 0, flops);
    }
  }
  __syncthreads();
  for (int j = tid; j < a.n_params; j += nth) partials[(size_t)blockIdx.x * a.n_params + j] = blk[j];
}

// L2 loss against target images on pixels interior to both hit masks:
//   L = sum w_t (t* - T)^2 + w_s (s - S)^2,  gt = dL/dt*, gs = dL/ds.
__device__ __forceinline__ void diff_loss_body(int width, int height, int n_views, const float* __restrict__ t,
                                               const float* __restrict__ s, const float* __restrict__ tt,
                                               const float* __restrict__ st, float wt, float ws,
                                               float* __restrict__ gt, float* __restrict__ gs,
                                               double* __restrict__ partials) {
  __shared__ double red[256];
  const size_t npix = (size_t)width * height;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (i < npix * n_views) {
    const size_t view = i / npix, pix = i - view * npix;
    const int y = (int)(pix / width), x = (int)(pix - (size_t)y * width);
    const float* tv = t + view * npix;
    const float* ttv = tt + view * npix;
    float a = 0.0f, b = 0.0f;
    if (interior_hit(tv, width, height, x, y) && interior_hit(ttv, width, height, x, y)) {
      const float dt = t[i] - tt[i], ds = s[i] - st[i];
      a = 2.0f * wt * dt;
      b = 2.0f * ws * ds;
      l = (double)wt * dt * dt + (double)ws * ds * ds;
    }
    gt[i] = a;
    gs[i] = b;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int m = blockDim.x / 2; m > 0; m >>= 1) {
    if ((int)threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// out[g][j] = sum of in[r][j] over rows r of group g (rows_per_group rows), fixed order.
__device__ __forceinline__ void diff_reduce_body(const double* __restrict__ in, int rows, int cols, int rows_per_group,
                                                 double* __restrict__ out) {
  const int j = blockIdx.y * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  const int r0 = blockIdx.x * rows_per_group;
  const int r1 = min(rows, r0 + rows_per_group);
  double s = 0.0;
  for (int r = r0; r < r1; ++r) s += in[(size_t)r * cols + j];
  out[(size_t)blockIdx.x * cols + j] = s;
}

// Adam on every parameter (one block), then the scene AABB from the new parameters.
__device__ __forceinline__ void diff_adam_body(float* __restrict__ th, const double* __restrict__ grad,
                                               double* __restrict__ m, double* __restrict__ v, int n_params,
                                               int n_prims, int step, float lr, float b1, float b2, float eps_adam,
                                               float* __restrict__ bounds) {
  for (int j = threadIdx.x; j < n_params; j += blockDim.x) {
    const double g = grad[j];
    m[j] = b1 * m[j] + (1.0 - b1) * g;
    v[j] = b2 * v[j] + (1.0 - b2) * g * g;
    const double mh = m[j] / (1.0 - pow((double)b1, (double)step));
    const double vh = v[j] / (1.0 - pow((double)b2, (double)step));
    th[j] = (float)(th[j] - lr * mh / (sqrt(vh) + eps_adam));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float kmax = 0.0f;
    for (int j = 4 * n_prims; j < n_params; ++j) kmax = fmaxf(kmax, fabsf(th[j]));
    const float mg = kmax + 1.0e-3f;
    float lo[3] = {__builtin_huge_valf(), __builtin_huge_valf(), __builtin_huge_valf()};
    float hi[3] = {-__builtin_huge_valf(), -__builtin_huge_valf(), -__builtin_huge_valf()};
    for (int i = 0; i < n_prims; ++i) {
      const float r = fabsf(th[4 * i + 3]) + mg;
      for (int k = 0; k < 3; ++k) {
        lo[k] = fminf(lo[k], th[4 * i + k] - r);
        hi[k] = fmaxf(hi[k], th[4 * i + k] + r);
      }
    }
    for (int k = 0; k < 3; ++k) {
      bounds[k] = lo[k];
      bounds[3 + k] = hi[k];
    }
  }
}

}  // namespace ftd

#define FT_DIFF_KERNEL extern "C" __global__ __launch_bounds__(128)
FT_DIFF_KERNEL void ft_k_diff_forward(ftd::DiffArgs a, float* __restrict__ t, float* __restrict__ s) {
  ftd::diff_forward_body<false>(a, t, s, nullptr);
}
FT_DIFF_KERNEL void ft_k_diff_adjoint(ftd::DiffArgs a, const float* __restrict__ t, const float* __restrict__ gt,
                                      const float* __restrict__ gs, double* __restrict__ partials) {
  ftd::diff_adjoint_body<false>(a, t, gt, gs, partials, nullptr);
}
// FLOP-counting variants (named apart so kernel traces keep them out of the timed runs)
FT_DIFF_KERNEL void ft_k_diff_forward_counted(ftd::DiffArgs a, float* __restrict__ t, float* __restrict__ s,
                                              unsigned long long* flops) {
  ftd::diff_forward_body<true>(a, t, s, flops);
}
FT_DIFF_KERNEL void ft_k_diff_adjoint_counted(ftd::DiffArgs a, const float* __restrict__ t,
                                              const float* __restrict__ gt, const float* __restrict__ gs,
                                              double* __restrict__ partials, unsigned long long* flops) {
  ftd::diff_adjoint_body<true>(a, t, gt, gs, partials, flops);
}
extern "C" __global__ __launch_bounds__(256) void ft_k_diff_loss(int width, int height, int n_views,
                                                                 const float* __restrict__ t, const float* __restrict__ s,
                                                                 const float* __restrict__ tt, const float* __restrict__ st,
                                                                 float wt, float ws, float* __restrict__ gt,
                                                                 float* __restrict__ gs, double* __restrict__ partials) {
  ftd::diff_loss_body(width, height, n_views, t, s, tt, st, wt, ws, gt, gs, partials);
}
extern "C" __global__ __launch_bounds__(256) void ft_k_diff_reduce(const double* __restrict__ in, int rows, int cols,
                                                                   int rows_per_group, double* __restrict__ out) {
  ftd::diff_reduce_body(in, rows, cols, rows_per_group, out);
}
extern "C" __global__ __launch_bounds__(256) void ft_k_diff_adam(float* __restrict__ th, const double* __restrict__ grad,
                                                                 double* __restrict__ m, double* __restrict__ v,
                                                                 int n_params, int n_prims, int step, float lr, float b1,
                                                                 float b2, float eps_adam, float* __restrict__ bounds) {
  ftd::diff_adam_body(th, grad, m, v, n_params, n_prims, step, lr, b1, b2, eps_adam, bounds);
}
