// ft_math.h — scalar/vector helpers, hg_sdf operators, noise and primitive SDFs shared by
// every gfx950 kernel (scene-independent: no scene table, no LDS).
#pragma once

#if !defined(__HIPCC_RTC__)
#include <hip/hip_runtime.h>
#endif

#include "ft_layout.h"

namespace ftd {

// ----------------------------------------------------------------- math
__device__ __forceinline__ float3 mk(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 add(float3 a, float3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 sub(float3 a, float3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 mul(float3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 cmul(float3 a, float3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Square root on the hardware instruction (v_sqrt_f32, within 1 ulp) rather than the
// correctly rounded sequence sqrtf expands to (about fifteen instructions, most of a
// primitive's cost).  Distances differ from IEEE sqrt by at most an ulp; arguments here are
// sums of squares far above the denormal range or exactly zero.
#if defined(FT_STRICT_MATH)
// strict arithmetic (FT_STRICT_MATH=1 at run time): correctly rounded square roots and no
// multiply-add contraction, the oracle's arithmetic
__device__ __forceinline__ float fsqrt(float x) { return sqrtf(x); }
#else
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
#endif
__device__ __forceinline__ float len(float3 a) { return fsqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ float3 cross(float3 a, float3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float3 normalize(float3 a) {  // Vec3::normalize (renderer.rs:317-323)
  float l = len(a);
  if (l <= 1.1920929e-7f) return a;
  return mul(a, 1.0f / l);
}
__device__ __forceinline__ float clampf(float x, float lo, float hi) {
  x = x < lo ? lo : x;
  return x > hi ? hi : x;
}
__device__ __forceinline__ float comp(float3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
__device__ __forceinline__ float deg2rad(float d) { return d * (3.14159265358979323846f / 180.0f); }
__device__ __forceinline__ float powi5(float x) { float x2 = x * x; return x * (x2 * x2); }
__device__ __forceinline__ float rem_euclid(float a, float b) {
  float r = fmodf(a, b);
  return r < 0.0f ? r + fabsf(b) : r;
}

__device__ __forceinline__ float3 rot_x(float3 v, float s, float c) { return mk(v.x, c * v.y - s * v.z, s * v.y + c * v.z); }
__device__ __forceinline__ float3 rot_y(float3 v, float s, float c) { return mk(c * v.x + s * v.z, v.y, -s * v.x + c * v.z); }
__device__ __forceinline__ float3 rot_z(float3 v, float s, float c) { return mk(c * v.x - s * v.y, s * v.x + c * v.y, v.z); }

constexpr unsigned kAabbFlops = kAabbFlopsHost;

// point_aabb_lower_bound (renderer.rs:5388-5411): per axis, lo - p below the box, p - hi
// above it, else 0.  Written as one select per axis: p > hi exactly when p - hi > 0 (a
// difference of distinct finite floats is never zero; NaN compares false and fmaxf drops
// it), so max(p - hi, 0) is the reference's inner branch up to the sign of a zero, which the
// square removes.  Every operand is computed, so the compiler emits v_cndmask rather than
// the exec-mask branches it made of the nested conditionals (eight scalar instructions per
// axis, which made the interpreted BVH walks scalar-issue bound).
__device__ __forceinline__ float aabb_gap(float v, float lo, float hi) {
  const float below = lo - v, above = fmaxf(v - hi, 0.0f);
  return v < lo ? below : above;
}
__device__ __forceinline__ float aabb_lb(float3 p, const float* mn, const float* mx) {
  const float dx = aabb_gap(p.x, mn[0], mx[0]);
  const float dy = aabb_gap(p.y, mn[1], mx[1]);
  const float dz = aabb_gap(p.z, mn[2], mx[2]);
  return fsqrt(dx * dx + dy * dy + dz * dz);
}

__device__ __forceinline__ float aabb_lb_lit(float3 p, float x0, float y0, float z0, float x1, float y1, float z1) {
  const float mn[3] = {x0, y0, z0}, mx[3] = {x1, y1, z1};
  return aabb_lb(p, mn, mx);
}

// --------------------------------------------------------- hg_sdf ops
// (renderer.rs:6011-6137)
__device__ __forceinline__ float op_smax(float a, float b, float k) {
  float h = fmaxf(k - fabsf(a - b), 0.0f);
  return fmaxf(a, b) + h * h * 0.25f / fmaxf(k, 1.0e-6f);
}
__device__ __forceinline__ float op_smin(float a, float b, float k) {
  float h = fmaxf(k - fabsf(a - b), 0.0f);
  return fminf(a, b) - h * h * 0.25f / fmaxf(k, 1.0e-6f);
}
__device__ __forceinline__ float op_union_round(float a, float b, float r) {
  r = fmaxf(fabsf(r), 1.0e-6f);
  float ux = fmaxf(r - a, 0.0f), uy = fmaxf(r - b, 0.0f);
  return fmaxf(r, fminf(a, b)) - fsqrt(ux * ux + uy * uy);
}
__device__ __forceinline__ float op_intersect_round(float a, float b, float r) {
  r = fmaxf(fabsf(r), 1.0e-6f);
  float ux = fmaxf(r + a, 0.0f), uy = fmaxf(r + b, 0.0f);
  return fminf(-r, fmaxf(a, b)) + fsqrt(ux * ux + uy * uy);
}
constexpr float kInvSqrt2 = 0.70710678118654752440f;
__device__ __forceinline__ float op_union_chamfer(float a, float b, float r) {
  return fminf(fminf(a, b), (a - r + b) * kInvSqrt2);
}
__device__ __forceinline__ float op_intersect_chamfer(float a, float b, float r) {
  return fmaxf(fmaxf(a, b), (a + r + b) * kInvSqrt2);
}
__device__ __forceinline__ float op_union_columns(float a, float b, float r, float n) {
  if (a < r && b < r) {
    n = fmaxf(roundf(n), 2.0f);
    const float s2 = 1.41421356237309504880f;
    float cr = r * s2 / ((n - 1.0f) * 2.0f + s2);
    float x = a + cr, y = b + cr;
    int ni = (int)n;
    if (ni % 2 == 1) x += cr;
    float rx = (x + y) * kInvSqrt2, ry = (y - x) * kInvSqrt2;
    x = rx;
    y = ry;
    x -= s2 * 0.5f * r;
    x += -cr * s2;
    if (ni % 2 == 1) y += cr;
    float size = fmaxf(cr * 2.0f, 1.0e-6f);
    y = rem_euclid(y + size * 0.5f, size) - size * 0.5f;
    return fsqrt(x * x + y * y) - cr;
  }
  return fminf(a, b);
}
__device__ __forceinline__ float op_union_stairs(float a, float b, float r, float n) {
  n = fmaxf(roundf(n), 1.0f);
  float s = r / n, u = b - r;
  return fminf(fminf(a, b), 0.5f * (u + a + fabsf(rem_euclid(u - a + s, 2.0f * s) - s)));
}
__device__ __forceinline__ float op_union_soft(float a, float b, float r) {
  r = fmaxf(fabsf(r), 1.0e-6f);
  float e = fmaxf(r - fabsf(a - b), 0.0f);
  return fminf(a, b) - e * e * 0.25f / r;
}

// noise modifier (renderer.rs:5957-6009)
__device__ __forceinline__ float fractf_(float x) { return x - floorf(x); }
constexpr float kInvPi = 0.31830988618379067154f;
// The hash multiplies lattice coordinates up to ~2.5e5 before taking the fraction, so one
// rounding step decides its value: no multiply-add fusion here (the reference's Rust never
// fuses), or the noise field would differ from the reference's almost everywhere.
__device__ __forceinline__ float hash_noise(float3 p) {
#pragma clang fp contract(off)
  float qx = fractf_(p.x * kInvPi + 0.11f) * 17.0f;
  float qy = fractf_(p.y * kInvPi + 0.17f) * 17.0f;
  float qz = fractf_(p.z * kInvPi + 0.13f) * 17.0f;
  return fractf_(qx * qy * qz * (qx + qy + qz));
}
__device__ inline float sd_base_noise(float3 p) {
  p = add(p, mk(0.5f, 0.5f, 0.5f));
  float3 i = mk(floorf(p.x), floorf(p.y), floorf(p.z));
  float3 f = mk(fractf_(p.x), fractf_(p.y), fractf_(p.z));
  float d = __builtin_huge_valf();
  for (int ix = 0; ix <= 1; ++ix)
    for (int iy = 0; iy <= 1; ++iy)
      for (int iz = 0; iz <= 1; ++iz) {
        float3 c = mk((float)ix, (float)iy, (float)iz);
        float r = hash_noise(add(i, c));
        r = r * r * 0.7f;
        d = fminf(d, len(sub(f, c)) - r);
      }
  return d;
}
__device__ inline float apply_noise(float3 p, float d, int octaves, float scale, float lac) {
  float3 q = mul(p, fmaxf(scale, 1.0e-4f));
  float os = fmaxf(lac, 1.0e-4f);
  float s = 1.0f, t = 0.0f;
  int n = octaves < 1 ? 1 : octaves;
  for (int k = 0; k < n; ++k) {
    float nn = s * sd_base_noise(q);
    nn = op_smax(nn, d - 0.1f * s, 0.3f * s);
    d = op_smin(nn, d, 0.3f * s);
    t += d;
    q = mul(mk(0.00f * q.x + 1.60f * q.y + 1.20f * q.z, -1.60f * q.x + 0.72f * q.y - 0.96f * q.z,
               -1.20f * q.x - 0.96f * q.y + 1.28f * q.z), os);
    s *= 0.415f;
    q.z += -4.33f * t * s;
  }
  return d;
}

// sd_regular_ngon (renderer.rs:6206-6214)
__device__ inline float sd_ngon(float3 p, int sides, float radius) {
  float n = (float)(sides < 3 ? 3 : sides);
  float an = 3.14159265358979323846f / n;
  float apothem = radius * cosf(an);
  float angle = atan2f(p.z, p.x);
  float wrapped = rem_euclid(angle + an, 2.0f * an) - an;
  return hypotf(p.x, p.z) * cosf(wrapped) - apothem;
}

// ------------------------------------------------------ Forge functions
// Expression DAG interpreter; builtin semantics eval.rs:2223-2690.
__device__ __forceinline__ float smooth01(float t) {
  t = clampf(t, 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
__device__ inline float hash_noise3(float3 p) { return hash_noise(p); }  // eval.rs:4286-4291 (same constants)
__device__ inline float lerpf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ inline float value_noise_3d(float3 p) {
  float3 i = mk(floorf(p.x), floorf(p.y), floorf(p.z));
  float3 f = mk(fractf_(p.x), fractf_(p.y), fractf_(p.z));
  float3 u = mk(smooth01(f.x), smooth01(f.y), smooth01(f.z));
  float n000 = hash_noise3(i), n001 = hash_noise3(mk(i.x, i.y, i.z + 1.0f));
  float n010 = hash_noise3(mk(i.x, i.y + 1.0f, i.z)), n011 = hash_noise3(mk(i.x, i.y + 1.0f, i.z + 1.0f));
  float n100 = hash_noise3(mk(i.x + 1.0f, i.y, i.z)), n101 = hash_noise3(mk(i.x + 1.0f, i.y, i.z + 1.0f));
  float n110 = hash_noise3(mk(i.x + 1.0f, i.y + 1.0f, i.z)), n111 = hash_noise3(mk(i.x + 1.0f, i.y + 1.0f, i.z + 1.0f));
  float nxy0 = lerpf(lerpf(n000, n100, u.x), lerpf(n010, n110, u.x), u.y);
  float nxy1 = lerpf(lerpf(n001, n101, u.x), lerpf(n011, n111, u.x), u.y);
  return lerpf(nxy0, nxy1, u.z) * 2.0f - 1.0f;
}
__device__ inline float fbm_3d(float3 p, int octaves, float scale, float lac) {
  float3 q = mul(p, scale);
  float amp = 0.5f, sum = 0.0f;
  float l = fabsf(lac) < 1.1920929e-7f ? 1.0f : lac;
  int n = octaves < 1 ? 1 : octaves;
  for (int k = 0; k < n; ++k) {
    sum += amp * value_noise_3d(q);
    q = mul(q, l);
    amp *= 0.55f;
  }
  return sum;
}


// Forge builtins shared by the expression interpreter below and the emitted functions.
__device__ __forceinline__ float3 splat3(float x) { return mk(x, x, x); }
__device__ __forceinline__ float saturatef_(float x) { return clampf(x, 0.0f, 1.0f); }
__device__ __forceinline__ float fclamp_any(float x, float a, float b) {  // clamp with either bound order
  return a <= b ? clampf(x, a, b) : clampf(x, b, a);
}
__device__ __forceinline__ float smoothstep_builtin(float e0, float e1, float x) {
  const float span = e1 - e0;
  float t;
  if (fabsf(span) < 1.1920929e-7f) t = x < e0 ? 0.0f : 1.0f;
  else t = clampf((x - e0) / span, 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
__device__ __forceinline__ float3 normalize_div(float3 a) {
  const float l = len(a);
  return l <= 1.1920929e-7f ? mk(0, 0, 0) : mk(a.x / l, a.y / l, a.z / l);
}
__device__ __forceinline__ float3 rotate_deg_x(float3 a, float deg) { const float r = deg2rad(deg); return rot_x(a, sinf(r), cosf(r)); }
__device__ __forceinline__ float3 rotate_deg_y(float3 a, float deg) { const float r = deg2rad(deg); return rot_y(a, sinf(r), cosf(r)); }
__device__ __forceinline__ float3 rotate_deg_z(float3 a, float deg) { const float r = deg2rad(deg); return rot_z(a, sinf(r), cosf(r)); }
__device__ __forceinline__ float box_dist_builtin(float3 a, float3 b) {
  const float3 q = mk(fabsf(a.x) - b.x, fabsf(a.y) - b.y, fabsf(a.z) - b.z);
  return len(mk(fmaxf(q.x, 0), fmaxf(q.y, 0), fmaxf(q.z, 0))) + fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f);
}
__device__ __forceinline__ float cylinder_dist_builtin(float3 a, float radius, float half_height) {
  const float radial = fsqrt(a.x * a.x + a.z * a.z) - radius, axial = fabsf(a.y) - half_height;
  const float ox = fmaxf(radial, 0.0f), oy = fmaxf(axial, 0.0f);
  return fsqrt(ox * ox + oy * oy) + fminf(fmaxf(radial, axial), 0.0f);
}
__device__ __forceinline__ float torus_dist_builtin(float3 a, float major, float minor) {
  const float qx = fsqrt(a.x * a.x + a.z * a.z) - major;
  return fsqrt(qx * qx + a.y * a.y) - minor;
}


__device__ __forceinline__ float shell_op(float d, float shell) { return shell > 0.0f ? fmaxf(d, -(d + shell)) : d; }

// primitive distances in the local frame (renderer.rs:5545-5663)
__device__ __forceinline__ float sdf_sphere(float3 q, float radius, float shell) {
  return shell_op(len(q) - radius, shell);
}
__device__ __forceinline__ float sdf_box(float3 q, float hx, float hy, float hz, float rd, float shell) {
  float3 a = mk(fabsf(q.x) - fmaxf(hx - rd, 0.0f), fabsf(q.y) - fmaxf(hy - rd, 0.0f), fabsf(q.z) - fmaxf(hz - rd, 0.0f));
  float outside = len(mk(fmaxf(a.x, 0.0f), fmaxf(a.y, 0.0f), fmaxf(a.z, 0.0f)));
  float inside = fminf(fmaxf(fmaxf(a.x, a.y), a.z), 0.0f);
  return shell_op(outside + inside - rd, shell);
}
__device__ __forceinline__ float sdf_cylinder(float3 q, float radius, float hh, float rd, float shell) {
  float dx = fsqrt(q.x * q.x + q.z * q.z) - fmaxf(radius - rd, 0.0f);
  float dy = fabsf(q.y) - fmaxf(hh - rd, 0.0f);
  float ox = fmaxf(dx, 0.0f), oy = fmaxf(dy, 0.0f);
  return shell_op(fsqrt(ox * ox + oy * oy) + fminf(fmaxf(dx, dy), 0.0f) - rd, shell);
}
__device__ __forceinline__ float sdf_torus(float3 q, float major, float minor) {
  float qx = fsqrt(q.x * q.x + q.z * q.z) - major;
  return fsqrt(qx * qx + q.y * q.y) - minor;
}
__device__ __noinline__ float sdf_ngon(float3 q, int sides, float radius, float hh, float rd, float shell) {
  float radial = sd_ngon(mk(q.x, 0.0f, q.z), sides, fmaxf(radius - rd, 0.0f));
  float dy = fabsf(q.y) - fmaxf(hh - rd, 0.0f);
  float ox = fmaxf(radial, 0.0f), oy = fmaxf(dy, 0.0f);
  return shell_op(fsqrt(ox * ox + oy * oy) + fminf(fmaxf(radial, dy), 0.0f) - rd, shell);
}

}  // namespace ftd
