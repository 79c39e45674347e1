// ft_march.h — sphere tracing of one ray against any distance evaluator (raymarch_hit,
// refine_hit_distance and estimate_normal of the reference).
#pragma once

#include "ft_math.h"

namespace ftd {

struct Hit {
  float t;
  float3 pos, n;
  bool front;
  unsigned obj, mat;
  int leaf;
};

// ray_aabb_intersection (renderer.rs:4141-4173)
__device__ __forceinline__ bool ray_aabb(float3 o, float3 d, const float* mn, const float* mx, float& t0, float& t1) {
  float tmin = -__builtin_huge_valf(), tmax = __builtin_huge_valf();
  for (int axis = 0; axis < 3; ++axis) {
    float oo = comp(o, axis), dd = comp(d, axis);
    if (fabsf(dd) <= 1.0e-8f) {
      if (oo < mn[axis] || oo > mx[axis]) return false;
      continue;
    }
    float inv = 1.0f / dd;
    float a = (mn[axis] - oo) * inv, c = (mx[axis] - oo) * inv;
    if (a > c) { float t = a; a = c; c = t; }
    tmin = fmaxf(tmin, a);
    tmax = fminf(tmax, c);
    if (tmax < tmin) return false;
  }
  t0 = tmin;
  t1 = tmax;
  return true;
}

// raymarch_hit + refine_hit_distance + estimate_normal (renderer.rs:4091-4215).
// One evaluation site for the distance: the march, the 10 bisection steps and the 6 central
// differences are phases of a single loop, so a specialised scene function is inlined once
// and lanes in different phases still share every evaluation; the final distance_info at
// the hit (which also names the primitive) is taken once per ray after the loop.
// No coarse distance: every march step evaluates the scene.
#ifndef FT_MARCH_UNROLL
#define FT_MARCH_UNROLL 1
#endif

struct NoCoarse {
  __device__ __forceinline__ bool enabled() const { return false; }
  __device__ __forceinline__ float operator()(float3) const { return -1.0f; }
  float switch_dist = 0.0f;
};

// `ev(p)` returns the scene's distance info at p (fields d, obj, mat, leaf), `ev.dist(p)`
// its distance alone; bmin/bmax are the scene bounds the ray is clipped to.  `coarse(p)`,
// when enabled, returns an approximate distance (or a negative value where it has none):
// while it exceeds
// coarse.switch_dist the march steps on it without evaluating the scene; from the first
// step below, the lane marches, refines and shades on exact evaluations only.
template <class Eval, class Coarse = NoCoarse>
__device__ __forceinline__ bool raymarch_core(const Eval& ev, const float* bmin, const float* bmax, float3 o, float3 d,
                                              int max_steps, float eps, float step_scale, float min_t, float max_t,
                                              Hit& hit, const Coarse& coarse = Coarse()) {
  float entry, exit;
  if (!ray_aabb(o, d, bmin, bmax, entry, exit)) return false;
  max_t = fminf(max_t, exit);
  float traveled = fmaxf(min_t, fmaxf(entry, 0.0f));
  float prev = traveled;
  const float sscale = clampf(step_scale, 0.05f, 1.0f);
  const float min_step = fmaxf(eps * 0.5f, 1.0e-5f);
  const float e = fmaxf(eps * 4.0f, 1.0e-5f);
  int phase = 0, k = 0, steps = 0;
  float lo = 0.0f, hi = 0.0f, mid = 0.0f, dplus = 0.0f;
  float3 pos = mk(0, 0, 0), grad = mk(0, 0, 0);
  bool exact = !coarse.enabled();
  // One march step (or bisection step, or normal sample): 0 go on, 1 miss, 2 done.
  auto step = [&]() -> int {
    float3 p;
    if (phase == 0) {
      if (steps >= max_steps || traveled > max_t) return 1;
      p = add(o, mul(d, traveled));
      if (!exact) {
        const float dc = coarse(p);
        if (dc > coarse.switch_dist) {
          prev = traveled;
          traveled += fmaxf(dc * sscale, min_step);
          ++steps;
          return 0;
        }
        exact = true;
      }
    } else if (phase == 1) {
      mid = 0.5f * (lo + hi);
      p = add(o, mul(d, mid));
    } else {
      const float off = (k & 1) ? -e : e;
      const int axis = k >> 1;
      p = mk(pos.x + (axis == 0 ? off : 0.0f), pos.y + (axis == 1 ? off : 0.0f), pos.z + (axis == 2 ? off : 0.0f));
    }
    const float dist = ev.dist(p);
    if (phase == 0) {
      if (fabsf(dist) < eps) {
        phase = 1;
        k = 0;
        lo = prev;
        hi = traveled;
      } else {
        prev = traveled;
        traveled += fmaxf(fabsf(dist) * sscale, min_step);
        ++steps;
      }
    } else if (phase == 1) {
      if (fabsf(dist) < eps) hi = mid; else lo = mid;
      if (++k == 10) {
        pos = add(o, mul(d, hi));
        phase = 2;
        k = 0;
      }
    } else {
      if ((k & 1) == 0) {
        dplus = dist;
      } else {
        const float g = dplus - dist;
        const int axis = k >> 1;
        grad = mk(axis == 0 ? g : grad.x, axis == 1 ? g : grad.y, axis == 2 ? g : grad.z);
      }
      if (++k == 6) return 2;
    }
    return 0;
  };
  // FT_MARCH_UNROLL steps per loop iteration (a launch-shape autotuner parameter); the
  // step sequence of every lane is the same for any value
  for (;;) {
    int r = 0;
#pragma unroll
    for (int u = 0; u < FT_MARCH_UNROLL; ++u) {
      r = step();
      if (r != 0) break;
    }
    if (r == 1) return false;
    if (r == 2) break;
  }
  const auto inf = ev(pos);
  const float3 nrm = len(grad) > 1.0e-8f ? normalize(grad) : mk(0.0f, 1.0f, 0.0f);
  hit.t = hi;
  hit.pos = pos;
  hit.n = nrm;
  hit.front = dot(nrm, d) < 0.0f;
  hit.obj = inf.obj;
  hit.mat = inf.mat;
  hit.leaf = inf.leaf;
  return true;
}

}  // namespace ftd
