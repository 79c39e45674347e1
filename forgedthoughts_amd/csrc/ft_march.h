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
// One ray's march as a resumable state machine, so that a lane can take a new ray between
// steps (the persistent batch kernel); raymarch_core runs it to the end for one ray.
struct MarchLane {
  float3 o, d;
  float traveled, prev, max_t, sscale, min_step, e, eps;
  int max_steps, phase, k, steps;
  float lo, hi, mid, dplus;
  float3 pos, grad;
  bool exact;

  // false: the ray misses the bounds (no step to take)
  template <class Coarse>
  __device__ __forceinline__ bool init(const float* bmin, const float* bmax, float3 o_, float3 d_, int max_steps_,
                                       float eps_, float step_scale, float min_t, float max_t_, const Coarse& coarse) {
    float entry, exit;
    if (!ray_aabb(o_, d_, bmin, bmax, entry, exit)) return false;
    o = o_;
    d = d_;
    max_steps = max_steps_;
    eps = eps_;
    max_t = fminf(max_t_, exit);
    traveled = fmaxf(min_t, fmaxf(entry, 0.0f));
    prev = traveled;
    sscale = clampf(step_scale, 0.05f, 1.0f);
    min_step = fmaxf(eps * 0.5f, 1.0e-5f);
    e = fmaxf(eps * 4.0f, 1.0e-5f);
    phase = 0;
    k = 0;
    steps = 0;
    lo = 0.0f;
    hi = 0.0f;
    mid = 0.0f;
    dplus = 0.0f;
    pos = mk(0, 0, 0);
    grad = mk(0, 0, 0);
    exact = !coarse.enabled();
    return true;
  }

  // The state a march reaches when its bisection ends at t = hi: the normal samples remain
  // (the batch Whitted pass resumes here from the depth pass's primary hits).
  __device__ __forceinline__ void resume_at_hit(float3 o_, float3 d_, float eps_, float hi_) {
    o = o_;
    d = d_;
    eps = eps_;
    e = fmaxf(eps * 4.0f, 1.0e-5f);
    hi = hi_;
    pos = add(o, mul(d, hi));
    phase = 2;
    k = 0;
    grad = mk(0, 0, 0);
  }

  // The hit record once step() has returned 2.
  template <class Eval>
  __device__ __forceinline__ void finish(const Eval& ev, Hit& hit) const {
    const auto inf = ev(pos);
    const float3 nrm = len(grad) > 1.0e-8f ? normalize(grad) : mk(0.0f, 1.0f, 0.0f);
    hit.t = hi;
    hit.pos = pos;
    hit.n = nrm;
    hit.front = dot(nrm, d) < 0.0f;
    hit.obj = inf.obj;
    hit.mat = inf.mat;
    hit.leaf = inf.leaf;
  }

  // One march step (or bisection step, or normal sample): 0 go on, 1 miss, 2 done (hit at
  // t = hi, position pos, unnormalised normal grad).
  template <class Eval, class Coarse>
  __device__ __forceinline__ int step(const Eval& ev, const Coarse& coarse) {
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
  }
};

template <class Eval, class Coarse = NoCoarse>
__device__ __forceinline__ bool raymarch_core(const Eval& ev, const float* bmin, const float* bmax, float3 o, float3 d,
                                              int max_steps, float eps, float step_scale, float min_t, float max_t,
                                              Hit& hit, const Coarse& coarse = Coarse()) {
  MarchLane m;
  if (!m.init(bmin, bmax, o, d, max_steps, eps, step_scale, min_t, max_t, coarse)) return false;
  // FT_MARCH_UNROLL steps per loop iteration (a launch-shape autotuner parameter); the
  // step sequence of every lane is the same for any value
  for (;;) {
    int r = 0;
#pragma unroll
    for (int u = 0; u < FT_MARCH_UNROLL; ++u) {
      r = m.step(ev, coarse);
      if (r != 0) break;
    }
    if (r == 1) return false;
    if (r == 2) break;
  }
  m.finish(ev, hit);
  return true;
}

// The hit of a ray whose march is known to end at t (t < 0: a miss) without marching it
// again: the normal samples and the final distance_info of raymarch_core at that t.
template <class Eval>
__device__ __forceinline__ bool hit_from_t(const Eval& ev, float3 o, float3 d, float eps, float t, Hit& hit) {
  if (!(t >= 0.0f)) return false;
  MarchLane m;
  m.resume_at_hit(o, d, eps, t);
  while (m.step(ev, NoCoarse()) == 0) {
  }
  m.finish(ev, hit);
  return true;
}

}  // namespace ftd
