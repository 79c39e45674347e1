// ftref.cpp — CPU oracle of the ForgedThoughts sphere-tracing path (TEST INFRASTRUCTURE).
//
// Restates, function by function, the reference's CPU renderer on the scene IR of
// include/ft_abi.h.  Every function cites the reference code it follows (paths relative
// to crates/forgedthoughts/src; R: = renderer.rs).  It must be compiled WITHOUT
// floating-point contraction (-ffp-contract=off): Rust never fuses a*b+c, so neither
// may the restatement.  The f32 instantiation is the parity oracle; the f64
// instantiation runs the same algorithm in double precision.
//
// Only tests/, bench.py's CPU-baseline leg and the package smoke check load this file's
// library; the HIP backend never does.
//
// Pinning (tests/test_oracle_pins.py, the reference fixtures, closed forms): every function
// is checked except these, which are parity unpinned: the lattice-sphere noise of the noise
// modifier (sd_base_noise and the octave rotation; the modifier is pinned only where the
// noise drops out), the fmin / fmax NaN behaviour of the reference's JIT, and the order of
// lights and materials the reference takes from a HashMap (this file keeps source order).

#include "ftref.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

thread_local std::string g_err;

// Planted mistakes (test infrastructure).  Compiling with -DFTREF_PLANT=k builds the oracle
// with mistake k at one site, marked planted(k) below; tests/test_oracle_pins.py builds each
// variant and checks that the pin of the function holding the mistake fails on it.  The
// default build (FTREF_PLANT = 0) contains none of them.
#ifndef FTREF_PLANT
#define FTREF_PLANT 0
#endif
constexpr bool planted(int k) { return FTREF_PLANT == k; }

// ------------------------------------------------------------------- math
template <class T> struct C;
template <> struct C<float> {
  static constexpr float EPS = 1.1920929e-7f;  // f32::EPSILON
  static float deg2rad(float d) { return d * (3.14159265358979323846f / 180.0f); }
};
template <> struct C<double> {
  static constexpr double EPS = 2.220446049250313e-16;
  static double deg2rad(double d) { return d * (3.14159265358979323846 / 180.0); }
};

template <class T> inline T rmax(T a, T b) { return std::fmax(a, b); }  // f32::max
template <class T> inline T rmin(T a, T b) { return std::fmin(a, b); }  // f32::min
template <class T> inline T rclamp(T x, T lo, T hi) {                   // f32::clamp
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
template <class T> inline T powi5(T x) { T x2 = x * x; return x * (x2 * x2); }  // llvm.powi(x, 5)
template <class T> inline T rem_euclid(T a, T b) {
  T r = std::fmod(a, b);
  return r < T(0) ? r + std::fabs(b) : r;
}

template <class T> struct V3 {
  T x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(T a, T b, T c) : x(a), y(b), z(c) {}
  static V3 splat(T v) { return V3(v, v, v); }
  V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
  V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
  V3 operator*(T s) const { return V3(x * s, y * s, z * s); }
  V3 cmul(const V3& o) const { return V3(x * o.x, y * o.y, z * o.z); }
  T dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
  T length() const { return std::sqrt(x * x + y * y + z * z); }
  V3 normalize() const {  // Vec3::normalize (R:317-323, render_api.rs:24-31)
    T len = length();
    if (len <= C<T>::EPS) return *this;
    return *this * (T(1) / len);
  }
  V3 cross(const V3& o) const {
    return V3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x);
  }
  V3 abs() const { return V3(std::fabs(x), std::fabs(y), std::fabs(z)); }
  T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class T> using Spec = V3<T>;  // Spectrum r,g,b

template <class T> struct Aabb { V3<T> min, max; };

template <class T> Aabb<T> aabb_union(const Aabb<T>& a, const Aabb<T>& b) {
  return {V3<T>(rmin(a.min.x, b.min.x), rmin(a.min.y, b.min.y), rmin(a.min.z, b.min.z)),
          V3<T>(rmax(a.max.x, b.max.x), rmax(a.max.y, b.max.y), rmax(a.max.z, b.max.z))};
}

// to_local / rotate_* (R:6216-6240)
template <class T> V3<T> rotate_x(V3<T> v, T deg) {
  T r = C<T>::deg2rad(deg), s = std::sin(r), c = std::cos(r);
  return V3<T>(v.x, c * v.y - s * v.z, s * v.y + c * v.z);
}
template <class T> V3<T> rotate_y(V3<T> v, T deg) {
  T r = C<T>::deg2rad(deg), s = std::sin(r), c = std::cos(r);
  return V3<T>(c * v.x + s * v.z, v.y, -s * v.x + c * v.z);
}
template <class T> V3<T> rotate_z(V3<T> v, T deg) {
  T r = C<T>::deg2rad(deg), s = std::sin(r), c = std::cos(r);
  return V3<T>(c * v.x - s * v.y, s * v.x + c * v.y, v.z);
}
template <class T> struct Xf { V3<T> center, rot; };
template <class T> V3<T> to_local(V3<T> p, const Xf<T>& t) {
  V3<T> q = p - t.center;
  if (planted(38)) return rotate_z(rotate_y(rotate_x(q, -t.rot.x), -t.rot.y), -t.rot.z);
  q = rotate_z(q, -t.rot.z);
  q = rotate_y(q, -t.rot.y);
  q = rotate_x(q, -t.rot.x);
  return q;
}
// transform_offset (R:2432-2437)
template <class T> V3<T> transform_offset(const Xf<T>& t, V3<T> off) {
  return rotate_z(rotate_y(rotate_x(off, t.rot.x), t.rot.y), t.rot.z);
}

// point_aabb_lower_bound (R:5388-5411)
template <class T> T point_aabb_lb(V3<T> p, const Aabb<T>& b) {
  T dx = p.x < b.min.x ? b.min.x - p.x : (p.x > b.max.x ? p.x - b.max.x : T(0));
  T dy = p.y < b.min.y ? b.min.y - p.y : (p.y > b.max.y ? p.y - b.max.y : T(0));
  T dz = p.z < b.min.z ? b.min.z - p.z : (p.z > b.max.z ? p.z - b.max.z : T(0));
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// ------------------------------------------------------------ hg_sdf ops (R:6019-6137)
template <class T> T op_smax(T a, T b, T k) {
  T h = rmax(k - std::fabs(a - b), T(0));
  return rmax(a, b) + h * h * T(0.25) / rmax(k, T(1.0e-6));
}
template <class T> T op_smin(T a, T b, T k) {
  T h = rmax(k - std::fabs(a - b), T(0));
  return rmin(a, b) - h * h * T(0.25) / rmax(k, T(1.0e-6));
}
template <class T> T op_union_round(T a, T b, T r) {
  r = rmax(std::fabs(r), T(1.0e-6));
  T ux = rmax(r - a, T(0)), uy = rmax(r - (planted(19) ? a : b), T(0));
  return rmax(r, rmin(a, b)) - std::sqrt(ux * ux + uy * uy);
}
template <class T> T op_intersect_round(T a, T b, T r) {
  r = rmax(std::fabs(r), T(1.0e-6));
  T ux = rmax(r + a, T(0)), uy = rmax(r + b, T(0));
  return rmin(planted(20) ? r : -r, rmax(a, b)) + std::sqrt(ux * ux + uy * uy);
}
template <class T> T op_diff_round(T a, T b, T r) {
  return planted(21) ? op_intersect_round(-a, b, r) : op_intersect_round(a, -b, r);
}
template <class T> T frac_1_sqrt_2() { return T(0.70710678118654752440); }
template <class T> T op_union_chamfer(T a, T b, T r) {
  return rmin(rmin(a, b), (planted(22) ? a + r + b : a - r + b) * frac_1_sqrt_2<T>());
}
template <class T> T op_intersect_chamfer(T a, T b, T r) {
  return rmax(rmax(a, b), (a + r + b) * (planted(23) ? T(0.5) : frac_1_sqrt_2<T>()));
}
template <class T> T op_diff_chamfer(T a, T b, T r) { return op_intersect_chamfer(a, -b, r); }
template <class T> T repeat_centered(T p, T size) {
  size = rmax(size, T(1.0e-6));
  return rem_euclid(p + size * T(0.5), size) - size * T(0.5);
}
template <class T> T op_union_columns(T a, T b, T r, T n) {
  if (a < r && b < r) {
    n = rmax(std::round(n), T(2));
    T s2 = std::sqrt(T(2));
    T column_radius = r * s2 / ((planted(31) ? n : n - T(1)) * T(2) + s2);
    T x = a + column_radius, y = b + column_radius;
    int ni = (int)n;
    if (ni % 2 == 1) x += column_radius;
    T rx = (x + y) * frac_1_sqrt_2<T>(), ry = (y - x) * frac_1_sqrt_2<T>();
    x = rx;
    y = ry;
    x -= s2 * T(0.5) * r;
    x += -column_radius * s2;
    if (ni % 2 == 1) y += column_radius;
    y = repeat_centered(y, column_radius * T(2));
    return std::sqrt(x * x + y * y) - column_radius;
  }
  return rmin(a, b);
}
template <class T> T op_intersect_columns(T a, T b, T r, T n) { return -op_union_columns(-a, -b, r, n); }
template <class T> T op_diff_columns(T a, T b, T r, T n) { return -op_union_columns(-a, b, r, n); }
template <class T> T op_union_stairs(T a, T b, T r, T n) {
  n = rmax(std::round(n), T(1));
  T s = r / (planted(24) ? n + T(1) : n), u = b - r;
  return rmin(rmin(a, b), T(0.5) * (u + a + std::fabs(rem_euclid(u - a + s, T(2) * s) - s)));
}
template <class T> T op_intersect_stairs(T a, T b, T r, T n) {
  return planted(25) ? -op_union_stairs(-a, b, r, n) : -op_union_stairs(-a, -b, r, n);
}
template <class T> T op_diff_stairs(T a, T b, T r, T n) { return -op_union_stairs(-a, b, r, n); }
template <class T> T op_union_soft(T a, T b, T r) {
  r = rmax(std::fabs(r), T(1.0e-6));
  T e = rmax(r - std::fabs(a - b), T(0));
  return rmin(a, b) - e * e * T(0.25) / (planted(26) ? T(2) * r : r);
}
template <class T> T op_pipe(T a, T b, T r) {
  return (planted(27) ? a * a + b * b : std::sqrt(a * a + b * b)) - std::fabs(r);
}
template <class T> T op_engrave(T a, T b, T r) {
  return rmax(a, (a + std::fabs(r) - (planted(28) ? b : std::fabs(b))) * frac_1_sqrt_2<T>());
}
template <class T> T op_groove(T a, T b, T ra, T rb) {
  return rmax(a, rmin(planted(29) ? a - std::fabs(ra) : a + std::fabs(ra), std::fabs(rb) - std::fabs(b)));
}
template <class T> T op_tongue(T a, T b, T ra, T rb) {
  return rmin(a, rmax(a - std::fabs(ra), (planted(30) ? b : std::fabs(b)) - std::fabs(rb)));
}

// noise modifier (R:5957-6009)
template <class T> T fractf(T x) { return x - std::floor(x); }
template <class T> T frac_1_pi() { return T(0.31830988618379067154); }
template <class T> T hash_noise(V3<T> p) {
  V3<T> q = V3<T>(fractf(p.x * frac_1_pi<T>() + T(0.11)), fractf(p.y * frac_1_pi<T>() + T(0.17)),
                  fractf(p.z * frac_1_pi<T>() + T(0.13))) * T(17);
  return fractf(q.x * q.y * q.z * (q.x + q.y + q.z));
}
template <class T> T sd_base_noise(V3<T> p) {
  p = p + V3<T>(T(0.5), T(0.5), T(0.5));
  V3<T> i(std::floor(p.x), std::floor(p.y), std::floor(p.z));
  V3<T> f(fractf(p.x), fractf(p.y), fractf(p.z));
  T d = std::numeric_limits<T>::infinity();
  for (int ix = 0; ix <= 1; ++ix)
    for (int iy = 0; iy <= 1; ++iy)
      for (int iz = 0; iz <= 1; ++iz) {
        V3<T> c((T)ix, (T)iy, (T)iz);
        T r = hash_noise(i + c);
        r = r * r * T(0.7);
        d = rmin(d, (f - c).length() - r);
      }
  return d;
}
template <class T> T apply_noise_modifier(V3<T> p, T d, uint32_t octaves, T scale, T lac) {
  const T m[3][3] = {{T(0.00), T(1.60), T(1.20)}, {T(-1.60), T(0.72), T(-0.96)}, {T(-1.20), T(-0.96), T(1.28)}};
  V3<T> q = p * rmax(scale, T(1.0e-4));
  T os = rmax(lac, T(1.0e-4));
  T s = T(1), t = T(0);
  uint32_t n = octaves < 1 ? 1 : octaves;
  for (uint32_t k = 0; k < n; ++k) {
    T nn = s * sd_base_noise(q);
    nn = op_smax(nn, d - (planted(44) ? T(0.2) : T(0.1)) * s, T(0.3) * s);
    d = op_smin(nn, d, T(0.3) * s);
    t += d;
    q = V3<T>(m[0][0] * q.x + m[0][1] * q.y + m[0][2] * q.z, m[1][0] * q.x + m[1][1] * q.y + m[1][2] * q.z,
              m[2][0] * q.x + m[2][1] * q.y + m[2][2] * q.z) * os;
    s *= planted(45) ? T(0.5) : T(0.415);
    q.z += T(-4.33) * t * s;
  }
  return d;
}

// sd_regular_ngon (R:6206-6214)
template <class T> T sd_regular_ngon(V3<T> p, uint32_t sides, T radius) {
  T n = (T)(sides < 3 ? 3 : sides);
  T an = T(3.14159265358979323846) / n;
  T apothem = radius * (planted(3) ? std::sin(an) : std::cos(an));
  T angle = std::atan2(p.z, p.x);
  T sector = T(2) * an;
  T wrapped = rem_euclid(angle + an, sector) - an;
  return std::hypot(p.x, p.z) * std::cos(wrapped) - apothem;
}

// ----------------------------------------------------------- expression DAG
// Builtin semantics: eval.rs:2223-2690 (interpreter), jit.rs:1621-1914 (JIT).
template <class T> T smoothstep01(T t) {
  t = rclamp(t, T(0), T(1));
  return t * t * (T(3) - T(2) * t);
}
template <class T> T hash_noise3(V3<T> p) {  // eval.rs:4286-4291
  T qx = fractf(p.x * frac_1_pi<T>() + T(0.11)) * T(17);
  T qy = fractf(p.y * frac_1_pi<T>() + T(0.17)) * T(17);
  T qz = fractf(p.z * frac_1_pi<T>() + T(0.13)) * T(17);
  return fractf(qx * qy * qz * (qx + qy + qz));
}
template <class T> T lerp64(T a, T b, T t) { return a * (T(1) - t) + b * t; }
template <class T> T value_noise_3d(V3<T> p) {  // eval.rs:4292-4317
  V3<T> i(std::floor(p.x), std::floor(p.y), std::floor(p.z));
  V3<T> f(fractf(p.x), fractf(p.y), fractf(p.z));
  V3<T> u = planted(34) ? f : V3<T>(smoothstep01(f.x), smoothstep01(f.y), smoothstep01(f.z));
  T n000 = hash_noise3(V3<T>(i.x, i.y, i.z)), n001 = hash_noise3(V3<T>(i.x, i.y, i.z + T(1)));
  T n010 = hash_noise3(V3<T>(i.x, i.y + T(1), i.z)), n011 = hash_noise3(V3<T>(i.x, i.y + T(1), i.z + T(1)));
  T n100 = hash_noise3(V3<T>(i.x + T(1), i.y, i.z)), n101 = hash_noise3(V3<T>(i.x + T(1), i.y, i.z + T(1)));
  T n110 = hash_noise3(V3<T>(i.x + T(1), i.y + T(1), i.z));
  T n111 = hash_noise3(V3<T>(i.x + T(1), i.y + T(1), i.z + T(1)));
  if (planted(35)) std::swap(n010, n100);
  T nx00 = lerp64(n000, n100, u.x), nx01 = lerp64(n001, n101, u.x);
  T nx10 = lerp64(n010, n110, u.x), nx11 = lerp64(n011, n111, u.x);
  T nxy0 = lerp64(nx00, nx10, u.y), nxy1 = lerp64(nx01, nx11, u.y);
  return lerp64(nxy0, nxy1, u.z) * T(2) - T(1);
}
template <class T> T fbm_3d(V3<T> p, uint32_t octaves, T scale, T lac) {  // eval.rs:4319-4336
  V3<T> q = p * scale;
  T amp = T(0.5), sum = T(0);
  T l = std::fabs(lac) < C<T>::EPS ? T(1) : lac;
  uint32_t n = octaves < 1 ? 1 : octaves;
  for (uint32_t k = 0; k < n; ++k) {
    sum += amp * value_noise_3d(q);
    q = q * l;
    amp *= planted(36) ? T(0.5) : T(0.55);
  }
  return sum;
}

template <class T> struct FnEval {
  const ft_scene_ir* ir;
  // Evaluates function `fn` with splatted params; returns the result (scalars splat).
  V3<T> operator()(int32_t fn, const V3<T>* params) const {
    const ft_function& F = ir->functions[fn];
    thread_local std::vector<V3<T>> vals;
    if (vals.size() < F.count) vals.resize(F.count);
    const ft_expr* E = ir->exprs + F.first;
    for (uint32_t k = 0; k < F.count; ++k) {
      const ft_expr& e = E[k];
      V3<T> a = e.a >= 0 ? vals[e.a] : V3<T>(), b = e.b >= 0 ? vals[e.b] : V3<T>(),
            c = e.c >= 0 ? vals[e.c] : V3<T>(), d = e.d >= 0 ? vals[e.d] : V3<T>();
      V3<T> r;
      auto map1 = [&](T (*f)(T)) { return V3<T>(f(a.x), f(a.y), f(a.z)); };
      switch (e.op) {
        case FT_EXPR_CONST: r = V3<T>::splat((T)e.value); break;
        case FT_EXPR_PARAM: r = params[e.a]; break;
        case FT_EXPR_VEC3: r = V3<T>(a.x, b.x, c.x); break;
        case FT_EXPR_GET_X: r = V3<T>::splat(a.x); break;
        case FT_EXPR_GET_Y: r = V3<T>::splat(a.y); break;
        case FT_EXPR_GET_Z: r = V3<T>::splat(a.z); break;
        case FT_EXPR_NEG: r = V3<T>(-a.x, -a.y, -a.z); break;
        case FT_EXPR_ADD: r = a + b; break;
        case FT_EXPR_SUB: r = a - b; break;
        case FT_EXPR_MUL: r = a.cmul(b); break;
        case FT_EXPR_DIV: r = V3<T>(a.x / b.x, a.y / b.y, a.z / b.z); break;
        case FT_EXPR_ABS: r = map1([](T x) { return std::fabs(x); }); break;
        case FT_EXPR_FLOOR: r = map1([](T x) { return std::floor(x); }); break;
        case FT_EXPR_CEIL: r = map1([](T x) { return std::ceil(x); }); break;
        case FT_EXPR_FRACT: r = map1([](T x) { return x - std::floor(x); }); break;
        case FT_EXPR_SQRT: r = map1([](T x) { return std::sqrt(x); }); break;
        case FT_EXPR_SIN: r = map1([](T x) { return std::sin(x); }); break;
        case FT_EXPR_COS: r = map1([](T x) { return std::cos(x); }); break;
        case FT_EXPR_SATURATE: r = map1([](T x) { return rclamp(x, T(0), T(1)); }); break;
        case FT_EXPR_MIN: r = V3<T>(rmin(a.x, b.x), rmin(a.y, b.y), rmin(a.z, b.z)); break;
        case FT_EXPR_MAX: r = V3<T>(rmax(a.x, b.x), rmax(a.y, b.y), rmax(a.z, b.z)); break;
        case FT_EXPR_POW: r = V3<T>(std::pow(a.x, b.x), std::pow(a.y, b.y), std::pow(a.z, b.z)); break;
        case FT_EXPR_CLAMP: {
          auto cl = [](T x, T lo, T hi) { return lo <= hi ? rclamp(x, lo, hi) : rclamp(x, hi, lo); };
          r = V3<T>(cl(a.x, b.x, c.x), cl(a.y, b.y, c.y), cl(a.z, b.z, c.z));
          break;
        }
        case FT_EXPR_MIX: {
          auto mx = [](T x, T y, T t) { return x * (T(1) - t) + y * t; };
          r = V3<T>(mx(a.x, b.x, c.x), mx(a.y, b.y, c.y), mx(a.z, b.z, c.z));
          break;
        }
        case FT_EXPR_STEP: r = V3<T>::splat(b.x < a.x ? T(0) : T(1)); break;
        case FT_EXPR_SMOOTHSTEP: {
          T span = b.x - a.x, t;
          if (std::fabs(span) < C<T>::EPS) t = c.x < a.x ? T(0) : T(1);
          else t = rclamp((c.x - a.x) / span, T(0), T(1));
          r = V3<T>::splat(t * t * (T(3) - T(2) * t));
          break;
        }
        case FT_EXPR_DOT: r = V3<T>::splat(a.x * b.x + a.y * b.y + a.z * b.z); break;
        case FT_EXPR_LENGTH: r = V3<T>::splat(std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z)); break;
        case FT_EXPR_NORMALIZE: {
          T len = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
          r = len <= C<T>::EPS ? V3<T>() : V3<T>(a.x / len, a.y / len, a.z / len);
          break;
        }
        case FT_EXPR_ROTATE_X: r = rotate_x(a, b.x); break;  // eval.rs:4733-4741
        case FT_EXPR_ROTATE_Y: r = rotate_y(a, b.x); break;
        case FT_EXPR_ROTATE_Z: r = rotate_z(a, b.x); break;
        case FT_EXPR_SPHERE_DIST: r = V3<T>::splat(a.length() - b.x); break;
        case FT_EXPR_BOX_DIST: {
          V3<T> q(std::fabs(a.x) - b.x, std::fabs(a.y) - b.y, std::fabs(a.z) - b.z);
          T out = V3<T>(rmax(q.x, T(0)), rmax(q.y, T(0)), rmax(q.z, T(0))).length();
          T in = rmin(rmax(q.x, rmax(q.y, q.z)), T(0));
          r = V3<T>::splat(out + in);
          break;
        }
        case FT_EXPR_CYLINDER_DIST: {
          T radial = std::sqrt(a.x * a.x + a.z * a.z) - b.x;
          T axial = std::fabs(a.y) - c.x;
          T ox = rmax(radial, T(0)), oy = rmax(axial, T(0));
          r = V3<T>::splat(std::sqrt(ox * ox + oy * oy) + rmin(rmax(radial, axial), T(0)));
          break;
        }
        case FT_EXPR_TORUS_DIST: {
          T qx = std::sqrt(a.x * a.x + a.z * a.z) - b.x;
          r = V3<T>::splat(std::sqrt(qx * qx + a.y * a.y) - c.x);
          break;
        }
        case FT_EXPR_VALUE_NOISE: r = V3<T>::splat(value_noise_3d(a)); break;
        case FT_EXPR_FBM: {
          uint32_t oct = b.x > T(0) ? (uint32_t)b.x : 0u;
          r = V3<T>::splat(fbm_3d(a, oct, c.x, d.x));
          break;
        }
        default: r = V3<T>(); break;
      }
      vals[k] = r;
    }
    return vals[F.count - 1];
  }
};

// ------------------------------------------------------------------- scene
template <class T> struct DistInfo {
  T distance;
  uint32_t object_id, material_id;
};

enum MatTag { TAG_STANDARD = 0, TAG_LAMBERT = 1, TAG_METAL = 2, TAG_DIELECTRIC = 3 };

template <class T> struct Params {  // MaterialParams (materials/mod.rs:67-87)
  Spec<T> color;
  T roughness, metallic, specular, specular_weight;
  Spec<T> specular_color;
  T ior, clearcoat, clearcoat_roughness, transmission;
  bool thin_walled;
  Spec<T> emission_color;
  T emission_strength;
  bool has_medium;
  T medium_ior;
  Spec<T> medium_absorption;
  T medium_density;
  int pattern;
  Spec<T> pattern_a, pattern_b;
  T pattern_scale;
  int hook = -1;  // hook table entry while the set keeps its dynamic id (blend_params drops it)
};
template <class T> struct Mat {  // Material (materials/mod.rs:214-221)
  bool blend;
  int model;  // when !blend
  Params<T> p;
  int a_model, b_model;
  Params<T> a, b;
  T t;
};

template <class T> struct Light {
  int kind;
  V3<T> position;
  Spec<T> radiance;
  T radius;
  uint32_t samples;
};

template <class T> struct Scene {
  const ft_scene_ir* ir;
  std::vector<Aabb<T>> bounds;
  std::vector<Xf<T>> xf;        // per node transform
  std::vector<Xf<T>> objects;   // object_transforms
  std::vector<Mat<T>> materials;
  std::vector<int> mat_hook;                  // per material: index into hooks, or -1
  std::vector<ft_material_hooks> hooks;
  std::vector<Light<T>> lights;
  FnEval<T> fn;
  uint32_t root;

  const ft_node& node(uint32_t i) const { return ir->nodes[i]; }
  T prm(uint32_t i, int k) const { return (T)ir->nodes[i].param[k]; }
};

template <class T> Aabb<T> cube(V3<T> c, T r) { return {c - V3<T>::splat(r), c + V3<T>::splat(r)}; }

// sdf_bounds (R:5241-5386)
template <class T> Aabb<T> compute_bounds(Scene<T>& S, uint32_t i) {
  const ft_node& n = S.node(i);
  V3<T> c = S.xf[i].center;
  switch (n.kind) {
    case FT_NODE_SPHERE: return cube(c, S.prm(i, 0));
    case FT_NODE_BOX: return cube(c, V3<T>(S.prm(i, 0), S.prm(i, 1), S.prm(i, 2)).length());
    case FT_NODE_CYLINDER: {
      T r = S.prm(i, 0), h = S.prm(i, 1);
      return cube(c, std::sqrt(r * r + h * h));
    }
    case FT_NODE_TORUS: return cube(c, S.prm(i, 0) + S.prm(i, 1));
    case FT_NODE_EXTRUDE_POLYGON: {
      T r = S.prm(i, 1), h = S.prm(i, 2);
      return cube(c, std::sqrt(r * r + h * h));
    }
    case FT_NODE_CUSTOM: {
      V3<T> h(S.prm(i, 0), S.prm(i, 1), S.prm(i, 2));
      T inf = std::numeric_limits<T>::infinity();
      Aabb<T> b{V3<T>::splat(inf), V3<T>::splat(-inf)};
      const T sx[8] = {-1, -1, -1, -1, 1, 1, 1, 1}, sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
              sz[8] = {-1, 1, -1, 1, -1, 1, -1, 1};
      for (int k = 0; k < 8; ++k) {
        V3<T> corner(sx[k] < 0 ? -h.x : h.x, sy[k] < 0 ? -h.y : h.y, sz[k] < 0 ? -h.z : h.z);
        V3<T> w = c + transform_offset(S.xf[i], corner);
        b.min = V3<T>(rmin(b.min.x, w.x), rmin(b.min.y, w.y), rmin(b.min.z, w.z));
        b.max = V3<T>(rmax(b.max.x, w.x), rmax(b.max.y, w.y), rmax(b.max.z, w.z));
      }
      return b;
    }
    case FT_NODE_DOMAIN_MODIFIER:
    case FT_NODE_DISTANCE_POST_MODIFIER:
    case FT_NODE_NOISE:
      return {V3<T>(n.bounds_min[0], n.bounds_min[1], n.bounds_min[2]),
              V3<T>(n.bounds_max[0], n.bounds_max[1], n.bounds_max[2])};
    case FT_NODE_SUBTRACT: return S.bounds[n.lhs];
    case FT_NODE_SLICE: {
      Aabb<T> b = S.bounds[n.lhs];
      int axis = (int)n.param[0];
      T mn = S.prm(i, 1), mx = S.prm(i, 2);
      T* lo = axis == 0 ? &b.min.x : (axis == 1 ? &b.min.y : &b.min.z);
      T* hi = axis == 0 ? &b.max.x : (axis == 1 ? &b.max.y : &b.max.z);
      *lo = rmax(*lo, mn);
      *hi = rmin(*hi, mx);
      if (*hi < *lo) *hi = *lo;
      return b;
    }
    case FT_NODE_SMOOTH: {
      Aabb<T> b = S.bounds[n.lhs];
      T d = S.prm(i, 0) * T(0.1);
      return {b.min - V3<T>::splat(d), b.max + V3<T>::splat(d)};
    }
    default: return aabb_union(S.bounds[n.lhs], S.bounds[n.rhs]);
  }
}

// sdf_center (R:4047-4089)
template <class T> V3<T> sdf_center(const Scene<T>& S, uint32_t i) {
  const ft_node& n = S.node(i);
  if (n.kind <= FT_NODE_NOISE) return S.xf[i].center;
  switch (n.kind) {
    case FT_NODE_SUBTRACT:
    case FT_NODE_DIFF_ROUND:
    case FT_NODE_DIFF_CHAMFER:
    case FT_NODE_DIFF_COLUMNS:
    case FT_NODE_DIFF_STAIRS:
    case FT_NODE_PIPE:
    case FT_NODE_ENGRAVE:
    case FT_NODE_GROOVE:
    case FT_NODE_TONGUE:
    case FT_NODE_SLICE:
    case FT_NODE_SMOOTH: return sdf_center(S, n.lhs);
    default: {
      V3<T> l = sdf_center(S, n.lhs), r = sdf_center(S, n.rhs);
      return V3<T>((l.x + r.x) * T(0.5), (l.y + r.y) * T(0.5), (l.z + r.z) * T(0.5));
    }
  }
}

template <class T> T shell_op(T d, T shell) { return shell > T(0) ? rmax(d, -(d + shell)) : d; }

// primitive distances shared by sdf_distance_info / sdf_lower_bound (R:5413-5690)
template <class T> T prim_distance(const Scene<T>& S, uint32_t i, V3<T> p) {
  const ft_node& n = S.node(i);
  switch (n.kind) {
    case FT_NODE_SPHERE: {
      T d = to_local(p, S.xf[i]).length() - S.prm(i, 0);
      return shell_op(d, S.prm(i, 1));
    }
    case FT_NODE_BOX: {
      T rd = S.prm(i, 3);
      V3<T> inner(rmax(S.prm(i, 0) - rd, T(0)), rmax(S.prm(i, 1) - rd, T(0)), rmax(S.prm(i, 2) - rd, T(0)));
      V3<T> q = to_local(p, S.xf[i]).abs() - inner;
      T outside = V3<T>(rmax(q.x, T(0)), rmax(q.y, T(0)), rmax(q.z, T(0))).length();
      T inside = rmin(rmax(rmax(q.x, q.y), q.z), T(0));
      return shell_op(outside + inside - rd, S.prm(i, 4));
    }
    case FT_NODE_CYLINDER: {
      V3<T> q = to_local(p, S.xf[i]);
      T rd = S.prm(i, 2);
      if (planted(2)) std::swap(q.y, q.z);
      T radial = std::sqrt(q.x * q.x + q.z * q.z);
      T dx = radial - rmax(S.prm(i, 0) - rd, T(0));
      T dy = std::fabs(q.y) - rmax(S.prm(i, 1) - rd, T(0));
      T ox = rmax(dx, T(0)), oy = rmax(dy, T(0));
      T outside = std::sqrt(ox * ox + oy * oy);
      T inside = rmin(rmax(dx, dy), T(0));
      return shell_op(outside + inside - rd, S.prm(i, 3));
    }
    case FT_NODE_TORUS: {
      V3<T> q = to_local(p, S.xf[i]);
      const T major = planted(1) ? S.prm(i, 1) : S.prm(i, 0), minor = planted(1) ? S.prm(i, 0) : S.prm(i, 1);
      T qx = std::sqrt(q.x * q.x + q.z * q.z) - major;
      return std::sqrt(qx * qx + q.y * q.y) - minor;
    }
    case FT_NODE_EXTRUDE_POLYGON: {
      V3<T> q = to_local(p, S.xf[i]);
      T rd = S.prm(i, 3);
      T radial = sd_regular_ngon(V3<T>(q.x, T(0), q.z), (uint32_t)n.param[0], rmax(S.prm(i, 1) - rd, T(0)));
      T dy = std::fabs(q.y) - rmax(S.prm(i, 2) - rd, T(0));
      T ox = rmax(radial, T(0)), oy = rmax(dy, T(0));
      T outside = std::sqrt(ox * ox + oy * oy);
      T inside = rmin(rmax(radial, dy), T(0));
      return shell_op(outside + inside - rd, S.prm(i, 4));
    }
    default: return T(0);
  }
}

// eval_custom_sdf_distance (R:6139-6179)
template <class T> T custom_distance(const Scene<T>& S, const ft_node& n, V3<T> p) {
  V3<T> dp = p;
  if (n.fn_domain >= 0) dp = S.fn(n.fn_domain, &p);
  T d = n.fn_distance >= 0 ? S.fn(n.fn_distance, &dp).x : T(1.0e6);
  if (n.fn_post >= 0) {
    V3<T> args[2] = {V3<T>::splat(d), dp};
    d = S.fn(n.fn_post, args).x;
  }
  return d;
}

// sdf_lower_bound (R:5413-5543)
template <class T> T sdf_lower_bound(const Scene<T>& S, uint32_t i, V3<T> p) {
  const ft_node& n = S.node(i);
  switch (n.kind) {
    case FT_NODE_SPHERE:
    case FT_NODE_BOX:
    case FT_NODE_CYLINDER:
    case FT_NODE_TORUS:
    case FT_NODE_EXTRUDE_POLYGON: return prim_distance(S, i, p);
    case FT_NODE_CUSTOM: {
      V3<T> h(S.prm(i, 0), S.prm(i, 1), S.prm(i, 2));
      return point_aabb_lb(to_local(p, S.xf[i]), Aabb<T>{h * T(-1), h});
    }
    case FT_NODE_UNION: return rmin(sdf_lower_bound(S, n.lhs, p), sdf_lower_bound(S, n.rhs, p));
    case FT_NODE_INTERSECT: return rmax(sdf_lower_bound(S, n.lhs, p), sdf_lower_bound(S, n.rhs, p));
    case FT_NODE_SUBTRACT: return sdf_lower_bound(S, n.lhs, p);
    case FT_NODE_SMOOTH: return sdf_lower_bound(S, n.lhs, p) - S.prm(i, 0) * T(0.1);
    default: return point_aabb_lb(p, S.bounds[i]);
  }
}

template <class T> bool is_intersect_family(uint32_t k) {
  return k == FT_NODE_INTERSECT_ROUND || k == FT_NODE_INTERSECT_CHAMFER || k == FT_NODE_INTERSECT_COLUMNS ||
         k == FT_NODE_INTERSECT_STAIRS;
}

// sdf_distance_info (R:5545-5955)
template <class T> DistInfo<T> sdf_distance_info(const Scene<T>& S, uint32_t i, V3<T> p) {
  const ft_node& n = S.node(i);
  switch (n.kind) {
    case FT_NODE_SPHERE:
    case FT_NODE_BOX:
    case FT_NODE_CYLINDER:
    case FT_NODE_TORUS:
    case FT_NODE_EXTRUDE_POLYGON: return {prim_distance(S, i, p), n.object_id, n.material_id};
    case FT_NODE_CUSTOM: return {custom_distance(S, n, to_local(p, S.xf[i])), n.object_id, n.material_id};
    case FT_NODE_DOMAIN_MODIFIER: {
      V3<T> q = to_local(p, S.xf[i]);
      V3<T> dq = n.fn_domain >= 0 ? S.fn(n.fn_domain, &q) : q;
      return sdf_distance_info(S, n.lhs, dq);
    }
    case FT_NODE_DISTANCE_POST_MODIFIER: {
      V3<T> q = to_local(p, S.xf[i]);
      DistInfo<T> info = sdf_distance_info(S, n.lhs, q);
      if (n.fn_post >= 0) {
        V3<T> args[2] = {V3<T>::splat(info.distance), q};
        info.distance = S.fn(n.fn_post, args).x;
      }
      return info;
    }
    case FT_NODE_NOISE: {
      V3<T> q = to_local(p, S.xf[i]);
      DistInfo<T> info = sdf_distance_info(S, n.lhs, q);
      info.distance = apply_noise_modifier(q, info.distance, (uint32_t)n.param[0], S.prm(i, 1), S.prm(i, 2));
      return info;
    }
    case FT_NODE_UNION: {
      T llb = sdf_lower_bound(S, n.lhs, p), rlb = sdf_lower_bound(S, n.rhs, p);
      if (llb <= rlb) {
        DistInfo<T> l = sdf_distance_info(S, n.lhs, p);
        if (l.distance <= rlb) return l;
        DistInfo<T> r = sdf_distance_info(S, n.rhs, p);
        return l.distance <= r.distance ? l : r;
      }
      DistInfo<T> r = sdf_distance_info(S, n.rhs, p);
      if (r.distance <= llb) return r;
      DistInfo<T> l = sdf_distance_info(S, n.lhs, p);
      return l.distance <= r.distance ? l : r;
    }
    case FT_NODE_INTERSECT: {
      DistInfo<T> l = sdf_distance_info(S, n.lhs, p), r = sdf_distance_info(S, n.rhs, p);
      return l.distance >= r.distance ? l : r;
    }
    case FT_NODE_SUBTRACT: {
      DistInfo<T> l = sdf_distance_info(S, n.lhs, p), r = sdf_distance_info(S, n.rhs, p);
      T rd = -r.distance;
      if (l.distance >= rd) return l;
      return {rd, r.object_id, r.material_id};
    }
    case FT_NODE_SLICE: {
      DistInfo<T> info = sdf_distance_info(S, n.lhs, p);
      int axis = (int)n.param[0];
      T coord = p[axis];
      T slab = planted(32) ? rmax(coord - S.prm(i, 1), S.prm(i, 2) - coord)
                           : rmax(S.prm(i, 1) - coord, coord - S.prm(i, 2));
      info.distance = rmax(info.distance, slab);
      return info;
    }
    case FT_NODE_SMOOTH: {
      DistInfo<T> info = sdf_distance_info(S, n.lhs, p);
      info.distance -= S.prm(i, 0) * (planted(33) ? T(1) : T(0.1));
      return info;
    }
    default: break;
  }
  DistInfo<T> l = sdf_distance_info(S, n.lhs, p), r = sdf_distance_info(S, n.rhs, p);
  T a = l.distance, b = r.distance, rr = S.prm(i, 0), nn = S.prm(i, 1);
  T d;
  switch (n.kind) {
    case FT_NODE_UNION_ROUND: d = op_union_round(a, b, rr); break;
    case FT_NODE_UNION_CHAMFER: d = op_union_chamfer(a, b, rr); break;
    case FT_NODE_UNION_COLUMNS: d = op_union_columns(a, b, rr, nn); break;
    case FT_NODE_UNION_STAIRS: d = op_union_stairs(a, b, rr, nn); break;
    case FT_NODE_UNION_SOFT: d = op_union_soft(a, b, rr); break;
    case FT_NODE_INTERSECT_ROUND: d = op_intersect_round(a, b, rr); break;
    case FT_NODE_INTERSECT_CHAMFER: d = op_intersect_chamfer(a, b, rr); break;
    case FT_NODE_INTERSECT_COLUMNS: d = op_intersect_columns(a, b, rr, nn); break;
    case FT_NODE_INTERSECT_STAIRS: d = op_intersect_stairs(a, b, rr, nn); break;
    case FT_NODE_DIFF_ROUND: d = op_diff_round(a, b, rr); break;
    case FT_NODE_DIFF_CHAMFER: d = op_diff_chamfer(a, b, rr); break;
    case FT_NODE_DIFF_COLUMNS: d = op_diff_columns(a, b, rr, nn); break;
    case FT_NODE_DIFF_STAIRS: d = op_diff_stairs(a, b, rr, nn); break;
    case FT_NODE_PIPE: return {op_pipe(a, b, rr), l.object_id, l.material_id};
    case FT_NODE_ENGRAVE: return {op_engrave(a, b, rr), l.object_id, l.material_id};
    case FT_NODE_GROOVE: return {op_groove(a, b, rr, nn), l.object_id, l.material_id};
    case FT_NODE_TONGUE: return {op_tongue(a, b, rr, nn), l.object_id, l.material_id};
    default: d = a; break;
  }
  if (n.kind >= FT_NODE_DIFF_ROUND && n.kind <= FT_NODE_DIFF_STAIRS) {
    if (a >= -b) return {d, l.object_id, l.material_id};
    return {d, r.object_id, r.material_id};
  }
  if (is_intersect_family<T>(n.kind)) return a >= b ? DistInfo<T>{d, l.object_id, l.material_id}
                                                     : DistInfo<T>{d, r.object_id, r.material_id};
  return a <= b ? DistInfo<T>{d, l.object_id, l.material_id} : DistInfo<T>{d, r.object_id, r.material_id};
}

// ------------------------------------------------------------ accelerators
template <class T> struct Leaf {
  Aabb<T> bounds;
  uint32_t node;
};

// accel_leaf_distance_info (R:629-657)
template <class T> DistInfo<T> leaf_distance_info(const Scene<T>& S, const Leaf<T>& leaf, V3<T> p) {
  const ft_node& n = S.node(leaf.node);
  if (n.kind == FT_NODE_CUSTOM) {
    T lb = point_aabb_lb(p, leaf.bounds);
    V3<T> ext = leaf.bounds.max - leaf.bounds.min;
    T far_t = rclamp(ext.length() * T(0.1), T(0.05), T(0.5));
    if (lb > far_t) return {lb, n.object_id, n.material_id};
  }
  return sdf_distance_info(S, leaf.node, p);
}

// collect_accel_leaves (R:659-670)
template <class T> void collect_leaves(const Scene<T>& S, uint32_t i, std::vector<Leaf<T>>& out) {
  const ft_node& n = S.node(i);
  if (n.kind == FT_NODE_UNION) {
    collect_leaves(S, n.lhs, out);
    collect_leaves(S, n.rhs, out);
  } else {
    out.push_back({S.bounds[i], i});
  }
}

template <class T> struct BvhNode {
  Aabb<T> bounds;
  int leaf;         // >= 0 for leaves
  int lhs, rhs;     // children for inner nodes
};

template <class T> struct Accel {
  const Scene<T>* S;
  uint32_t mode;
  Aabb<T> scene_bounds;
  std::vector<Leaf<T>> leaves;
  std::vector<BvhNode<T>> bvh;
  int bvh_root = -1;
  // BrickGrid (R:824-954)
  bool has_grid = false;
  Aabb<T> gbounds;
  size_t dims = 0;
  std::vector<std::vector<uint32_t>> cells;
};

// build_bvh (R:693-732)
template <class T> int build_bvh(Accel<T>& A, std::vector<uint32_t> idx) {
  if (idx.empty()) return -1;
  if (idx.size() == 1) {
    A.bvh.push_back({A.leaves[idx[0]].bounds, (int)idx[0], -1, -1});
    return (int)A.bvh.size() - 1;
  }
  Aabb<T> b = A.leaves[idx[0]].bounds;
  for (size_t k = 1; k < idx.size(); ++k) b = aabb_union(b, A.leaves[idx[k]].bounds);
  V3<T> e = b.max - b.min;
  int axis = (e.x >= e.y && e.x >= e.z) ? 0 : (e.y >= e.z ? 1 : 2);
  auto cen = [&](uint32_t l) {
    const Aabb<T>& lb = A.leaves[l].bounds;
    V3<T> c = (lb.min + lb.max) * T(0.5);
    return c[axis];
  };
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t bb) { return cen(a) < cen(bb); });
  size_t mid = idx.size() / 2;
  int l = build_bvh(A, std::vector<uint32_t>(idx.begin(), idx.begin() + mid));
  int r = build_bvh(A, std::vector<uint32_t>(idx.begin() + mid, idx.end()));
  A.bvh.push_back({b, -1, l, r});
  return (int)A.bvh.size() - 1;
}

// bvh_distance_info (R:743-786); returns false for None
template <class T> bool bvh_distance_info(const Accel<T>& A, int ni, V3<T> p, T& best, DistInfo<T>& out) {
  const BvhNode<T>& nd = A.bvh[ni];
  if (point_aabb_lb(p, nd.bounds) > best) return false;
  if (nd.leaf >= 0) {
    DistInfo<T> info = leaf_distance_info(*A.S, A.leaves[nd.leaf], p);
    if (info.distance < best) best = info.distance;
    out = info;
    return true;
  }
  T llb = point_aabb_lb(p, A.bvh[nd.lhs].bounds), rlb = point_aabb_lb(p, A.bvh[nd.rhs].bounds);
  int first = nd.lhs, second = nd.rhs;
  T flb = llb, slb = rlb;
  if (planted(42) ? !(llb < rlb) : !(llb <= rlb)) { first = nd.rhs; second = nd.lhs; flb = rlb; slb = llb; }
  bool have = false;
  DistInfo<T> bi{};
  if (flb <= best) have = bvh_distance_info(A, first, p, best, bi);
  DistInfo<T> si{};
  if (slb <= best && bvh_distance_info(A, second, p, best, si)) {
    if (!(have && bi.distance <= si.distance)) bi = si;
    have = true;
  }
  if (have) out = bi;
  return have;
}

// bvh_lower_bound (R:788-822)
template <class T> bool bvh_lower_bound(const Accel<T>& A, int ni, V3<T> p, T& best, T& out) {
  const BvhNode<T>& nd = A.bvh[ni];
  if (point_aabb_lb(p, nd.bounds) > best) return false;
  if (nd.leaf >= 0) {
    T lb = point_aabb_lb(p, nd.bounds);
    if (lb < best) best = lb;
    out = lb;
    return true;
  }
  T llb = point_aabb_lb(p, A.bvh[nd.lhs].bounds), rlb = point_aabb_lb(p, A.bvh[nd.rhs].bounds);
  int first = llb <= rlb ? nd.lhs : nd.rhs, second = llb <= rlb ? nd.rhs : nd.lhs;
  bool have = false;
  T bl = T(0);
  if (point_aabb_lb(p, A.bvh[first].bounds) <= best) have = bvh_lower_bound(A, first, p, best, bl);
  T sl;
  if (point_aabb_lb(p, A.bvh[second].bounds) <= best && bvh_lower_bound(A, second, p, best, sl)) {
    bl = have ? rmin(bl, sl) : sl;
    have = true;
  }
  if (have) out = bl;
  return have;
}

// brick grid helpers (R:956-1022)
template <class T> void cell_coords(const Accel<T>& A, V3<T> p, size_t out[3]) {
  V3<T> ext = A.gbounds.max - A.gbounds.min;
  for (int k = 0; k < 3; ++k) {
    T e = ext[k];
    if (e <= T(1.0e-6)) { out[k] = 0; continue; }
    T v = rclamp((p[k] - A.gbounds.min[k]) / e, T(0), T(0.999999)) * (T)A.dims;
    out[k] = (size_t)v;
  }
}
template <class T> size_t cell_index(const Accel<T>& A, size_t x, size_t y, size_t z) {
  return x + A.dims * (y + A.dims * z);
}
template <class T> Aabb<T> cell_aabb(const Accel<T>& A, size_t cx, size_t cy, size_t cz) {
  V3<T> ext = A.gbounds.max - A.gbounds.min;
  T dm = (T)(A.dims ? A.dims : 1);
  V3<T> size(ext.x / dm, ext.y / dm, ext.z / dm);
  V3<T> mn(A.gbounds.min.x + size.x * (T)cx, A.gbounds.min.y + size.y * (T)cy, A.gbounds.min.z + size.z * (T)cz);
  return {mn, mn + size};
}
template <class T> T shell_lb(const Accel<T>& A, const size_t o[3], size_t shell, V3<T> p) {
  if (shell == 0) return T(0);
  size_t d1 = A.dims - 1;
  size_t mn[3], mx[3];
  for (int k = 0; k < 3; ++k) {
    mn[k] = o[k] >= shell ? o[k] - shell : 0;
    mx[k] = std::min(o[k] + shell, d1);
  }
  Aabb<T> a = cell_aabb(A, mn[0], mn[1], mn[2]), b = cell_aabb(A, mx[0], mx[1], mx[2]);
  return point_aabb_lb(p, Aabb<T>{a.min, b.max});
}

template <class T, class F> void grid_shell_walk(const Accel<T>& A, V3<T> p, F&& visit_leaf, T& best) {
  size_t o[3];
  cell_coords(A, p, o);
  size_t max_shell = A.dims;
  std::unordered_set<uint32_t> visited;
  size_t d1 = A.dims - 1;
  for (size_t shell = 0; shell < max_shell; ++shell) {
    if (shell_lb(A, o, shell, p) > best) break;
    size_t lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = o[k] >= shell ? o[k] - shell : 0;
      hi[k] = std::min(o[k] + shell, d1);
    }
    for (size_t z = lo[2]; z <= hi[2]; ++z)
      for (size_t y = lo[1]; y <= hi[1]; ++y)
        for (size_t x = lo[0]; x <= hi[0]; ++x) {
          if (shell > 0 && x > lo[0] && x < hi[0] && y > lo[1] && y < hi[1] && z > lo[2] && z < hi[2]) continue;
          for (uint32_t li : A.cells[cell_index(A, x, y, z)]) {
            if (!visited.insert(li).second) continue;
            visit_leaf(li);
          }
        }
  }
}

template <class T> DistInfo<T> accel_distance_info(const Accel<T>& A, V3<T> p) {
  const Scene<T>& S = *A.S;
  if (A.mode == FT_ACCEL_BVH) {
    T best = std::numeric_limits<T>::infinity();
    DistInfo<T> out;
    if (A.bvh_root >= 0 && bvh_distance_info(A, A.bvh_root, p, best, out)) return out;
    return sdf_distance_info(S, S.root, p);
  }
  if (A.mode == FT_ACCEL_BRICKS && A.has_grid) {
    T best = std::numeric_limits<T>::infinity();
    bool have = false;
    DistInfo<T> bi{};
    grid_shell_walk(A, p, [&](uint32_t li) {
      const Leaf<T>& leaf = A.leaves[li];
      if (point_aabb_lb(p, leaf.bounds) > best) return;
      DistInfo<T> info = leaf_distance_info(S, leaf, p);
      if (info.distance < best) { best = info.distance; bi = info; have = true; }
    }, best);
    if (have) return bi;
    return sdf_distance_info(S, A.leaves[0].node, p);
  }
  return sdf_distance_info(S, S.root, p);
}

template <class T> T accel_lower_bound(const Accel<T>& A, V3<T> p) {
  const Scene<T>& S = *A.S;
  if (A.mode == FT_ACCEL_BVH) {
    T best = std::numeric_limits<T>::infinity(), out;
    if (A.bvh_root >= 0 && bvh_lower_bound(A, A.bvh_root, p, best, out)) return out;
    return sdf_lower_bound(S, S.root, p);
  }
  if (A.mode == FT_ACCEL_BRICKS && A.has_grid) {
    T best = std::numeric_limits<T>::infinity();
    grid_shell_walk(A, p, [&](uint32_t li) {
      T lb = point_aabb_lb(p, A.leaves[li].bounds);
      if (lb < best) best = lb;
    }, best);
    return best;
  }
  return sdf_lower_bound(S, S.root, p);
}

template <class T> T accel_distance(const Accel<T>& A, V3<T> p) { return accel_distance_info(A, p).distance; }

// --------------------------------------------------------- scene loading
template <class T> Params<T> params_from_ir(const ft_material& m) {
  Params<T> p;
  p.color = Spec<T>(m.color[0], m.color[1], m.color[2]);
  p.roughness = m.roughness; p.metallic = m.metallic; p.specular = m.specular;
  p.specular_weight = m.specular_weight;
  p.specular_color = Spec<T>(m.specular_color[0], m.specular_color[1], m.specular_color[2]);
  p.ior = m.ior; p.clearcoat = m.clearcoat; p.clearcoat_roughness = m.clearcoat_roughness;
  p.transmission = m.transmission; p.thin_walled = m.thin_walled != 0;
  p.emission_color = Spec<T>(m.emission_color[0], m.emission_color[1], m.emission_color[2]);
  p.emission_strength = m.emission_strength;
  p.has_medium = m.has_medium != 0; p.medium_ior = m.medium_ior;
  p.medium_absorption = Spec<T>(m.medium_absorption[0], m.medium_absorption[1], m.medium_absorption[2]);
  p.medium_density = m.medium_density;
  p.pattern = (int)m.pattern;
  p.pattern_a = Spec<T>(m.pattern_color_a[0], m.pattern_color_a[1], m.pattern_color_a[2]);
  p.pattern_b = Spec<T>(m.pattern_color_b[0], m.pattern_color_b[1], m.pattern_color_b[2]);
  p.pattern_scale = m.pattern_scale;
  return p;
}

bool validate(const ft_scene_ir* ir) {
  if (!ir) { g_err = "null scene"; return false; }
  if (ir->abi_version != FT_ABI_VERSION) { g_err = "abi version mismatch"; return false; }
  if (ir->n_nodes == 0 || ir->root >= ir->n_nodes) { g_err = "no suitable scene root"; return false; }
  if (ir->n_material_hooks > 0 && !ir->material_hooks) { g_err = "material hook table is null"; return false; }
  for (uint32_t h = 0; h < ir->n_material_hooks; ++h) {
    const ft_material_hooks& H = ir->material_hooks[h];
    if (H.material >= ir->n_materials) { g_err = "material hook names no material"; return false; }
    for (int32_t f : {H.color, H.roughness, H.ior, H.thin_walled, H.emission_color, H.emission_strength,
                      H.metallic, H.specular, H.specular_weight, H.transmission, H.specular_color,
                      H.clearcoat_roughness, H.clearcoat, H.coat_mask, H.bump, H.normal,
                      H.medium_ior, H.medium_absorption, H.medium_density, H.eval})
      if (f >= (int32_t)ir->n_functions) { g_err = "material hook names no function"; return false; }
  }
  for (uint32_t i = 0; i < ir->n_nodes; ++i) {
    const ft_node& n = ir->nodes[i];
    if (n.kind >= FT_NODE_KIND_COUNT) { g_err = "bad node kind"; return false; }
    bool binary = n.kind >= FT_NODE_UNION && n.kind <= FT_NODE_TONGUE;
    bool unary = n.kind == FT_NODE_DOMAIN_MODIFIER || n.kind == FT_NODE_DISTANCE_POST_MODIFIER ||
                 n.kind == FT_NODE_NOISE || n.kind == FT_NODE_SLICE || n.kind == FT_NODE_SMOOTH;
    if ((binary || unary) && (n.lhs < 0 || (uint32_t)n.lhs >= i)) { g_err = "bad child index"; return false; }
    if (binary && (n.rhs < 0 || (uint32_t)n.rhs >= i)) { g_err = "bad child index"; return false; }
    for (int32_t f : {n.fn_domain, n.fn_distance, n.fn_post})
      if (f >= (int32_t)ir->n_functions) { g_err = "bad function index"; return false; }
  }
  for (uint32_t f = 0; f < ir->n_functions; ++f) {
    const ft_function& F = ir->functions[f];
    if (F.count == 0 || F.first + F.count > ir->n_exprs) { g_err = "bad function range"; return false; }
    for (uint32_t k = 0; k < F.count; ++k) {
      const ft_expr& e = ir->exprs[F.first + k];
      for (int32_t o : {e.a, e.b, e.c, e.d})
        if (e.op != FT_EXPR_PARAM && o >= (int32_t)k) { g_err = "bad expression operand"; return false; }
      if (e.op == FT_EXPR_PARAM && (e.a < 0 || (uint32_t)e.a >= 16)) { g_err = "bad parameter"; return false; }
    }
  }
  if (ir->env_fn >= (int32_t)ir->n_functions) { g_err = "bad environment function"; return false; }
  return true;
}

template <class T> void load_scene(const ft_scene_ir* ir, Scene<T>& S) {
  S.ir = ir;
  S.fn.ir = ir;
  S.root = ir->root;
  S.xf.resize(ir->n_nodes);
  S.bounds.resize(ir->n_nodes);
  for (uint32_t i = 0; i < ir->n_nodes; ++i) {
    const ft_node& n = ir->nodes[i];
    S.xf[i].center = V3<T>(n.center[0], n.center[1], n.center[2]);
    S.xf[i].rot = V3<T>(n.rot_deg[0], n.rot_deg[1], n.rot_deg[2]);
  }
  for (uint32_t i = 0; i < ir->n_nodes; ++i) S.bounds[i] = compute_bounds(S, i);
  for (uint32_t o = 0; o < ir->n_objects; ++o) {
    const float* t = ir->object_transforms + 6 * o;
    S.objects.push_back({V3<T>(t[0], t[1], t[2]), V3<T>(t[3], t[4], t[5])});
  }
  for (uint32_t m = 0; m < ir->n_materials; ++m) {
    Mat<T> mm{};
    mm.blend = false;
    mm.model = (int)ir->materials[m].model;
    mm.p = params_from_ir<T>(ir->materials[m]);
    S.materials.push_back(mm);
    S.mat_hook.push_back(-1);
  }
  for (uint32_t h = 0; h < ir->n_material_hooks; ++h) {
    const ft_material_hooks& H = ir->material_hooks[h];
    S.mat_hook[H.material] = (int)S.hooks.size();
    S.hooks.push_back(H);
  }
  for (uint32_t l = 0; l < ir->n_lights; ++l) {
    const ft_light& L = ir->lights[l];
    S.lights.push_back({(int)L.kind, V3<T>(L.position[0], L.position[1], L.position[2]),
                        Spec<T>(L.radiance[0], L.radiance[1], L.radiance[2]), (T)L.radius, L.samples});
  }
}

template <class T> void build_accel(const Scene<T>& S, uint32_t mode, Accel<T>& A) {
  A.S = &S;
  A.mode = mode;
  A.scene_bounds = S.bounds[S.root];
  if (mode == FT_ACCEL_BVH || mode == FT_ACCEL_BRICKS) {
    collect_leaves(S, S.root, A.leaves);
  }
  if (mode == FT_ACCEL_BVH) {
    std::vector<uint32_t> idx(A.leaves.size());
    for (size_t k = 0; k < idx.size(); ++k) idx[k] = (uint32_t)k;
    A.bvh.reserve(2 * idx.size());
    A.bvh_root = build_bvh(A, idx);
  }
  if (mode == FT_ACCEL_BRICKS && !A.leaves.empty()) {
    A.has_grid = true;
    A.gbounds = A.scene_bounds;
    float lc = (float)std::max<size_t>(A.leaves.size(), 1);
    size_t dim = (size_t)std::max(std::ceil(std::cbrt(lc)), 1.0f);
    A.dims = dim;
    A.cells.assign(dim * dim * dim, {});
    for (uint32_t li = 0; li < A.leaves.size(); ++li) {
      size_t mn[3], mx[3];
      cell_coords(A, A.leaves[li].bounds.min, mn);
      cell_coords(A, A.leaves[li].bounds.max, mx);
      for (size_t z = mn[2]; z <= mx[2]; ++z)
        for (size_t y = mn[1]; y <= mx[1]; ++y)
          for (size_t x = mn[0]; x <= mx[0]; ++x) A.cells[cell_index(A, x, y, z)].push_back(li);
    }
    for (auto& c : A.cells) {
      if (planted(43)) std::sort(c.begin(), c.end(), [](uint32_t a, uint32_t b) { return a > b; });
      else std::sort(c.begin(), c.end());
      c.erase(std::unique(c.begin(), c.end()), c.end());
    }
  }
}

// ------------------------------------------------------------------- march
template <class T> struct Opts {
  uint32_t width, height, max_steps;
  T max_dist, epsilon, step_scale, camera_z, fov;
};
template <class T> Opts<T> opts_from(const ft_render_options* o) {
  return {o->width, o->height, o->max_steps, (T)o->max_dist, (T)o->epsilon, (T)o->step_scale,
          (T)o->camera_z, (T)o->fov_y_degrees};
}

template <class T> struct Hit {
  T t;
  V3<T> position, normal;
  bool front_face;
  uint32_t object_id, material_id;
  V3<T> view_dir;  // normalize(-dir) of the ray (the shading context of material hooks)
};

// ray_aabb_intersection (R:4141-4173)
template <class T> bool ray_aabb(V3<T> o, V3<T> d, const Aabb<T>& b, T& t0, T& t1) {
  T tmin = -std::numeric_limits<T>::infinity(), tmax = std::numeric_limits<T>::infinity();
  for (int axis = 0; axis < 3; ++axis) {
    T oo = o[axis], dd = d[axis], mn = b.min[axis], mx = b.max[axis];
    if (std::fabs(dd) <= T(1.0e-8)) {
      if (oo < mn || oo > mx) return false;
      continue;
    }
    T inv = T(1) / dd;
    T a = (mn - oo) * inv, c = (mx - oo) * inv;
    if (a > c) std::swap(a, c);
    tmin = rmax(tmin, a);
    tmax = rmin(tmax, c);
    if (tmax < tmin) return false;
  }
  t0 = tmin;
  t1 = tmax;
  return true;
}

thread_local uint64_t tl_evals = 0;
thread_local bool tl_count = false;

template <class T> DistInfo<T> ev_info(const Accel<T>& A, V3<T> p) {
  if (tl_count) ++tl_evals;
  return accel_distance_info(A, p);
}

// estimate_normal (R:4202-4215)
template <class T> V3<T> estimate_normal(const Accel<T>& A, V3<T> p, T e) {
  T dx = ev_info(A, V3<T>(p.x + e, p.y, p.z)).distance - ev_info(A, V3<T>(p.x - e, p.y, p.z)).distance;
  T dy = ev_info(A, V3<T>(p.x, p.y + e, p.z)).distance - ev_info(A, V3<T>(p.x, p.y - e, p.z)).distance;
  T dz = ev_info(A, V3<T>(p.x, p.y, p.z + e)).distance - ev_info(A, V3<T>(p.x, p.y, p.z - e)).distance;
  V3<T> n(dx, dy, dz);
  if (n.length() > T(1.0e-8)) return n.normalize();
  return V3<T>(0, 1, 0);
}

// raymarch_hit (R:4091-4139) + refine_hit_distance (R:4180-4200)
template <class T> bool raymarch_hit(const Accel<T>& A, V3<T> o, V3<T> d, const Opts<T>& op, T min_t, T max_t,
                                     Hit<T>& hit) {
  T entry, exit;
  if (!ray_aabb(o, d, A.scene_bounds, entry, exit)) return false;
  max_t = rmin(max_t, exit);
  T traveled = rmax(min_t, rmax(entry, T(0)));
  T prev = traveled;
  for (uint32_t s = 0; s < op.max_steps; ++s) {
    if (traveled > max_t) return false;
    V3<T> p = o + d * traveled;
    T dist = ev_info(A, p).distance;
    if (std::fabs(dist) < op.epsilon) {
      T lo = prev, hi = traveled;
      for (int k = 0; k < 10; ++k) {
        T mid = T(0.5) * (lo + hi);
        T dm = std::fabs(ev_info(A, o + d * mid).distance);
        if (dm < op.epsilon) hi = mid; else lo = mid;
      }
      V3<T> pos = o + d * hi;
      V3<T> nrm = estimate_normal(A, pos, rmax(op.epsilon * T(4), T(1.0e-5)));
      DistInfo<T> fin = ev_info(A, pos);
      hit = {hi, pos, nrm, nrm.dot(d) < T(0), fin.object_id, fin.material_id, V3<T>()};
      return true;
    }
    prev = traveled;
    T step = rmax(std::fabs(dist) * rclamp(op.step_scale, T(0.05), T(1)), rmax(op.epsilon * T(0.5), T(1.0e-5)));
    traveled += step;
  }
  return false;
}

// PinholeCamera::generate_ray (render_api.rs:132-144)
template <class T> struct Cam {
  V3<T> origin, target, up;
  T fov;
  void gen(T ndc_x, T ndc_y, V3<T>& o, V3<T>& d) const {
    V3<T> forward = (target - origin).normalize();
    V3<T> right = forward.cross(up).normalize();
    V3<T> upv = right.cross(forward).normalize();
    T tan_fov = std::tan(T(0.5) * C<T>::deg2rad(fov));
    d = (forward + right * (ndc_x * tan_fov) + upv * (ndc_y * tan_fov)).normalize();
    o = origin;
  }
};

// parse_camera (R:6263-6286)
template <class T> Cam<T> make_camera(const Scene<T>& S, const Opts<T>& op) {
  V3<T> center = sdf_center(S, S.root);
  const ft_camera& c = S.ir->camera;
  Cam<T> cam;
  cam.up = V3<T>(0, 1, 0);
  if (c.present) {
    cam.origin = c.has_origin ? V3<T>(c.origin[0], c.origin[1], c.origin[2]) : center + V3<T>(0, 0, op.camera_z);
    cam.target = c.has_target ? V3<T>(c.target[0], c.target[1], c.target[2]) : center;
    cam.fov = c.has_fov ? (T)c.fov_y_degrees : op.fov;
  } else {
    cam.origin = center + V3<T>(0, 0, op.camera_z);
    cam.target = center;
    cam.fov = op.fov;
  }
  return cam;
}

// pixel_sample_offsets (R:4838-4854)
template <class T> std::vector<std::pair<T, T>> sample_offsets(uint32_t n) {
  n = n < 1 ? 1 : n;
  uint32_t grid = (uint32_t)std::ceil(std::sqrt((float)n));
  std::vector<std::pair<T, T>> out;
  for (uint32_t iy = 0; iy < grid; ++iy)
    for (uint32_t ix = 0; ix < grid; ++ix) {
      if (out.size() >= n) break;
      out.push_back({((T)ix + T(0.5)) / (T)grid, ((T)iy + T(0.5)) / (T)grid});
    }
  return out;
}

// spectrum_to_rgb8 / _reinhard (R:5221-5239)
template <class T> uint8_t to_u8(T v) {
  T mapped = rmin(std::pow(rmax(v, T(0)), T(1) / T(2.2)), T(1));
  T s = mapped * T(255);
  if (!(s > T(0))) return 0;
  if (s >= T(255)) return 255;
  return (uint8_t)s;
}
template <class T> uint8_t to_u8_reinhard(T v) {
  T x = rmax(v, T(0));
  return to_u8(x / (T(1) + x));
}

// depth_preview_value (R:5215-5219)
template <class T> T depth_preview_value(T t, T max_dist) {
  T d = rclamp(t / rmax(max_dist, T(1.0e-6)), T(0), T(1));
  T s = T(1) - d;
  return s * s;
}

// ----------------------------------------------------------- Whitted
template <class T> Spec<T> lerp_spectrum(Spec<T> a, Spec<T> b, T t) {  // R:5206-5213
  T tt = rclamp(t, T(0), T(1));
  return Spec<T>(a.x * (T(1) - tt) + b.x * tt, a.y * (T(1) - tt) + b.y * tt, a.z * (T(1) - tt) + b.z * tt);
}
template <class T> T smoothstepf(T e0, T e1, T x) {  // R:7538-7541
  T t = rclamp((x - e0) / rmax(e1 - e0, T(1.0e-6)), T(0), T(1));
  return t * t * (T(3) - T(2) * t);
}

// materials/*.rs eval
template <class T> T ggx_d(T ndoth, T alpha) {
  T a2 = planted(13) ? alpha : alpha * alpha, n2 = ndoth * ndoth;
  T denom = rmax(n2 * (a2 - T(1)) + T(1), T(1.0e-5));
  return a2 / (T(3.14159265358979323846) * denom * denom);
}
template <class T> T smith_g1(T ndotx, T alpha) {
  T k = (planted(14) ? alpha : alpha * alpha) * T(0.5);
  return ndotx / rmax(ndotx * (T(1) - k) + k, T(1.0e-5));
}
template <class T> Spec<T> metal_eval(const Params<T>& m, V3<T> n, V3<T> wi, V3<T> wo) {  // metal.rs:7-21
  T ndotl = rmax(n.dot(wi), T(0)), ndotv = rmax(n.dot(wo), T(0));
  if (ndotl <= T(0) || ndotv <= T(0)) return Spec<T>();
  V3<T> h = (wi + wo).normalize();
  T ndoth = rmax(n.dot(h), T(0)), vdoth = rmax(wo.dot(h), T(0));
  T rr = rclamp(m.roughness, T(0.02), T(1));
  T alpha = rr * rr;
  T d = ggx_d(ndoth, alpha);
  T g = smith_g1(ndotl, alpha) * smith_g1(ndotv, alpha);
  T mm = planted(15) ? std::pow(T(1) - rclamp(vdoth, T(0), T(1)), T(4)) : powi5(T(1) - rclamp(vdoth, T(0), T(1)));
  Spec<T> fres = m.color + Spec<T>(1, 1, 1) * mm + m.color * (-mm);
  return fres * ((d * g) / rmax(T(4) * ndotl * ndotv, T(1.0e-5)));
}
template <class T> Spec<T> lambert_eval(const Params<T>& m, V3<T> n, V3<T> wi) {  // lambert.rs:7-13
  if (n.dot(wi) <= T(0)) return Spec<T>();
  return m.color * (T(1) / (T(3.14159265358979323846) * (planted(12) ? T(2) : T(1))));
}
template <class T> T fresnel_f0_from_ior(T ior) {
  T i = rmax(ior, T(1.0e-3));
  T r = (i - T(1)) / (i + T(1));
  return r * r;
}
template <class T> Params<T> metal_params(Spec<T> color, T roughness) {
  Params<T> p{};
  p.color = color; p.roughness = roughness; p.metallic = 1; p.specular = 1; p.specular_weight = 1;
  p.specular_color = Spec<T>(1, 1, 1); p.ior = T(1.5); p.clearcoat = 0; p.clearcoat_roughness = T(0.1);
  return p;
}
template <class T> Spec<T> standard_eval(const Params<T>& m, V3<T> n, V3<T> wi, V3<T> wo) {  // standard.rs:7-45
  T ndotl = rmax(n.dot(wi), T(0)), ndotv = rmax(n.dot(wo), T(0));
  if (ndotl <= T(0) || ndotv <= T(0)) return Spec<T>();
  T metallic = rclamp(m.metallic, T(0), T(1)), transmission = rclamp(m.transmission, T(0), T(1));
  T clearcoat = rclamp(m.clearcoat, T(0), T(1));
  T mw = rclamp(T(1) - rclamp(ndotv, T(0), T(1)), T(0), T(1));
  T m2 = mw * mw;
  T schlick = m2 * m2 * mw;
  T coat_cover = rclamp(clearcoat * (T(0.2) + T(0.8) * schlick) * (planted(18) ? T(0.5) : T(0.65)), T(0), T(0.95));
  T base_kd = rclamp((T(1) - metallic) * (planted(16) ? T(1) : T(1) - transmission), T(0), T(1)) * (T(1) - coat_cover);
  Spec<T> diffuse = m.color * (base_kd / T(3.14159265358979323846));
  T f0d = rclamp(fresnel_f0_from_ior(m.ior), T(0), T(1));
  Spec<T> base_f0 = Spec<T>(f0d, f0d, f0d) * (rclamp(m.specular, T(0), T(1)) * rclamp(m.specular_weight, T(0), T(1)));
  base_f0 = base_f0.cmul(m.specular_color);
  Spec<T> spec_f0 = planted(17) ? base_f0 * metallic + m.color * (T(1) - metallic)
                                 : base_f0 * (T(1) - metallic) + m.color * metallic;  // standard.rs lerp_spectrum
  Spec<T> specular = metal_eval(metal_params(spec_f0, m.roughness), n, wi, wo) * (T(1) - coat_cover * T(0.85));
  Spec<T> cc;
  if (clearcoat > T(1.0e-5))
    cc = metal_eval(metal_params(Spec<T>(1, 1, 1), rclamp(m.clearcoat_roughness, T(0.02), T(1))), n, wi, wo) *
         (clearcoat * T(0.18));
  return diffuse + specular + cc;
}
template <class T> Spec<T> model_eval(int model, const Params<T>& p, V3<T> n, V3<T> wi, V3<T> wo) {
  switch (model) {
    case TAG_STANDARD: return standard_eval(p, n, wi, wo);
    case TAG_LAMBERT: return lambert_eval(p, n, wi);
    case TAG_METAL: return metal_eval(p, n, wi, wo);
    default: return Spec<T>();  // dielectric.rs:5-7
  }
}
template <class T> Spec<T> mat_eval(const Mat<T>& m, V3<T> n, V3<T> wi, V3<T> wo) {  // materials/mod.rs:291-305
  if (!m.blend) return model_eval(m.model, m.p, n, wi, wo);
  return lerp_spectrum(model_eval(m.a_model, m.a, n, wi, wo), model_eval(m.b_model, m.b, n, wi, wo), m.t);
}
template <class T> Spec<T> emission_of(const Params<T>& p) { return p.emission_color * rmax(p.emission_strength, T(0)); }
template <class T> Spec<T> mat_emission(const Mat<T>& m) {
  if (!m.blend) return emission_of(m.p);
  return lerp_spectrum(emission_of(m.a), emission_of(m.b), m.t);
}
template <class T> int dominant_model(const Mat<T>& m) {  // R:7653-7667
  if (!m.blend) return m.model;
  return m.t < T(0.5) ? m.a_model : m.b_model;
}
template <class T> const Params<T>& dominant_params(const Mat<T>& m) {  // R:7669-7683
  if (!m.blend) return m.p;
  return m.t < T(0.5) ? m.a : m.b;
}

// blend_params / blend_materials (R:7602-7651)
template <class T> Params<T> blend_params(Params<T> a, const Params<T>& b, T t) {
  a.color = lerp_spectrum(a.color, b.color, t);
  a.roughness = a.roughness + (b.roughness - a.roughness) * t;
  a.metallic = a.metallic + (b.metallic - a.metallic) * t;
  a.specular = a.specular + (b.specular - a.specular) * t;
  a.specular_weight = a.specular_weight + (b.specular_weight - a.specular_weight) * t;
  a.specular_color = lerp_spectrum(a.specular_color, b.specular_color, t);
  a.ior = a.ior + (b.ior - a.ior) * t;
  a.clearcoat = a.clearcoat + (b.clearcoat - a.clearcoat) * t;
  a.clearcoat_roughness = a.clearcoat_roughness + (b.clearcoat_roughness - a.clearcoat_roughness) * t;
  a.transmission = a.transmission + (b.transmission - a.transmission) * t;
  a.thin_walled = t < T(0.5) ? a.thin_walled : b.thin_walled;
  a.hook = -1;  // a.dynamic_material_id = None (R:7624)
  a.emission_color = lerp_spectrum(a.emission_color, b.emission_color, t);
  a.emission_strength = a.emission_strength + (b.emission_strength - a.emission_strength) * t;
  // blend_medium (R:7543-7566)
  if (a.has_medium && b.has_medium) {
    a.medium_ior = a.medium_ior + (b.medium_ior - a.medium_ior) * t;
    a.medium_absorption = lerp_spectrum(a.medium_absorption, b.medium_absorption, t);
    a.medium_density = a.medium_density + (b.medium_density - a.medium_density) * t;
  } else if (a.has_medium) {
    a.has_medium = t < T(0.5);
  } else if (b.has_medium) {
    if (t >= T(0.5)) {
      a.has_medium = true;
      a.medium_ior = b.medium_ior;
      a.medium_absorption = b.medium_absorption;
      a.medium_density = b.medium_density;
    }
  }
  a.pattern = FT_PATTERN_NONE;
  return a;
}
template <class T> Mat<T> blend_materials(const Mat<T>& a, const Mat<T>& b, T t) {
  Mat<T> out{};
  if (!a.blend && !b.blend && a.model == b.model) {
    out.blend = false;
    out.model = a.model;
    out.p = blend_params(a.p, b.p, t);
    return out;
  }
  out.blend = true;
  out.a_model = dominant_model(a);
  out.a = dominant_params(a);
  out.b_model = dominant_model(b);
  out.b = dominant_params(b);
  out.t = rclamp(t, T(0), T(1));
  return out;
}

template <class T> struct Ctx {
  const Scene<T>* S;
  const Accel<T>* A;
  Opts<T> op;
  uint32_t max_depth;
  bool has_env_fn;
  Spec<T> env_radiance;
};

template <class T> Mat<T> material_for_id(const Scene<T>& S, uint32_t id) {
  if (id < S.materials.size()) return S.materials[id];
  return S.materials[0];
}

// resolve_pattern_color (R:7855-7873)
template <class T> Spec<T> pattern_color(const Params<T>& p, V3<T> lp) {
  if (p.pattern != FT_PATTERN_CHECKER3D) return p.color;
  int sx = (int)std::floor(lp.x * p.pattern_scale), sy = (int)std::floor(lp.y * p.pattern_scale),
      sz = (int)std::floor(lp.z * p.pattern_scale);
  return ((sx + sy + sz) & 1) == 0 ? p.pattern_a : p.pattern_b;
}

// resolve_leaf_material (R:7164-7536): the static material, or for a definition with per-hit
// hooks the arm of its model: each hook is a function of the shading context
// (make_shading_context R:7875-7911) replacing the static value before the clamps.  The
// Standard / Layered block (R:7259-7470) also reads metallic, specular, specular_weight and
// transmission, then its coat hooks (listed under the parameter each replaces); its
// clearcoat is clamp(coat_weight * coat_mask), the weight static where it has no hook.
template <class T> Mat<T> resolve_leaf_material(const Ctx<T>& c, Mat<T> m, const Hit<T>& hit) {
  if (m.blend) return m;
  const Scene<T>& S = *c.S;
  Xf<T> xf = hit.object_id < S.objects.size() ? S.objects[hit.object_id] : Xf<T>{};
  V3<T> lp = to_local(hit.position, xf);
  const uint32_t mi = hit.material_id < S.materials.size() ? hit.material_id : 0u;
  const int hk = S.mat_hook.empty() ? -1 : S.mat_hook[mi];
  m.p.hook = hk;
  if (hk >= 0) {
    const ft_material_hooks& H = S.hooks[hk];
    V3<T> args[FT_HOOK_PARAMS];
    args[0] = hit.position;
    args[1] = lp;
    args[2] = hit.front_face ? hit.normal.normalize() : (hit.normal * T(-1)).normalize();
    args[3] = hit.view_dir;
    args[4] = V3<T>::splat(hit.front_face ? T(1) : T(0));
    args[5] = V3<T>::splat((T)hit.object_id);
    args[6] = V3<T>::splat((T)hit.material_id);
    for (uint32_t k = 7; k < FT_HOOK_PARAMS; ++k) args[k] = V3<T>();
    if (H.color >= 0) {
      V3<T> v = S.fn(H.color, args);
      m.p.color = Spec<T>(v.x, v.y, v.z);
    } else {
      m.p.color = pattern_color(m.p, lp);
    }
    if (H.roughness >= 0) m.p.roughness = S.fn(H.roughness, args).x;
    if (H.ior >= 0) m.p.ior = S.fn(H.ior, args).x;
    if (H.thin_walled >= 0) m.p.thin_walled = S.fn(H.thin_walled, args).x >= T(0.5);
    if (H.emission_color >= 0) {
      V3<T> v = S.fn(H.emission_color, args);
      m.p.emission_color = Spec<T>(v.x, v.y, v.z);
    }
    if (H.emission_strength >= 0) m.p.emission_strength = S.fn(H.emission_strength, args).x;
    if (H.metallic >= 0) m.p.metallic = rclamp(S.fn(H.metallic, args).x, T(0), T(1));
    if (H.specular >= 0) m.p.specular = rclamp(S.fn(H.specular, args).x, T(0), T(1));
    if (H.specular_weight >= 0) m.p.specular_weight = rclamp(S.fn(H.specular_weight, args).x, T(0), T(1));
    if (H.transmission >= 0) m.p.transmission = rclamp(S.fn(H.transmission, args).x, T(0), T(1));
    if (H.specular_color >= 0) {
      V3<T> v = S.fn(H.specular_color, args);
      m.p.specular_color = Spec<T>(v.x, v.y, v.z);
    }
    if (H.clearcoat_roughness >= 0) m.p.clearcoat_roughness = rclamp(S.fn(H.clearcoat_roughness, args).x, T(0), T(1));
    if (H.clearcoat >= 0) m.p.clearcoat = rclamp(S.fn(H.clearcoat, args).x, T(0), T(1));
    if (H.coat_mask >= 0) m.p.clearcoat = rclamp(m.p.clearcoat * rclamp(S.fn(H.coat_mask, args).x, T(0), T(1)), T(0), T(1));
    if (H.medium_ior >= 0) {  // medium_from_value (R:6834-6845) of the hook's Medium object
      const V3<T> ab = S.fn(H.medium_absorption, args);
      m.p.has_medium = true;
      m.p.medium_ior = rclamp(S.fn(H.medium_ior, args).x, T(1), T(3));
      m.p.medium_absorption = Spec<T>(ab.x, ab.y, ab.z);
      m.p.medium_density = rmax(S.fn(H.medium_density, args).x, T(0));
    }
  } else {
    m.p.color = pattern_color(m.p, lp);
  }
  if (m.model != TAG_LAMBERT) m.p.roughness = rclamp(m.p.roughness, T(0), T(1));
  if (m.model == TAG_STANDARD || m.model == TAG_DIELECTRIC) m.p.ior = rclamp(m.p.ior, T(1), T(3));
  m.p.emission_strength = rmax(m.p.emission_strength, T(0));
  return m;
}

// resolve_dynamic_bump / resolve_dynamic_normal / orthonormal_tangent (R:7749-7827): the
// material of the hit's id reshapes the shading normal `gn` of build_bsdf_context
// (R:4397-4449) through its bump hook (finite differences at 0.01 along a tangent frame)
// and then its normal hook, each result flipped to the side of the normal it started from.
// The context is make_shading_context_with_normal (R:7884-7911): normal = the normal being
// reshaped, view_dir = -dir as build_bsdf_context receives it.
template <class T> V3<T> apply_normal_hooks(const Scene<T>& S, const Hit<T>& hit, V3<T> lp, V3<T> view, V3<T> gn) {
  const uint32_t mi = hit.material_id < S.materials.size() ? hit.material_id : 0u;
  const int hk = S.mat_hook.empty() ? -1 : S.mat_hook[mi];
  if (hk < 0) return gn;
  const ft_material_hooks& H = S.hooks[hk];
  if (H.bump < 0 && H.normal < 0) return gn;
  V3<T> args[FT_HOOK_PARAMS];
  args[0] = hit.position;
  args[1] = lp;
  args[2] = gn;
  args[3] = view;
  args[4] = V3<T>::splat(hit.front_face ? T(1) : T(0));
  args[5] = V3<T>::splat((T)hit.object_id);
  args[6] = V3<T>::splat((T)hit.material_id);
  for (uint32_t k = 7; k < FT_HOOK_PARAMS; ++k) args[k] = V3<T>();
  V3<T> n = gn;
  if (H.bump >= 0) {
    const T center = S.fn(H.bump, args).x;
    const V3<T> up = std::fabs(gn.y) < T(0.999) ? V3<T>(0, 1, 0) : V3<T>(1, 0, 0);
    const V3<T> tangent = gn.cross(up).normalize();
    const V3<T> bitangent = gn.cross(tangent).normalize();
    const T eps = T(0.01);
    args[1] = lp + tangent * eps;
    const T dx = S.fn(H.bump, args).x - center;
    args[1] = lp + bitangent * eps;
    const T dy = S.fn(H.bump, args).x - center;
    args[1] = lp;
    V3<T> nb = (gn - tangent * (dx / eps) - bitangent * (dy / eps)).normalize();
    if (nb.length() > T(1.0e-6)) n = nb.dot(gn) < T(0) ? nb * T(-1) : nb;
  }
  if (H.normal >= 0) {
    args[2] = n;
    const V3<T> nn = S.fn(H.normal, args).normalize();
    if (nn.length() > T(1.0e-6)) n = nn.dot(n) < T(0) ? nn * T(-1) : nn;
  }
  return n;
}

// resolve_material_from_node (R:6945-7059)
template <class T> Mat<T> resolve_material_from_node(const Ctx<T>& c, uint32_t i, const Hit<T>& hit);
template <class T> Mat<T> blend_node_materials(const Ctx<T>& c, const ft_node& n, const Hit<T>& hit, T radius, bool diff) {
  const Scene<T>& S = *c.S;
  DistInfo<T> l = sdf_distance_info(S, n.lhs, hit.position), r = sdf_distance_info(S, n.rhs, hit.position);
  Mat<T> left = resolve_material_from_node(c, n.lhs, hit), right = resolve_material_from_node(c, n.rhs, hit);
  T k = rmax(std::fabs(radius), T(1.0e-6));
  T t = diff ? smoothstepf(T(0), T(1), T(0.5) + T(0.5) * (((-r.distance) - l.distance) / k))
             : smoothstepf(T(0), T(1), T(0.5) + T(0.5) * ((l.distance - r.distance) / k));
  return blend_materials(left, right, t);
}
template <class T> T blend_radius(const ft_node& n) {
  if (n.kind == FT_NODE_GROOVE || n.kind == FT_NODE_TONGUE) return rmax((T)n.param[0], (T)n.param[1]);
  return (T)n.param[0];
}
template <class T> int blend_class(uint32_t k) {  // 0 none, 1 smooth union/intersect, 2 difference-mode
  switch (k) {
    case FT_NODE_UNION_ROUND: case FT_NODE_UNION_CHAMFER: case FT_NODE_UNION_SOFT:
    case FT_NODE_INTERSECT_ROUND: case FT_NODE_INTERSECT_CHAMFER: case FT_NODE_UNION_COLUMNS:
    case FT_NODE_UNION_STAIRS: case FT_NODE_INTERSECT_COLUMNS: case FT_NODE_INTERSECT_STAIRS: return 1;
    case FT_NODE_DIFF_ROUND: case FT_NODE_DIFF_CHAMFER: case FT_NODE_PIPE: case FT_NODE_ENGRAVE:
    case FT_NODE_DIFF_COLUMNS: case FT_NODE_DIFF_STAIRS: case FT_NODE_GROOVE: case FT_NODE_TONGUE: return 2;
    default: return 0;
  }
}
template <class T> Mat<T> resolve_material_from_node(const Ctx<T>& c, uint32_t i, const Hit<T>& hit) {
  const Scene<T>& S = *c.S;
  const ft_node& n = S.node(i);
  if (n.kind <= FT_NODE_CUSTOM) {
    Hit<T> lh = hit;
    lh.object_id = n.object_id;
    lh.material_id = n.material_id;
    return resolve_leaf_material(c, material_for_id(S, n.material_id), lh);
  }
  if (n.kind == FT_NODE_DOMAIN_MODIFIER || n.kind == FT_NODE_DISTANCE_POST_MODIFIER || n.kind == FT_NODE_NOISE ||
      n.kind == FT_NODE_SLICE || n.kind == FT_NODE_SMOOTH)
    return resolve_material_from_node(c, n.lhs, hit);
  if (n.kind == FT_NODE_UNION || n.kind == FT_NODE_INTERSECT || n.kind == FT_NODE_SUBTRACT) {
    DistInfo<T> l = sdf_distance_info(S, n.lhs, hit.position), r = sdf_distance_info(S, n.rhs, hit.position);
    bool left = n.kind == FT_NODE_UNION ? l.distance <= r.distance
              : (n.kind == FT_NODE_INTERSECT ? l.distance >= r.distance : l.distance >= -r.distance);
    return resolve_material_from_node(c, left ? n.lhs : n.rhs, hit);
  }
  return blend_node_materials(c, n, hit, blend_radius<T>(n), blend_class<T>(n.kind) == 2);
}

// resolve_split_material_from_node (R:7061-7120); returns false for None
template <class T> bool resolve_split(const Ctx<T>& c, uint32_t i, const Hit<T>& hit, Mat<T>& a, Mat<T>& b, T& t) {
  const Scene<T>& S = *c.S;
  const ft_node& n = S.node(i);
  if (n.kind == FT_NODE_UNION || n.kind == FT_NODE_INTERSECT || n.kind == FT_NODE_SUBTRACT) {
    DistInfo<T> l = sdf_distance_info(S, n.lhs, hit.position), r = sdf_distance_info(S, n.rhs, hit.position);
    bool left = n.kind == FT_NODE_UNION ? l.distance <= r.distance
              : (n.kind == FT_NODE_INTERSECT ? l.distance >= r.distance : l.distance >= -r.distance);
    return resolve_split(c, left ? n.lhs : n.rhs, hit, a, b, t);
  }
  if (n.kind == FT_NODE_SMOOTH || n.kind == FT_NODE_NOISE) return resolve_split(c, n.lhs, hit, a, b, t);
  int cls = blend_class<T>(n.kind);
  if (cls == 0) return false;
  DistInfo<T> l = sdf_distance_info(S, n.lhs, hit.position), r = sdf_distance_info(S, n.rhs, hit.position);
  a = resolve_material_from_node(c, n.lhs, hit);
  b = resolve_material_from_node(c, n.rhs, hit);
  T k = rmax(std::fabs(blend_radius<T>(n)), T(1.0e-6));
  t = cls == 2 ? smoothstepf(T(0), T(1), T(0.5) + T(0.5) * (((-r.distance) - l.distance) / k))
               : smoothstepf(T(0), T(1), T(0.5) + T(0.5) * ((l.distance - r.distance) / k));
  return true;
}

// resolve_surface_normal_from_node (R:4462-4570)
template <class T> V3<T> estimate_node_normal(const Scene<T>& S, uint32_t i, V3<T> p, T eps) {
  T e = rmax(eps, T(1.0e-5));
  V3<T> k1(1, -1, -1), k2(-1, -1, 1), k3(-1, 1, -1), k4(1, 1, 1);
  V3<T> n = k1 * sdf_distance_info(S, i, p + k1 * e).distance + k2 * sdf_distance_info(S, i, p + k2 * e).distance +
            k3 * sdf_distance_info(S, i, p + k3 * e).distance + k4 * sdf_distance_info(S, i, p + k4 * e).distance;
  if (n.length() > T(1.0e-6)) return n.normalize();
  return V3<T>(0, 1, 0);
}
template <class T> V3<T> surface_normal(const Scene<T>& S, uint32_t i, V3<T> p, T eps) {
  const ft_node& n = S.node(i);
  if (n.kind <= FT_NODE_NOISE) return estimate_node_normal(S, i, p, eps);
  if (n.kind == FT_NODE_SLICE || n.kind == FT_NODE_SMOOTH) return surface_normal(S, n.lhs, p, eps);
  if (n.kind == FT_NODE_UNION || n.kind == FT_NODE_INTERSECT || n.kind == FT_NODE_SUBTRACT) {
    DistInfo<T> l = sdf_distance_info(S, n.lhs, p), r = sdf_distance_info(S, n.rhs, p);
    if (n.kind == FT_NODE_UNION) return surface_normal(S, l.distance <= r.distance ? n.lhs : n.rhs, p, eps);
    if (n.kind == FT_NODE_INTERSECT) return surface_normal(S, l.distance >= r.distance ? n.lhs : n.rhs, p, eps);
    if (l.distance >= -r.distance) return surface_normal(S, n.lhs, p, eps);
    return surface_normal(S, n.rhs, p, eps) * T(-1);
  }
  // blend_surface_normals (R:4525-4552)
  bool diff = blend_class<T>(n.kind) == 2;
  DistInfo<T> l = sdf_distance_info(S, n.lhs, p), r = sdf_distance_info(S, n.rhs, p);
  V3<T> left = surface_normal(S, n.lhs, p, eps), right = surface_normal(S, n.rhs, p, eps);
  T k = rmax(std::fabs(blend_radius<T>(n)), T(1.0e-6));
  T t;
  if (diff) {
    right = right * T(-1);
    t = smoothstepf(T(0), T(1), T(0.5) + T(0.5) * (((-r.distance) - l.distance) / k));
  } else {
    t = smoothstepf(T(0), T(1), T(0.5) + T(0.5) * ((l.distance - r.distance) / k));
  }
  V3<T> nn = left * (T(1) - t) + right * t;
  if (nn.length() > T(1.0e-6)) return nn.normalize();
  return t < T(0.5) ? left : right;
}

// offset_ray_origin (R:5067-5078)
template <class T> V3<T> offset_ray_origin(V3<T> pos, V3<T> gn, V3<T> dir, T eps) {
  V3<T> n = gn.length() > T(1.0e-8) ? gn.normalize() : dir.normalize();
  T sign = dir.dot(n) >= T(0) ? T(1) : T(-1);
  V3<T> no = n * (sign * rmax(eps * T(16), T(2.0e-4)));
  T ds = rmax(eps * T(8), T(1.0e-4));
  return pos + no + dir.normalize() * ds;
}

// shadow_visibility (R:4706-4733)
template <class T> T shadow_visibility(const Accel<T>& A, V3<T> o, V3<T> d, T max_t, T eps) {
  T t = eps * T(2);
  T vis = T(1);
  for (int k = 0; k < 80; ++k) {
    if (t >= max_t) return rclamp(vis, T(0), T(1));
    V3<T> p = o + d * t;
    T lower = std::fabs(accel_lower_bound(A, p));
    T h = lower > rmax(eps * T(8), T(0.02)) ? lower : std::fabs(accel_distance(A, p));
    if (h < eps * T(4)) return T(0);
    vis = rmin(vis, rclamp(T(10) * h / t, T(0), T(1)));
    t += rmax(h, eps * T(2));
  }
  return rclamp(vis, T(0), T(1));
}

// hammersley_2d / concentric_disk_sample (render_api.rs:258-289)
template <class T> void hammersley(uint32_t i, uint32_t n, T& u, T& v) {
  u = ((T)i + T(0.5)) / (T)(n < 1 ? 1 : n);
  uint32_t bits = (i >> 16) | (i << 16);
  bits = ((bits & 0x55555555u) << 1) | ((bits & 0xAAAAAAAAu) >> 1);
  bits = ((bits & 0x33333333u) << 2) | ((bits & 0xCCCCCCCCu) >> 2);
  if (!planted(10)) bits = ((bits & 0x0F0F0F0Fu) << 4) | ((bits & 0xF0F0F0F0u) >> 4);
  bits = ((bits & 0x00FF00FFu) << 8) | ((bits & 0xFF00FF00u) >> 8);
  v = (T)(float)bits * T(2.3283064e-10f);
}
template <class T> V3<T> concentric_disk(T u1, T u2) {
  T sx = T(2) * u1 - T(1), sy = T(2) * u2 - T(1);
  if (std::fabs(sx) <= C<T>::EPS && std::fabs(sy) <= C<T>::EPS) return V3<T>();
  T r, th;
  const T pi4 = T(0.78539816339744830962), pi2 = T(1.57079632679489661923);
  if (std::fabs(sx) > std::fabs(sy)) { r = sx; th = (planted(11) ? pi2 : pi4) * (sy / sx); }
  else { r = sy; th = pi2 - pi4 * (sx / sy); }
  return V3<T>(r * std::cos(th), r * std::sin(th), T(0));
}

template <class T> struct LightSample { V3<T> wi; Spec<T> radiance; T distance; };
template <class T> LightSample<T> sample_light(const Light<T>& L, V3<T> point, uint32_t idx, uint32_t count) {
  if (L.kind == FT_LIGHT_ENV)
    return {planted(7) ? V3<T>(0, 0, 1) : V3<T>(0, 1, 0), L.radiance, std::numeric_limits<T>::infinity()};
  V3<T> target = L.position;
  if (L.kind == FT_LIGHT_SPHERE && L.radius > T(1.0e-5)) {
    V3<T> normal = (point - L.position).normalize();
    V3<T> up = std::fabs(normal.y) < T(0.99) ? V3<T>(0, 1, 0) : V3<T>(1, 0, 0);
    V3<T> tangent = normal.cross(up).normalize();
    V3<T> bitangent = tangent.cross(normal).normalize();
    T u1, u2;
    hammersley(idx, count < 1 ? 1 : count, u1, u2);
    V3<T> disk = concentric_disk(u1, u2);
    const T rad = planted(9) ? L.radius * T(0.5) : L.radius;
    target = L.position + tangent * (disk.x * rad) + bitangent * (disk.y * rad);
  }
  V3<T> to = target - point;
  T dist = rmax(to.length(), T(1.0e-4));
  T att = planted(8) ? T(1) / dist : T(1) / (dist * dist);
  return {to * (T(1) / dist), L.radiance * att, dist};
}

template <class T> struct BsdfCtx {
  Hit<T> hit;
  V3<T> local_position, normal, wo;
  T current_ior;
};

// shade_color (R:4217-4276)
template <class T> Spec<T> shade_color(const Ctx<T>& c, const Mat<T>& mat, const BsdfCtx<T>& b) {
  const Scene<T>& S = *c.S;
  if (S.lights.empty()) return Spec<T>(1, 1, 1);
  V3<T> p = b.hit.position;
  V3<T> n = b.normal.normalize();
  V3<T> gn = b.hit.front_face ? b.hit.normal.normalize() : (b.hit.normal * T(-1)).normalize();
  Spec<T> color;
  for (const Light<T>& L : S.lights) {
    uint32_t ns = L.kind == FT_LIGHT_SPHERE ? (L.samples < 1 ? 1 : L.samples) : 1;
    T inv = T(1) / (T)ns;
    for (uint32_t si = 0; si < ns; ++si) {
      LightSample<T> smp = sample_light(L, p, si, ns);
      V3<T> wi = smp.wi.normalize();
      T ndotl = n.x * wi.x + n.y * wi.y + n.z * wi.z;
      if (ndotl <= T(0)) continue;
      V3<T> so = offset_ray_origin(b.hit.position, gn, wi, c.op.epsilon);
      T max_t = std::isfinite(smp.distance) ? rmax(smp.distance - c.op.epsilon * T(8), T(0)) : c.op.max_dist;
      if (max_t <= T(0)) continue;
      T sh = shadow_visibility(*c.A, so, wi, max_t, c.op.epsilon);
      if (sh <= T(0)) continue;
      // eval_bsdf (R:4905-4920): the material's eval hook, with the context of
      // bsdf_context_value (R:5003-5030), replaces the BSDF while it keeps its dynamic id
      const Params<T>& dp = dominant_params(mat);
      const int ev = dp.hook >= 0 ? S.hooks[dp.hook].eval : -1;
      Spec<T> f;
      if (ev >= 0) {
        V3<T> args[FT_HOOK_PARAMS];
        args[0] = b.hit.position;
        args[1] = b.local_position;
        args[2] = b.normal;
        args[3] = b.wo;
        args[4] = V3<T>::splat(b.hit.front_face ? T(1) : T(0));
        args[5] = V3<T>::splat((T)b.hit.object_id);
        args[6] = V3<T>::splat((T)b.hit.material_id);
        args[7] = b.wo;
        args[8] = wi;
        args[9] = V3<T>::splat(b.current_ior);
        for (uint32_t k = 10; k < FT_HOOK_PARAMS; ++k) args[k] = V3<T>();
        const V3<T> v = S.fn(ev, args);
        f = Spec<T>(v.x, v.y, v.z);
      } else {
        f = mat_eval(mat, b.normal, wi, b.wo);
      }
      color = color + f.cmul(smp.radiance) * (ndotl * sh * inv);
    }
  }
  return color;
}

template <class T> struct Medium {
  T ior;
  Spec<T> absorption;
  T density;
  static Medium air() { return {T(1), Spec<T>(1, 1, 1), T(0)}; }
};

// medium_transmittance / apply_medium_attenuation (R:5109-5125)
template <class T> Spec<T> attenuate(Spec<T> c, const Medium<T>& m, T dist) {
  if (m.density <= T(0) || dist <= T(0)) return c;
  T ex = m.density * rmax(dist, T(0)) * (planted(6) ? T(0.5) : T(1));
  return c.cmul(Spec<T>(std::pow(rclamp(m.absorption.x, T(0), T(1)), ex), std::pow(rclamp(m.absorption.y, T(0), T(1)), ex),
                        std::pow(rclamp(m.absorption.z, T(0), T(1)), ex)));
}

// transition_medium / medium_state_from_material (R:5127-5175)
template <class T> bool is_transmissive(const Mat<T>& m) {
  int model = dominant_model(m);
  return model == TAG_DIELECTRIC || (model == TAG_STANDARD && dominant_params(m).transmission > T(0.01));
}
template <class T> Medium<T> transition_medium(const Mat<T>& m, bool front, const Medium<T>& cur) {
  if (!is_transmissive(m)) return cur;
  if (!front) return Medium<T>::air();
  const Params<T>& p = dominant_params(m);
  if (p.has_medium) return {rclamp(p.medium_ior, T(1), T(3)), p.medium_absorption, rmax(p.medium_density, T(0))};
  return {rclamp(p.ior, T(1), T(3)), Spec<T>(1, 1, 1), T(0)};
}

template <class T> T fresnel_dielectric(T cos_i, T eta_i, T eta_t) {  // R:5186-5204
  T ei = rmax(eta_i, T(1.0e-4)), et = rmax(eta_t, T(1.0e-4));
  if (std::fabs(ei - et) < T(1.0e-6)) return T(0);
  T ci = rclamp(cos_i, T(0), T(1));
  T eta = ei / et;
  T sin2t = eta * eta * (T(1) - ci * ci);
  if (sin2t >= T(1)) return T(1);
  T ct = std::sqrt(T(1) - sin2t);
  T rs = ((ei * ci) - (et * ct)) / rmax((ei * ci) + (et * ct), T(1.0e-6));
  T rp = planted(5) ? rs : ((et * ci) - (ei * ct)) / rmax((et * ci) + (ei * ct), T(1.0e-6));
  return T(0.5) * (rs * rs + rp * rp);
}
template <class T> T fresnel_schlick(T f0, T cos_t) { return f0 + (T(1) - f0) * powi5(T(1) - rclamp(cos_t, T(0), T(1))); }
template <class T> V3<T> reflect(V3<T> v, V3<T> n) { return v - n * (T(2) * v.dot(n)); }
template <class T> bool refract(V3<T> incident, V3<T> normal, T eta, V3<T>& out) {  // R:5080-5095
  V3<T> i = incident.normalize();
  V3<T> n = normal.normalize();
  T cosi = rclamp(i.dot(n), T(-1), T(1));
  if (cosi > T(0)) n = n * T(-1); else cosi = -cosi;
  T sin2t = eta * eta * (T(1) - cosi * cosi);
  if (sin2t > T(1)) return false;
  T cost = std::sqrt(T(1) - sin2t);
  out = (i * eta + n * (eta * cosi - cost)).normalize();
  return true;
}

template <class T> Spec<T> environment(const Ctx<T>& c, V3<T> dir) {
  if (c.has_env_fn) return c.S->fn(c.S->ir->env_fn, &dir);
  return c.env_radiance;
}

template <class T> Spec<T> trace_ray(const Ctx<T>& c, V3<T> o, V3<T> d, const Medium<T>& med, uint32_t depth);

// rough_lobe_directions (ray.rs:383-417)
template <class T> std::vector<V3<T>> lobe_dirs(V3<T> ideal_dir, T rough, uint32_t count) {
  V3<T> ideal = ideal_dir.normalize();
  V3<T> seed = std::fabs(ideal.y) < T(0.99) ? V3<T>(0, 1, 0) : V3<T>(1, 0, 0);
  V3<T> tangent = ideal.cross(seed).normalize();
  V3<T> bitangent = tangent.cross(ideal).normalize();
  T spread = rclamp(rough, T(0), T(0.5)) * (planted(40) ? T(0.25) : T(0.35));
  std::vector<V3<T>> dirs{ideal};
  if (count <= 1 || spread <= T(1.0e-4)) return dirs;
  const T off[6][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {T(0.707), T(0.707)},
                       {T(-0.707), planted(41) ? T(-0.707) : T(0.707)}};
  for (uint32_t k = 0; k + 1 < count && k < 6; ++k)
    dirs.push_back((ideal + tangent * (off[k][0] * spread) + bitangent * (off[k][1] * spread)).normalize());
  return dirs;
}
template <class T> Spec<T> trace_lobes(const Ctx<T>& c, V3<T> hp, V3<T> gn, V3<T> ideal, const Medium<T>& med,
                                       uint32_t depth, T rough, uint32_t count) {
  std::vector<V3<T>> dirs = lobe_dirs(ideal, rough, count);
  T inv = T(1) / (T)dirs.size();
  Spec<T> sum;
  for (const V3<T>& dd : dirs) {
    V3<T> o = offset_ray_origin(hp, gn, dd, c.op.epsilon * (T(8) + rough * T(12)));
    sum = sum + trace_ray(c, o, dd, med, depth + 1) * inv;
  }
  return sum;
}

// trace_hit_with_material (ray.rs:41-222)
template <class T> Spec<T> trace_hit(const Ctx<T>& c, const Hit<T>& hit, V3<T> dir, const Medium<T>& med, uint32_t depth,
                                     const Mat<T>& mat) {
  const Scene<T>& S = *c.S;
  V3<T> hp = hit.position;
  V3<T> normal = hit.front_face ? hit.normal.normalize() : (hit.normal * T(-1)).normalize();
  // build_bsdf_context (R:4397-4449)
  BsdfCtx<T> b;
  b.hit = hit;
  Xf<T> xf = hit.object_id < S.objects.size() ? S.objects[hit.object_id] : Xf<T>{};
  b.local_position = to_local(hit.position, xf);
  V3<T> sn = surface_normal(S, S.root, hit.position, T(1.0e-4));
  sn = hit.front_face ? sn : sn * T(-1);
  b.normal = apply_normal_hooks(S, hit, b.local_position, dir * T(-1), sn.normalize());
  b.wo = (dir * T(-1)).normalize();
  b.current_ior = med.ior;

  const Params<T>& pr = dominant_params(mat);
  T rough = rclamp(pr.roughness, T(0), T(1));
  T transmission, ior, metallic, clearcoat, spec_strength;
  bool thin;
  Spec<T> spec_color, base_color;
  switch (dominant_model(mat)) {
    case TAG_STANDARD:
      transmission = rclamp(pr.transmission, T(0), T(1)); ior = rclamp(pr.ior, T(1), T(3)); thin = pr.thin_walled;
      spec_color = pr.specular_color; base_color = pr.color; metallic = rclamp(pr.metallic, T(0), T(1));
      clearcoat = rclamp(pr.clearcoat, T(0), T(1)); spec_strength = rclamp(pr.specular * pr.specular_weight, T(0), T(1));
      break;
    case TAG_LAMBERT:
      transmission = 0; ior = 1; thin = false; spec_color = Spec<T>(); base_color = pr.color; metallic = 0;
      clearcoat = 0; spec_strength = 0;
      break;
    case TAG_METAL:
      transmission = 0; ior = 1; thin = false; spec_color = pr.color; base_color = pr.color; metallic = 1;
      clearcoat = 0; spec_strength = 1;
      break;
    default:
      transmission = 1; ior = rclamp(pr.ior, T(1), T(3)); thin = pr.thin_walled; spec_color = Spec<T>(1, 1, 1);
      base_color = pr.color; metallic = 0; clearcoat = 0; spec_strength = 1;
      break;
  }
  T eta_i, eta_t, next_ior;
  if (thin) { eta_i = 1; eta_t = 1; next_ior = med.ior; }
  else if (hit.front_face) { eta_i = med.ior; eta_t = ior; next_ior = ior; }
  else { eta_i = med.ior; eta_t = 1; next_ior = 1; }
  const T eta_ratio = planted(4) ? eta_t / eta_i : eta_i / eta_t;
  V3<T> wo = b.wo;
  V3<T> gn = normal;
  T fresnel = transmission > T(1.0e-4) ? fresnel_dielectric(std::fabs(normal.dot(wo)), eta_i, eta_t)
                                       : rclamp(fresnel_schlick(rclamp(fresnel_f0_from_ior(ior), T(0), T(1)),
                                                                std::fabs(normal.dot(wo))), T(0), T(1));
  T reflect_weight;
  if (transmission > T(1.0e-4)) reflect_weight = rclamp(fresnel, T(0), T(1));
  else {
    T base_reflect = metallic > T(0.5) ? T(1) : rclamp(spec_strength, T(0), T(1));
    reflect_weight = rclamp(base_reflect * fresnel + metallic * (planted(39) ? T(0.6) : T(0.8)) + clearcoat * T(0.2),
                            T(0), T(1));
  }
  T local_weight = rclamp((T(1) - transmission) * (T(1) - reflect_weight), T(0), T(1));
  Spec<T> color = shade_color(c, mat, b) * local_weight;
  color = color + mat_emission(mat);

  if (reflect_weight > T(1.0e-4)) {
    Spec<T> reflected;
    if (transmission > T(1.0e-4) && rough > T(0.06)) {
      V3<T> ideal = reflect(dir, normal).normalize();
      uint32_t n = rough < T(0.12) ? 1 : (rough < T(0.28) ? 2 : 3);
      reflected = trace_lobes(c, hp, gn, ideal, med, depth, rough, n);
    } else {
      V3<T> rd = reflect(dir, normal).normalize();
      V3<T> ro = offset_ray_origin(hp, gn, rd, c.op.epsilon * (T(1) + rough * T(6)));
      reflected = trace_ray(c, ro, rd, med, depth + 1);
    }
    Spec<T> tint = lerp_spectrum(spec_color, base_color, rclamp(metallic, T(0), T(1)));
    color = color + reflected.cmul(tint) * reflect_weight;
  }
  if (transmission > T(1.0e-4)) {
    Medium<T> next = thin ? med : transition_medium(mat, hit.front_face, med);
    next.ior = rclamp(next_ior, T(1), T(3));
    Spec<T> refracted;
    V3<T> rdir;
    if (rough > T(0.06)) {
      if (refract(dir, normal, eta_ratio, rdir)) {
        uint32_t n = rough < T(0.1) ? 2 : (rough < T(0.24) ? 3 : 4);
        refracted = trace_lobes(c, hp, gn, rdir.normalize(), next, depth, rough, n);
      }
    } else if (refract(dir, normal, eta_ratio, rdir)) {
      V3<T> ro = offset_ray_origin(hp, gn, rdir, c.op.epsilon * (T(8) + rough * T(12)));
      refracted = trace_ray(c, ro, rdir.normalize(), next, depth + 1);
    }
    if (!(refracted.x == T(0) && refracted.y == T(0) && refracted.z == T(0))) {
      Spec<T> tt(T(1) - (T(1) - rclamp(base_color.x, T(0), T(1))) * transmission,
                 T(1) - (T(1) - rclamp(base_color.y, T(0), T(1))) * transmission,
                 T(1) - (T(1) - rclamp(base_color.z, T(0), T(1))) * transmission);
      color = color + refracted.cmul(tt) * (transmission * (planted(37) ? T(1) : T(1) - fresnel));
    }
  }
  return attenuate(color, med, hit.t);
}

// trace_ray_recursive (ray.rs:3-38)
template <class T> Spec<T> trace_ray(const Ctx<T>& c, V3<T> o, V3<T> d, const Medium<T>& med, uint32_t depth) {
  if (depth >= c.max_depth) return Spec<T>();
  T min_t = depth == 0 ? T(0) : c.op.epsilon * T(2);
  Hit<T> hit;
  if (!raymarch_hit(*c.A, o, d, c.op, min_t, c.op.max_dist, hit))
    return attenuate(environment(c, d), med, c.op.max_dist);
  hit.view_dir = (d * T(-1)).normalize();
  Mat<T> a, b;
  T t;
  if (resolve_split(c, c.S->root, hit, a, b, t)) {
    Spec<T> ca = trace_hit(c, hit, d, med, depth, a);
    Spec<T> cb = trace_hit(c, hit, d, med, depth, b);
    return lerp_spectrum(ca, cb, t);
  }
  Mat<T> m = resolve_material_from_node(c, c.S->root, hit);
  return trace_hit(c, hit, d, med, depth, m);
}

// trace_ray_debug_aov (ray.rs:224-285)
template <class T> Spec<T> trace_aov(const Ctx<T>& c, V3<T> o, V3<T> d, int aov) {
  Hit<T> hit;
  if (!raymarch_hit(*c.A, o, d, c.op, T(0), c.op.max_dist, hit)) return Spec<T>();
  hit.view_dir = (d * T(-1)).normalize();
  Mat<T> mat = resolve_material_from_node(c, c.S->root, hit);
  V3<T> normal = hit.front_face ? hit.normal.normalize() : (hit.normal * T(-1)).normalize();
  int model = dominant_model(mat);
  T ior = (model == TAG_DIELECTRIC || model == TAG_STANDARD) ? rclamp(dominant_params(mat).ior, T(1), T(3)) : T(1);
  switch (aov) {
    case FT_AOV_DEPTH: { T v = rclamp(T(1) - hit.t / c.op.max_dist, T(0), T(1)); return Spec<T>(v, v, v); }
    case FT_AOV_NORMAL: return Spec<T>(normal.x * T(0.5) + T(0.5), normal.y * T(0.5) + T(0.5), normal.z * T(0.5) + T(0.5));
    case FT_AOV_MATERIAL_ID: {
      uint32_t x = hit.material_id * 0x9E3779B9u + 0x7F4A7C15u;
      x ^= x >> 16;
      x *= 0x85EBCA6Bu;
      x ^= x >> 13;
      return Spec<T>((T)(x & 0xFF) / T(255), (T)((x >> 8) & 0xFF) / T(255), (T)((x >> 16) & 0xFF) / T(255));
    }
    case FT_AOV_IOR: { T v = rclamp((ior - T(1)) / T(2), T(0), T(1)); return Spec<T>(v, v, v); }
    case FT_AOV_TRANSMISSION: { T v = (model == TAG_DIELECTRIC || model == TAG_STANDARD) ? T(1) : T(0); return Spec<T>(v, v, v); }
    case FT_AOV_FRESNEL: {
      V3<T> wo = (d * T(-1)).normalize();
      T f = fresnel_dielectric(std::fabs(normal.dot(wo)), T(1), ior);
      return Spec<T>(f, f, f);
    }
    default: { T v = rclamp(hit.t / c.op.max_dist, T(0), T(1)); return Spec<T>(v, v, v); }
  }
}

// ------------------------------------------------------------ tile runtime
// Work distribution mirrors render_preview_tiled / render_ray_with_accel_progressive
// (R:3659-3717, R:3771-3827): worker threads pull rows from a shared queue.
template <class F> void parallel_rows(int32_t row0, int32_t row1, int32_t threads, F&& fn) {
  int n = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  std::atomic<int32_t> next(row0);
  auto worker = [&]() {
    for (;;) {
      int32_t y = next.fetch_add(1);
      if (y >= row1) break;
      fn(y);
    }
  };
  if (n <= 1) { worker(); return; }
  std::vector<std::thread> pool;
  for (int k = 0; k < n; ++k) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
}

template <class T> int render_depth_t(const ft_scene_ir* ir, uint32_t accel, const ft_render_options* o, uint32_t aa,
                                      int32_t threads, int32_t row0, int32_t row1, uint8_t* rgb, float* t_out,
                                      uint64_t* evals) {
  Scene<T> S;
  load_scene(ir, S);
  Accel<T> A;
  build_accel(S, accel, A);
  Opts<T> op = opts_from<T>(o);
  Cam<T> cam = make_camera(S, op);
  T aspect = (T)op.width / (T)op.height;
  auto offs = sample_offsets<T>(aa);
  std::atomic<uint64_t> total(0);
  parallel_rows(row0, row1, threads, [&](int32_t y) {
    tl_count = evals != nullptr;
    tl_evals = 0;
    for (uint32_t x = 0; x < op.width; ++x) {
      T sum = T(0);
      bool first = true;
      for (auto& so : offs) {
        T px = (((T)x + so.first) / (T)op.width) * T(2) - T(1);
        T py = T(1) - (((T)y + so.second) / (T)op.height) * T(2);
        V3<T> ro, rd;
        cam.gen(px * aspect, py, ro, rd);
        rd = rd.normalize();
        Hit<T> hit;
        bool h = raymarch_hit(A, ro, rd, op, T(0), op.max_dist, hit);
        T depth = h ? depth_preview_value(hit.t, op.max_dist) : T(0);
        sum = sum + depth;
        if (first && t_out) t_out[(size_t)y * op.width + x] = h ? (float)hit.t : -1.0f;
        first = false;
      }
      T avg = sum * (T(1) / (T)offs.size());
      uint8_t v = to_u8(avg);
      uint8_t* px = rgb + ((size_t)y * op.width + x) * 3;
      px[0] = px[1] = px[2] = v;
    }
    if (evals) total += tl_evals;
    tl_count = false;
  });
  if (evals) *evals = total.load();
  return FT_OK;
}

template <class T> int render_whitted_t(const ft_scene_ir* ir, uint32_t accel, const ft_render_options* o,
                                        const ft_ray_settings* rs, int32_t threads, int32_t row0, int32_t row1,
                                        uint8_t* rgb, float* hdr) {
  Scene<T> S;
  load_scene(ir, S);
  Accel<T> A;
  build_accel(S, accel, A);
  Ctx<T> c;
  c.S = &S;
  c.A = &A;
  c.op = opts_from<T>(o);
  c.max_depth = rs->max_depth < 1 ? 1 : rs->max_depth;
  c.has_env_fn = ir->env_fn >= 0;
  for (const Light<T>& L : S.lights)
    if (L.kind == FT_LIGHT_ENV) c.env_radiance = c.env_radiance + L.radiance;
  Cam<T> cam = make_camera(S, c.op);
  T aspect = (T)c.op.width / (T)c.op.height;
  uint32_t aa = rs->aa_samples < 1 ? 1 : rs->aa_samples;
  auto offs = sample_offsets<T>(aa);
  parallel_rows(row0, row1, threads, [&](int32_t y) {
    for (uint32_t x = 0; x < c.op.width; ++x) {
      Spec<T> sum;
      for (auto& so : offs) {
        T px = (((T)x + so.first) / (T)c.op.width) * T(2) - T(1);
        T py = T(1) - (((T)y + so.second) / (T)c.op.height) * T(2);
        V3<T> ro, rd;
        cam.gen(px * aspect, py, ro, rd);
        rd = rd.normalize();
        Spec<T> col = rs->debug_aov >= 0 ? trace_aov(c, ro, rd, rs->debug_aov) : trace_ray(c, ro, rd, Medium<T>::air(), 0);
        sum = sum + col;
      }
      Spec<T> avg = sum * (T(1) / (T)aa);
      size_t pix = (size_t)y * c.op.width + x;
      uint8_t* px = rgb + pix * 3;
      if (rs->debug_aov >= 0) { px[0] = to_u8(avg.x); px[1] = to_u8(avg.y); px[2] = to_u8(avg.z); }
      else { px[0] = to_u8_reinhard(avg.x); px[1] = to_u8_reinhard(avg.y); px[2] = to_u8_reinhard(avg.z); }
      if (hdr) { hdr[pix * 3] = (float)avg.x; hdr[pix * 3 + 1] = (float)avg.y; hdr[pix * 3 + 2] = (float)avg.z; }
    }
  });
  return FT_OK;
}

bool check_rows(const ft_render_options* o, int32_t& row0, int32_t& row1) {
  if (!o || o->width == 0 || o->height == 0) { g_err = "invalid render options"; return false; }
  if (row1 < 0 || row1 > (int32_t)o->height) row1 = (int32_t)o->height;
  if (row0 < 0) row0 = 0;
  if (row0 > row1) row0 = row1;
  return true;
}

}  // namespace

extern "C" {

const char* ftref_last_error(void) { return g_err.c_str(); }

int ftref_render_depth(const ft_scene_ir* ir, uint32_t accel, const ft_render_options* opts, uint32_t aa,
                       int32_t threads, int32_t precision, int32_t row0, int32_t row1, uint8_t* rgb, float* t_out) {
  if (!validate(ir) || !check_rows(opts, row0, row1) || !rgb) return FT_ERR_INVALID_ARGUMENT;
  if (precision == 1) return render_depth_t<double>(ir, accel, opts, aa, threads, row0, row1, rgb, t_out, nullptr);
  return render_depth_t<float>(ir, accel, opts, aa, threads, row0, row1, rgb, t_out, nullptr);
}

int ftref_depth_eval_count(const ft_scene_ir* ir, uint32_t accel, const ft_render_options* opts, uint32_t aa,
                           int32_t threads, int32_t row0, int32_t row1, uint64_t* evals) {
  if (!validate(ir) || !check_rows(opts, row0, row1) || !evals) return FT_ERR_INVALID_ARGUMENT;
  std::vector<uint8_t> rgb((size_t)opts->width * opts->height * 3);
  return render_depth_t<float>(ir, accel, opts, aa, threads, row0, row1, rgb.data(), nullptr, evals);
}

int ftref_render_whitted(const ft_scene_ir* ir, uint32_t accel, const ft_render_options* opts,
                         const ft_ray_settings* settings, int32_t threads, int32_t precision, int32_t row0,
                         int32_t row1, uint8_t* rgb, float* hdr) {
  if (!validate(ir) || !check_rows(opts, row0, row1) || !rgb || !settings) return FT_ERR_INVALID_ARGUMENT;
  if (precision == 1) return render_whitted_t<double>(ir, accel, opts, settings, threads, row0, row1, rgb, hdr);
  return render_whitted_t<float>(ir, accel, opts, settings, threads, row0, row1, rgb, hdr);
}

int ftref_distance_info(const ft_scene_ir* ir, uint32_t accel, int32_t precision, const float* p, int32_t n,
                        float* dist, uint32_t* object_id, uint32_t* material_id) {
  if (!validate(ir)) return FT_ERR_INVALID_ARGUMENT;
  auto run = [&](auto tag) {
    using T = decltype(tag);
    Scene<T> S;
    load_scene(ir, S);
    Accel<T> A;
    build_accel(S, accel, A);
    for (int32_t k = 0; k < n; ++k) {
      DistInfo<T> d = accel_distance_info(A, V3<T>(p[3 * k], p[3 * k + 1], p[3 * k + 2]));
      dist[k] = (float)d.distance;
      if (object_id) object_id[k] = d.object_id;
      if (material_id) material_id[k] = d.material_id;
    }
  };
  if (precision == 1) run(double(0)); else run(float(0));
  return FT_OK;
}

// Brute-force view of the accelerator (test support, pins accel_distance_info): per point
// and accel leaf the leaf's accel_leaf_distance_info (R:629-657) distance and object, and
// per leaf its rank in the order the traversal meets leaves whose bounds contain a point:
// the BVH's left-to-right leaf order (R:743-786 descends lhs first when both children's
// bounds contain the point), the leaf index for the bricks grid (its cell lists ascend,
// R:824-954).  With d_out null only *n_leaves is written.
int ftref_leaf_table(const ft_scene_ir* ir, uint32_t accel, const float* p, int32_t n, float* d_out,
                     uint32_t* obj_out, int32_t* rank_out, int32_t* n_leaves) {
  if (!validate(ir) || !n_leaves) return FT_ERR_INVALID_ARGUMENT;
  Scene<float> S;
  load_scene(ir, S);
  Accel<float> A;
  build_accel(S, accel, A);
  const int32_t L = (int32_t)A.leaves.size();
  *n_leaves = L;
  if (!d_out) return FT_OK;
  if (rank_out) {
    for (int32_t li = 0; li < L; ++li) rank_out[li] = li;
    if (accel == FT_ACCEL_BVH && A.bvh_root >= 0) {
      int32_t next = 0;
      std::vector<int> stack{A.bvh_root};
      while (!stack.empty()) {
        const BvhNode<float>& nd = A.bvh[stack.back()];
        stack.pop_back();
        if (nd.leaf >= 0) { rank_out[nd.leaf] = next++; continue; }
        stack.push_back(nd.rhs);
        stack.push_back(nd.lhs);
      }
    }
  }
  for (int32_t k = 0; k < n; ++k) {
    const V3<float> q(p[3 * k], p[3 * k + 1], p[3 * k + 2]);
    for (int32_t li = 0; li < L; ++li) {
      DistInfo<float> d = leaf_distance_info(S, A.leaves[li], q);
      d_out[(size_t)k * L + li] = d.distance;
      if (obj_out) obj_out[(size_t)k * L + li] = d.object_id;
    }
  }
  return FT_OK;
}

int ftref_lower_bound(const ft_scene_ir* ir, uint32_t accel, int32_t precision, const float* p, int32_t n, float* lb) {
  if (!validate(ir)) return FT_ERR_INVALID_ARGUMENT;
  auto run = [&](auto tag) {
    using T = decltype(tag);
    Scene<T> S;
    load_scene(ir, S);
    Accel<T> A;
    build_accel(S, accel, A);
    for (int32_t k = 0; k < n; ++k) lb[k] = (float)accel_lower_bound(A, V3<T>(p[3 * k], p[3 * k + 1], p[3 * k + 2]));
  };
  if (precision == 1) run(double(0)); else run(float(0));
  return FT_OK;
}

int ftref_eval_function(const ft_scene_ir* ir, int32_t fn, const float* params, int32_t n, float* out) {
  if (!validate(ir) || fn < 0 || fn >= (int32_t)ir->n_functions) return FT_ERR_INVALID_ARGUMENT;
  FnEval<float> F{ir};
  uint32_t np = ir->functions[fn].n_params;
  std::vector<V3<float>> args(np < 1 ? 1 : np);
  for (int32_t k = 0; k < n; ++k) {
    for (uint32_t a = 0; a < np; ++a)
      args[a] = V3<float>(params[(k * np + a) * 3], params[(k * np + a) * 3 + 1], params[(k * np + a) * 3 + 2]);
    V3<float> r = F(fn, args.data());
    out[3 * k] = r.x;
    out[3 * k + 1] = r.y;
    out[3 * k + 2] = r.z;
  }
  return FT_OK;
}

// Unit-level views of single oracle functions, so each can be checked against closed forms.

// sample_light (render_api.rs:171-289) of the scene's light `light` at `point`.
int ftref_light_sample(const ft_scene_ir* ir, uint32_t light, const float* point, uint32_t sample_index,
                       uint32_t sample_count, float* wi3, float* radiance3, float* distance) {
  if (!validate(ir)) return FT_ERR_INVALID_ARGUMENT;
  Scene<float> S;
  load_scene(ir, S);
  if (light >= S.lights.size()) { g_err = "light index out of range"; return FT_ERR_INVALID_ARGUMENT; }
  LightSample<float> ls = sample_light(S.lights[light], V3<float>(point[0], point[1], point[2]), sample_index,
                                       sample_count);
  wi3[0] = ls.wi.x; wi3[1] = ls.wi.y; wi3[2] = ls.wi.z;
  radiance3[0] = ls.radiance.x; radiance3[1] = ls.radiance.y; radiance3[2] = ls.radiance.z;
  *distance = ls.distance;
  return FT_OK;
}

// The BSDF value (eval_bsdf R:4905-4920 without hooks) of static material `material` for n
// direction triples (normal, wi, wo), 9 floats each; out: 3 floats each.
int ftref_material_eval(const ft_scene_ir* ir, uint32_t material, const float* nwiwo, int32_t n, float* out) {
  if (!validate(ir)) return FT_ERR_INVALID_ARGUMENT;
  Scene<float> S;
  load_scene(ir, S);
  if (material >= S.materials.size()) { g_err = "material index out of range"; return FT_ERR_INVALID_ARGUMENT; }
  for (int32_t k = 0; k < n; ++k) {
    const float* v = nwiwo + 9 * k;
    Spec<float> f = mat_eval(S.materials[material], V3<float>(v[0], v[1], v[2]), V3<float>(v[3], v[4], v[5]),
                             V3<float>(v[6], v[7], v[8]));
    out[3 * k] = f.x; out[3 * k + 1] = f.y; out[3 * k + 2] = f.z;
  }
  return FT_OK;
}

// fresnel_dielectric_scalar (R:5186-5204) for n (cos_i, eta_i, eta_t) triples.
int ftref_fresnel_dielectric(const float* args, int32_t n, float* out) {
  for (int32_t k = 0; k < n; ++k) out[k] = fresnel_dielectric(args[3 * k], args[3 * k + 1], args[3 * k + 2]);
  return FT_OK;
}

// refract (R:5080-5095) for n (incident, normal, eta) records of 7 floats; out: 3 floats
// each, all zero when refract returns None.
int ftref_refract(const float* args, int32_t n, float* out) {
  for (int32_t k = 0; k < n; ++k) {
    const float* a = args + 7 * k;
    V3<float> r;
    if (!refract(V3<float>(a[0], a[1], a[2]), V3<float>(a[3], a[4], a[5]), a[6], r)) r = V3<float>();
    out[3 * k] = r.x; out[3 * k + 1] = r.y; out[3 * k + 2] = r.z;
  }
  return FT_OK;
}

// rough_lobe_directions (ray.rs:383-417) of `ideal` at roughness `rough` with `count`
// lobes; out: up to 7 directions (3 floats each), *n_out of them.
int ftref_lobe_dirs(const float* ideal, float rough, uint32_t count, float* out, int32_t* n_out) {
  std::vector<V3<float>> d = lobe_dirs(V3<float>(ideal[0], ideal[1], ideal[2]), rough, count);
  for (size_t k = 0; k < d.size(); ++k) { out[3 * k] = d[k].x; out[3 * k + 1] = d[k].y; out[3 * k + 2] = d[k].z; }
  *n_out = (int32_t)d.size();
  return FT_OK;
}

int ftref_scene_bounds(const ft_scene_ir* ir, float* min3, float* max3, float* center3) {
  if (!validate(ir)) return FT_ERR_INVALID_ARGUMENT;
  Scene<float> S;
  load_scene(ir, S);
  Aabb<float> b = S.bounds[S.root];
  V3<float> c = sdf_center(S, S.root);
  for (int k = 0; k < 3; ++k) {
    if (min3) min3[k] = b.min[k];
    if (max3) max3[k] = b.max[k];
    if (center3) center3[k] = c[k];
  }
  return FT_OK;
}

}  // extern "C"

// ------------------------------------------------------- differentiable sphere tree
// The workload of the ft_diff_* entry points, which have no counterpart in the reference:
// n spheres joined by a balanced tree of union_round (op_union_round R:6029-6034), split
// [lo, mid) + [mid, hi) with mid = lo + (hi - lo) / 2, joins numbered in post order.
// Parameters: centre xyz and radius per sphere, then the blend radius k per join.
// Per pixel: march (raymarch_hit R:4091-4139 with its bisection), one Newton step on the
// hit distance, t* = t - f(x)/(grad f(x) . d), then the Lambert shade max(0, n . l) with n
// from central differences of step `normal_step` at x* = o + t* d.
namespace {

struct DiffNode {
  int leaf;      // sphere index, or -1
  int join;      // join index, or -1
  int lhs, rhs;  // node indices of the children of a join
};

void diff_build(int lo, int hi, std::vector<DiffNode>& out, int& joins) {
  if (hi - lo == 1) {
    out.push_back({lo, -1, -1, -1});
    return;
  }
  const int mid = lo + (hi - lo) / 2;
  diff_build(lo, mid, out, joins);
  const int l = (int)out.size() - 1;
  diff_build(mid, hi, out, joins);
  const int r = (int)out.size() - 1;
  out.push_back({-1, joins++, l, r});
}

template <class T> struct DiffTree {
  std::vector<DiffNode> nodes;
  std::vector<T> th;  // parameters
  int n = 0;
  Aabb<T> bounds;

  V3<T> centre(int i) const { return V3<T>(th[4 * i], th[4 * i + 1], th[4 * i + 2]); }
  T radius(int i) const { return th[4 * i + 3]; }
  T k(int j) const { return th[4 * n + j]; }

  T eval(V3<T> p, std::vector<T>& v) const {
    for (size_t a = 0; a < nodes.size(); ++a) {
      const DiffNode& d = nodes[a];
      if (d.leaf >= 0) v[a] = (p - centre(d.leaf)).length() - radius(d.leaf);
      else v[a] = op_union_round(v[d.lhs], v[d.rhs], k(d.join));
    }
    return v.back();
  }
  T value(V3<T> p) const {
    std::vector<T> v(nodes.size());
    return eval(p, v);
  }
  // value and analytic gradient with respect to p (reverse mode through the tree)
  T value_grad(V3<T> p, V3<T>& g) const {
    std::vector<T> v(nodes.size()), adj(nodes.size(), T(0));
    const T f = eval(p, v);
    adj.back() = T(1);
    g = V3<T>(0, 0, 0);
    for (size_t a = nodes.size(); a-- > 0;) {
      const DiffNode& d = nodes[a];
      if (d.leaf >= 0) {
        V3<T> q = p - centre(d.leaf);
        T l = q.length();
        if (l > T(0)) g = g + q * (adj[a] / l);
        continue;
      }
      const T av = v[d.lhs], bv = v[d.rhs];
      const T r = rmax(std::fabs(k(d.join)), T(1.0e-6));
      const T ux = rmax(r - av, T(0)), uy = rmax(r - bv, T(0));
      const T m = rmin(av, bv);
      const T L = std::sqrt(ux * ux + uy * uy);
      T da = (m > r && av <= bv) ? T(1) : T(0), db = (m > r && !(av <= bv)) ? T(1) : T(0);
      if (L > T(0)) {
        da = da + ux / L;
        db = db + uy / L;
      }
      adj[d.lhs] = adj[d.lhs] + adj[a] * da;
      adj[d.rhs] = adj[d.rhs] + adj[a] * db;
    }
    return f;
  }
};

template <class T> DiffTree<T> make_diff_tree(uint32_t n, const double* params) {
  DiffTree<T> D;
  D.n = (int)n;
  int joins = 0;
  diff_build(0, (int)n, D.nodes, joins);
  D.th.resize(4 * n + (n - 1));
  for (size_t a = 0; a < D.th.size(); ++a) D.th[a] = (T)params[a];
  T kmax = T(0);
  for (int j = 0; j < joins; ++j) kmax = rmax(kmax, std::fabs(D.k(j)));
  const T m = kmax + T(1.0e-3);
  for (int i = 0; i < (int)n; ++i) {
    const V3<T> c = D.centre(i);
    const T r = std::fabs(D.radius(i)) + m;
    Aabb<T> b{c - V3<T>::splat(r), c + V3<T>::splat(r)};
    D.bounds = i == 0 ? b : aabb_union(D.bounds, b);
  }
  return D;
}

// The march of raymarch_hit (R:4091-4139 + bisection R:4180-4200) against the tree.
template <class T> bool diff_march(const DiffTree<T>& D, V3<T> o, V3<T> d, const Opts<T>& op, T& t_hit) {
  T entry, exit;
  if (!ray_aabb(o, d, D.bounds, entry, exit)) return false;
  const T max_t = rmin(op.max_dist, exit);
  T traveled = rmax(T(0), rmax(entry, T(0)));
  T prev = traveled;
  for (uint32_t s = 0; s < op.max_steps; ++s) {
    if (traveled > max_t) return false;
    const T dist = D.value(o + d * traveled);
    if (std::fabs(dist) < op.epsilon) {
      T lo = prev, hi = traveled;
      for (int k = 0; k < 10; ++k) {
        const T mid = T(0.5) * (lo + hi);
        if (std::fabs(D.value(o + d * mid)) < op.epsilon) hi = mid; else lo = mid;
      }
      t_hit = hi;
      return true;
    }
    prev = traveled;
    traveled += rmax(std::fabs(dist) * rclamp(op.step_scale, T(0.05), T(1)), rmax(op.epsilon * T(0.5), T(1.0e-5)));
  }
  return false;
}

template <class T>
int diff_forward_t(uint32_t n, const double* params, const ft_view* views, uint32_t n_views, const ft_render_options* o,
                   const float light[3], float normal_step, int32_t threads, double* t_out, double* s_out) {
  const DiffTree<T> D = make_diff_tree<T>(n, params);
  const Opts<T> op = opts_from<T>(o);
  const T aspect = (T)op.width / (T)op.height;
  const V3<T> l = V3<T>(light[0], light[1], light[2]).normalize();
  const T e = (T)normal_step;
  const size_t npix = (size_t)op.width * op.height;
  for (uint32_t v = 0; v < n_views; ++v) {
    Cam<T> cam;
    cam.up = V3<T>(0, 1, 0);
    cam.origin = V3<T>(views[v].origin[0], views[v].origin[1], views[v].origin[2]);
    cam.target = V3<T>(views[v].target[0], views[v].target[1], views[v].target[2]);
    cam.fov = (T)views[v].fov_y_degrees;
    parallel_rows(0, (int32_t)op.height, threads, [&](int32_t y) {
      for (uint32_t x = 0; x < op.width; ++x) {
        const size_t idx = (size_t)v * npix + (size_t)y * op.width + x;
        const T px = (((T)x + T(0.5)) / (T)op.width) * T(2) - T(1);
        const T py = T(1) - (((T)y + T(0.5)) / (T)op.height) * T(2);
        V3<T> ro, rd;
        cam.gen(px * aspect, py, ro, rd);
        rd = rd.normalize();
        T th;
        if (!diff_march(D, ro, rd, op, th)) {
          t_out[idx] = -1.0;
          s_out[idx] = 0.0;
          continue;
        }
        V3<T> g;
        const T f = D.value_grad(ro + rd * th, g);
        const T gd = g.dot(rd);
        const T ts = std::fabs(gd) > T(1.0e-6) ? th - f / gd : th;
        const V3<T> xs = ro + rd * ts;
        const V3<T> G(D.value(V3<T>(xs.x + e, xs.y, xs.z)) - D.value(V3<T>(xs.x - e, xs.y, xs.z)),
                      D.value(V3<T>(xs.x, xs.y + e, xs.z)) - D.value(V3<T>(xs.x, xs.y - e, xs.z)),
                      D.value(V3<T>(xs.x, xs.y, xs.z + e)) - D.value(V3<T>(xs.x, xs.y, xs.z - e)));
        const T gl = G.length();
        const V3<T> nrm = gl > T(1.0e-12) ? G * (T(1) / gl) : V3<T>(0, 1, 0);
        t_out[idx] = (double)ts;
        s_out[idx] = (double)rmax(nrm.dot(l), T(0));
      }
    });
  }
  return FT_OK;
}

}  // namespace

extern "C" {

int ftref_diff_forward(uint32_t n_prims, const double* params, const ft_view* views, uint32_t n_views,
                       const ft_render_options* opts, const float light_dir[3], float normal_step, int32_t threads,
                       int32_t precision, double* t_out, double* s_out) {
  if (n_prims < 1 || !params || !views || !opts || !light_dir || !t_out || !s_out) return FT_ERR_INVALID_ARGUMENT;
  if (precision == 1)
    return diff_forward_t<double>(n_prims, params, views, n_views, opts, light_dir, normal_step, threads, t_out, s_out);
  return diff_forward_t<float>(n_prims, params, views, n_views, opts, light_dir, normal_step, threads, t_out, s_out);
}

}  // extern "C"
