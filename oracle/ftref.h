/*
 * ftref.h — CPU oracle of the ForgedThoughts sphere-tracing path (TEST INFRASTRUCTURE).
 *
 * A restatement of the reference's CPU renderer (crates/forgedthoughts/src/renderer.rs,
 * renderer/ray.rs, materials/ *.rs, render_api.rs) over the scene IR of include/ft_abi.h.
 * It is the checker for the HIP backend and the timed CPU baseline; nothing in the
 * product links, loads or calls it.  Built by oracle/Makefile into oracle/libftref.so.
 *
 * `precision`: 0 = f32 (the reference's arithmetic), 1 = f64 (the same algorithm in
 * double precision, used to state the error of the f32 march).
 */
#ifndef FTREF_H
#define FTREF_H

#include <stdint.h>
#include "../include/ft_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Depth preview of rows [row0, row1) (render_depth_tile renderer.rs:3931-3971).
 * rgb: width*height*3 (only the requested rows are written).  t_out (nullable):
 * refined hit t of the first AA sample per pixel, -1 on a miss.  threads <= 0: all cores. */
int ftref_render_depth(const ft_scene_ir* ir, uint32_t accel, const ft_render_options* opts,
                       uint32_t aa, int32_t threads, int32_t precision, int32_t row0, int32_t row1,
                       uint8_t* rgb, float* t_out);

/* Whitted trace of rows [row0, row1) (render_ray_tile renderer.rs:3874-3929).
 * hdr (nullable): AA-averaged radiance per pixel before tone mapping. */
int ftref_render_whitted(const ft_scene_ir* ir, uint32_t accel, const ft_render_options* opts,
                         const ft_ray_settings* settings, int32_t threads, int32_t precision,
                         int32_t row0, int32_t row1, uint8_t* rgb, float* hdr);

/* accel.distance_info(p) for n points (xyz interleaved). */
int ftref_distance_info(const ft_scene_ir* ir, uint32_t accel, int32_t precision, const float* p,
                        int32_t n, float* dist, uint32_t* object_id, uint32_t* material_id);

/* accel.lower_bound(p) for n points. */
/* Per point and accel leaf the leaf's distance and object; per leaf its traversal rank
   (test support).  With d_out null only *n_leaves is written. */
int ftref_leaf_table(const ft_scene_ir* ir, uint32_t accel, const float* p, int32_t n, float* d_out,
                     uint32_t* obj_out, int32_t* rank_out, int32_t* n_leaves);
int ftref_lower_bound(const ft_scene_ir* ir, uint32_t accel, int32_t precision, const float* p,
                      int32_t n, float* lb);

/* Evaluate expression function `fn` at n argument sets; params: n * n_params * 3 floats
 * (scalars splatted); out: n * 3 floats. */
int ftref_eval_function(const ft_scene_ir* ir, int32_t fn, const float* params, int32_t n, float* out);

/* Scene AABB (accel.scene_bounds) and centre (sdf_center). */
int ftref_scene_bounds(const ft_scene_ir* ir, float* min3, float* max3, float* center3);

/* Per-ray march statistics used for work accounting: for rows [row0,row1) of a depth
 * render, the number of accel.distance_info evaluations (march steps + refinement +
 * normal + final info) summed over all pixels and samples. */
int ftref_depth_eval_count(const ft_scene_ir* ir, uint32_t accel, const ft_render_options* opts,
                           uint32_t aa, int32_t threads, int32_t row0, int32_t row1,
                           uint64_t* evals);

/* Differentiable sphere-tree workload (no reference counterpart; restates the forward of
 * the ft_diff_* entry points): n_prims spheres in a balanced union_round tree, params =
 * 4 * n_prims (centre xyz, radius) + n_prims - 1 (blend radius per join, post order).
 * Per view (1 spp, pixel centres): t_out = Newton-refined hit distance (-1 on a miss),
 * s_out = max(0, n . light) with n from central differences of step normal_step. */
int ftref_diff_forward(uint32_t n_prims, const double* params, const ft_view* views, uint32_t n_views,
                       const ft_render_options* opts, const float light_dir[3], float normal_step, int32_t threads,
                       int32_t precision, double* t_out, double* s_out);

/* Unit-level views of single functions (for closed-form checks):
 * sample_light of light `light` (wi, radiance, distance), eval_bsdf of a static material
 * for n (normal, wi, wo) triples, fresnel_dielectric for n (cos_i, eta_i, eta_t), refract
 * for n (incident[3], normal[3], eta) records (zero vector when it returns None). */
int ftref_light_sample(const ft_scene_ir* ir, uint32_t light, const float* point, uint32_t sample_index,
                       uint32_t sample_count, float* wi3, float* radiance3, float* distance);
int ftref_material_eval(const ft_scene_ir* ir, uint32_t material, const float* nwiwo, int32_t n, float* out);
int ftref_fresnel_dielectric(const float* args, int32_t n, float* out);
int ftref_refract(const float* args, int32_t n, float* out);
/* rough_lobe_directions: up to 7 directions into out (3 floats each), their number in *n_out. */
int ftref_lobe_dirs(const float* ideal, float rough, uint32_t count, float* out, int32_t* n_out);

const char* ftref_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
