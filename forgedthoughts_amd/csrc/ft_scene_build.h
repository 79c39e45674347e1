// ft_scene_build.h — host-side preparation of the device scene layout (see ft_device.h).
#pragma once

#include <string>
#include <vector>

#include "ft_layout.h"

namespace fth {

struct HostScene {
  std::vector<ftd::DNode> nodes;
  std::vector<unsigned> prog;
  std::vector<ftd::DOp> ops;  // prog as records, plus one padding record
  std::vector<unsigned> lbprog;
  std::vector<ftd::DParams> mats;
  std::vector<ftd::DHooks> hooks;  // per-hit hook functions, indexed by DParams::hook
  std::vector<float> objects;  // 9 per object
  std::vector<ftd::DLight> lights;
  std::vector<ft_function> fns;
  std::vector<ft_expr> exprs;
  std::vector<ftd::DLeaf> leaves;
  std::vector<ftd::DBvh> bvh;
  int root = 0;
  int env_fn = -1;
  int accel = 0;
  float bmin[3], bmax[3];
  float center[3];
  float env_rad[3];
  ft_camera camera;
  int max_stack = 0;        // value-stack depth the program needs
  int max_points = 0;       // modifier nesting
  bool simple = true;       // no custom SDF, modifier or noise node: the program calls no function
  bool root_walks = false;  // a columns boolean: its distance can fall below its bounds
  int max_walk_frames = 0;  // frames of the per-hit walks
  int max_walk_mats = 0;    // material results held by the material walk
  int max_fn_exprs = 0;
  int max_children = 2;     // rays a single hit can spawn (both split materials together)
  int max_bvh_depth = 0;
  int dynamic_material = -1;  // first material with per-hit hooks the GPU does not evaluate, or -1
};

// Validates `ir` and builds the device layout.  Returns false with `err` set on failure.
bool build_scene(const ft_scene_ir* ir, uint32_t accel, HostScene& out, std::string& err);

// Camera basis for one render (parse_camera renderer.rs:6263-6286 + PinholeCamera::generate_ray
// render_api.rs:132-144, everything except the per-pixel terms).
struct CamBasis {
  float origin[3], forward[3], right[3], up[3];
  float tan_fov;
};
CamBasis camera_basis(const HostScene& s, const ft_render_options& o);
// The same for an explicit camera (a Camera binding with origin, target and fov_y).
CamBasis camera_basis(const float origin[3], const float target[3], float fov_y_degrees);

}  // namespace fth
