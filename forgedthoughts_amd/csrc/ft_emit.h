// ft_emit.h — scene emitter (scene IR -> specialised HIP source for hiprtc).
#pragma once

#include <string>

#include "ft_scene_build.h"

namespace fth {

// Root subtrees are emitted as straight-line code within these sizes (node count of the
// program range): below the lower limit the interpreter is as fast, above the upper one
// compile time grows past what a scene load should cost.
constexpr int kMinEmittedNodes = 1;
constexpr int kMaxEmittedNodes = 4096;

struct EmitPlan {
  bool functions = false;  // Forge functions become ftx_fn_<k>
  bool scene = false;      // the scene distance / lower bound become ftx_scene_info / ftx_lower_bound
  bool sweep = false;      // the shading sweep of the root becomes ftx_sweep (scene and no modifier nodes)
  bool any() const { return functions || scene; }
};

EmitPlan plan_emission(const HostScene& S, const ft_scene_ir* ir);

// Default policy: specialise when the interpreter's per-node dispatch dominates — any Forge
// function (interpreted per expression otherwise) or a root program of at least
// kAutoEmitNodes nodes.
constexpr int kAutoEmitNodes = 16;
bool emission_pays(const HostScene& S, const ft_scene_ir* ir, const EmitPlan& plan);

// Complete translation unit: the emitted functions followed by the kernels of ft_kernels.h.
std::string emit_scene_source(const HostScene& S, const ft_scene_ir* ir, const EmitPlan& plan);

}  // namespace fth
