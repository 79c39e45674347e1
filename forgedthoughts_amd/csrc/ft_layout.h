// ft_layout.h — memory layout of the device scene (shared by host builder and kernels).
// Plain C++: no HIP headers, so the host-side builder compiles with any C++ compiler.
#pragma once

#if !defined(__HIPCC_RTC__)
#include <stdint.h>
#endif

#include "ft_abi.h"

namespace ftd {

constexpr int kMaxStack = 24;      // value-stack depth of the node program
constexpr int kMaxPointStack = 4;  // nested modifier depth
constexpr int kMaxExpr = 96;       // values live at once in a Forge function
constexpr int kMaxWalk = 48;       // frames of the per-hit tree walks
constexpr int kMaxWalkMats = 12;   // pending materials of the material walk
constexpr int kMaxBvhStack = 48;
constexpr unsigned kBeginFlag = 0x80000000u;
constexpr unsigned kAabbFlopsHost = 12;  // FLOPs of one point-AABB lower bound

struct DNode {
  int kind, lhs, rhs, flags;            // flags bit0: rotated, bit1: parent needs the lower bound
  unsigned obj, mat;
  int leaf, fn_domain;
  int fn_distance, fn_post, prog_first, prog_last;
  float cx, cy, cz, cost;               // cost: FLOPs of one evaluation of this node
  float szs, szc, sys, syc;             // sin/cos of (-rot_z), (-rot_y)
  float sxs, sxc;                       // sin/cos of (-rot_x)
  int walk_split;                       // shading sweep from this node: smooth node whose children's
                                        // materials are traced apart, -1 none, -2 use the per-node walks
  int walk_ok;                          // 1: the shading sweep from this node fits its stacks
  float p[8];
  float bmin[4];
  float bmax[4];
};
static_assert(sizeof(DNode) == 160, "DNode layout");

// One word of the node program as a self-contained 64-byte record (one scalar load), in
// program order, so the evaluator streams records without a dependent node fetch.
//   code: node kind | kOpRotated | kOpNeedsLb, or kBeginFlag for a modifier's BEGIN word
//   node: the DNode index (ids, materials, the slow path of modifiers and custom SDFs)
//   primitives: c = centre, p = parameters, r = sin/cos of (-rot_z), (-rot_y), (-rot_x)
//   booleans, slice, smooth: p = parameters; with kOpNeedsLb, c / r[0..2] = AABB min / max
struct DOp {
  int code;
  int node;
  float c[3];
  float p[5];
  float r[6];
};
static_assert(sizeof(DOp) == 64, "DOp layout");
constexpr int kOpKindMask = 0xff;
constexpr int kOpRotated = 0x100;
constexpr int kOpNeedsLb = 0x200;

struct DParams {  // MaterialParams (materials/mod.rs:67-87)
  float color[3];
  float roughness;
  float metallic, specular, specular_weight, ior;
  float specular_color[3];
  float clearcoat;
  float clearcoat_roughness, transmission, emission_strength, thin_walled;
  float emission_color[3];
  float has_medium;
  float medium_ior, medium_density, pattern, pattern_scale;
  float medium_absorption[3];
  float model;
  float pattern_a[3];
  float hook;       // index into the hook table (DHooks), -1: no per-hit hooks (also after a
                    // lerp of two sets: blend_params drops the dynamic id, renderer.rs:7624)
  float pattern_b[3];
  float pad1;
};
static_assert(sizeof(DParams) == 144, "DParams layout");

// Per-hit hook functions of one material (ft_material_hooks without the material index).
struct DHooks {
  int color, roughness, ior, thin_walled, emission_color, emission_strength;
  int metallic, specular, specular_weight, transmission, specular_color, clearcoat_roughness, clearcoat, coat_mask;
  int bump, normal;
  int medium_ior, medium_absorption, medium_density;
  int eval;
};
static_assert(sizeof(DHooks) == 80, "DHooks layout");

struct DLight {
  int kind;
  unsigned samples;
  float radius, pad;
  float pos[3], pad1;
  float rad[3], pad2;
};

struct DBvh {
  float bmin[3];
  int left;   // >= 0 inner: child index; < 0 leaf: -(leaf+1)
  float bmax[3];
  int right;
};

struct DLeaf {
  float bmin[3];
  int node;
  float bmax[3];
  int is_custom;
  int op;       // the leaf's program record when the leaf is one primitive, else -1
  int pad[3];
};
static_assert(sizeof(DLeaf) == 48, "DLeaf layout");

}  // namespace ftd
