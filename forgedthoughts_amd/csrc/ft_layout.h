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
constexpr int kMaxExpr = 96;       // nodes per Forge function
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
  float pad0;
  float pattern_b[3];
  float pad1;
};
static_assert(sizeof(DParams) == 144, "DParams layout");

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
};

}  // namespace ftd
