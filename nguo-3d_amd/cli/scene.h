// The reference's scene JSON (engine-api/src/scene.rs Scene, params.rs SolverConfig,
// material.rs Material), read with its field names and defaults and lowered onto
// nguo_scene_desc as Scene::build_implicit lowers it onto ImplicitSolver (scene.rs:89-229).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nguo_hip.h"
#include "json.h"

namespace mguso {

// Input the scene format allows but this build does not simulate: the message names it.
struct Unsupported : std::runtime_error {
  using std::runtime_error::runtime_error;
};
// Malformed input (what serde would reject): the message names the field.
struct BadScene : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// SolverConfig (params.rs:41-122) fields the implicit path and the frame loop read.
struct SolverConfig {
  float dt = 1.0e-3f;
  double fps = 60.0;
  int32_t frames = 300;
  float gravity[3] = {0.0f, -9.8f, 0.0f};
  uint32_t cg_max_iter = 10000;
  float cg_tol = 1.0e-3f;
  float friction_eps = 1.0e-5f;
  uint32_t iterations = 4;
  bool self_collision = false;
  bool has_floor = false;
  float floor_y = 0.0f;
};

// Material (material.rs:9-73) with its defaults.
struct Material {
  float young_mod = 1000.0f, poisson = 0.35f, density = 1.0f, bend = 10.0f;
  float strain_limit = 0.0f, strain_limit_weft = 0.0f;
  float contact_gap = 1.0e-3f, contact_dynamic = 0.0f, contact_offset = 0.0f, friction = 0.0f;
  float membrane_warp = 0.0f, membrane_weft = 0.0f, membrane_shear = 0.0f;
};

// A lowered scene: the concatenated vertex space and the solver parameters, owning every array
// the descriptor points at.
struct LoweredScene {
  SolverConfig config;
  std::vector<float> positions, velocities, inv_mass;
  std::vector<uint32_t> tris;
  std::vector<uint8_t> static_mask;  // empty when no object is static
  std::vector<nguo_collider> colliders;
  nguo_params params{};
  nguo_scene_desc desc{};
  uint32_t n_verts() const { return (uint32_t)(positions.size() / 3); }
  void bind();  // points desc at the arrays (call after any move)
};

// Parses scene JSON text (BadScene on malformed input, Unsupported on a feature this build
// does not run) and lowers it; precond is NGUO_PRECOND_BLOCK or NGUO_PRECOND_SCHWARZ (the
// latter with 16-vertex domains and coarse weight 0.7, the CLI's schwarz-ml).
LoweredScene lower_scene_json(const std::string &text, int precond);

}  // namespace mguso
