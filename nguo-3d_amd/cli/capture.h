// Reader for the reference capture (the recorded inputs and per-frame vertex checkpoints of
// the CUDA ancestor, engine-io/src/golden.rs) and the mapping of a captured scene onto the
// device step (engine-cli/src/gauge_driver.rs build_from_scene), for `mguso gauge-scene`.
// The capture corpus is not distributed; the format is little-endian binaries under
// <scene>/input_bin plus param.toml, info.toml and checkpoints/vert_<N>.bin.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nguo_hip.h"

namespace mguso {

// A malformed or unreadable capture (the reference's anyhow errors): exit 1.
struct CaptureError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// golden.rs:53-170 decoders over whole little-endian buffers. Indices are u64 on disk and
// narrowed to u32 only when they fit (a larger word is an error, never truncated).
std::vector<float> decode_vec3_f64(const std::vector<uint8_t> &b);  // xyz per item
std::vector<float> decode_vec3_f32(const std::vector<uint8_t> &b);
std::vector<uint32_t> decode_index_u64(const std::vector<uint8_t> &b, size_t width, const char *what);

// golden.rs:35-50, 313-396 load_golden_input: vert.bin (f64, with the displacement groups
// added), tri.bin (u64, optional), vel.bin (f32), tet.bin (optional, may be empty),
// rod.bin (optional).
struct CaptureInput {
  std::vector<float> positions, velocities;  // xyz per vertex
  std::vector<uint32_t> tris, tets, rods;    // 3, 4, 2 per element
  uint32_t n_verts() const { return (uint32_t)(positions.size() / 3); }
};
CaptureInput load_capture_input(const std::string &dir);

// golden.rs:399-557 GoldenParams / read_param_toml: the [param] table with the reference's
// defaults for absent keys.
struct CaptureParams {
  float dt = 1.0e-3f, fps = 60.0f;
  uint32_t frames = 0;
  float gravity[3] = {0.0f, -9.8f, 0.0f};
  float world_scaling = 1.0f;
  bool include_face_mass = false;
  float air_density = 1.0e-3f, air_friction = 0.0f, isotropic_air_friction = 0.0f;
  float wind[3] = {0.0f, 0.0f, 0.0f};
  float eiganalysis_eps = 1.0e-2f, cg_tol = 1.0e-3f;
  uint32_t cg_max_iter = 10000, ccd_max_iter = 4096, min_newton_steps = 0;
  float target_toi = 0.25f, max_dx = 1.0f, friction_eps = 1.0e-5f;
  bool cubic_barrier = false, disable_contact = false;
  int friction_mode = 0;  // 0 min, 1 max, 2 mean
  float stitch_length_factor = 10.0f, constraint_ghat = 1.0e-3f, constraint_tol = 1.0e-2f;
};
CaptureParams parse_param_toml(const std::string &text);
CaptureParams read_param_toml(const std::string &dir);  // CaptureError when param.toml is absent

// golden.rs:562-598 pin-ind-<i>.bin: u64 index lists, one per pin block.
std::vector<std::vector<uint32_t>> read_pin_blocks(const std::string &dir);

// golden.rs:154-300 per-element parameter arrays of one prefix ("tri", "tet", "rod"); absent
// files give empty arrays. bend selects the bend-plasticity files.
enum RefModel : uint8_t { MODEL_ARAP = 0, MODEL_STVK = 1, MODEL_BW = 2, MODEL_SNHK = 3, MODEL_PDRD = 4 };
struct ElementParams {
  std::vector<uint8_t> model;
  std::vector<float> plasticity, plasticity_threshold, young_mod, poiss_rat, friction, strain_limit, contact_gap, density, bend,
      bend_rest_from_geometry, bending_damping, length_factor, contact_offset, shrink_x, shrink_y, pressure, deformation_damping;
};
ElementParams load_element_params(const std::string &dir, const std::string &prefix, bool bend);

// golden.rs:636-800 analytic colliders from info.toml [wall-<i>] / [sphere-<i>] and their
// keyframe binaries; sample() interpolates linearly, clamped at both ends.
struct SceneCollider {
  bool wall = false;
  float normal[3] = {0.0f, 1.0f, 0.0f};
  std::vector<float> times, positions, radii;  // positions xyz per keyframe
  float ghat = 1.0e-3f, friction = 0.0f, thickness = 1.0f, active_duration = -1.0f;
  bool animated() const { return positions.size() > 3 || radii.size() > 1; }
  void sample(float time, float p[3], float &r) const;
};
std::vector<SceneCollider> read_colliders(const std::string &dir);

// golden.rs:1265-1271 checkpoints/vert_<frame>.bin (f32 xyz).
std::vector<float> load_checkpoint(const std::string &dir, uint32_t frame);

// golden.rs:1280-1387 divergence of positions from a checkpoint, normalised by the
// checkpoint's bounding-box diagonal; gates max_rel < 8e-2, rms < 2e-2, centroid_rel < 6e-3.
struct GateReport {
  float max_rel = 0, rms = 0, centroid_rel = 0, diagonal = 1;
  bool passes() const { return max_rel < 8.0e-2f && rms < 2.0e-2f && centroid_rel < 6.0e-3f; }
};
GateReport compare_frames(const std::vector<float> &ours, const std::vector<float> &ref);

// gauge_driver.rs:28-70 DriverReport: what the mapping applied, and every declared feature it
// could not reproduce on the device step.
struct DriverReport {
  size_t n_verts = 0, n_tris = 0, n_tets = 0, n_stitches = 0, n_rods = 0, n_rod_hinges = 0;
  size_t n_baraff_witkin = 0, n_shell = 0, n_strain_limited = 0, n_pins_applied = 0, n_pins_kinematic = 0, n_pulls_applied = 0;
  size_t n_colliders = 0, n_colliders_animated = 0, n_solid_surface = 0;
  std::vector<std::string> unsupported;
};

// A captured scene lowered onto nguo_scene_desc (gauge_driver.rs:318-1200 build_from_scene),
// owning every array the descriptor points at.
struct CaptureScene {
  CaptureParams toml;
  bool has_toml = false;
  bool has_tets = false;
  std::vector<float> positions, velocities, inv_mass;
  std::vector<uint32_t> tris, shell_faces, rod_verts, rod_hinge_verts;
  std::vector<float> rod_rest, rod_k, rod_hinge_rest, rod_hinge_stiff;
  std::vector<nguo_collider> colliders;
  nguo_params params{};
  nguo_scene_desc desc{};
  uint32_t n_verts() const { return (uint32_t)(positions.size() / 3); }
  void bind();
};
CaptureScene lower_capture(const std::string &dir, DriverReport &rep);

}  // namespace mguso
