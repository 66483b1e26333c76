// The reference capture's reader (engine-io/src/golden.rs) and the mapping of a captured
// scene onto the device step (engine-cli/src/gauge_driver.rs:318-1200). The mapping keeps the
// reference driver's rules (mass lumping from densities, the tet-boundary faces carrying no
// in-plane material, rod stretch and bend coefficients, pins, analytic colliders, the
// [param] table) and, where the device step has no counterpart (per-element stiffness,
// Baraff-Witkin membranes, tets, soft and animated pins, strain limits, stitches, air), it does
// what the reference driver does for a gap: applies the scene-wide value and reports it.
#include "capture.h"

#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <tuple>

namespace mguso {
namespace {

bool exists(const std::string &p) {
  struct stat sb;
  return ::stat(p.c_str(), &sb) == 0;
}

std::vector<uint8_t> read_bytes(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw CaptureError("reading golden binary " + path + ": " + std::strerror(errno));
  return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

std::string read_text(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return std::string();
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

template <class T> T le(const uint8_t *p) {
  T v;
  std::memcpy(&v, p, sizeof v);  // the capture is little-endian, as is every host this builds for
  return v;
}

// golden.rs:236-254: absent -> empty; a length that is not a whole number of f32 is an error.
std::vector<float> read_f32_array(const std::string &path) {
  if (!exists(path)) return {};
  const auto b = read_bytes(path);
  if (b.size() % 4) throw CaptureError("f32 param " + path + " length " + std::to_string(b.size()) + " not a multiple of 4");
  std::vector<float> out(b.size() / 4);
  for (size_t i = 0; i < out.size(); ++i) out[i] = le<float>(&b[4 * i]);
  return out;
}

std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r\n"), z = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, z - a + 1);
}
std::string unquote(const std::string &s) {
  std::string t = trim(s);
  while (!t.empty() && t.front() == '"') t.erase(t.begin());
  while (!t.empty() && t.back() == '"') t.pop_back();
  return t;
}
// Rust's str::parse for f32/u32: the whole token must parse; otherwise the default stays.
bool parse_f32(const std::string &v, float &out) {
  std::string t = trim(v);
  if (t.empty()) return false;
  char *end = nullptr;
  errno = 0;
  float x = std::strtof(t.c_str(), &end);
  if (end != t.c_str() + t.size() || errno == ERANGE) return false;
  out = x;
  return true;
}
bool parse_u32(const std::string &v, uint32_t &out) {
  std::string t = trim(v);
  if (t.empty() || t.find_first_not_of("0123456789") != std::string::npos) return false;
  unsigned long long x = std::strtoull(t.c_str(), nullptr, 10);
  if (x > 0xffffffffULL) return false;
  out = (uint32_t)x;
  return true;
}
bool parse_vec3(const std::string &v, float out[3]) {
  std::string t = trim(v);
  while (!t.empty() && (t.front() == '[')) t.erase(t.begin());
  while (!t.empty() && (t.back() == ']')) t.pop_back();
  std::vector<float> nums;
  std::stringstream ss(t);
  std::string tok;
  while (std::getline(ss, tok, ',')) {
    float x;
    if (parse_f32(tok, x)) nums.push_back(x);
  }
  if (nums.size() != 3) return false;
  for (int k = 0; k < 3; ++k) out[k] = nums[k];
  return true;
}

// golden.rs:1205-1233 toml_section: key = value pairs of one [section], quotes stripped.
std::map<std::string, std::string> toml_section(const std::string &text, const std::string &want) {
  std::map<std::string, std::string> m;
  bool inside = false;
  std::stringstream ss(text);
  std::string line;
  while (std::getline(ss, line)) {
    std::string t = trim(line);
    if (t.size() >= 2 && t.front() == '[' && t.back() == ']') {
      inside = t.substr(1, t.size() - 2) == want;
      continue;
    }
    if (!inside) continue;
    size_t eq = t.find('=');
    if (eq == std::string::npos) continue;
    m[trim(t.substr(0, eq))] = unquote(t.substr(eq + 1));
  }
  return m;
}

// material.rs:31-44 lame_from_young_poisson
void lame(float young, float poisson, float &mu, float &lambda) {
  const float nu = std::isfinite(poisson) ? std::min(std::max(poisson, 1.0e-6f), 0.5f - 1.0e-4f) : 0.35f;
  mu = young / (2.0f * (1.0f + nu));
  lambda = young * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
}

// gauge_driver.rs:254-298 helpers
float mean_or(const std::vector<float> &v, float fallback) {
  if (v.empty()) return fallback;
  float s = 0.0f;
  for (float x : v) s += x;
  return s / (float)v.size();
}
bool varies(const std::vector<float> &v) {
  if (v.size() < 2) return false;
  float lo = INFINITY, hi = -INFINITY;
  for (float x : v) { lo = std::min(lo, x); hi = std::max(hi, x); }
  return hi - lo > 1e-6f * std::max(std::fabs(hi), 1.0f);
}
float poisson_representative(const std::vector<float> &v) {
  float s = 0.0f;
  size_t n = 0;
  for (float x : v)
    if (std::isfinite(x) && x > 0.0f && x < 0.5f) { s += x; ++n; }
  return n ? s / (float)n : 0.35f;
}

std::string sci(double v, int digits = 3) {
  char b[48];
  std::snprintf(b, sizeof b, "%.*e", digits, v);
  return b;
}

}  // namespace

// ---------------------------------------------------------------- decoders (golden.rs:53-170)
std::vector<float> decode_vec3_f64(const std::vector<uint8_t> &b) {
  if (b.size() % 24) throw CaptureError("f64 vec3 buffer length " + std::to_string(b.size()) + " not a multiple of 24");
  std::vector<float> out(b.size() / 8);
  for (size_t i = 0; i < out.size(); ++i) out[i] = (float)le<double>(&b[8 * i]);
  return out;
}

std::vector<float> decode_vec3_f32(const std::vector<uint8_t> &b) {
  if (b.size() % 12) throw CaptureError("f32 vec3 buffer length " + std::to_string(b.size()) + " not a multiple of 12");
  std::vector<float> out(b.size() / 4);
  for (size_t i = 0; i < out.size(); ++i) out[i] = le<float>(&b[4 * i]);
  return out;
}

std::vector<uint32_t> decode_index_u64(const std::vector<uint8_t> &b, size_t width, const char *what) {
  if (b.size() % (8 * width))
    throw CaptureError(std::string(what) + " buffer length " + std::to_string(b.size()) + " not a multiple of " + std::to_string(8 * width));
  std::vector<uint32_t> out(b.size() / 8);
  for (size_t i = 0; i < out.size(); ++i) {
    const uint64_t v = le<uint64_t>(&b[8 * i]);
    if (v > 0xffffffffULL) throw CaptureError(std::string(what) + " index " + std::to_string(v) + " at element " + std::to_string(i / width) + " exceeds u32::MAX");
    out[i] = (uint32_t)v;
  }
  return out;
}

// ---------------------------------------------------------------- inputs (golden.rs:313-396, 600-633)
CaptureInput load_capture_input(const std::string &dir) {
  const std::string ib = dir + "/input_bin/";
  CaptureInput in;
  in.positions = decode_vec3_f64(read_bytes(ib + "vert.bin"));
  // vert.bin holds the template; each vertex's displacement group is added on top
  if (exists(ib + "vert_dmap.bin") && exists(ib + "displacement.bin")) {
    const std::vector<float> disp = decode_vec3_f64(read_bytes(ib + "displacement.bin"));
    const auto raw = read_bytes(ib + "vert_dmap.bin");
    if (raw.size() % 4) throw CaptureError("vert_dmap.bin is " + std::to_string(raw.size()) + " bytes, not a whole number of u32");
    if (raw.size() / 4 != in.n_verts())
      throw CaptureError("vert_dmap has " + std::to_string(raw.size() / 4) + " entries but there are " + std::to_string(in.n_verts()) + " vertices");
    for (uint32_t i = 0; i < in.n_verts(); ++i) {
      const uint32_t g = le<uint32_t>(&raw[4 * i]);
      if ((size_t)g * 3 + 2 >= disp.size()) throw CaptureError("vert_dmap group " + std::to_string(g) + " exceeds the displacement table");
      for (int k = 0; k < 3; ++k) in.positions[3 * i + k] += disp[3 * (size_t)g + k];
    }
  }
  if (exists(ib + "tri.bin")) in.tris = decode_index_u64(read_bytes(ib + "tri.bin"), 3, "triangle");
  in.velocities = decode_vec3_f32(read_bytes(ib + "vel.bin"));
  if (in.velocities.size() != in.positions.size())
    throw CaptureError("vel count " + std::to_string(in.velocities.size() / 3) + " != vert count " + std::to_string(in.n_verts()));
  const uint32_t nv = in.n_verts();
  for (size_t f = 0; f < in.tris.size(); f += 3)
    if (in.tris[f] >= nv || in.tris[f + 1] >= nv || in.tris[f + 2] >= nv)
      throw CaptureError("triangle " + std::to_string(f / 3) + " references a vertex >= " + std::to_string(nv));
  if (exists(ib + "tet.bin")) {
    const auto raw = read_bytes(ib + "tet.bin");
    if (!raw.empty()) in.tets = decode_index_u64(raw, 4, "tet");
  }
  for (size_t t = 0; t < in.tets.size(); ++t)
    if (in.tets[t] >= nv) throw CaptureError("tet " + std::to_string(t / 4) + " references a vertex >= " + std::to_string(nv));
  if (exists(ib + "rod.bin")) in.rods = decode_index_u64(read_bytes(ib + "rod.bin"), 2, "rod");
  for (size_t r = 0; r < in.rods.size(); ++r)
    if (in.rods[r] >= nv) throw CaptureError("rod " + std::to_string(r / 2) + " references a vertex >= " + std::to_string(nv));
  return in;
}

CaptureParams parse_param_toml(const std::string &text) {
  CaptureParams p;
  std::stringstream ss(text);
  std::string line;
  while (std::getline(ss, line)) {
    const std::string t = trim(line);
    const size_t eq = t.find('=');
    if (eq == std::string::npos) continue;
    const std::string key = trim(t.substr(0, eq)), val = trim(t.substr(eq + 1));
    if (key == "dt") parse_f32(val, p.dt);
    else if (key == "fps") parse_f32(val, p.fps);
    else if (key == "frames") parse_u32(val, p.frames);
    else if (key == "world_scaling") parse_f32(val, p.world_scaling);
    else if (key == "include_face_mass") p.include_face_mass = val == "true";
    else if (key == "air_density") parse_f32(val, p.air_density);
    else if (key == "air_friction") parse_f32(val, p.air_friction);
    else if (key == "eiganalysis_eps") parse_f32(val, p.eiganalysis_eps);
    else if (key == "target_toi") parse_f32(val, p.target_toi);
    else if (key == "max_dx") parse_f32(val, p.max_dx);
    else if (key == "friction_eps") parse_f32(val, p.friction_eps);
    else if (key == "cg_tol") parse_f32(val, p.cg_tol);
    else if (key == "cg_max_iter") parse_u32(val, p.cg_max_iter);
    else if (key == "ccd_max_iter") parse_u32(val, p.ccd_max_iter);
    else if (key == "min_newton_steps") parse_u32(val, p.min_newton_steps);
    else if (key == "barrier") p.cubic_barrier = unquote(val) == "cubic";
    else if (key == "disable_contact") p.disable_contact = val == "true";
    else if (key == "isotropic_air_friction") parse_f32(val, p.isotropic_air_friction);
    else if (key == "wind") parse_vec3(val, p.wind);
    else if (key == "friction_mode") {
      const std::string m = unquote(val);
      p.friction_mode = m == "max" ? 1 : m == "mean" ? 2 : 0;  // unknown spellings fall to min
    } else if (key == "stitch_length_factor") parse_f32(val, p.stitch_length_factor);
    else if (key == "constraint_ghat") parse_f32(val, p.constraint_ghat);
    else if (key == "constraint_tol") parse_f32(val, p.constraint_tol);
    else if (key == "gravity") parse_vec3(val, p.gravity);
  }
  return p;
}

CaptureParams read_param_toml(const std::string &dir) {
  const std::string path = dir + "/param.toml";
  std::ifstream f(path, std::ios::binary);
  if (!f) throw CaptureError("reading param.toml: " + std::string(std::strerror(errno)));
  std::stringstream ss;
  ss << f.rdbuf();
  return parse_param_toml(ss.str());
}

std::vector<std::vector<uint32_t>> read_pin_blocks(const std::string &dir) {
  std::vector<std::vector<uint32_t>> out;
  for (uint32_t n = 0;; ++n) {
    const std::string p = dir + "/input_bin/pin-ind-" + std::to_string(n) + ".bin";
    if (!exists(p)) break;
    const auto b = read_bytes(p);
    if (b.size() % 8) throw CaptureError("pin-ind-" + std::to_string(n) + ".bin is " + std::to_string(b.size()) + " bytes, not a whole number of u64");
    std::vector<uint32_t> idx(b.size() / 8);
    for (size_t i = 0; i < idx.size(); ++i) {
      const uint64_t v = le<uint64_t>(&b[8 * i]);
      if (v > 0xffffffffULL) throw CaptureError("pin index " + std::to_string(v) + " exceeds u32::MAX");
      idx[i] = (uint32_t)v;
    }
    out.push_back(std::move(idx));
  }
  return out;
}

ElementParams load_element_params(const std::string &dir, const std::string &prefix, bool bend) {
  const std::string d = dir + "/input_bin/param/" + prefix + "-";
  ElementParams e;
  if (exists(d + "model.bin")) e.model = read_bytes(d + "model.bin");
  e.plasticity = read_f32_array(d + (bend ? "bend-plasticity.bin" : "plasticity.bin"));
  e.plasticity_threshold = read_f32_array(d + (bend ? "bend-plasticity-threshold.bin" : "plasticity-threshold.bin"));
  e.young_mod = read_f32_array(d + "young-mod.bin");
  e.poiss_rat = read_f32_array(d + "poiss-rat.bin");
  e.friction = read_f32_array(d + "friction.bin");
  e.strain_limit = read_f32_array(d + "strain-limit.bin");
  e.contact_gap = read_f32_array(d + "contact-gap.bin");
  e.density = read_f32_array(d + "density.bin");
  e.bend = read_f32_array(d + "bend.bin");
  e.bend_rest_from_geometry = read_f32_array(d + "bend-rest-from-geometry.bin");
  e.bending_damping = read_f32_array(d + "bending-damping.bin");
  e.length_factor = read_f32_array(d + "length-factor.bin");
  e.contact_offset = read_f32_array(d + "contact-offset.bin");
  e.shrink_x = read_f32_array(d + "shrink-x.bin");
  e.shrink_y = read_f32_array(d + "shrink-y.bin");
  e.pressure = read_f32_array(d + "pressure.bin");
  e.deformation_damping = read_f32_array(d + "deformation-damping.bin");
  return e;
}

void SceneCollider::sample(float time, float p[3], float &r) const {
  const size_t nk = positions.size() / 3;
  if (nk == 0) { p[0] = p[1] = p[2] = 0.0f; r = 0.0f; return; }
  auto pos = [&](size_t i, int k) { return positions[3 * std::min(i, nk - 1) + k]; };
  auto rad = [&](size_t i) { return radii.empty() ? 0.0f : radii[std::min(i, radii.size() - 1)]; };
  auto at = [&](size_t i) { for (int k = 0; k < 3; ++k) p[k] = pos(i, k); r = rad(i); };
  if (times.size() < 2 || time <= times[0]) { at(0); return; }
  const size_t last = times.size() - 1;
  if (time >= times[last]) { at(last); return; }
  size_t k = 0;
  for (size_t w = 0; w + 1 < times.size(); ++w)
    if (time >= times[w] && time <= times[w + 1]) { k = w; break; }
  const float t0 = times[k], t1 = times[k + 1];
  const float u = std::fabs(t1 - t0) > 0.0f ? (time - t0) / (t1 - t0) : 0.0f;
  for (int c = 0; c < 3; ++c) p[c] = pos(k, c) + (pos(k + 1, c) - pos(k, c)) * u;
  r = rad(k) + (rad(k + 1) - rad(k)) * u;
}

std::vector<SceneCollider> read_colliders(const std::string &dir) {
  const std::string info = read_text(dir + "/info.toml"), ib = dir + "/input_bin/";
  std::vector<SceneCollider> out;
  for (const char *kind : {"wall", "sphere"}) {
    for (int i = 0; i < 64; ++i) {
      const std::string tag = std::string(kind) + "-" + std::to_string(i);
      const std::string pos_p = ib + std::string(kind) + "-pos-" + std::to_string(i) + ".bin";
      if (!exists(pos_p)) break;
      const auto sec = toml_section(info, tag);
      auto num = [&](const char *k, float d) {
        auto it = sec.find(k);
        float v = d;
        if (it != sec.end() && parse_f32(it->second, v)) return v;
        return d;
      };
      SceneCollider c;
      c.positions = decode_vec3_f64(read_bytes(pos_p));
      const std::string tp = ib + std::string(kind) + "-timing-" + std::to_string(i) + ".bin";
      if (exists(tp)) {
        const auto b = read_bytes(tp);
        if (b.size() % 8) throw CaptureError(std::string(kind) + "-timing-" + std::to_string(i) + ".bin is not a whole number of f64");
        for (size_t k = 0; k < b.size(); k += 8) c.times.push_back((float)le<double>(&b[k]));
      }
      const std::string rp = ib + std::string(kind) + "-radius-" + std::to_string(i) + ".bin";
      if (exists(rp)) {
        const auto b = read_bytes(rp);
        if (b.size() % 4) throw CaptureError(std::string(kind) + "-radius-" + std::to_string(i) + ".bin is not a whole number of f32");
        for (size_t k = 0; k < b.size(); k += 4) c.radii.push_back(le<float>(&b[k]));
      }
      c.wall = std::strcmp(kind, "wall") == 0;
      if (c.wall) { c.normal[0] = num("nx", 0.0f); c.normal[1] = num("ny", 1.0f); c.normal[2] = num("nz", 0.0f); }
      c.ghat = num("contact-gap", 1.0e-3f);
      c.friction = num("friction", 0.0f);
      c.thickness = num("thickness", 1.0f);
      c.active_duration = num("active-duration", -1.0f);
      out.push_back(std::move(c));
    }
  }
  return out;
}

std::vector<float> load_checkpoint(const std::string &dir, uint32_t frame) {
  return decode_vec3_f32(read_bytes(dir + "/checkpoints/vert_" + std::to_string(frame) + ".bin"));
}

GateReport compare_frames(const std::vector<float> &ours, const std::vector<float> &ref) {
  if (ours.size() != ref.size())
    throw CaptureError("frame vertex count mismatch: ours " + std::to_string(ours.size() / 3) + " vs reference " + std::to_string(ref.size() / 3));
  const size_t nv = ref.size() / 3, n = std::max<size_t>(nv, 1);
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  double rc[3] = {0, 0, 0}, oc[3] = {0, 0, 0};
  for (int k = 0; k < 3; ++k) {
    for (size_t i = 0; i < nv; ++i) {
      lo[k] = std::min(lo[k], ref[3 * i + k]);
      hi[k] = std::max(hi[k], ref[3 * i + k]);
      rc[k] += (double)ref[3 * i + k];
    }
    for (size_t i = 0; i < nv; ++i) oc[k] += (double)ours[3 * i + k];
  }
  float d2 = 0.0f;
  for (int k = 0; k < 3; ++k) d2 += (hi[k] - lo[k]) * (hi[k] - lo[k]);
  float diag = std::sqrt(d2);
  if (!(diag > 1.0e-12f)) diag = 1.0f;
  float max_d = 0.0f;
  double sum_sq = 0.0;
  for (size_t i = 0; i < nv; ++i) {
    float e2 = 0.0f;
    for (int k = 0; k < 3; ++k) e2 += (ours[3 * i + k] - ref[3 * i + k]) * (ours[3 * i + k] - ref[3 * i + k]);
    max_d = std::max(max_d, std::sqrt(e2));
    sum_sq += (double)e2;
  }
  double c2 = 0.0;
  for (int k = 0; k < 3; ++k) c2 += ((oc[k] - rc[k]) / (double)n) * ((oc[k] - rc[k]) / (double)n);
  GateReport g;
  g.max_rel = max_d / diag;
  g.rms = (float)std::sqrt(sum_sq / (double)n) / diag;
  g.centroid_rel = (float)std::sqrt(c2) / diag;
  g.diagonal = diag;
  return g;
}

// ---------------------------------------------------------------- mapping (gauge_driver.rs:318-1200)
void CaptureScene::bind() {
  desc = nguo_scene_desc{};
  desc.n_verts = n_verts();
  desc.positions = positions.data();
  desc.velocities = velocities.data();
  desc.inv_mass = inv_mass.data();
  desc.n_tris = (uint32_t)(tris.size() / 3);
  desc.tris = tris.empty() ? nullptr : tris.data();
  desc.static_collider = nullptr;
  desc.n_shell_faces = (uint32_t)(shell_faces.size() / 3);
  desc.shell_faces = shell_faces.empty() ? nullptr : shell_faces.data();
  desc.n_rods = (uint32_t)(rod_verts.size() / 2);
  desc.rod_verts = rod_verts.empty() ? nullptr : rod_verts.data();
  desc.rod_rest = rod_rest.empty() ? nullptr : rod_rest.data();
  desc.rod_k = rod_k.empty() ? nullptr : rod_k.data();
  desc.n_rod_hinges = (uint32_t)(rod_hinge_verts.size() / 3);
  desc.rod_hinge_verts = rod_hinge_verts.empty() ? nullptr : rod_hinge_verts.data();
  desc.rod_hinge_rest = rod_hinge_rest.empty() ? nullptr : rod_hinge_rest.data();
  desc.rod_hinge_stiff = rod_hinge_stiff.empty() ? nullptr : rod_hinge_stiff.data();
  desc.n_colliders = (uint32_t)colliders.size();
  desc.colliders = colliders.empty() ? nullptr : colliders.data();
  desc.params = params;
}

CaptureScene lower_capture(const std::string &dir, DriverReport &rep) {
  const CaptureInput in = load_capture_input(dir);
  if (in.n_verts() == 0) throw CaptureError("scene has no vertices");
  const ElementParams tri_p = load_element_params(dir, "tri", false), bend_p = load_element_params(dir, "tri", true);
  const ElementParams tet_p = load_element_params(dir, "tet", false), rod_p = load_element_params(dir, "rod", true);
  CaptureScene S;
  S.has_toml = exists(dir + "/param.toml");
  if (S.has_toml) S.toml = read_param_toml(dir);
  const CaptureParams &T = S.toml;
  const uint32_t nv = in.n_verts();
  const size_t nf = in.tris.size() / 3, nt = in.tets.size() / 4, nr = in.rods.size() / 2;
  rep = DriverReport{};
  rep.n_verts = nv;
  rep.n_tris = nf;
  rep.n_tets = nt;
  S.has_tets = nt > 0;
  if (nt) rep.unsupported.push_back(std::to_string(nt) + " tets (volumetric solids) are not on the device step; the solids are driven by their surface faces' contact only");

  // a face shared by exactly one tet is a solid's surface: contact geometry, no in-plane material
  std::vector<uint8_t> solid(nf, 0);
  if (nt) {
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, uint32_t> count;
    auto key = [](uint32_t a, uint32_t b, uint32_t c) {
      uint32_t k[3] = {a, b, c};
      std::sort(k, k + 3);
      return std::make_tuple(k[0], k[1], k[2]);
    };
    static const int tf[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
    for (size_t t = 0; t < nt; ++t)
      for (auto &f : tf) ++count[key(in.tets[4 * t + f[0]], in.tets[4 * t + f[1]], in.tets[4 * t + f[2]])];
    for (size_t f = 0; f < nf; ++f) {
      auto it = count.find(key(in.tris[3 * f], in.tris[3 * f + 1], in.tris[3 * f + 2]));
      solid[f] = it != count.end() && it->second == 1;
    }
  }
  for (uint8_t s : solid) rep.n_solid_surface += s;

  // faces by constitutive model: ARAP / StVK are the device's isotropic shells (driven as ARAP,
  // as the reference driver drives them); Baraff-Witkin faces are driven as the same shells, a
  // substituted material; SNHk / Pdrd faces carry no in-plane material, as in the reference
  std::vector<uint8_t> elastic(nf, 0);
  size_t n_bw = 0, n_iso = 0, n_snhk = 0, n_pdrd = 0;
  if (tri_p.model.size() == nf) {
    for (size_t f = 0; f < nf; ++f) {
      if (solid[f]) continue;
      const uint8_t m = tri_p.model[f];
      if (m == MODEL_ARAP || m == MODEL_STVK) { elastic[f] = 1; ++n_iso; }
      else if (m == MODEL_BW) { elastic[f] = 1; ++n_bw; }
      else if (m == MODEL_SNHK) ++n_snhk;
      else if (m == MODEL_PDRD) ++n_pdrd;
    }
  } else {
    rep.unsupported.push_back("no per-face model array; all faces treated as Baraff-Witkin");
    for (size_t f = 0; f < nf; ++f)
      if (!solid[f]) { elastic[f] = 1; ++n_bw; }
  }
  if (n_snhk) rep.unsupported.push_back(std::to_string(n_snhk) + " ELASTIC faces use the SNHk model (not implemented)");
  if (n_pdrd) rep.unsupported.push_back(std::to_string(n_pdrd) + " ELASTIC faces use the Pdrd model (not implemented)");
  if (n_bw)
    rep.unsupported.push_back(std::to_string(n_bw) + " faces use the Baraff-Witkin membrane, which is not on the device step; they are driven as ARAP shells (a substituted material)");
  rep.n_baraff_witkin = n_bw;
  rep.n_shell = n_iso;
  if (!tet_p.model.empty() && nt) {
    bool uniform = std::all_of(tet_p.model.begin(), tet_p.model.end(), [&](uint8_t m) { return m == tet_p.model[0]; });
    if (!uniform) rep.unsupported.push_back("per-tet model VARIES within the scene");
  }

  S.positions = in.positions;
  S.velocities = in.velocities;
  S.tris = in.tris;
  for (size_t f = 0; f < nf; ++f)
    if (elastic[f]) S.shell_faces.insert(S.shell_faces.end(), &in.tris[3 * f], &in.tris[3 * f] + 3);

  // per-vertex mass: density * area / 3 per face (solid surfaces only with include_face_mass),
  // density * volume / 4 per tet, density * length / 2 per rod segment (builder.rs:485-529)
  auto P = [&](uint32_t v, int k) { return S.positions[3 * (size_t)v + k]; };
  std::vector<float> mass(nv, 0.0f), lumped;
  const bool have_tri_rho = tri_p.density.size() == nf, have_tet_rho = tet_p.density.size() == nt, have_rod_rho = rod_p.density.size() == nr;
  bool massed = false;
  if (have_tri_rho || have_tet_rho || (have_rod_rho && nr)) {
    massed = true;
    if (have_tri_rho) {
      for (size_t f = 0; f < nf; ++f) {
        if (solid[f] && !T.include_face_mass) continue;
        const uint32_t a = in.tris[3 * f], b = in.tris[3 * f + 1], c = in.tris[3 * f + 2];
        const float e1[3] = {P(b, 0) - P(a, 0), P(b, 1) - P(a, 1), P(b, 2) - P(a, 2)};
        const float e2[3] = {P(c, 0) - P(a, 0), P(c, 1) - P(a, 1), P(c, 2) - P(a, 2)};
        const float cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
        const float area = 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
        const float m = tri_p.density[f] * area / 3.0f;
        mass[a] += m; mass[b] += m; mass[c] += m;
      }
    } else if (nf) {
      rep.unsupported.push_back("no per-face density array; face mass omitted");
    }
    if (have_tet_rho) {
      for (size_t t = 0; t < nt; ++t) {
        const uint32_t *q = &in.tets[4 * t];
        float d[3][3];
        for (int r = 0; r < 3; ++r)
          for (int k = 0; k < 3; ++k) d[r][k] = P(q[r + 1], k) - P(q[0], k);
        const float cx = d[0][1] * d[1][2] - d[0][2] * d[1][1], cy = d[0][2] * d[1][0] - d[0][0] * d[1][2], cz = d[0][0] * d[1][1] - d[0][1] * d[1][0];
        const float vol = std::fabs(cx * d[2][0] + cy * d[2][1] + cz * d[2][2]) / 6.0f;
        const float m = tet_p.density[t] * vol / 4.0f;
        for (int k = 0; k < 4; ++k) mass[q[k]] += m;
      }
    } else if (nt) {
      rep.unsupported.push_back("no per-tet density array; tet mass omitted");
    }
    if (have_rod_rho) {
      for (size_t r = 0; r < nr; ++r) {
        const uint32_t a = in.rods[2 * r], b = in.rods[2 * r + 1];
        const float dx = P(b, 0) - P(a, 0), dy = P(b, 1) - P(a, 1), dz = P(b, 2) - P(a, 2);
        const float m = rod_p.density[r] * std::sqrt(dx * dx + dy * dy + dz * dz) / 2.0f;
        mass[a] += m; mass[b] += m;
      }
    } else if (nr) {
      rep.unsupported.push_back("no per-rod density array; rod mass omitted");
    }
    float sum = 0.0f;
    size_t npos = 0, nzero = 0;
    for (float m : mass) {
      if (m > 0.0f) { sum += m; ++npos; }
      else ++nzero;
    }
    const float fallback = npos ? sum / (float)npos : 1.0f;
    if (nzero)
      rep.unsupported.push_back(std::to_string(nzero) + " vertices carry no element mass; each was given the mean mass " + sci(fallback, 4) +
                                " so it stays dynamic rather than acting as a pin");
    lumped = mass;  // the rod bend coefficient reads the lumped mass before the fallback (gauge_driver.rs:542-546)
    for (float &m : mass)
      if (!(m > 0.0f)) m = fallback;
  } else {
    rep.unsupported.push_back("scene ships no density arrays; a uniform vertex mass of 1.0 is used");
    std::fill(mass.begin(), mass.end(), 1.0f);
  }
  S.inv_mass.resize(nv);
  for (uint32_t i = 0; i < nv; ++i) S.inv_mass[i] = 1.0f / mass[i];

  // stiffness: the device shells carry one (mu, lambda) per scene
  const float tri_young = mean_or(tri_p.young_mod, 1.0e3f);
  float mu = 0.0f, lam = 0.0f;
  lame(tri_young, poisson_representative(tri_p.poiss_rat), mu, lam);
  if (varies(tri_p.young_mod)) rep.unsupported.push_back("per-face young-mod varies; the device shells take the scene-wide mean " + sci(tri_young));
  if (varies(tri_p.poiss_rat)) rep.unsupported.push_back("per-face poisson ratio varies; the device shells take the scene-wide representative value");
  // plasticity and the features the reference reports only when active (gauge_driver.rs:880-896)
  for (auto [name, arr] : {std::pair<const char *, const std::vector<float> *>{"face plasticity", &tri_p.plasticity},
                           {"bend plasticity", &bend_p.plasticity}, {"tet plasticity", &tet_p.plasticity}})
    if (std::any_of(arr->begin(), arr->end(), [](float v) { return v > 0.0f; })) rep.unsupported.push_back(std::string(name) + " is ACTIVE in this scene but is not modeled");
  for (auto [name, arr, identity] :
       {std::tuple<const char *, const std::vector<float> *, float>{"shrink-x", &tri_p.shrink_x, 1.0f}, {"shrink-y", &tri_p.shrink_y, 1.0f},
        {"pressure", &tri_p.pressure, 0.0f}, {"deformation-damping", &tri_p.deformation_damping, 0.0f}, {"contact-offset", &tri_p.contact_offset, 0.0f}})
    if (std::any_of(arr->begin(), arr->end(), [&](float v) { return std::fabs(v - identity) > 1e-6f; }))
      rep.unsupported.push_back(std::string(name) + " is ACTIVE in this scene but is not modeled");
  if (std::any_of(tri_p.strain_limit.begin(), tri_p.strain_limit.end(), [](float v) { return v > 0.0f; })) {
    size_t n = 0;
    for (size_t f = 0; f < std::min(nf, tri_p.strain_limit.size()); ++f) n += !solid[f] && tri_p.strain_limit[f] > 0.0f;
    rep.unsupported.push_back(std::to_string(n) + " faces carry the strain-limit barrier, which is not on the device step");
  }
  const float gap_mean = mean_or(tri_p.contact_gap, 1.0e-2f);
  if (varies(tri_p.contact_gap))
    rep.unsupported.push_back("per-face contact-gap varies; our dhat is scene-wide so the mean " + sci(gap_mean) + " is used");
  const float bend_k = mean_or(bend_p.bend, 1.0f);
  if (!bend_p.bend.empty())
    rep.unsupported.push_back("per-hinge bending coefficients from the per-face bend array (area-weighted, implicit.rs:998-1070) are not on the device step; the scene-wide mean " +
                              sci(bend_k) + " is used with the device's hinge coefficient");
  float fric = 0.0f;
  if (tri_p.friction.size() == nf && nf) {
    fric = mean_or(tri_p.friction, 0.0f);
    if (varies(tri_p.friction)) rep.unsupported.push_back("per-face friction varies; the device's contact friction is scene-wide (mean " + sci(fric) + ")");
  } else if (!tri_p.friction.empty()) {
    rep.unsupported.push_back("per-face friction array length does not match the face count; friction disabled");
  }
  if (S.has_toml && T.friction_mode != 0) rep.unsupported.push_back("friction_mode combines per-participant coefficients; with one scene-wide coefficient every mode gives it");

  // rod segments and rod bending (gauge_driver.rs:956-1095)
  if (nr) {
    const auto have = [&](const std::vector<float> &v) { return v.size() == nr; };
    if (have(rod_p.young_mod) && have(rod_p.density) && have(rod_p.length_factor)) {
      for (size_t r = 0; r < nr; ++r) {
        const uint32_t a = in.rods[2 * r], b = in.rods[2 * r + 1];
        const float dx = P(b, 0) - P(a, 0), dy = P(b, 1) - P(a, 1), dz = P(b, 2) - P(a, 2);
        const float l = std::sqrt(dx * dx + dy * dy + dz * dz), lf = rod_p.length_factor[r];
        S.rod_verts.push_back(a);
        S.rod_verts.push_back(b);
        S.rod_rest.push_back(std::max(l * lf, 1e-9f));
        S.rod_k.push_back(rod_p.young_mod[r] * rod_p.density[r] / std::max(lf, 1e-9f));
      }
      rep.n_rods = nr;
    } else {
      rep.unsupported.push_back(std::to_string(nr) + " rod segments ship no complete young-mod/density/length-factor triple; rod stretch omitted");
    }
    if (rod_p.bend.size() == nr && massed) {
      std::vector<uint8_t> has_face(nv, 0);
      for (uint32_t v : in.tris) has_face[v] = 1;
      std::vector<uint32_t> deg(nv, 0);
      std::vector<uint32_t> nbr(2 * (size_t)nv, 0xffffffffu);
      std::vector<float> bend_sum(nv, 0.0f);
      for (size_t r = 0; r < nr; ++r) {
        const uint32_t a = in.rods[2 * r], b = in.rods[2 * r + 1];
        for (auto [v, other] : {std::pair<uint32_t, uint32_t>{a, b}, {b, a}}) {
          if (deg[v] < 2) nbr[2 * (size_t)v + deg[v]] = other;
          ++deg[v];
        }
        bend_sum[a] += rod_p.bend[r];
        bend_sum[b] += rod_p.bend[r];
      }
      for (uint32_t v = 0; v < nv; ++v) {
        if (deg[v] != 2 || has_face[v]) continue;
        const float stiff = 0.5f * bend_sum[v] * lumped[v];  // mean segment bend x vertex mass (energy.cu:508-511)
        S.rod_hinge_verts.push_back(nbr[2 * (size_t)v]);
        S.rod_hinge_verts.push_back(v);
        S.rod_hinge_verts.push_back(nbr[2 * (size_t)v + 1]);
        S.rod_hinge_rest.push_back(3.14159265358979323846f);
        S.rod_hinge_stiff.push_back(stiff);
        rep.n_rod_hinges += stiff > 0.0f;
      }
    } else if (std::any_of(rod_p.bend.begin(), rod_p.bend.end(), [](float b) { return b > 0.0f; })) {
      rep.unsupported.push_back("rod segments carry bending but the bend/density arrays are incomplete; rod bending omitted");
    }
    const size_t geo = std::count_if(rod_p.bend_rest_from_geometry.begin(), rod_p.bend_rest_from_geometry.end(), [](float g) { return g != 0.0f; });
    if (geo) rep.unsupported.push_back(std::to_string(geo) + " rod segments request bend-rest-from-geometry; the straight rest angle pi is used instead");
    const size_t damp = std::count_if(rod_p.bending_damping.begin(), rod_p.bending_damping.end(), [](float d) { return d > 0.0f; });
    if (damp) rep.unsupported.push_back(std::to_string(damp) + " rod segments set bending-damping; the reference's lagged bending damping is not modeled");
    if (std::any_of(rod_p.contact_offset.begin(), rod_p.contact_offset.end(), [](float o) { return o > 0.0f; }))
      rep.unsupported.push_back("rod contact-offset (strand radius) is not on the device step; strands are zero-thickness");
    if (std::any_of(rod_p.contact_gap.begin(), rod_p.contact_gap.end(), [&](float g) { return g > 0.0f && std::fabs(g - gap_mean) > 1e-9f; }))
      rep.unsupported.push_back("rod contact-gap differs from the scene-wide contact_dhat, which is used");
    for (auto [name, v] : {std::pair<const char *, const std::vector<float> *>{"strain-limit", &rod_p.strain_limit}, {"bend-plasticity", &rod_p.plasticity},
                           {"deformation-damping", &rod_p.deformation_damping}, {"pressure", &rod_p.pressure}}) {
      const size_t on = std::count_if(v->begin(), v->end(), [](float x) { return x > 0.0f; });
      if (on) rep.unsupported.push_back(std::to_string(on) + " rod segments set " + name + "; not modeled");
    }
  }

  // pins: the device holds a pinned vertex fixed (inverse mass 0). The reference's FixPair is a
  // soft constraint of radius constraint_ghat and its blocks may animate (info.toml operations)
  // or pull (pull > 0); those are reported and held at their t = 0 positions.
  const std::string info = read_text(dir + "/info.toml");
  const auto blocks = read_pin_blocks(dir);
  size_t n_anim = 0, n_pull = 0;
  for (size_t b = 0; b < blocks.size(); ++b) {
    const auto head = toml_section(info, "pin-" + std::to_string(b));
    uint32_t ops = 0;
    float pull = 0.0f;
    if (head.count("operation_count")) parse_u32(head.at("operation_count"), ops);
    if (head.count("pull")) parse_f32(head.at("pull"), pull);
    for (uint32_t v : blocks[b]) {
      if (v >= nv) throw CaptureError("pin index " + std::to_string(v) + " is not a vertex of the scene");
      if (pull > 0.0f) { ++n_pull; continue; }
      S.inv_mass[v] = 0.0f;
      ++rep.n_pins_applied;
      if (ops) ++n_anim;
    }
  }
  if (rep.n_pins_applied)
    rep.unsupported.push_back(std::to_string(rep.n_pins_applied) + " pins are held fixed; the reference's pins are soft constraints confining each vertex to a ball of radius constraint_ghat");
  if (n_anim) rep.unsupported.push_back(std::to_string(n_anim) + " pinned vertices belong to animated blocks; they are held at their t = 0 positions");
  if (n_pull) rep.unsupported.push_back(std::to_string(n_pull) + " vertices are pulled (pull > 0); pulls are not on the device step");
  rep.n_pulls_applied = 0;
  // stitches (golden.rs:839-878: stitch_ind.bin is u64[S][6])
  if (exists(dir + "/input_bin/stitch_ind.bin")) {
    const auto sb = read_bytes(dir + "/input_bin/stitch_ind.bin");
    if (sb.size() % 48) throw CaptureError("stitch_ind.bin is not a whole number of u64[6] records");
    rep.n_stitches = sb.size() / 48;
    if (rep.n_stitches) rep.unsupported.push_back(std::to_string(rep.n_stitches) + " stitches are not modeled; the seams stay unsewn");
  }
  // analytic colliders at t = 0 (gauge_driver.rs:157-180, 1195-1230)
  for (const SceneCollider &c : read_colliders(dir)) {
    float p[3], r;
    c.sample(0.0f, p, r);
    nguo_collider o{};
    if (c.wall) {
      const float n = std::sqrt(c.normal[0] * c.normal[0] + c.normal[1] * c.normal[1] + c.normal[2] * c.normal[2]);
      if (!(n > 1.0e-6f)) continue;
      o.kind = NGUO_COLLIDER_HALFSPACE;
      for (int k = 0; k < 3; ++k) { o.a[k] = p[k]; o.b[k] = c.normal[k] / n; }
    } else {
      if (!(r > 0.0f)) continue;
      o.kind = NGUO_COLLIDER_SPHERE;
      for (int k = 0; k < 3; ++k) o.a[k] = p[k];
      o.radius = r;
    }
    S.colliders.push_back(o);
    ++rep.n_colliders;
    rep.n_colliders_animated += c.animated();
  }
  if (rep.n_colliders_animated)
    rep.unsupported.push_back(std::to_string(rep.n_colliders_animated) + " analytic colliders animate; they are held at their t = 0 keyframe");
  // air (gauge_driver.rs:700-712): the device step has no aerodynamic damper
  if (S.has_toml && (T.air_friction > 0.0f || T.isotropic_air_friction > 0.0f || T.wind[0] != 0.0f || T.wind[1] != 0.0f || T.wind[2] != 0.0f))
    rep.unsupported.push_back("air friction / wind is ACTIVE in this scene but the device step has no aerodynamic damper");
  if (S.has_toml && T.world_scaling != 1.0f) rep.unsupported.push_back("world_scaling != 1 is not applied to the displacement groups");

  // solver parameters (gauge_driver.rs:680-760)
  nguo_params &p = S.params;
  p.dt = S.has_toml ? T.dt : 0.01f;
  for (int k = 0; k < 3; ++k) p.gravity[k] = S.has_toml ? T.gravity[k] : (k == 1 ? -9.8f : 0.0f);
  p.newton_iters = 32;
  p.bend_stiffness = bend_k;
  p.shell_mu = S.shell_faces.empty() ? 0.0f : mu;
  p.shell_lambda = S.shell_faces.empty() ? 0.0f : lam;
  p.shell_model = NGUO_SHELL_ARAP;
  p.eiganalysis_eps = S.has_toml ? T.eiganalysis_eps : 1.0e-6f;
  p.contact_dhat = gap_mean;
  p.contact_kappa = 1.0e2f;  // ImplicitParams::default (implicit.rs:202): the Rayleigh stiffness's fallback
  p.friction = fric;
  p.friction_eps = S.has_toml ? T.friction_eps : 1.0e-5f;
  p.barrier_kind = (S.has_toml && T.cubic_barrier) ? NGUO_BARRIER_CUBIC : NGUO_BARRIER_LOG;
  p.self_collision = !(S.has_toml && T.disable_contact);
  p.max_dx = S.has_toml ? T.max_dx : INFINITY;
  p.target_toi = S.has_toml ? T.target_toi : 0.0f;
  p.min_newton_steps = S.has_toml ? T.min_newton_steps : 0;
  p.ccd_max_iter = S.has_toml ? T.ccd_max_iter : 256;
  p.cg_tol = S.has_toml ? T.cg_tol : 1.0e-4f;
  p.cg_max_iter = S.has_toml ? T.cg_max_iter : 2000;
  p.floor_y = -INFINITY;
  p.precond = NGUO_PRECOND_SCHWARZ;
  p.schwarz_domain = 32;
  p.schwarz_weight = 0.0f;
  p.stretch_k = 0.0f;
  S.bind();
  return S;
}

}  // namespace mguso
