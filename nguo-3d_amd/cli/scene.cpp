// Scene JSON -> nguo_scene_desc, following engine-api/src/scene.rs:89-229 (lower and
// build_implicit), params.rs:41-122 (SolverConfig) and material.rs:9-122 (Material, presets).
#include "scene.h"

#include <cerrno>
#include <cmath>
#include <limits>

namespace mguso {
namespace {

const char *kind_name(Json::Kind k) {
  switch (k) {
    case Json::Null: return "null";
    case Json::Bool: return "a boolean";
    case Json::Num: return "a number";
    case Json::Str: return "a string";
    case Json::Arr: return "an array";
    default: return "an object";
  }
}

[[noreturn]] void bad(const std::string &field, const std::string &what) { throw BadScene("`" + field + "`: " + what); }

void want(const Json &v, Json::Kind k, const std::string &field) {
  if (v.kind != k) bad(field, std::string("expected ") + kind_name(k) + ", found " + kind_name(v.kind));
}

// serde_json reads an f32 as the nearest f64 cast to f32.
float f32_of(const Json &v, const std::string &field) {
  want(v, Json::Num, field);
  return (float)v.num;
}
double f64_of(const Json &v, const std::string &field) {
  want(v, Json::Num, field);
  return v.num;
}
// Integers must be integer literals (serde_json rejects 1.0 or 1e2 for an integer field).
int64_t int_of(const Json &v, const std::string &field, int64_t lo, int64_t hi) {
  want(v, Json::Num, field);
  if (v.text.find_first_of(".eE") != std::string::npos) bad(field, "expected an integer, found " + v.text);
  errno = 0;
  long long x = std::strtoll(v.text.c_str(), nullptr, 10);
  if (errno == ERANGE || x < lo || x > hi) bad(field, "integer " + v.text + " out of range");
  return (int64_t)x;
}
uint32_t u32_of(const Json &v, const std::string &field) { return (uint32_t)int_of(v, field, 0, 0xffffffffLL); }
bool bool_of(const Json &v, const std::string &field) {
  want(v, Json::Bool, field);
  return v.b;
}
void vec_of(const Json &v, const std::string &field, float *out, size_t n) {
  want(v, Json::Arr, field);
  if (v.arr.size() != n) bad(field, "expected " + std::to_string(n) + " elements, found " + std::to_string(v.arr.size()));
  for (size_t k = 0; k < n; ++k) out[k] = f32_of(v.arr[k], field);
}
std::string enum_of(const Json &v, const std::string &field, std::initializer_list<const char *> names) {
  want(v, Json::Str, field);
  for (const char *n : names)
    if (v.text == n) return v.text;
  bad(field, "unknown variant `" + v.text + "`");
}

// params.rs:41-122: every field optional, kebab-case names.
SolverConfig read_config(const Json *c) {
  SolverConfig s;
  if (!c) return s;
  want(*c, Json::Obj, "config");
  float tmp3[3];
  for (auto &kv : c->obj) {
    const std::string &k = kv.first, f = "config." + k;
    const Json &v = kv.second;
    if (k == "dt") s.dt = f32_of(v, f);
    else if (k == "fps") s.fps = f64_of(v, f);
    else if (k == "frames") s.frames = (int32_t)int_of(v, f, INT32_MIN, INT32_MAX);
    else if (k == "gravity") vec_of(v, f, s.gravity, 3);
    else if (k == "wind") vec_of(v, f, tmp3, 3);
    else if (k == "air-friction" || k == "air-density" || k == "target-toi" || k == "max-dx" || k == "ccd-reduction" ||
             k == "line-search-max-t" || k == "damping")
      (void)f32_of(v, f);
    else if (k == "min-newton-steps" || k == "ccd-max-iter" || k == "substeps") (void)u32_of(v, f);
    else if (k == "cg-max-iter") s.cg_max_iter = u32_of(v, f);
    else if (k == "cg-tol") s.cg_tol = f32_of(v, f);
    else if (k == "disable-contact" || k == "preview-aero") (void)bool_of(v, f);
    else if (k == "barrier") (void)enum_of(v, f, {"cubic", "quad", "log"});
    else if (k == "friction-eps") s.friction_eps = f32_of(v, f);
    else if (k == "iterations") s.iterations = u32_of(v, f);
    else if (k == "self-collision") s.self_collision = bool_of(v, f);
    else if (k == "floor-y") {
      s.has_floor = v.kind != Json::Null;
      if (s.has_floor) s.floor_y = f32_of(v, f);
    } else if (k == "bend-compliance") {
      if (v.kind != Json::Null) (void)f32_of(v, f);
    }
    // unknown fields are ignored, as serde does without deny_unknown_fields
  }
  return s;
}

// material.rs:9-73: every field optional, kebab-case names.
Material read_material(const Json *m, const std::string &where) {
  Material out;
  if (!m) return out;
  want(*m, Json::Obj, where);
  struct F { const char *name; float Material::*p; };
  static const F fields[] = {
      {"young-mod", &Material::young_mod}, {"poisson", &Material::poisson}, {"density", &Material::density},
      {"bend", &Material::bend}, {"strain-limit", &Material::strain_limit}, {"strain-limit-weft", &Material::strain_limit_weft},
      {"contact-gap", &Material::contact_gap}, {"contact-dynamic", &Material::contact_dynamic},
      {"contact-offset", &Material::contact_offset}, {"friction", &Material::friction},
      {"membrane-warp", &Material::membrane_warp}, {"membrane-weft", &Material::membrane_weft},
      {"membrane-shear", &Material::membrane_shear}};
  for (auto &kv : m->obj) {
    if (kv.first == "model") {
      (void)enum_of(kv.second, where + ".model", {"arap", "stvk", "baraff-witkin", "snhk", "pdrd"});
      continue;
    }
    for (auto &fd : fields)
      if (kv.first == fd.name) out.*fd.p = f32_of(kv.second, where + "." + kv.first);
  }
  return out;
}

struct Object {
  std::vector<float> pos;
  std::vector<uint32_t> tris, pins;
  Material material;
  bool has_preset = false;
  std::string preset;
  bool is_static = false;
};

// scene.rs:37-71 Object: name, positions and tris are required; the rest default. The static
// flag is the field `is_static` (the struct renames nothing).
Object read_object(const Json &o, size_t index) {
  const std::string where = "objects[" + std::to_string(index) + "]";
  want(o, Json::Obj, where);
  Object out;
  const Json *name = o.get("name"), *pos = o.get("positions"), *tris = o.get("tris");
  if (!name) bad(where, "missing field `name`");
  want(*name, Json::Str, where + ".name");
  if (!pos) bad(where, "missing field `positions`");
  if (!tris) bad(where, "missing field `tris`");
  want(*pos, Json::Arr, where + ".positions");
  out.pos.resize(3 * pos->arr.size());
  for (size_t i = 0; i < pos->arr.size(); ++i) vec_of(pos->arr[i], where + ".positions", &out.pos[3 * i], 3);
  want(*tris, Json::Arr, where + ".tris");
  for (auto &t : tris->arr) {
    want(t, Json::Arr, where + ".tris");
    if (t.arr.size() != 3) bad(where + ".tris", "expected 3 elements, found " + std::to_string(t.arr.size()));
    for (auto &i : t.arr) out.tris.push_back(u32_of(i, where + ".tris"));
  }
  out.material = read_material(o.get("material"), where + ".material");
  if (const Json *p = o.get("preset"); p && p->kind != Json::Null) {
    out.preset = enum_of(*p, where + ".preset", {"cotton", "denim", "silk", "jersey", "wool", "leather"});
    out.has_preset = true;
  }
  if (const Json *p = o.get("pins")) {
    want(*p, Json::Arr, where + ".pins");
    for (auto &i : p->arr) out.pins.push_back(u32_of(i, where + ".pins"));
  }
  if (const Json *s = o.get("is_static")) out.is_static = bool_of(*s, where + ".is_static");
  return out;
}

}  // namespace

void LoweredScene::bind() {
  desc = nguo_scene_desc{};
  desc.n_verts = n_verts();
  desc.positions = positions.data();
  desc.velocities = velocities.data();
  desc.inv_mass = inv_mass.data();
  desc.n_tris = (uint32_t)(tris.size() / 3);
  desc.tris = tris.data();
  desc.static_collider = static_mask.empty() ? nullptr : static_mask.data();
  desc.n_colliders = (uint32_t)colliders.size();
  desc.colliders = colliders.empty() ? nullptr : colliders.data();
  desc.params = params;
}

LoweredScene lower_scene_json(const std::string &text, int precond) {
  Json root;
  try {
    root = JsonParser(text).parse();
  } catch (const std::runtime_error &e) {
    throw BadScene(e.what());
  }
  want(root, Json::Obj, "scene");
  LoweredScene L;
  L.config = read_config(root.get("config"));
  const Json *objs = root.get("objects");
  if (!objs) bad("scene", "missing field `objects`");
  want(*objs, Json::Arr, "objects");
  std::vector<Object> objects;
  for (size_t i = 0; i < objs->arr.size(); ++i) objects.push_back(read_object(objs->arr[i], i));
  // scene.rs:15-35 ColliderSpec, tagged by "type"
  if (const Json *cs = root.get("colliders")) {
    want(*cs, Json::Arr, "colliders");
    for (size_t i = 0; i < cs->arr.size(); ++i) {
      const Json &c = cs->arr[i];
      const std::string where = "colliders[" + std::to_string(i) + "]";
      want(c, Json::Obj, where);
      const Json *type = c.get("type");
      if (!type) bad(where, "missing field `type`");
      std::string t = enum_of(*type, where + ".type", {"plane", "sphere"});
      nguo_collider k{};
      if (t == "plane") {
        const Json *p = c.get("point"), *n = c.get("normal");
        if (!p) bad(where, "missing field `point`");
        if (!n) bad(where, "missing field `normal`");
        k.kind = NGUO_COLLIDER_HALFSPACE;
        vec_of(*p, where + ".point", k.a, 3);
        vec_of(*n, where + ".normal", k.b, 3);  // normalize_or_zero happens at step creation
      } else {
        const Json *ctr = c.get("center"), *r = c.get("radius");
        if (!ctr) bad(where, "missing field `center`");
        if (!r) bad(where, "missing field `radius`");
        k.kind = NGUO_COLLIDER_SPHERE;
        vec_of(*ctr, where + ".center", k.a, 3);
        k.radius = f32_of(*r, where + ".radius");
      }
      L.colliders.push_back(k);
    }
  }

  // scene.rs:89-152 lower: one vertex space, static objects pinned, the last dynamic
  // object's material (its preset overriding it) for the whole scene
  std::vector<uint8_t> pin_flags, static_obj;
  Material material;
  std::string preset;
  bool any_static = false;
  for (auto &o : objects) {
    const uint32_t base = (uint32_t)(L.positions.size() / 3);
    L.positions.insert(L.positions.end(), o.pos.begin(), o.pos.end());
    pin_flags.insert(pin_flags.end(), o.pos.size() / 3, o.is_static ? 1 : 0);
    static_obj.insert(static_obj.end(), o.pos.size() / 3, o.is_static ? 1 : 0);
    any_static |= o.is_static && !o.pos.empty();
    for (uint32_t t : o.tris) L.tris.push_back(t + base);
    if (!o.is_static) {
      material = o.material;
      preset = o.has_preset ? o.preset : std::string();
    }
    for (uint32_t p : o.pins) {
      const uint64_t idx = (uint64_t)base + p;
      if (idx < pin_flags.size()) pin_flags[idx] = 1;
    }
  }
  const uint32_t V = (uint32_t)(L.positions.size() / 3);
  for (uint32_t t : L.tris)
    if (t >= V) throw BadScene("triangle index " + std::to_string(t) + " out of range (" + std::to_string(V) + " vertices)");

  // what the implicit path would run and this build does not
  if (!preset.empty())
    throw Unsupported("fabric preset `" + preset + "` selects the Baraff-Witkin membrane (material.rs:90-125), which this build does not simulate");
  if (material.membrane_warp > 0.0f || material.membrane_weft > 0.0f || material.membrane_shear > 0.0f)
    throw Unsupported("membrane stiffness selects the Baraff-Witkin membrane (implicit.rs:2532-2598), which this build does not simulate");
  if (material.strain_limit > 0.0f || material.strain_limit_weft > 0.0f)
    throw Unsupported("strain limiting (implicit.rs:2508-2530) is not simulated by this build");
  if (material.contact_dynamic != 0.0f && !L.colliders.empty())
    throw Unsupported("contact-dynamic != 0 scales the analytic-collider stiffness (contact.rs dynamic_stiffness), which this build does not simulate");

  // scene.rs:261-271 total_area (FP32, triangle order) and the uniform lumped mass
  float area = 0.0f;
  for (size_t f = 0; f + 2 < L.tris.size(); f += 3) {
    const float *a = &L.positions[3 * (size_t)L.tris[f]], *b = &L.positions[3 * (size_t)L.tris[f + 1]], *c = &L.positions[3 * (size_t)L.tris[f + 2]];
    const float u0 = b[0] - a[0], u1 = b[1] - a[1], u2 = b[2] - a[2];
    const float w0 = c[0] - a[0], w1 = c[1] - a[1], w2 = c[2] - a[2];
    const float x = u1 * w2 - u2 * w1, y = u2 * w0 - u0 * w2, z = u0 * w1 - u1 * w0;
    area += 0.5f * std::sqrt(x * x + y * y + z * z);
  }
  const float n = (float)std::max<uint32_t>(V, 1u);
  const float vertex_mass = std::max(area * material.density / n, 1.0e-6f);
  const float inv = vertex_mass > 0.0f ? 1.0f / vertex_mass : 0.0f;
  L.inv_mass.assign(V, inv);
  for (uint32_t i = 0; i < V; ++i)
    if (pin_flags[i]) L.inv_mass[i] = 0.0f;
  L.velocities.assign(3 * (size_t)V, 0.0f);
  if (any_static) L.static_mask = static_obj;

  // scene.rs:165-229 build_implicit: ImplicitParams from the config and material, the rest at
  // ImplicitParams::default() (implicit.rs:158-229)
  nguo_params &p = L.params;
  const SolverConfig &c = L.config;
  p.dt = c.dt;
  for (int k = 0; k < 3; ++k) p.gravity[k] = c.gravity[k];
  p.newton_iters = std::max(c.iterations, 1u);
  p.stretch_k = std::max(material.young_mod, 1.0f);
  p.bend_stiffness = material.bend;
  p.shell_mu = 0.0f;
  p.shell_lambda = 0.0f;
  p.shell_model = NGUO_SHELL_ARAP;
  p.eiganalysis_eps = 1.0e-6f;
  p.contact_dhat = std::max(material.contact_gap, 1.0e-3f);
  p.contact_kappa = std::max(material.young_mod, 1.0f);
  p.friction = material.friction;
  p.friction_eps = c.friction_eps;
  p.barrier_kind = NGUO_BARRIER_LOG;
  p.self_collision = c.self_collision ? 1 : 0;
  p.max_dx = std::numeric_limits<float>::infinity();
  p.target_toi = 0.0f;
  p.min_newton_steps = 0;
  p.ccd_max_iter = 256;
  p.cg_tol = c.cg_tol;
  p.cg_max_iter = c.cg_max_iter;
  p.floor_y = c.has_floor ? c.floor_y : -std::numeric_limits<float>::infinity();
  // the CLI's preconditioners (main.rs:729-776): block Jacobi, or schwarz-ml with 16-vertex
  // domains and coarse weight 0.7
  p.precond = precond;
  p.schwarz_domain = precond == NGUO_PRECOND_SCHWARZ ? 16 : 32;
  p.schwarz_weight = precond == NGUO_PRECOND_SCHWARZ ? 0.7f : 0.0f;
  L.bind();
  return L;
}

}  // namespace mguso
