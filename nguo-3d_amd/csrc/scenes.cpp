// Seeded synthetic workloads in the solver's vertex space: the stand-in for the scene JSON
// that engine-api lowers (engine-api/src/scene.rs:89-229). Host code only; no device.
//
// World is y-up (ImplicitParams default gravity (0, -9.8, 0), implicit.rs:181). Cloth
// vertices come first, then rod vertices, then the static body (inv_mass 0, static mask 1).
// Lumped mass: areal density x cloth area / cloth vertex count (scene.rs:130-132), given to
// every free vertex, rod vertices included.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "nguo_hip.h"
#include "nguo_internal.h"

struct nguo_scene_buf {
  std::vector<float> pos, vel, inv_mass, rod_rest, rod_k, rh_rest, rh_stiff;
  std::vector<uint32_t> tris, shell_faces, rods, rod_hinges;
  std::vector<uint8_t> static_mask;
  std::vector<nguo_collider> colliders;
  nguo_params params{};
};

namespace {

using V3d = std::array<double, 3>;

struct Mesh {
  std::vector<V3d> v;
  std::vector<std::array<uint32_t, 3>> f;
};

double signed_volume(const Mesh &m) {
  double vol = 0;
  for (auto &t : m.f) {
    const V3d &a = m.v[t[0]], &b = m.v[t[1]], &c = m.v[t[2]];
    vol += (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0])) / 6.0;
  }
  return vol;
}

void orient_outward(Mesh &m) {
  if (signed_volume(m) < 0)
    for (auto &t : m.f) std::swap(t[1], t[2]);
}

// Icosphere by midpoint subdivision: level 4 gives 2562 vertices / 5120 triangles.
Mesh icosphere(int level, double radius, V3d c) {
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  Mesh m;
  m.v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  m.f = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
         {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
         {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
  auto norm = [](V3d p) {
    double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    return V3d{p[0] / l, p[1] / l, p[2] / l};
  };
  for (auto &p : m.v) p = norm(p);
  for (int lv = 0; lv < level; ++lv) {
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
    auto midpoint = [&](uint32_t a, uint32_t b) {
      auto key = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      V3d p = norm({(m.v[a][0] + m.v[b][0]) * 0.5, (m.v[a][1] + m.v[b][1]) * 0.5, (m.v[a][2] + m.v[b][2]) * 0.5});
      uint32_t id = (uint32_t)m.v.size();
      m.v.push_back(p);
      mid.emplace(key, id);
      return id;
    };
    std::vector<std::array<uint32_t, 3>> nf;
    for (auto &f : m.f) {
      uint32_t a = midpoint(f[0], f[1]), b = midpoint(f[1], f[2]), cc = midpoint(f[2], f[0]);
      nf.push_back({f[0], a, cc});
      nf.push_back({f[1], b, a});
      nf.push_back({f[2], cc, b});
      nf.push_back({a, b, cc});
    }
    m.f.swap(nf);
  }
  for (auto &p : m.v) p = {c[0] + radius * p[0], c[1] + radius * p[1], c[2] + radius * p[2]};
  orient_outward(m);
  return m;
}

// Capsule (sphere-swept segment p0-p1, radius r) as a closed ring mesh: `ns` vertices per
// ring, `h` rings per hemisphere, `cyl` rings on the straight part, two poles.
Mesh capsule(V3d p0, V3d p1, double r, int ns, int h, int cyl) {
  V3d a{p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
  double la = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  a = {a[0] / la, a[1] / la, a[2] / la};
  V3d ref = std::fabs(a[0]) < 0.9 ? V3d{1, 0, 0} : V3d{0, 1, 0};
  V3d u{a[1] * ref[2] - a[2] * ref[1], a[2] * ref[0] - a[0] * ref[2], a[0] * ref[1] - a[1] * ref[0]};
  double lu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  u = {u[0] / lu, u[1] / lu, u[2] / lu};
  V3d w{a[1] * u[2] - a[2] * u[1], a[2] * u[0] - a[0] * u[2], a[0] * u[1] - a[1] * u[0]};
  Mesh m;
  auto point = [&](V3d ctr, double phi, double th) {
    double cp = std::cos(phi), sp = std::sin(phi), ct = std::cos(th), st = std::sin(th);
    return V3d{ctr[0] + r * (cp * (ct * u[0] + st * w[0]) + sp * a[0]), ctr[1] + r * (cp * (ct * u[1] + st * w[1]) + sp * a[1]),
               ctr[2] + r * (cp * (ct * u[2] + st * w[2]) + sp * a[2])};
  };
  const double PI = 3.14159265358979323846;
  std::vector<std::pair<V3d, double>> rings;  // (centre, latitude)
  for (int k = 0; k < h; ++k) rings.push_back({p0, -PI / 2 + PI / 2 * (double)(k + 1) / h});
  for (int c = 1; c <= cyl; ++c) {
    double t = (double)c / (cyl + 1);
    rings.push_back({{p0[0] + t * (p1[0] - p0[0]), p0[1] + t * (p1[1] - p0[1]), p0[2] + t * (p1[2] - p0[2])}, 0.0});
  }
  for (int k = 0; k < h; ++k) rings.push_back({p1, PI / 2 * (double)k / h});
  m.v.push_back(point(p0, -PI / 2, 0));  // bottom pole
  for (auto &rg : rings)
    for (int s = 0; s < ns; ++s) m.v.push_back(point(rg.first, rg.second, 2 * PI * s / ns));
  uint32_t top = (uint32_t)m.v.size();
  m.v.push_back(point(p1, PI / 2, 0));
  uint32_t K = (uint32_t)rings.size();
  auto id = [&](uint32_t ring, uint32_t s) { return 1 + ring * ns + (s % ns); };
  for (int s = 0; s < ns; ++s) m.f.push_back({0u, id(0, s + 1), id(0, s)});
  for (uint32_t k = 0; k + 1 < K; ++k)
    for (int s = 0; s < ns; ++s) {
      m.f.push_back({id(k, s), id(k, s + 1), id(k + 1, s + 1)});
      m.f.push_back({id(k, s), id(k + 1, s + 1), id(k + 1, s)});
    }
  for (int s = 0; s < ns; ++s) m.f.push_back({top, id(K - 1, s), id(K - 1, s + 1)});
  orient_outward(m);
  return m;
}

void append(Mesh &dst, const Mesh &src) {
  uint32_t base = (uint32_t)dst.v.size();
  dst.v.insert(dst.v.end(), src.v.begin(), src.v.end());
  for (auto f : src.f) dst.f.push_back({f[0] + base, f[1] + base, f[2] + base});
}

// Twelve capsules (torso, hips, neck, head, upper arms, forearms, thighs, shins), 48 x 44
// rings each: 12 x 4224 = 50 688 triangles. `s` scales the body in x and z.
Mesh capsule_body(double s) {
  struct C { V3d p0, p1; double r; };
  std::vector<C> caps = {
      {{0, 1.00, 0}, {0, 1.40, 0}, 0.13},       {{-0.15, 0.95, 0}, {0.15, 0.95, 0}, 0.11},
      {{0, 1.45, 0}, {0, 1.55, 0}, 0.05},       {{0, 1.62, 0}, {0, 1.70, 0}, 0.09},
      {{-0.25, 1.38, 0}, {-0.27, 1.10, 0}, 0.045}, {{0.25, 1.38, 0}, {0.27, 1.10, 0}, 0.045},
      {{-0.27, 1.10, 0}, {-0.28, 0.85, 0.03}, 0.04}, {{0.27, 1.10, 0}, {0.28, 0.85, 0.03}, 0.04},
      {{-0.10, 0.90, 0}, {-0.11, 0.50, 0}, 0.08}, {{0.10, 0.90, 0}, {0.11, 0.50, 0}, 0.08},
      {{-0.11, 0.50, 0}, {-0.11, 0.08, 0}, 0.06}, {{0.11, 0.50, 0}, {0.11, 0.08, 0}, 0.06}};
  Mesh body;
  for (auto &c : caps) {
    V3d p0{c.p0[0] * s, c.p0[1], c.p0[2] * s}, p1{c.p1[0] * s, c.p1[1], c.p1[2] * s};
    append(body, capsule(p0, p1, c.r * s, 48, 20, 4));
  }
  return body;
}

// Closed-in-theta tube around the y axis: `na` columns around, `nr` rows from y0 to y0+height.
Mesh tube(double radius, double y0, double height, int na, int nr, std::mt19937_64 *rng, double noise_u, double noise_n, double theta0) {
  Mesh m;
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  std::normal_distribution<double> N(0.0, 1.0);
  const double PI = 3.14159265358979323846;
  for (int j = 0; j < nr; ++j)
    for (int i = 0; i < na; ++i) {
      double th = theta0 + 2 * PI * i / na;
      double r = radius;
      if (noise_u > 0) r += noise_u * U(*rng);
      if (noise_n > 0) r += noise_n * N(*rng);
      m.v.push_back({r * std::cos(th), y0 + height * j / (nr - 1), r * std::sin(th)});
    }
  for (int j = 0; j + 1 < nr; ++j)
    for (int i = 0; i < na; ++i) {
      uint32_t v = (uint32_t)(j * na + i), v1 = (uint32_t)(j * na + (i + 1) % na);
      uint32_t r0 = v + (uint32_t)na, r1 = v1 + (uint32_t)na;
      m.f.push_back({v, v1, r0});
      m.f.push_back({r0, v1, r1});
    }
  return m;
}

double mesh_area(const Mesh &m) {
  double a = 0;
  for (auto &t : m.f) {
    const V3d &p = m.v[t[0]], &q = m.v[t[1]], &r = m.v[t[2]];
    V3d e1{q[0] - p[0], q[1] - p[1], q[2] - p[2]}, e2{r[0] - p[0], r[1] - p[1], r[2] - p[2]};
    V3d c{e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    a += 0.5 * std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  }
  return a;
}

nguo_params default_params() {
  nguo_params p{};
  p.dt = 1.0f / 60.0f;
  p.gravity[0] = 0.0f; p.gravity[1] = -9.8f; p.gravity[2] = 0.0f;
  p.newton_iters = 4;
  p.bend_stiffness = 1.0e-5f;
  // StVK shells with (mu, lambda) = Lame(E * thickness, nu), E = 1e5 Pa, h = 0.5 mm, nu = 0.3
  // (material.rs:31-44 lame_from_young_poisson applied to the membrane modulus E*h).
  float young = 1.0e5f * 5.0e-4f, nu = 0.3f;
  p.shell_mu = young / (2.0f * (1.0f + nu));
  p.shell_lambda = young * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
  p.shell_model = NGUO_SHELL_STVK;
  p.eiganalysis_eps = 1.0e-6f;
  p.contact_dhat = 1.0e-3f;
  p.contact_kappa = 1.0e5f;
  p.friction = 0.3f;
  p.friction_eps = 1.0e-5f;
  p.barrier_kind = NGUO_BARRIER_LOG;
  p.self_collision = 1;
  p.max_dx = INFINITY;
  p.target_toi = 0.0f;
  p.min_newton_steps = 0;
  p.ccd_max_iter = 256;
  p.cg_tol = 1.0e-6f;
  p.cg_max_iter = 10000;
  p.floor_y = -INFINITY;
  p.precond = NGUO_PRECOND_SCHWARZ;
  p.schwarz_domain = 32;
  p.schwarz_weight = 0.0f;
  p.stretch_k = 0.0f;
  return p;
}

// Concatenate cloth (free), rods (free) and body (static) into the buffer.
void finish(nguo_scene_buf &b, const Mesh &cloth, const std::vector<V3d> &rodv, const std::vector<std::vector<uint32_t>> &rod_lines,
            const Mesh &body, double density, double rod_k, double rod_bend) {
  double area = mesh_area(cloth);
  uint32_t ncloth = (uint32_t)cloth.v.size();
  float vmass = (float)std::max(area * density / (double)std::max(ncloth, 1u), 1.0e-6);
  float inv = 1.0f / vmass;
  auto push_v = [&](const V3d &p, float im, uint8_t st) {
    b.pos.push_back((float)p[0]); b.pos.push_back((float)p[1]); b.pos.push_back((float)p[2]);
    b.inv_mass.push_back(im);
    b.static_mask.push_back(st);
  };
  for (auto &p : cloth.v) push_v(p, inv, 0);
  uint32_t rbase = (uint32_t)(b.inv_mass.size());
  for (auto &p : rodv) push_v(p, inv, 0);
  uint32_t bbase = (uint32_t)(b.inv_mass.size());
  for (auto &p : body.v) push_v(p, 0.0f, 1);
  for (auto &t : cloth.f) { b.tris.insert(b.tris.end(), {t[0], t[1], t[2]}); b.shell_faces.insert(b.shell_faces.end(), {t[0], t[1], t[2]}); }
  for (auto &t : body.f) b.tris.insert(b.tris.end(), {t[0] + bbase, t[1] + bbase, t[2] + bbase});
  for (auto &line : rod_lines) {
    for (size_t k = 0; k + 1 < line.size(); ++k) {
      uint32_t a = rbase + line[k], c = rbase + line[k + 1];
      b.rods.push_back(a); b.rods.push_back(c);
      const float *pa = &b.pos[3 * a], *pc = &b.pos[3 * c];
      float dx = pa[0] - pc[0], dy = pa[1] - pc[1], dz = pa[2] - pc[2];
      b.rod_rest.push_back(std::sqrt(dx * dx + dy * dy + dz * dz));
      b.rod_k.push_back((float)rod_k);
    }
    for (size_t k = 1; k + 1 < line.size(); ++k) {
      b.rod_hinges.push_back(rbase + line[k - 1]);
      b.rod_hinges.push_back(rbase + line[k]);
      b.rod_hinges.push_back(rbase + line[k + 1]);
      b.rh_rest.push_back(3.14159265358979323846f);
      b.rh_stiff.push_back((float)rod_bend);
    }
  }
  b.vel.assign(b.pos.size(), 0.0f);
}

// Vertical seam rods just outside a tube of radius `r`: `nseg` segments from y0 to y0+h.
void seam_rods(double r, double y0, double h, int nrods, int nseg, std::vector<V3d> &v, std::vector<std::vector<uint32_t>> &lines) {
  const double PI = 3.14159265358979323846;
  for (int k = 0; k < nrods; ++k) {
    double th = 2 * PI * k / nrods + 0.013;
    std::vector<uint32_t> line;
    for (int s = 0; s <= nseg; ++s) {
      line.push_back((uint32_t)v.size());
      v.push_back({r * std::cos(th), y0 + h * s / nseg, r * std::sin(th)});
    }
    lines.push_back(line);
  }
}

}  // namespace

extern "C" {

int nguo_scene_generate(int config, uint64_t seed, double param, nguo_scene_buf **out) {
  if (!out) return nguo::fail(NGUO_E_INVALID, "nguo_scene_generate: null output");
  auto *b = new nguo_scene_buf();
  b->params = default_params();
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  Mesh none;
  std::vector<V3d> rodv;
  std::vector<std::vector<uint32_t>> lines;
  switch (config) {
    case 0:
    case 100: {
      // 1 m x 1 m sheet (32 x 32, or 10 x 10 for tests) in the x-z plane at y = 0.5 with the
      // height jittered by U(+-1 mm), dropped onto a static r = 0.3 m icosphere at the origin.
      int n = config == 0 ? 32 : 10;
      Mesh sheet;
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) sheet.v.push_back({-0.5 + (double)i / (n - 1), 0.5 + 1.0e-3 * U(rng), -0.5 + (double)j / (n - 1)});
      for (int j = 0; j + 1 < n; ++j)
        for (int i = 0; i + 1 < n; ++i) {
          uint32_t v = (uint32_t)(j * n + i), r = v + (uint32_t)n;
          sheet.f.push_back({v, v + 1, r});
          sheet.f.push_back({r, v + 1, r + 1});
        }
      Mesh sphere = icosphere(config == 0 ? 4 : 2, 0.3, {0, 0, 0});
      finish(*b, sheet, rodv, lines, sphere, 0.2, 0, 0);
      break;
    }
    case 1:
    case 3: {
      // Tube garment r = 0.18 m x 1 m (150 x 133, +-2 mm radial noise) above the hips of a
      // 12-capsule body, three 200-segment seam rods 4 mm outside it. Config 3 is the graded
      // size with circumference factor `param` and a mesh density that scales as param^1.25.
      double f = config == 3 ? param : 1.0;
      double dens = config == 3 ? std::pow(f, 1.25) : 1.0;
      int na = (int)std::lround(150 * f * dens), nr = (int)std::lround(133 * dens);
      Mesh t = tube(0.18 * f, 1.08, 1.0, na, nr, &rng, 2.0e-3, 0.0, 0.0);
      seam_rods(0.18 * f + 4.0e-3, 1.08, 1.0, 3, 200, rodv, lines);
      Mesh body = capsule_body(f);
      finish(*b, t, rodv, lines, body, 0.2, 1.0e3, 1.0e-4);
      break;
    }
    case 2: {
      // Two concentric 170 x 147 tubes; the inner one sits 4.9 mm inside the outer one with a
      // per-vertex radial displacement N(0, 3 mm), so about 5 % of it starts interpenetrating.
      Mesh outer = tube(0.18, 1.08, 1.0, 170, 147, &rng, 0.0, 0.0, 0.0);
      Mesh inner = tube(0.18 - 4.9e-3, 1.08, 1.0, 170, 147, &rng, 0.0, 3.0e-3, 0.5 * 2 * 3.14159265358979323846 / 170);
      Mesh both = outer;
      append(both, inner);
      Mesh body = capsule_body(1.0);
      finish(*b, both, rodv, lines, body, 0.2, 0, 0);
      b->params.contact_dhat = 5.0e-4f;
      b->params.friction = 0.4f;
      break;
    }
    case 101: {
      // Small tube (24 x 20) over the hip capsule only, with one 24-segment rod.
      Mesh t = tube(0.18, 1.08, 0.3, 24, 20, &rng, 2.0e-3, 0.0, 0.0);
      seam_rods(0.184, 1.08, 0.3, 1, 24, rodv, lines);
      Mesh body = capsule({-0.15, 0.95, 0}, {0.15, 0.95, 0}, 0.11, 16, 6, 2);
      finish(*b, t, rodv, lines, body, 0.2, 1.0e3, 1.0e-4);
      break;
    }
    default:
      delete b;
      return nguo::fail(NGUO_E_INVALID, "nguo_scene_generate: unknown config " + std::to_string(config));
  }
  *out = b;
  return NGUO_OK;
}

int nguo_scene_view(const nguo_scene_buf *b, nguo_scene_desc *d) {
  if (!b || !d) return nguo::fail(NGUO_E_INVALID, "nguo_scene_view: null argument");
  std::memset(d, 0, sizeof(*d));
  d->n_verts = (uint32_t)b->inv_mass.size();
  d->positions = b->pos.data();
  d->velocities = b->vel.data();
  d->inv_mass = b->inv_mass.data();
  d->n_tris = (uint32_t)(b->tris.size() / 3);
  d->tris = b->tris.data();
  d->static_collider = b->static_mask.data();
  d->n_shell_faces = (uint32_t)(b->shell_faces.size() / 3);
  d->shell_faces = b->shell_faces.data();
  d->n_rods = (uint32_t)(b->rods.size() / 2);
  d->rod_verts = b->rods.data();
  d->rod_rest = b->rod_rest.data();
  d->rod_k = b->rod_k.data();
  d->n_rod_hinges = (uint32_t)(b->rod_hinges.size() / 3);
  d->rod_hinge_verts = b->rod_hinges.data();
  d->rod_hinge_rest = b->rh_rest.data();
  d->rod_hinge_stiff = b->rh_stiff.data();
  d->n_colliders = (uint32_t)b->colliders.size();
  d->colliders = b->colliders.data();
  d->params = b->params;
  return NGUO_OK;
}

void nguo_scene_free(nguo_scene_buf *b) { delete b; }

}  // extern "C"
