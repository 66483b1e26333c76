// mguso: the headless runner of the reference (engine-cli/src/main.rs:92-187, 617-899) over the
// HIP step library, keeping its process contract (core-engine/docs/PROCESS_CONTRACT.md:10-27):
// one pretty JSON object on stdout per command, logs on stderr, exit 0 on success, 1 on
// failure (I/O, malformed or unsupported input, divergence, no device), 2 when `verify` ran
// but a solver invariant failed (and, as clap does, on a usage error).
//
//   mguso [--log-format text|json] [-v] gpu-info
//   mguso run <scene.json> [--frames N] [--out DIR] [--solver xpbd|implicit] [--cpu] [--precond block|schwarz-ml]
//   mguso bench <scene.json> [--steps N] [--solver implicit] [--cpu] [--precond block|schwarz-ml]
//   mguso verify <scene.json> [--steps N]
//   mguso gauge-scene <scene-dir> [--steps N] [--compare-frame N]
//
// The implicit solver runs on the GPU (its linear solve included). The XPBD preview tier, the
// CPU linear solvers (--cpu) and the capture-fitting command (golden) are not part of this
// build and exit 1 saying so. gauge-scene prints its report as text lines, as the reference's
// does (main.rs:264-450); loading and reporting need no device.
#include <sys/stat.h>

#include <cctype>
#include <cerrno>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/nguo_hip.h"
#include "capture.h"
#include "json.h"
#include "scene.h"

namespace mguso {
namespace {

const char *VERSION = "0.1.0";

// ---------------------------------------------------------------- logging (stderr only)
enum Level { ERROR = 0, WARN = 1, INFO = 2, DEBUG = 3 };
struct Log {
  bool json = false;
  int level = INFO;
} g_log;

std::string now_rfc3339() {
  auto t = std::chrono::system_clock::now();
  std::time_t s = std::chrono::system_clock::to_time_t(t);
  auto us = std::chrono::duration_cast<std::chrono::microseconds>(t.time_since_epoch()).count() % 1000000;
  std::tm tm{};
  gmtime_r(&s, &tm);
  char buf[64];
  std::snprintf(buf, sizeof buf, "%04d-%02d-%02dT%02d:%02d:%02d.%06lldZ", tm.tm_year + 1900, tm.tm_mon + 1, tm.tm_mday, tm.tm_hour,
                tm.tm_min, tm.tm_sec, (long long)us);
  return buf;
}

void log(int level, const std::string &msg, const std::vector<std::pair<std::string, std::string>> &fields = {}) {
  if (level > g_log.level) return;
  static const char *names[] = {"ERROR", "WARN", "INFO", "DEBUG"};
  std::string line;
  if (g_log.json) {
    line = "{\"timestamp\":" + quote(now_rfc3339()) + ",\"level\":" + quote(names[level]) + ",\"fields\":{\"message\":" + quote(msg);
    for (auto &f : fields) line += "," + quote(f.first) + ":" + quote(f.second);
    line += "},\"target\":\"mguso\"}";
  } else {
    line = now_rfc3339() + " " + std::string(level == INFO || level == WARN ? " " : "") + names[level] + " mguso: " + msg;
    for (auto &f : fields) line += " " + f.first + "=" + f.second;
  }
  std::fprintf(stderr, "%s\n", line.c_str());
}

// ---------------------------------------------------------------- errors
struct Failure : std::runtime_error {  // exit 1
  using std::runtime_error::runtime_error;
};
struct Usage : std::runtime_error {  // exit 2
  using std::runtime_error::runtime_error;
};

void check(int rc, const std::string &what) {
  if (rc != NGUO_OK) throw Failure(what + ": " + nguo_last_error());
}

// ---------------------------------------------------------------- arguments
struct Args {
  std::string command, scene, out;
  bool has_steps = false, has_frames = false, has_precond = false, cpu = false, has_compare = false;
  uint32_t steps = 200, frames = 0, compare_frame = 0;
  std::string solver, precond = "block";
};

const char *USAGE =
    "Usage: mguso [--log-format <text|json>] [-v] <COMMAND>\n\n"
    "Commands:\n"
    "  gpu-info  Print the device and run a no-op kernel probe\n"
    "  run       Simulate a scene (JSON scene file)\n"
    "  bench     Benchmark solver throughput (steps/sec) on a scene\n"
    "  verify    Run a scene with the implicit solver and check the solver invariants\n"
    "  gauge-scene  Inspect a scene from the reference capture; optionally step it or compare a frame\n\n"
    "Options:\n"
    "  --log-format <text|json>  [default: text]\n"
    "  -v, --verbose             more log output\n"
    "  -h, --help                Print help\n"
    "  -V, --version             Print version\n";

uint32_t parse_u32(const std::string &flag, const std::string &v) {
  if (v.empty() || v.find_first_not_of("0123456789") != std::string::npos)
    throw Usage("invalid value '" + v + "' for '" + flag + " <N>': invalid digit found in string");
  unsigned long long x = std::strtoull(v.c_str(), nullptr, 10);
  if (x > 0xffffffffULL) throw Usage("invalid value '" + v + "' for '" + flag + " <N>': number too large to fit in target type");
  return (uint32_t)x;
}

Args parse_args(int argc, char **argv, bool &done) {
  Args a;
  done = false;
  std::vector<std::string> rest;
  for (int i = 1; i < argc; ++i) rest.emplace_back(argv[i]);
  auto value = [&](size_t &i, const std::string &flag, std::string inline_v, bool has_inline) -> std::string {
    if (has_inline) return inline_v;
    if (i + 1 >= rest.size()) throw Usage("a value is required for '" + flag + "' but none was supplied");
    return rest[++i];
  };
  for (size_t i = 0; i < rest.size(); ++i) {
    std::string arg = rest[i], inline_v;
    bool has_inline = false;
    if (arg.rfind("--", 0) == 0) {
      size_t eq = arg.find('=');
      if (eq != std::string::npos) { inline_v = arg.substr(eq + 1); arg = arg.substr(0, eq); has_inline = true; }
    }
    if (arg == "-h" || arg == "--help") { std::fputs(USAGE, stdout); done = true; return a; }
    if (arg == "-V" || arg == "--version") { std::printf("mguso %s\n", VERSION); done = true; return a; }
    if (arg == "--log-format") {
      std::string v = value(i, arg, inline_v, has_inline);
      if (v != "text" && v != "json") throw Usage("invalid value '" + v + "' for '--log-format <LOG_FORMAT>'");
      g_log.json = v == "json";
    } else if (arg.size() > 1 && arg[0] == '-' && arg[1] != '-' && arg.find_first_not_of('v', 1) == std::string::npos) {
      g_log.level = std::min<int>(DEBUG, g_log.level + (int)(arg.size() - 1));
    } else if (arg == "--verbose") {
      g_log.level = std::min<int>(DEBUG, g_log.level + 1);
    } else if (arg == "--steps" || arg == "--frames") {
      uint32_t n = parse_u32(arg, value(i, arg, inline_v, has_inline));
      if (arg == "--steps") { a.steps = n; a.has_steps = true; }
      else { a.frames = n; a.has_frames = true; }
    } else if (arg == "--compare-frame") {
      a.compare_frame = parse_u32(arg, value(i, arg, inline_v, has_inline));
      a.has_compare = true;
    } else if (arg == "--out") {
      a.out = value(i, arg, inline_v, has_inline);
    } else if (arg == "--solver") {
      a.solver = value(i, arg, inline_v, has_inline);
      if (a.solver != "xpbd" && a.solver != "implicit") throw Usage("invalid value '" + a.solver + "' for '--solver <SOLVER>'");
    } else if (arg == "--precond") {
      a.precond = value(i, arg, inline_v, has_inline);
      a.has_precond = true;
      if (a.precond != "block" && a.precond != "schwarz-ml") throw Usage("invalid value '" + a.precond + "' for '--precond <PRECOND>'");
    } else if (arg == "--cpu") {
      a.cpu = true;
    } else if (!arg.empty() && arg[0] == '-') {
      throw Usage("unexpected argument '" + rest[i] + "' found");
    } else if (a.command.empty()) {
      a.command = arg;
    } else if (a.scene.empty()) {
      a.scene = arg;
    } else {
      throw Usage("unexpected argument '" + arg + "' found");
    }
  }
  static const char *known[] = {"gpu-info", "run", "bench", "verify", "golden", "gauge-scene"};
  if (a.command.empty()) throw Usage("'mguso' requires a subcommand but one was not provided");
  bool ok = false;
  for (const char *k : known) ok |= a.command == k;
  if (!ok) throw Usage("unrecognized subcommand '" + a.command + "'");
  if (a.command != "gpu-info" && a.scene.empty()) throw Usage("the following required arguments were not provided: <SCENE>");
  if (a.command == "gpu-info" && !a.scene.empty()) throw Usage("unexpected argument '" + a.scene + "' found");
  // per-command flags (main.rs:92-187)
  const bool is_run = a.command == "run", is_bench = a.command == "bench", is_verify = a.command == "verify";
  const bool is_gauge = a.command == "gauge-scene";
  if (a.has_steps && !(is_bench || is_verify || is_gauge)) throw Usage("unexpected argument '--steps' found");
  if (a.has_compare && !is_gauge) throw Usage("unexpected argument '--compare-frame' found");
  if (is_gauge && !a.has_steps) a.steps = 0;  // main.rs:103-105: 0 = load and report only
  if (a.has_frames && !is_run) throw Usage("unexpected argument '--frames' found");
  if (!a.out.empty() && !is_run) throw Usage("unexpected argument '--out' found");
  if ((!a.solver.empty() || a.cpu || a.has_precond) && !(is_run || is_bench))
    throw Usage("unexpected argument for '" + a.command + "'");
  if (a.solver.empty()) a.solver = is_run ? "xpbd" : "implicit";  // run defaults to the XPBD tier
  return a;
}

// ---------------------------------------------------------------- the device session
std::string read_file(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Failure("reading scene " + path + ": " + std::strerror(errno));
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

LoweredScene load_scene(const std::string &path, int precond) {
  const size_t dot = path.find_last_of('.');
  std::string ext = dot == std::string::npos ? "" : path.substr(dot + 1);
  for (auto &c : ext) c = (char)std::tolower((unsigned char)c);
  if (ext != "json") {
    if (ext == "obj" || ext == "stl" || ext == "gltf" || ext == "glb" || ext == "usda" || ext == "dae" || ext == "fbx")
      throw Failure("unsupported input " + path + ": a bare mesh is simulated pinned at its first two vertices (main.rs:854-866), and pinned cloth vertices are not part of this build; pass a scene .json");
    throw Failure("unsupported scene file extension: " + path);
  }
  std::string text = read_file(path);
  try {
    return lower_scene_json(text, precond);
  } catch (const BadScene &e) {
    throw Failure(std::string("parsing scene JSON: ") + e.what());
  } catch (const Unsupported &e) {
    throw Failure(std::string("unsupported input: ") + e.what());
  }
}

struct Session {
  nguo_ctx *ctx = nullptr;
  nguo_step *st = nullptr;
  std::string device;
  ~Session() {
    if (st) nguo_step_destroy(st);
    if (ctx) nguo_ctx_destroy(ctx);
  }
  void open_ctx() {
    check(nguo_ctx_create(0, &ctx), "GPU bring-up");
    char name[256] = {0};
    int cus = 0;
    check(nguo_ctx_info(ctx, name, sizeof name, &cus), "GPU bring-up");
    device = name;
  }
  void create(const LoweredScene &L) { check(nguo_step_create(ctx, &L.desc, &st), "building the implicit solver"); }
  void advance(uint32_t n) { check(nguo_step_advance(st, n), "implicit step"); }
  nguo_step_metrics metrics() {
    nguo_step_metrics m{};
    check(nguo_step_get_metrics(st, &m), "reading step metrics");
    return m;
  }
  std::vector<float> positions(uint32_t V) {
    std::vector<float> x(3 * (size_t)V);
    check(nguo_step_get_positions(st, x.data()), "reading positions");
    return x;
  }
  std::vector<float> velocities(uint32_t V) {
    std::vector<float> v(3 * (size_t)V);
    check(nguo_step_get_velocities(st, v.data()), "reading velocities");
    return v;
  }
};

void require_implicit_gpu(const Args &a) {
  if (a.solver == "xpbd")
    throw Failure("unsupported solver tier `xpbd`: the XPBD preview tier is not part of this build; pass --solver implicit");
  if (a.cpu) throw Failure("unsupported option --cpu: this build's linear solve runs on the GPU only");
}

int precond_of(const Args &a) { return a.precond == "schwarz-ml" ? NGUO_PRECOND_SCHWARZ : NGUO_PRECOND_BLOCK; }

std::string backend_of(const Args &a, const Session &s) {
  return std::string("implicit/gpu-pcg-") + (a.precond == "schwarz-ml" ? "schwarz-ml" : "block-jacobi") + "(hip:" + s.device + ")";
}

// (1 / fps) / dt rounded, at least 1 (main.rs:620-623, 804-806).
uint32_t steps_per_frame(const SolverConfig &c) {
  double spf = std::round((1.0 / c.fps) / (double)c.dt);
  if (!(spf >= 1.0)) spf = 1.0;
  return spf > 4294967295.0 ? 0xffffffffu : (uint32_t)spf;
}

// ---------------------------------------------------------------- commands
int cmd_gpu_info() {
  Session s;
  s.open_ctx();
  // the no-op probe: y = 2 I x through the SpMV kernel, [1, 2, 3] -> [2, 4, 6] (main.rs:778-790
  // runs a SAXPY 2x + 0 for the same purpose)
  nguo_ls *ls = nullptr;
  check(nguo_ls_create(s.ctx, NGUO_PRECOND_BLOCK, 1, 0.0f, &ls), "GPU probe");
  const uint32_t rp[4] = {0, 1, 2, 3}, ci[3] = {0, 1, 2};
  const float va[3] = {2.0f, 2.0f, 2.0f}, x[3] = {1.0f, 2.0f, 3.0f};
  float y[3] = {0, 0, 0};
  int rc = nguo_ls_spmv(ls, 3, rp, ci, va, x, y);
  nguo_ls_destroy(ls);
  check(rc, "GPU probe");
  PrettyObject o;
  o.put("backend", quote("hip"));
  o.put("adapter", quote(s.device));
  o.put("device_type", quote("DiscreteGpu"));
  o.put("noop_saxpy_ok", (y[0] == 2.0f && y[1] == 4.0f && y[2] == 6.0f) ? "true" : "false");
  std::printf("%s\n", o.render().c_str());
  return 0;
}

int cmd_bench(const Args &a) {
  require_implicit_gpu(a);
  LoweredScene L = load_scene(a.scene, precond_of(a));
  Session s;
  s.open_ctx();
  s.create(L);
  const std::string backend = backend_of(a, s);
  const uint32_t warmup = std::min<uint32_t>(10, a.steps);  // main.rs:661-664
  for (uint32_t i = 0; i < warmup; ++i) s.advance(1);
  auto t0 = std::chrono::steady_clock::now();
  for (uint32_t i = 0; i < a.steps; ++i) s.advance(1);
  const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  PrettyObject o;
  o.put("backend", quote(backend));
  o.put("verts", std::to_string(L.n_verts()));
  o.put("steps", std::to_string(a.steps));
  o.put("elapsed_s", num_f64(elapsed));
  o.put("steps_per_sec", num_f64(a.steps / elapsed));
  o.put("ms_per_step", num_f64(1000.0 * elapsed / a.steps));
  std::printf("%s\n", o.render().c_str());
  return 0;
}

// engine-io obj.rs:322-347 write_obj: a comment line, then v and f records (1-based).
void write_frame(const std::string &dir, uint32_t frame, const std::vector<float> &x, const std::vector<uint32_t> &tris) {
  char name[32];
  std::snprintf(name, sizeof name, "frame_%04u", frame);
  const std::string path = dir + "/" + name + ".obj";
  std::string s = std::string("# ppf-mguso-solver OBJ export: ") + name + "\n";
  s.reserve(x.size() * 10 + tris.size() * 8);
  for (size_t i = 0; i + 2 < x.size(); i += 3) {
    char buf[64];
    s += "v ";
    auto r = std::to_chars(buf, buf + sizeof buf, x[i]);
    s.append(buf, r.ptr);
    s += ' ';
    r = std::to_chars(buf, buf + sizeof buf, x[i + 1]);
    s.append(buf, r.ptr);
    s += ' ';
    r = std::to_chars(buf, buf + sizeof buf, x[i + 2]);
    s.append(buf, r.ptr);
    s += '\n';
  }
  for (size_t f = 0; f + 2 < tris.size(); f += 3)
    s += "f " + std::to_string(tris[f] + 1) + " " + std::to_string(tris[f + 1] + 1) + " " + std::to_string(tris[f + 2] + 1) + "\n";
  std::ofstream o(path, std::ios::binary);
  if (!o || !o.write(s.data(), (std::streamsize)s.size())) throw Failure("writing OBJ " + path);
}

bool all_finite(const std::vector<float> &x) {
  for (float v : x)
    if (!std::isfinite(v)) return false;
  return true;
}

int cmd_run(const Args &a) {
  require_implicit_gpu(a);
  LoweredScene L = load_scene(a.scene, precond_of(a));
  const uint32_t total_frames = a.has_frames ? a.frames : (uint32_t)std::max<int32_t>(L.config.frames, 0);
  const uint32_t spf = steps_per_frame(L.config);
  if (!a.out.empty()) {
    // create_dir_all
    std::string p;
    std::stringstream ss(a.out);
    std::string part;
    if (!a.out.empty() && a.out[0] == '/') p = "/";
    while (std::getline(ss, part, '/')) {
      if (part.empty()) continue;
      p += part;
      if (mkdir(p.c_str(), 0755) != 0 && errno != EEXIST) throw Failure("mkdir " + a.out + ": " + std::strerror(errno));
      p += "/";
    }
  }
  Session s;
  s.open_ctx();
  s.create(L);
  const std::string backend = backend_of(a, s);
  log(INFO, "simulation start", {{"total_frames", std::to_string(total_frames)}, {"steps_per_frame", std::to_string(spf)}, {"backend", backend}});
  const uint32_t V = L.n_verts();
  double time = 0.0;
  std::vector<float> x = L.positions;
  for (uint32_t frame = 0; frame < total_frames; ++frame) {
    s.advance(spf);
    time = s.metrics().time;
    x = s.positions(V);
    if (!all_finite(x)) throw Failure("solver diverged at frame " + std::to_string(frame) + " (non-finite positions)");
    if (!a.out.empty()) write_frame(a.out, frame, x, L.tris);
    log(DEBUG, "frame done", {{"frame", std::to_string(frame)}});
  }
  // state.rs:84-92 kinetic energy over the free vertices, FP32 in vertex order; main.rs:891-899
  // bounds
  std::vector<float> v = total_frames ? s.velocities(V) : L.velocities;
  float ke = 0.0f;
  for (uint32_t i = 0; i < V; ++i) {
    const float im = L.inv_mass[i];
    if (!(im > 0.0f)) continue;
    const float vx = v[3 * i], vy = v[3 * i + 1], vz = v[3 * i + 2];
    ke += 0.5f * (1.0f / im) * (vx * vx + vy * vy + vz * vz);
  }
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < V; ++i)
    for (int c = 0; c < 3; ++c) {
      lo[c] = std::min(lo[c], x[3 * i + c]);
      hi[c] = std::max(hi[c], x[3 * i + c]);
    }
  PrettyObject o;
  o.put("backend", quote(backend));
  o.put("frames", std::to_string(total_frames));
  o.put("verts", std::to_string(V));
  o.put("final_time", num_f64(time));
  o.put("final_kinetic_energy", num_f32(ke));
  o.put("bbox_min", pretty_array({num_f32(lo[0]), num_f32(lo[1]), num_f32(lo[2])}, 2));
  o.put("bbox_max", pretty_array({num_f32(hi[0]), num_f32(hi[1]), num_f32(hi[2])}, 2));
  std::printf("%s\n", o.render().c_str());
  return 0;
}

// engine-core/src/metrics.rs:84-183 check_invariants over the run's per-step metrics (the
// metrics' strain limit is 0 for a scene-built solver, implicit.rs:544-547, so that check
// auto-passes).
struct Check {
  std::string name, detail;
  bool ok;
};
std::vector<Check> check_invariants(const std::vector<nguo_step_metrics> &steps, float cg_tol, uint32_t cg_max_iter) {
  std::vector<Check> out;
  size_t bad = 0, nonfin = 0;
  for (auto &m : steps) {
    bad += m.penetration_free ? 0 : 1;
    nonfin += m.all_finite ? 0 : 1;
  }
  out.push_back({"penetration_free", std::to_string(bad) + " step(s) with penetration/CCD failure", bad == 0});
  out.push_back({"all_frames_finite", std::to_string(nonfin) + " non-finite step(s)", nonfin == 0});
  out.push_back({"strain_limit_respected", "strain limiting disabled (auto-pass)", true});
  double worst = 0.0;
  uint32_t max_it = 0;
  for (auto &m : steps) {
    worst = std::max(worst, m.max_reresid);
    max_it = std::max(max_it, m.max_cg_iters);
  }
  const double tol = (double)cg_tol * 1.5;
  out.push_back({"cg_converged", "worst reresid " + rust_exp(worst, 3) + " vs " + rust_exp(tol, 3), worst <= tol});
  out.push_back({"cg_iter_under_cap", "max cg iters " + std::to_string(max_it) + " vs cap " + std::to_string(cg_max_iter), max_it < cg_max_iter});
  uint64_t ta = 0, tf = 0;
  bool vacuous = true;
  for (size_t k = 0; k < steps.size(); ++k) {
    if (k >= steps.size() / 2) { ta += steps[k].device_alloc; tf += steps[k].device_free; }
    vacuous &= steps[k].device_alloc == 0 && steps[k].device_free == 0;
  }
  out.push_back({"zero_steady_state_alloc",
                 vacuous ? "no device-pool counters recorded (CPU tier: auto-pass, host alloc unchecked)"
                         : "steady-state device alloc " + std::to_string(ta) + " free " + std::to_string(tf),
                 ta == 0 && tf == 0});
  return out;
}

int cmd_verify(const Args &a) {
  LoweredScene L = load_scene(a.scene, NGUO_PRECOND_BLOCK);
  uint32_t total = a.steps;
  if (!a.has_steps) {  // main.rs:618-626: frames x steps per frame, at most 2000
    const uint64_t t = (uint64_t)(uint32_t)std::max<int32_t>(L.config.frames, 1) * steps_per_frame(L.config);
    total = (uint32_t)std::min<uint64_t>(t, 2000);
  }
  Session s;
  s.open_ctx();
  s.create(L);
  std::vector<nguo_step_metrics> steps;
  steps.reserve(total);
  for (uint32_t i = 0; i < total; ++i) {
    s.advance(1);
    steps.push_back(s.metrics());
  }
  auto checks = check_invariants(steps, L.params.cg_tol, L.params.cg_max_iter);
  bool all_ok = true;
  std::vector<std::string> items;
  for (auto &c : checks) {
    all_ok &= c.ok;
    PrettyObject co;
    co.put("name", quote(c.name));
    co.put("ok", c.ok ? "true" : "false");
    co.put("detail", quote(c.detail));
    items.push_back(co.render(4));
  }
  PrettyObject o;
  o.put("scene", quote(a.scene));
  o.put("steps", std::to_string(total));
  o.put("verts", std::to_string(L.n_verts()));
  o.put("all_ok", all_ok ? "true" : "false");
  o.put("checks", pretty_array(items, 2));
  std::printf("%s\n", o.render().c_str());
  return all_ok ? 0 : 2;
}

// main.rs:264-450 GaugeScene: the driver report, then either --compare-frame N (step until the
// clock reaches frame N / fps and compare with checkpoints/vert_N.bin, against the null model of
// frame 0 held still) or --steps N (step N times and summarise).
int cmd_gauge_scene(const Args &a) {
  DriverReport rep;
  CaptureScene S;
  try {
    S = lower_capture(a.scene, rep);
  } catch (const CaptureError &e) {
    throw Failure("loading gauge scene " + a.scene + ": " + e.what());
  }
  S.bind();
  std::printf("scene: %s\n", a.scene.c_str());
  std::printf("  verts %zu  tris %zu  tets %zu  rods %zu (bend hinges %zu)\n", rep.n_verts, rep.n_tris, rep.n_tets, rep.n_rods, rep.n_rod_hinges);
  std::printf("  model split: baraff-witkin %zu  isotropic-shell %zu\n", rep.n_baraff_witkin, rep.n_shell);
  std::printf("  solid-surface faces %zu  strain-limited faces %zu  pins %zu (%zu kinematic)  pulls %zu  colliders %zu (%zu animated)\n",
              rep.n_solid_surface, rep.n_strain_limited, rep.n_pins_applied, rep.n_pins_kinematic, rep.n_pulls_applied, rep.n_colliders,
              rep.n_colliders_animated);
  const nguo_params &p = S.params;
  std::printf("  solver: dt %s  gravity [%s, %s, %s]  shell mu %s lambda %s  bend %s  dhat %s  friction %s  barrier %s  cg_tol %s  newton cap %u\n",
              num_f32(p.dt).c_str(), num_f32(p.gravity[0]).c_str(), num_f32(p.gravity[1]).c_str(), num_f32(p.gravity[2]).c_str(),
              num_f32(p.shell_mu).c_str(), num_f32(p.shell_lambda).c_str(), num_f32(p.bend_stiffness).c_str(), num_f32(p.contact_dhat).c_str(),
              num_f32(p.friction).c_str(), p.barrier_kind == NGUO_BARRIER_CUBIC ? "cubic" : "log", num_f32(p.cg_tol).c_str(), p.newton_iters);
  if (rep.unsupported.empty()) {
    std::printf("  every declared feature of this scene is modeled\n");
  } else {
    std::printf("  NOT reproduced by this build:\n");
    for (auto &u : rep.unsupported) std::printf("    - %s\n", u.c_str());
  }
  std::fflush(stdout);
  if (!a.has_compare && a.steps == 0) return 0;
  if (S.has_tets) throw Failure("stepping a scene with tets: volumetric solids are not on the device step");
  const uint32_t V = S.n_verts();
  if (a.has_compare) {
    if (!S.has_toml) throw Failure("param.toml is required to map a frame to a time");
    const double target_t = (double)a.compare_frame / (double)S.toml.fps;
    std::vector<float> reference;
    std::vector<float> frame0;
    try {
      reference = load_checkpoint(a.scene, a.compare_frame);
      frame0 = load_checkpoint(a.scene, 0);
    } catch (const CaptureError &e) {
      throw Failure(std::string("checkpoints/vert_") + std::to_string(a.compare_frame) + ".bin: " + e.what());
    }
    // step until the clock reaches the frame's time (dt need not divide the frame interval)
    std::vector<float> ours = S.positions;
    double time = 0.0;
    uint64_t n_steps = 0;
    if (time < target_t) {
      Session s;
      s.open_ctx();
      check(nguo_step_create(s.ctx, &S.desc, &s.st), "building the implicit solver");
      while (time < target_t) {
        s.advance(1);
        ++n_steps;
        time = s.metrics().time;
        if (!s.metrics().all_finite) throw Failure("non-finite state at step " + std::to_string(n_steps));
      }
      ours = s.positions(V);
    }
    GateReport gate, null;
    try {
      gate = compare_frames(ours, reference);
      null = compare_frames(frame0, reference);
    } catch (const CaptureError &e) {
      throw Failure(e.what());
    }
    auto e4 = [](double v) { char b[32]; std::snprintf(b, sizeof b, "%.4e", v); return std::string(b); };
    std::printf("  frame %u: target t %.6f, reached %.6f in %llu steps\n", a.compare_frame, target_t, time, (unsigned long long)n_steps);
    std::printf("  divergence vs reference: max_rel %s  rms %s  centroid_rel %s  (diagonal %.4f)\n", e4(gate.max_rel).c_str(), e4(gate.rms).c_str(),
                e4(gate.centroid_rel).c_str(), gate.diagonal);
    std::printf("  gates (max_rel < 8e-2, rms < 2e-2, centroid_rel < 6e-3): %s\n", gate.passes() ? "PASS" : "FAIL");
    std::printf("  null model (predict frame 0, no simulation): max_rel %s  rms %s  gates %s\n", e4(null.max_rel).c_str(), e4(null.rms).c_str(),
                null.passes() ? "PASS  <-- the gates cannot distinguish a solver from no solver here" : "FAIL");
    auto skill = [](float o, float n) { return n > 0.0f ? 1.0 - (double)o / (double)n : NAN; };
    std::printf("  SKILL vs null (1 = exact, 0 = no better than not simulating): max %.3f  rms %.3f\n", skill(gate.max_rel, null.max_rel),
                skill(gate.rms, null.rms));
    // displacement fields from frame 0: magnitude ratio and direction cosine
    double mo = 0, mr = 0, dot = 0, la = 0, lb = 0;
    for (uint32_t i = 0; i < V; ++i) {
      float od[3], rd[3];
      for (int k = 0; k < 3; ++k) { od[k] = ours[3 * i + k] - frame0[3 * i + k]; rd[k] = reference[3 * i + k] - frame0[3 * i + k]; }
      const float oo = od[0] * od[0] + od[1] * od[1] + od[2] * od[2], rr = rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2];
      mo += std::sqrt((double)oo);
      mr += std::sqrt((double)rr);
      dot += (double)(od[0] * rd[0] + od[1] * rd[1] + od[2] * rd[2]);
      la += (double)oo;
      lb += (double)rr;
    }
    mo /= std::max<uint32_t>(V, 1);
    mr /= std::max<uint32_t>(V, 1);
    la = std::sqrt(la);
    lb = std::sqrt(lb);
    const double cosv = (la > 0 && lb > 0) ? dot / (la * lb) : NAN;
    std::printf("  displacement field: ours %.6f  reference %.6f  ratio %.3f  cosine %.3f\n", mo, mr, mr > 0 ? mo / mr : NAN, cosv);
    return 0;
  }
  Session s;
  s.open_ctx();
  check(nguo_step_create(s.ctx, &S.desc, &s.st), "building the implicit solver");
  uint32_t ni_max = 0, cg_max = 0, nc = 0;
  for (uint32_t i = 0; i < a.steps; ++i) {
    s.advance(1);
    const nguo_step_metrics m = s.metrics();
    if (!m.all_finite) throw Failure("non-finite state at step " + std::to_string(i));
    ni_max = std::max(ni_max, m.newton_iters);
    cg_max = std::max(cg_max, m.max_cg_iters);
    nc = std::max(nc, m.num_contacts);
  }
  const std::vector<float> x = s.positions(V);
  float min_y = INFINITY;
  for (uint32_t i = 0; i < V; ++i) min_y = std::min(min_y, x[3 * i + 1]);
  std::printf("  max per step: newton %u  cg %u  contacts %u\n", ni_max, cg_max, nc);
  std::printf("  stepped %u step(s) to t = %.4f: all finite, min y = %.6f\n", a.steps, s.metrics().time, min_y);
  return 0;
}

}  // namespace
}  // namespace mguso

int main(int argc, char **argv) {
  using namespace mguso;
  Args a;
  try {
    bool done = false;
    a = parse_args(argc, argv, done);
    if (done) return 0;
  } catch (const Usage &e) {
    std::fprintf(stderr, "error: %s\n\n%s", e.what(), USAGE);
    return 2;
  }
  try {
    if (a.command == "gpu-info") return cmd_gpu_info();
    if (a.command == "bench") return cmd_bench(a);
    if (a.command == "run") return cmd_run(a);
    if (a.command == "verify") return cmd_verify(a);
    if (a.command == "gauge-scene") return cmd_gauge_scene(a);
    throw Failure("`" + a.command + "` fits solver parameters to the reference's capture corpus, which is not distributed with this source and not part of this build");
  } catch (const std::exception &e) {
    log(ERROR, a.command == "verify" ? "verify failed" : "command failed", {{"error", e.what()}});
    return 1;
  }
}
