// Runs the C++ facade (include/parallax_hip/splat_trainer.hpp) on inputs
// written by tests/test_gpu_golden_driver.py and prints its results, which
// that test compares with the reference build's golden fixtures. Test
// infrastructure; needs a gfx950 device.
//
//   golden_driver trainer IN OUT       HipSplatTrainer, step by step
//   golden_driver stage-seed IN OUT    splatStageSeedPoints + initializeFromPointCloud
//   golden_driver stage-config IN OUT  splatStageTrainingConfig + validate()
//
// IN and OUT are whitespace-separated text. Floats are written with 9
// significant digits (exact for float32) and read as double, then narrowed.
#include <parallax_hip/splat_trainer.hpp>

#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

using namespace parallax_hip;

namespace {

struct Reader {
  std::ifstream in;
  explicit Reader(const char* path) : in(path) {}
  double d() {
    double v = 0.0;
    if (!(in >> v)) throw std::runtime_error("input ends early");
    return v;
  }
  float f() { return static_cast<float>(d()); }
  int i() { return static_cast<int>(d()); }
  void expect(const char* word) {
    std::string w;
    if (!(in >> w) || w != word) throw std::runtime_error(std::string("expected ") + word + ", read " + w);
  }
};

GaussianCloud read_cloud(Reader& r) {
  r.expect("cloud");
  const int n = r.i();
  GaussianCloud c;
  c.sh_degree = r.i();
  const int rs = r.i();
  c.gaussians.resize(static_cast<size_t>(n));
  for (Gaussian& g : c.gaussians) {
    for (float& v : g.position) v = r.f();
    for (float& v : g.log_scale) v = r.f();
    for (float& v : g.rotation) v = r.f();
    g.opacity_logit = r.f();
    for (float& v : g.dc) v = r.f();
    g.sh_rest.resize(static_cast<size_t>(rs));
    for (float& v : g.sh_rest) v = r.f();
  }
  return c;
}

void write_cloud(FILE* out, const GaussianCloud& c) {
  const size_t rs = c.gaussians.empty() ? 0 : c.gaussians[0].sh_rest.size();
  std::fprintf(out, "cloud %zu %d %zu\n", c.size(), c.sh_degree, rs);
  for (const Gaussian& g : c.gaussians) {
    for (float v : g.position) std::fprintf(out, "%.9g ", v);
    for (float v : g.log_scale) std::fprintf(out, "%.9g ", v);
    for (float v : g.rotation) std::fprintf(out, "%.9g ", v);
    std::fprintf(out, "%.9g ", g.opacity_logit);
    for (float v : g.dc) std::fprintf(out, "%.9g ", v);
    for (float v : g.sh_rest) std::fprintf(out, "%.9g ", v);
    std::fprintf(out, "\n");
  }
}

int run_trainer(Reader& r, FILE* out) {
  r.expect("config");
  SplatTrainingConfig cfg;
  cfg.iterations = r.i();
  cfg.sh_degree = r.i();
  cfg.sh_increase_interval = r.i();
  cfg.densify_from_iteration = r.i();
  cfg.densify_until_iteration = r.i();
  cfg.densify_interval = r.i();
  cfg.densify_grad_threshold = r.f();
  cfg.opacity_reset_interval = r.i();
  cfg.ssim_weight = r.f();
  cfg.max_gaussians = r.i();
  cfg.deterministic = r.i() != 0;
  cfg.graph_step = r.i() != 0;
  cfg.batch_views = r.i();
  cfg.autotune = false;
  GaussianCloud start = read_cloud(r);
  r.expect("views");
  const int n_views = r.i();
  std::vector<SplatTrainingView> views(static_cast<size_t>(n_views));
  for (SplatTrainingView& v : views) {
    for (double& x : v.view.camera_from_world.R) x = r.d();
    for (double& x : v.view.camera_from_world.t) x = r.d();
    v.view.intrinsics.fx = r.d();
    v.view.intrinsics.fy = r.d();
    v.view.intrinsics.cx = r.d();
    v.view.intrinsics.cy = r.d();
    v.view.intrinsics.width = r.i();
    v.view.intrinsics.height = r.i();
    v.image.resize(static_cast<size_t>(v.view.intrinsics.width) * v.view.intrinsics.height * 3);
    for (float& x : v.image) x = r.f();
  }
  const std::string invalid = cfg.validate();
  if (!invalid.empty()) {
    std::fprintf(stderr, "invalid config: %s\n", invalid.c_str());
    return 2;
  }
  HipSplatTrainer trainer(std::move(start), std::move(views), cfg);
  if (!trainer.ok()) {
    std::fprintf(stderr, "trainer: %s\n", trainer.error().c_str());
    return 3;
  }
  std::vector<float> losses;
  while (trainer.step()) losses.push_back(trainer.lastLoss());
  if (!trainer.ok()) {
    std::fprintf(stderr, "trainer: %s\n", trainer.error().c_str());
    return 3;
  }
  std::fprintf(out, "losses %zu\n", losses.size());
  for (float l : losses) std::fprintf(out, "%.9g\n", l);
  std::fprintf(out, "smoothed %.9g extent %.9g\n", trainer.smoothedLoss(), trainer.sceneExtent());
  write_cloud(out, trainer.cloud());
  return 0;
}

int run_stage_seed(Reader& r, FILE* out) {
  const int max_gaussians = r.i();
  const int sh_degree = r.i();
  const int n = r.i();
  const int n_colors = r.i();
  std::vector<std::array<float, 3>> points(static_cast<size_t>(n));
  std::vector<std::array<uint8_t, 3>> colors(static_cast<size_t>(n_colors));
  for (auto& p : points)
    for (float& v : p) v = r.f();
  for (auto& c : colors)
    for (uint8_t& v : c) v = static_cast<uint8_t>(r.i());
  std::vector<std::array<float, 3>> seed_points;
  std::vector<std::array<uint8_t, 3>> seed_colors;
  splatStageSeedPoints(points, colors, max_gaussians, &seed_points, &seed_colors);
  std::string error;
  const GaussianCloud cloud = initializeFromPointCloud(seed_points, seed_colors, sh_degree, 0.1f, &error);
  if (cloud.empty()) {
    std::fprintf(stderr, "seed: %s\n", error.c_str());
    return 3;
  }
  write_cloud(out, cloud);
  return 0;
}

int run_stage_config(Reader& r, FILE* out) {
  const int cases = r.i();
  for (int k = 0; k < cases; ++k) {
    SplatStageSettings s;
    s.iterations = r.i();
    s.sh_degree = r.i();
    s.max_gaussians = r.i();
    const SplatTrainingConfig c = splatStageTrainingConfig(s);
    // the engine's SplatTrainingConfig fields, in declaration order (SplatTrainer.h:25-94)
    std::fprintf(out, "%d %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %d %d %d %d %d %.9g %.9g %.9g %d %.9g "
                 "%.9g %.9g %d %u |%s|\n",
                 c.iterations, c.position_lr_initial, c.position_lr_final, c.dc_lr, c.sh_rest_lr_divider, c.opacity_lr,
                 c.scale_lr, c.rotation_lr, c.adam_beta1, c.adam_beta2, c.adam_epsilon, c.ssim_weight, c.sh_degree,
                 c.sh_increase_interval, c.densify_from_iteration, c.densify_until_iteration, c.densify_interval,
                 c.densify_grad_threshold, c.percent_dense, c.split_scale_divider, c.opacity_reset_interval,
                 c.opacity_reset_value, c.prune_opacity, c.prune_scale_fraction, c.max_gaussians, c.seed,
                 c.validate().c_str());
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: golden_driver trainer|stage-seed|stage-config IN OUT\n");
    return 2;
  }
  const std::string mode = argv[1];
  Reader r(argv[2]);
  if (!r.in) {
    std::fprintf(stderr, "cannot read %s\n", argv[2]);
    return 2;
  }
  FILE* out = std::fopen(argv[3], "w");
  if (!out) {
    std::fprintf(stderr, "cannot write %s\n", argv[3]);
    return 2;
  }
  int rc = 2;
  try {
    if (mode == "trainer") rc = run_trainer(r, out);
    else if (mode == "stage-seed") rc = run_stage_seed(r, out);
    else if (mode == "stage-config") rc = run_stage_config(r, out);
    else std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    rc = 2;
  }
  std::fclose(out);
  return rc;
}
