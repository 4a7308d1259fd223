// Compiles include/parallax_hip/engine_adapter.hpp against the engine's own
// headers and its vendored Eigen (where the engine source is present) and
// checks that every conversion round-trips exactly: a seeded SH3
// parallax::GaussianCloud (with the antialiased flag and unnormalised
// quaternions), SplatViews, SplatTrainingViews, a SplatTrainingConfig with
// every field off its default, and SplatTrainingProgress. Without a usable
// device it also checks trainSplatsOnHip's failure contract: false, a
// reason, the cloud unchanged. Test infrastructure, run by
// tests/test_engine_adapter.py.
#include <parallax_hip/engine_adapter.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

namespace {

int failures = 0;

void check(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAIL %s\n", what);
    ++failures;
  }
}

bool same(float a, float b) { return std::memcmp(&a, &b, sizeof a) == 0; }
bool same(double a, double b) { return std::memcmp(&a, &b, sizeof a) == 0; }

bool same(const parallax::Gaussian& a, const parallax::Gaussian& b) {
  for (int i = 0; i < 3; ++i)
    if (!same(a.position[i], b.position[i]) || !same(a.log_scale[i], b.log_scale[i]) || !same(a.dc[i], b.dc[i]))
      return false;
  if (!same(a.rotation.w(), b.rotation.w()) || !same(a.rotation.x(), b.rotation.x()) ||
      !same(a.rotation.y(), b.rotation.y()) || !same(a.rotation.z(), b.rotation.z()))
    return false;
  if (!same(a.opacity_logit, b.opacity_logit) || a.sh_rest.size() != b.sh_rest.size()) return false;
  for (size_t k = 0; k < a.sh_rest.size(); ++k)
    if (!same(a.sh_rest[k], b.sh_rest[k])) return false;
  return true;
}

parallax::GaussianCloud seeded_cloud(int n, int degree, unsigned seed) {
  std::mt19937 rng(seed);
  std::normal_distribution<float> normal(0.f, 1.f);
  parallax::GaussianCloud c;
  c.sh_degree = degree;
  c.antialiased = true;
  const int rest = 3 * parallax::shRestPerChannel(degree);
  for (int i = 0; i < n; ++i) {
    parallax::Gaussian g;
    g.position = parallax::Vec3f(normal(rng), normal(rng), normal(rng));
    g.log_scale = parallax::Vec3f(normal(rng) - 4.f, normal(rng) - 4.f, normal(rng) - 4.f);
    g.rotation = Eigen::Quaternionf(normal(rng), normal(rng), normal(rng), normal(rng));  // not normalised
    g.opacity_logit = normal(rng);
    g.dc = {normal(rng), normal(rng), normal(rng)};
    for (int k = 0; k < rest; ++k) g.sh_rest.push_back(0.2f * normal(rng));
    c.gaussians.push_back(std::move(g));
  }
  return c;
}

parallax::SplatView seeded_view(unsigned seed) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  parallax::SplatView v;
  v.intrinsics.fx = 1000.0 + 100.0 * u(rng);
  v.intrinsics.fy = 1000.0 + 100.0 * u(rng);
  v.intrinsics.cx = 640.0 + u(rng);
  v.intrinsics.cy = 360.0 + u(rng);
  v.intrinsics.width = 1280;
  v.intrinsics.height = 720;
  const Eigen::Quaterniond q = Eigen::Quaterniond(u(rng), u(rng), u(rng), u(rng)).normalized();
  v.camera_from_world.R = q.toRotationMatrix();
  v.camera_from_world.t = parallax::Vec3(u(rng), u(rng), 5.0 + u(rng));
  return v;
}

bool same(const parallax::SplatView& a, const parallax::SplatView& b) {
  const auto& i = a.intrinsics;
  const auto& j = b.intrinsics;
  if (!same(i.fx, j.fx) || !same(i.fy, j.fy) || !same(i.cx, j.cx) || !same(i.cy, j.cy) || i.width != j.width ||
      i.height != j.height)
    return false;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c)
      if (!same(a.camera_from_world.R(r, c), b.camera_from_world.R(r, c))) return false;
    if (!same(a.camera_from_world.t[r], b.camera_from_world.t[r])) return false;
  }
  return true;
}

}  // namespace

int main() {
  using namespace parallax_hip::engine;
  // clouds: every degree, field by field, and the hip side's layout
  for (int degree = 0; degree <= 3; ++degree) {
    const parallax::GaussianCloud c = seeded_cloud(257, degree, 11u + degree);
    const parallax_hip::GaussianCloud h = toHip(c);
    check(h.sh_degree == degree && h.antialiased && h.size() == c.size(), "cloud header");
    check(h.gaussians[5].rotation[0] == c.gaussians[5].rotation.w() &&
              h.gaussians[5].rotation[1] == c.gaussians[5].rotation.x(),
          "quaternion stored (w, x, y, z)");
    check(h.gaussians[7].sh_rest == c.gaussians[7].sh_rest, "sh_rest channel-major as stored");
    const parallax::GaussianCloud back = fromHip(h);
    bool all = back.sh_degree == c.sh_degree && back.antialiased == c.antialiased && back.size() == c.size();
    for (size_t i = 0; all && i < c.size(); ++i) all = same(back.gaussians[i], c.gaussians[i]);
    check(all, "cloud round trip");
  }
  // views and training views
  for (unsigned s = 0; s < 8; ++s) {
    const parallax::SplatView v = seeded_view(s);
    check(same(fromHip(toHip(v)), v), "view round trip");
    const parallax_hip::SplatView hv = toHip(v);
    check(hv.camera_from_world.R[1] == v.camera_from_world.R(0, 1) &&
              hv.camera_from_world.R[3] == v.camera_from_world.R(1, 0),
          "rotation row-major");
    parallax::SplatTrainingView tv;
    tv.view = v;
    tv.image.assign(12, 0.25f * static_cast<float>(s));
    const parallax::SplatTrainingView back = fromHip(toHip(tv));
    check(same(back.view, v) && back.image == tv.image, "training view round trip");
  }
  // every config field off its default
  parallax::SplatTrainingConfig c;
  c.iterations = 1234;
  c.position_lr_initial = 1.1e-4f;
  c.position_lr_final = 1.2e-6f;
  c.dc_lr = 2.1e-3f;
  c.sh_rest_lr_divider = 17.f;
  c.opacity_lr = 0.04f;
  c.scale_lr = 4e-3f;
  c.rotation_lr = 9e-4f;
  c.adam_beta1 = 0.85f;
  c.adam_beta2 = 0.995f;
  c.adam_epsilon = 1e-14f;
  c.ssim_weight = 0.3f;
  c.sh_degree = 2;
  c.sh_increase_interval = 321;
  c.densify_from_iteration = 111;
  c.densify_until_iteration = 999;
  c.densify_interval = 77;
  c.densify_grad_threshold = 3e-4f;
  c.percent_dense = 0.02f;
  c.split_scale_divider = 1.7f;
  c.opacity_reset_interval = 555;
  c.opacity_reset_value = 0.02f;
  c.prune_opacity = 0.006f;
  c.prune_scale_fraction = 0.12f;
  c.max_gaussians = 76543;
  c.seed = 99u;
  const parallax::SplatTrainingConfig b = fromHip(toHip(c));
  check(b.iterations == c.iterations && same(b.position_lr_initial, c.position_lr_initial) &&
            same(b.position_lr_final, c.position_lr_final) && same(b.dc_lr, c.dc_lr) &&
            same(b.sh_rest_lr_divider, c.sh_rest_lr_divider) && same(b.opacity_lr, c.opacity_lr) &&
            same(b.scale_lr, c.scale_lr) && same(b.rotation_lr, c.rotation_lr) &&
            same(b.adam_beta1, c.adam_beta1) && same(b.adam_beta2, c.adam_beta2) &&
            same(b.adam_epsilon, c.adam_epsilon) && same(b.ssim_weight, c.ssim_weight) &&
            b.sh_degree == c.sh_degree && b.sh_increase_interval == c.sh_increase_interval &&
            b.densify_from_iteration == c.densify_from_iteration &&
            b.densify_until_iteration == c.densify_until_iteration && b.densify_interval == c.densify_interval &&
            same(b.densify_grad_threshold, c.densify_grad_threshold) && same(b.percent_dense, c.percent_dense) &&
            same(b.split_scale_divider, c.split_scale_divider) &&
            b.opacity_reset_interval == c.opacity_reset_interval &&
            same(b.opacity_reset_value, c.opacity_reset_value) && same(b.prune_opacity, c.prune_opacity) &&
            same(b.prune_scale_fraction, c.prune_scale_fraction) && b.max_gaussians == c.max_gaussians &&
            b.seed == c.seed,
        "config round trip");
  check(toHip(c).validate() == c.validate(), "validate() agrees");
  parallax::SplatTrainingConfig bad;
  bad.sh_degree = 7;
  check(toHip(bad).validate() == bad.validate(), "validate() agrees on an invalid config");
  check(toHip(parallax::SplatTrainingConfig()).batch_views == 1, "one view per step by default");
  parallax_hip::SplatTrainingProgress p;
  p.iteration = 3;
  p.total_iterations = 9;
  p.gaussian_count = 42;
  p.loss = 0.125f;
  p.smoothed_loss = 0.25f;
  const parallax::SplatTrainingProgress q = fromHip(p);
  check(q.iteration == 3 && q.total_iterations == 9 && q.gaussian_count == 42 && q.loss == 0.125f &&
            q.smoothed_loss == 0.25f,
        "progress");
  // without a device: trainSplatsOnGpu's failure contract
  const parallax::SplatGpuAvailability a = splatHipAvailability();
  if (!a.available) {
    parallax::GaussianCloud cloud = seeded_cloud(64, 1, 5u);
    const parallax::GaussianCloud before = cloud;
    std::vector<parallax::SplatTrainingView> views(2);
    views[0].view = seeded_view(1);
    views[1].view = seeded_view(2);
    for (auto& v : views) v.image.assign(static_cast<size_t>(v.view.intrinsics.width) * v.view.intrinsics.height * 3, 0.5f);
    std::string error;
    int calls = 0;
    const bool ok = trainSplatsOnHip(&cloud, views, parallax::SplatTrainingConfig(),
                                     [&calls](const parallax::SplatTrainingProgress&) { return ++calls > 0; }, &error);
    check(!ok && !error.empty() && calls == 0, "no device: false with a reason, no callback");
    bool unchanged = cloud.size() == before.size();
    for (size_t i = 0; unchanged && i < cloud.size(); ++i) unchanged = same(cloud.gaussians[i], before.gaussians[i]);
    check(unchanged, "no device: cloud unchanged");
    std::printf("no device: \"%s\" / \"%s\"\n", a.reason.c_str(), error.c_str());
  }
  std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "ok", failures);
  return failures ? 1 : 0;
}
