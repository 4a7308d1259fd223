// C++ facade over the C ABI: the engine's splat-trainer interface with the
// model resident on the device. The schedule follows the engine's
// SplatTrainer (engine/native/cpp/src/splat/SplatTrainer.cpp:134-383):
// per-epoch mt19937 + std::shuffle view order, SH band unlock, the geometric
// position learning-rate decay scaled by the camera extent, the smoothed
// loss, densify every `densify_interval` inside [from, until], opacity reset,
// and finish = prune + stable sort by opacity.
#include "../../include/parallax_hip/splat_trainer.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>

#include "../../include/splat_hip.h"

namespace parallax_hip {
namespace {

splat_view toAbi(const SplatView& v) {
  splat_view s{};
  for (int i = 0; i < 9; ++i) s.R[i] = v.camera_from_world.R[static_cast<size_t>(i)];
  for (int i = 0; i < 3; ++i) s.t[i] = v.camera_from_world.t[static_cast<size_t>(i)];
  s.fx = v.intrinsics.fx;
  s.fy = v.intrinsics.fy;
  s.cx = v.intrinsics.cx;
  s.cy = v.intrinsics.cy;
  s.width = v.intrinsics.width;
  s.height = v.intrinsics.height;
  return s;
}

splat_raster_options toAbi(const RasterOptions& o) {
  splat_raster_options r{};
  for (int c = 0; c < 3; ++c) r.background[c] = o.background[static_cast<size_t>(c)];
  r.sh_degree = o.sh_degree;
  r.min_alpha = o.min_alpha;
  r.transmittance_cutoff = o.transmittance_cutoff;
  r.screen_dilation = o.screen_dilation;
  r.deterministic = o.deterministic ? 1 : 0;
  return r;
}

// The cloud in structure-of-arrays form for the ABI.
struct SoA {
  int n = 0, rs = 0, degree = 0;
  std::vector<float> pos, ls, rot, op, dc, rest;
  explicit SoA(const GaussianCloud& c) {
    n = static_cast<int>(c.size());
    degree = c.sh_degree;
    rs = n > 0 ? static_cast<int>(c.gaussians.front().sh_rest.size()) : 3 * shRestPerChannel(c.sh_degree);
    pos.resize(3 * static_cast<size_t>(n));
    ls.resize(3 * static_cast<size_t>(n));
    rot.resize(4 * static_cast<size_t>(n));
    op.resize(static_cast<size_t>(n));
    dc.resize(3 * static_cast<size_t>(n));
    rest.resize(static_cast<size_t>(rs) * n);
    for (int i = 0; i < n; ++i) {
      const Gaussian& g = c.gaussians[static_cast<size_t>(i)];
      for (int k = 0; k < 3; ++k) {
        pos[3 * i + k] = g.position[static_cast<size_t>(k)];
        ls[3 * i + k] = g.log_scale[static_cast<size_t>(k)];
        dc[3 * i + k] = g.dc[static_cast<size_t>(k)];
      }
      for (int k = 0; k < 4; ++k) rot[4 * i + k] = g.rotation[static_cast<size_t>(k)];
      op[static_cast<size_t>(i)] = g.opacity_logit;
      std::copy(g.sh_rest.begin(), g.sh_rest.begin() + std::min<size_t>(g.sh_rest.size(), static_cast<size_t>(rs)),
                rest.begin() + static_cast<long>(rs) * i);
    }
  }
  splat_cloud abi() {
    return {n, degree, rs, pos.data(), ls.data(), rot.data(), op.data(), dc.data(), rest.data()};
  }
};

GaussianCloud download(splat_context* ctx, int sh_degree) {
  GaussianCloud out;
  out.sh_degree = sh_degree;
  const int n = splat_count(ctx);
  const int rs = 3 * shRestPerChannel(sh_degree);
  std::vector<float> pos(3 * static_cast<size_t>(n)), ls(3 * static_cast<size_t>(n)), rot(4 * static_cast<size_t>(n)),
      op(static_cast<size_t>(n)), dc(3 * static_cast<size_t>(n)), rest(static_cast<size_t>(rs) * n);
  splat_cloud c{n, sh_degree, rs, pos.data(), ls.data(), rot.data(), op.data(), dc.data(), rest.data()};
  if (splat_download_cloud(ctx, &c) != SPLAT_OK) return out;
  out.gaussians.resize(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    Gaussian& g = out.gaussians[static_cast<size_t>(i)];
    for (int k = 0; k < 3; ++k) {
      g.position[static_cast<size_t>(k)] = pos[3 * i + k];
      g.log_scale[static_cast<size_t>(k)] = ls[3 * i + k];
      g.dc[static_cast<size_t>(k)] = dc[3 * i + k];
    }
    for (int k = 0; k < 4; ++k) g.rotation[static_cast<size_t>(k)] = rot[4 * i + k];
    g.opacity_logit = op[static_cast<size_t>(i)];
    g.sh_rest.assign(rest.begin() + static_cast<long>(rs) * i, rest.begin() + static_cast<long>(rs) * (i + 1));
  }
  return out;
}

// cameraExtent, SplatTrainer.cpp:74-85.
float cameraExtent(const std::vector<SplatTrainingView>& views) {
  if (views.empty()) return 1.f;
  auto center = [](const SplatView& v, float c[3]) {
    const auto& R = v.camera_from_world.R;
    const auto& t = v.camera_from_world.t;
    for (int i = 0; i < 3; ++i) {
      const double d = (-R[static_cast<size_t>(0 * 3 + i)]) * t[0] + (-R[static_cast<size_t>(1 * 3 + i)]) * t[1] +
                       (-R[static_cast<size_t>(2 * 3 + i)]) * t[2];
      c[i] = static_cast<float>(d);
    }
  };
  float centroid[3] = {0.f, 0.f, 0.f};
  for (const SplatTrainingView& v : views) {
    float c[3];
    center(v.view, c);
    for (int i = 0; i < 3; ++i) centroid[i] += c[i];
  }
  for (float& c : centroid) c /= static_cast<float>(views.size());
  float furthest = 0.f;
  for (const SplatTrainingView& v : views) {
    float c[3];
    center(v.view, c);
    const float d[3] = {c[0] - centroid[0], c[1] - centroid[1], c[2] - centroid[2]};
    furthest = std::max(furthest, std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
  }
  return std::max(furthest * 1.1f, 1e-3f);
}

std::string lastError(splat_context* ctx) { return std::string(splat_last_error(ctx)); }

class HipBackend final : public SplatBackend {
 public:
  explicit HipBackend(splat_context* ctx) : ctx_(ctx) {}
  ~HipBackend() override { splat_destroy(ctx_); }
  std::string name() const override { return "hip"; }

  SplatImage render(const GaussianCloud& cloud, const SplatView& view, const RasterOptions& options) override {
    SplatImage img;
    img.width = view.intrinsics.width;
    img.height = view.intrinsics.height;
    if (!prepare(cloud, view, nullptr)) return img;
    const splat_raster_options o = toAbi(options);
    const int32_t id = 0;
    if (splat_render_forward(ctx_, 1, &id, &o) != SPLAT_OK) return img;
    const size_t px = static_cast<size_t>(img.width) * img.height;
    img.rgb.resize(px * 3);
    img.transmittance.resize(px);
    splat_download_image(ctx_, 0, img.rgb.data(), img.transmittance.data(), nullptr);
    return img;
  }

  float renderAndAccumulate(const GaussianCloud& cloud, const SplatView& view, const RasterOptions& options,
                            const std::vector<float>& target, float ssim_weight, SplatGradients* gradients) override {
    if (!prepare(cloud, view, &target)) return 0.f;
    const splat_raster_options o = toAbi(options);
    const int32_t id = 0;
    float loss = 0.f;
    if (splat_render_forward(ctx_, 1, &id, &o) != SPLAT_OK || splat_loss(ctx_, ssim_weight, &loss) != SPLAT_OK ||
        splat_zero_gradients(ctx_) != SPLAT_OK || splat_render_backward(ctx_) != SPLAT_OK)
      return 0.f;
    const size_t n = cloud.size();
    const int rs = n > 0 ? static_cast<int>(cloud.gaussians.front().sh_rest.size()) : 0;
    if (gradients->position.size() != n) gradients->resize(n, rs);
    std::vector<float> pos(3 * n), ls(3 * n), rot(4 * n), op(n), dc(3 * n), rest(static_cast<size_t>(rs) * n), sg(n);
    std::vector<int32_t> vc(n);
    splat_gradients g{pos.data(), ls.data(), rot.data(), op.data(), dc.data(), rest.data(), sg.data(), vc.data()};
    if (splat_download_gradients(ctx_, &g) != SPLAT_OK) return loss;
    // the engine's backward adds into the caller's accumulators
    for (size_t i = 0; i < n; ++i) {
      for (int k = 0; k < 3; ++k) {
        gradients->position[i][static_cast<size_t>(k)] += pos[3 * i + static_cast<size_t>(k)];
        gradients->log_scale[i][static_cast<size_t>(k)] += ls[3 * i + static_cast<size_t>(k)];
        gradients->dc[i][static_cast<size_t>(k)] += dc[3 * i + static_cast<size_t>(k)];
      }
      for (int k = 0; k < 4; ++k) gradients->rotation[i][static_cast<size_t>(k)] += rot[4 * i + static_cast<size_t>(k)];
      gradients->opacity_logit[i] += op[i];
      gradients->screen_gradient[i] += sg[i];
      gradients->visible_count[i] += vc[i];
    }
    for (size_t k = 0; k < rest.size() && k < gradients->sh_rest.size(); ++k) gradients->sh_rest[k] += rest[k];
    return loss;
  }

 private:
  bool prepare(const GaussianCloud& cloud, const SplatView& view, const std::vector<float>* target) {
    SoA soa(cloud);
    const int degree = cloud.sh_degree;
    if (soa.rs != 3 * shRestPerChannel(degree)) return false;
    if (splat_set_capacity(ctx_, std::max(1, soa.n), degree) != SPLAT_OK) return false;
    splat_cloud c = soa.abi();
    if (splat_upload_cloud(ctx_, &c) != SPLAT_OK) return false;
    const splat_view v = toAbi(view);
    const float* img = target ? target->data() : nullptr;
    return splat_upload_views(ctx_, 1, &v, target ? &img : nullptr) == SPLAT_OK;
  }

  splat_context* ctx_;
};

}  // namespace

void SplatGradients::resize(size_t count, int rest_stride) {
  sh_rest_stride = rest_stride;
  position.assign(count, {0.f, 0.f, 0.f});
  log_scale.assign(count, {0.f, 0.f, 0.f});
  rotation.assign(count, {0.f, 0.f, 0.f, 0.f});
  opacity_logit.assign(count, 0.f);
  dc.assign(count, {0.f, 0.f, 0.f});
  sh_rest.assign(count * static_cast<size_t>(std::max(0, rest_stride)), 0.f);
  screen_gradient.assign(count, 0.f);
  visible_count.assign(count, 0);
}

std::string SplatTrainingConfig::validate() const {
  // SplatTrainer.cpp:118-126, plus the batch size
  if (iterations <= 0) return "iterations must be positive";
  if (sh_degree < 0 || sh_degree > 3) return "sh_degree must be between 0 and 3";
  if (densify_interval <= 0) return "densify_interval must be positive";
  if (split_scale_divider <= 1.f) return "split_scale_divider must be greater than one";
  if (max_gaussians <= 0) return "max_gaussians must be positive";
  if (ssim_weight < 0.f || ssim_weight > 1.f) return "ssim_weight must be in [0,1]";
  if (batch_views < 1 || batch_views > 64) return "batch_views must be between 1 and 64";
  return {};
}

std::unique_ptr<SplatBackend> makeHipSplatBackend(std::string* error, int device) {
  splat_context* ctx = nullptr;
  if (splat_create(device, &ctx) != SPLAT_OK) {
    if (error) *error = lastError(nullptr);
    return nullptr;
  }
  return std::unique_ptr<SplatBackend>(new HipBackend(ctx));
}

SplatGpuAvailability splatGpuAvailability() {
  SplatGpuAvailability a;
  char name[256] = {0}, reason[256] = {0};
  a.available = splat_availability(name, sizeof name, reason, sizeof reason) == SPLAT_OK;
  a.device_name = name;
  a.reason = reason;
  return a;
}

// ------------------------------------------------------------ trainer --
struct HipSplatTrainer::Impl {
  splat_context* ctx = nullptr;
  std::string error;
  SplatTrainingConfig config;
  int n_views = 0;
  int sh_degree = 0;  // the cloud's degree (capacity layout)
  std::mt19937 engine;
  int iteration = 0;
  float extent = 1.f;
  float last_loss = 0.f, smoothed_loss = 0.f;
  uint32_t densify_rng = 1u;
  std::vector<int> order;
  size_t cursor = 0;
  bool empty_cloud = false;

  ~Impl() {
    if (ctx) splat_destroy(ctx);
  }

  bool fail(const std::string& what) {
    error = what + (ctx ? std::string(": ") + splat_last_error(ctx) : std::string());
    return false;
  }

  int shDegreeInUse() const {
    const int unlocked = config.sh_increase_interval > 0 ? iteration / config.sh_increase_interval : config.sh_degree;
    return std::min(config.sh_degree, std::max(0, unlocked));
  }

  float positionLearningRate() const {
    const float t = config.iterations <= 1 ? 1.f
                                           : static_cast<float>(iteration) / static_cast<float>(config.iterations - 1);
    const float ratio = config.position_lr_final / std::max(config.position_lr_initial, 1e-30f);
    return config.position_lr_initial * std::pow(ratio, std::clamp(t, 0.f, 1.f)) * extent;
  }

  int nextView() {
    if (cursor >= order.size()) {
      std::shuffle(order.begin(), order.end(), engine);
      cursor = 0;
    }
    return order[cursor++];
  }

  splat_adam_params adam() const {
    splat_adam_params a{};
    a.lr_position = positionLearningRate();
    a.lr_log_scale = config.scale_lr;
    a.lr_rotation = config.rotation_lr;
    a.lr_opacity = config.opacity_lr;
    a.lr_dc = config.dc_lr;
    a.lr_rest = config.dc_lr / std::max(1.f, config.sh_rest_lr_divider);
    a.beta1 = config.adam_beta1;
    a.beta2 = config.adam_beta2;
    a.epsilon = config.adam_epsilon;
    const int t = iteration + 1;
    a.correction1 = 1.f - std::pow(config.adam_beta1, static_cast<float>(t));
    a.correction2 = 1.f - std::pow(config.adam_beta2, static_cast<float>(t));
    return a;
  }
};

HipSplatTrainer::HipSplatTrainer(GaussianCloud initial, std::vector<SplatTrainingView> views,
                                 SplatTrainingConfig config)
    : impl_(new Impl) {
  Impl& m = *impl_;
  m.config = config;
  m.engine.seed(config.seed == 0u ? 1u : config.seed);
  m.densify_rng = config.seed == 0u ? 1u : config.seed;
  m.extent = cameraExtent(views);
  m.n_views = static_cast<int>(views.size());
  m.sh_degree = initial.sh_degree;
  m.empty_cloud = initial.empty();
  m.order.resize(views.size());
  std::iota(m.order.begin(), m.order.end(), 0);
  std::shuffle(m.order.begin(), m.order.end(), m.engine);
  const std::string invalid = config.validate();
  if (!invalid.empty()) {
    m.error = invalid;
    return;
  }
  if (splat_create(config.device, &m.ctx) != SPLAT_OK) {
    m.error = lastError(nullptr);
    m.ctx = nullptr;
    return;
  }
  SoA soa(initial);
  if (soa.rs != 3 * shRestPerChannel(initial.sh_degree)) {
    m.error = "sh_rest size does not match the cloud's SH degree";
    return;
  }
  const int capacity = std::max({1, config.max_gaussians, soa.n});
  if (splat_set_capacity(m.ctx, capacity, initial.sh_degree) != SPLAT_OK) {
    m.fail("reserving device memory");
    return;
  }
  splat_set_graph_mode(m.ctx, config.graph_step ? 1 : 0);
  splat_set_autotune(m.ctx, config.autotune ? 1 : 0);
  splat_cloud c = soa.abi();
  if (splat_upload_cloud(m.ctx, &c) != SPLAT_OK) {
    m.fail("uploading the cloud");
    return;
  }
  std::vector<splat_view> vs;
  std::vector<const float*> imgs;
  for (const SplatTrainingView& v : views) {
    vs.push_back(toAbi(v.view));
    imgs.push_back(v.image.data());
  }
  if (splat_upload_views(m.ctx, static_cast<int32_t>(vs.size()), vs.data(), imgs.data()) != SPLAT_OK) {
    m.fail("uploading the training views");
    return;
  }
}

HipSplatTrainer::~HipSplatTrainer() = default;

bool HipSplatTrainer::ok() const { return impl_->ctx != nullptr && impl_->error.empty(); }
const std::string& HipSplatTrainer::error() const { return impl_->error; }

bool HipSplatTrainer::step() {
  Impl& m = *impl_;
  if (!ok()) return false;
  if (m.iteration >= m.config.iterations) return false;
  if (m.n_views == 0 || splat_count(m.ctx) == 0) {
    m.iteration = m.config.iterations;
    return false;
  }
  RasterOptions options;
  options.sh_degree = m.shDegreeInUse();
  options.deterministic = m.config.deterministic;
  const splat_raster_options o = toAbi(options);
  std::vector<int32_t> ids;
  for (int k = 0; k < m.config.batch_views; ++k) ids.push_back(m.nextView());
  const splat_adam_params a = m.adam();
  std::vector<float> losses(ids.size());
  if (splat_train_step(m.ctx, static_cast<int32_t>(ids.size()), ids.data(), &o, m.config.ssim_weight, &a,
                       losses.data()) != SPLAT_OK)
    return m.fail("training step");
  float loss = 0.f;
  for (float l : losses) loss += l;
  m.last_loss = loss / static_cast<float>(losses.size());
  m.smoothed_loss = m.iteration == 0 ? m.last_loss : 0.99f * m.smoothed_loss + 0.01f * m.last_loss;
  ++m.iteration;
  const SplatTrainingConfig& c = m.config;
  if (m.iteration >= c.densify_from_iteration && m.iteration <= c.densify_until_iteration &&
      m.iteration % c.densify_interval == 0) {
    splat_densify_params d{};
    d.densify_grad_threshold = c.densify_grad_threshold;
    d.percent_dense = c.percent_dense;
    d.split_scale_divider = c.split_scale_divider;
    d.opacity_reset_interval = c.opacity_reset_interval;
    d.prune_opacity = c.prune_opacity;
    d.prune_scale_fraction = c.prune_scale_fraction;
    d.max_gaussians = c.max_gaussians;
    if (splat_densify(m.ctx, &d, m.extent, m.iteration, &m.densify_rng, nullptr, nullptr) != SPLAT_OK)
      return m.fail("densify");
  }
  if (c.opacity_reset_interval > 0 && m.iteration % c.opacity_reset_interval == 0 &&
      m.iteration < c.densify_until_iteration) {
    // resetOpacity, SplatTrainer.cpp:301-309
    const float value = std::clamp(c.opacity_reset_value, 1e-6f, 1.f - 1e-6f);
    if (splat_reset_opacity(m.ctx, std::log(value / (1.f - value))) != SPLAT_OK) return m.fail("opacity reset");
  }
  return true;
}

bool HipSplatTrainer::train(const SplatTrainingCallback& callback) {
  while (step()) {
    if (callback == nullptr) continue;
    SplatTrainingProgress p;
    p.iteration = impl_->iteration;
    p.total_iterations = impl_->config.iterations;
    p.gaussian_count = splat_count(impl_->ctx);
    p.loss = impl_->last_loss;
    p.smoothed_loss = impl_->smoothed_loss;
    if (!callback(p)) return false;
  }
  return ok();
}

int HipSplatTrainer::iteration() const { return impl_->iteration; }
int HipSplatTrainer::shDegreeInUse() const { return impl_->shDegreeInUse(); }
float HipSplatTrainer::lastLoss() const { return impl_->last_loss; }
float HipSplatTrainer::smoothedLoss() const { return impl_->smoothed_loss; }
float HipSplatTrainer::sceneExtent() const { return impl_->extent; }

GaussianCloud HipSplatTrainer::cloud() const {
  if (!impl_->ctx) return {};
  return download(impl_->ctx, impl_->sh_degree);
}

GaussianCloud HipSplatTrainer::finish() {
  Impl& m = *impl_;
  if (!m.ctx) return {};
  if (splat_prune_and_sort(m.ctx, m.config.prune_opacity) != SPLAT_OK) m.fail("finish");
  GaussianCloud out = download(m.ctx, m.sh_degree);
  out.sh_degree = m.config.sh_degree;
  return out;
}

bool trainSplatsOnGpu(GaussianCloud* cloud, const std::vector<SplatTrainingView>& views,
                      const SplatTrainingConfig& config, const SplatTrainingCallback& callback, std::string* error) {
  if (!cloud) {
    if (error) *error = "null cloud";
    return false;
  }
  HipSplatTrainer trainer(*cloud, views, config);
  if (!trainer.ok()) {
    if (error) *error = trainer.error();
    return false;
  }
  trainer.train(callback);  // false on cancel: the partial field is still kept
  if (!trainer.error().empty()) {
    if (error) *error = trainer.error();
    return false;
  }
  *cloud = trainer.finish();
  if (!trainer.error().empty()) {
    if (error) *error = trainer.error();
    return false;
  }
  return true;
}

SplatImage renderSplats(const GaussianCloud& cloud, const SplatView& view, const RasterOptions& options,
                        std::string* error) {
  std::unique_ptr<SplatBackend> backend = makeHipSplatBackend(error);
  if (!backend) return SplatImage{};
  return backend->render(cloud, view, options);
}

namespace {

// A context sized for `n` gaussians of `degree` (null and `error` set if the
// device cannot be used).
splat_context* makeContext(int n, int degree, int device, std::string* error) {
  splat_context* ctx = nullptr;
  if (splat_create(device, &ctx) != SPLAT_OK) {
    if (error) *error = splat_last_error(nullptr);
    return nullptr;
  }
  if (splat_set_capacity(ctx, std::max(1, n), degree) != SPLAT_OK) {
    if (error) *error = splat_last_error(ctx);
    splat_destroy(ctx);
    return nullptr;
  }
  return ctx;
}

}  // namespace

GaussianCloud initializeFromPointCloud(const std::vector<std::array<float, 3>>& points,
                                       const std::vector<std::array<uint8_t, 3>>& colors, int sh_degree,
                                       float initial_opacity, std::string* error) {
  GaussianCloud cloud;
  cloud.sh_degree = std::max(0, std::min(3, sh_degree));
  if (points.empty()) return cloud;
  const int n = static_cast<int>(points.size());
  splat_context* ctx = makeContext(n, cloud.sh_degree, 0, error);
  if (!ctx) return cloud;
  std::vector<float> xyz(3 * points.size());
  for (size_t i = 0; i < points.size(); ++i)
    for (int k = 0; k < 3; ++k) xyz[3 * i + static_cast<size_t>(k)] = points[i][static_cast<size_t>(k)];
  std::vector<uint8_t> rgb;
  if (colors.size() == points.size()) {
    rgb.resize(3 * colors.size());
    for (size_t i = 0; i < colors.size(); ++i)
      for (int k = 0; k < 3; ++k) rgb[3 * i + static_cast<size_t>(k)] = colors[i][static_cast<size_t>(k)];
  }
  if (splat_init_from_points(ctx, n, xyz.data(), rgb.empty() ? nullptr : rgb.data(), cloud.sh_degree,
                             initial_opacity, nullptr) != SPLAT_OK) {
    if (error) *error = splat_last_error(ctx);
  } else {
    cloud = download(ctx, cloud.sh_degree);
  }
  splat_destroy(ctx);
  return cloud;
}

bool writeGaussianPly(const std::string& path, const GaussianCloud& cloud, std::string* error) {
  if (cloud.empty()) {
    if (error) *error = "the cloud is empty";
    return false;
  }
  SoA soa(cloud);
  splat_context* ctx = makeContext(soa.n, cloud.sh_degree, 0, error);
  if (!ctx) return false;
  const splat_cloud c = soa.abi();
  const bool ok = splat_upload_cloud(ctx, &c) == SPLAT_OK && splat_write_ply(ctx, path.c_str()) == SPLAT_OK;
  if (!ok && error) *error = splat_last_error(ctx);
  splat_destroy(ctx);
  return ok;
}

bool readGaussianPly(const std::string& path, GaussianCloud* out, std::string* error) {
  int32_t n = 0, degree = 0;
  char message[256] = {0};
  if (splat_ply_info(path.c_str(), &n, &degree, message, sizeof message) != SPLAT_OK) {
    if (error) *error = message;
    return false;
  }
  splat_context* ctx = makeContext(n, degree, 0, error);
  if (!ctx) return false;
  const bool ok = splat_read_ply(ctx, path.c_str()) == SPLAT_OK;
  if (!ok) {
    if (error) *error = splat_last_error(ctx);
  } else if (out) {
    *out = download(ctx, degree);
  }
  splat_destroy(ctx);
  return ok;
}

namespace {

// Uploads `cloud` into a fresh context and runs `write` on it.
bool writeWith(const GaussianCloud& cloud, std::string* error,
               const std::function<splat_status(splat_context*)>& write) {
  if (cloud.empty()) {
    if (error) *error = "the cloud is empty";
    return false;
  }
  SoA soa(cloud);
  splat_context* ctx = makeContext(soa.n, cloud.sh_degree, 0, error);
  if (!ctx) return false;
  const splat_cloud c = soa.abi();
  const bool ok = splat_upload_cloud(ctx, &c) == SPLAT_OK && write(ctx) == SPLAT_OK;
  if (!ok && error) *error = splat_last_error(ctx);
  splat_destroy(ctx);
  return ok;
}

// Sizes a fresh context for `n` gaussians of `degree`, runs `read` on it and
// downloads the result.
bool readWith(int n, int degree, GaussianCloud* out, std::string* error,
              const std::function<splat_status(splat_context*)>& read) {
  splat_context* ctx = makeContext(n, degree, 0, error);
  if (!ctx) return false;
  const bool ok = read(ctx) == SPLAT_OK;
  if (!ok) {
    if (error) *error = splat_last_error(ctx);
  } else if (out) {
    *out = download(ctx, degree);
  }
  splat_destroy(ctx);
  return ok;
}

}  // namespace

bool writeSplat(const std::string& path, const GaussianCloud& cloud, std::string* error) {
  return writeWith(cloud, error, [&](splat_context* ctx) { return splat_write_splat(ctx, path.c_str()); });
}

bool readSplat(const std::string& path, GaussianCloud* out, std::string* error) {
  int32_t n = 0;
  char message[256] = {0};
  if (splat_splat_info(path.c_str(), &n, message, sizeof message) != SPLAT_OK) {
    if (error) *error = message;
    return false;
  }
  return readWith(n, 0, out, error, [&](splat_context* ctx) { return splat_read_splat(ctx, path.c_str()); });
}

bool writeSpz(const std::string& path, const GaussianCloud& cloud, int fractional_bits, std::string* error) {
  if (cloud.sh_degree < 0 || cloud.sh_degree > 3) {
    if (error) *error = "sh_degree must be 0..3";
    return false;
  }
  return writeWith(cloud, error, [&](splat_context* ctx) {
    return splat_write_spz(ctx, path.c_str(), fractional_bits, cloud.antialiased ? 1 : 0);
  });
}

bool readSpz(const std::string& path, GaussianCloud* out, std::string* error) {
  int32_t n = 0, degree = 0, antialiased = 0;
  char message[256] = {0};
  if (splat_spz_info(path.c_str(), &n, &degree, &antialiased, message, sizeof message) != SPLAT_OK) {
    if (error) *error = message;
    return false;
  }
  const bool ok = readWith(n, degree, out, error, [&](splat_context* ctx) {
    return splat_read_spz(ctx, path.c_str(), nullptr);
  });
  if (ok && out) out->antialiased = antialiased != 0;
  return ok;
}

void splatStageSeedPoints(const std::vector<std::array<float, 3>>& points,
                          const std::vector<std::array<uint8_t, 3>>& colors, int max_gaussians,
                          std::vector<std::array<float, 3>>* seed_points,
                          std::vector<std::array<uint8_t, 3>>* seed_colors) {
  const size_t budget = std::max<size_t>(1, max_gaussians / 4);
  if (points.size() <= budget) {
    *seed_points = points;
    *seed_colors = colors;
    return;
  }
  const size_t step = (points.size() + budget - 1) / budget;
  seed_points->clear();
  seed_colors->clear();
  for (size_t i = 0; i < points.size(); i += step) {
    seed_points->push_back(points[i]);
    seed_colors->push_back(i < colors.size() ? colors[i] : std::array<uint8_t, 3>{128, 128, 128});
  }
}

SplatTrainingConfig splatStageTrainingConfig(const SplatStageSettings& settings) {
  SplatTrainingConfig config;
  config.iterations = std::max(1, settings.iterations);
  config.sh_degree = settings.sh_degree;
  config.max_gaussians = settings.max_gaussians;
  config.densify_until_iteration = config.iterations / 2;
  config.opacity_reset_interval = std::max(1, config.iterations / 3);
  config.sh_increase_interval = std::max(1, config.iterations / 8);
  config.batch_views = std::max(1, settings.batch_views);
  config.device = settings.device;
  return config;
}

bool runSplatStageOnGpu(const std::vector<std::array<float, 3>>& points,
                        const std::vector<std::array<uint8_t, 3>>& colors, std::vector<SplatTrainingView> views,
                        const SplatStageSettings& settings, const std::function<bool()>& cancelled,
                        const std::vector<std::string>& ply_paths, SplatStageResult* result, std::string* error) {
  SplatStageResult local;
  SplatStageResult& r = result ? *result : local;
  r = SplatStageResult{};
  const auto fail = [error](const std::string& m) {
    if (error) *error = m;
    return false;
  };
  if (points.empty()) return fail("no point cloud to seed the splat field from");
  if (views.size() < 2) return fail("at least two posed views are needed to train splats");
  std::vector<std::array<float, 3>> seed_points;
  std::vector<std::array<uint8_t, 3>> seed_colors;
  splatStageSeedPoints(points, colors, settings.max_gaussians, &seed_points, &seed_colors);
  std::string init_error;
  GaussianCloud cloud = initializeFromPointCloud(seed_points, seed_colors, settings.sh_degree, 0.1f, &init_error);
  if (cloud.empty()) return fail(init_error.empty() ? "the point cloud produced no gaussians" : init_error);
  const SplatTrainingConfig config = splatStageTrainingConfig(settings);
  const std::string invalid = config.validate();
  if (!invalid.empty()) return fail(invalid);
  // cancel and time budget through the per-iteration callback (:1929-1954)
  const auto started = std::chrono::steady_clock::now();
  const SplatTrainingCallback callback = [&](const SplatTrainingProgress& p) {
    if (cancelled && cancelled()) {
      r.cancelled = true;
      return false;
    }
    r.iterations_done = p.iteration;
    if (settings.time_budget_millis > 0) {
      const auto elapsed =
          std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - started).count();
      if (elapsed >= settings.time_budget_millis) {
        r.ran_out_of_time = true;
        return false;
      }
    }
    return true;
  };
  std::string train_error;
  if (!trainSplatsOnGpu(&cloud, views, config, callback, &train_error)) return fail(train_error);
  if (r.cancelled) return fail("cancelled");
  r.backend = "hip";
  for (const std::string& path : ply_paths) {
    std::string write_error;
    if (!writeGaussianPly(path, cloud, &write_error)) return fail("could not write " + path + ": " + write_error);
  }
  r.note = std::to_string(cloud.size()) + " gaussians (hip)";
  if (r.ran_out_of_time)
    r.note += ", stopped at iteration " + std::to_string(r.iterations_done) + " when the time budget ran out";
  r.cloud = std::move(cloud);
  return true;
}

}  // namespace parallax_hip
