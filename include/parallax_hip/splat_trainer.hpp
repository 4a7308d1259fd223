// C++ facade of the MI355X splat trainer, mirroring the engine's splat
// interface so that host code written against it switches by changing the
// namespace:
//
//   engine (reference)                                   here
//   SplatView, RasterOptions, SplatImage, SplatGradients parallax_hip:: same names, same fields,
//     (SplatRaster.h:26-114)                             Eigen types replaced by std::array
//   Gaussian, GaussianCloud (GaussianCloud.h:41-72)      same
//   SplatTrainingView/Config/Progress/Callback           same, plus batch_views / deterministic
//     (SplatTrainer.h:20-130)
//   SplatBackend, makeCpuSplatBackend (:114-127)         SplatBackend, makeHipSplatBackend
//   SplatTrainer (:132-161)                              HipSplatTrainer (same methods)
//   splatGpuAvailability, trainSplatsOnGpu               same signatures
//     (VulkanSplatTrainer.h:15-37)
//
// Everything here is host C++ over the C ABI in include/splat_hip.h; the
// model stays in device memory between steps. Errors are returned, never
// thrown across this boundary (bool + error string, as the engine does).
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

struct splat_context;

namespace parallax_hip {

constexpr float kShC0 = 0.28209479177387814f;

inline int shRestPerChannel(int degree) {
  switch (degree) {
    case 1: return 3;
    case 2: return 8;
    case 3: return 15;
    default: return 0;
  }
}

struct Gaussian {
  std::array<float, 3> position{0.f, 0.f, 0.f};
  std::array<float, 3> log_scale{0.f, 0.f, 0.f};
  std::array<float, 4> rotation{1.f, 0.f, 0.f, 0.f};  // (w, x, y, z)
  float opacity_logit = 0.f;
  std::array<float, 3> dc{0.f, 0.f, 0.f};
  std::vector<float> sh_rest;  // channel-major
};

struct GaussianCloud {
  std::vector<Gaussian> gaussians;
  int sh_degree = 0;
  /// Trained with the antialiased rasteriser: carried by .spz as a flag
  /// (GaussianCloud.h:58-60).
  bool antialiased = false;
  size_t size() const { return gaussians.size(); }
  bool empty() const { return gaussians.empty(); }
};

struct CameraIntrinsics {
  double fx = 0.0, fy = 0.0, cx = 0.0, cy = 0.0;
  int width = 0, height = 0;
};

/// Camera-from-world transform: p_cam = R p + t, R row-major.
struct Pose {
  std::array<double, 9> R{1, 0, 0, 0, 1, 0, 0, 0, 1};
  std::array<double, 3> t{0, 0, 0};
};

struct SplatView {
  CameraIntrinsics intrinsics;
  Pose camera_from_world;
};

struct RasterOptions {
  std::array<float, 3> background{0.f, 0.f, 0.f};
  int sh_degree = 0;
  float min_alpha = 1.f / 255.f;
  float transmittance_cutoff = 1e-4f;
  float screen_dilation = 0.3f;
  /// Fixed-order gradient reduction: bitwise repeatable runs.
  bool deterministic = false;
};

struct SplatImage {
  int width = 0;
  int height = 0;
  std::vector<float> rgb;
  std::vector<float> transmittance;
  float at(int x, int y, int channel) const { return rgb[(static_cast<size_t>(y) * width + x) * 3 + channel]; }
};

struct SplatGradients {
  std::vector<std::array<float, 3>> position;
  std::vector<std::array<float, 3>> log_scale;
  std::vector<std::array<float, 4>> rotation;  // (w, x, y, z)
  std::vector<float> opacity_logit;
  std::vector<std::array<float, 3>> dc;
  std::vector<float> sh_rest;
  int sh_rest_stride = 0;
  std::vector<float> screen_gradient;
  std::vector<int> visible_count;
  void resize(size_t count, int rest_stride);
  void clear() { resize(position.size(), sh_rest_stride); }
};

struct SplatTrainingView {
  SplatView view;
  std::vector<float> image;
};

struct SplatTrainingConfig {
  int iterations = 7000;
  float position_lr_initial = 1.6e-4f;
  float position_lr_final = 1.6e-6f;
  float dc_lr = 2.5e-3f;
  float sh_rest_lr_divider = 20.f;
  float opacity_lr = 0.05f;
  float scale_lr = 5e-3f;
  float rotation_lr = 1e-3f;
  float adam_beta1 = 0.9f;
  float adam_beta2 = 0.999f;
  float adam_epsilon = 1e-15f;
  float ssim_weight = 0.2f;
  int sh_degree = 3;
  int sh_increase_interval = 1000;
  int densify_from_iteration = 500;
  int densify_until_iteration = 5000;
  int densify_interval = 100;
  float densify_grad_threshold = 2e-4f;
  float percent_dense = 0.01f;
  float split_scale_divider = 1.6f;
  int opacity_reset_interval = 3000;
  float opacity_reset_value = 0.01f;
  float prune_opacity = 0.005f;
  float prune_scale_fraction = 0.1f;
  int max_gaussians = 600000;
  unsigned seed = 7u;
  /// Views per step (one multi-view launch sequence, gradients summed, one
  /// Adam step). 1 reproduces the engine's schedule exactly.
  int batch_views = 1;
  bool deterministic = false;
  int device = 0;
  /// Replay each step from a captured HIP graph (same results as plain launches).
  bool graph_step = true;
  /// Time the compositing launch shapes on the first batch of each resolution
  /// mix and keep the fastest (same results whichever is kept).
  bool autotune = true;

  std::string validate() const;
};

struct SplatTrainingProgress {
  int iteration = 0;
  int total_iterations = 0;
  int gaussian_count = 0;
  float loss = 0.f;
  float smoothed_loss = 0.f;
};

using SplatTrainingCallback = std::function<bool(const SplatTrainingProgress&)>;

/// The per-view seam of the engine's trainer (SplatTrainer.h:114-125).
class SplatBackend {
 public:
  virtual ~SplatBackend() = default;
  virtual std::string name() const = 0;
  virtual float renderAndAccumulate(const GaussianCloud& cloud, const SplatView& view, const RasterOptions& options,
                                    const std::vector<float>& target, float ssim_weight,
                                    SplatGradients* gradients) = 0;
  virtual SplatImage render(const GaussianCloud& cloud, const SplatView& view, const RasterOptions& options) = 0;
};

/// A SplatBackend on the GPU: uploads the cloud and view, runs the HIP
/// forward / loss / adjoint, and adds the gradients into `gradients`.
/// Returns null (and sets `error`) if no usable device is present.
std::unique_ptr<SplatBackend> makeHipSplatBackend(std::string* error = nullptr, int device = 0);

/// The engine's training schedule (SplatTrainer.cpp:134-383) with the model
/// resident on the device: view order from mt19937 + std::shuffle per epoch,
/// SH bands unlocked on schedule, Adam with the decaying position rate,
/// densify / prune and opacity reset on the device.
class HipSplatTrainer {
 public:
  HipSplatTrainer(GaussianCloud initial, std::vector<SplatTrainingView> views, SplatTrainingConfig config);
  ~HipSplatTrainer();
  HipSplatTrainer(const HipSplatTrainer&) = delete;
  HipSplatTrainer& operator=(const HipSplatTrainer&) = delete;

  /// False if the device could not be set up; `error()` says why.
  bool ok() const;
  const std::string& error() const;

  bool step();
  bool train(const SplatTrainingCallback& callback = nullptr);
  int iteration() const;
  int shDegreeInUse() const;
  float lastLoss() const;
  float smoothedLoss() const;
  float sceneExtent() const;
  /// Downloads the current cloud.
  GaussianCloud cloud() const;
  /// Prune, sort by opacity, hand the cloud over.
  GaussianCloud finish();

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

struct SplatGpuAvailability {
  bool available = false;
  std::string device_name;
  std::string reason;
};

SplatGpuAvailability splatGpuAvailability();

/// Train `cloud` in place on the GPU (VulkanSplatTrainer.h:27-37 semantics):
/// false with `error` set if the device cannot do it; a callback returning
/// false cancels and the partially trained cloud is still written back; the
/// result is pruned and sorted by opacity.
bool trainSplatsOnGpu(GaussianCloud* cloud, const std::vector<SplatTrainingView>& views,
                      const SplatTrainingConfig& config, const SplatTrainingCallback& callback, std::string* error);

/// Free functions of the engine's rasteriser on the device, one call each
/// (SplatRaster.h:123-142).
SplatImage renderSplats(const GaussianCloud& cloud, const SplatView& view, const RasterOptions& options,
                        std::string* error = nullptr);

/// initializeFromPointCloud (GaussianCloud.h:90-93), the nearest-neighbour
/// spacing and the per-point fill on the device. `colors` empty or one per
/// point. An empty cloud (and `error` set) if the device cannot do it.
GaussianCloud initializeFromPointCloud(const std::vector<std::array<float, 3>>& points,
                                       const std::vector<std::array<uint8_t, 3>>& colors, int sh_degree = 0,
                                       float initial_opacity = 0.1f, std::string* error = nullptr);

/// writeGaussianPly / readGaussianPly (GaussianCloud.h:102-103): the 3DGS
/// binary PLY, rows packed and unpacked on the device, bytes as the engine
/// writes them.
bool writeGaussianPly(const std::string& path, const GaussianCloud& cloud, std::string* error = nullptr);
bool readGaussianPly(const std::string& path, GaussianCloud* out, std::string* error = nullptr);

/// writeSplat / readSplat (GaussianCloud.h:117-118): antimatter15's 32-byte
/// records, quantised and decoded on the device. Higher SH bands are dropped
/// on write; a read cloud has SH degree 0.
bool writeSplat(const std::string& path, const GaussianCloud& cloud, std::string* error = nullptr);
bool readSplat(const std::string& path, GaussianCloud* out, std::string* error = nullptr);

/// writeSpz / readSpz (GaussianCloud.h:130-131): Niantic's .spz v2. The
/// attribute planes are quantised and decoded on the device; the gzip stream
/// is deflated and inflated on the host with the engine's zlib settings.
bool writeSpz(const std::string& path, const GaussianCloud& cloud, int fractional_bits = 12,
              std::string* error = nullptr);
bool readSpz(const std::string& path, GaussianCloud* out, std::string* error = nullptr);

/// The device half of the splat stage (Pipeline.cpp:1844-2010) from the
/// point cloud on: seed at most a quarter of the gaussian budget from the
/// points (every step-th point), derive the schedule from the iteration count
/// as the stage does, train on the GPU under the stage's cancel / time-budget
/// callback rules, and write the field to each of `ply_paths`. Returns false
/// with `error` set when the device cannot do it -- the caller then falls back
/// to the CPU trainer as the stage does today -- or when cancelled.
struct SplatStageSettings {
  int iterations = 7000;
  int sh_degree = 3;
  int max_gaussians = 600000;
  long long time_budget_millis = 0;  // 0: none
  int batch_views = 1;
  int device = 0;
};
struct SplatStageResult {
  GaussianCloud cloud;
  std::string backend;  // "hip"
  int iterations_done = 0;
  bool ran_out_of_time = false;
  bool cancelled = false;
  std::string note;
};
/// The stage's seed set (Pipeline.cpp:1860-1876): with more points than a
/// quarter of the gaussian budget, every step-th point, step = ceil(points /
/// budget), a missing colour filled with grey (128, 128, 128); otherwise the
/// points and colours as given.
void splatStageSeedPoints(const std::vector<std::array<float, 3>>& points,
                          const std::vector<std::array<uint8_t, 3>>& colors, int max_gaussians,
                          std::vector<std::array<float, 3>>* seed_points,
                          std::vector<std::array<uint8_t, 3>>* seed_colors);
/// The stage's training configuration (Pipeline.cpp:1921-1926): the
/// defaults, with the iteration count (at least 1), SH degree and budget of
/// `settings`, densify until half way, opacity reset every third and an SH
/// band every eighth of the run (each at least 1).
SplatTrainingConfig splatStageTrainingConfig(const SplatStageSettings& settings);
bool runSplatStageOnGpu(const std::vector<std::array<float, 3>>& points,
                        const std::vector<std::array<uint8_t, 3>>& colors, std::vector<SplatTrainingView> views,
                        const SplatStageSettings& settings, const std::function<bool()>& cancelled,
                        const std::vector<std::string>& ply_paths, SplatStageResult* result, std::string* error);

}  // namespace parallax_hip
