// Engine-side adapter between the engine's splat types (namespace parallax,
// Eigen vectors and quaternions) and the HIP facade's (namespace
// parallax_hip, std::array), plus trainSplatsOnHip: trainSplatsOnGpu's
// signature and contract (engine/native/cpp/include/parallax/splat/
// VulkanSplatTrainer.h:27-37) over the HIP trainer, so the splat stage's
// call (Pipeline.cpp:1961) switches by name:
//
//   #include <parallax_hip/engine_adapter.hpp>
//   ...
//   if (parallax_hip::engine::trainSplatsOnHip(&cloud, views, training_config, callback, &gpu_error)) {
//
// For engine code only: it includes the engine's headers and Eigen. Every
// conversion is exact (the same floats and doubles, in the facade's order);
// the camera's distortion terms k1, k2 are not carried, as the engine's
// rasteriser projects with fx, fy, cx, cy alone (SplatRaster.cpp:201-263).
#pragma once

#include <parallax/splat/VulkanSplatTrainer.h>
#include <parallax_hip/splat_trainer.hpp>

#include <string>
#include <vector>

namespace parallax_hip {
namespace engine {

// ---- Gaussian, GaussianCloud (GaussianCloud.h:41-72) ----
inline Gaussian toHip(const parallax::Gaussian& g) {
  Gaussian h;
  h.position = {g.position.x(), g.position.y(), g.position.z()};
  h.log_scale = {g.log_scale.x(), g.log_scale.y(), g.log_scale.z()};
  h.rotation = {g.rotation.w(), g.rotation.x(), g.rotation.y(), g.rotation.z()};
  h.opacity_logit = g.opacity_logit;
  h.dc = g.dc;
  h.sh_rest = g.sh_rest;
  return h;
}

inline parallax::Gaussian fromHip(const Gaussian& h) {
  parallax::Gaussian g;
  g.position = parallax::Vec3f(h.position[0], h.position[1], h.position[2]);
  g.log_scale = parallax::Vec3f(h.log_scale[0], h.log_scale[1], h.log_scale[2]);
  g.rotation = Eigen::Quaternionf(h.rotation[0], h.rotation[1], h.rotation[2], h.rotation[3]);
  g.opacity_logit = h.opacity_logit;
  g.dc = h.dc;
  g.sh_rest = h.sh_rest;
  return g;
}

inline GaussianCloud toHip(const parallax::GaussianCloud& c) {
  GaussianCloud h;
  h.sh_degree = c.sh_degree;
  h.antialiased = c.antialiased;
  h.gaussians.reserve(c.gaussians.size());
  for (const parallax::Gaussian& g : c.gaussians) h.gaussians.push_back(toHip(g));
  return h;
}

inline parallax::GaussianCloud fromHip(const GaussianCloud& h) {
  parallax::GaussianCloud c;
  c.sh_degree = h.sh_degree;
  c.antialiased = h.antialiased;
  c.gaussians.reserve(h.gaussians.size());
  for (const Gaussian& g : h.gaussians) c.gaussians.push_back(fromHip(g));
  return c;
}

// ---- SplatView (SplatRaster.h:26-29), CameraIntrinsics (Types.h:30-94), Pose (SE3.h:36-49) ----
inline SplatView toHip(const parallax::SplatView& v) {
  SplatView h;
  h.intrinsics.fx = v.intrinsics.fx;
  h.intrinsics.fy = v.intrinsics.fy;
  h.intrinsics.cx = v.intrinsics.cx;
  h.intrinsics.cy = v.intrinsics.cy;
  h.intrinsics.width = v.intrinsics.width;
  h.intrinsics.height = v.intrinsics.height;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) h.camera_from_world.R[static_cast<size_t>(3 * r + c)] = v.camera_from_world.R(r, c);
    h.camera_from_world.t[static_cast<size_t>(r)] = v.camera_from_world.t[r];
  }
  return h;
}

inline parallax::SplatView fromHip(const SplatView& h) {
  parallax::SplatView v;
  v.intrinsics.fx = h.intrinsics.fx;
  v.intrinsics.fy = h.intrinsics.fy;
  v.intrinsics.cx = h.intrinsics.cx;
  v.intrinsics.cy = h.intrinsics.cy;
  v.intrinsics.width = h.intrinsics.width;
  v.intrinsics.height = h.intrinsics.height;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) v.camera_from_world.R(r, c) = h.camera_from_world.R[static_cast<size_t>(3 * r + c)];
    v.camera_from_world.t[r] = h.camera_from_world.t[static_cast<size_t>(r)];
  }
  return v;
}

// ---- SplatTrainingView, SplatTrainingConfig, SplatTrainingProgress (SplatTrainer.h:20-108) ----
inline SplatTrainingView toHip(const parallax::SplatTrainingView& v) { return {toHip(v.view), v.image}; }

inline parallax::SplatTrainingView fromHip(const SplatTrainingView& h) {
  parallax::SplatTrainingView v;
  v.view = fromHip(h.view);
  v.image = h.image;
  return v;
}

// The engine's fields; the facade's own (batch_views, deterministic, device,
// graph_step, autotune) keep their defaults: one view per step, the engine's
// schedule exactly.
inline SplatTrainingConfig toHip(const parallax::SplatTrainingConfig& c) {
  SplatTrainingConfig h;
  h.iterations = c.iterations;
  h.position_lr_initial = c.position_lr_initial;
  h.position_lr_final = c.position_lr_final;
  h.dc_lr = c.dc_lr;
  h.sh_rest_lr_divider = c.sh_rest_lr_divider;
  h.opacity_lr = c.opacity_lr;
  h.scale_lr = c.scale_lr;
  h.rotation_lr = c.rotation_lr;
  h.adam_beta1 = c.adam_beta1;
  h.adam_beta2 = c.adam_beta2;
  h.adam_epsilon = c.adam_epsilon;
  h.ssim_weight = c.ssim_weight;
  h.sh_degree = c.sh_degree;
  h.sh_increase_interval = c.sh_increase_interval;
  h.densify_from_iteration = c.densify_from_iteration;
  h.densify_until_iteration = c.densify_until_iteration;
  h.densify_interval = c.densify_interval;
  h.densify_grad_threshold = c.densify_grad_threshold;
  h.percent_dense = c.percent_dense;
  h.split_scale_divider = c.split_scale_divider;
  h.opacity_reset_interval = c.opacity_reset_interval;
  h.opacity_reset_value = c.opacity_reset_value;
  h.prune_opacity = c.prune_opacity;
  h.prune_scale_fraction = c.prune_scale_fraction;
  h.max_gaussians = c.max_gaussians;
  h.seed = c.seed;
  return h;
}

inline parallax::SplatTrainingConfig fromHip(const SplatTrainingConfig& h) {
  parallax::SplatTrainingConfig c;
  c.iterations = h.iterations;
  c.position_lr_initial = h.position_lr_initial;
  c.position_lr_final = h.position_lr_final;
  c.dc_lr = h.dc_lr;
  c.sh_rest_lr_divider = h.sh_rest_lr_divider;
  c.opacity_lr = h.opacity_lr;
  c.scale_lr = h.scale_lr;
  c.rotation_lr = h.rotation_lr;
  c.adam_beta1 = h.adam_beta1;
  c.adam_beta2 = h.adam_beta2;
  c.adam_epsilon = h.adam_epsilon;
  c.ssim_weight = h.ssim_weight;
  c.sh_degree = h.sh_degree;
  c.sh_increase_interval = h.sh_increase_interval;
  c.densify_from_iteration = h.densify_from_iteration;
  c.densify_until_iteration = h.densify_until_iteration;
  c.densify_interval = h.densify_interval;
  c.densify_grad_threshold = h.densify_grad_threshold;
  c.percent_dense = h.percent_dense;
  c.split_scale_divider = h.split_scale_divider;
  c.opacity_reset_interval = h.opacity_reset_interval;
  c.opacity_reset_value = h.opacity_reset_value;
  c.prune_opacity = h.prune_opacity;
  c.prune_scale_fraction = h.prune_scale_fraction;
  c.max_gaussians = h.max_gaussians;
  c.seed = h.seed;
  return c;
}

inline parallax::SplatTrainingProgress fromHip(const SplatTrainingProgress& p) {
  parallax::SplatTrainingProgress q;
  q.iteration = p.iteration;
  q.total_iterations = p.total_iterations;
  q.gaussian_count = p.gaussian_count;
  q.loss = p.loss;
  q.smoothed_loss = p.smoothed_loss;
  return q;
}

// ---- splatGpuAvailability / trainSplatsOnGpu (VulkanSplatTrainer.h:15-37) ----
inline parallax::SplatGpuAvailability splatHipAvailability() {
  const SplatGpuAvailability a = splatGpuAvailability();
  parallax::SplatGpuAvailability out;
  out.available = a.available;
  out.device_name = a.device_name;
  out.reason = a.reason;
  return out;
}

// trainSplatsOnGpu's contract: false with `error` set when the device cannot
// train (the stage then falls back to the CPU trainer); a callback returning
// false cancels, and the partially trained field is still written back; the
// result is pruned and sorted by opacity. The cloud is converted back on
// every return: where the device never started, that reproduces it exactly.
inline bool trainSplatsOnHip(parallax::GaussianCloud* cloud, const std::vector<parallax::SplatTrainingView>& views,
                             const parallax::SplatTrainingConfig& config,
                             const parallax::SplatTrainingCallback& callback, std::string* error) {
  if (!cloud) {
    if (error) *error = "trainSplatsOnHip: null cloud";
    return false;
  }
  GaussianCloud hip_cloud = toHip(*cloud);
  std::vector<SplatTrainingView> hip_views;
  hip_views.reserve(views.size());
  for (const parallax::SplatTrainingView& v : views) hip_views.push_back(toHip(v));
  SplatTrainingCallback hip_callback;
  if (callback) hip_callback = [&callback](const SplatTrainingProgress& p) { return callback(fromHip(p)); };
  const bool ok = trainSplatsOnGpu(&hip_cloud, hip_views, toHip(config), hip_callback, error);
  *cloud = fromHip(hip_cloud);
  return ok;
}

}  // namespace engine
}  // namespace parallax_hip
