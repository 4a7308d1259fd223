// C shim over the reference engine's own splat sources -- TEST INFRASTRUCTURE.
//
// oracle/Makefile.ref compiles this file together with the reference's
// SplatRaster.cpp, SplatTrainer.cpp, SplatDensify.cpp, GaussianCloud.cpp and
// SE3.cpp, straight from where they lie under the reference checkout, into
// oracle/_ref/libparallax_ref.so. It converts the structure-of-arrays form the
// tests use into the engine's own `GaussianCloud` / `SplatView` types and
// calls the engine's functions unchanged. It is used to write the golden
// vectors under tests/golden/ and as the reference CPU trainer in the
// benchmark's CPU-baseline leg.
#include "parallax/splat/GaussianCloud.h"
#include "parallax/splat/SplatDensify.h"
#include "parallax/splat/SplatRaster.h"
#include "parallax/splat/SplatTrainer.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <string>
#include <cstdint>
#include <cstring>
#include <vector>

using namespace parallax;

extern "C" {

typedef struct {
  int n;
  int sh_degree;
  int rest_stride;
  float* pos;
  float* log_scale;
  float* rot;  // (w, x, y, z)
  float* opacity;
  float* dc;
  float* rest;
} ref_cloud;

typedef struct {
  double R[9];
  double t[3];
  double fx, fy, cx, cy;
  int width, height;
} ref_view;

typedef struct {
  float background[3];
  int sh_degree;
  float min_alpha;
  float transmittance_cutoff;
  float screen_dilation;
} ref_options;

typedef struct {
  int iterations;
  float position_lr_initial, position_lr_final, dc_lr, sh_rest_lr_divider;
  float opacity_lr, scale_lr, rotation_lr;
  float adam_beta1, adam_beta2, adam_epsilon;
  float ssim_weight;
  int sh_degree, sh_increase_interval;
  int densify_from_iteration, densify_until_iteration, densify_interval;
  float densify_grad_threshold, percent_dense, split_scale_divider;
  int opacity_reset_interval;
  float opacity_reset_value, prune_opacity, prune_scale_fraction;
  int max_gaussians;
  unsigned seed;
} ref_config;

}  // extern "C"

namespace {

GaussianCloud toCloud(const ref_cloud& c) {
  GaussianCloud cloud;
  cloud.sh_degree = c.sh_degree;
  cloud.gaussians.resize(static_cast<size_t>(c.n));
  for (int i = 0; i < c.n; ++i) {
    Gaussian& g = cloud.gaussians[static_cast<size_t>(i)];
    g.position = Vec3f(c.pos[3 * i], c.pos[3 * i + 1], c.pos[3 * i + 2]);
    g.log_scale = Vec3f(c.log_scale[3 * i], c.log_scale[3 * i + 1], c.log_scale[3 * i + 2]);
    g.rotation = Eigen::Quaternionf(c.rot[4 * i], c.rot[4 * i + 1], c.rot[4 * i + 2], c.rot[4 * i + 3]);
    g.opacity_logit = c.opacity[i];
    for (int k = 0; k < 3; ++k) g.dc[k] = c.dc[3 * i + k];
    g.sh_rest.assign(c.rest + static_cast<size_t>(c.rest_stride) * i,
                     c.rest + static_cast<size_t>(c.rest_stride) * (i + 1));
  }
  return cloud;
}

void fromCloud(const GaussianCloud& cloud, ref_cloud* c) {
  c->n = static_cast<int>(cloud.size());
  for (int i = 0; i < c->n; ++i) {
    const Gaussian& g = cloud.gaussians[static_cast<size_t>(i)];
    for (int k = 0; k < 3; ++k) {
      c->pos[3 * i + k] = g.position[k];
      c->log_scale[3 * i + k] = g.log_scale[k];
      c->dc[3 * i + k] = g.dc[k];
    }
    c->rot[4 * i] = g.rotation.w();
    c->rot[4 * i + 1] = g.rotation.x();
    c->rot[4 * i + 2] = g.rotation.y();
    c->rot[4 * i + 3] = g.rotation.z();
    c->opacity[i] = g.opacity_logit;
    for (size_t k = 0; k < g.sh_rest.size(); ++k) c->rest[static_cast<size_t>(c->rest_stride) * i + k] = g.sh_rest[k];
  }
}

SplatView toView(const ref_view& v) {
  SplatView view;
  view.intrinsics.fx = v.fx;
  view.intrinsics.fy = v.fy;
  view.intrinsics.cx = v.cx;
  view.intrinsics.cy = v.cy;
  view.intrinsics.width = v.width;
  view.intrinsics.height = v.height;
  Mat3 R;
  R << v.R[0], v.R[1], v.R[2], v.R[3], v.R[4], v.R[5], v.R[6], v.R[7], v.R[8];
  view.camera_from_world = Pose(R, Vec3(v.t[0], v.t[1], v.t[2]));
  return view;
}

RasterOptions toOptions(const ref_options& o) {
  RasterOptions r;
  r.background = {o.background[0], o.background[1], o.background[2]};
  r.sh_degree = o.sh_degree;
  r.min_alpha = o.min_alpha;
  r.transmittance_cutoff = o.transmittance_cutoff;
  r.screen_dilation = o.screen_dilation;
  return r;
}

SplatTrainingConfig toConfig(const ref_config& c) {
  SplatTrainingConfig s;
  s.iterations = c.iterations;
  s.position_lr_initial = c.position_lr_initial;
  s.position_lr_final = c.position_lr_final;
  s.dc_lr = c.dc_lr;
  s.sh_rest_lr_divider = c.sh_rest_lr_divider;
  s.opacity_lr = c.opacity_lr;
  s.scale_lr = c.scale_lr;
  s.rotation_lr = c.rotation_lr;
  s.adam_beta1 = c.adam_beta1;
  s.adam_beta2 = c.adam_beta2;
  s.adam_epsilon = c.adam_epsilon;
  s.ssim_weight = c.ssim_weight;
  s.sh_degree = c.sh_degree;
  s.sh_increase_interval = c.sh_increase_interval;
  s.densify_from_iteration = c.densify_from_iteration;
  s.densify_until_iteration = c.densify_until_iteration;
  s.densify_interval = c.densify_interval;
  s.densify_grad_threshold = c.densify_grad_threshold;
  s.percent_dense = c.percent_dense;
  s.split_scale_divider = c.split_scale_divider;
  s.opacity_reset_interval = c.opacity_reset_interval;
  s.opacity_reset_value = c.opacity_reset_value;
  s.prune_opacity = c.prune_opacity;
  s.prune_scale_fraction = c.prune_scale_fraction;
  s.max_gaussians = c.max_gaussians;
  s.seed = c.seed;
  return s;
}

}  // namespace

extern "C" {

/* projectSplats: returns the projected count; arrays sized for cloud->n. */
int ref_project(const ref_cloud* cloud, const ref_view* view, const ref_options* opt, int* index, float* mean2,
                float* cov2d3, float* conic3, float* depth, float* radius, float* color3, float* opacity,
                int* clamped3) {
  const std::vector<ProjectedSplat> p = projectSplats(toCloud(*cloud), toView(*view), toOptions(*opt));
  for (size_t s = 0; s < p.size(); ++s) {
    index[s] = p[s].index;
    mean2[2 * s] = p[s].mean.x();
    mean2[2 * s + 1] = p[s].mean.y();
    for (int k = 0; k < 3; ++k) {
      cov2d3[3 * s + k] = p[s].cov2d[k];
      conic3[3 * s + k] = p[s].conic[k];
      color3[3 * s + k] = p[s].color[k];
      clamped3[3 * s + k] = p[s].color_clamped[k] ? 1 : 0;
    }
    depth[s] = p[s].depth;
    radius[s] = p[s].radius;
    opacity[s] = p[s].opacity;
  }
  return static_cast<int>(p.size());
}

/* renderSplats. */
void ref_render(const ref_cloud* cloud, const ref_view* view, const ref_options* opt, float* rgb, float* T) {
  const SplatImage img = renderSplats(toCloud(*cloud), toView(*view), toOptions(*opt));
  std::memcpy(rgb, img.rgb.data(), img.rgb.size() * sizeof(float));
  std::memcpy(T, img.transmittance.data(), img.transmittance.size() * sizeof(float));
}

/* photometricLoss on a rendered image. */
float ref_photometric_loss(const float* rendered, const float* target, int width, int height, float ssim_weight,
                           float* dL_drgb) {
  SplatImage img;
  img.width = width;
  img.height = height;
  const size_t values = static_cast<size_t>(width) * height * 3;
  img.rgb.assign(rendered, rendered + values);
  img.transmittance.assign(static_cast<size_t>(width) * height, 0.f);
  std::vector<float> t(target, target + values), dL;
  const float loss = photometricLoss(img, t, ssim_weight, dL_drgb ? &dL : nullptr);
  if (dL_drgb) std::memcpy(dL_drgb, dL.data(), values * sizeof(float));
  return loss;
}

/* project + rasterize + rasterizeSplatsBackward with a given dL/drgb;
 * gradients are written (not accumulated) into the group arrays. */
void ref_backward(const ref_cloud* cloud, const ref_view* view, const ref_options* opt, const float* dL_drgb,
                  float* g_pos, float* g_ls, float* g_rot, float* g_op, float* g_dc, float* g_rest, float* g_sg,
                  int* g_vc) {
  const GaussianCloud cl = toCloud(*cloud);
  const SplatView v = toView(*view);
  const RasterOptions o = toOptions(*opt);
  const std::vector<ProjectedSplat> p = projectSplats(cl, v, o);
  const SplatImage img = rasterizeSplats(p, v, o);
  std::vector<float> dL(dL_drgb, dL_drgb + img.rgb.size());
  SplatGradients g;
  g.resize(cl.size(), cloud->rest_stride);
  rasterizeSplatsBackward(cl, v, o, p, img, dL, &g);
  for (int i = 0; i < cloud->n; ++i) {
    for (int k = 0; k < 3; ++k) {
      g_pos[3 * i + k] = g.position[static_cast<size_t>(i)][k];
      g_ls[3 * i + k] = g.log_scale[static_cast<size_t>(i)][k];
      g_dc[3 * i + k] = g.dc[static_cast<size_t>(i)][k];
    }
    for (int k = 0; k < 4; ++k) g_rot[4 * i + k] = g.rotation[static_cast<size_t>(i)][k];
    g_op[i] = g.opacity_logit[static_cast<size_t>(i)];
    g_sg[i] = g.screen_gradient[static_cast<size_t>(i)];
    g_vc[i] = g.visible_count[static_cast<size_t>(i)];
  }
  std::memcpy(g_rest, g.sh_rest.data(), g.sh_rest.size() * sizeof(float));
}

/* densifyAndPrune on a copy of `cloud`; the result is written to `out`
 * (arrays sized for `capacity`). Returns the new count. */
int ref_densify(const ref_cloud* cloud, const float* screen_gradient, const int* visible_count,
                const ref_config* config, float extent, int iteration, uint32_t* rng_state, ref_cloud* out,
                int* source_index, int* counts) {
  GaussianCloud cl = toCloud(*cloud);
  SplatDensifyStats stats;
  stats.screen_gradient.assign(screen_gradient, screen_gradient + cloud->n);
  stats.visible_count.assign(visible_count, visible_count + cloud->n);
  const SplatDensifyResult r = densifyAndPrune(&cl, stats, toConfig(*config), extent, iteration, rng_state);
  fromCloud(cl, out);
  std::memcpy(source_index, r.source_index.data(), r.source_index.size() * sizeof(int));
  counts[0] = r.cloned;
  counts[1] = r.split;
  counts[2] = r.pruned;
  return static_cast<int>(cl.size());
}

/* The reference trainer: construct, run `steps` steps, copy the cloud and the
 * per-step losses out. Returns the gaussian count after the last step. */
int ref_train(const ref_cloud* initial, int n_views, const ref_view* views, const float* const* images,
              const ref_config* config, int steps, float* losses, ref_cloud* out, int finish) {
  std::vector<SplatTrainingView> training;
  for (int k = 0; k < n_views; ++k) {
    SplatTrainingView tv;
    tv.view = toView(views[k]);
    tv.image.assign(images[k], images[k] + static_cast<size_t>(views[k].width) * views[k].height * 3);
    training.push_back(std::move(tv));
  }
  SplatTrainer trainer(toCloud(*initial), std::move(training), toConfig(*config));
  for (int s = 0; s < steps; ++s) {
    if (!trainer.step()) break;
    if (losses) losses[s] = trainer.lastLoss();
  }
  if (finish) {
    const GaussianCloud done = trainer.finish();
    fromCloud(done, out);
    return static_cast<int>(done.size());
  }
  fromCloud(trainer.cloud(), out);
  return static_cast<int>(trainer.cloud().size());
}

/* Wall time in seconds of `steps` reference training steps (fwd + loss +
 * bwd + Adam on one view each), the trainer's construction excluded. */
double ref_time_steps(const ref_cloud* initial, int n_views, const ref_view* views, const float* const* images,
                      const ref_config* config, int steps, float* last_loss) {
  std::vector<SplatTrainingView> training;
  for (int k = 0; k < n_views; ++k) {
    SplatTrainingView tv;
    tv.view = toView(views[k]);
    tv.image.assign(images[k], images[k] + static_cast<size_t>(views[k].width) * views[k].height * 3);
    training.push_back(std::move(tv));
  }
  SplatTrainer trainer(toCloud(*initial), std::move(training), toConfig(*config));
  const auto t0 = std::chrono::steady_clock::now();
  for (int s = 0; s < steps; ++s) trainer.step();
  const auto t1 = std::chrono::steady_clock::now();
  if (last_loss) *last_loss = trainer.lastLoss();
  return std::chrono::duration<double>(t1 - t0).count();
}

/* initializeFromPointCloud (GaussianCloud.cpp:199-256): n points (xyz),
 * optional 8-bit colours (rgb, nullable); `out` arrays sized for n. */
int ref_init_from_points(int n, const float* xyz, const unsigned char* rgb, int sh_degree, float initial_opacity,
                         ref_cloud* out) {
  std::vector<Vec3f> points(static_cast<size_t>(n));
  std::vector<std::array<uint8_t, 3>> colors;
  for (int i = 0; i < n; ++i) points[static_cast<size_t>(i)] = Vec3f(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  if (rgb) {
    colors.resize(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) colors[static_cast<size_t>(i)] = {rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]};
  }
  const GaussianCloud cloud = initializeFromPointCloud(points, colors, sh_degree, initial_opacity);
  out->sh_degree = cloud.sh_degree;
  fromCloud(cloud, out);
  return out->n;
}

/* The splat stage's seeding (Pipeline.cpp:1860-1876), restated here because
 * Pipeline.cpp builds only with the whole engine: with more than a quarter of
 * max_gaussians points, every step-th point (step = ceil(n / budget)), grey
 * where the point has no colour; then the engine's own
 * initializeFromPointCloud at the stage's SH degree and default opacity.
 * n_colors colours are given (0 <= n_colors <= n). `out` sized for n. */
int ref_stage_seed(int n, const float* xyz, int n_colors, const unsigned char* rgb, int max_gaussians, int sh_degree,
                   ref_cloud* out) {
  std::vector<Vec3f> points(static_cast<size_t>(n));
  std::vector<std::array<uint8_t, 3>> colors(static_cast<size_t>(n_colors));
  for (int i = 0; i < n; ++i) points[static_cast<size_t>(i)] = Vec3f(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  for (int i = 0; i < n_colors; ++i) colors[static_cast<size_t>(i)] = {rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]};
  const size_t seed_budget = std::max<size_t>(1, max_gaussians / 4);
  if (points.size() > seed_budget) {
    const size_t step = (points.size() + seed_budget - 1) / seed_budget;
    std::vector<Vec3f> kept_points;
    std::vector<std::array<uint8_t, 3>> kept_colors;
    for (size_t i = 0; i < points.size(); i += step) {
      kept_points.push_back(points[i]);
      kept_colors.push_back(i < colors.size() ? colors[i] : std::array<uint8_t, 3>{128, 128, 128});
    }
    points.swap(kept_points);
    colors.swap(kept_colors);
  }
  const GaussianCloud cloud = initializeFromPointCloud(points, colors, sh_degree);
  out->sh_degree = cloud.sh_degree;
  fromCloud(cloud, out);
  return out->n;
}

/* The splat stage's SplatTrainingConfig (Pipeline.cpp:1921-1926) on the
 * engine's own defaults (SplatTrainer.h:25-94), with validate()'s verdict in
 * `error` (empty when valid). Returns 1 when valid. */
int ref_stage_config(int iterations, int sh_degree, int max_gaussians, ref_config* out, char* error, int error_len) {
  SplatTrainingConfig c;
  c.iterations = std::max(1, iterations);
  c.sh_degree = sh_degree;
  c.max_gaussians = max_gaussians;
  c.densify_until_iteration = c.iterations / 2;
  c.opacity_reset_interval = std::max(1, c.iterations / 3);
  c.sh_increase_interval = std::max(1, c.iterations / 8);
  out->iterations = c.iterations;
  out->position_lr_initial = c.position_lr_initial;
  out->position_lr_final = c.position_lr_final;
  out->dc_lr = c.dc_lr;
  out->sh_rest_lr_divider = c.sh_rest_lr_divider;
  out->opacity_lr = c.opacity_lr;
  out->scale_lr = c.scale_lr;
  out->rotation_lr = c.rotation_lr;
  out->adam_beta1 = c.adam_beta1;
  out->adam_beta2 = c.adam_beta2;
  out->adam_epsilon = c.adam_epsilon;
  out->ssim_weight = c.ssim_weight;
  out->sh_degree = c.sh_degree;
  out->sh_increase_interval = c.sh_increase_interval;
  out->densify_from_iteration = c.densify_from_iteration;
  out->densify_until_iteration = c.densify_until_iteration;
  out->densify_interval = c.densify_interval;
  out->densify_grad_threshold = c.densify_grad_threshold;
  out->percent_dense = c.percent_dense;
  out->split_scale_divider = c.split_scale_divider;
  out->opacity_reset_interval = c.opacity_reset_interval;
  out->opacity_reset_value = c.opacity_reset_value;
  out->prune_opacity = c.prune_opacity;
  out->prune_scale_fraction = c.prune_scale_fraction;
  out->max_gaussians = c.max_gaussians;
  out->seed = c.seed;
  const std::string invalid = c.validate();
  if (error && error_len > 0) {
    std::strncpy(error, invalid.c_str(), static_cast<size_t>(error_len - 1));
    error[error_len - 1] = 0;
  }
  return invalid.empty() ? 1 : 0;
}

/* writeGaussianPly (GaussianCloud.cpp:259-303). Returns 1 on success. */
int ref_write_ply(const ref_cloud* cloud, const char* path) { return writeGaussianPly(path, toCloud(*cloud)) ? 1 : 0; }

/* readGaussianPly (GaussianCloud.cpp:305-403): first call with out == NULL
 * for the count and degree, then with arrays sized for them. Returns 1 on
 * success; the engine's error message goes to `error` (nullable). */
int ref_read_ply(const char* path, int* n, int* sh_degree, ref_cloud* out, char* error, int error_len) {
  GaussianCloud cloud;
  std::string message;
  if (!readGaussianPly(path, &cloud, &message)) {
    if (error && error_len > 0) {
      std::strncpy(error, message.c_str(), static_cast<size_t>(error_len - 1));
      error[error_len - 1] = 0;
    }
    return 0;
  }
  *n = static_cast<int>(cloud.size());
  *sh_degree = cloud.sh_degree;
  if (out) {
    out->sh_degree = cloud.sh_degree;
    fromCloud(cloud, out);
  }
  return 1;
}

/* writeSplat (GaussianCloud.cpp:416-446) and writeSpz (:497-560). Return 1 on
 * success. The cloud's `antialiased` flag is set from the argument. */
int ref_write_splat(const ref_cloud* cloud, const char* path) { return writeSplat(path, toCloud(*cloud)) ? 1 : 0; }

int ref_write_spz(const ref_cloud* cloud, const char* path, int fractional_bits, int antialiased) {
  GaussianCloud c = toCloud(*cloud);
  c.antialiased = antialiased != 0;
  return writeSpz(path, c, fractional_bits) ? 1 : 0;
}

/* readSplat (:448-493) and readSpz (:562-641), called as ref_read_ply is:
 * first with out == NULL for the count and degree. `antialiased` (nullable)
 * receives the .spz flag. */
static int read_with(bool (*reader)(const std::string&, GaussianCloud*, std::string*), const char* path, int* n,
                     int* sh_degree, int* antialiased, ref_cloud* out, char* error, int error_len) {
  GaussianCloud cloud;
  std::string message;
  if (!reader(path, &cloud, &message)) {
    if (error && error_len > 0) {
      std::strncpy(error, message.c_str(), static_cast<size_t>(error_len - 1));
      error[error_len - 1] = 0;
    }
    return 0;
  }
  *n = static_cast<int>(cloud.size());
  *sh_degree = cloud.sh_degree;
  if (antialiased) *antialiased = cloud.antialiased ? 1 : 0;
  if (out) {
    out->sh_degree = cloud.sh_degree;
    fromCloud(cloud, out);
  }
  return 1;
}

int ref_read_splat(const char* path, int* n, int* sh_degree, ref_cloud* out, char* error, int error_len) {
  return read_with(readSplat, path, n, sh_degree, nullptr, out, error, error_len);
}

int ref_read_spz(const char* path, int* n, int* sh_degree, int* antialiased, ref_cloud* out, char* error,
                 int error_len) {
  return read_with(readSpz, path, n, sh_degree, antialiased, out, error, error_len);
}

}  // extern "C"
