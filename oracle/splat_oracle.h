/*
 * CPU oracle for the Gaussian-splat training step -- TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch restatement of the engine's CPU trainer, used by the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark as the
 * checker of the HIP path. The product never links or calls it.
 *
 * It follows, operation for operation:
 *   engine/native/cpp/src/splat/SplatRaster.cpp   (projection, SH, binning,
 *       compositing, adjoint, L1 + D-SSIM loss)
 *   engine/native/cpp/src/splat/SplatTrainer.cpp  (Adam, learning rates,
 *       schedule, statistics, moment remap, opacity reset)
 *   engine/native/cpp/src/splat/SplatDensify.cpp  (clone / split / prune)
 *   engine/native/cpp/src/splat/GaussianCloud.cpp:177-197 (prune, sort)
 * Each function in splat_oracle.cpp cites the lines it restates.
 *
 * Two precisions:
 *   - float: the reference's arithmetic, compiled strict (no fast-math, no
 *     FMA contraction), sums evaluated left to right;
 *   - double: the same equations evaluated in FP64 while *replaying the
 *     discrete decisions of the float pass* (which splats are projected, the
 *     per-tile lists, every per-pixel skip / cap / termination, the SH and
 *     frustum clamps). This is the FP64 gradient reference and the function
 *     that finite differences are taken of.
 *
 * Pins (tests/test_oracle_*.py): golden vectors written by the reference's own
 * sources compiled into oracle/_ref (tests/golden/), the reference's property
 * tests re-expressed, FP64 finite differences of the frozen-decision forward,
 * brute-force tile enumeration and closed-form Adam.
 */
#ifndef PARALLAX_SPLAT_ORACLE_H
#define PARALLAX_SPLAT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Cloud in structure-of-arrays form. rot is (w, x, y, z) per gaussian; rest
 * is channel-major per gaussian (all R, then G, then B), rest_stride floats. */
typedef struct {
  int n;
  int sh_degree;
  int rest_stride;
  float* pos;       /* 3n */
  float* log_scale; /* 3n */
  float* rot;       /* 4n */
  float* opacity;   /* n, logits */
  float* dc;        /* 3n */
  float* rest;      /* rest_stride * n */
} orc_cloud;

typedef struct {
  int n;
  int sh_degree;
  int rest_stride;
  const double* pos;
  const double* log_scale;
  const double* rot;
  const double* opacity;
  const double* dc;
  const double* rest;
} orc_cloud_d;

/* SplatView: camera-from-world pose (p_cam = R p + t, R row-major) plus
 * pinhole intrinsics. */
typedef struct {
  double R[9];
  double t[3];
  double fx, fy, cx, cy;
  int width, height;
} orc_view;

/* RasterOptions (SplatRaster.h:31-50). */
typedef struct {
  float background[3];
  int sh_degree;
  float min_alpha;
  float transmittance_cutoff;
  float screen_dilation;
} orc_options;

/* SplatGradients (SplatRaster.h:87-114), accumulated into. */
typedef struct {
  float* pos;
  float* log_scale;
  float* rot;
  float* opacity;
  float* dc;
  float* rest;
  float* screen_gradient;
  int* visible_count;
} orc_grads;

typedef struct {
  double* pos;
  double* log_scale;
  double* rot;
  double* opacity;
  double* dc;
  double* rest;
  double* screen_gradient;
  int* visible_count;
} orc_grads_d;

/* SplatTrainingConfig (SplatTrainer.h:25-94). */
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
} orc_config;

void orc_default_options(orc_options* o);
void orc_default_config(orc_config* c);

/* ---- the deterministic exponential shared with the kernels ---- */
float orc_det_exp(float x);

/* ---- one rendered frame: projection, binning, compositing ---- */
typedef struct orc_frame orc_frame;

orc_frame* orc_frame_create(const orc_cloud* cloud, const orc_view* view, const orc_options* opt);
void orc_frame_destroy(orc_frame* f);
int orc_frame_projected_count(const orc_frame* f);
/* Per projected splat in depth order. Any pointer may be null. */
void orc_frame_get_projected(const orc_frame* f, int* index, float* mean2, float* cov2d3,
                             float* conic3, float* depth, float* radius, float* color3,
                             float* opacity, int* clamped3, int* rect4);
/* Tile grid of 16x16 logical tiles. Returns the entry count. */
int orc_frame_tile_grid(const orc_frame* f, int* cols, int* rows);
/* offsets: cols*rows+1; entries: gaussian index of each list entry. */
void orc_frame_get_tiles(const orc_frame* f, int* offsets, int* entries_gaussian);
/* rgb: 3*W*H; T: W*H; last: W*H list end (exclusive, relative to the tile's
 * first entry) of the last contributor. evals[0] = forward (pixel, entry)
 * evaluations, evals[1] = backward ones. */
void orc_frame_get_image(const orc_frame* f, float* rgb, float* T, int* last, int64_t* evals);

/* rasterizeSplatsBackward in float; adds into g (which must hold n entries). */
void orc_frame_backward(orc_frame* f, const orc_cloud* cloud, const float* dL_drgb,
                        orc_grads* g);
/* The same adjoint in FP64 with the float pass's decisions replayed. */
void orc_frame_backward_f64(orc_frame* f, const orc_cloud_d* cloud, const float* dL_drgb,
                            orc_grads_d* g);
/* FP64 forward with frozen decisions: writes the double image (3*W*H). */
void orc_frame_render_f64(orc_frame* f, const orc_cloud_d* cloud, double* rgb);
/* Central difference of sum(weights * rgb) with respect to one parameter of
 * one gaussian, using the FP64 frozen-decision forward restricted to the
 * tiles that gaussian touches. group: 0 pos, 1 log_scale, 2 rot, 3 opacity,
 * 4 dc, 5 rest; comp indexes within the group. Returns NaN if the gaussian
 * was not projected. */
double orc_frame_fd_f64(orc_frame* f, const orc_cloud_d* cloud, const double* weights, int gaussian,
                        int group, int comp, double h);

/* ---- loss (SplatRaster.cpp:580-738) ---- */
float orc_photometric_loss(const float* rendered, const float* target, int width, int height,
                           float ssim_weight, float* dL_drgb /* nullable */);
float orc_mean_ssim(const float* rendered, const float* target, int width, int height,
                    float* d_rendered /* nullable */);
void orc_blur(const float* src, int width, int height, float* dst);
/* SH colour for one gaussian along a unit direction (evaluateSh). */
void orc_evaluate_sh(const float* dc3, const float* rest, int rest_stride, int degree,
                     const float* direction3, float* color3, int* clamped3);

/* ---- trainer (SplatTrainer.cpp) ---- */
typedef struct orc_trainer orc_trainer;
orc_trainer* orc_trainer_create(const orc_cloud* initial, int n_views, const orc_view* views,
                                const float* const* images, const orc_config* config);
void orc_trainer_destroy(orc_trainer* t);
/* Runs one schedule step over `batch` views taken from the shuffled order
 * (batch 1 is SplatTrainer::step). Returns 0 once the schedule is done. */
int orc_trainer_step(orc_trainer* t, int batch);
/* As orc_trainer_step, with the views given explicitly. */
int orc_trainer_step_views(orc_trainer* t, int n, const int* view_ids);
int orc_trainer_iteration(const orc_trainer* t);
int orc_trainer_count(const orc_trainer* t);
int orc_trainer_sh_degree_in_use(const orc_trainer* t);
float orc_trainer_last_loss(const orc_trainer* t);
float orc_trainer_smoothed_loss(const orc_trainer* t);
float orc_trainer_extent(const orc_trainer* t);
int orc_trainer_last_views(const orc_trainer* t, int* out);
/* Copies the cloud out (arrays sized for count). */
void orc_trainer_get_cloud(const orc_trainer* t, orc_cloud* out);
/* moments: group-concatenated [pos 3n | log_scale 3n | rot 4n | opacity n |
 * dc 3n | rest rs n] for m and v. */
void orc_trainer_get_moments(const orc_trainer* t, float* m, float* v);
void orc_trainer_get_stats(const orc_trainer* t, float* screen_gradient, int* visible_count);
void orc_trainer_get_last_grads(const orc_trainer* t, orc_grads* g);
uint32_t orc_trainer_densify_rng(const orc_trainer* t);
/* Prune + sort by opacity (SplatTrainer::finish). Returns the count. */
int orc_trainer_finish(orc_trainer* t);

/* ---- standalone pieces for unit checks ---- */
/* One Adam step (applyGradients, SplatTrainer.cpp:218-256) over a cloud with
 * group-concatenated moments; iteration is the 0-based step index. */
void orc_adam_apply(orc_cloud* cloud, const orc_grads* g, float* m, float* v, int iteration,
                    float extent, const orc_config* c);
/* densifyAndPrune (SplatDensify.cpp:28-109). The cloud arrays must have room
 * for `capacity` gaussians; cloud->n is updated. source_index gets the new
 * count entries. counts = {cloned, split, pruned}. */
int orc_densify(orc_cloud* cloud, int capacity, const float* screen_gradient,
                const int* visible_count, const orc_config* c, float extent, int iteration,
                uint32_t* rng_state, int* source_index, int* counts);
float orc_camera_extent(int n_views, const orc_view* views);
/* Stable sort by descending opacity logit after pruning (finish). */
int orc_prune_and_sort(orc_cloud* cloud, float min_opacity, int* order_out);

#ifdef __cplusplus
}
#endif
#endif
