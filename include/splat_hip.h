/*
 * splat_hip.h -- C ABI of the MI355X Gaussian-splat training step.
 *
 * A thin, POD-only boundary (no C++ or torch types) between host code and the
 * HIP kernels in libparallax_splat.so. It replaces the engine's device path
 * for splat training and exposes, one call each, the stages of the engine's
 * CPU trainer so they can be run and checked separately:
 *
 *   reference interface                                  entry point here
 *   -------------------                                  ----------------
 *   splatGpuAvailability()  VulkanSplatTrainer.h:15-25   splat_availability
 *   Context::create         vulkan/VulkanContext.cpp:65  splat_create / splat_destroy
 *   GpuTrainer::setUp/allocate                           splat_set_capacity
 *     VulkanSplatTrainer.cpp:290-405
 *   HostParameters upload / download                     splat_upload_cloud / splat_download_cloud
 *     VulkanSplatTrainer.cpp:147-195, 621-629, 863-866
 *   SplatTrainingView images  SplatTrainer.h:20-23       splat_upload_views
 *   projectSplats + rasterizeSplats  SplatRaster.h:123-131
 *                                                        splat_render_forward (+ splat_download_image,
 *                                                        splat_download_projection, splat_download_tiles)
 *   photometricLoss  SplatRaster.h:153-154                splat_loss
 *   rasterizeSplatsBackward  SplatRaster.h:138-142       splat_render_backward (+ splat_set_image_gradient,
 *                                                        splat_download_gradients, splat_zero_gradients)
 *   SplatTrainer applyGradients  SplatTrainer.cpp:218-256 splat_adam_step
 *   SplatBackend::renderAndAccumulate + Adam              splat_train_step (+ splat_set_graph_mode)
 *     SplatTrainer.cpp:337-349
 *   densifyAndPrune + remapState                          splat_densify
 *     SplatDensify.h:55-57, SplatTrainer.cpp:267-299
 *   resetOpacity  SplatTrainer.cpp:301-309                splat_reset_opacity
 *   accumulated statistics  SplatTrainer.cpp:153-154      splat_download_stats
 *   initializeFromPointCloud  GaussianCloud.h:90-93        splat_init_from_points
 *   writeGaussianPly / readGaussianPly  GaussianCloud.h:102-103
 *                                                        splat_write_ply / splat_read_ply (+ splat_ply_info)
 *   writeSplat / readSplat      GaussianCloud.h:117-118    splat_write_splat / splat_read_splat
 *                                                        (+ splat_splat_info)
 *   writeSpz / readSpz          GaussianCloud.h:130-131    splat_write_spz / splat_read_spz (+ splat_spz_info)
 *
 * Conventions follow the engine: results are returned through out-parameters,
 * every call returns a status and never throws, and on failure
 * splat_last_error() holds a sentence for a person to read
 * (engine/native/cpp CLAUDE.md:130-131, VulkanSplatTrainer.h:27-37). One
 * context is used by one host thread at a time.
 *
 * Device data layout (HBM), all float32 unless noted:
 *   parameters, gradients and both Adam moments each live in one buffer laid
 *   out by capacity Cap: [pos 3Cap | log_scale 3Cap | rot 4Cap (w,x,y,z) |
 *   opacity Cap | dc 3Cap | rest rs*Cap], rest channel-major per gaussian.
 */
#ifndef PARALLAX_SPLAT_HIP_H
#define PARALLAX_SPLAT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SPLAT_OK = 0,
  SPLAT_ERROR_INVALID_ARGUMENT = 1,
  SPLAT_ERROR_NO_DEVICE = 2,
  SPLAT_ERROR_OUT_OF_MEMORY = 3,
  SPLAT_ERROR_CAPACITY = 4,
  SPLAT_ERROR_DEVICE = 5,
  SPLAT_ERROR_STATE = 6
} splat_status;

typedef struct splat_context splat_context;

/* SplatView (SplatRaster.h:26-29): camera-from-world pose p_cam = R p + t with
 * R row-major, pinhole intrinsics (CameraIntrinsics, Types.h:30-94). */
typedef struct {
  double R[9];
  double t[3];
  double fx, fy, cx, cy;
  int32_t width, height;
} splat_view;

/* RasterOptions (SplatRaster.h:31-50) plus the launch mode. */
typedef struct {
  float background[3];
  int32_t sh_degree;
  float min_alpha;             /* default 1/255 */
  float transmittance_cutoff;  /* default 1e-4 */
  float screen_dilation;       /* default 0.3 */
  int32_t deterministic;       /* 1: fixed-order gradient reduction, bitwise repeatable */
} splat_raster_options;

/* A cloud in host memory, structure of arrays (Gaussian, GaussianCloud.h:41-72). */
typedef struct {
  int32_t n;
  int32_t sh_degree;
  int32_t rest_stride; /* 3 * shRestPerChannel(sh_degree) */
  float* pos;          /* 3n */
  float* log_scale;    /* 3n */
  float* rot;          /* 4n, (w, x, y, z) */
  float* opacity;      /* n, logits */
  float* dc;           /* 3n */
  float* rest;         /* rest_stride * n, channel-major per gaussian */
} splat_cloud;

/* SplatGradients (SplatRaster.h:87-114) in host memory; any pointer may be null. */
typedef struct {
  float* pos;
  float* log_scale;
  float* rot; /* (w, x, y, z) */
  float* opacity;
  float* dc;
  float* rest;
  float* screen_gradient;
  int32_t* visible_count;
} splat_gradients;

/* The scalars of one Adam step (SplatTrainer.cpp:56-65, 218-256). */
typedef struct {
  float lr_position, lr_log_scale, lr_rotation, lr_opacity, lr_dc, lr_rest;
  float beta1, beta2, epsilon;
  float correction1; /* 1 - beta1^t */
  float correction2; /* 1 - beta2^t */
} splat_adam_params;

/* SplatTrainingConfig fields used by densify (SplatTrainer.h:61-89). */
typedef struct {
  float densify_grad_threshold;
  float percent_dense;
  float split_scale_divider;
  int32_t opacity_reset_interval;
  float prune_opacity;
  float prune_scale_fraction;
  int32_t max_gaussians;
} splat_densify_params;

/* ---- availability, lifecycle ---- */
splat_status splat_availability(char* device_name, size_t name_len, char* reason, size_t reason_len);
splat_status splat_create(int32_t device, splat_context** out);
void splat_destroy(splat_context* ctx);
const char* splat_last_error(const splat_context* ctx);
void splat_default_options(splat_raster_options* options);

/* ---- scene ---- */
/* Reserve device memory for `max_gaussians` at `sh_degree`; clears the cloud. */
splat_status splat_set_capacity(splat_context* ctx, int32_t max_gaussians, int32_t sh_degree);
/* Upload a cloud (parameters; gradients, moments and statistics are zeroed). */
splat_status splat_upload_cloud(splat_context* ctx, const splat_cloud* cloud);
/* Download into arrays sized for splat_count(). */
splat_status splat_download_cloud(splat_context* ctx, splat_cloud* out);
int32_t splat_count(const splat_context* ctx);
/* Camera set and (optional) target images, 3*W*H floats each, linear RGB. */
splat_status splat_upload_views(splat_context* ctx, int32_t n, const splat_view* views,
                                const float* const* images);

/* ---- forward / backward pair ---- */
/* Project, bin, sort and composite the listed views in one batch. The batch
 * keeps its sorted lists, final transmittance and last contributor per pixel
 * for splat_render_backward; nothing per (splat, pixel) is stored. */
splat_status splat_render_forward(splat_context* ctx, int32_t n_views, const int32_t* view_ids,
                                  const splat_raster_options* options);
/* Image of batch slot `slot` (3*W*H rgb, W*H transmittance, W*H last
 * contributor relative to the tile's list); any pointer may be null. */
splat_status splat_download_image(splat_context* ctx, int32_t slot, float* rgb, float* transmittance,
                                  int32_t* last_contributor);
/* Per projected splat of batch slot `slot` in depth order (gaussian index,
 * mean, conic, depth, radius, colour, opacity, clamp flags, 16x16 tile rect);
 * returns the count through *count. Arrays sized for splat_count(). */
splat_status splat_download_projection(splat_context* ctx, int32_t slot, int32_t* count, int32_t* index,
                                       float* mean2, float* conic3, float* depth, float* radius,
                                       float* color3, float* opacity, int32_t* clamped3, int32_t* rect4);
/* Per-tile lists of batch slot `slot` over 16x16 tiles: offsets (tiles+1),
 * entries (gaussian indices); *entry_count set; entries may be null to query. */
splat_status splat_download_tiles(splat_context* ctx, int32_t slot, int32_t* offsets, int32_t* entries,
                                  int32_t* entry_count);
/* Photometric loss (1-w) L1 + w (1-SSIM) of every batch view against its
 * target; writes dL/drgb for the backward and the per-view losses. */
splat_status splat_loss(splat_context* ctx, float ssim_weight, float* losses);
/* No counterpart in the engine: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). splat_loss of one image held in the caller's device
 * memory: rgb and target (3*W*H floats each, interleaved linear RGB) in,
 * dL_drgb (3*W*H) and the loss (one float) out, all device pointers. Every
 * intermediate lives in one fixed device array of the library (about 76 MB
 * per device, held until the process ends), so the call allocates nothing;
 * W*H <= 2,097,152 (1920x1080 fits) and each side <= 8192, else
 * SPLAT_ERROR_CAPACITY. Results are bit-identical to splat_loss on the same
 * image. Enqueued on the context's stream with no host synchronisation;
 * calls that overlap on one device (any context, any stream) race on the
 * shared array. */
splat_status splat_loss_noalloc(splat_context* ctx, const float* rgb, const float* target, int32_t width,
                                int32_t height, float ssim_weight, float* dL_drgb, float* loss);
/* Replace dL/drgb of batch slot `slot` (3*W*H) -- for objectives other than
 * the photometric loss. */
splat_status splat_set_image_gradient(splat_context* ctx, int32_t slot, const float* dL_drgb);
/* dL/drgb of batch slot `slot` as the loss (or the caller) left it (3*W*H). */
splat_status splat_download_image_gradient(splat_context* ctx, int32_t slot, float* dL_drgb);
/* Adjoint of the last forward: adds parameter gradients, screen gradients and
 * visible counts into the device gradient buffer (rasterizeSplatsBackward
 * semantics: gradients are added, not overwritten). */
splat_status splat_render_backward(splat_context* ctx);
splat_status splat_download_gradients(splat_context* ctx, splat_gradients* out);
splat_status splat_zero_gradients(splat_context* ctx);

/* ---- optimiser and fused step ---- */
/* One Adam step over every parameter; gradients are zeroed after use and the
 * screen gradients / visible counts are added into the densify statistics. */
splat_status splat_adam_step(splat_context* ctx, const splat_adam_params* params);
/* renderAndAccumulate over the listed views + Adam: forward, loss, adjoint,
 * Adam. losses: n_views floats (nullable). */
splat_status splat_train_step(splat_context* ctx, int32_t n_views, const int32_t* view_ids,
                              const splat_raster_options* options, float ssim_weight,
                              const splat_adam_params* adam, float* losses);
/* enabled != 0: splat_train_step replays the whole step (forward, loss,
 * adjoint, Adam) as a captured HIP graph, one per batch shape and options,
 * with no host synchronisation inside the step. Pair buffers then run at
 * capacities learnt from earlier steps (the first step runs plain); a step
 * that outgrows them is detected on the device and redone. Results equal the
 * plain launch sequence's. No counterpart in the engine, whose Vulkan loop
 * records its command buffers per iteration (VulkanSplatTrainer.cpp:748-861).
 * Default: off. */
splat_status splat_set_graph_mode(splat_context* ctx, int32_t enabled);

/* ---- compositing launch shape (performance only) ---- */
/* How the forward and adjoint compositing kernels are launched. Every choice
 * gives the same images, and the same gradients in deterministic mode. */
typedef struct splat_raster_config {
  int32_t forward_shape;   /* workgroup of the forward: 0 16x16 pixels (one tile,
                              four 8x8 waves), 1 8x8 (one wave), 2 32x8 (the top
                              or bottom halves of two adjacent tiles) */
  int32_t forward_queue;   /* 1: persistent grid pulling tiles, longest list
                              first, from a work queue; 0: one workgroup per unit */
  int32_t backward_batch;  /* splats staged per adjoint batch: 64 or 128 */
  int32_t backward_queue;  /* as forward_queue, for the adjoint */
} splat_raster_config;
/* enabled != 0: the first batch of each new shape (the resolutions of its
 * views) times every launch shape on that batch and keeps the fastest for the
 * shape. Default: off (16x16, 128, no queue). No counterpart in the engine. */
splat_status splat_set_autotune(splat_context* ctx, int32_t enabled);
/* Use `config` for every batch (NULL: back to default / tuned). */
splat_status splat_set_raster_config(splat_context* ctx, const splat_raster_config* config);
/* The launch shape of the current batch. */
splat_status splat_get_raster_config(const splat_context* ctx, splat_raster_config* out);

/* ---- adaptive density ---- */
/* densifyAndPrune + remapState on the device. rng_state is advanced exactly
 * as the CPU trainer advances it. source_index (nullable, sized for the new
 * count) receives the old index or -1; counts = {cloned, split, pruned}. */
splat_status splat_densify(splat_context* ctx, const splat_densify_params* params, float extent,
                           int32_t iteration, uint32_t* rng_state, int32_t* source_index, int32_t* counts);
/* resetOpacity: logit = min(logit, ceiling); opacity moments zeroed. */
splat_status splat_reset_opacity(splat_context* ctx, float ceiling);
/* Accumulated densify statistics since the last densify (or upload). */
splat_status splat_download_stats(splat_context* ctx, float* screen_gradient, int32_t* visible_count);
/* Restore accumulated densify statistics (resume; n = splat_count() each). */
splat_status splat_upload_stats(splat_context* ctx, const float* screen_gradient, const int32_t* visible_count);
/* Adam moments, group-concatenated by count: [pos 3n | log_scale 3n | rot 4n
 * | opacity n | dc 3n | rest rs n] (snapshot / resume of the optimiser). */
splat_status splat_download_moments(splat_context* ctx, float* m, float* v);
splat_status splat_upload_moments(splat_context* ctx, const float* m, const float* v);
/* Prune (opacity < min_opacity, non-finite or > 1e3 scale) and stable sort by
 * descending opacity logit (GaussianCloud.cpp:177-197), on the device. */
splat_status splat_prune_and_sort(splat_context* ctx, float min_opacity);

/* ---- seeding and the saved field (the stages either side of training) ---- */
/* initializeFromPointCloud (GaussianCloud.cpp:199-256) into the context: one
 * isotropic gaussian per point (xyz: 3n floats; rgb: 3n bytes or NULL) with
 * log sigma from the sampled nearest-neighbour spacing, identity rotation,
 * `initial_opacity` as a logit (the engine's default is 0.1) and DC from the
 * colour. The capacity (splat_set_capacity) must hold n points at sh_degree;
 * moments and statistics are cleared. log_sigma_out (nullable) receives the
 * shared log sigma. */
splat_status splat_init_from_points(splat_context* ctx, int32_t n, const float* xyz, const uint8_t* rgb,
                                    int32_t sh_degree, float initial_opacity, float* log_sigma_out);
/* writeGaussianPly (GaussianCloud.cpp:259-303): the context's cloud as a
 * binary little-endian 3DGS PLY file, byte for byte as the engine writes it. */
splat_status splat_write_ply(splat_context* ctx, const char* path);
/* The gaussian count and SH degree of a 3DGS PLY file (to size
 * splat_set_capacity before splat_read_ply); on failure the engine's message
 * goes to `error` (nullable). Needs no context. */
splat_status splat_ply_info(const char* path, int32_t* count, int32_t* sh_degree, char* error, size_t error_len);
/* readGaussianPly (GaussianCloud.cpp:305-403) into the context: properties
 * located by name, missing DC / scale read as zero, a missing rotation as the
 * identity, rotations of norm > 1e-6 renormalised; moments and statistics
 * cleared. The capacity must hold the count at the file's SH degree. */
splat_status splat_read_ply(splat_context* ctx, const char* path);

/* writeSplat (GaussianCloud.cpp:416-446): the context's cloud as .splat
 * records (32 bytes each: position, exp(log sigma), base colour and opacity
 * as bytes, the normalised quaternion as q * 128 + 128), quantised on the
 * device. The bytes are the engine's but for the sigma floats, which are the
 * correctly rounded exponential (the engine's libm expf is within 1 ulp). */
splat_status splat_write_splat(splat_context* ctx, const char* path);
/* The record count of a .splat file (its SH degree is always 0); on failure
 * the engine's message goes to `error` (nullable). Needs no context. */
splat_status splat_splat_info(const char* path, int32_t* count, char* error, size_t error_len);
/* readSplat (GaussianCloud.cpp:448-493) into the context, decoded on the
 * device. The capacity must be at SH degree 0 and hold the count. */
splat_status splat_read_splat(splat_context* ctx, const char* path);
/* writeSpz (GaussianCloud.cpp:497-560): .spz v2 -- the attribute planes are
 * quantised on the device, then the host deflates them with the engine's
 * zlib settings; the file is byte for byte the engine's. `fractional_bits`
 * is clamped to 0..20 (the engine's default is 12); `antialiased` sets the
 * header flag. */
splat_status splat_write_spz(splat_context* ctx, const char* path, int32_t fractional_bits, int32_t antialiased);
/* The count, SH degree and antialiased flag of a .spz file, after readSpz's
 * checks of the whole inflated stream; on failure the engine's message goes to
 * `error` (nullable). Needs no context. */
splat_status splat_spz_info(const char* path, int32_t* count, int32_t* sh_degree, int32_t* antialiased, char* error,
                            size_t error_len);
/* readSpz (GaussianCloud.cpp:562-641) into the context: inflated on the host,
 * decoded on the device. The capacity must hold the count at the file's SH
 * degree; `antialiased` (nullable) receives the header flag. */
splat_status splat_read_spz(splat_context* ctx, const char* path, int32_t* antialiased);

/* ---- measurement ---- */
/* Per-kernel timing with HIP events on the context's stream: enable, then
 * read the accumulated milliseconds and launch counts by kernel name. */
splat_status splat_set_timing(splat_context* ctx, int32_t enabled);
int32_t splat_timing_count(const splat_context* ctx);
splat_status splat_timing_entry(splat_context* ctx, int32_t i, const char** name, double* total_ms,
                                int64_t* launches);
splat_status splat_reset_timing(splat_context* ctx);
/* Number of (tile, splat) pairs of the last forward batch. */
int64_t splat_last_pair_count(const splat_context* ctx);
/* Device memory the context holds (its buffers only grow, so this is also
 * its peak), and the device's used memory as the runtime reports it. */
splat_status splat_memory(splat_context* ctx, uint64_t* context_bytes, uint64_t* device_used_bytes);
splat_status splat_synchronize(splat_context* ctx);

#ifdef __cplusplus
}
#endif
#endif
