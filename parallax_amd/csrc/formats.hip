// The stages either side of the training step that touch the whole cloud:
// seeding it from a point cloud, and the 3DGS PLY file it is saved to and
// resumed from.
//
// initializeFromPointCloud (engine/native/cpp/src/splat/GaussianCloud.cpp:
// 199-256): the point spacing is the mean, over a strided sample of at most
// ~512 points, of each sample point's nearest other sample point (float
// distances, double sum); every gaussian is isotropic with log sigma =
// log(max(1e-4, spacing / 2)), identity rotation, the given opacity as a
// logit, and DC from the point colour. The O(s^2) nearest-neighbour pass and
// the per-point fill run on the device, writing straight into the capacity
// layout; the scalar steps between them (the double sum over <= ~1k values
// in sample order, and the two logarithms) run on the host, as the engine
// evaluates them.
//
// writeGaussianPly / readGaussianPly (GaussianCloud.cpp:259-403): the binary
// little-endian 3DGS layout x, y, z, nx, ny, nz, f_dc_0..2, f_rest_* (channel
// major), opacity, scale_0..2, rot_0..3 (w first). The device interleaves the
// structure-of-arrays parameters into PLY rows (and back, by property index,
// renormalising the quaternion as the engine does), so the host only moves
// one contiguous block and the header text. HBM-bound: 4 * (17 + rest) bytes
// read and written per gaussian.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/splat_hip.h"
#include "splat_context.h"
#include "splat_kernels.h"

using namespace splat;

namespace {

constexpr int kNnThreads = 256;

// Nearest other sample point of every sample point (points 0, stride,
// 2 stride, ... < n), as the engine computes it: the Vec3f difference, its
// squared norm summed x, y, z in order, the float square root; the minimum
// over the sample (+inf when the sample has one point).
__global__ __launch_bounds__(kNnThreads) void k_nn_sample(const float* __restrict__ xyz, long long n, long long stride,
                                                         int m, float* __restrict__ nearest) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  const long long i = (long long)s * stride;
  const float xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
  float best = __builtin_inff();
  for (int t = 0; t < m; ++t) {
    const long long j = (long long)t * stride;
    if (j == i) continue;
    const float dx = xi - xyz[3 * j], dy = yi - xyz[3 * j + 1], dz = zi - xyz[3 * j + 2];
    const float d = __builtin_sqrtf((dx * dx + dy * dy) + dz * dz);
    best = d < best ? d : best;
  }
  nearest[s] = best;
}

// One gaussian per point, written into the capacity layout.
__global__ __launch_bounds__(256) void k_init_points(const float* __restrict__ xyz, const uint8_t* __restrict__ rgb,
                                                     long long n, long long cap, int rs, float log_sigma, float logit,
                                                     float* __restrict__ P) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    P[3 * i + k] = xyz[3 * i + k];
    P[3 * cap + 3 * i + k] = log_sigma;
    // the viewer's DC-to-colour transform inverted (GaussianCloud.cpp:246-248)
    P[11 * cap + 3 * i + k] = rgb ? ((float)rgb[3 * i + k] / 255.0f - 0.5f) / kShC0 : 0.f;
  }
  P[6 * cap + 4 * i] = 1.f;
  P[6 * cap + 4 * i + 1] = 0.f;
  P[6 * cap + 4 * i + 2] = 0.f;
  P[6 * cap + 4 * i + 3] = 0.f;
  P[10 * cap + i] = logit;
  for (int k = 0; k < rs; ++k) P[14 * cap + (long long)rs * i + k] = 0.f;
}

// PLY rows (17 + rs floats each) from the capacity layout.
__global__ __launch_bounds__(256) void k_pack_ply(const float* __restrict__ P, long long n, long long cap, int rs,
                                                  float* __restrict__ rows) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float* r = rows + (long long)(17 + rs) * i;
  r[0] = P[3 * i];
  r[1] = P[3 * i + 1];
  r[2] = P[3 * i + 2];
  r[3] = r[4] = r[5] = 0.f;  // normals: unused by 3DGS, kept for positional readers
  r[6] = P[11 * cap + 3 * i];
  r[7] = P[11 * cap + 3 * i + 1];
  r[8] = P[11 * cap + 3 * i + 2];
  for (int k = 0; k < rs; ++k) r[9 + k] = P[14 * cap + (long long)rs * i + k];
  r[9 + rs] = P[10 * cap + i];
  r[10 + rs] = P[3 * cap + 3 * i];
  r[11 + rs] = P[3 * cap + 3 * i + 1];
  r[12 + rs] = P[3 * cap + 3 * i + 2];
  r[13 + rs] = P[6 * cap + 4 * i];  // rot_0 is w
  r[14 + rs] = P[6 * cap + 4 * i + 1];
  r[15 + rs] = P[6 * cap + 4 * i + 2];
  r[16 + rs] = P[6 * cap + 4 * i + 3];
}

// Where each field of a PLY row lives (-1: absent).
struct PlyIndex {
  int x, y, z, opacity;
  int dc[3], scale[3], rot[4];
  int rest_base;  // f_rest_k at rest_map[k]
};

// The capacity layout from PLY rows of `stride` floats (readGaussianPly's
// field rules: absent dc / scale read 0, absent rotation is the identity, a
// rotation of norm > 1e-6 is normalised -- the norm over (x, y, z, w) in that
// order, as the trainer's own renormalisation).
__global__ __launch_bounds__(256) void k_unpack_ply(const float* __restrict__ rows, long long n, int stride,
                                                    PlyIndex ix, const int* __restrict__ rest_map, long long cap,
                                                    int rs, float* __restrict__ P) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* r = rows + (long long)stride * i;
  P[3 * i] = r[ix.x];
  P[3 * i + 1] = r[ix.y];
  P[3 * i + 2] = r[ix.z];
  P[10 * cap + i] = r[ix.opacity];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    P[11 * cap + 3 * i + k] = ix.dc[k] >= 0 ? r[ix.dc[k]] : 0.f;
    P[3 * cap + 3 * i + k] = ix.scale[k] >= 0 ? r[ix.scale[k]] : 0.f;
  }
  float w = 1.f, x = 0.f, y = 0.f, z = 0.f;
  if (ix.rot[0] >= 0 && ix.rot[1] >= 0 && ix.rot[2] >= 0 && ix.rot[3] >= 0) {
    w = r[ix.rot[0]];
    x = r[ix.rot[1]];
    y = r[ix.rot[2]];
    z = r[ix.rot[3]];
    const float norm = __builtin_sqrtf(x * x + y * y + z * z + w * w);
    if (norm > 1e-6f) {
      w = w / norm;
      x = x / norm;
      y = y / norm;
      z = z / norm;
    }
  }
  P[6 * cap + 4 * i] = w;
  P[6 * cap + 4 * i + 1] = x;
  P[6 * cap + 4 * i + 2] = y;
  P[6 * cap + 4 * i + 3] = z;
  for (int k = 0; k < rs; ++k) P[14 * cap + (long long)rs * i + k] = r[rest_map[k]];
}

int rest_per_channel(int degree) { return degree <= 0 ? 0 : (degree == 1 ? 3 : (degree == 2 ? 8 : 15)); }

// A parsed 3DGS PLY header (readGaussianPly's rules, GaussianCloud.cpp:311-365).
struct PlyHeader {
  size_t count = 0;
  size_t body = 0;  // byte offset of the first row
  std::vector<std::string> properties;
  int degree = 0;
  int rest_count = 0;
};

bool parse_ply_header(const std::vector<uint8_t>& bytes, PlyHeader* h, std::string* error) {
  const std::string text(bytes.begin(), bytes.begin() + std::min<size_t>(bytes.size(), 65536));
  const size_t header_end = text.find("end_header\n");
  if (text.compare(0, 3, "ply") != 0 || header_end == std::string::npos) {
    *error = "not a PLY file";
    return false;
  }
  if (text.find("binary_little_endian") == std::string::npos) {
    *error = "only binary little-endian 3DGS PLY is supported";
    return false;
  }
  std::istringstream lines(text.substr(0, header_end));
  std::string line;
  while (std::getline(lines, line)) {
    std::istringstream ls(line);
    std::string keyword;
    ls >> keyword;
    if (keyword == "element") {
      std::string name;
      ls >> name >> h->count;
    } else if (keyword == "property") {
      std::string type, name;
      ls >> type >> name;
      if (type != "float" && type != "float32") {
        *error = "3DGS PLY properties must be float";
        return false;
      }
      h->properties.push_back(name);
    }
  }
  if (h->count == 0 || h->properties.empty()) {
    *error = "3DGS PLY has no vertices";
    return false;
  }
  auto index_of = [h](const std::string& name) {
    const auto it = std::find(h->properties.begin(), h->properties.end(), name);
    return it == h->properties.end() ? -1 : static_cast<int>(it - h->properties.begin());
  };
  if (index_of("x") < 0 || index_of("y") < 0 || index_of("z") < 0 || index_of("opacity") < 0) {
    *error = "3DGS PLY is missing core properties";
    return false;
  }
  while (index_of("f_rest_" + std::to_string(h->rest_count)) >= 0) ++h->rest_count;
  if (h->rest_count % 3 != 0) {
    *error = "f_rest count is not a multiple of 3";
    return false;
  }
  const int per_channel = h->rest_count / 3;
  h->degree = 0;
  for (int d = 0; d <= 3; ++d)
    if (rest_per_channel(d) == per_channel) h->degree = d;
  if (rest_per_channel(h->degree) != per_channel) {
    *error = "f_rest count does not correspond to an SH degree";
    return false;
  }
  h->body = header_end + std::strlen("end_header\n");
  const size_t stride = h->properties.size() * sizeof(float);
  if (bytes.size() < h->body + stride * h->count) {
    *error = "3DGS PLY body is truncated";
    return false;
  }
  return true;
}

bool read_file(const char* path, std::vector<uint8_t>* out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out->resize(size > 0 ? static_cast<size_t>(size) : 0);
  const bool ok = size >= 0 && std::fread(out->data(), 1, out->size(), f) == out->size();
  std::fclose(f);
  return ok;
}

splat_status fmt_fail(splat_context* ctx, splat_status st, const std::string& msg) {
  ctx->error = msg;
  return st;
}

#define CKF(expr, what)                                                \
  do {                                                                 \
    hipError_t _e = (expr);                                            \
    if (_e != hipSuccess) {                                            \
      ctx->error = std::string(what) + ": " + hipGetErrorString(_e);  \
      return SPLAT_ERROR_DEVICE;                                       \
    }                                                                  \
  } while (0)

// Scratch of the format calls: released at the end of each call.
struct Scratch {
  std::vector<void*> ptrs;
  ~Scratch() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  hipError_t alloc(void** p, size_t bytes) {
    const hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 4));
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
  }
};

// The bookkeeping splat_upload_cloud does after new parameters land.
splat_status reset_state(splat_context* ctx, int n, int degree) {
  const long long floats = ctx->cap * (14 + ctx->rs);
  CKF(launch_fill_f32(ctx->grads.f(), 0.f, floats, ctx->stream), "clear");
  CKF(launch_fill_f32(ctx->m.f(), 0.f, floats, ctx->stream), "clear");
  CKF(launch_fill_f32(ctx->v.f(), 0.f, floats, ctx->stream), "clear");
  CKF(launch_fill_f32(ctx->sg.f(), 0.f, ctx->cap, ctx->stream), "clear");
  CKF(launch_fill_u32(ctx->vc.u(), 0u, ctx->cap, ctx->stream), "clear");
  CKF(launch_fill_f32(ctx->stat_sg.f(), 0.f, ctx->cap, ctx->stream), "clear");
  CKF(launch_fill_u32(ctx->stat_vc.u(), 0u, ctx->cap, ctx->stream), "clear");
  CKF(hipStreamSynchronize(ctx->stream), "sync");
  ctx->n = n;
  ctx->cloud_degree = degree;
  ctx->forward_done = ctx->loss_done = false;
  return SPLAT_OK;
}

}  // namespace

extern "C" {

splat_status splat_init_from_points(splat_context* ctx, int32_t n, const float* xyz, const uint8_t* rgb,
                                    int32_t sh_degree, float initial_opacity, float* log_sigma_out) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  if (ctx->cap <= 0) return fmt_fail(ctx, SPLAT_ERROR_STATE, "init_from_points: no capacity reserved");
  if (n < 0 || (n > 0 && !xyz)) return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "init_from_points: bad points");
  const int degree = std::max(0, std::min(3, (int)sh_degree));
  if (3 * rest_per_channel(degree) != ctx->rs)
    return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT,
                    "init_from_points: the SH degree does not match the capacity's (splat_set_capacity)");
  if (n > ctx->cap) return fmt_fail(ctx, SPLAT_ERROR_CAPACITY, "init_from_points: more points than the capacity");
  CKF(hipSetDevice(ctx->device), "set device");
  if (n == 0) return reset_state(ctx, 0, degree);
  Scratch scratch;
  float* d_xyz = nullptr;
  uint8_t* d_rgb = nullptr;
  float* d_near = nullptr;
  CKF(scratch.alloc(reinterpret_cast<void**>(&d_xyz), sizeof(float) * 3 * (size_t)n), "alloc points");
  CKF(hipMemcpyAsync(d_xyz, xyz, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, ctx->stream), "upload");
  if (rgb) {
    CKF(scratch.alloc(reinterpret_cast<void**>(&d_rgb), 3 * (size_t)n), "alloc colours");
    CKF(hipMemcpyAsync(d_rgb, rgb, 3 * (size_t)n, hipMemcpyHostToDevice, ctx->stream), "upload");
  }
  // the strided sample of GaussianCloud.cpp:211-213
  const long long sample = std::min<long long>(n, 512);
  const long long stride = std::max<long long>(1, n / sample);
  const int m = (int)((n + stride - 1) / stride);
  CKF(scratch.alloc(reinterpret_cast<void**>(&d_near), sizeof(float) * (size_t)m), "alloc");
  hipLaunchKernelGGL(k_nn_sample, dim3((m + kNnThreads - 1) / kNnThreads), dim3(kNnThreads), 0, ctx->stream, d_xyz,
                     (long long)n, stride, m, d_near);
  CKF(hipGetLastError(), "nearest neighbours");
  std::vector<float> nearest((size_t)m);
  CKF(hipMemcpyAsync(nearest.data(), d_near, sizeof(float) * (size_t)m, hipMemcpyDeviceToHost, ctx->stream),
      "download");
  CKF(hipStreamSynchronize(ctx->stream), "sync");
  double spacing = 0.0;
  int counted = 0;
  for (int s = 0; s < m; ++s) {
    const double d = std::min(1e30, static_cast<double>(nearest[(size_t)s]));
    if (d < 1e29) {
      spacing += d;
      ++counted;
    }
  }
  spacing = counted > 0 ? spacing / counted : 0.01;
  const double sigma = std::max(1e-4, spacing * 0.5);
  const float log_sigma = static_cast<float>(std::log(sigma));
  const float logit = std::log(std::max(1e-6f, initial_opacity) / std::max(1e-6f, 1.f - initial_opacity));
  hipLaunchKernelGGL(k_init_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, d_xyz, d_rgb,
                     (long long)n, ctx->cap, ctx->rs, log_sigma, logit, ctx->params.f());
  CKF(hipGetLastError(), "init points");
  if (log_sigma_out) *log_sigma_out = log_sigma;
  return reset_state(ctx, n, degree);
}

splat_status splat_write_ply(splat_context* ctx, const char* path) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  if (!path) return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "write_ply: null path");
  if (ctx->n <= 0) return fmt_fail(ctx, SPLAT_ERROR_STATE, "write_ply: the cloud is empty");
  CKF(hipSetDevice(ctx->device), "set device");
  const long long n = ctx->n;
  // rows carry the cloud's own degree; the capacity's rest stride is the same
  const int rs = ctx->rs;
  std::ostringstream header;
  header << "ply\nformat binary_little_endian 1.0\n";
  header << "comment generated by Parallax\n";
  header << "element vertex " << n << "\n";
  header << "property float x\nproperty float y\nproperty float z\n";
  header << "property float nx\nproperty float ny\nproperty float nz\n";
  for (int i = 0; i < 3; ++i) header << "property float f_dc_" << i << "\n";
  for (int i = 0; i < rs; ++i) header << "property float f_rest_" << i << "\n";
  header << "property float opacity\n";
  for (int i = 0; i < 3; ++i) header << "property float scale_" << i << "\n";
  for (int i = 0; i < 4; ++i) header << "property float rot_" << i << "\n";
  header << "end_header\n";
  const std::string text = header.str();
  const size_t row_floats = 17 + (size_t)rs;
  Scratch scratch;
  float* d_rows = nullptr;
  CKF(scratch.alloc(reinterpret_cast<void**>(&d_rows), sizeof(float) * row_floats * (size_t)n), "alloc rows");
  hipLaunchKernelGGL(k_pack_ply, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->params.f(), n,
                     ctx->cap, rs, d_rows);
  CKF(hipGetLastError(), "pack ply");
  std::vector<uint8_t> bytes(text.size() + sizeof(float) * row_floats * (size_t)n);
  std::memcpy(bytes.data(), text.data(), text.size());
  CKF(hipMemcpyAsync(bytes.data() + text.size(), d_rows, sizeof(float) * row_floats * (size_t)n,
                     hipMemcpyDeviceToHost, ctx->stream),
      "download rows");
  CKF(hipStreamSynchronize(ctx->stream), "sync");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, std::string("write_ply: cannot open ") + path);
  const bool ok = std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
  if (std::fclose(f) != 0 || !ok)
    return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, std::string("write_ply: cannot write ") + path);
  return SPLAT_OK;
}

splat_status splat_ply_info(const char* path, int32_t* count, int32_t* sh_degree, char* error, size_t error_len) {
  std::vector<uint8_t> bytes;
  PlyHeader h;
  std::string message;
  bool ok = path && read_file(path, &bytes);
  if (!ok) message = std::string("cannot read ") + (path ? path : "(null)");
  if (ok) ok = parse_ply_header(bytes, &h, &message);
  if (!ok) {
    if (error && error_len > 0) {
      std::strncpy(error, message.c_str(), error_len - 1);
      error[error_len - 1] = 0;
    }
    return SPLAT_ERROR_INVALID_ARGUMENT;
  }
  if (h.count > 0x7fffffffULL) return SPLAT_ERROR_CAPACITY;
  if (count) *count = (int32_t)h.count;
  if (sh_degree) *sh_degree = h.degree;
  return SPLAT_OK;
}

splat_status splat_read_ply(splat_context* ctx, const char* path) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  if (!path) return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "read_ply: null path");
  if (ctx->cap <= 0) return fmt_fail(ctx, SPLAT_ERROR_STATE, "read_ply: no capacity reserved");
  std::vector<uint8_t> bytes;
  if (!read_file(path, &bytes)) return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, std::string("cannot read ") + path);
  PlyHeader h;
  std::string message;
  if (!parse_ply_header(bytes, &h, &message)) return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, message);
  if ((long long)h.count > ctx->cap)
    return fmt_fail(ctx, SPLAT_ERROR_CAPACITY, "read_ply: more gaussians than the capacity (splat_ply_info first)");
  if (h.rest_count != ctx->rs)
    return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT,
                    "read_ply: the file's SH degree does not match the capacity's (splat_ply_info first)");
  auto index_of = [&h](const std::string& name) {
    const auto it = std::find(h.properties.begin(), h.properties.end(), name);
    return it == h.properties.end() ? -1 : static_cast<int>(it - h.properties.begin());
  };
  PlyIndex ix{};
  ix.x = index_of("x");
  ix.y = index_of("y");
  ix.z = index_of("z");
  ix.opacity = index_of("opacity");
  for (int c = 0; c < 3; ++c) {
    ix.dc[c] = index_of("f_dc_" + std::to_string(c));
    ix.scale[c] = index_of("scale_" + std::to_string(c));
  }
  for (int c = 0; c < 4; ++c) ix.rot[c] = index_of("rot_" + std::to_string(c));
  std::vector<int> rest_map((size_t)std::max(1, h.rest_count));
  for (int k = 0; k < h.rest_count; ++k) rest_map[(size_t)k] = index_of("f_rest_" + std::to_string(k));
  CKF(hipSetDevice(ctx->device), "set device");
  const long long n = (long long)h.count;
  const int stride = (int)h.properties.size();
  Scratch scratch;
  float* d_rows = nullptr;
  int* d_map = nullptr;
  CKF(scratch.alloc(reinterpret_cast<void**>(&d_rows), sizeof(float) * (size_t)stride * (size_t)n), "alloc rows");
  CKF(scratch.alloc(reinterpret_cast<void**>(&d_map), sizeof(int) * rest_map.size()), "alloc");
  CKF(hipMemcpyAsync(d_rows, bytes.data() + h.body, sizeof(float) * (size_t)stride * (size_t)n,
                     hipMemcpyHostToDevice, ctx->stream),
      "upload rows");
  CKF(hipMemcpyAsync(d_map, rest_map.data(), sizeof(int) * rest_map.size(), hipMemcpyHostToDevice, ctx->stream),
      "upload");
  hipLaunchKernelGGL(k_unpack_ply, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, d_rows, n, stride,
                     ix, d_map, ctx->cap, ctx->rs, ctx->params.f());
  CKF(hipGetLastError(), "unpack ply");
  return reset_state(ctx, (int)n, h.degree);
}

}  // extern "C"
