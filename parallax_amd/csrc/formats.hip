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
//
// writeSplat / readSplat (GaussianCloud.cpp:416-493): antimatter15's 32-byte
// records (position, linear sigma, RGBA bytes, quaternion bytes), and
// writeSpz / readSpz (:497-641): Niantic's .spz v2, a gzip stream around a
// 16-byte header and attribute planes (24-bit fixed-point positions, then
// scale, rotation, alpha, colour and SH bytes). The device quantises the
// structure-of-arrays parameters into the records or planes and decodes them
// back; the host only deflates or inflates the stream (zlib, with the
// engine's parameters, so the file is byte for byte the engine's) and moves
// one block. The float forms are the ones the engine's optimised build
// evaluates: divisions by constants decode as multiplies by the rounded
// reciprocal, and a quaternion's squared norm is Eigen's 4-float packet
// reduction.
#include <zlib.h>

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

// Every float operation below is rounded on its own, as on the host that
// writes the engine's files (no fused multiply-add there).
#pragma clang fp contract(off)

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

// The norm of a Quaternionf as Eigen evaluates it on the host: the
// coefficients are stored (x, y, z, w) in one 4-float packet, whose
// horizontal sum is (x^2 + z^2) + (y^2 + w^2).
__device__ inline float quat_norm(float w, float x, float y, float z) {
  return __builtin_sqrtf((x * x + z * z) + (y * y + w * w));
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
// rotation of norm > 1e-6 is normalised, dividing by Eigen's norm).
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
    const float norm = quat_norm(w, x, y, z);
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

// ---- .splat and .spz -------------------------------------------------------

// round(v) clamped to a byte: the engine's quantisers all have this form.
__device__ inline uint32_t to_byte(float v) { return (uint32_t)fmaxf(0.f, fminf(255.f, roundf(v))); }

// gaussianOpacity (GaussianCloud.h:80-82) quantised to a byte: the engine's
// build folds round(255 * (1 / (1 + e))) into round(255 / (1 + e)). The
// exponential in double, rounded once, is the correctly rounded expf.
__device__ inline uint32_t opacity_byte(float logit) {
  return to_byte(255.0f / (1.0f + (float)exp(-(double)logit)));
}

// The opacity logit of an alpha byte, as both readers decode it (the clamp
// keeps the logit finite).
__device__ inline float logit_of_byte(uint32_t b) {
  const float alpha = fmaxf(1e-6f, fminf(1.f - 1e-6f, (float)b * (1.0f / 255.0f)));
  return (float)log((double)(alpha / (1.f - alpha)));
}

// A colour byte back to the DC coefficient.
__device__ inline float dc_of_byte(uint32_t b) { return ((float)b * (1.0f / 255.0f) - 0.5f) * (1.0f / kShC0); }

// One 32-byte .splat record per gaussian (writeSplat, GaussianCloud.cpp:
// 416-446): position, exp(log sigma), the clamped base colour and opacity as
// bytes, the normalised quaternion (w first) as q * 128 + 128.
__global__ __launch_bounds__(256) void k_pack_splat(const float* __restrict__ P, long long n, long long cap,
                                                    uint4* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint4 a, b;
  a.x = __float_as_uint(P[3 * i]);
  a.y = __float_as_uint(P[3 * i + 1]);
  a.z = __float_as_uint(P[3 * i + 2]);
  a.w = __float_as_uint((float)exp((double)P[3 * cap + 3 * i]));
  b.x = __float_as_uint((float)exp((double)P[3 * cap + 3 * i + 1]));
  b.y = __float_as_uint((float)exp((double)P[3 * cap + 3 * i + 2]));
  uint32_t rgba = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float c = fmaxf(0.f, fminf(1.f, 0.5f + kShC0 * P[11 * cap + 3 * i + k]));
    rgba |= to_byte(c * 255.f) << (8 * k);
  }
  rgba |= opacity_byte(P[10 * cap + i]) << 24;
  b.z = rgba;
  float q[4] = {P[6 * cap + 4 * i], P[6 * cap + 4 * i + 1], P[6 * cap + 4 * i + 2], P[6 * cap + 4 * i + 3]};
  const float s = (q[1] * q[1] + q[3] * q[3]) + (q[2] * q[2] + q[0] * q[0]);
  uint32_t rot = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    // Eigen's normalized() leaves a zero quaternion as it is
    const float c = s > 0.f ? q[k] / __builtin_sqrtf(s) : q[k];
    rot |= to_byte(c * 128.0f + 128.0f) << (8 * k);
  }
  b.w = rot;
  out[2 * i] = a;
  out[2 * i + 1] = b;
}

// .splat records back into the capacity layout (readSplat, :448-493): log
// sigma of max(1e-20, sigma), DC from the colour byte, the opacity logit from
// the alpha byte, the quaternion bytes dequantised and normalised (the
// identity when its norm is <= 1e-6); no higher SH bands.
__global__ __launch_bounds__(256) void k_unpack_splat(const uint4* __restrict__ in, long long n, long long cap,
                                                      float* __restrict__ P) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint4 a = in[2 * i], b = in[2 * i + 1];
  P[3 * i] = __uint_as_float(a.x);
  P[3 * i + 1] = __uint_as_float(a.y);
  P[3 * i + 2] = __uint_as_float(a.z);
  const float sigma[3] = {__uint_as_float(a.w), __uint_as_float(b.x), __uint_as_float(b.y)};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    P[3 * cap + 3 * i + k] = (float)log((double)fmaxf(1e-20f, sigma[k]));
    P[11 * cap + 3 * i + k] = dc_of_byte((b.z >> (8 * k)) & 0xffu);
  }
  P[10 * cap + i] = logit_of_byte(b.z >> 24);
  float q[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] = ((float)((b.w >> (8 * k)) & 0xffu) - 128.0f) / 128.0f;
  const float norm = quat_norm(q[0], q[1], q[2], q[3]);
#pragma unroll
  for (int k = 0; k < 4; ++k) P[6 * cap + 4 * i + k] = norm > 1e-6f ? q[k] / norm : (k == 0 ? 1.f : 0.f);
}

// Where each .spz attribute plane starts after the 16-byte header, for n
// gaussians: positions 9 bytes each, scales 3, rotations 3, alphas 1,
// colours 3, then 3 * rest-per-channel SH bytes.
struct SpzPlanes {
  long long pos, scale, rot, alpha, colour, sh, end;
  __host__ __device__ SpzPlanes(long long n, int rpc)
      : pos(0), scale(9 * n), rot(12 * n), alpha(15 * n), colour(16 * n), sh(19 * n), end((19 + 3LL * rpc) * n) {}
};

// The .spz planes of every gaussian (writeSpz, :497-560).
__global__ __launch_bounds__(256) void k_pack_spz(const float* __restrict__ P, long long n, long long cap, int rpc,
                                                  int fractional_bits, uint8_t* __restrict__ body) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SpzPlanes at(n, rpc);
  // positions: 24-bit signed fixed point, little-endian, clamped to the range
  const float unit = (float)(1 << fractional_bits);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float r = roundf(P[3 * i + k] * unit);
    const int32_t fixed = r >= 8388607.f ? 8388607 : (r <= -8388608.f ? -8388608 : (int32_t)r);
    const uint32_t bits = (uint32_t)fixed & 0xffffffu;
    uint8_t* p = body + at.pos + 9 * i + 3 * k;
    p[0] = (uint8_t)bits;
    p[1] = (uint8_t)(bits >> 8);
    p[2] = (uint8_t)(bits >> 16);
  }
  // log sigma on a byte: (log sigma + 10) * 16
#pragma unroll
  for (int k = 0; k < 3; ++k) body[at.scale + 3 * i + k] = (uint8_t)to_byte((P[3 * cap + 3 * i + k] + 10.0f) * 16.0f);
  // x, y, z of the normalised quaternion in the w >= 0 hemisphere
  float q[4] = {P[6 * cap + 4 * i], P[6 * cap + 4 * i + 1], P[6 * cap + 4 * i + 2], P[6 * cap + 4 * i + 3]};
  const float s = (q[1] * q[1] + q[3] * q[3]) + (q[2] * q[2] + q[0] * q[0]);
  if (s > 0.f) {
    const float norm = __builtin_sqrtf(s);
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = q[k] / norm;
  }
  // v * 127.5 + 127.5, which the engine's build factors into (v + 1) * 127.5
  const float flip = q[0] < 0.f ? -1.f : 1.f;
#pragma unroll
  for (int k = 1; k < 4; ++k) body[at.rot + 3 * i + k - 1] = (uint8_t)to_byte((flip * q[k] + 1.0f) * 127.5f);
  body[at.alpha + i] = (uint8_t)opacity_byte(P[10 * cap + i]);
#pragma unroll
  for (int k = 0; k < 3; ++k)
    body[at.colour + 3 * i + k] = (uint8_t)to_byte((0.5f + kShC0 * P[11 * cap + 3 * i + k]) * 255.f);
  // SH: coefficient-major, colour channel fastest (memory is channel-major)
  const float* rest = P + 14 * cap + 3LL * rpc * i;
  uint8_t* sh = body + at.sh + 3LL * rpc * i;
  for (int k = 0; k < rpc; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) sh[3 * k + c] = (uint8_t)to_byte(rest[c * rpc + k] * 128.0f + 128.0f);
}

// .spz planes back into the capacity layout (readSpz, :562-641): w recovered
// from the unit norm, the quaternion normalised (the identity when its norm is
// <= 1e-6).
__global__ __launch_bounds__(256) void k_unpack_spz(const uint8_t* __restrict__ body, long long n, int rpc,
                                                    int fractional_bits, long long cap, float* __restrict__ P) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SpzPlanes at(n, rpc);
  const float unit = (float)(1 << fractional_bits);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const uint8_t* p = body + at.pos + 9 * i + 3 * k;
    uint32_t bits = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16);
    if (bits & 0x800000u) bits |= 0xff000000u;
    P[3 * i + k] = (float)(int32_t)bits / unit;
    P[3 * cap + 3 * i + k] = (float)body[at.scale + 3 * i + k] / 16.0f - 10.0f;
    P[11 * cap + 3 * i + k] = dc_of_byte(body[at.colour + 3 * i + k]);
  }
  const float x = ((float)body[at.rot + 3 * i] - 127.5f) * (1.0f / 127.5f);
  const float y = ((float)body[at.rot + 3 * i + 1] - 127.5f) * (1.0f / 127.5f);
  const float z = ((float)body[at.rot + 3 * i + 2] - 127.5f) * (1.0f / 127.5f);
  const float w = __builtin_sqrtf(fmaxf(0.f, 1.f - (x * x + (y * y + z * z))));
  const float norm = quat_norm(w, x, y, z);
  const bool unit_q = norm > 1e-6f;
  P[6 * cap + 4 * i] = unit_q ? w / norm : 1.f;
  P[6 * cap + 4 * i + 1] = unit_q ? x / norm : 0.f;
  P[6 * cap + 4 * i + 2] = unit_q ? y / norm : 0.f;
  P[6 * cap + 4 * i + 3] = unit_q ? z / norm : 0.f;
  P[10 * cap + i] = logit_of_byte(body[at.alpha + i]);
  // undo the colour-channel interleave
  const uint8_t* sh = body + at.sh + 3LL * rpc * i;
  float* rest = P + 14 * cap + 3LL * rpc * i;
  for (int k = 0; k < rpc; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) rest[c * rpc + k] = ((float)sh[3 * k + c] - 128.0f) / 128.0f;
}

int rest_per_channel(int degree) { return degree <= 0 ? 0 : (degree == 1 ? 3 : (degree == 2 ? 8 : 15)); }

// A parsed 3DGS PLY header: which inputs readGaussianPly accepts and the
// error it reports first (GaussianCloud.cpp:305-374). The header is read from
// the first 64 KiB of the file; `index` maps a property name to its first
// position in the row.
struct PlyHeader {
  size_t count = 0;
  size_t body = 0;  // byte offset of the first row
  std::vector<std::string> properties;
  std::unordered_map<std::string, int> first;
  int degree = 0;
  int rest_count = 0;
  int index(const std::string& name) const {
    const auto it = first.find(name);
    return it == first.end() ? -1 : it->second;
  }
};

// The whitespace-separated words of header bytes [b, e), as a stream reads
// strings (the C locale's isspace).
std::vector<std::string> header_words(const char* b, const char* e) {
  std::vector<std::string> words;
  const char* p = b;
  for (;;) {
    while (p < e && std::isspace(static_cast<unsigned char>(*p))) ++p;
    if (p == e) return words;
    const char* q = p;
    while (q < e && !std::isspace(static_cast<unsigned char>(*q))) ++q;
    words.emplace_back(p, q);
    p = q;
  }
}

bool parse_ply_header(const std::vector<uint8_t>& bytes, PlyHeader* h, std::string* error) {
  enum Fault { kNone, kNotPly, kNotLittleEndian, kNotFloat, kNoVertices, kNoCore, kRestNot3, kRestNoDegree, kShort };
  static const char* const kMessage[] = {"",
                                         "not a PLY file",
                                         "only binary little-endian 3DGS PLY is supported",
                                         "3DGS PLY properties must be float",
                                         "3DGS PLY has no vertices",
                                         "3DGS PLY is missing core properties",
                                         "f_rest count is not a multiple of 3",
                                         "f_rest count does not correspond to an SH degree",
                                         "3DGS PLY body is truncated"};
  auto verdict = [&](Fault f) {
    *error = kMessage[f];
    return f == kNone;
  };
  static const char kEnd[] = "end_header\n";
  const size_t scanned = std::min<size_t>(bytes.size(), 65536);
  const char* text = reinterpret_cast<const char*>(bytes.data());
  const char* stop = text + scanned;
  const char* end = std::search(text, stop, kEnd, kEnd + sizeof(kEnd) - 1);
  if (scanned < 3 || std::memcmp(text, "ply", 3) != 0 || end == stop) return verdict(kNotPly);
  static const char kLe[] = "binary_little_endian";
  if (std::search(text, stop, kLe, kLe + sizeof(kLe) - 1) == stop) return verdict(kNotLittleEndian);

  // line by line up to end_header: every `element` line sets the count (when
  // it has a third word; a word with no leading digits reads as 0), and every
  // `property` line of any element must be a float, in row order
  for (const char* line = text; line < end;) {
    const char* eol = std::find(line, end, '\n');
    const std::vector<std::string> w = header_words(line, eol);
    line = eol + (eol < end ? 1 : 0);
    if (w.empty()) continue;
    if (w[0] == "element" && w.size() >= 3) {
      h->count = std::strtoull(w[2].c_str(), nullptr, 10);
    } else if (w[0] == "property") {
      const std::string type = w.size() > 1 ? w[1] : std::string();
      if (type != "float" && type != "float32") return verdict(kNotFloat);
      const std::string name = w.size() > 2 ? w[2] : std::string();
      h->first.emplace(name, static_cast<int>(h->properties.size()));
      h->properties.push_back(name);
    }
  }
  if (h->count == 0 || h->properties.empty()) return verdict(kNoVertices);
  for (const char* core : {"x", "y", "z", "opacity"})
    if (h->index(core) < 0) return verdict(kNoCore);
  h->rest_count = 0;
  while (h->index("f_rest_" + std::to_string(h->rest_count)) >= 0) ++h->rest_count;
  if (h->rest_count % 3 != 0) return verdict(kRestNot3);
  // SH degree d carries (d + 1)^2 - 1 coefficients per channel
  h->degree = -1;
  for (int d = 0; d <= 3; ++d)
    if ((d + 1) * (d + 1) - 1 == h->rest_count / 3) h->degree = d;
  if (h->degree < 0) {
    h->degree = 0;
    return verdict(kRestNoDegree);
  }
  h->body = static_cast<size_t>(end - text) + sizeof(kEnd) - 1;
  const size_t row = h->properties.size() * sizeof(float);
  const size_t room = bytes.size() - h->body;  // bytes.size() >= body: end_header lies inside the file
  if (h->count > room / row) return verdict(kShort);
  return verdict(kNone);
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

bool write_file(const char* path, const uint8_t* data, size_t size) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  const bool ok = std::fwrite(data, 1, size, f) == size;
  return std::fclose(f) == 0 && ok;
}

// One gzip member holding `size` bytes, deflated with the engine's settings
// (default level, 15-bit window behind a gzip wrapper, memory level 8,
// default strategy) in a single Z_FINISH call: for the same zlib, the same
// bytes as the engine's writer.
bool gzip_deflate(const uint8_t* in, size_t size, std::vector<uint8_t>* out) {
  if (size > 0x7fffffffULL) return false;
  z_stream zs{};
  if (deflateInit2(&zs, Z_DEFAULT_COMPRESSION, Z_DEFLATED, MAX_WBITS + 16, 8, Z_DEFAULT_STRATEGY) != Z_OK)
    return false;
  out->resize(deflateBound(&zs, static_cast<uLong>(size)) + 64);
  zs.next_in = const_cast<Bytef*>(in);
  zs.avail_in = static_cast<uInt>(size);
  zs.next_out = out->data();
  zs.avail_out = static_cast<uInt>(out->size());
  const int rc = deflate(&zs, Z_FINISH);
  out->resize(out->size() - zs.avail_out);
  deflateEnd(&zs);
  return rc == Z_STREAM_END;
}

// The first gzip (or zlib) member of `in`, refusing more than 2 GiB of output
// as the engine's reader does.
bool gzip_inflate(const std::vector<uint8_t>& in, std::vector<uint8_t>* out) {
  constexpr size_t kLimit = size_t(1) << 31;
  if (in.size() > 0xffffffffULL) return false;
  z_stream zs{};
  if (inflateInit2(&zs, MAX_WBITS + 32) != Z_OK) return false;
  zs.next_in = const_cast<Bytef*>(in.data());
  zs.avail_in = static_cast<uInt>(in.size());
  size_t have = 0;
  int rc = Z_OK;
  out->resize(std::max<size_t>(1 << 16, 4 * in.size()));
  while (rc != Z_STREAM_END && have <= kLimit) {
    if (out->size() - have < (1 << 16)) out->resize(2 * out->size());
    const size_t room = std::min<size_t>(out->size() - have, 1u << 30);
    zs.next_out = out->data() + have;
    zs.avail_out = static_cast<uInt>(room);
    rc = inflate(&zs, Z_NO_FLUSH);
    have += room - zs.avail_out;
    if (rc != Z_OK && rc != Z_STREAM_END) break;
  }
  inflateEnd(&zs);
  out->resize(have);
  return rc == Z_STREAM_END && have <= kLimit;
}

constexpr uint32_t kSpzMagic = 0x5053474eu;  // "NGSP", little-endian
constexpr uint32_t kSpzVersion = 2u;

struct SpzHeader {
  uint32_t count = 0;
  int degree = 0;
  int fractional_bits = 0;
  bool antialiased = false;
};

// readSpz's checks on the inflated stream (GaussianCloud.cpp:572-597), with
// the engine's messages.
bool parse_spz(const std::vector<uint8_t>& raw, SpzHeader* h, std::string* error) {
  if (raw.size() < 16) {
    *error = ".spz header is truncated";
    return false;
  }
  uint32_t magic, version;
  std::memcpy(&magic, raw.data(), 4);
  std::memcpy(&version, raw.data() + 4, 4);
  std::memcpy(&h->count, raw.data() + 8, 4);
  if (magic != kSpzMagic) {
    *error = "bad .spz magic";
    return false;
  }
  if (version != kSpzVersion) {
    *error = "unsupported .spz version";
    return false;
  }
  h->degree = raw[12];
  h->fractional_bits = raw[13];
  h->antialiased = (raw[14] & 0x1) != 0;
  if (h->degree > 3) {
    *error = ".spz sh_degree out of range";
    return false;
  }
  if (raw.size() - 16 != SpzPlanes(h->count, rest_per_channel(h->degree)).end) {
    *error = ".spz payload size does not match the declared point count";
    return false;
  }
  return true;
}

// A .spz file inflated and checked.
bool load_spz(const char* path, std::vector<uint8_t>* raw, SpzHeader* h, std::string* error) {
  std::vector<uint8_t> compressed;
  if (!path || !read_file(path, &compressed)) {
    *error = std::string("cannot read ") + (path ? path : "(null)");
    return false;
  }
  if (!gzip_inflate(compressed, raw)) {
    *error = ".spz is not a valid gzip stream";
    return false;
  }
  return parse_spz(*raw, h, error);
}

// A .splat file read and checked (readSplat, :454-457).
bool load_splat(const char* path, std::vector<uint8_t>* bytes, std::string* error) {
  if (!path || !read_file(path, bytes)) {
    *error = std::string("cannot read ") + (path ? path : "(null)");
    return false;
  }
  if (bytes->empty() || bytes->size() % 32 != 0) {
    *error = ".splat size is not a multiple of the 32-byte record";
    return false;
  }
  return true;
}

void copy_error(const std::string& message, char* error, size_t error_len) {
  if (error && error_len > 0) {
    std::strncpy(error, message.c_str(), error_len - 1);
    error[error_len - 1] = 0;
  }
}

// The SH degree a capacity's rest stride stands for (-1: none).
int degree_of_rest(int rs) {
  for (int d = 0; d <= 3; ++d)
    if (3 * rest_per_channel(d) == rs) return d;
  return -1;
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
    copy_error(message, error, error_len);
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
  auto index_of = [&h](const std::string& name) { return h.index(name); };
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

splat_status splat_write_splat(splat_context* ctx, const char* path) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  if (!path) return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "write_splat: null path");
  if (ctx->n <= 0) return fmt_fail(ctx, SPLAT_ERROR_STATE, "write_splat: the cloud is empty");
  CKF(hipSetDevice(ctx->device), "set device");
  const long long n = ctx->n;
  Scratch scratch;
  uint4* d_out = nullptr;
  CKF(scratch.alloc(reinterpret_cast<void**>(&d_out), 32 * (size_t)n), "alloc records");
  hipLaunchKernelGGL(k_pack_splat, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->params.f(), n,
                     ctx->cap, d_out);
  CKF(hipGetLastError(), "pack splat");
  std::vector<uint8_t> bytes(32 * (size_t)n);
  CKF(hipMemcpyAsync(bytes.data(), d_out, bytes.size(), hipMemcpyDeviceToHost, ctx->stream), "download records");
  CKF(hipStreamSynchronize(ctx->stream), "sync");
  if (!write_file(path, bytes.data(), bytes.size()))
    return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, std::string("write_splat: cannot write ") + path);
  return SPLAT_OK;
}

splat_status splat_splat_info(const char* path, int32_t* count, char* error, size_t error_len) {
  std::vector<uint8_t> bytes;
  std::string message;
  if (!load_splat(path, &bytes, &message)) {
    copy_error(message, error, error_len);
    return SPLAT_ERROR_INVALID_ARGUMENT;
  }
  if (bytes.size() / 32 > 0x7fffffffULL) return SPLAT_ERROR_CAPACITY;
  if (count) *count = (int32_t)(bytes.size() / 32);
  return SPLAT_OK;
}

splat_status splat_read_splat(splat_context* ctx, const char* path) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  if (ctx->cap <= 0) return fmt_fail(ctx, SPLAT_ERROR_STATE, "read_splat: no capacity reserved");
  std::vector<uint8_t> bytes;
  std::string message;
  if (!load_splat(path, &bytes, &message)) return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, message);
  const long long n = (long long)(bytes.size() / 32);
  if (n > ctx->cap)
    return fmt_fail(ctx, SPLAT_ERROR_CAPACITY, "read_splat: more gaussians than the capacity (splat_splat_info first)");
  if (ctx->rs != 0)
    return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT,
                    "read_splat: a .splat file carries SH degree 0 only; reserve the capacity at degree 0");
  CKF(hipSetDevice(ctx->device), "set device");
  Scratch scratch;
  uint4* d_in = nullptr;
  CKF(scratch.alloc(reinterpret_cast<void**>(&d_in), bytes.size()), "alloc records");
  CKF(hipMemcpyAsync(d_in, bytes.data(), bytes.size(), hipMemcpyHostToDevice, ctx->stream), "upload records");
  hipLaunchKernelGGL(k_unpack_splat, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, d_in, n, ctx->cap,
                     ctx->params.f());
  CKF(hipGetLastError(), "unpack splat");
  return reset_state(ctx, (int)n, 0);
}

splat_status splat_write_spz(splat_context* ctx, const char* path, int32_t fractional_bits, int32_t antialiased) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  if (!path) return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "write_spz: null path");
  if (ctx->n <= 0) return fmt_fail(ctx, SPLAT_ERROR_STATE, "write_spz: the cloud is empty");
  const int degree = degree_of_rest(ctx->rs);
  if (degree < 0) return fmt_fail(ctx, SPLAT_ERROR_STATE, "write_spz: the capacity's SH degree is not 0..3");
  const int bits = std::max(0, std::min(20, (int)fractional_bits));
  CKF(hipSetDevice(ctx->device), "set device");
  const long long n = ctx->n;
  const int rpc = rest_per_channel(degree);
  const size_t body = (size_t)SpzPlanes(n, rpc).end;
  Scratch scratch;
  uint8_t* d_body = nullptr;
  CKF(scratch.alloc(reinterpret_cast<void**>(&d_body), body), "alloc planes");
  hipLaunchKernelGGL(k_pack_spz, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->params.f(), n,
                     ctx->cap, rpc, bits, d_body);
  CKF(hipGetLastError(), "pack spz");
  std::vector<uint8_t> raw(16 + body);
  const uint32_t words[3] = {kSpzMagic, kSpzVersion, (uint32_t)n};
  std::memcpy(raw.data(), words, 12);
  raw[12] = (uint8_t)degree;
  raw[13] = (uint8_t)bits;
  raw[14] = antialiased ? 0x1 : 0x0;
  raw[15] = 0;  // reserved
  CKF(hipMemcpyAsync(raw.data() + 16, d_body, body, hipMemcpyDeviceToHost, ctx->stream), "download planes");
  CKF(hipStreamSynchronize(ctx->stream), "sync");
  std::vector<uint8_t> compressed;
  if (!gzip_deflate(raw.data(), raw.size(), &compressed))
    return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "write_spz: deflate failed");
  if (!write_file(path, compressed.data(), compressed.size()))
    return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, std::string("write_spz: cannot write ") + path);
  return SPLAT_OK;
}

splat_status splat_spz_info(const char* path, int32_t* count, int32_t* sh_degree, int32_t* antialiased, char* error,
                            size_t error_len) {
  std::vector<uint8_t> raw;
  SpzHeader h;
  std::string message;
  if (!load_spz(path, &raw, &h, &message)) {
    copy_error(message, error, error_len);
    return SPLAT_ERROR_INVALID_ARGUMENT;
  }
  if (h.count > 0x7fffffffu) return SPLAT_ERROR_CAPACITY;
  if (count) *count = (int32_t)h.count;
  if (sh_degree) *sh_degree = h.degree;
  if (antialiased) *antialiased = h.antialiased ? 1 : 0;
  return SPLAT_OK;
}

splat_status splat_read_spz(splat_context* ctx, const char* path, int32_t* antialiased) {
  if (!ctx) return SPLAT_ERROR_INVALID_ARGUMENT;
  if (ctx->cap <= 0) return fmt_fail(ctx, SPLAT_ERROR_STATE, "read_spz: no capacity reserved");
  std::vector<uint8_t> raw;
  SpzHeader h;
  std::string message;
  if (!load_spz(path, &raw, &h, &message)) return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, message);
  const long long n = (long long)h.count;
  if (n > ctx->cap)
    return fmt_fail(ctx, SPLAT_ERROR_CAPACITY, "read_spz: more gaussians than the capacity (splat_spz_info first)");
  if (3 * rest_per_channel(h.degree) != ctx->rs)
    return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT,
                    "read_spz: the file's SH degree does not match the capacity's (splat_spz_info first)");
  if (h.fractional_bits > 30)
    return fmt_fail(ctx, SPLAT_ERROR_INVALID_ARGUMENT, "read_spz: more than 30 fractional position bits");
  CKF(hipSetDevice(ctx->device), "set device");
  if (antialiased) *antialiased = h.antialiased ? 1 : 0;
  if (n == 0) return reset_state(ctx, 0, h.degree);
  Scratch scratch;
  uint8_t* d_body = nullptr;
  CKF(scratch.alloc(reinterpret_cast<void**>(&d_body), raw.size() - 16), "alloc planes");
  CKF(hipMemcpyAsync(d_body, raw.data() + 16, raw.size() - 16, hipMemcpyHostToDevice, ctx->stream), "upload planes");
  hipLaunchKernelGGL(k_unpack_spz, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, d_body, n,
                     rest_per_channel(h.degree), h.fractional_bits, ctx->cap, ctx->params.f());
  CKF(hipGetLastError(), "unpack spz");
  return reset_state(ctx, (int)n, h.degree);
}

}  // extern "C"
