// EWA projection + spherical-harmonic colour + tile-rect count.
//
// One thread per gaussian; it reads the gaussian's parameters once and loops
// over the views of the batch, writing one 64-byte record, one depth key and
// one tile count per (view, gaussian) slot. The arithmetic is the engine's
// projectSplats (engine/native/cpp/src/splat/SplatRaster.cpp:201-263) and
// evaluateSh (:179-199) with its covariance3d (:82-87), projectionJacobian
// (:91-102) and tile rect (:121-126), in IEEE float with no contraction so that
// culling, the radius and the tile rect -- the inputs of the binning -- match
// the CPU checker bit for bit. HBM-bound: compulsory traffic is the parameter
// read (56 + 4*rs bytes per gaussian) plus, per (view, gaussian), the record
// and the depth key and -- where every view has at most 256 x 256 tiles -- the
// packed tile rect (72 bytes), else the tile count, the super-tile count and
// the 8-byte rect (84 bytes).
#include "splat_common.h"
#include "splat_kernels.h"

namespace splat {
namespace {

constexpr float kShC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                            -1.0925484305920792f, 0.5462742152960396f};
constexpr float kShC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                            0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                            -0.5900435899266435f};

__device__ __forceinline__ float fmax_std(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float fmin_std(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float clamp_std(float v, float lo, float hi) {
  return (v < lo) ? lo : ((hi < v) ? hi : v);
}

// The 15 non-constant real SH basis values along unit direction d
// (SplatRaster.cpp:29-52).
__device__ __forceinline__ int sh_basis(int degree, float x, float y, float z, float* b) {
  if (degree < 1) return 0;
  b[0] = -kShC1 * y;
  b[1] = kShC1 * z;
  b[2] = -kShC1 * x;
  if (degree < 2) return 3;
  const float xx = x * x, yy = y * y, zz = z * z;
  const float xy = x * y, yz = y * z, xz = x * z;
  b[3] = kShC2[0] * xy;
  b[4] = kShC2[1] * yz;
  b[5] = kShC2[2] * (2.f * zz - xx - yy);
  b[6] = kShC2[3] * xz;
  b[7] = kShC2[4] * (xx - yy);
  if (degree < 3) return 8;
  b[8] = kShC3[0] * y * (3.f * xx - yy);
  b[9] = kShC3[1] * xy * z;
  b[10] = kShC3[2] * y * (4.f * zz - xx - yy);
  b[11] = kShC3[3] * z * (2.f * zz - 3.f * xx - 3.f * yy);
  b[12] = kShC3[4] * x * (4.f * zz - xx - yy);
  b[13] = kShC3[5] * z * (xx - yy);
  b[14] = kShC3[6] * x * (xx - 3.f * yy);
  return 15;
}

// Half extents of a box that holds every pixel centre where the compositing
// loop can find alpha >= min_alpha for this splat, so a workgroup whose pixels
// lie outside the box may skip the splat without changing any result.
//
// alpha >= min_alpha needs o * exp(power) >= min_alpha, power = -Q(d)/2 with
// Q(d) = a dx^2 + 2 b dx dy + c dy^2 (the conic), i.e. Q <= q = 2 ln(o/min_alpha).
// The box of that ellipse is |dx| <= sqrt(q c / det), |dy| <= sqrt(q a / det),
// det = ac - b^2. The loop evaluates Q in float: each term carries at most
// ~3e-7 relative error, and a dx^2 + c dy^2 <= Q * 2ac/det, so the float Q is
// within 1.2e-6 * (ac/det) * Q of the exact one. q is therefore raised by 1e-5
// (the exponential's and the product's rounding), the box widened by
// 1/(1 - 4e-6 ac/det) in q and by 1e-3 px, and a splat too elongated for the
// bound to be meaningful gets no box (+inf).
//
// q depends on the splat only (cull_q, once per gaussian, in double); the box
// is evaluated per view in float. det comes from Kahan's fused form
// fma(a, c, -b b) + fma(-b, b, b b), within 2 ulp of the exact a c - b^2 for
// every conic the shrink test keeps (ac/det < 1.9e5, far from the 1/ulp where
// it would fail); with the hardware reciprocal and square root (1 ulp each)
// the float result is within 1e-6 of the exact one, and the factor 1 + 1e-5
// (against 1 + 1e-6 in a double evaluation) keeps the box containing the
// exact box. tests/test_cull_rules.py checks it against the double form.
__device__ __forceinline__ float cull_q(float opacity, float min_alpha) {
  if (!(min_alpha > 0.f)) return -1.f;  // thresholds off: no box
  const double q = 2.0 * (log((double)opacity / (double)min_alpha) + 1e-5);
  return q < 0.0 ? 0.f : (float)(q * (1.0 + 1e-7));  // rounded to float, up
}
__device__ __forceinline__ void cull_extents(float a, float b, float c, float q, float* ex, float* ey) {
  *ex = *ey = __builtin_inff();
  if (!(q >= 0.f)) return;
  const float bb = b * b;
  const float det = __builtin_fmaf(a, c, -bb) + __builtin_fmaf(-b, b, bb);
  if (!(det > 0.f)) return;
  const float rdet = __builtin_amdgcn_rcpf(det);
  const float shrink = 1.f - 4e-6f * ((a * c) * rdet);
  if (!(shrink > 0.25f)) return;
  const float s = q * rdet * __builtin_amdgcn_rcpf(shrink);
  *ex = __builtin_amdgcn_sqrtf(s * c) * 1.00001f + 1e-3f;
  *ey = __builtin_amdgcn_sqrtf(s * a) * 1.00001f + 1e-3f;
}

template <int kDeg>
__global__ __launch_bounds__(256) void k_project(const float* __restrict__ params, long long cap, int n,
                                                 int rest_stride, const ViewDesc* __restrict__ views,
                                                 int n_views, float dilation, float min_alpha,
                                                 SplatRec* __restrict__ rec, uint32_t* __restrict__ keys,
                                                 uint32_t* __restrict__ counts, uint32_t* __restrict__ scounts,
                                                 uint2* __restrict__ rects, uint32_t* __restrict__ prects,
                                                 uint32_t* __restrict__ key_ranges) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;  // the lanes left are a prefix of the wave
  const float* P = params;
  const float p0 = P[3 * g], p1 = P[3 * g + 1], p2 = P[3 * g + 2];
  const float* lsp = P + 3 * cap + 3 * g;
  const float* rp = P + 6 * cap + 4 * (long long)g;
  const float logit = P[10 * cap + g];
  const float* dcp = P + 11 * cap + 3 * (long long)g;
  const float* rest = P + 14 * cap + (long long)rest_stride * g;

  // covariance3d: R from the normalised quaternion (Eigen normalized() and
  // toRotationMatrix(), Quaternion.h:592-622), scale = exp(log_scale).
  float qw = rp[0], qx = rp[1], qy = rp[2], qz = rp[3];
  const float qsq = qx * qx + qy * qy + qz * qz + qw * qw;
  if (qsq > 0.f) {
    const float s = __builtin_sqrtf(qsq);
    qx = ((qx) / (s)); qy = ((qy) / (s)); qz = ((qz) / (s)); qw = ((qw) / (s));
  }
  const float tx = 2.f * qx, ty = 2.f * qy, tz = 2.f * qz;
  const float twx = tx * qw, twy = ty * qw, twz = tz * qw;
  const float txx = tx * qx, txy = ty * qx, txz = tz * qx;
  const float tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
  float R[3][3];
  R[0][0] = 1.f - (tyy + tzz); R[0][1] = txy - twz; R[0][2] = txz + twy;
  R[1][0] = txy + twz; R[1][1] = 1.f - (txx + tzz); R[1][2] = tyz - twx;
  R[2][0] = txz - twy; R[2][1] = tyz + twx; R[2][2] = 1.f - (txx + tyy);
  const float s0 = det_exp(lsp[0]), s1 = det_exp(lsp[1]), s2 = det_exp(lsp[2]);
  float M[3][3];
  for (int i = 0; i < 3; ++i) {
    M[i][0] = R[i][0] * s0;
    M[i][1] = R[i][1] * s1;
    M[i][2] = R[i][2] * s2;
  }
  float S[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) S[i][j] = M[i][0] * M[j][0] + M[i][1] * M[j][1] + M[i][2] * M[j][2];
  const float opacity = ((1.0f) / (1.0f + det_exp(-logit)));
  const float q_cull = cull_q(opacity, min_alpha);  // the cull ellipse's level, the same in every view
  const float dc0 = dcp[0], dc1 = dcp[1], dc2 = dcp[2];
  const int per_channel = rest_stride / 3;
  // the SH coefficients this degree uses, loaded once for all views
  constexpr int kCount = kDeg == 0 ? 0 : (kDeg == 1 ? 3 : (kDeg == 2 ? 8 : 15));
  float rc[3][kCount > 0 ? kCount : 1];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int k = 0; k < kCount; ++k) rc[c][k] = rest[c * per_channel + k];

  uint32_t key_min = 0xFFFFFFFFu, key_max = 0u;  // over this gaussian's kept slots
  // two views at a time: their chains (transform, Jacobian, 2D covariance,
  // conic, SH colour, each with its IEEE divisions) are independent, and
  // interleaving them beats the occupancy they cost (three waves per SIMD at
  // SH3 instead of four)
#pragma unroll 2
  for (int v = 0; v < n_views; ++v) {
    const ViewDesc& V = views[v];
    const long long slot = (long long)v * n + g;
    SplatRec o;
    o.flags = 0u;
    uint32_t count = 0, scount = 0;
    uint32_t key = 0xFFFFFFFFu;
    bool ok = true;
    float pc[3];
    for (int i = 0; i < 3; ++i) pc[i] = V.W[3 * i] * p0 + V.W[3 * i + 1] * p1 + V.W[3 * i + 2] * p2 + V.t[i];
    if (pc[2] < 0.01f) ok = false;
    float det = 0.f, cov00 = 0.f, cov01 = 0.f, cov11 = 0.f;
    if (ok) {
      const float z = pc[2];
      const float lx = V.thx * kSlack, ly = V.thy * kSlack;
      const float jx = clamp_std(((pc[0]) / (z)), -lx, lx) * z;
      const float jy = clamp_std(((pc[1]) / (z)), -ly, ly) * z;
      const float zz = z * z;
      const float J00 = ((V.fx) / (z)), J02 = ((-V.fx * jx) / (zz));
      const float J11 = ((V.fy) / (z)), J12 = ((-V.fy * jy) / (zz));
      float T[2][3];
      for (int c = 0; c < 3; ++c) {
        T[0][c] = J00 * V.W[c] + 0.f * V.W[3 + c] + J02 * V.W[6 + c];
        T[1][c] = 0.f * V.W[c] + J11 * V.W[3 + c] + J12 * V.W[6 + c];
      }
      float A[2][3];
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) A[r][c] = T[r][0] * S[0][c] + T[r][1] * S[1][c] + T[r][2] * S[2][c];
      cov00 = A[0][0] * T[0][0] + A[0][1] * T[0][1] + A[0][2] * T[0][2];
      cov01 = A[0][0] * T[1][0] + A[0][1] * T[1][1] + A[0][2] * T[1][2];
      const float cov10 = A[1][0] * T[0][0] + A[1][1] * T[0][1] + A[1][2] * T[0][2];
      cov11 = A[1][0] * T[1][0] + A[1][1] * T[1][1] + A[1][2] * T[1][2];
      cov00 += dilation;
      cov11 += dilation;
      det = cov00 * cov11 - cov01 * cov10;
      if (det <= 1e-12f) ok = false;
      if (ok) {
        o.mx = ((V.fx * pc[0]) / (pc[2])) + V.cx;
        o.my = ((V.fy * pc[1]) / (pc[2])) + V.cy;
        o.ca = ((cov11) / (det));
        o.cb = ((-cov01) / (det));
        o.cc = ((cov00) / (det));
        o.depth = pc[2];
        const float mid = 0.5f * (cov00 + cov11);
        const float spread = __builtin_sqrtf(fmax_std(0.f, mid * mid - det));
        o.radius = 3.f * __builtin_sqrtf(fmax_std(mid + spread, 1e-6f));
        if (o.mx + o.radius < 0.f || o.mx - o.radius >= (float)V.width || o.my + o.radius < 0.f ||
            o.my - o.radius >= (float)V.height)
          ok = false;
        if (!(o.radius == o.radius) || !(o.mx == o.mx) || !(o.my == o.my)) ok = false;
        if (opacity < min_alpha) ok = false;
        if (ok) {
          o.opacity = opacity;
          cull_extents(o.ca, o.cb, o.cc, q_cull, &o.ex, &o.ey);
          const float lx2 = V.thx * kSlack, ly2 = V.thy * kSlack;
          if (fabsf(((pc[0]) / (z))) > lx2) o.flags |= 8u;
          if (fabsf(((pc[1]) / (z))) > ly2) o.flags |= 16u;
          // view direction and SH colour
          float d0 = p0 - V.C[0], d1 = p1 - V.C[1], d2 = p2 - V.C[2];
          const float dsq = d0 * d0 + d1 * d1 + d2 * d2;
          if (dsq > 0.f) {
            const float nrm = __builtin_sqrtf(dsq);
            d0 = ((d0) / (nrm)); d1 = ((d1) / (nrm)); d2 = ((d2) / (nrm));
          }
          float basis[15];
          sh_basis(kDeg, d0, d1, d2, basis);
          float col[3] = {0.5f + kShC0 * dc0, 0.5f + kShC0 * dc1, 0.5f + kShC0 * dc2};
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            float sum = 0.f;
#pragma unroll
            for (int k = 0; k < kCount; ++k) sum += basis[k] * rc[c][k];
            col[c] += sum;
          }
          for (int c = 0; c < 3; ++c) {
            if (col[c] < 0.f) {
              col[c] = 0.f;
              o.flags |= (1u << c);
            }
          }
          o.cr = col[0];
          o.cg = col[1];
          o.cbl = col[2];
          // tile rect (SplatRaster.cpp:121-126)
          const int fx0 = (int)floorf((o.mx - o.radius) / (float)kTile);
          const int fx1 = (int)floorf((o.mx + o.radius) / (float)kTile);
          const int fy0 = (int)floorf((o.my - o.radius) / (float)kTile);
          const int fy1 = (int)floorf((o.my + o.radius) / (float)kTile);
          const int x0 = max(0, fx0), x1 = min(V.cols - 1, fx1);
          const int y0 = max(0, fy0), y1 = min(V.rows - 1, fy1);
          o.x0 = (int16_t)x0; o.x1 = (int16_t)x1; o.y0 = (int16_t)y0; o.y1 = (int16_t)y1;
          const int w = x1 - x0 + 1, h = y1 - y0 + 1;
          count = (w > 0 && h > 0) ? (uint32_t)(w * h) : 0u;
          if (count)
            scount = (uint32_t)((x1 / kSuper - x0 / kSuper + 1) * (y1 / kSuper - y0 / kSuper + 1));
          o.flags |= kProjected;
          key = depth_key(o.depth);
        }
      }
    }
    if (!ok) {
      o.mx = o.my = 0.f;
      o.ca = o.cb = o.cc = 0.f;
      o.opacity = 0.f;
      o.cr = o.cg = o.cbl = 0.f;
      o.depth = 0.f;
      o.radius = 0.f;
      o.x0 = o.x1 = o.y0 = o.y1 = 0;
      o.ex = o.ey = 0.f;
    }
    rec[slot] = o;
    keys[slot] = key;
    if (key != 0xFFFFFFFFu) {
      key_min = min(key_min, key);
      key_max = max(key_max, key);
    }
    if (prects) {  // wave-uniform
      prects[slot] = count ? pack_rect(o.x0, o.x1, o.y0, o.y1) : kNoRect;
    } else {
      counts[slot] = count;
      scounts[slot] = scount;
      rects[slot] = make_uint2((uint32_t)(uint16_t)o.x0 | (uint32_t)(uint16_t)o.x1 << 16,
                               (uint32_t)(uint16_t)o.y0 | (uint32_t)(uint16_t)o.y1 << 16);
    }
  }
  if (key_ranges) {  // uniform: the wave's extremes over its active lanes, into the wave's own pair
    key_min = __ockl_wfred_min_u32(key_min);
    key_max = __ockl_wfred_max_u32(key_max);
    if ((threadIdx.x & 63) == 0) {
      const int w = g >> 6;
      key_ranges[2 * w] = key_min;
      key_ranges[2 * w + 1] = key_max;
    }
  }
}

}  // namespace

hipError_t launch_project(const float* params, long long cap, int n, int rest_stride, int sh_use,
                          const ViewDesc* views, int n_views, float dilation, float min_alpha, SplatRec* rec,
                          uint32_t* keys, uint32_t* counts, uint32_t* scounts, uint2* rects, uint32_t* prects,
                          hipStream_t stream, uint32_t* key_ranges) {
  if (n <= 0 || n_views <= 0) return hipSuccess;
  const int blocks = (n + 255) / 256;
  // the cloud's coefficients must cover the degree in use
  if (3 * (sh_use == 0 ? 0 : (sh_use == 1 ? 3 : (sh_use == 2 ? 8 : 15))) > rest_stride) return hipErrorInvalidValue;
  switch (sh_use) {
    case 0:
      hipLaunchKernelGGL(k_project<0>, dim3(blocks), dim3(256), 0, stream, params, cap, n, rest_stride, views,
                         n_views, dilation, min_alpha, rec, keys, counts, scounts, rects, prects, key_ranges);
      break;
    case 1:
      hipLaunchKernelGGL(k_project<1>, dim3(blocks), dim3(256), 0, stream, params, cap, n, rest_stride, views,
                         n_views, dilation, min_alpha, rec, keys, counts, scounts, rects, prects, key_ranges);
      break;
    case 2:
      hipLaunchKernelGGL(k_project<2>, dim3(blocks), dim3(256), 0, stream, params, cap, n, rest_stride, views,
                         n_views, dilation, min_alpha, rec, keys, counts, scounts, rects, prects, key_ranges);
      break;
    default:
      hipLaunchKernelGGL(k_project<3>, dim3(blocks), dim3(256), 0, stream, params, cap, n, rest_stride, views,
                         n_views, dilation, min_alpha, rec, keys, counts, scounts, rects, prects, key_ranges);
      break;
  }
  return hipGetLastError();
}

}  // namespace splat
