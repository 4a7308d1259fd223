// Per-gaussian adjoint: screen-space gradients back to the parameters.
//
// The per-gaussian half of the engine's rasterizeSplatsBackward
// (engine/native/cpp/src/splat/SplatRaster.cpp:435-576): opacity through the
// sigmoid, colour through the SH clamp into the coefficients and the view
// direction, the mean through the pinhole, the conic through -A^-1 dA A^-1
// into the 2D covariance, then into the 3D covariance and the frustum-clamped
// Jacobian (with the halved z-derivative of a clamped entry), scale, and the
// normalised quaternion. One thread per gaussian; it loops over the views of
// the batch, so gradients from all views are reduced in registers and written
// once -- no atomics on parameter gradients. Gradients are added (the engine's
// accumulate semantics). HBM-bound: reads params (56 + 4 rs B) and 48 B of
// screen terms per (view, gaussian), read-modify-writes gradients.
//
// In deterministic mode the screen terms of each (view, gaussian) are first
// summed from the per-tile partials the compositing adjoint wrote, in a fixed
// order that does not depend on scheduling: the partials lie in depth-sorted
// slot order, each slot's run contiguous, so a wave takes 64 consecutive
// records, sums each slot's records with a segmented scan across its lanes,
// and the pieces of runs crossing a 64-record boundary are added in chunk
// order by a second pass. Balanced by records, however uneven the runs.
#include "splat_common.h"
#include "splat_kernels.h"

namespace splat {
namespace {

constexpr float kShC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                            -1.0925484305920792f, 0.5462742152960396f};
constexpr float kShC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                            0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                            -0.5900435899266435f};

__device__ __forceinline__ int sh_basis(int degree, const float d[3], float* b) {
  if (degree < 1) return 0;
  const float x = d[0], y = d[1], z = d[2];
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

// d(basis_k)/d(direction), SplatRaster.cpp:55-76.
__device__ __forceinline__ void sh_jacobian(int count, const float d[3], float j[15][3]) {
  const float x = d[0], y = d[1], z = d[2];
  auto set = [&](int k, float s, float a, float b, float c) {
    j[k][0] = s * a;
    j[k][1] = s * b;
    j[k][2] = s * c;
  };
  set(0, 1.f, 0.f, -kShC1, 0.f);
  set(1, 1.f, 0.f, 0.f, kShC1);
  set(2, 1.f, -kShC1, 0.f, 0.f);
  if (count < 8) return;
  const float xx = x * x, yy = y * y, zz = z * z;
  set(3, kShC2[0], y, x, 0.f);
  set(4, kShC2[1], 0.f, z, y);
  set(5, kShC2[2], -2.f * x, -2.f * y, 4.f * z);
  set(6, kShC2[3], z, 0.f, x);
  set(7, kShC2[4], 2.f * x, -2.f * y, 0.f);
  if (count < 15) return;
  set(8, kShC3[0], 6.f * x * y, 3.f * xx - 3.f * yy, 0.f);
  set(9, kShC3[1], y * z, x * z, x * y);
  set(10, kShC3[2], -2.f * x * y, 4.f * zz - xx - 3.f * yy, 8.f * y * z);
  set(11, kShC3[3], -6.f * x * z, -6.f * y * z, 6.f * zz - 3.f * xx - 3.f * yy);
  set(12, kShC3[4], 4.f * zz - 3.f * xx - yy, -2.f * x * y, 8.f * x * z);
  set(13, kShC3[5], 2.f * x * z, -2.f * y * z, xx - yy);
  set(14, kShC3[6], 3.f * xx - 3.f * yy, -6.f * x * y, 0.f);
}

// The Adam step of one parameter, as k_adam takes it (SplatTrainer.cpp:56-65).
__device__ __forceinline__ float adam_delta(float& m, float& v, float g, float lr, const AdamStep& p) {
  m = p.beta1 * m + (1.f - p.beta1) * g;
  v = p.beta2 * v + (1.f - p.beta2) * g * g;
  const float m_hat = m / p.c1;
  const float v_hat = v / p.c2;
  return -lr * m_hat / (__builtin_sqrtf(v_hat) + p.eps);
}

// Adam on kN consecutive parameters at o: every operand loaded first (the
// four arrays are distinct, so the loads go out together), then the update,
// then the stores. `touched`: this adjoint added gradient (the plain adjoint
// leaves an untouched gaussian's stored gradient as it is). kClear: the
// stored gradient is known to be all zeros, so it is neither read nor cleared
// (0 + g is still added, which keeps the bits of the general form).
template <int kN, bool kClear>
__device__ __forceinline__ void adam_block(long long o, const float* gnew, bool touched, float lr, const AdamStep& P,
                                           float* __restrict__ pw, float* __restrict__ gw, float* __restrict__ mw,
                                           float* __restrict__ vw) {
  float pp[kN], gg[kN], mm[kN], vv[kN];
#pragma unroll
  for (int i = 0; i < kN; ++i) {
    pp[i] = pw[o + i];
    gg[i] = kClear ? 0.f : gw[o + i];
    mm[i] = mw[o + i];
    vv[i] = vw[o + i];
  }
#pragma unroll
  for (int i = 0; i < kN; ++i) {
    const float g = touched && gnew ? gg[i] + gnew[i] : gg[i];
    pp[i] = pp[i] + adam_delta(mm[i], vv[i], g, lr, P);
  }
#pragma unroll
  for (int i = 0; i < kN; ++i) {
    pw[o + i] = pp[i];
    mw[o + i] = mm[i];
    vw[o + i] = vv[i];
    if (!kClear) gw[o + i] = 0.f;
  }
}

// With kAdam (a training step): unless *skip is set, the gaussian's gradient
// -- the stored one plus this adjoint's, added exactly as the plain adjoint
// adds it -- goes straight into the Adam update of all its parameters (the
// quaternion renormalised after), the stored gradient is cleared and the
// step's densify statistics are folded, exactly as k_adam would do after the
// plain adjoint; the parameters are read once and the gradients never
// written back. kClear (with kAdam): the stored gradient and the step's
// stored statistics are known to be zeros (the host tracks it: every step,
// Adam step, densify, upload and clear leaves them so, a standalone adjoint
// does not), so they are neither read nor cleared: 0.49 KB less traffic per
// SH3 gaussian, and the same bits.
template <int kDeg, bool kAdam, bool kClear>
__global__ __launch_bounds__(256) void k_adjoint(AdjointParams p, const ViewDesc* __restrict__ views,
                                                 const float* params_in, const SplatRec* __restrict__ rec,
                                                 const float* __restrict__ screen, float* __restrict__ grads,
                                                 float* __restrict__ sg, int* __restrict__ vc, AdamFuse fuse) {
  const float* params = params_in;  // not restrict: the fused Adam writes the same buffer
  // the fused Adam's rows of SH rest gradients, one slice per wave
  __shared__ float s_rest[kAdam ? 4 : 1][kAdam ? 64 * 45 : 1];
  __shared__ int s_used[kAdam ? 4 : 1][kAdam ? 64 : 1];
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= p.n) return;
  const long long cap = p.cap;
  const int rs = p.rest_stride;
  const float pos[3] = {params[3 * g], params[3 * g + 1], params[3 * g + 2]};
  const float ls[3] = {params[3 * cap + 3 * g], params[3 * cap + 3 * g + 1], params[3 * cap + 3 * g + 2]};
  const float* rp = params + 6 * cap + 4 * (long long)g;
  const float w0 = rp[0], x0 = rp[1], y0 = rp[2], z0 = rp[3];
  const float* rest = params + 14 * cap + (long long)rs * g;
  float* grest = grads + 14 * cap + (long long)rs * g;

  // covariance3d pieces (SplatRaster.cpp:78-87, 539-541)
  const float sqn = x0 * x0 + y0 * y0 + z0 * z0 + w0 * w0;
  float qw = w0, qx = x0, qy = y0, qz = z0;
  if (sqn > 0.f) {
    const float s = __builtin_sqrtf(sqn);
    qx = qx / s; qy = qy / s; qz = qz / s; qw = qw / s;
  }
  float R[3][3];
  {
    const float tx = 2.f * qx, ty = 2.f * qy, tz = 2.f * qz;
    const float twx = tx * qw, twy = ty * qw, twz = tz * qw;
    const float txx = tx * qx, txy = ty * qx, txz = tz * qx;
    const float tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
    R[0][0] = 1.f - (tyy + tzz); R[0][1] = txy - twz; R[0][2] = txz + twy;
    R[1][0] = txy + twz; R[1][1] = 1.f - (txx + tzz); R[1][2] = tyz - twx;
    R[2][0] = txz - twy; R[2][1] = tyz + twx; R[2][2] = 1.f - (txx + tyy);
  }
  const float sc[3] = {det_exp(ls[0]), det_exp(ls[1]), det_exp(ls[2])};
  float M[3][3], S[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M[i][j] = R[i][j] * sc[j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) S[i][j] = M[i][0] * M[j][0] + M[i][1] * M[j][1] + M[i][2] * M[j][2];
  const float raw_norm = fmaxf(__builtin_sqrtf(sqn), 1e-12f);
  const int per_channel = rs / 3;
  constexpr int kCount = kDeg == 0 ? 0 : (kDeg == 1 ? 3 : (kDeg == 2 ? 8 : 15));
  constexpr int kC = kCount > 0 ? kCount : 1;
  float rc[3][kC], grc[3][kC];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int k = 0; k < kCount; ++k) {
      rc[c][k] = rest[c * per_channel + k];
      grc[c][k] = 0.f;
    }

  float gpos[3] = {0.f, 0.f, 0.f}, gls[3] = {0.f, 0.f, 0.f}, grot[4] = {0.f, 0.f, 0.f, 0.f};
  float gop = 0.f, gdc[3] = {0.f, 0.f, 0.f};
  float sg_sum = 0.f;
  int vc_sum = 0;

  // The next view's record fields and screen gradients are loaded while this
  // view is processed (both buffers hold every slot of the batch).
  struct Next {
    float4 r0, r1, s0, s1, s2;
    uint32_t flags;
  };
  auto fetch = [&](int v) {
    Next x;
    const long long slot = (long long)v * p.n + g;
    const float4* rr = reinterpret_cast<const float4*>(rec + slot);
    x.r0 = rr[0];  // mx, my, ca, cb
    x.r1 = rr[1];  // cc, opacity, cr, cg
    x.flags = rec[slot].flags;
    const float4* src = reinterpret_cast<const float4*>(screen + slot * kScreenStride);
    x.s0 = src[0];
    x.s1 = src[1];
    x.s2 = src[2];
    return x;
  };
  Next next = fetch(0);
  for (int v = 0; v < p.n_views; ++v) {
    const Next cur = next;
    if (v + 1 < p.n_views) next = fetch(v + 1);
    struct {
      float ca, cb, cc, opacity;
      uint32_t flags;
    } r = {cur.r0.z, cur.r0.w, cur.r1.x, cur.r1.y, cur.flags};
    if (!(r.flags & kProjected)) continue;
    const ViewDesc& V = views[v];
    vc_sum += 1;
    float acc[11];
    {
      const float4 a = cur.s0, b = cur.s1, c = cur.s2;
      acc[0] = a.x;
      acc[1] = a.y;
      acc[2] = a.z;
      acc[3] = a.w;
      acc[4] = b.x;
      acc[5] = b.y;
      acc[6] = b.z;
      acc[7] = b.w;
      acc[8] = c.x;
      acc[9] = c.y;
      acc[10] = c.z;
    }
    sg_sum += acc[6];
    if (acc[10] == 0.f) continue;  // not touched

    const float opacity = r.opacity;
    gop += acc[3] * opacity * (1.f - opacity);
    float dcol[3] = {acc[0], acc[1], acc[2]};
    for (int c = 0; c < 3; ++c) {
      if (r.flags & (1u << c)) dcol[c] = 0.f;
      gdc[c] += kShC0 * dcol[c];
    }
    float ddir[3] = {0.f, 0.f, 0.f};
    const float off[3] = {pos[0] - V.C[0], pos[1] - V.C[1], pos[2] - V.C[2]};
    const float dist = __builtin_sqrtf(off[0] * off[0] + off[1] * off[1] + off[2] * off[2]);
    if (kDeg >= 1 && dist > 1e-9f && per_channel > 0) {
      const float dir[3] = {off[0] / dist, off[1] / dist, off[2] / dist};
      float basis[15], jac[15][3];
      sh_basis(kDeg, dir, basis);
      sh_jacobian(kCount, dir, jac);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (dcol[c] == 0.f) continue;
#pragma unroll
        for (int b = 0; b < kCount; ++b) {
          grc[c][b] += basis[b] * dcol[c];
          const float k = rc[c][b] * dcol[c];
          ddir[0] += jac[b][0] * k;
          ddir[1] += jac[b][1] * k;
          ddir[2] += jac[b][2] * k;
        }
      }
    }
    const float dmx = acc[4], dmy = acc[5];
    float pc[3];
    for (int i = 0; i < 3; ++i) pc[i] = V.W[3 * i] * pos[0] + V.W[3 * i + 1] * pos[1] + V.W[3 * i + 2] * pos[2] + V.t[i];
    const float z = pc[2];
    float dp[3] = {V.fx / z * dmx, V.fy / z * dmy, -(V.fx * pc[0] * dmx + V.fy * pc[1] * dmy) / (z * z)};
    // conic -> cov2d: D = -(C G C)
    const float C[2][2] = {{r.ca, r.cb}, {r.cb, r.cc}};
    const float Gm[2][2] = {{acc[7], 0.5f * acc[8]}, {0.5f * acc[8], acc[9]}};
    float CG[2][2], D[2][2];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) CG[i][j] = C[i][0] * Gm[0][j] + C[i][1] * Gm[1][j];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) D[i][j] = -(CG[i][0] * C[0][j] + CG[i][1] * C[1][j]);
    // projection Jacobian at the clamped point
    const float lx = V.thx * kSlack, ly = V.thy * kSlack;
    const float qx_ = pc[0] / z, qy_ = pc[1] / z;
    const float jxJ = ((qx_ < -lx) ? -lx : ((lx < qx_) ? lx : qx_)) * z;
    const float jyJ = ((qy_ < -ly) ? -ly : ((ly < qy_) ? ly : qy_)) * z;
    const float J00 = V.fx / z, J02 = -V.fx * jxJ / (z * z);
    const float J11 = V.fy / z, J12 = -V.fy * jyJ / (z * z);
    float T[2][3];
    for (int c = 0; c < 3; ++c) {
      T[0][c] = J00 * V.W[c] + 0.f * V.W[3 + c] + J02 * V.W[6 + c];
      T[1][c] = 0.f * V.W[c] + J11 * V.W[3 + c] + J12 * V.W[6 + c];
    }
    float TtD[3][2], dS3[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 2; ++j) TtD[i][j] = T[0][i] * D[0][j] + T[1][i] * D[1][j];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) dS3[i][j] = TtD[i][0] * T[0][j] + TtD[i][1] * T[1][j];
    float D2T[2][3], dT[2][3], dJ[2][3];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 3; ++j) D2T[i][j] = (2.f * D[i][0]) * T[0][j] + (2.f * D[i][1]) * T[1][j];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 3; ++j) dT[i][j] = D2T[i][0] * S[0][j] + D2T[i][1] * S[1][j] + D2T[i][2] * S[2][j];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 3; ++j)
        dJ[i][j] = dT[i][0] * V.W[3 * j] + dT[i][1] * V.W[3 * j + 1] + dT[i][2] * V.W[3 * j + 2];
    const bool cx = (r.flags & 8u) != 0, cy = (r.flags & 16u) != 0;
    const float jx = cx ? copysignf(lx, pc[0]) * z : pc[0];
    const float jy = cy ? copysignf(ly, pc[1]) * z : pc[1];
    if (!cx) dp[0] += dJ[0][2] * (-V.fx / (z * z));
    if (!cy) dp[1] += dJ[1][2] * (-V.fy / (z * z));
    const float dz_x = (cx ? 1.f : 2.f) * V.fx * jx / (z * z * z);
    const float dz_y = (cy ? 1.f : 2.f) * V.fy * jy / (z * z * z);
    dp[2] += dJ[0][0] * (-V.fx / (z * z)) + dJ[1][1] * (-V.fy / (z * z)) + dJ[0][2] * dz_x + dJ[1][2] * dz_y;
    float dpos[3];
    for (int i = 0; i < 3; ++i) dpos[i] = V.W[i] * dp[0] + V.W[3 + i] * dp[1] + V.W[6 + i] * dp[2];
    const float ddsq = ddir[0] * ddir[0] + ddir[1] * ddir[1] + ddir[2] * ddir[2];
    if (dist > 1e-9f && ddsq > 0.f) {
      const float dir[3] = {off[0] / dist, off[1] / dist, off[2] / dist};
      const float dot = dir[0] * ddir[0] + dir[1] * ddir[1] + dir[2] * ddir[2];
      for (int i = 0; i < 3; ++i) dpos[i] += (ddir[i] - dir[i] * dot) / dist;
    }
    for (int i = 0; i < 3; ++i) gpos[i] += dpos[i];
    // cov3d -> scale and rotation
    float dM[3][3];
    {
      float sym[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) sym[i][j] = 0.5f * (dS3[i][j] + dS3[j][i]);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          dM[i][j] = 2.f * sym[i][0] * M[0][j] + 2.f * sym[i][1] * M[1][j] + 2.f * sym[i][2] * M[2][j];
    }
    for (int a = 0; a < 3; ++a) gls[a] += (R[0][a] * dM[0][a] + R[1][a] * dM[1][a] + R[2][a] * dM[2][a]) * sc[a];
    float dR[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) dR[i][j] = dM[i][j] * sc[j];
    float dq[4];
    dq[0] = 2.f * (-qz * dR[0][1] + qy * dR[0][2] + qz * dR[1][0] - qx * dR[1][2] - qy * dR[2][0] + qx * dR[2][1]);
    dq[1] = 2.f * (qy * dR[0][1] + qz * dR[0][2] + qy * dR[1][0] - 2.f * qx * dR[1][1] - qw * dR[1][2] +
                   qz * dR[2][0] + qw * dR[2][1] - 2.f * qx * dR[2][2]);
    dq[2] = 2.f * (-2.f * qy * dR[0][0] + qx * dR[0][1] + qw * dR[0][2] + qx * dR[1][0] + qz * dR[1][2] -
                   qw * dR[2][0] + qz * dR[2][1] - 2.f * qy * dR[2][2]);
    dq[3] = 2.f * (-2.f * qz * dR[0][0] - qw * dR[0][1] + qx * dR[0][2] + qw * dR[1][0] - 2.f * qz * dR[1][1] +
                   qy * dR[1][2] + qx * dR[2][0] + qy * dR[2][1]);
    const float unit[4] = {qw, qx, qy, qz};
    const float ud = unit[0] * dq[0] + unit[1] * dq[1] + unit[2] * dq[2] + unit[3] * dq[3];
    for (int a = 0; a < 4; ++a) grot[a] += (dq[a] - unit[a] * ud) / raw_norm;
  }
  if (kAdam && !(fuse.skip && *fuse.skip)) {
    const AdamStep P = *fuse.step;
    float* __restrict__ pw = fuse.params;  // the same buffer as params, written here
    float* __restrict__ gw = grads;
    float* __restrict__ mw = fuse.m;
    float* __restrict__ vw = fuse.v;
    const bool touched = vc_sum != 0;
    adam_block<3, kClear>(3 * (long long)g, gpos, touched, P.lr[0], P, pw, gw, mw, vw);
    adam_block<3, kClear>(3 * cap + 3 * (long long)g, gls, touched, P.lr[1], P, pw, gw, mw, vw);
    {
      const long long o = 6 * cap + 4 * (long long)g;
      float q[4], gg[4], mm[4], vv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        q[a] = pw[o + a];
        gg[a] = kClear ? 0.f : gw[o + a];
        mm[a] = mw[o + a];
        vv[a] = vw[o + a];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) q[a] = q[a] + adam_delta(mm[a], vv[a], touched ? gg[a] + grot[a] : gg[a], P.lr[2], P);
      // Eigen norm() / normalize() over the (x, y, z, w) coefficients
      const float sq = q[1] * q[1] + q[2] * q[2] + q[3] * q[3] + q[0] * q[0];
      const float nrm = __builtin_sqrtf(sq);
      if (nrm > 1e-6f) {
#pragma unroll
        for (int a = 0; a < 4; ++a) q[a] = q[a] / nrm;
      } else {
        q[0] = 1.f;
        q[1] = q[2] = q[3] = 0.f;
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        pw[o + a] = q[a];
        mw[o + a] = mm[a];
        vw[o + a] = vv[a];
        if (!kClear) gw[o + a] = 0.f;
      }
    }
    adam_block<1, kClear>(10 * cap + g, &gop, touched, P.lr[3], P, pw, gw, mw, vw);
    adam_block<3, kClear>(11 * cap + 3 * (long long)g, gdc, touched, P.lr[4], P, pw, gw, mw, vw);
    // The SH rest coefficients, every stored band: the wave's gaussians hold
    // one contiguous block (rs floats each), so each lane puts its new
    // gradient row in the wave's LDS slice and the wave then updates the block
    // with coalesced accesses. Active lanes are a prefix of the wave (threads
    // past n have left), and they stride the block together.
    if (rs > 0) {
      const int lane = threadIdx.x & 63;
      float* row = s_rest[threadIdx.x >> 6];
      int* used = s_used[threadIdx.x >> 6];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int k = 0; k < kCount; ++k) row[lane * rs + c * per_channel + k] = grc[c][k];
      }
      used[lane] = touched ? 1 : 0;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const int nv = __popcll(__ballot(true));
      const long long base = 14 * cap + (long long)rs * (g - lane);
      const float inv_rs = 1.f / (float)rs;
      const int total = nv * rs;
      // kU elements a round per lane (e, e + nv, ..), every load of the round
      // issued before any update: a round keeps 3 kU loads in flight where
      // one element at a time left two waves per SIMD waiting on each
      constexpr int kU = 9;
      for (int e0 = lane; e0 < total; e0 += kU * nv) {
        float pp[kU], gg[kU], mm[kU], vv[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const int e = e0 + u * nv;
          const long long o = base + (e < total ? e : lane);  // past the block: a valid element, not stored
          pp[u] = pw[o];
          gg[u] = kClear ? 0.f : gw[o];
          mm[u] = mw[o];
          vv[u] = vw[o];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const int e = e0 + u * nv;
          if (e >= total) break;
          int j = (int)((float)e * inv_rs);  // gaussian of element e (exact after the correction)
          if (j * rs > e) --j;
          else if ((j + 1) * rs <= e) ++j;
          int k = e - j * rs;  // coefficient within the gaussian's row, then within its channel
          k = k >= 2 * per_channel ? k - 2 * per_channel : (k >= per_channel ? k - per_channel : k);
          const bool add = used[j] != 0 && k < kCount;  // as the plain adjoint adds (used bands, touched gaussians)
          const long long o = base + e;
          const float gt = add ? gg[u] + row[e] : gg[u];
          pw[o] = pp[u] + adam_delta(mm[u], vv[u], gt, P.lr[5], P);
          mw[o] = mm[u];
          vw[o] = vv[u];
          if (!kClear) gw[o] = 0.f;
        }
      }
    }
    const float sg_stored = kClear ? 0.f : sg[g];
    const float sg_step = touched ? sg_stored + sg_sum : sg_stored;
    const int vc_step = (kClear ? 0 : vc[g]) + vc_sum;
    fuse.stat_sg[g] += sg_step;
    fuse.stat_vc[g] += vc_step;
    if (!kClear) {
      sg[g] = 0.f;
      vc[g] = 0;
    }
    return;
  }
  if (vc_sum == 0) return;
  for (int i = 0; i < 3; ++i) {
    grads[3 * g + i] += gpos[i];
    grads[3 * cap + 3 * g + i] += gls[i];
    grads[11 * cap + 3 * g + i] += gdc[i];
  }
  for (int a = 0; a < 4; ++a) grads[6 * cap + 4 * (long long)g + a] += grot[a];
  grads[10 * cap + g] += gop;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int k = 0; k < kCount; ++k) grest[c * per_channel + k] += grc[c][k];
  sg[g] += sg_sum;
  vc[g] += vc_sum;
}

}  // namespace

namespace {

constexpr int kDrChunk = 64;  // partial records per wave (one per lane)

struct Vals11 {
  float v[11];
};

__device__ __forceinline__ void store_screen(float* dst_f, const float (&v)[11]) {
  float4* dst = reinterpret_cast<float4*>(dst_f);
  dst[0] = make_float4(v[0], v[1], v[2], v[3]);
  dst[1] = make_float4(v[4], v[5], v[6], v[7]);
  dst[2] = make_float4(v[8], v[9], v[10], 0.f);
}

// Level 0: the owner of every chunk's first record -- the sorted position
// whose run [offsets[o], offsets[o + 1]) holds record 64 c. Each chunk start
// lies in exactly one non-empty run, so every entry below the record total is
// written once, by that run's thread.
// The depth-sorted positions in use: min(m, *m_dev).
__device__ __forceinline__ long long sorted_count(long long m, const uint32_t* m_dev) {
  return m_dev ? min(m, (long long)*m_dev) : m;
}

__global__ __launch_bounds__(256) void k_det_chunk_owners(const uint32_t* __restrict__ offsets,
                                                          const uint32_t* __restrict__ p_total, long long m_cap,
                                                          const uint32_t* __restrict__ m_dev, uint32_t p_cap,
                                                          uint32_t* __restrict__ chunk_owner) {
  const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long m = sorted_count(m_cap, m_dev);
  if (o >= m) return;
  const uint32_t total = min(*p_total, p_cap);
  const uint32_t a = min(offsets[o], total), b = o + 1 < m ? min(offsets[o + 1], total) : total;
  for (uint32_t c = (a + kDrChunk - 1) / kDrChunk; c * kDrChunk < b; ++c) chunk_owner[c] = (uint32_t)o;
}

// Level 1: a wave per 64 consecutive partial records. Records are ordered by
// depth-sorted slot position, each slot's run contiguous, so the lanes' owners
// are non-decreasing; a segmented inclusive scan across the wave (fixed shape)
// sums each owner's records in the chunk. An owner whose run lies inside the
// chunk is written to its screen slot; the pieces of runs that cross a chunk
// edge go to carry_first (the chunk's first owner) / carry_last (its last).
__global__ __launch_bounds__(256) void k_det_chunks(const uint32_t* __restrict__ sorted_slots,
                                                    const uint32_t* __restrict__ offsets,
                                                    const uint32_t* __restrict__ p_total, long long m_cap,
                                                    const uint32_t* __restrict__ m_dev, uint32_t p_cap,
                                                    const uint32_t* __restrict__ chunk_owner,
                                                    const float* __restrict__ partial,
                                                    const uint8_t* __restrict__ visit,
                                                    const uint32_t* __restrict__ epoch, float* __restrict__ screen,
                                                    float* __restrict__ carry_first, float* __restrict__ carry_last) {
  const int lane = threadIdx.x & 63;
  const long long m = sorted_count(m_cap, m_dev);
  const long long c = (long long)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const uint32_t total = min(*p_total, p_cap);  // records past p_cap exist only in a flagged step
  const uint32_t tag = *epoch;
  const long long r0 = c * kDrChunk;
  if (r0 >= total || m <= 0) return;  // wave-uniform
  // owner of r0: the last sorted position whose run starts at or before it
  const long long o0 = chunk_owner[c];  // owners of this chunk: o0 .. o0 + 63 (every run holds a record)
  auto off_at = [&](long long k) -> uint32_t { return k < m ? min(offsets[k], total) : total; };
  const uint32_t my_off = off_at(o0 + lane);      // run start of candidate owner o0 + lane
  const uint32_t next_off = off_at(o0 + lane + 1);
  const long long r = r0 + lane;
  const bool valid = r < total;
  // this lane's record's owner: the last candidate whose run starts at or before r
  int own = 0;
#pragma unroll
  for (int step = 32; step > 0; step >>= 1) {
    const int cand = own + step;
    const uint32_t oc = __shfl(my_off, cand & 63, 64);
    if (cand < 64 && oc <= (uint32_t)r) own = cand;
  }
  float v[11];
#pragma unroll
  for (int k = 0; k < 11; ++k) v[k] = 0.f;
  // a pair the adjoint did not visit this pass holds an older tag: zeros.
  // Its visit byte is then (but for a wrap of the byte) another than the
  // tag's low byte, and the record is not loaded; the tag decides.
  if (valid && visit[r] == (uint8_t)tag) {
    const float4* src = reinterpret_cast<const float4*>(partial + r * kPartialStride);
    const float4 a = src[0], b = src[1], cc = src[2];
    if (__float_as_uint(cc.w) == tag) {
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
      v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
      v[8] = cc.x; v[9] = cc.y; v[10] = cc.z;
    }
  }
  const int key = valid ? own : 64 + lane;  // invalid lanes: segments of their own
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int other = __shfl_up(key, d, 64);
    const bool join = lane >= d && other == key;
#pragma unroll
    for (int k = 0; k < 11; ++k) {
      const float x = __shfl_up(v[k], d, 64);
      if (join) v[k] += x;
    }
  }
  const int next_key = __shfl_down(key, 1, 64);
  // the owner's run bounds, fetched while every lane is still active
  const uint32_t start = __shfl(my_off, own, 64), stop = __shfl(next_off, own, 64);
  const bool seg_last = valid && (lane == 63 || next_key != key);
  if (!seg_last) return;
  const bool starts_inside = start >= (uint32_t)r0;
  const bool ends_inside = stop <= (uint32_t)(r0 + kDrChunk);
  if (starts_inside && ends_inside) {
    store_screen(screen + (long long)sorted_slots[o0 + own] * kScreenStride, v);
  } else {
    if (!starts_inside) store_screen(carry_first + c * kScreenStride, v);
    if (!ends_inside) store_screen(carry_last + c * kScreenStride, v);
  }
}

// Level 2: one thread per depth-sorted position. Runs that crossed a chunk
// edge are summed from the carries in chunk order; slots without records get
// zeros.
__global__ __launch_bounds__(256) void k_det_owners(const uint32_t* __restrict__ sorted_slots,
                                                    const uint32_t* __restrict__ offsets,
                                                    const uint32_t* __restrict__ p_total, long long m_cap,
                                                    const uint32_t* __restrict__ m_dev, uint32_t p_cap,
                                                    const float* __restrict__ carry_first,
                                                    const float* __restrict__ carry_last, float* __restrict__ screen) {
  const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long m = sorted_count(m_cap, m_dev);
  if (o >= m) return;
  const uint32_t total = min(*p_total, p_cap);
  const uint32_t a = min(offsets[o], total), b = o + 1 < m ? min(offsets[o + 1], total) : total;
  float v[11];
#pragma unroll
  for (int k = 0; k < 11; ++k) v[k] = 0.f;
  float* dst = screen + (long long)sorted_slots[o] * kScreenStride;
  if (a == b) {
    store_screen(dst, v);
    return;
  }
  const long long ca = a / kDrChunk, cb = (b - 1) / kDrChunk;
  if (ca == cb) return;  // written by the chunk pass
  auto add = [&](const float* src_f) {
    const float4* src = reinterpret_cast<const float4*>(src_f);
    const float4 x = src[0], y = src[1], z = src[2];
    v[0] += x.x; v[1] += x.y; v[2] += x.z; v[3] += x.w;
    v[4] += y.x; v[5] += y.y; v[6] += y.z; v[7] += y.w;
    v[8] += z.x; v[9] += z.y; v[10] += z.z;
  };
  add(carry_last + ca * kScreenStride);
  for (long long c = ca + 1; c <= cb; ++c) add(carry_first + c * kScreenStride);
  store_screen(dst, v);
}

__global__ void k_bump_epoch(uint32_t* epoch) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const uint32_t e = *epoch + 1u;
    *epoch = e ? e : 1u;
  }
}

}  // namespace

hipError_t launch_bump_epoch(uint32_t* epoch, hipStream_t stream) {
  hipLaunchKernelGGL(k_bump_epoch, dim3(1), dim3(64), 0, stream, epoch);
  return hipGetLastError();
}

// carry_first, carry_last (kScreenStride floats a chunk each), then the
// chunk owners (one u32 a chunk)
long long det_reduce_carry_floats(unsigned long long p_cap) {
  return (2LL * kScreenStride + 1) * (long long)((p_cap + kDrChunk - 1) / kDrChunk + 1);
}

hipError_t launch_det_reduce(const uint32_t* sorted_slots, const uint32_t* offsets, const uint32_t* p_total,
                             long long m, const uint32_t* m_dev, uint32_t p_cap, const float* partial,
                             const uint8_t* visit, const uint32_t* epoch, float* screen, float* carry,
                             hipStream_t stream) {
  if (m <= 0) return hipSuccess;
  const long long chunks = ((long long)p_cap + kDrChunk - 1) / kDrChunk + 1;
  float* carry_first = carry;
  float* carry_last = carry + kScreenStride * chunks;
  uint32_t* chunk_owner = reinterpret_cast<uint32_t*>(carry + 2LL * kScreenStride * chunks);
  if (p_cap > 0) {
    hipLaunchKernelGGL(k_det_chunk_owners, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream, offsets, p_total,
                       m, m_dev, p_cap, chunk_owner);
    SPLAT_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_det_chunks, dim3((unsigned)((chunks + 3) / 4)), dim3(256), 0, stream, sorted_slots, offsets,
                       p_total, m, m_dev, p_cap, chunk_owner, partial, visit, epoch, screen, carry_first,
                       carry_last);
    SPLAT_HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_det_owners, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream, sorted_slots, offsets,
                     p_total, m, m_dev, p_cap, carry_first, carry_last, screen);
  return hipGetLastError();
}

namespace {
template <int kDeg>
void adjoint_variant(const AdjointParams& p, const ViewDesc* views, const float* params, const SplatRec* rec,
                     const float* screen, float* grads, float* sg, int* vc, const AdamFuse* fuse,
                     hipStream_t stream) {
  const dim3 grid((p.n + 255) / 256);
  if (fuse && fuse->stored_clear)
    hipLaunchKernelGGL((k_adjoint<kDeg, true, true>), grid, dim3(256), 0, stream, p, views, params, rec, screen, grads,
                       sg, vc, *fuse);
  else if (fuse)
    hipLaunchKernelGGL((k_adjoint<kDeg, true, false>), grid, dim3(256), 0, stream, p, views, params, rec, screen,
                       grads, sg, vc, *fuse);
  else
    hipLaunchKernelGGL((k_adjoint<kDeg, false, false>), grid, dim3(256), 0, stream, p, views, params, rec, screen,
                       grads, sg, vc, AdamFuse{});
}
}  // namespace

hipError_t launch_adjoint(const AdjointParams& p, const ViewDesc* views, const float* params, const SplatRec* rec,
                          const float* screen, float* grads, float* sg, int* vc, hipStream_t stream,
                          const AdamFuse* fuse) {
  if (p.n <= 0) return hipSuccess;
  if (3 * (p.sh_use == 0 ? 0 : (p.sh_use == 1 ? 3 : (p.sh_use == 2 ? 8 : 15))) > p.rest_stride)
    return hipErrorInvalidValue;
  switch (p.sh_use) {
    case 0: adjoint_variant<0>(p, views, params, rec, screen, grads, sg, vc, fuse, stream); break;
    case 1: adjoint_variant<1>(p, views, params, rec, screen, grads, sg, vc, fuse, stream); break;
    case 2: adjoint_variant<2>(p, views, params, rec, screen, grads, sg, vc, fuse, stream); break;
    default: adjoint_variant<3>(p, views, params, rec, screen, grads, sg, vc, fuse, stream); break;
  }
  return hipGetLastError();
}

}  // namespace splat
