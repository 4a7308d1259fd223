// CPU oracle for the Gaussian-splat training step -- TEST INFRASTRUCTURE ONLY.
// See splat_oracle.h for what it restates and how it is pinned. Compiled
// strict: g++ -O2 -ffp-contract=off, no fast-math. All sums of products are
// evaluated left to right; the reference build (-ffast-math) fixes no order,
// so those choices are ulp-level and are shared with the HIP kernels.
#include "splat_oracle.h"

#include "../parallax_amd/csrc/splat_exp.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <type_traits>
#include <vector>

namespace {

constexpr int kTile = 16;             // SplatRaster.cpp:18
constexpr float kMaxAlpha = 0.99f;    // SplatRaster.cpp:22
constexpr float kSlack = 1.3f;        // SplatRaster.cpp:26
constexpr float kShC0 = 0.28209479177387814f;  // GaussianCloud.h:54
constexpr float kShC1 = 0.4886025119029199f;   // SplatRaster.cpp:11-16
constexpr float kShC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                            -1.0925484305920792f, 0.5462742152960396f};
constexpr float kShC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                            0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                            -0.5900435899266435f};

inline float dexp(float x) { return splat_det::exp(x); }
inline double dexp(double x) { return std::exp(x); }
inline float sigm(float x) { return 1.0f / (1.0f + dexp(-x)); }
inline double sigm(double x) { return 1.0 / (1.0 + std::exp(-x)); }

// ---------------------------------------------------------------- SH basis --
// shBasis / shBasisJacobian, SplatRaster.cpp:29-76. Templated so the FP64
// path uses the same expressions with double constants.
template <typename R>
int shBasis(int degree, const R d[3], R* basis) {
  if (degree < 1) return 0;
  const R C1 = R(kShC1);
  const R x = d[0], y = d[1], z = d[2];
  basis[0] = -C1 * y;
  basis[1] = C1 * z;
  basis[2] = -C1 * x;
  if (degree < 2) return 3;
  const R xx = x * x, yy = y * y, zz = z * z;
  const R xy = x * y, yz = y * z, xz = x * z;
  basis[3] = R(kShC2[0]) * xy;
  basis[4] = R(kShC2[1]) * yz;
  basis[5] = R(kShC2[2]) * (R(2) * zz - xx - yy);
  basis[6] = R(kShC2[3]) * xz;
  basis[7] = R(kShC2[4]) * (xx - yy);
  if (degree < 3) return 8;
  basis[8] = R(kShC3[0]) * y * (R(3) * xx - yy);
  basis[9] = R(kShC3[1]) * xy * z;
  basis[10] = R(kShC3[2]) * y * (R(4) * zz - xx - yy);
  basis[11] = R(kShC3[3]) * z * (R(2) * zz - R(3) * xx - R(3) * yy);
  basis[12] = R(kShC3[4]) * x * (R(4) * zz - xx - yy);
  basis[13] = R(kShC3[5]) * z * (xx - yy);
  basis[14] = R(kShC3[6]) * x * (xx - R(3) * yy);
  return 15;
}

template <typename R>
void shBasisJacobian(int count, const R d[3], R jac[15][3]) {
  if (count < 1) return;
  const R x = d[0], y = d[1], z = d[2];
  const R C1 = R(kShC1);
  auto set = [&](int k, R s, R a, R b, R c) {
    jac[k][0] = s * a;
    jac[k][1] = s * b;
    jac[k][2] = s * c;
  };
  jac[0][0] = R(0); jac[0][1] = -C1; jac[0][2] = R(0);
  jac[1][0] = R(0); jac[1][1] = R(0); jac[1][2] = C1;
  jac[2][0] = -C1; jac[2][1] = R(0); jac[2][2] = R(0);
  if (count < 8) return;
  const R xx = x * x, yy = y * y, zz = z * z;
  set(3, R(kShC2[0]), y, x, R(0));
  set(4, R(kShC2[1]), R(0), z, y);
  set(5, R(kShC2[2]), R(-2) * x, R(-2) * y, R(4) * z);
  set(6, R(kShC2[3]), z, R(0), x);
  set(7, R(kShC2[4]), R(2) * x, R(-2) * y, R(0));
  if (count < 15) return;
  set(8, R(kShC3[0]), R(6) * x * y, R(3) * xx - R(3) * yy, R(0));
  set(9, R(kShC3[1]), y * z, x * z, x * y);
  set(10, R(kShC3[2]), R(-2) * x * y, R(4) * zz - xx - R(3) * yy, R(8) * y * z);
  set(11, R(kShC3[3]), R(-6) * x * z, R(-6) * y * z, R(6) * zz - R(3) * xx - R(3) * yy);
  set(12, R(kShC3[4]), R(4) * zz - R(3) * xx - yy, R(-2) * x * y, R(8) * x * z);
  set(13, R(kShC3[5]), R(2) * x * z, R(-2) * y * z, xx - yy);
  set(14, R(kShC3[6]), R(3) * xx - R(3) * yy, R(-6) * x * y, R(0));
}

// ------------------------------------------------------------- view setup --
template <typename R>
struct ViewR {
  R W[3][3];
  R t[3];
  R C[3];
  R fx, fy, cx, cy;
  R thx, thy;  // tan of the half field of view
  int width, height;
};

// SplatRaster.cpp:206-212 (float) -- the pose is cast per view. The FP64 path
// keeps the pose in double.
template <typename R>
ViewR<R> makeView(const orc_view& v) {
  ViewR<R> o;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) o.W[i][j] = static_cast<R>(v.R[i * 3 + j]);
    o.t[i] = static_cast<R>(v.t[i]);
    // center() = -R^T t (SE3.h:49), evaluated in double, then cast.
    const double c = (-v.R[0 * 3 + i]) * v.t[0] + (-v.R[1 * 3 + i]) * v.t[1] +
                     (-v.R[2 * 3 + i]) * v.t[2];
    o.C[i] = static_cast<R>(c);
  }
  o.fx = static_cast<R>(v.fx);
  o.fy = static_cast<R>(v.fy);
  o.cx = static_cast<R>(v.cx);
  o.cy = static_cast<R>(v.cy);
  o.thx = static_cast<R>(0.5 * v.width / v.fx);
  o.thy = static_cast<R>(0.5 * v.height / v.fy);
  o.width = v.width;
  o.height = v.height;
  return o;
}

// Normalised quaternion to rotation matrix: Eigen normalized() then
// toRotationMatrix() (Quaternion.h:592-622). q is stored (w, x, y, z).
template <typename R>
void rotationOf(const R q[4], R Rm[3][3]) {
  R w = q[0], x = q[1], y = q[2], z = q[3];
  const R sq = x * x + y * y + z * z + w * w;
  if (sq > R(0)) {
    const R s = std::sqrt(sq);
    x = x / s; y = y / s; z = z / s; w = w / s;
  }
  const R tx = R(2) * x, ty = R(2) * y, tz = R(2) * z;
  const R twx = tx * w, twy = ty * w, twz = tz * w;
  const R txx = tx * x, txy = ty * x, txz = tz * x;
  const R tyy = ty * y, tyz = tz * y, tzz = tz * z;
  Rm[0][0] = R(1) - (tyy + tzz);
  Rm[0][1] = txy - twz;
  Rm[0][2] = txz + twy;
  Rm[1][0] = txy + twz;
  Rm[1][1] = R(1) - (txx + tzz);
  Rm[1][2] = tyz - twx;
  Rm[2][0] = txz - twy;
  Rm[2][1] = tyz + twx;
  Rm[2][2] = R(1) - (txx + tyy);
}

// covariance3d, SplatRaster.cpp:82-87.
template <typename R>
void covariance3d(const R q[4], const R log_scale[3], R Sigma[3][3], R Rm[3][3], R s[3],
                  R M[3][3]) {
  rotationOf(q, Rm);
  for (int i = 0; i < 3; ++i) s[i] = dexp(log_scale[i]);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M[i][j] = Rm[i][j] * s[j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Sigma[i][j] = M[i][0] * M[j][0] + M[i][1] * M[j][1] + M[i][2] * M[j][2];
}

// ------------------------------------------------------------- projection --
struct ProjF {
  int index = -1;
  float mean[2] = {0, 0};
  float cov2d[3] = {0, 0, 0};
  float conic[3] = {0, 0, 0};
  float depth = 0.f;
  float radius = 0.f;
  float color[3] = {0, 0, 0};
  float opacity = 0.f;
  bool clamped[3] = {false, false, false};
  bool fclamp[2] = {false, false};  // frustum clamp of the Jacobian, per axis
  int rect[4] = {0, -1, 0, -1};     // x0, x1, y0, y1 in 16x16 tiles, inclusive
};

template <typename R>
struct ProjR {
  R mean[2];
  R conic[3];
  R opacity;
  R color[3];
};

struct CloudView {
  int n, sh_degree, rs;
  const float *pos, *ls, *rot, *op, *dc, *rest;
};

CloudView cv(const orc_cloud* c) {
  return {c->n, c->sh_degree, c->rest_stride, c->pos, c->log_scale, c->rot, c->opacity, c->dc, c->rest};
}

// evaluateSh, SplatRaster.cpp:179-199.
template <typename R>
void evalSh(const R dc[3], const R* rest, int rest_stride, int degree, const R dir[3], R color[3],
            bool clamped[3]) {
  for (int c = 0; c < 3; ++c) color[c] = R(0.5f) + R(kShC0) * dc[c];
  R basis[15];
  const int count = shBasis(degree, dir, basis);
  const int stride = rest_stride / 3;
  const int usable = std::min(count, stride);
  for (int c = 0; c < 3; ++c) {
    const R* r = rest + static_cast<size_t>(c) * stride;
    R sum = R(0);
    for (int k = 0; k < usable; ++k) sum += basis[k] * r[k];
    color[c] += sum;
  }
  for (int c = 0; c < 3; ++c) {
    const bool negative = color[c] < R(0);
    if (clamped) clamped[c] = negative;
    if (negative) color[c] = R(0);
  }
}

// Camera transform p_cam = W p + t.
template <typename R>
void toCamera(const ViewR<R>& v, const R p[3], R pc[3]) {
  for (int i = 0; i < 3; ++i) pc[i] = v.W[i][0] * p[0] + v.W[i][1] * p[1] + v.W[i][2] * p[2] + v.t[i];
}

// J * W for the 2x3 pinhole Jacobian whose x-, y-entries are `jx`, `jy`.
template <typename R>
void jacobianTimesW(const ViewR<R>& v, R z, R jx, R jy, R J[2][3], R T[2][3]) {
  J[0][0] = v.fx / z; J[0][1] = R(0); J[0][2] = -v.fx * jx / (z * z);
  J[1][0] = R(0); J[1][1] = v.fy / z; J[1][2] = -v.fy * jy / (z * z);
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c) T[r][c] = J[r][0] * v.W[0][c] + J[r][1] * v.W[1][c] + J[r][2] * v.W[2][c];
}

// T Sigma T^T, evaluated as (T Sigma) T^T.
template <typename R>
void projectCov(const R T[2][3], const R S[3][3], R cov[2][2]) {
  R A[2][3];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c) A[r][c] = T[r][0] * S[0][c] + T[r][1] * S[1][c] + T[r][2] * S[2][c];
  for (int r = 0; r < 2; ++r)
    for (int q = 0; q < 2; ++q) cov[r][q] = A[r][0] * T[q][0] + A[r][1] * T[q][1] + A[r][2] * T[q][2];
}

// projectSplats for one gaussian, SplatRaster.cpp:214-257. Returns false when
// culled.
bool projectOne(const CloudView& cl, int i, const ViewR<float>& v, const orc_options& opt,
                int sh_use, ProjF* out) {
  const float* p = cl.pos + 3 * i;
  float pc[3];
  toCamera(v, p, pc);
  if (pc[2] < 0.01f) return false;
  const float z = pc[2];
  const float lx = v.thx * kSlack, ly = v.thy * kSlack;
  const float jx = std::clamp(pc[0] / z, -lx, lx) * z;
  const float jy = std::clamp(pc[1] / z, -ly, ly) * z;
  float J[2][3], T[2][3];
  jacobianTimesW(v, z, jx, jy, J, T);
  float S[3][3], Rm[3][3], s[3], M[3][3];
  covariance3d(cl.rot + 4 * i, cl.ls + 3 * i, S, Rm, s, M);
  float cov[2][2];
  projectCov(T, S, cov);
  cov[0][0] += opt.screen_dilation;
  cov[1][1] += opt.screen_dilation;
  const float det = cov[0][0] * cov[1][1] - cov[0][1] * cov[1][0];
  if (det <= 1e-12f) return false;
  ProjF o;
  o.index = i;
  o.mean[0] = v.fx * pc[0] / pc[2] + v.cx;
  o.mean[1] = v.fy * pc[1] / pc[2] + v.cy;
  o.cov2d[0] = cov[0][0];
  o.cov2d[1] = cov[0][1];
  o.cov2d[2] = cov[1][1];
  o.conic[0] = cov[1][1] / det;
  o.conic[1] = -cov[0][1] / det;
  o.conic[2] = cov[0][0] / det;
  o.depth = pc[2];
  const float mid = 0.5f * (cov[0][0] + cov[1][1]);
  const float spread = std::sqrt(std::max(0.f, mid * mid - det));
  o.radius = 3.f * std::sqrt(std::max(mid + spread, 1e-6f));
  if (o.mean[0] + o.radius < 0.f || o.mean[0] - o.radius >= static_cast<float>(v.width) ||
      o.mean[1] + o.radius < 0.f || o.mean[1] - o.radius >= static_cast<float>(v.height))
    return false;
  o.opacity = sigm(cl.op[i]);
  if (o.opacity < opt.min_alpha) return false;
  float d[3] = {p[0] - v.C[0], p[1] - v.C[1], p[2] - v.C[2]};
  const float sq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  if (sq > 0.f) {
    const float nrm = std::sqrt(sq);
    for (float& e : d) e = e / nrm;
  }
  evalSh(cl.dc + 3 * i, cl.rest + static_cast<size_t>(cl.rs) * i, cl.rs, sh_use, d, o.color,
         o.clamped);
  // The backward's frustum-clamp decision (SplatRaster.cpp:512-515), kept
  // for the FP64 replay.
  o.fclamp[0] = std::abs(pc[0] / z) > v.thx * kSlack;
  o.fclamp[1] = std::abs(pc[1] / z) > v.thy * kSlack;
  *out = o;
  return true;
}

// The FP64 projection of one gaussian with the float pass's decisions frozen
// (SH clamps, frustum clamps); `pf` is the float record of the same gaussian.
ProjR<double> projectFrozen(const orc_cloud_d& cl, int i, const ViewR<double>& v,
                            const orc_options& opt, int sh_use, const ProjF& pf) {
  const double* p = cl.pos + 3 * i;
  double pc[3];
  toCamera(v, p, pc);
  const double z = pc[2];
  const double lx = v.thx * kSlack, ly = v.thy * kSlack;
  const double jx = pf.fclamp[0] ? std::copysign(lx, pc[0]) * z : pc[0];
  const double jy = pf.fclamp[1] ? std::copysign(ly, pc[1]) * z : pc[1];
  double J[2][3], T[2][3];
  jacobianTimesW(v, z, jx, jy, J, T);
  double S[3][3], Rm[3][3], s[3], M[3][3];
  covariance3d(cl.rot + 4 * i, cl.log_scale + 3 * i, S, Rm, s, M);
  double cov[2][2];
  projectCov(T, S, cov);
  cov[0][0] += opt.screen_dilation;
  cov[1][1] += opt.screen_dilation;
  const double det = cov[0][0] * cov[1][1] - cov[0][1] * cov[1][0];
  ProjR<double> o;
  o.mean[0] = v.fx * pc[0] / pc[2] + v.cx;
  o.mean[1] = v.fy * pc[1] / pc[2] + v.cy;
  o.conic[0] = cov[1][1] / det;
  o.conic[1] = -cov[0][1] / det;
  o.conic[2] = cov[0][0] / det;
  o.opacity = sigm(cl.opacity[i]);
  double d[3] = {p[0] - v.C[0], p[1] - v.C[1], p[2] - v.C[2]};
  const double nrm = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (nrm > 0) for (double& e : d) e = e / nrm;
  const double* dc = cl.dc + 3 * i;
  const double* rest = cl.rest + static_cast<size_t>(cl.rest_stride) * i;
  double basis[15];
  const int count = shBasis(sh_use, d, basis);
  const int stride = cl.rest_stride / 3;
  const int usable = std::min(count, stride);
  for (int c = 0; c < 3; ++c) {
    double col = 0.5 + double(kShC0) * dc[c];
    for (int k = 0; k < usable; ++k) col += basis[k] * rest[c * stride + k];
    o.color[c] = pf.clamped[c] ? 0.0 : col;
  }
  return o;
}

// ---------------------------------------------------------------- binning --
// buildTiles, SplatRaster.cpp:114-151.
struct Grid {
  int cols = 0, rows = 0;
  std::vector<int> entries;  // indices into the projected list
  std::vector<int> offsets;
};

void tileRect(const ProjF& s, int cols, int rows, int r[4]) {
  r[0] = std::max(0, static_cast<int>(std::floor((s.mean[0] - s.radius) / kTile)));
  r[1] = std::min(cols - 1, static_cast<int>(std::floor((s.mean[0] + s.radius) / kTile)));
  r[2] = std::max(0, static_cast<int>(std::floor((s.mean[1] - s.radius) / kTile)));
  r[3] = std::min(rows - 1, static_cast<int>(std::floor((s.mean[1] + s.radius) / kTile)));
}

Grid buildTiles(std::vector<ProjF>& proj, int width, int height) {
  Grid g;
  g.cols = (width + kTile - 1) / kTile;
  g.rows = (height + kTile - 1) / kTile;
  const int tiles = g.cols * g.rows;
  std::vector<int> counts(static_cast<size_t>(tiles) + 1, 0);
  for (ProjF& s : proj) {
    tileRect(s, g.cols, g.rows, s.rect);
    for (int ty = s.rect[2]; ty <= s.rect[3]; ++ty)
      for (int tx = s.rect[0]; tx <= s.rect[1]; ++tx) counts[static_cast<size_t>(ty) * g.cols + tx + 1]++;
  }
  g.offsets.assign(counts.begin(), counts.end());
  for (int i = 0; i < tiles; ++i) g.offsets[i + 1] += g.offsets[i];
  g.entries.resize(static_cast<size_t>(g.offsets[tiles]));
  std::vector<int> cursor(g.offsets.begin(), g.offsets.end() - 1);
  for (size_t i = 0; i < proj.size(); ++i) {
    const ProjF& s = proj[i];
    for (int ty = s.rect[2]; ty <= s.rect[3]; ++ty)
      for (int tx = s.rect[0]; tx <= s.rect[1]; ++tx)
        g.entries[static_cast<size_t>(cursor[static_cast<size_t>(ty) * g.cols + tx]++)] = static_cast<int>(i);
  }
  return g;
}

// ------------------------------------------------------ per-pixel evaluation --
// One (pixel, splat) evaluation of SplatRaster.cpp:294-300 in float: false
// when the splat is skipped at this pixel.
inline bool evalF(const ProjF& s, float px, float py, float min_alpha, float& dx, float& dy,
                  float& G, float& raw, float& alpha) {
  dx = px - s.mean[0];
  dy = py - s.mean[1];
  const float power = -0.5f * (s.conic[0] * dx * dx + s.conic[2] * dy * dy) - s.conic[1] * dx * dy;
  if (power > 0.f) return false;
  G = dexp(power);
  raw = s.opacity * G;
  alpha = raw < kMaxAlpha ? raw : kMaxAlpha;
  return !(alpha < min_alpha);
}

}  // namespace

// ------------------------------------------------------------------ frame --
struct orc_frame {
  orc_view view;
  orc_options opt;
  int n = 0;
  int rest_stride = 0;
  int sh_use = 0;
  ViewR<float> vf;
  std::vector<ProjF> proj;
  std::vector<int> slot_of;  // gaussian -> projected slot or -1
  Grid grid;
  std::vector<float> rgb, T;
  std::vector<int> last;
  int64_t evals[2] = {0, 0};
  // FP64 frozen projection of every projected splat, cached per cloud.
  std::vector<ProjR<double>> pr_cache;
  const void* pr_key = nullptr;
};

namespace {

// rasterizeSplats, SplatRaster.cpp:265-315, plus the last-contributor record.
void rasterize(orc_frame& f) {
  const int W = f.view.width, H = f.view.height;
  const size_t px = static_cast<size_t>(W) * H;
  f.rgb.assign(px * 3, 0.f);
  f.T.assign(px, 1.f);
  f.last.assign(px, 0);
  f.evals[0] = 0;
  if (px == 0) return;
  const Grid& g = f.grid;
  for (int ty = 0; ty < g.rows; ++ty) {
    for (int tx = 0; tx < g.cols; ++tx) {
      const int tile = ty * g.cols + tx;
      const int begin = g.offsets[tile], end = g.offsets[tile + 1];
      const int y_end = std::min(H, (ty + 1) * kTile), x_end = std::min(W, (tx + 1) * kTile);
      for (int y = ty * kTile; y < y_end; ++y) {
        for (int x = tx * kTile; x < x_end; ++x) {
          const size_t pixel = static_cast<size_t>(y) * W + x;
          const float pxf = static_cast<float>(x) + 0.5f, pyf = static_cast<float>(y) + 0.5f;
          float T = 1.f;
          float acc[3] = {0.f, 0.f, 0.f};
          int last = 0;
          for (int e = begin; e < end; ++e) {
            ++f.evals[0];
            const ProjF& s = f.proj[static_cast<size_t>(g.entries[e])];
            float dx, dy, G, raw, alpha;
            if (!evalF(s, pxf, pyf, f.opt.min_alpha, dx, dy, G, raw, alpha)) continue;
            for (int c = 0; c < 3; ++c) acc[c] += T * alpha * s.color[c];
            T *= 1.f - alpha;
            last = e + 1 - begin;
            if (T < f.opt.transmittance_cutoff) break;
          }
          for (int c = 0; c < 3; ++c) f.rgb[pixel * 3 + c] = acc[c] + T * f.opt.background[c];
          f.T[pixel] = T;
          f.last[pixel] = last;
        }
      }
    }
  }
}

// Screen-space accumulators per projected splat (SplatRaster.cpp:336-342).
template <typename R>
struct ScreenAcc {
  std::vector<R> dmean, dconic, dopacity, dcolor, dmean_abs;
  std::vector<int> touched;
  void init(size_t n) {
    dmean.assign(2 * n, R(0));
    dconic.assign(3 * n, R(0));
    dopacity.assign(n, R(0));
    dcolor.assign(3 * n, R(0));
    dmean_abs.assign(n, R(0));
    touched.assign(n, 0);
  }
};

// The pixel half of rasterizeSplatsBackward (SplatRaster.cpp:346-433) over
// one tile. Decisions always come from the float records `pf`; values come
// from `pr` (float: the same records; double: the frozen FP64 projection).
template <typename R>
void backwardTile(const orc_frame& f, int tx, int ty, const std::vector<ProjR<R>>& pr,
                  const float* dL_drgb, ScreenAcc<R>& acc, int64_t* evals) {
  const Grid& g = f.grid;
  const int W = f.view.width, H = f.view.height;
  const int tile = ty * g.cols + tx;
  const int begin = g.offsets[tile], end = g.offsets[tile + 1];
  const int y_end = std::min(H, (ty + 1) * kTile), x_end = std::min(W, (tx + 1) * kTile);
  const float min_alpha = f.opt.min_alpha, cutoff = f.opt.transmittance_cutoff;
  for (int y = ty * kTile; y < y_end; ++y) {
    for (int x = tx * kTile; x < x_end; ++x) {
      const size_t pixel = static_cast<size_t>(y) * W + x;
      const float pxf = static_cast<float>(x) + 0.5f, pyf = static_cast<float>(y) + 0.5f;
      const R px = R(pxf), py = R(pyf);
      const float* dLf = dL_drgb + pixel * 3;
      const R dL[3] = {R(dLf[0]), R(dLf[1]), R(dLf[2])};
      // Replay the forward to find the last contributor (:362-377).
      float Tf = 1.f;
      R T = R(1);
      int last = begin;
      for (int e = begin; e < end; ++e) {
        const size_t s = static_cast<size_t>(g.entries[e]);
        float dx, dy, G, raw, alpha;
        if (!evalF(f.proj[s], pxf, pyf, min_alpha, dx, dy, G, raw, alpha)) continue;
        Tf *= 1.f - alpha;
        if (std::is_same<R, float>::value) {
          T = R(Tf);
        } else {
          const ProjR<R>& q = pr[s];
          const R ddx = px - q.mean[0], ddy = py - q.mean[1];
          const R power = R(-0.5) * (q.conic[0] * ddx * ddx + q.conic[2] * ddy * ddy) - q.conic[1] * ddx * ddy;
          const R a = raw < kMaxAlpha ? q.opacity * dexp(power) : R(kMaxAlpha);
          T *= R(1) - a;
        }
        last = e + 1;
        if (Tf < cutoff) break;
      }
      const R final_T = T;
      R suffix[3] = {R(0), R(0), R(0)};
      R last_alpha = R(0);
      R last_color[3] = {R(0), R(0), R(0)};
      R bg_term = R(0);
      for (int c = 0; c < 3; ++c) bg_term += R(f.opt.background[c]) * dL[c];
      for (int e = last - 1; e >= begin; --e) {
        const size_t s = static_cast<size_t>(g.entries[e]);
        float dxf, dyf, Gf, rawf, alphaf;
        if (!evalF(f.proj[s], pxf, pyf, min_alpha, dxf, dyf, Gf, rawf, alphaf)) continue;
        if (evals) ++*evals;
        const ProjR<R>& q = pr[s];
        R dx, dy, G, alpha;
        const bool capped = rawf >= kMaxAlpha;
        if (std::is_same<R, float>::value) {
          dx = R(dxf); dy = R(dyf); G = R(Gf); alpha = R(alphaf);
        } else {
          dx = px - q.mean[0];
          dy = py - q.mean[1];
          const R power = R(-0.5) * (q.conic[0] * dx * dx + q.conic[2] * dy * dy) - q.conic[1] * dx * dy;
          G = dexp(power);
          alpha = capped ? R(kMaxAlpha) : q.opacity * G;
        }
        T /= R(1) - alpha;
        const R weight = alpha * T;
        R dL_dalpha = R(0);
        for (int c = 0; c < 3; ++c) {
          acc.dcolor[3 * s + c] += weight * dL[c];
          suffix[c] = last_alpha * last_color[c] + (R(1) - last_alpha) * suffix[c];
          last_color[c] = q.color[c];
          dL_dalpha += (q.color[c] - suffix[c]) * T * dL[c];
        }
        last_alpha = alpha;
        dL_dalpha += -final_T / (R(1) - alpha) * bg_term;
        if (capped) continue;
        acc.dopacity[s] += G * dL_dalpha;
        const R dL_dweight = q.opacity * dL_dalpha;
        const R gx = G * dx, gy = G * dy;
        const R mx = dL_dweight * (-gx * q.conic[0] - gy * q.conic[1]);
        const R my = dL_dweight * (-gy * q.conic[2] - gx * q.conic[1]);
        acc.dmean[2 * s + 0] -= mx;
        acc.dmean[2 * s + 1] -= my;
        acc.dmean_abs[s] += std::sqrt(mx * mx + my * my);
        acc.dconic[3 * s + 0] += R(-0.5) * dL_dweight * gx * dx;
        acc.dconic[3 * s + 1] += R(-1.0) * dL_dweight * gx * dy;
        acc.dconic[3 * s + 2] += R(-0.5) * dL_dweight * gy * dy;
        acc.touched[s] = 1;
      }
    }
  }
}

template <typename R>
struct GradOut {
  R *pos, *ls, *rot, *op, *dc, *rest, *sg;
  int* vc;
};

// The per-gaussian half of rasterizeSplatsBackward, SplatRaster.cpp:446-575.
// `params` gives the gaussian's parameters in R; decisions from `pf`.
template <typename R>
void perGaussian(const ViewR<R>& v, int sh_use, int rest_stride, const R* pos, const R* ls,
                 const R* rot, const R* rest, const ProjF& pf, const ProjR<R>& q, size_t s,
                 const ScreenAcc<R>& acc, GradOut<R>& out, int g) {
  out.vc[g] += 1;
  out.sg[g] += acc.dmean_abs[s];
  if (acc.touched[s] == 0) return;
  const R opacity = q.opacity;
  out.op[g] += acc.dopacity[s] * opacity * (R(1) - opacity);
  R dcol[3];
  for (int c = 0; c < 3; ++c) {
    dcol[c] = acc.dcolor[3 * s + c];
    if (pf.clamped[c]) dcol[c] = R(0);
    out.dc[3 * g + c] += R(kShC0) * dcol[c];
  }
  R ddir[3] = {R(0), R(0), R(0)};
  const int per_channel = rest_stride / 3;
  const R offset[3] = {pos[0] - v.C[0], pos[1] - v.C[1], pos[2] - v.C[2]};
  const R distance = std::sqrt(offset[0] * offset[0] + offset[1] * offset[1] + offset[2] * offset[2]);
  if (sh_use >= 1 && distance > R(1e-9f) && per_channel > 0) {
    const R dir[3] = {offset[0] / distance, offset[1] / distance, offset[2] / distance};
    R basis[15];
    const int count = shBasis(sh_use, dir, basis);
    const int usable = std::min(count, per_channel);
    R jac[15][3];
    shBasisJacobian(count, dir, jac);
    for (int c = 0; c < 3; ++c) {
      if (dcol[c] == R(0)) continue;
      const R* r = rest + static_cast<size_t>(c) * per_channel;
      R* dst = out.rest + static_cast<size_t>(g) * rest_stride + static_cast<size_t>(c) * per_channel;
      for (int b = 0; b < usable; ++b) {
        dst[b] += basis[b] * dcol[c];
        const R k = r[b] * dcol[c];
        for (int a = 0; a < 3; ++a) ddir[a] += jac[b][a] * k;
      }
    }
  }
  const R dmx = acc.dmean[2 * s], dmy = acc.dmean[2 * s + 1];
  R pc[3];
  toCamera(v, pos, pc);
  const R z = pc[2];
  R dp[3] = {v.fx / z * dmx, v.fy / z * dmy, -(v.fx * pc[0] * dmx + v.fy * pc[1] * dmy) / (z * z)};
  // conic -> cov2d
  const R C[2][2] = {{q.conic[0], q.conic[1]}, {q.conic[1], q.conic[2]}};
  const R Gm[2][2] = {{acc.dconic[3 * s], R(0.5f) * acc.dconic[3 * s + 1]},
                      {R(0.5f) * acc.dconic[3 * s + 1], acc.dconic[3 * s + 2]}};
  R CG[2][2], D[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) CG[i][j] = C[i][0] * Gm[0][j] + C[i][1] * Gm[1][j];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) D[i][j] = -(CG[i][0] * C[0][j] + CG[i][1] * C[1][j]);
  // cov2d -> cov3d and J
  const R lx = v.thx * R(kSlack), ly = v.thy * R(kSlack);
  R jxJ, jyJ;
  if (std::is_same<R, float>::value) {
    jxJ = std::clamp(pc[0] / z, -lx, lx) * z;
    jyJ = std::clamp(pc[1] / z, -ly, ly) * z;
  } else {
    jxJ = pf.fclamp[0] ? std::copysign(lx, pc[0]) * z : pc[0];
    jyJ = pf.fclamp[1] ? std::copysign(ly, pc[1]) * z : pc[1];
  }
  R J[2][3], T[2][3];
  jacobianTimesW(v, z, jxJ, jyJ, J, T);
  R S[3][3], Rm[3][3], sc[3], M[3][3];
  covariance3d(rot, ls, S, Rm, sc, M);
  // dcov3d = T^T D T, evaluated as (T^T D) T
  R TtD[3][2], dS3[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 2; ++j) TtD[i][j] = T[0][i] * D[0][j] + T[1][i] * D[1][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) dS3[i][j] = TtD[i][0] * T[0][j] + TtD[i][1] * T[1][j];
  // dT = ((2 D) T) Sigma
  R D2[2][2], D2T[2][3], dT[2][3], dJ[2][3];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) D2[i][j] = R(2) * D[i][j];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 3; ++j) D2T[i][j] = D2[i][0] * T[0][j] + D2[i][1] * T[1][j];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 3; ++j) dT[i][j] = D2T[i][0] * S[0][j] + D2T[i][1] * S[1][j] + D2T[i][2] * S[2][j];
  // dJ = dT W^T
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 3; ++j) dJ[i][j] = dT[i][0] * v.W[j][0] + dT[i][1] * v.W[j][1] + dT[i][2] * v.W[j][2];
  const bool cx = pf.fclamp[0], cy = pf.fclamp[1];
  const R jx = cx ? std::copysign(lx, pc[0]) * z : pc[0];
  const R jy = cy ? std::copysign(ly, pc[1]) * z : pc[1];
  if (!cx) dp[0] += dJ[0][2] * (-v.fx / (z * z));
  if (!cy) dp[1] += dJ[1][2] * (-v.fy / (z * z));
  const R dz_x = (cx ? R(1) : R(2)) * v.fx * jx / (z * z * z);
  const R dz_y = (cy ? R(1) : R(2)) * v.fy * jy / (z * z * z);
  dp[2] += dJ[0][0] * (-v.fx / (z * z)) + dJ[1][1] * (-v.fy / (z * z)) + dJ[0][2] * dz_x + dJ[1][2] * dz_y;
  R dpos[3];
  for (int i = 0; i < 3; ++i) dpos[i] = v.W[0][i] * dp[0] + v.W[1][i] * dp[1] + v.W[2][i] * dp[2];
  const R ddsq = ddir[0] * ddir[0] + ddir[1] * ddir[1] + ddir[2] * ddir[2];
  if (distance > R(1e-9f) && ddsq > R(0)) {
    const R dir[3] = {offset[0] / distance, offset[1] / distance, offset[2] / distance};
    const R dot = dir[0] * ddir[0] + dir[1] * ddir[1] + dir[2] * ddir[2];
    for (int i = 0; i < 3; ++i) dpos[i] += (ddir[i] - dir[i] * dot) / distance;
  }
  for (int i = 0; i < 3; ++i) out.pos[3 * g + i] += dpos[i];
  // cov3d -> scale and rotation
  R sym[3][3], dM[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) sym[i][j] = R(0.5f) * (dS3[i][j] + dS3[j][i]);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) dM[i][j] = R(2) * sym[i][0] * M[0][j] + R(2) * sym[i][1] * M[1][j] + R(2) * sym[i][2] * M[2][j];
  for (int a = 0; a < 3; ++a) {
    const R dSaa = Rm[0][a] * dM[0][a] + Rm[1][a] * dM[1][a] + Rm[2][a] * dM[2][a];
    out.ls[3 * g + a] += dSaa * sc[a];
  }
  R dR[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) dR[r][c] = dM[r][c] * sc[c];
  const R w0 = rot[0], x0 = rot[1], y0 = rot[2], z0 = rot[3];
  const R sqn = x0 * x0 + y0 * y0 + z0 * z0 + w0 * w0;
  const R raw_norm = std::max(std::sqrt(sqn), R(1e-12f));
  R qw = w0, qx = x0, qy = y0, qz = z0;
  if (sqn > R(0)) {
    const R s0 = std::sqrt(sqn);
    qx = qx / s0; qy = qy / s0; qz = qz / s0; qw = qw / s0;
  }
  R dq[4];
  dq[0] = R(2) * (-qz * dR[0][1] + qy * dR[0][2] + qz * dR[1][0] - qx * dR[1][2] - qy * dR[2][0] + qx * dR[2][1]);
  dq[1] = R(2) * (qy * dR[0][1] + qz * dR[0][2] + qy * dR[1][0] - R(2) * qx * dR[1][1] - qw * dR[1][2] +
                  qz * dR[2][0] + qw * dR[2][1] - R(2) * qx * dR[2][2]);
  dq[2] = R(2) * (R(-2) * qy * dR[0][0] + qx * dR[0][1] + qw * dR[0][2] + qx * dR[1][0] + qz * dR[1][2] -
                  qw * dR[2][0] + qz * dR[2][1] - R(2) * qy * dR[2][2]);
  dq[3] = R(2) * (R(-2) * qz * dR[0][0] - qw * dR[0][1] + qx * dR[0][2] + qw * dR[1][0] -
                  R(2) * qz * dR[1][1] + qy * dR[1][2] + qx * dR[2][0] + qy * dR[2][1]);
  const R unit[4] = {qw, qx, qy, qz};
  const R ud = unit[0] * dq[0] + unit[1] * dq[1] + unit[2] * dq[2] + unit[3] * dq[3];
  for (int a = 0; a < 4; ++a) out.rot[4 * g + a] += (dq[a] - unit[a] * ud) / raw_norm;
}

std::vector<ProjR<float>> floatValues(const std::vector<ProjF>& proj) {
  std::vector<ProjR<float>> v(proj.size());
  for (size_t i = 0; i < proj.size(); ++i) {
    const ProjF& p = proj[i];
    v[i] = {{p.mean[0], p.mean[1]}, {p.conic[0], p.conic[1], p.conic[2]}, p.opacity,
            {p.color[0], p.color[1], p.color[2]}};
  }
  return v;
}

std::vector<ProjR<double>> frozenValues(const orc_frame& f, const orc_cloud_d& cl) {
  const ViewR<double> vd = makeView<double>(f.view);
  std::vector<ProjR<double>> v(f.proj.size());
  for (size_t s = 0; s < f.proj.size(); ++s)
    v[s] = projectFrozen(cl, f.proj[s].index, vd, f.opt, f.sh_use, f.proj[s]);
  return v;
}

// Composite one tile in FP64 with frozen decisions; returns sum(w * rgb).
double renderTileFrozen(const orc_frame& f, int tx, int ty, const std::vector<ProjR<double>>& pr,
                        const double* weights, double* rgb_out) {
  const Grid& g = f.grid;
  const int W = f.view.width, H = f.view.height;
  const int tile = ty * g.cols + tx;
  const int begin = g.offsets[tile], end = g.offsets[tile + 1];
  const int y_end = std::min(H, (ty + 1) * kTile), x_end = std::min(W, (tx + 1) * kTile);
  double total = 0.0;
  for (int y = ty * kTile; y < y_end; ++y) {
    for (int x = tx * kTile; x < x_end; ++x) {
      const size_t pixel = static_cast<size_t>(y) * W + x;
      const float pxf = static_cast<float>(x) + 0.5f, pyf = static_cast<float>(y) + 0.5f;
      const double px = pxf, py = pyf;
      float Tf = 1.f;
      double T = 1.0, acc[3] = {0, 0, 0};
      for (int e = begin; e < end; ++e) {
        const size_t s = static_cast<size_t>(g.entries[e]);
        float dxf, dyf, Gf, rawf, alphaf;
        if (!evalF(f.proj[s], pxf, pyf, f.opt.min_alpha, dxf, dyf, Gf, rawf, alphaf)) continue;
        const ProjR<double>& q = pr[s];
        const double dx = px - q.mean[0], dy = py - q.mean[1];
        const double power = -0.5 * (q.conic[0] * dx * dx + q.conic[2] * dy * dy) - q.conic[1] * dx * dy;
        const double a = rawf >= kMaxAlpha ? double(kMaxAlpha) : q.opacity * std::exp(power);
        for (int c = 0; c < 3; ++c) acc[c] += T * a * q.color[c];
        T *= 1.0 - a;
        Tf *= 1.f - alphaf;
        if (Tf < f.opt.transmittance_cutoff) break;
      }
      for (int c = 0; c < 3; ++c) {
        const double v = acc[c] + T * double(f.opt.background[c]);
        if (rgb_out) rgb_out[pixel * 3 + c] = v;
        if (weights) total += weights[pixel * 3 + c] * v;
      }
    }
  }
  return total;
}

}  // namespace

extern "C" {

void orc_default_options(orc_options* o) {
  o->background[0] = o->background[1] = o->background[2] = 0.f;
  o->sh_degree = 0;
  o->min_alpha = 1.f / 255.f;
  o->transmittance_cutoff = 1e-4f;
  o->screen_dilation = 0.3f;
}

void orc_default_config(orc_config* c) {
  c->iterations = 7000;
  c->position_lr_initial = 1.6e-4f;
  c->position_lr_final = 1.6e-6f;
  c->dc_lr = 2.5e-3f;
  c->sh_rest_lr_divider = 20.f;
  c->opacity_lr = 0.05f;
  c->scale_lr = 5e-3f;
  c->rotation_lr = 1e-3f;
  c->adam_beta1 = 0.9f;
  c->adam_beta2 = 0.999f;
  c->adam_epsilon = 1e-15f;
  c->ssim_weight = 0.2f;
  c->sh_degree = 3;
  c->sh_increase_interval = 1000;
  c->densify_from_iteration = 500;
  c->densify_until_iteration = 5000;
  c->densify_interval = 100;
  c->densify_grad_threshold = 2e-4f;
  c->percent_dense = 0.01f;
  c->split_scale_divider = 1.6f;
  c->opacity_reset_interval = 3000;
  c->opacity_reset_value = 0.01f;
  c->prune_opacity = 0.005f;
  c->prune_scale_fraction = 0.1f;
  c->max_gaussians = 600000;
  c->seed = 7u;
}

float orc_det_exp(float x) { return splat_det::exp(x); }

orc_frame* orc_frame_create(const orc_cloud* cloud, const orc_view* view, const orc_options* opt) {
  orc_frame* f = new orc_frame;
  f->view = *view;
  f->opt = *opt;
  f->n = cloud->n;
  f->rest_stride = cloud->rest_stride;
  f->sh_use = std::min(opt->sh_degree, cloud->sh_degree);
  f->vf = makeView<float>(*view);
  const CloudView cl = cv(cloud);
  f->proj.reserve(static_cast<size_t>(cloud->n));
  for (int i = 0; i < cloud->n; ++i) {
    ProjF p;
    if (projectOne(cl, i, f->vf, *opt, f->sh_use, &p)) f->proj.push_back(p);
  }
  // SplatRaster.cpp:260-261 sorts by depth with an unstable std::sort; the
  // oracle fixes exact ties by gaussian index (stable sort of index order).
  std::stable_sort(f->proj.begin(), f->proj.end(),
                   [](const ProjF& a, const ProjF& b) { return a.depth < b.depth; });
  f->slot_of.assign(static_cast<size_t>(cloud->n), -1);
  for (size_t s = 0; s < f->proj.size(); ++s) f->slot_of[static_cast<size_t>(f->proj[s].index)] = static_cast<int>(s);
  f->grid = buildTiles(f->proj, view->width, view->height);
  rasterize(*f);
  return f;
}

void orc_frame_destroy(orc_frame* f) { delete f; }

int orc_frame_projected_count(const orc_frame* f) { return static_cast<int>(f->proj.size()); }

void orc_frame_get_projected(const orc_frame* f, int* index, float* mean2, float* cov2d3, float* conic3,
                             float* depth, float* radius, float* color3, float* opacity, int* clamped3,
                             int* rect4) {
  for (size_t s = 0; s < f->proj.size(); ++s) {
    const ProjF& p = f->proj[s];
    if (index) index[s] = p.index;
    for (int k = 0; k < 2; ++k) if (mean2) mean2[2 * s + k] = p.mean[k];
    for (int k = 0; k < 3; ++k) {
      if (cov2d3) cov2d3[3 * s + k] = p.cov2d[k];
      if (conic3) conic3[3 * s + k] = p.conic[k];
      if (color3) color3[3 * s + k] = p.color[k];
      if (clamped3) clamped3[3 * s + k] = p.clamped[k] ? 1 : 0;
    }
    if (depth) depth[s] = p.depth;
    if (radius) radius[s] = p.radius;
    if (opacity) opacity[s] = p.opacity;
    for (int k = 0; k < 4; ++k) if (rect4) rect4[4 * s + k] = p.rect[k];
  }
}

int orc_frame_tile_grid(const orc_frame* f, int* cols, int* rows) {
  if (cols) *cols = f->grid.cols;
  if (rows) *rows = f->grid.rows;
  return static_cast<int>(f->grid.entries.size());
}

void orc_frame_get_tiles(const orc_frame* f, int* offsets, int* entries_gaussian) {
  if (offsets) std::copy(f->grid.offsets.begin(), f->grid.offsets.end(), offsets);
  if (entries_gaussian)
    for (size_t e = 0; e < f->grid.entries.size(); ++e)
      entries_gaussian[e] = f->proj[static_cast<size_t>(f->grid.entries[e])].index;
}

void orc_frame_get_image(const orc_frame* f, float* rgb, float* T, int* last, int64_t* evals) {
  if (rgb) std::copy(f->rgb.begin(), f->rgb.end(), rgb);
  if (T) std::copy(f->T.begin(), f->T.end(), T);
  if (last) std::copy(f->last.begin(), f->last.end(), last);
  if (evals) { evals[0] = f->evals[0]; evals[1] = f->evals[1]; }
}

void orc_frame_backward(orc_frame* f, const orc_cloud* cloud, const float* dL_drgb, orc_grads* g) {
  if (f->proj.empty() || f->view.width == 0 || f->view.height == 0) return;
  const std::vector<ProjR<float>> pr = floatValues(f->proj);
  ScreenAcc<float> acc;
  acc.init(f->proj.size());
  f->evals[1] = 0;
  for (int ty = 0; ty < f->grid.rows; ++ty)
    for (int tx = 0; tx < f->grid.cols; ++tx) backwardTile<float>(*f, tx, ty, pr, dL_drgb, acc, &f->evals[1]);
  GradOut<float> out{g->pos, g->log_scale, g->rot, g->opacity, g->dc, g->rest, g->screen_gradient, g->visible_count};
  const int rs = cloud->rest_stride;
  for (size_t s = 0; s < f->proj.size(); ++s) {
    const int i = f->proj[s].index;
    perGaussian<float>(f->vf, f->sh_use, rs, cloud->pos + 3 * i, cloud->log_scale + 3 * i, cloud->rot + 4 * i,
                       cloud->rest + static_cast<size_t>(rs) * i, f->proj[s], pr[s], s, acc, out, i);
  }
}

void orc_frame_backward_f64(orc_frame* f, const orc_cloud_d* cloud, const float* dL_drgb, orc_grads_d* g) {
  if (f->proj.empty() || f->view.width == 0 || f->view.height == 0) return;
  const std::vector<ProjR<double>> pr = frozenValues(*f, *cloud);
  ScreenAcc<double> acc;
  acc.init(f->proj.size());
  for (int ty = 0; ty < f->grid.rows; ++ty)
    for (int tx = 0; tx < f->grid.cols; ++tx) backwardTile<double>(*f, tx, ty, pr, dL_drgb, acc, nullptr);
  const ViewR<double> vd = makeView<double>(f->view);
  GradOut<double> out{g->pos, g->log_scale, g->rot, g->opacity, g->dc, g->rest, g->screen_gradient, g->visible_count};
  const int rs = cloud->rest_stride;
  for (size_t s = 0; s < f->proj.size(); ++s) {
    const int i = f->proj[s].index;
    perGaussian<double>(vd, f->sh_use, rs, cloud->pos + 3 * i, cloud->log_scale + 3 * i, cloud->rot + 4 * i,
                        cloud->rest + static_cast<size_t>(rs) * i, f->proj[s], pr[s], s, acc, out, i);
  }
}

void orc_frame_render_f64(orc_frame* f, const orc_cloud_d* cloud, double* rgb) {
  const size_t px = static_cast<size_t>(f->view.width) * f->view.height;
  std::vector<ProjR<double>> pr = frozenValues(*f, *cloud);
  for (size_t p = 0; p < px; ++p)
    for (int c = 0; c < 3; ++c) rgb[p * 3 + c] = f->opt.background[c];
  for (int ty = 0; ty < f->grid.rows; ++ty)
    for (int tx = 0; tx < f->grid.cols; ++tx) renderTileFrozen(*f, tx, ty, pr, nullptr, rgb);
}

double orc_frame_fd_f64(orc_frame* f, const orc_cloud_d* cloud, const double* weights, int gaussian, int group,
                        int comp, double h) {
  if (gaussian < 0 || gaussian >= f->n) return std::nan("");
  const int s = f->slot_of[static_cast<size_t>(gaussian)];
  if (s < 0) return std::nan("");
  const int rs = cloud->rest_stride;
  // Private copy of this gaussian's parameters, perturbed in place.
  std::vector<double> pos(cloud->pos + 3 * gaussian, cloud->pos + 3 * gaussian + 3);
  std::vector<double> ls(cloud->log_scale + 3 * gaussian, cloud->log_scale + 3 * gaussian + 3);
  std::vector<double> rot(cloud->rot + 4 * gaussian, cloud->rot + 4 * gaussian + 4);
  double op = cloud->opacity[gaussian];
  std::vector<double> dc(cloud->dc + 3 * gaussian, cloud->dc + 3 * gaussian + 3);
  std::vector<double> rest(cloud->rest + static_cast<size_t>(rs) * gaussian,
                           cloud->rest + static_cast<size_t>(rs) * gaussian + rs);
  double* target = nullptr;
  switch (group) {
    case 0: target = &pos[static_cast<size_t>(comp)]; break;
    case 1: target = &ls[static_cast<size_t>(comp)]; break;
    case 2: target = &rot[static_cast<size_t>(comp)]; break;
    case 3: target = &op; break;
    case 4: target = &dc[static_cast<size_t>(comp)]; break;
    default: target = &rest[static_cast<size_t>(comp)]; break;
  }
  const double original = *target;
  if (f->pr_key != static_cast<const void*>(cloud->pos) || f->pr_cache.size() != f->proj.size()) {
    f->pr_cache = frozenValues(*f, *cloud);
    f->pr_key = cloud->pos;
  }
  std::vector<ProjR<double>>& pr = f->pr_cache;
  const ProjR<double> saved = pr[static_cast<size_t>(s)];
  const ViewR<double> vd = makeView<double>(f->view);
  orc_cloud_d one{1, cloud->sh_degree, rs, pos.data(), ls.data(), rot.data(), &op, dc.data(), rest.data()};
  const ProjF& pf = f->proj[static_cast<size_t>(s)];
  double value[2];
  for (int side = 0; side < 2; ++side) {
    *target = original + (side == 0 ? h : -h);
    pr[static_cast<size_t>(s)] = projectFrozen(one, 0, vd, f->opt, f->sh_use, pf);
    double total = 0.0;
    for (int ty = pf.rect[2]; ty <= pf.rect[3]; ++ty)
      for (int tx = pf.rect[0]; tx <= pf.rect[1]; ++tx) total += renderTileFrozen(*f, tx, ty, pr, weights, nullptr);
    value[side] = total;
  }
  *target = original;
  pr[static_cast<size_t>(s)] = saved;
  return (value[0] - value[1]) / (2.0 * h);
}

// ------------------------------------------------------------------ loss --
namespace {
constexpr int kSsimRadius = 5;
constexpr float kSsimSigma = 1.5f;
constexpr float kSsimC1 = 0.01f * 0.01f;
constexpr float kSsimC2 = 0.03f * 0.03f;

// ssimKernel, SplatRaster.cpp:587-600 (libm expf, as the reference).
const std::vector<float>& ssimKernel() {
  static const std::vector<float> kernel = [] {
    std::vector<float> values(2 * kSsimRadius + 1);
    float total = 0.f;
    for (int i = -kSsimRadius; i <= kSsimRadius; ++i) {
      const float v = std::exp(-0.5f * static_cast<float>(i * i) / (kSsimSigma * kSsimSigma));
      values[static_cast<size_t>(i + kSsimRadius)] = v;
      total += v;
    }
    for (float& v : values) v /= total;
    return values;
  }();
  return kernel;
}
}  // namespace

// blur, SplatRaster.cpp:614-642: separable, zero padded. Each tap is one
// fused multiply-add, in tap order -5..5 (a fixed rounding both this checker
// and the device kernels follow; the engine's own build leaves the
// association to the compiler).
void orc_blur(const float* src, int width, int height, float* dst) {
  const std::vector<float>& k = ssimKernel();
  const size_t n = static_cast<size_t>(width) * height;
  std::vector<float> horizontal(n, 0.f);
  for (int y = 0; y < height; ++y)
    for (int x = 0; x < width; ++x) {
      float sum = 0.f;
      for (int i = -kSsimRadius; i <= kSsimRadius; ++i) {
        const int sx = x + i;
        if (sx < 0 || sx >= width) continue;
        sum = std::fma(k[static_cast<size_t>(i + kSsimRadius)], src[static_cast<size_t>(y) * width + sx], sum);
      }
      horizontal[static_cast<size_t>(y) * width + x] = sum;
    }
  for (int y = 0; y < height; ++y)
    for (int x = 0; x < width; ++x) {
      float sum = 0.f;
      for (int i = -kSsimRadius; i <= kSsimRadius; ++i) {
        const int sy = y + i;
        if (sy < 0 || sy >= height) continue;
        sum = std::fma(k[static_cast<size_t>(i + kSsimRadius)], horizontal[static_cast<size_t>(sy) * width + x], sum);
      }
      dst[static_cast<size_t>(y) * width + x] = sum;
    }
}

// meanSsim, SplatRaster.cpp:646-710.
float orc_mean_ssim(const float* rendered, const float* target, int width, int height, float* d_rendered) {
  const size_t pixels = static_cast<size_t>(width) * height;
  if (pixels == 0) return 0.f;
  if (d_rendered) std::fill(d_rendered, d_rendered + pixels * 3, 0.f);
  const float scale = 1.f / static_cast<float>(pixels * 3);
  float total = 0.f;
  std::vector<float> x(pixels), y(pixels), xx(pixels), yy(pixels), xy(pixels);
  std::vector<float> mu_x(pixels), mu_y(pixels), exx(pixels), eyy(pixels), exy(pixels);
  std::vector<float> term_mu(pixels), term_var(pixels), term_cov(pixels), blurred(pixels), scratch(pixels);
  for (int channel = 0; channel < 3; ++channel) {
    for (size_t p = 0; p < pixels; ++p) {
      x[p] = rendered[p * 3 + channel];
      y[p] = target[p * 3 + channel];
      xx[p] = x[p] * x[p];
      yy[p] = y[p] * y[p];
      xy[p] = x[p] * y[p];
    }
    orc_blur(x.data(), width, height, mu_x.data());
    orc_blur(y.data(), width, height, mu_y.data());
    orc_blur(xx.data(), width, height, exx.data());
    orc_blur(yy.data(), width, height, eyy.data());
    orc_blur(xy.data(), width, height, exy.data());
    for (size_t p = 0; p < pixels; ++p) {
      const float var_x = exx[p] - mu_x[p] * mu_x[p];
      const float var_y = eyy[p] - mu_y[p] * mu_y[p];
      const float cov = exy[p] - mu_x[p] * mu_y[p];
      const float a = 2.f * mu_x[p] * mu_y[p] + kSsimC1;
      const float b = 2.f * cov + kSsimC2;
      const float c = mu_x[p] * mu_x[p] + mu_y[p] * mu_y[p] + kSsimC1;
      const float d = var_x + var_y + kSsimC2;
      total += (a * b) / (c * d);
      if (!d_rendered) continue;
      term_mu[p] = scale * (2.f * mu_y[p] * b / (c * d) - 2.f * mu_x[p] * a * b / (c * c * d));
      term_var[p] = scale * (-(a * b) / (c * d * d));
      term_cov[p] = scale * (2.f * a / (c * d));
    }
    if (!d_rendered) continue;
    orc_blur(term_mu.data(), width, height, blurred.data());
    for (size_t p = 0; p < pixels; ++p) d_rendered[p * 3 + channel] += blurred[p];
    orc_blur(term_var.data(), width, height, blurred.data());
    for (size_t p = 0; p < pixels; ++p) d_rendered[p * 3 + channel] += 2.f * x[p] * blurred[p];
    for (size_t p = 0; p < pixels; ++p) scratch[p] = term_var[p] * mu_x[p];
    orc_blur(scratch.data(), width, height, blurred.data());
    for (size_t p = 0; p < pixels; ++p) d_rendered[p * 3 + channel] -= 2.f * blurred[p];
    orc_blur(term_cov.data(), width, height, blurred.data());
    for (size_t p = 0; p < pixels; ++p) d_rendered[p * 3 + channel] += y[p] * blurred[p];
    for (size_t p = 0; p < pixels; ++p) scratch[p] = term_cov[p] * mu_y[p];
    orc_blur(scratch.data(), width, height, blurred.data());
    for (size_t p = 0; p < pixels; ++p) d_rendered[p * 3 + channel] -= blurred[p];
  }
  return total * scale;
}

// photometricLoss, SplatRaster.cpp:712-738.
float orc_photometric_loss(const float* rendered, const float* target, int width, int height, float ssim_weight,
                           float* dL_drgb) {
  const size_t values = static_cast<size_t>(width) * height * 3;
  if (dL_drgb) std::fill(dL_drgb, dL_drgb + values, 0.f);
  if (values == 0) return 0.f;
  const float inverse = 1.f / static_cast<float>(values);
  float l1 = 0.f;
  for (size_t i = 0; i < values; ++i) {
    const float difference = rendered[i] - target[i];
    l1 += std::abs(difference);
    if (dL_drgb) {
      const float sign = difference > 0.f ? 1.f : (difference < 0.f ? -1.f : 0.f);
      dL_drgb[i] = (1.f - ssim_weight) * sign * inverse;
    }
  }
  l1 *= inverse;
  if (ssim_weight <= 0.f) return l1;
  std::vector<float> d_ssim(dL_drgb ? values : 0);
  const float ssim = orc_mean_ssim(rendered, target, width, height, dL_drgb ? d_ssim.data() : nullptr);
  if (dL_drgb)
    for (size_t i = 0; i < values; ++i) dL_drgb[i] -= ssim_weight * d_ssim[i];
  return (1.f - ssim_weight) * l1 + ssim_weight * (1.f - ssim);
}

void orc_evaluate_sh(const float* dc3, const float* rest, int rest_stride, int degree, const float* direction3,
                     float* color3, int* clamped3) {
  bool cl[3];
  evalSh(dc3, rest, rest_stride, degree, direction3, color3, cl);
  if (clamped3)
    for (int c = 0; c < 3; ++c) clamped3[c] = cl[c] ? 1 : 0;
}

// -------------------------------------------------------------- training --
float orc_camera_extent(int n_views, const orc_view* views) {
  // cameraExtent, SplatTrainer.cpp:74-85
  if (n_views <= 0) return 1.f;
  auto center = [](const orc_view& v, float c[3]) {
    for (int i = 0; i < 3; ++i) {
      const double d = (-v.R[0 * 3 + i]) * v.t[0] + (-v.R[1 * 3 + i]) * v.t[1] + (-v.R[2 * 3 + i]) * v.t[2];
      c[i] = static_cast<float>(d);
    }
  };
  float centroid[3] = {0.f, 0.f, 0.f};
  for (int k = 0; k < n_views; ++k) {
    float c[3];
    center(views[k], c);
    for (int i = 0; i < 3; ++i) centroid[i] += c[i];
  }
  for (int i = 0; i < 3; ++i) centroid[i] /= static_cast<float>(n_views);
  float furthest = 0.f;
  for (int k = 0; k < n_views; ++k) {
    float c[3];
    center(views[k], c);
    const float d[3] = {c[0] - centroid[0], c[1] - centroid[1], c[2] - centroid[2]};
    furthest = std::max(furthest, std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
  }
  return std::max(furthest * 1.1f, 1e-3f);
}

namespace {

struct HostCloud {
  int n = 0, sh_degree = 0, rs = 0;
  std::vector<float> pos, ls, rot, op, dc, rest;
  void resize(int count) {
    n = count;
    pos.resize(3 * static_cast<size_t>(count));
    ls.resize(3 * static_cast<size_t>(count));
    rot.resize(4 * static_cast<size_t>(count));
    op.resize(static_cast<size_t>(count));
    dc.resize(3 * static_cast<size_t>(count));
    rest.resize(static_cast<size_t>(rs) * count);
  }
  orc_cloud view() {
    return {n, sh_degree, rs, pos.data(), ls.data(), rot.data(), op.data(), dc.data(), rest.data()};
  }
  void load(const orc_cloud& c) {
    sh_degree = c.sh_degree;
    rs = c.rest_stride;
    resize(c.n);
    std::copy(c.pos, c.pos + 3 * c.n, pos.begin());
    std::copy(c.log_scale, c.log_scale + 3 * c.n, ls.begin());
    std::copy(c.rot, c.rot + 4 * c.n, rot.begin());
    std::copy(c.opacity, c.opacity + c.n, op.begin());
    std::copy(c.dc, c.dc + 3 * c.n, dc.begin());
    std::copy(c.rest, c.rest + static_cast<size_t>(rs) * c.n, rest.begin());
  }
};

struct Grads {
  std::vector<float> pos, ls, rot, op, dc, rest, sg;
  std::vector<int> vc;
  void resize(int n, int rs) {
    pos.assign(3 * static_cast<size_t>(n), 0.f);
    ls.assign(3 * static_cast<size_t>(n), 0.f);
    rot.assign(4 * static_cast<size_t>(n), 0.f);
    op.assign(static_cast<size_t>(n), 0.f);
    dc.assign(3 * static_cast<size_t>(n), 0.f);
    rest.assign(static_cast<size_t>(rs) * n, 0.f);
    sg.assign(static_cast<size_t>(n), 0.f);
    vc.assign(static_cast<size_t>(n), 0);
  }
  orc_grads view() {
    return {pos.data(), ls.data(), rot.data(), op.data(), dc.data(), rest.data(), sg.data(), vc.data()};
  }
};

// adamDelta, SplatTrainer.cpp:56-65.
inline float adamDelta(float& m, float& v, float g, float lr, float b1, float b2, float eps, float c1, float c2) {
  m = b1 * m + (1.f - b1) * g;
  v = b2 * v + (1.f - b2) * g * g;
  const float m_hat = m / c1;
  const float v_hat = v / c2;
  return -lr * m_hat / (std::sqrt(v_hat) + eps);
}

// Learning rates of one step (SplatTrainer.cpp:200-208, 219-225).
struct StepRates {
  float position, sh_rest, c1, c2;
};

StepRates stepRates(const orc_config& c, int iteration, float extent) {
  StepRates r;
  const float t = c.iterations <= 1 ? 1.f : static_cast<float>(iteration) / static_cast<float>(c.iterations - 1);
  const float ratio = c.position_lr_final / std::max(c.position_lr_initial, 1e-30f);
  r.position = c.position_lr_initial * std::pow(ratio, std::clamp(t, 0.f, 1.f)) * extent;
  r.sh_rest = c.dc_lr / std::max(1.f, c.sh_rest_lr_divider);
  const int step = iteration + 1;
  r.c1 = 1.f - std::pow(c.adam_beta1, static_cast<float>(step));
  r.c2 = 1.f - std::pow(c.adam_beta2, static_cast<float>(step));
  return r;
}

// applyGradients, SplatTrainer.cpp:218-256. Moments are group-concatenated.
void applyAdam(orc_cloud& cl, const orc_grads& g, float* m, float* v, int iteration, float extent,
               const orc_config& c) {
  const StepRates r = stepRates(c, iteration, extent);
  const size_t n = static_cast<size_t>(cl.n), rs = static_cast<size_t>(cl.rest_stride);
  const size_t o_pos = 0, o_ls = 3 * n, o_rot = 6 * n, o_op = 10 * n, o_dc = 11 * n, o_rest = 14 * n;
  const float b1 = c.adam_beta1, b2 = c.adam_beta2, eps = c.adam_epsilon;
  for (size_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      const size_t k = 3 * i + static_cast<size_t>(a);
      cl.pos[k] += adamDelta(m[o_pos + k], v[o_pos + k], g.pos[k], r.position, b1, b2, eps, r.c1, r.c2);
      cl.log_scale[k] += adamDelta(m[o_ls + k], v[o_ls + k], g.log_scale[k], c.scale_lr, b1, b2, eps, r.c1, r.c2);
      cl.dc[k] += adamDelta(m[o_dc + k], v[o_dc + k], g.dc[k], c.dc_lr, b1, b2, eps, r.c1, r.c2);
    }
    cl.opacity[i] += adamDelta(m[o_op + i], v[o_op + i], g.opacity[i], c.opacity_lr, b1, b2, eps, r.c1, r.c2);
    float* q = cl.rot + 4 * i;
    for (int a = 0; a < 4; ++a) {
      const size_t k = 4 * i + static_cast<size_t>(a);
      q[a] += adamDelta(m[o_rot + k], v[o_rot + k], g.rot[k], c.rotation_lr, b1, b2, eps, r.c1, r.c2);
    }
    // Eigen norm()/normalize() over the (x, y, z, w) coefficients.
    const float sq = q[1] * q[1] + q[2] * q[2] + q[3] * q[3] + q[0] * q[0];
    if (std::sqrt(sq) > 1e-6f) {
      const float s = std::sqrt(sq);
      for (int a = 0; a < 4; ++a) q[a] = q[a] / s;
    } else {
      q[0] = 1.f; q[1] = q[2] = q[3] = 0.f;
    }
    for (size_t k = 0; k < rs; ++k) {
      const size_t idx = i * rs + k;
      cl.rest[idx] += adamDelta(m[o_rest + idx], v[o_rest + idx], g.rest[idx], r.sh_rest, b1, b2, eps, r.c1, r.c2);
    }
  }
}

// nextNormal, SplatDensify.cpp:11-24.
float nextNormal(uint32_t* state) {
  const auto next = [&]() {
    uint32_t x = *state;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    *state = x == 0u ? 1u : x;
    return *state;
  };
  const float u1 = std::max(1e-7f, static_cast<float>(next()) / 4294967296.f);
  const float u2 = static_cast<float>(next()) / 4294967296.f;
  return std::sqrt(-2.f * std::log(u1)) * std::cos(2.f * 3.14159265358979f * u2);
}

float maxScale(const float* ls) {
  return std::max(std::max(dexp(ls[0]), dexp(ls[1])), dexp(ls[2]));
}

}  // namespace

// densifyAndPrune, SplatDensify.cpp:28-109.
int orc_densify(orc_cloud* cloud, int capacity, const float* screen_gradient, const int* visible_count,
                const orc_config* c, float extent, int iteration, uint32_t* rng_state, int* source_index,
                int* counts) {
  const int before = cloud->n;
  int cloned = 0, split = 0, pruned = 0;
  if (counts) counts[0] = counts[1] = counts[2] = 0;
  if (before == 0) return 0;
  const int rs = cloud->rest_stride;
  const float size_threshold = c->percent_dense * extent;
  std::vector<char> keep(static_cast<size_t>(before), 1);
  HostCloud add;
  add.rs = rs;
  add.sh_degree = cloud->sh_degree;
  auto push = [&](int g) {
    const int k = add.n;
    add.resize(k + 1);
    std::copy(cloud->pos + 3 * g, cloud->pos + 3 * g + 3, add.pos.begin() + 3 * k);
    std::copy(cloud->log_scale + 3 * g, cloud->log_scale + 3 * g + 3, add.ls.begin() + 3 * k);
    std::copy(cloud->rot + 4 * g, cloud->rot + 4 * g + 4, add.rot.begin() + 4 * k);
    add.op[static_cast<size_t>(k)] = cloud->opacity[g];
    std::copy(cloud->dc + 3 * g, cloud->dc + 3 * g + 3, add.dc.begin() + 3 * k);
    std::copy(cloud->rest + static_cast<size_t>(rs) * g, cloud->rest + static_cast<size_t>(rs) * (g + 1),
              add.rest.begin() + static_cast<long>(rs) * k);
    return k;
  };
  const bool has_room = before < c->max_gaussians;
  if (has_room) {
    const float log_div = std::log(c->split_scale_divider);
    for (int g = 0; g < before; ++g) {
      const int seen = visible_count[g];
      if (seen <= 0) continue;
      const float mean_gradient = screen_gradient[g] / static_cast<float>(seen);
      if (mean_gradient < c->densify_grad_threshold) continue;
      const float* ls = cloud->log_scale + 3 * g;
      if (maxScale(ls) <= size_threshold) {
        push(g);
        ++cloned;
      } else {
        float scale[3] = {dexp(ls[0]), dexp(ls[1]), dexp(ls[2])};
        float Rm[3][3];
        rotationOf(cloud->rot + 4 * g, Rm);
        for (int child = 0; child < 2; ++child) {
          const int k = push(g);
          // The reference draws the three normals as the arguments of one
          // Vec3f constructor (SplatDensify.cpp:63-64), whose evaluation order
          // C++ leaves unspecified; its GCC build evaluates them last to first,
          // which the golden vectors pin.
          float sample[3];
          sample[2] = nextNormal(rng_state) * scale[2];
          sample[1] = nextNormal(rng_state) * scale[1];
          sample[0] = nextNormal(rng_state) * scale[0];
          for (int i = 0; i < 3; ++i)
            add.pos[3 * static_cast<size_t>(k) + i] =
                cloud->pos[3 * g + i] + (Rm[i][0] * sample[0] + Rm[i][1] * sample[1] + Rm[i][2] * sample[2]);
          for (int i = 0; i < 3; ++i) add.ls[3 * static_cast<size_t>(k) + i] = ls[i] - log_div;
        }
        keep[static_cast<size_t>(g)] = 0;
        ++split;
      }
    }
  }
  const bool prune_oversized = iteration > c->opacity_reset_interval;
  const float max_scale = c->prune_scale_fraction * extent;
  auto prunable = [&](float logit, const float* ls) {
    if (sigm(logit) < c->prune_opacity) return true;
    return prune_oversized && maxScale(ls) > max_scale;
  };
  HostCloud next;
  next.rs = rs;
  next.sh_degree = cloud->sh_degree;
  std::vector<int> src;
  auto copyInto = [&](const float* p, const float* l, const float* r, float o, const float* d, const float* re) {
    const int k = next.n;
    next.resize(k + 1);
    std::copy(p, p + 3, next.pos.begin() + 3 * k);
    std::copy(l, l + 3, next.ls.begin() + 3 * k);
    std::copy(r, r + 4, next.rot.begin() + 4 * k);
    next.op[static_cast<size_t>(k)] = o;
    std::copy(d, d + 3, next.dc.begin() + 3 * k);
    std::copy(re, re + rs, next.rest.begin() + static_cast<long>(rs) * k);
  };
  next.pos.reserve(3 * static_cast<size_t>(before + add.n));
  for (int g = 0; g < before; ++g) {
    if (!keep[static_cast<size_t>(g)]) continue;
    if (prunable(cloud->opacity[g], cloud->log_scale + 3 * g)) {
      ++pruned;
      continue;
    }
    copyInto(cloud->pos + 3 * g, cloud->log_scale + 3 * g, cloud->rot + 4 * g, cloud->opacity[g], cloud->dc + 3 * g,
             cloud->rest + static_cast<size_t>(rs) * g);
    src.push_back(g);
  }
  for (int k = 0; k < add.n; ++k) {
    if (next.n >= c->max_gaussians) break;
    if (prunable(add.op[static_cast<size_t>(k)], add.ls.data() + 3 * k)) {
      ++pruned;
      continue;
    }
    copyInto(add.pos.data() + 3 * k, add.ls.data() + 3 * k, add.rot.data() + 4 * k, add.op[static_cast<size_t>(k)],
             add.dc.data() + 3 * k, add.rest.data() + static_cast<size_t>(rs) * k);
    src.push_back(-1);
  }
  if (next.n > capacity) return -1;
  cloud->n = next.n;
  std::copy(next.pos.begin(), next.pos.end(), cloud->pos);
  std::copy(next.ls.begin(), next.ls.end(), cloud->log_scale);
  std::copy(next.rot.begin(), next.rot.end(), cloud->rot);
  std::copy(next.op.begin(), next.op.end(), cloud->opacity);
  std::copy(next.dc.begin(), next.dc.end(), cloud->dc);
  std::copy(next.rest.begin(), next.rest.end(), cloud->rest);
  if (source_index) std::copy(src.begin(), src.end(), source_index);
  if (counts) { counts[0] = cloned; counts[1] = split; counts[2] = pruned; }
  return next.n;
}

void orc_adam_apply(orc_cloud* cloud, const orc_grads* g, float* m, float* v, int iteration, float extent,
                    const orc_config* c) {
  applyAdam(*cloud, *g, m, v, iteration, extent, *c);
}

// GaussianCloud::prune (default max_scale 1e3) then sortByOpacity,
// GaussianCloud.cpp:177-197.
int orc_prune_and_sort(orc_cloud* cloud, float min_opacity, int* order_out) {
  HostCloud h;
  h.load(*cloud);
  std::vector<int> kept;
  for (int g = 0; g < h.n; ++g) {
    if (sigm(h.op[static_cast<size_t>(g)]) < min_opacity) continue;
    const float* ls = h.ls.data() + 3 * g;
    const float largest = dexp(std::max(std::max(ls[0], ls[1]), ls[2]));
    if (!std::isfinite(largest) || largest > 1e3f) continue;
    kept.push_back(g);
  }
  std::stable_sort(kept.begin(), kept.end(),
                   [&](int a, int b) { return h.op[static_cast<size_t>(a)] > h.op[static_cast<size_t>(b)]; });
  const int rs = h.rs;
  for (size_t k = 0; k < kept.size(); ++k) {
    const int g = kept[k];
    std::copy(h.pos.begin() + 3 * g, h.pos.begin() + 3 * g + 3, cloud->pos + 3 * k);
    std::copy(h.ls.begin() + 3 * g, h.ls.begin() + 3 * g + 3, cloud->log_scale + 3 * k);
    std::copy(h.rot.begin() + 4 * g, h.rot.begin() + 4 * g + 4, cloud->rot + 4 * k);
    cloud->opacity[k] = h.op[static_cast<size_t>(g)];
    std::copy(h.dc.begin() + 3 * g, h.dc.begin() + 3 * g + 3, cloud->dc + 3 * k);
    std::copy(h.rest.begin() + static_cast<long>(rs) * g, h.rest.begin() + static_cast<long>(rs) * (g + 1),
              cloud->rest + static_cast<size_t>(rs) * k);
    if (order_out) order_out[k] = g;
  }
  cloud->n = static_cast<int>(kept.size());
  return cloud->n;
}

}  // extern "C"

// ---------------------------------------------------------------- trainer --
struct orc_trainer {
  HostCloud cloud;
  std::vector<orc_view> views;
  std::vector<std::vector<float>> images;
  orc_config config;
  std::mt19937 engine;
  int iteration = 0;
  float extent = 1.f;
  float last_loss = 0.f, smoothed_loss = 0.f;
  std::vector<float> m, v;  // group-concatenated moments
  Grads grads;
  std::vector<float> acc_gradient;
  std::vector<int> acc_views;
  uint32_t densify_rng = 1u;
  std::vector<int> order;
  size_t cursor = 0;
  std::vector<int> last_views;

  size_t paramCount() const { return static_cast<size_t>(cloud.n) * (14 + static_cast<size_t>(cloud.rs)); }
  int shDegreeInUse() const {
    const int unlocked = config.sh_increase_interval > 0 ? iteration / config.sh_increase_interval : config.sh_degree;
    return std::min(config.sh_degree, std::max(0, unlocked));
  }
  int nextView() {
    if (cursor >= order.size()) {
      std::shuffle(order.begin(), order.end(), engine);
      cursor = 0;
    }
    return order[cursor++];
  }
  // remapState, SplatTrainer.cpp:267-288, on group-concatenated moments.
  void remap(const std::vector<int>& source, int old_n) {
    const size_t n0 = static_cast<size_t>(old_n), n1 = source.size(), rs = static_cast<size_t>(cloud.rs);
    const size_t strides[6] = {3, 3, 4, 1, 3, rs};
    std::vector<float> m1(n1 * (14 + rs), 0.f), v1(n1 * (14 + rs), 0.f);
    size_t off0 = 0, off1 = 0;
    for (size_t gidx = 0; gidx < 6; ++gidx) {
      const size_t st = strides[gidx];
      for (size_t i = 0; i < n1; ++i) {
        if (source[i] < 0) continue;
        const size_t from = static_cast<size_t>(source[i]) * st;
        for (size_t k = 0; k < st; ++k) {
          m1[off1 + i * st + k] = m[off0 + from + k];
          v1[off1 + i * st + k] = v[off0 + from + k];
        }
      }
      off0 += n0 * st;
      off1 += n1 * st;
    }
    m.swap(m1);
    v.swap(v1);
  }
};

extern "C" {

orc_trainer* orc_trainer_create(const orc_cloud* initial, int n_views, const orc_view* views,
                                const float* const* images, const orc_config* config) {
  orc_trainer* t = new orc_trainer;
  t->cloud.load(*initial);
  t->config = *config;
  t->engine.seed(config->seed == 0u ? 1u : config->seed);
  t->densify_rng = config->seed == 0u ? 1u : config->seed;
  for (int k = 0; k < n_views; ++k) {
    t->views.push_back(views[k]);
    const size_t values = static_cast<size_t>(views[k].width) * views[k].height * 3;
    t->images.emplace_back(images[k], images[k] + values);
  }
  t->extent = orc_camera_extent(n_views, views);
  t->m.assign(t->paramCount(), 0.f);
  t->v.assign(t->paramCount(), 0.f);
  t->acc_gradient.assign(static_cast<size_t>(t->cloud.n), 0.f);
  t->acc_views.assign(static_cast<size_t>(t->cloud.n), 0);
  t->order.resize(t->views.size());
  std::iota(t->order.begin(), t->order.end(), 0);
  std::shuffle(t->order.begin(), t->order.end(), t->engine);
  return t;
}

void orc_trainer_destroy(orc_trainer* t) { delete t; }

int orc_trainer_step_views(orc_trainer* t, int n, const int* view_ids) {
  // SplatTrainer::step, SplatTrainer.cpp:326-362, over n views whose
  // gradients are summed before the one Adam step.
  if (t->iteration >= t->config.iterations) return 0;
  if (t->views.empty() || t->cloud.n == 0) {
    t->iteration = t->config.iterations;
    return 0;
  }
  orc_options opt;
  orc_default_options(&opt);
  opt.sh_degree = t->shDegreeInUse();
  t->grads.resize(t->cloud.n, t->cloud.rs);
  orc_cloud cl = t->cloud.view();
  orc_grads g = t->grads.view();
  float loss = 0.f;
  t->last_views.assign(view_ids, view_ids + n);
  for (int k = 0; k < n; ++k) {
    const orc_view& view = t->views[static_cast<size_t>(view_ids[k])];
    orc_frame* f = orc_frame_create(&cl, &view, &opt);
    std::vector<float> dL(f->rgb.size());
    loss += orc_photometric_loss(f->rgb.data(), t->images[static_cast<size_t>(view_ids[k])].data(), view.width,
                                 view.height, t->config.ssim_weight, dL.data());
    orc_frame_backward(f, &cl, dL.data(), &g);
    orc_frame_destroy(f);
  }
  t->last_loss = n > 0 ? loss / static_cast<float>(n) : 0.f;
  t->smoothed_loss = t->iteration == 0 ? t->last_loss : 0.99f * t->smoothed_loss + 0.01f * t->last_loss;
  for (int i = 0; i < t->cloud.n; ++i) {
    t->acc_gradient[static_cast<size_t>(i)] += t->grads.sg[static_cast<size_t>(i)];
    t->acc_views[static_cast<size_t>(i)] += t->grads.vc[static_cast<size_t>(i)];
  }
  applyAdam(cl, g, t->m.data(), t->v.data(), t->iteration, t->extent, t->config);
  ++t->iteration;
  const orc_config& c = t->config;
  if (t->iteration >= c.densify_from_iteration && t->iteration <= c.densify_until_iteration &&
      t->iteration % c.densify_interval == 0) {
    const int old_n = t->cloud.n;
    const int cap = std::max(old_n * 3 + 8, c.max_gaussians + 8);
    HostCloud work;
    work.rs = t->cloud.rs;
    work.sh_degree = t->cloud.sh_degree;
    work.resize(cap);
    std::copy(t->cloud.pos.begin(), t->cloud.pos.end(), work.pos.begin());
    std::copy(t->cloud.ls.begin(), t->cloud.ls.end(), work.ls.begin());
    std::copy(t->cloud.rot.begin(), t->cloud.rot.end(), work.rot.begin());
    std::copy(t->cloud.op.begin(), t->cloud.op.end(), work.op.begin());
    std::copy(t->cloud.dc.begin(), t->cloud.dc.end(), work.dc.begin());
    std::copy(t->cloud.rest.begin(), t->cloud.rest.end(), work.rest.begin());
    work.n = old_n;
    orc_cloud wc = work.view();
    std::vector<int> source(static_cast<size_t>(cap));
    const int n1 = orc_densify(&wc, cap, t->acc_gradient.data(), t->acc_views.data(), &c, t->extent, t->iteration,
                               &t->densify_rng, source.data(), nullptr);
    source.resize(static_cast<size_t>(n1));
    work.n = n1;
    work.resize(n1);
    t->cloud = work;
    t->remap(source, old_n);
    t->acc_gradient.assign(static_cast<size_t>(n1), 0.f);
    t->acc_views.assign(static_cast<size_t>(n1), 0);
  }
  if (c.opacity_reset_interval > 0 && t->iteration % c.opacity_reset_interval == 0 &&
      t->iteration < c.densify_until_iteration) {
    // resetOpacity, SplatTrainer.cpp:301-309
    const float value = std::clamp(c.opacity_reset_value, 1e-6f, 1.f - 1e-6f);
    const float ceiling = std::log(value / (1.f - value));
    const size_t n = static_cast<size_t>(t->cloud.n);
    for (size_t i = 0; i < n; ++i) {
      t->cloud.op[i] = std::min(t->cloud.op[i], ceiling);
      t->m[10 * n + i] = 0.f;
      t->v[10 * n + i] = 0.f;
    }
  }
  return 1;
}

int orc_trainer_step(orc_trainer* t, int batch) {
  if (t->iteration >= t->config.iterations) return 0;
  if (t->views.empty() || t->cloud.n == 0) {
    t->iteration = t->config.iterations;
    return 0;
  }
  std::vector<int> ids;
  for (int k = 0; k < std::max(1, batch); ++k) ids.push_back(t->nextView());
  return orc_trainer_step_views(t, static_cast<int>(ids.size()), ids.data());
}

int orc_trainer_iteration(const orc_trainer* t) { return t->iteration; }
int orc_trainer_count(const orc_trainer* t) { return t->cloud.n; }
int orc_trainer_sh_degree_in_use(const orc_trainer* t) { return t->shDegreeInUse(); }
float orc_trainer_last_loss(const orc_trainer* t) { return t->last_loss; }
float orc_trainer_smoothed_loss(const orc_trainer* t) { return t->smoothed_loss; }
float orc_trainer_extent(const orc_trainer* t) { return t->extent; }
uint32_t orc_trainer_densify_rng(const orc_trainer* t) { return t->densify_rng; }

int orc_trainer_last_views(const orc_trainer* t, int* out) {
  if (out) std::copy(t->last_views.begin(), t->last_views.end(), out);
  return static_cast<int>(t->last_views.size());
}

void orc_trainer_get_cloud(const orc_trainer* t, orc_cloud* out) {
  const HostCloud& c = t->cloud;
  std::copy(c.pos.begin(), c.pos.end(), out->pos);
  std::copy(c.ls.begin(), c.ls.end(), out->log_scale);
  std::copy(c.rot.begin(), c.rot.end(), out->rot);
  std::copy(c.op.begin(), c.op.end(), out->opacity);
  std::copy(c.dc.begin(), c.dc.end(), out->dc);
  std::copy(c.rest.begin(), c.rest.end(), out->rest);
  out->n = c.n;
}

void orc_trainer_get_moments(const orc_trainer* t, float* m, float* v) {
  if (m) std::copy(t->m.begin(), t->m.end(), m);
  if (v) std::copy(t->v.begin(), t->v.end(), v);
}

void orc_trainer_get_stats(const orc_trainer* t, float* screen_gradient, int* visible_count) {
  if (screen_gradient) std::copy(t->acc_gradient.begin(), t->acc_gradient.end(), screen_gradient);
  if (visible_count) std::copy(t->acc_views.begin(), t->acc_views.end(), visible_count);
}

void orc_trainer_get_last_grads(const orc_trainer* t, orc_grads* g) {
  const Grads& s = t->grads;
  if (g->pos) std::copy(s.pos.begin(), s.pos.end(), g->pos);
  if (g->log_scale) std::copy(s.ls.begin(), s.ls.end(), g->log_scale);
  if (g->rot) std::copy(s.rot.begin(), s.rot.end(), g->rot);
  if (g->opacity) std::copy(s.op.begin(), s.op.end(), g->opacity);
  if (g->dc) std::copy(s.dc.begin(), s.dc.end(), g->dc);
  if (g->rest) std::copy(s.rest.begin(), s.rest.end(), g->rest);
  if (g->screen_gradient) std::copy(s.sg.begin(), s.sg.end(), g->screen_gradient);
  if (g->visible_count) std::copy(s.vc.begin(), s.vc.end(), g->visible_count);
}

int orc_trainer_finish(orc_trainer* t) {
  orc_cloud cl = t->cloud.view();
  const int n = orc_prune_and_sort(&cl, t->config.prune_opacity, nullptr);
  t->cloud.resize(n);
  t->cloud.sh_degree = t->config.sh_degree;
  return n;
}

}  // extern "C"
