// The engine's trainer tests (engine/native/cpp/tests/SplatTrainerTest.cpp
// and the accumulation / background tests of SplatRasterTest.cpp),
// re-expressed against the HIP facade. Prints one PASS/FAIL line per test;
// the exit status is the number of failures.
#include <parallax_hip/splat_trainer.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <vector>

using namespace parallax_hip;

namespace {

int failures = 0;

void check(bool ok, const char* test, const std::string& what) {
  if (!ok) {
    ++failures;
    std::printf("FAIL %s: %s\n", test, what.c_str());
  }
}

// xorshift64 of the engine's tests (SyntheticScene.h:80-107)
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed ? seed : 1) {}
  uint64_t next() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
  }
  double uniform(double lo, double hi) { return lo + (hi - lo) * ((next() >> 11) * (1.0 / 9007199254740992.0)); }
};

std::vector<SplatView> ringOfViews(int count, double radius, int width, int height) {
  std::vector<SplatView> views;
  for (int i = 0; i < count; ++i) {
    const double a = 2.0 * M_PI * i / count;
    const double eye[3] = {radius * std::sin(a), 0.0, -radius * std::cos(a)};
    const double len = std::sqrt(eye[0] * eye[0] + eye[1] * eye[1] + eye[2] * eye[2]);
    const double f[3] = {-eye[0] / len, -eye[1] / len, -eye[2] / len};
    // right = UnitY x forward, down = forward x right
    double r[3] = {1.0 * f[2] - 0.0 * f[1], 0.0 * f[0] - 0.0 * f[2], 0.0 * f[1] - 1.0 * f[0]};
    const double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (double& x : r) x /= rl;
    const double d[3] = {f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2], f[0] * r[1] - f[1] * r[0]};
    SplatView v;
    v.intrinsics.width = width;
    v.intrinsics.height = height;
    v.intrinsics.fx = v.intrinsics.fy = 1.1 * width;
    v.intrinsics.cx = 0.5 * width;
    v.intrinsics.cy = 0.5 * height;
    const double R[9] = {r[0], r[1], r[2], d[0], d[1], d[2], f[0], f[1], f[2]};
    for (int k = 0; k < 9; ++k) v.camera_from_world.R[static_cast<size_t>(k)] = R[k];
    for (int k = 0; k < 3; ++k)
      v.camera_from_world.t[static_cast<size_t>(k)] = -(R[3 * k] * eye[0] + R[3 * k + 1] * eye[1] + R[3 * k + 2] * eye[2]);
    views.push_back(v);
  }
  return views;
}

GaussianCloud truthCloud() {
  GaussianCloud c;
  const float pos[3][3] = {{-0.25f, 0.f, 0.f}, {0.25f, 0.f, 0.f}, {0.f, 0.2f, 0.2f}};
  const float col[3][3] = {{0.9f, 0.2f, 0.2f}, {0.2f, 0.8f, 0.3f}, {0.3f, 0.3f, 0.9f}};
  for (int i = 0; i < 3; ++i) {
    Gaussian g;
    for (int k = 0; k < 3; ++k) {
      g.position[static_cast<size_t>(k)] = pos[i][k];
      g.log_scale[static_cast<size_t>(k)] = std::log(0.12f);
      g.dc[static_cast<size_t>(k)] = (col[i][k] - 0.5f) / kShC0;
    }
    g.opacity_logit = 4.f;
    c.gaussians.push_back(g);
  }
  return c;
}

GaussianCloud roughStart(size_t count = 24, unsigned seed = 3) {
  GaussianCloud c;
  Rng rng(seed);
  for (size_t i = 0; i < count; ++i) {
    Gaussian g;
    g.position = {static_cast<float>(rng.uniform(-0.4, 0.4)), static_cast<float>(rng.uniform(-0.3, 0.3)),
                  static_cast<float>(rng.uniform(-0.3, 0.3))};
    g.log_scale = {std::log(0.1f), std::log(0.1f), std::log(0.1f)};
    c.gaussians.push_back(g);
  }
  return c;
}

std::vector<SplatTrainingView> renderTrainingSet(const GaussianCloud& cloud, const std::vector<SplatView>& views) {
  std::vector<SplatTrainingView> set;
  for (const SplatView& v : views) set.push_back({v, renderSplats(cloud, v, {}).rgb});
  return set;
}

SplatTrainingConfig shortSchedule(int iterations) {
  SplatTrainingConfig c;
  c.iterations = iterations;
  c.sh_degree = 0;
  c.sh_increase_interval = 0;
  c.densify_from_iteration = iterations + 1;
  c.densify_until_iteration = iterations + 1;
  c.opacity_reset_interval = 0;
  c.ssim_weight = 0.2f;
  return c;
}

float opacityOf(const Gaussian& g) { return 1.f / (1.f + std::exp(-g.opacity_logit)); }

void run(const char* name, const std::function<void(const char*)>& body) {
  const int before = failures;
  body(name);
  if (failures == before) std::printf("PASS %s\n", name);
  std::fflush(stdout);
}

}  // namespace

int main() {
  const SplatGpuAvailability avail = splatGpuAvailability();
  if (!avail.available) {
    std::printf("FAIL availability: %s\n", avail.reason.c_str());
    return 1;
  }
  std::printf("device %s\n", avail.device_name.c_str());

  run("ConfigValidation", [](const char* t) {
    check(SplatTrainingConfig{}.validate().empty(), t, "defaults");
    SplatTrainingConfig c;
    c.iterations = 0;
    check(!c.validate().empty(), t, "iterations");
    c = {};
    c.sh_degree = 7;
    check(!c.validate().empty(), t, "degree");
    c = {};
    c.split_scale_divider = 0.8f;
    check(!c.validate().empty(), t, "divider");
    c = {};
    c.ssim_weight = 1.5f;
    check(!c.validate().empty(), t, "ssim");
  });

  run("EmptyAreasShowTheBackground", [](const char* t) {
    SplatView v = ringOfViews(1, 2.0, 48, 36)[0];
    RasterOptions o;
    o.background = {0.25f, 0.5f, 0.75f};
    GaussianCloud empty;
    SplatImage img = renderSplats(empty, v, o);
    bool ok = img.rgb.size() == 48u * 36u * 3u;
    for (size_t p = 0; ok && p < img.transmittance.size(); ++p)
      ok = img.transmittance[p] == 1.f && img.rgb[3 * p] == 0.25f && img.rgb[3 * p + 1] == 0.5f &&
           img.rgb[3 * p + 2] == 0.75f;
    check(ok, t, "background and T = 1 exactly");
  });

  run("SceneExtentFollowsTheCameraRig", [](const char* t) {
    for (double radius : {1.0, 4.0}) {
      HipSplatTrainer tr(roughStart(), renderTrainingSet(truthCloud(), ringOfViews(4, radius, 16, 16)),
                         shortSchedule(1));
      check(std::abs(tr.sceneExtent() - radius * 1.1) < 1e-3, t, "extent");
    }
  });

  run("TrainingDrivesTheLossDown", [](const char* t) {
    const auto training = renderTrainingSet(truthCloud(), ringOfViews(6, 1.6, 32, 32));
    HipSplatTrainer tr(roughStart(), training, shortSchedule(400));
    check(tr.ok(), t, tr.error());
    check(tr.step(), t, "first step");
    const float first = tr.lastLoss();
    check(tr.train(), t, "train");
    check(tr.smoothedLoss() < 0.5f * first, t,
          "loss " + std::to_string(first) + " -> " + std::to_string(tr.smoothedLoss()));
  });

  run("TrainingIsDeterministicForAGivenSeed", [](const char* t) {
    const auto training = renderTrainingSet(truthCloud(), ringOfViews(4, 1.6, 16, 16));
    SplatTrainingConfig c = shortSchedule(60);
    c.densify_from_iteration = 20;
    c.densify_until_iteration = 60;
    c.densify_interval = 20;
    c.densify_grad_threshold = 0.f;
    c.deterministic = true;
    auto once = [&]() {
      HipSplatTrainer tr(roughStart(8), training, c);
      tr.train();
      return tr.finish();
    };
    const GaussianCloud a = once(), b = once();
    bool same = a.size() == b.size() && a.size() > 0;
    for (size_t i = 0; same && i < a.size(); ++i)
      same = a.gaussians[i].position == b.gaussians[i].position &&
             a.gaussians[i].opacity_logit == b.gaussians[i].opacity_logit;
    check(same, t, "two seeded runs differ");
  });

  run("CancellingStopsTrainingWhereItWasAsked", [](const char* t) {
    const auto training = renderTrainingSet(truthCloud(), ringOfViews(4, 1.6, 16, 16));
    HipSplatTrainer tr(roughStart(8), training, shortSchedule(500));
    int seen = 0;
    const bool done = tr.train([&](const SplatTrainingProgress& p) {
      seen = p.iteration;
      return p.iteration < 25;
    });
    check(!done && seen == 25 && tr.iteration() == 25, t, "stopped at " + std::to_string(tr.iteration()));
  });

  run("SphericalHarmonicBandsUnlockOnSchedule", [](const char* t) {
    const auto training = renderTrainingSet(truthCloud(), ringOfViews(3, 1.6, 12, 12));
    GaussianCloud start = roughStart(4);
    start.sh_degree = 3;
    for (Gaussian& g : start.gaussians) g.sh_rest.assign(45, 0.f);
    SplatTrainingConfig c = shortSchedule(31);
    c.sh_degree = 3;
    c.sh_increase_interval = 10;
    HipSplatTrainer tr(start, training, c);
    check(tr.shDegreeInUse() == 0, t, "starts at 0");
    for (int target = 1; target <= 3; ++target) {
      while (tr.iteration() < target * 10) tr.step();
      check(tr.shDegreeInUse() == target, t, "degree " + std::to_string(target));
    }
  });

  run("DensificationAddsGaussiansAndRespectsTheCeiling", [](const char* t) {
    const auto training = renderTrainingSet(truthCloud(), ringOfViews(4, 1.6, 24, 24));
    SplatTrainingConfig c = shortSchedule(20);
    c.densify_from_iteration = 10;
    c.densify_until_iteration = 20;
    c.densify_interval = 10;
    c.densify_grad_threshold = 0.f;
    HipSplatTrainer grown(roughStart(8), training, c);
    grown.train();
    check(grown.cloud().size() > 8, t, "grows");
    c.densify_grad_threshold = 1e9f;
    HipSplatTrainer same(roughStart(8), training, c);
    same.train();
    check(same.cloud().size() == 8, t, "unchanged");
    SplatTrainingConfig cap = shortSchedule(40);
    cap.densify_from_iteration = 5;
    cap.densify_until_iteration = 40;
    cap.densify_interval = 5;
    cap.densify_grad_threshold = 0.f;
    cap.max_gaussians = 12;
    HipSplatTrainer capped(roughStart(10), training, cap);
    capped.train();
    check(capped.cloud().size() <= 12, t, "cap");
  });

  run("TransparentGaussiansArePrunedAway", [](const char* t) {
    const auto training = renderTrainingSet(truthCloud(), ringOfViews(4, 1.6, 16, 16));
    GaussianCloud start = roughStart(10);
    for (Gaussian& g : start.gaussians) g.opacity_logit = -20.f;
    SplatTrainingConfig c = shortSchedule(10);
    c.densify_from_iteration = 10;
    c.densify_until_iteration = 10;
    c.densify_interval = 10;
    c.densify_grad_threshold = 1e9f;
    c.opacity_lr = 0.f;
    HipSplatTrainer tr(start, training, c);
    tr.train();
    check(tr.cloud().empty(), t, "not pruned");
  });

  run("OpacityResetCapsEveryGaussian", [](const char* t) {
    const auto training = renderTrainingSet(truthCloud(), ringOfViews(4, 1.6, 16, 16));
    GaussianCloud start = roughStart(6);
    for (Gaussian& g : start.gaussians) g.opacity_logit = 5.f;
    SplatTrainingConfig c = shortSchedule(60);
    c.opacity_reset_interval = 20;
    c.densify_until_iteration = 40;
    c.densify_from_iteration = 41;
    c.opacity_lr = 0.f;
    HipSplatTrainer tr(start, training, c);
    while (tr.iteration() < 20) tr.step();
    bool ok = true;
    for (const Gaussian& g : tr.cloud().gaussians) ok = ok && opacityOf(g) < 0.011f;
    check(ok, t, "opacity above 0.011");
  });

  run("AnEmptyCloudOrViewSetFinishes", [](const char* t) {
    const auto training = renderTrainingSet(truthCloud(), ringOfViews(2, 1.6, 8, 8));
    HipSplatTrainer none(GaussianCloud{}, training, shortSchedule(10));
    check(none.train() && none.cloud().empty(), t, "empty cloud");
    HipSplatTrainer noviews(roughStart(4), {}, shortSchedule(10));
    check(noviews.train() && noviews.cloud().size() == 4, t, "no views");
  });

  run("FinishPrunesSortsAndRecordsTheDegree", [](const char* t) {
    const auto training = renderTrainingSet(truthCloud(), ringOfViews(3, 1.6, 12, 12));
    GaussianCloud start = roughStart(6);
    start.gaussians[0].opacity_logit = -30.f;
    start.gaussians[3].opacity_logit = -30.f;
    SplatTrainingConfig c = shortSchedule(1);
    c.opacity_lr = 0.f;
    HipSplatTrainer tr(start, training, c);
    tr.train();
    const GaussianCloud done = tr.finish();
    bool sorted = done.size() == 4 && done.sh_degree == 0;
    for (size_t i = 1; sorted && i < done.size(); ++i)
      sorted = done.gaussians[i - 1].opacity_logit >= done.gaussians[i].opacity_logit;
    check(sorted, t, "size " + std::to_string(done.size()));
  });

  run("TrainSplatsOnGpuTrainsInPlace", [](const char* t) {
    const auto training = renderTrainingSet(truthCloud(), ringOfViews(4, 1.6, 16, 16));
    GaussianCloud cloud = roughStart(8);
    std::string err;
    const bool ok = trainSplatsOnGpu(&cloud, training, shortSchedule(30), nullptr, &err);
    check(ok && cloud.size() == 8, t, err);
    GaussianCloud cancelled = roughStart(8);
    const bool ok2 = trainSplatsOnGpu(&cancelled, training, shortSchedule(30),
                                      [](const SplatTrainingProgress& p) { return p.iteration < 5; }, &err);
    check(ok2 && cancelled.size() == 8, t, "cancel keeps the field");
  });

  run("BackendIsHipAndAccumulates", [](const char* t) {
    std::string err;
    std::unique_ptr<SplatBackend> b = makeHipSplatBackend(&err);
    check(b != nullptr && b->name() == "hip", t, err);
    if (!b) return;
    const SplatView v = ringOfViews(1, 1.6, 24, 20)[0];
    const GaussianCloud cloud = roughStart(3);
    std::vector<float> target(24 * 20 * 3, 0.3f);
    RasterOptions o;
    o.deterministic = true;
    SplatGradients once;
    b->renderAndAccumulate(cloud, v, o, target, 0.f, &once);
    SplatGradients twice = once;
    b->renderAndAccumulate(cloud, v, o, target, 0.f, &twice);
    bool ok = true;
    for (size_t i = 0; i < cloud.size(); ++i) {
      ok = ok && std::abs(2.f * once.opacity_logit[i] - twice.opacity_logit[i]) < 1e-5f;
      ok = ok && 2 * once.visible_count[i] == twice.visible_count[i];
    }
    check(ok, t, "twice != 2x once");
  });

  run("init_from_points_spacing", [](const char* t) {
    // a 6x6x6 grid of pitch 0.05: every sampled point's nearest neighbour is
    // one pitch away, so log sigma = log(0.025)
    std::vector<std::array<float, 3>> pts;
    std::vector<std::array<uint8_t, 3>> cols;
    for (int z = 0; z < 6; ++z)
      for (int y = 0; y < 6; ++y)
        for (int x = 0; x < 6; ++x) {
          pts.push_back({0.05f * x, 0.05f * y, 0.05f * z});
          cols.push_back({static_cast<uint8_t>(40 * x), 128, 255});
        }
    std::string err;
    const GaussianCloud c = initializeFromPointCloud(pts, cols, 1, 0.1f, &err);
    check(c.size() == pts.size() && c.sh_degree == 1, t, "size " + std::to_string(c.size()) + " " + err);
    bool ok = !c.empty();
    for (const Gaussian& g : c.gaussians) {
      ok = ok && std::abs(g.log_scale[0] - std::log(0.025f)) < 1e-5f && g.rotation[0] == 1.f &&
           g.sh_rest.size() == 9 && std::abs(g.opacity_logit - std::log(0.1f / 0.9f)) < 1e-6f;
    }
    check(ok, t, "fields");
    check(std::abs(c.gaussians[0].dc[2] - (1.f - 0.5f) / kShC0) < 1e-6f, t, "colour");
  });

  run("ply_round_trip", [](const char* t) {
    GaussianCloud c = roughStart(40, 9);
    c.sh_degree = 1;
    for (Gaussian& g : c.gaussians) g.sh_rest.assign(9, 0.25f);
    const std::string path = "/tmp/parallax_hip_facade_test.ply";
    std::string err;
    check(writeGaussianPly(path, c, &err), t, "write: " + err);
    GaussianCloud back;
    check(readGaussianPly(path, &back, &err), t, "read: " + err);
    bool same = back.size() == c.size() && back.sh_degree == 1;
    for (size_t i = 0; same && i < c.size(); ++i)
      same = back.gaussians[i].position == c.gaussians[i].position &&
             back.gaussians[i].sh_rest == c.gaussians[i].sh_rest;
    check(same, t, "round trip");
    std::remove(path.c_str());
    check(!readGaussianPly(path, &back, &err) && !err.empty(), t, "missing file reports an error");
  });

  run("compact_round_trips", [](const char* t) {
    GaussianCloud c = roughStart(40, 11);
    c.sh_degree = 1;
    c.antialiased = true;
    for (Gaussian& g : c.gaussians) g.sh_rest.assign(9, 0.5f);
    std::string err;
    // .spz: positions on the 12-bit grid, SH bytes exact for multiples of 1/128
    const std::string spz = "/tmp/parallax_hip_facade_test.spz";
    check(writeSpz(spz, c, 12, &err), t, "write spz: " + err);
    GaussianCloud back;
    check(readSpz(spz, &back, &err), t, "read spz: " + err);
    bool ok = back.size() == c.size() && back.sh_degree == 1 && back.antialiased;
    for (size_t i = 0; ok && i < c.size(); ++i) {
      for (int k = 0; k < 3; ++k)
        ok = ok && std::fabs(back.gaussians[i].position[k] - c.gaussians[i].position[k]) <= 0.5f / 4096.f;
      ok = ok && back.gaussians[i].sh_rest == c.gaussians[i].sh_rest;
    }
    check(ok, t, ".spz round trip");
    std::remove(spz.c_str());
    // .splat: positions exact, SH dropped
    const std::string splat = "/tmp/parallax_hip_facade_test.splat";
    check(writeSplat(splat, c, &err), t, "write splat: " + err);
    check(readSplat(splat, &back, &err), t, "read splat: " + err);
    ok = back.size() == c.size() && back.sh_degree == 0;
    for (size_t i = 0; ok && i < c.size(); ++i)
      ok = back.gaussians[i].position == c.gaussians[i].position && back.gaussians[i].sh_rest.empty();
    check(ok, t, ".splat round trip");
    std::remove(splat.c_str());
    check(!readSpz(splat, &back, &err) && err.find("cannot read") != std::string::npos, t,
          "missing file reports the engine's message");
  });

  run("splat_stage_on_gpu", [](const char* t) {
    const GaussianCloud truth = truthCloud();
    const auto set = renderTrainingSet(truth, ringOfViews(4, 2.5, 48, 40));
    std::vector<std::array<float, 3>> pts;
    std::vector<std::array<uint8_t, 3>> cols;
    Rng rng(5);
    for (int i = 0; i < 400; ++i) {
      pts.push_back({static_cast<float>(rng.uniform(-0.4, 0.4)), static_cast<float>(rng.uniform(-0.3, 0.3)),
                     static_cast<float>(rng.uniform(-0.3, 0.3))});
      cols.push_back({200, 100, 50});
    }
    SplatStageSettings s;
    s.iterations = 60;
    s.sh_degree = 0;
    s.max_gaussians = 400;  // seeds a quarter: every 4th point
    const std::string path = "/tmp/parallax_hip_stage_test.ply";
    SplatStageResult r;
    std::string err;
    const bool ok = runSplatStageOnGpu(pts, cols, set, s, nullptr, {path}, &r, &err);
    check(ok && r.backend == "hip" && r.iterations_done == 60, t, err + " " + r.note);
    check(!r.cloud.empty() && r.cloud.size() <= 100, t, "seeded from a quarter: " + std::to_string(r.cloud.size()));
    GaussianCloud back;
    check(readGaussianPly(path, &back, &err) && back.size() == r.cloud.size(), t, "field written: " + err);
    std::remove(path.c_str());
    // a zero time budget stops at once and keeps the field; cancellation fails
    s.time_budget_millis = 1;
    SplatStageResult timed;
    check(runSplatStageOnGpu(pts, cols, set, s, nullptr, {}, &timed, &err) && timed.ran_out_of_time, t,
          "time budget: " + timed.note);
    check(!runSplatStageOnGpu(pts, cols, set, s, [] { return true; }, {}, &timed, &err) && timed.cancelled, t,
          "cancel");
    check(!runSplatStageOnGpu({}, cols, set, s, nullptr, {}, &timed, &err), t, "no points");
  });

  std::printf("%d failure(s)\n", failures);
  return failures;
}
