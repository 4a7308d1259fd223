"""include/parallax_hip/engine_adapter.hpp compiled against the engine's own
headers and vendored Eigen (engine/native/cpp/include,
engine/native/cpp/third_party/eigen), linked with the HIP library and the
reference build (oracle/_ref, for SplatTrainingConfig::validate,
SplatTrainer.cpp:118), and run: exact round trips of an SH0-SH3
GaussianCloud, SplatViews, SplatTrainingViews, a SplatTrainingConfig with
every field off its default and SplatTrainingProgress; validate() agreeing
on both sides; and, with no device, trainSplatsOnHip returning false with a
reason and the cloud unchanged (the contract of trainSplatsOnGpu,
VulkanSplatTrainer.h:27-37, that Pipeline.cpp:1961 relies on to fall back).
Skipped where the engine source or either library is absent.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENGINE = "/root/reference/engine/native/cpp"
LIB = os.path.join(ROOT, "parallax_amd")
REF_LIB = os.path.join(ROOT, "oracle", "_ref")


def test_engine_adapter_round_trips(tmp_path):
    if not os.path.isdir(os.path.join(ENGINE, "include", "parallax")):
        pytest.skip("engine headers not present")
    for so in (os.path.join(LIB, "libparallax_splat.so"), os.path.join(REF_LIB, "libparallax_ref.so")):
        if not os.path.exists(so):
            pytest.skip(f"{so} not built")
    exe = str(tmp_path / "adapter_test")
    build = subprocess.run(
        ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
         "-I" + os.path.join(ROOT, "include"),
         "-I" + os.path.join(ENGINE, "include"),
         "-I" + os.path.join(ENGINE, "third_party", "eigen"),
         "-o", exe, os.path.join(ROOT, "tests", "cpp", "adapter_test.cpp"),
         "-L" + LIB, "-lparallax_splat", "-L" + REF_LIB, "-lparallax_ref",
         "-Wl,-rpath," + LIB + ":" + REF_LIB],
        capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-4000:]
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(run.stdout)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "ok: 0 failure(s)" in run.stdout
