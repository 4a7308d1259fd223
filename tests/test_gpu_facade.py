"""Runs the C++ facade tests (tests/cpp/facade_test.cpp): the engine's
trainer tests against HipSplatTrainer / trainSplatsOnGpu / the HIP backend."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_test")


def test_facade_cpp_suite():
    assert os.path.exists(BIN), "build it with __graft_entry__.build() (make -C tests/cpp)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=170)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout[-3000:]
