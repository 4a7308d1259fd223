"""The C-ABI library loads and exports every entry point include/*.h
declares, and the Python binding's table covers them (no device needed)."""
import ctypes
import os
import re

from parallax_amd import splat as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fname in os.listdir(os.path.join(ROOT, "include")):
        if not fname.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fname)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(splat_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_abi():
    names = declared_functions()
    for required in ("splat_create", "splat_render_forward", "splat_render_backward", "splat_train_step",
                     "splat_adam_step", "splat_densify", "splat_loss"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(S.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert declared_functions() <= set(S.SIGNATURES), sorted(declared_functions() - set(S.SIGNATURES))
    S.lib()  # every signature resolves


def test_no_device_is_reported_not_crashed():
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        return  # a device may exist here; the GPU tests cover that path
    ok, _, reason = S.availability()
    assert not ok and reason
