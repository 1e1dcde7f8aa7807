"""The C ABI library loads without a GPU and exports every symbol include/raccoon_gpu.h declares."""
import ctypes
import os
import re

import numpy as np

from raccoongpurender_amd import _lib
from raccoongpurender_amd import guide as _guide  # noqa: F401  (registers the rgpu_guide_* signatures)

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "raccoon_gpu.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rgpu_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    syms = declared_symbols()
    assert len(syms) >= 40
    for s in syms:
        assert hasattr(_lib.lib, s), s
    assert sorted(set(_lib.EXPORTED)) == syms


def test_host_only_entry_points_work_without_gpu():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.array([[0, 1, 2]], np.uint32)
    b = _lib.BVH(v, f)
    assert b.nodes == 1 and b.slots == 1
    e = _lib.Envmap(None, np.ones((4, 8, 3), np.float32))
    assert e.export(0)["pdf"].shape == (32,)
    cam = _lib.camera_look_at((0, 0, 5), (0, 0, 0), (0, 1, 0), 45.0, 64, 32)
    assert cam.width == 64 and abs(np.linalg.norm(cam.as_array()[3:6]) - 1.0) < 1e-6


def test_invalid_arguments_are_status_codes():
    h = ctypes.c_void_p()
    assert _lib.lib.rgpu_bvh_build(None, 3, 0, None, 1, None, 0, None, ctypes.byref(h)) == -1
    assert _lib.lib.rgpu_create(0, None, None) == -1
    opt = _lib.rgpu_bvh_options(99, 4, 16, 63)
    v = np.zeros((3, 3), np.float32)
    f = np.array([0, 1, 2], np.uint32)
    assert _lib.lib.rgpu_bvh_build(_lib.ptr(v), 3, 3, _lib.ptr(f), 1, None, 0, ctypes.byref(opt),
                                   ctypes.byref(h)) == -1
    bad = np.array([0, 1, 7], np.uint32)  # index out of range
    assert _lib.lib.rgpu_bvh_build(_lib.ptr(v), 3, 3, _lib.ptr(bad), 1, None, 0, None, ctypes.byref(h)) == -1
    cam = _lib.rgpu_camera()
    z = np.zeros(3, np.float32)
    assert _lib.lib.rgpu_camera_look_at(_lib.ptr(z), _lib.ptr(z), _lib.ptr(z), 45.0, 0.0, 8, 8, 0.0,
                                        ctypes.byref(cam)) == -1


def test_no_device_is_reported_not_fatal():
    if _lib.device_count() > 0:
        return
    h = ctypes.c_void_p()
    assert _lib.lib.rgpu_create(0, None, ctypes.byref(h)) == -4
