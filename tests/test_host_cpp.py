"""The C++ host surface (raccoongpurender_amd/host/raccoon_host.hpp) through its test program
tests/host_cpp/host_test (built by __graft_entry__.build()).

CPU: scene assembly, BVH build, argument checks and the missing-device error.
GPU: SceneManager -> WavefrontPathTracing.launch_fixed/re_start and the two BVH ray casters;
the program dumps its scene, and the Python binding redoes the same render and ray casts
through the C ABI: sample counts equal, images within one 8-bit step (float atomics add the
pixel sums in arbitrary order), hits bit-identical.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from raccoongpurender_amd import _lib

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host_cpp", "host_test")
needs_bin = pytest.mark.skipif(not os.path.exists(BIN), reason="host_test not built (run __graft_entry__.build())")


@needs_bin
def test_host_surface_on_the_cpu():
    out = json.loads(subprocess.run([BIN, "cpu"], check=True, capture_output=True, text=True).stdout)
    assert out["triangles"] == 12 and out["spheres"] == 2
    assert out["slots"] == 14 and out["nodes"] == 2 * out["leaves"] - 1
    assert out["bad_index_rejected"]
    if _lib.device_count() == 0:
        assert out["no_device"] and out["code"] == -4


def _load(d, name, dtype):
    return np.fromfile(os.path.join(d, name), dtype=dtype)


@needs_bin
@pytest.mark.gpu
def test_host_surface_renders_like_the_binding(ctx, tmp_path):
    d = str(tmp_path)
    res = subprocess.run([BIN, "gpu", d], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    out = json.loads(res.stdout)
    w, h, n = out["width"], out["height"], out["paths"]
    points = _load(d, "points.f32", np.float32).reshape(-1, 3)
    indices = _load(d, "indices.u32", np.uint32).reshape(-1, 3)
    spheres = _load(d, "spheres.f32", np.float32).reshape(-1, 4)
    lam = _load(d, "lambertians.f32", np.float32).reshape(-1, 7)
    ms = _lib.MaterialSet(_load(d, "types.i32", np.int32), _load(d, "idx.i32", np.int32),
                          lambertians=[(r[0:3], r[3:6]) for r in lam],
                          dielectrics=[float(x) for x in _load(d, "iors.f32", np.float32)])
    bvh = _lib.BVH(points, indices, spheres)
    scene = _lib.Scene(ctx, bvh)
    mats = _lib.Materials(ctx, ms)
    env = _lib.Envmap(ctx, _load(d, "env.f32", np.float32).reshape(8, 16, 3))
    c = _load(d, "camera.f32", np.float32)
    cam = _lib.rgpu_camera()
    for k, name in enumerate(("eye", "forward", "up", "right", "imageplane_o", "imageplane_r", "imageplane_b")):
        getattr(cam, name)[:] = [float(x) for x in c[3 * k:3 * k + 3]]
    cam.width, cam.height, cam.lens_radius = w, h, float(c[21])
    wf = _lib.Wavefront(ctx, scene, mats, env, n, seed=7)
    wf.initialize(cam)
    wf.step(12)
    img = wf.tonemap().reshape(-1).astype(np.int32)
    assert abs(wf.stat() / (w * h) - out["spp"]) < 1e-6 and out["steps"] == 12
    for name in ("image.u8", "image_again.u8"):
        host = _load(d, name, np.uint8).astype(np.int32)
        assert host.size == img.size and np.abs(host - img).max() <= 1, name
    rays = _load(d, "rays.f32", np.float32).view(_lib.RAY_DTYPE)
    for variant in ("stackless", "threaded"):
        g = _lib.trace(ctx, scene, rays, variant)
        host = _load(d, f"hits_{variant}.i32", np.int32).view(_lib.HIT_DTYPE)
        assert np.array_equal(g.view(np.int32), host.view(np.int32)), variant
