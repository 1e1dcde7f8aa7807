"""The C++ host surface (raccoongpurender_amd/host/raccoon_host.hpp) through its test program
tests/host_cpp/host_test (built by __graft_entry__.build()).

CPU: scene assembly, BVH build, argument checks and the missing-device error; the OBJ reader
(rt::load_obj) against the Python one (scenes.load_obj) on one file, bit for bit, and both
refusing the same bad records.
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

from raccoongpurender_amd import _lib, scenes

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
    # continuous rendering: images arrived while the step thread ran, their totals never fell,
    # and after stop() the preview holds a published snapshot no newer than the lane
    assert out["live_images"] >= 8 and out["live_steps"] > 4
    assert out["live_monotonic"] and out["live_alpha"]
    assert 0 < out["live_last_total"] <= out["live_final_total"] <= out["live_stat_total"]
    for name in ("image.u8", "image_again.u8"):
        host = _load(d, name, np.uint8).astype(np.int32)
        assert host.size == img.size and np.abs(host - img).max() <= 1, name
    rays = _load(d, "rays.f32", np.float32).view(_lib.RAY_DTYPE)
    for variant in ("stackless", "threaded", "threaded_single"):
        g = _lib.trace(ctx, scene, rays, variant)
        host = _load(d, f"hits_{variant}.i32", np.int32).view(_lib.HIT_DTYPE)
        assert np.array_equal(g.view(np.int32), host.view(np.int32)), variant


OBJ_TEXT = """# a quad, a pentagon and mixed corner forms
mtllib none.mtl
o thing
v 0 0 0
v 1.0 0 0
v 1 1 0 1.0
v 0 1 0
vt 0 0
vn 0 0 1
v 0.1 0.2 0.30000001192092896
v -2.5e-3 7 1e1
g part
usemtl m
s off
f 1 2 3 4
f 1/1 2//1 5/1/1
f -1 -2 -3 -4 -5   # pentagon, counted back from the last vertex
"""


@needs_bin
def test_obj_loader_cpp_and_python_agree(tmp_path):
    path = tmp_path / "mesh.obj"
    path.write_text(OBJ_TEXT)
    out = json.loads(subprocess.run([BIN, "obj", str(path)], check=True, capture_output=True, text=True).stdout)
    pts, tri = scenes.load_obj(str(path))
    assert np.array_equal(np.asarray(out["points_bits"], np.uint32), pts.reshape(-1).view(np.uint32))
    assert np.array_equal(np.asarray(out["indices"], np.uint32), tri.reshape(-1))
    # fans of (0, k, k+1), reversed to clockwise; negative indices from the end
    ccw = scenes.load_obj(OBJ_TEXT, clockwise=False)[1]
    assert ccw.tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 4], [5, 4, 3], [5, 3, 2], [5, 2, 1]]
    assert tri.tolist() == [[a, c, b] for a, b, c in ccw.tolist()]
    assert pts.shape == (6, 3) and pts[5].tolist() == [np.float32(-2.5e-3), 7.0, 10.0]


@needs_bin
@pytest.mark.parametrize("bad", ["v 1 2\n", "v 1 2 x\n", "v 0 0 0\nf 1 2\n", "v 0 0 0\nf 1 2 3\n",
                                 "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 0 2\n", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 -4\n"])
def test_obj_loader_rejects_bad_records(tmp_path, bad):
    path = tmp_path / "bad.obj"
    path.write_text(bad)
    out = json.loads(subprocess.run([BIN, "obj", str(path)], check=True, capture_output=True, text=True).stdout)
    assert out == {"error": -1}
    with pytest.raises(ValueError):
        scenes.load_obj(bad)


REF_SCENES = "/root/reference/scenes"


@needs_bin
@pytest.mark.parametrize("name,ntris", [("bunny", 34916), ("dragon", 100000)])
def test_obj_readers_agree_on_the_reference_meshes(name, ntris):
    """The reference's own meshes (scenes/bunny.obj, scenes/dragon.obj), read by the C++ reader
    (rt::load_obj) and the Python one (scenes.load_obj): the same points bit for bit and the same
    triangles. Runs where the reference's scene files are present."""
    path = os.path.join(REF_SCENES, name + ".obj")
    if not os.path.exists(path):
        pytest.skip("the reference's scene files are not on this machine")
    out = json.loads(subprocess.run([BIN, "obj", path], check=True, capture_output=True, text=True).stdout)
    pts, tri = scenes.load_obj(path)
    assert len(tri) == ntris
    assert np.array_equal(np.asarray(out["points_bits"], np.uint32), pts.reshape(-1).view(np.uint32))
    assert np.array_equal(np.asarray(out["indices"], np.uint32), tri.reshape(-1))


PACK_BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host_cpp", "pack_test")


@pytest.mark.skipif(not os.path.exists(PACK_BIN), reason="pack_test not built (run __graft_entry__.build())")
def test_stackless_leaf_links():
    """The packer marks exactly the child and sibling links that name leaves, and the ids under
    the marks are the unmarked packing's (tests/host_cpp/pack_test.cpp checks every record of five
    trees: leaf limits 1, 4 and 8, with and without spheres, and a single-leaf tree)."""
    res = subprocess.run([PACK_BIN], capture_output=True, text=True, timeout=120)
    trees = [json.loads(line) for line in res.stdout.splitlines()]
    assert res.returncode == 0 and len(trees) == 5 and all(t["ok"] for t in trees)
    for t in trees:
        # every leaf but a leaf root is some branch's child, so it is marked exactly once
        assert t["marks"] == (t["leaves"] if t["nodes"] > 1 else 0)
