"""GPU parity of extension ray casting (every traversal variant) against the oracle.

Bit-exact bar: primitive index, t, u and v of every ray equal the oracle's bits when
the oracle runs the reference algorithm (kernels/extension_ray_cast_stackless.cl,
common/threaded_bvh.hpp:527-569 and 366-395) on the reference layouts restated from the same tree.
"""
import numpy as np
import pytest

import oracle as O
from raccoongpurender_amd import _lib, scenes
from tests import helpers as H

pytestmark = pytest.mark.gpu


VARIANTS = ("stackless", "threaded", "threaded_single")


def _oracle(tree, v, f, sph, rays, variant):
    if variant == "stackless":
        return O.intersect_stackless(O.RefStackless(tree), v, f, rays, sph)
    if variant == "threaded":
        return O.intersect_mtbvh(O.RefMTBVH(tree), v, f, rays, sph)
    return O.intersect_tbvh(O.RefMTBVH(tree), v, f, rays, sph)


def _ray_sets(v):
    lo, hi = v.min(0), v.max(0)
    return {
        "incoherent": scenes.incoherent_rays(20000, lo, hi, seed=11),
        "primary": scenes.primary_rays(96, 96),
        "vertex": H.degenerate_rays(v, 4000, 3),
        "clipped": H.with_tmax(scenes.incoherent_rays(4000, lo, hi, seed=12), np.float32(0.25)),
        "zero_tmax": H.with_tmax(scenes.incoherent_rays(500, lo, hi, seed=13), np.float32(0.0)),
    }


def _meshes():
    v, f = scenes.displaced_icosphere(level=5)
    yield "icosphere5", v, f, None
    v, f = H.triangle_soup(4000, 1)
    yield "soup", v, f, None
    v, f = H.grid_plane(32)
    yield "grid", v, f, None
    v, f = H.triangle_soup(300, 2)
    rng = np.random.default_rng(4)
    sph = np.concatenate([rng.uniform(-1, 1, (60, 3)), rng.uniform(0.02, 0.15, (60, 1))], 1).astype(np.float32)
    yield "soup+spheres", v, f, sph


@pytest.mark.parametrize("variant", VARIANTS)
def test_bit_exact_small_scenes(ctx, variant):
    for name, v, f, sph in _meshes():
        bvh = _lib.BVH(v, f, sph)
        scene = _lib.Scene(ctx, bvh)
        tree = bvh.export()
        for rname, rays in _ray_sets(v).items():
            print(f"case {name}/{rname}/{variant} n={len(rays)}", flush=True)
            g = _lib.trace(ctx, scene, rays, variant)
            o, _ = _oracle(tree, v, f, sph, rays, variant)
            assert H.compare_hits(g, o) == dict(prim=0, t=0, uv=0), (name, rname, variant)
        scene.close()


def test_variants_agree_and_block_sizes_identical(ctx):
    v, f = scenes.displaced_icosphere(level=5)
    bvh = _lib.BVH(v, f)
    scene = _lib.Scene(ctx, bvh)
    rays = scenes.incoherent_rays(50000, v.min(0), v.max(0), seed=21)
    base = _lib.trace(ctx, scene, rays, "stackless")
    for block in (64, 128, 256):
        for variant in VARIANTS:
            g = _lib.trace(ctx, scene, rays, variant, block=block)
            ties, wrong = H.tie_excepted_mismatch(g, base)
            assert wrong == 0
            if variant == "stackless":
                assert H.compare_hits(g, base) == dict(prim=0, t=0, uv=0)


def test_persistent_launch_shapes_match_per_ray_kernel(ctx):
    # block 0 runs the persistent kernel (waves refilling idle lanes from their ray chunk);
    # every shape must give the one-thread-per-ray kernel's bits
    v, f = scenes.displaced_icosphere(level=5)
    scene = _lib.Scene(ctx, _lib.BVH(v, f))
    for rays in (scenes.incoherent_rays(30011, v.min(0), v.max(0), seed=23), scenes.primary_rays(64, 48),
                 scenes.incoherent_rays(37, v.min(0), v.max(0), seed=24)):
        for variant in VARIANTS:
            ref = _lib.trace(ctx, scene, rays, variant, block=256)
            try:
                for wps, refill, lb, tl in ((1, 1, 1, 1), (4, 16, 64, 7), (8, 64, 8, 1024), (16, 8, 32, 0),
                                            (0, 0, 0, _lib.TREELET_OFF), (0, 0, 0, 0)):
                    # tl: top-of-tree records in LDS (stackless), up to the 1024 maximum
                    ctx.set_trace_tuning(wps, refill, lb, tl)
                    # block 0: the persistent kernel; 64: one thread per ray, leaves batched when lb > 1
                    for block in (0, 64):
                        g = _lib.trace(ctx, scene, rays, variant, block=block)
                        assert H.compare_hits(g, ref) == dict(prim=0, t=0, uv=0), (variant, wps, refill, lb, tl,
                                                                                   block, len(rays))
            finally:
                ctx.set_trace_tuning(0, 0, 0)
    with pytest.raises(_lib.RgpuError):
        ctx.set_trace_tuning(17, 0)
    with pytest.raises(_lib.RgpuError):
        ctx.set_trace_tuning(0, 0, 65)
    with pytest.raises(_lib.RgpuError):
        ctx.set_trace_tuning(0, 0, 0, 1025)
    # a tree smaller than the treelet: every record comes from LDS
    v, f = scenes.displaced_icosphere(level=1)
    small = _lib.Scene(ctx, _lib.BVH(v, f))
    rays = scenes.incoherent_rays(5000, v.min(0), v.max(0), seed=25)
    ref = _lib.trace(ctx, small, rays, "stackless", block=256)
    try:
        ctx.set_trace_tuning(0, 0, 0, 1024)
        assert H.compare_hits(_lib.trace(ctx, small, rays, "stackless", block=0), ref) == dict(prim=0, t=0, uv=0)
    finally:
        ctx.set_trace_tuning(0, 0, 0, 0)


def test_nan_guarded_slabs_in_mixed_waves(ctx):
    # Waves take the slab test without its NaN substitution only when no active lane can form a
    # NaN product. Rays that can (zero direction components from box-plane origins, a NaN origin)
    # are spread one in seven among ordinary rays, so most waves hold both kinds; a scene with
    # an infinite vertex has non-finite boxes and always keeps the substitution.
    v, f = scenes.displaced_icosphere(level=4)
    lo, hi = v.min(0), v.max(0)
    rays = scenes.incoherent_rays(14000, lo, hi, seed=31)
    deg = H.degenerate_rays(v, 2000, 5)
    rays[::7] = deg[:len(rays[::7])]
    rays["ro"][3::97] = np.nan
    bad_v = np.concatenate([v, np.array([[np.inf, 0.0, 0.0]], np.float32)])
    bad_f = np.concatenate([f, np.array([[0, 1, len(v)]], np.uint32)])
    for name, vv, ff in (("finite", v, f), ("infinite vertex", bad_v, bad_f)):
        bvh = _lib.BVH(vv, ff)
        scene = _lib.Scene(ctx, bvh)
        tree = bvh.export()
        for variant in VARIANTS:
            o, _ = _oracle(tree, vv, ff, None, rays, variant)
            for block in (0, 64, 256):
                print(f"case {name}/{variant}/block {block}", flush=True)
                g = _lib.trace(ctx, scene, rays, variant, block=block)
                assert H.compare_hits(g, o) == dict(prim=0, t=0, uv=0), (name, variant, block)
        scene.close()


def test_empty_and_single_triangle(ctx):
    v = np.zeros((0, 3), np.float32)
    f = np.zeros((0, 3), np.uint32)
    bvh = _lib.BVH(v, f)
    scene = _lib.Scene(ctx, bvh)
    rays = scenes.incoherent_rays(1000, [-1, -1, -1], [1, 1, 1], seed=1)
    for variant in VARIANTS:
        g = _lib.trace(ctx, scene, rays, variant)
        assert np.all(g["prim"] == -1) and np.array_equal(g["t"], rays["tmax"])
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.array([[0, 1, 2]], np.uint32)
    bvh = _lib.BVH(v, f)
    scene = _lib.Scene(ctx, bvh)
    rays = scenes.make_rays(np.array([[0.2, 0.2, 1.0], [2, 2, 1.0], [0.1, 0.1, -1.0]]),
                            np.array([[0, 0, -1.0], [0, 0, -1.0], [0, 0, 1.0]]))
    for variant in VARIANTS:
        g = _lib.trace(ctx, scene, rays, variant)
        assert g["prim"].tolist() == [0, -1, 0]
        assert g["t"][0] == 1.0 and g["t"][2] == 1.0


def test_invalid_arguments_return_status(ctx):
    v, f = scenes.displaced_icosphere(level=2)
    scene = _lib.Scene(ctx, _lib.BVH(v, f))
    rays = ctx.to_device(scenes.incoherent_rays(64, [-1] * 3, [1] * 3))
    hits = ctx.alloc(64 * 16)
    with pytest.raises(_lib.RgpuError) as e:
        scene.extension_ray_cast(7, rays, hits, 64)
    assert e.value.code == -1
    with pytest.raises(_lib.RgpuError) as e:
        scene.extension_ray_cast("stackless", rays, hits, 64, block=100)
    assert e.value.code == -1


@pytest.fixture(scope="module")
def c2():
    d = scenes.raycast_benchmark_scene(level=8)
    d["bvh"] = _lib.BVH(d["points"], d["indices"])
    d["tree"] = d["bvh"].export()
    return d


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("rayset", ["coherent", "incoherent"])
def test_bit_exact_raycast_benchmark_full_size(ctx, c2, variant, rayset):
    """1,310,720 triangles, 4,194,304 rays per set: GPU == oracle bit for bit."""
    v, f = c2["points"], c2["indices"]
    assert len(f) == 1310720
    scene = _lib.Scene(ctx, c2["bvh"])
    rays = c2[rayset]
    assert len(rays) == 4194304
    print(f"case c2/{rayset}/{variant}", flush=True)
    g = _lib.trace(ctx, scene, rays, variant)
    o, cnt = _oracle(c2["tree"], v, f, None, rays, variant)
    assert H.compare_hits(g, o) == dict(prim=0, t=0, uv=0)
    # the instrumented kernel (whose counts price the roofline) counts the oracle's events
    rd = ctx.to_device(np.ascontiguousarray(rays))
    hd = ctx.alloc(len(rays) * 16)
    gcnt = scene.extension_ray_cast_stats(variant, rd, hd, len(rays))
    gh = hd.download(_lib.HIT_DTYPE, (len(rays),))
    print(f"counters gpu {gcnt} oracle {cnt}", flush=True)
    assert gcnt == cnt
    assert (cnt["branch_visits"] > 0) if variant == "stackless" else (cnt["node_visits"] > 0)
    assert H.compare_hits(gh, o) == dict(prim=0, t=0, uv=0)
    rd.free()
    hd.free()
    # the oracle itself against brute force on a sample
    idx = np.random.default_rng(0).choice(len(rays), 256, replace=False)
    bf = O.intersect_bruteforce(v, f, rays[idx])
    ties, wrong = H.tie_excepted_mismatch(o[idx], bf)
    assert wrong == 0 and ties <= 2
    scene.close()
