"""Headless renderer: python -m raccoongpurender_amd.render --scene cornell --steps 200 --out img.png

The counterpart of the reference's Render app (Render/src/ofApp.cpp:25-134), which renders
until a time budget and writes a PNG; here the budget is a step count or seconds.
"""
import argparse
import time

import numpy as np

from . import _lib, render_scenes, render_workloads, scenes
from .png import write_png

SCENES = {
    "cornell": lambda w, h: render_scenes.cornell_box(w or 256, h or 256),
    "materials": lambda w, h: render_scenes.material_test_scene(w or 320, h or 240, env_size=(256, 128)),
    "full": lambda w, h: render_workloads.full_render_scene(w or 1920, h or 1080),
}


def build_lane(ctx, sd, path_count, **kw):
    """Device scene + wavefront for a scene dict; the scene's max_depth is the default."""
    kw.setdefault("max_depth", sd.get("max_depth", 0))
    bvh = _lib.BVH(sd["points"], sd["indices"])
    scene = _lib.Scene(ctx, bvh)
    mats = _lib.Materials(ctx, sd["materials"])
    env = _lib.Envmap(ctx, sd["env"])
    wf = _lib.Wavefront(ctx, scene, mats, env, path_count, **kw)
    c = sd["camera"]
    wf.initialize(_lib.camera_look_at(c["eye"], c["target"], c["up"], c["fov"], c["width"], c["height"],
                                      c.get("lens_radius", 0.0)))
    wf._keep = (bvh, scene, mats, env)
    return wf


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scene", choices=sorted(SCENES), default="cornell")
    ap.add_argument("--obj", default="", help="render this Wavefront OBJ mesh on a ground plane under a sky")
    ap.add_argument("--width", type=int, default=0)
    ap.add_argument("--height", type=int, default=0)
    ap.add_argument("--paths", type=int, default=1 << 22)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--seconds", type=float, default=0.0)
    ap.add_argument("--variant", default="stackless")
    ap.add_argument("--trace-block", type=int, default=256, help="0: persistent trace; 64/128/256: one thread per slot")
    ap.add_argument("--queue-blocks", type=int, default=0, help="workgroups of the queue kernels (0: default)")
    ap.add_argument("--out", default="render.png")
    a = ap.parse_args()
    if a.obj:
        sd = render_scenes.mesh_scene(*scenes.load_obj(a.obj), a.width or 320, a.height or 240)
    else:
        sd = SCENES[a.scene](a.width, a.height)
    ctx = _lib.Context(0)
    wf = build_lane(ctx, sd, a.paths, variant=a.variant)
    if a.trace_block != 256 or a.queue_blocks:
        wf.set_launch(a.variant, a.trace_block, a.queue_blocks or 1024)
        c = sd["camera"]
        wf.initialize(_lib.camera_look_at(c["eye"], c["target"], c["up"], c["fov"], c["width"], c["height"],
                                          c.get("lens_radius", 0.0)))
    t0 = time.perf_counter()
    steps = 0
    while True:
        wf.step(10)
        steps += 10
        ctx.synchronize()
        el = time.perf_counter() - t0
        if (a.seconds and el >= a.seconds) or (not a.seconds and steps >= a.steps):
            break
    total = wf.stat()
    w, h = sd["camera"]["width"], sd["camera"]["height"]
    write_png(a.out, wf.tonemap(), w, h)
    print(f"{steps} steps in {el:.2f} s, {total} samples ({total / (w * h):.1f} spp), "
          f"{total / el / 1e6:.2f} Msamples/s -> {a.out}")


if __name__ == "__main__":
    main()
