"""Per-stage timing and roofline of the wavefront bounce on the full-render workload.

Each stage of WavefrontLane::step (common/wavefront_path_tracing.hpp:606-875) is timed in-band with
HIP events between its launches (rgpu_wavefront_time_stages), and priced with algorithmic bytes:
the per-slot records the stage's kernels read and write for the items in its queue (the record
sizes of the device layouts: ray 32 B, path 32 B, extension result 32 B, shading result 32 B,
medium list 32 B, RNG state 16 B, queue item 4 B), plus, for the extension ray cast, the node
and primitive bytes of the traversal itself, counted by the instrumented ray cast on the rays the
bounce actually traced. Achieved = bytes / stage time, against the 8 TB/s HBM peak.

    python -m raccoongpurender_amd.stage_profile [--steps 16] [--guided] [--trace-block 64]

prints one JSON object. The same run under rocprofv3 gives the per-kernel durations and HBM
traffic the stage figures are checked against.
"""
import argparse
import json
import sys

import numpy as np

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per queue item (new path, shading stages) or per slot (ray cast, logic,
# queues), from the records each kernel loads and stores.
BYTES_PER_ITEM = {
    # item 4, RNG 16 in + 16 out, ray 32, path 32, extension 32 in + 32 out, shading 32,
    # medium count 4
    "new_path": 4 + 32 + 32 + 32 + 64 + 32 + 4,
    # item 4, RNG 32, extension 32, ray 32 in + 32 out, shading 32, material 8 + 32, one envmap
    # texel record 64 + alias bucket 8
    "lambertian": 4 + 32 + 32 + 64 + 32 + 40 + 72,
    # as lambertian with the 48-byte Ward record
    "ward": 4 + 32 + 32 + 64 + 32 + 56 + 72,
    # item 4, RNG 32 (dielectric), extension 32, ray 32 in + 32 out, shading 32, material 8 + 4
    "mirror_dielectric": 4 + 32 + 32 + 64 + 32 + 12,
    # item 4, RNG 32, extension 32, ray 32 in + 32 out, medium list 32 in + 32 out, shading 32,
    # material 8 + 16
    "medium": 4 + 32 + 32 + 64 + 64 + 32 + 24,
    # per slot besides the traversal: medium list 32, RNG 32, ray 32, extension result 32 out
    "ray_cast": 32 + 32 + 32 + 32,
    # per slot: path 32 in + 32 out, extension 32, shading 32, ray 32, RNG 32, queue choice 1
    "logic": 64 + 32 + 32 + 32 + 32 + 1,
    # per slot: queue choice 1, queue item 4 (block counts are negligible)
    "queues": 5,
}
# what each stage's item count is: a queue (index into the 7 queue lengths), or every slot
QUEUE_OF = {"new_path": [0], "lambertian": [1], "ward": [4], "mirror_dielectric": [2, 3], "medium": [5, 6]}


def traversal_bytes_per_ray(ctx, scene, st, variant, sample=1 << 20):
    """Node and primitive bytes per ray of the extension rays a bounce traced (stackless: 64 B per
    branch visit, 16 B per leaf visit and backtrack step; threaded: 32 B per visit; 48 B per
    primitive test), from the instrumented ray cast on a sample of those rays."""
    from . import _lib

    rays = st["rays"]
    ext = st["ext"]
    n = min(sample, len(rays))
    pick = np.linspace(0, len(rays) - 1, n).astype(np.int64)
    r = np.zeros(n, _lib.RAY_DTYPE)
    r["ro"] = rays["ro"][pick]
    r["rd"] = rays["rd"][pick]
    # rays that scattered in a medium were cast with the free-flight distance as far clip
    medium = (ext["hit_surface"][pick] < 0) & (ext["hit_volume"][pick] >= 0)
    r["tmax"] = np.where(medium, ext["tmin"][pick], np.float32(3.4028235e38))
    rd = ctx.to_device(r)
    hd = ctx.alloc(n * 16)
    c = scene.extension_ray_cast_stats(variant, rd, hd, n)
    rd.free()
    hd.free()
    if variant == "stackless":
        b = 64 * c["branch_visits"] + 16 * c["leaf_visits"] + 16 * c["parent_steps"]
    else:
        b = 32 * c["node_visits"]
    return (b + 48 * c["prim_tests"]) / n, {k: v / n for k, v in c.items()}


def profile(steps=16, warmup=16, guided=False, trace_block=64, queue_blocks=4096, variant="stackless", paths=1 << 22,
            ctx=None, train_every=4):
    from . import _lib, render_workloads
    from .render import build_lane

    own = ctx is None
    ctx = ctx or _lib.Context(0)
    sd = render_workloads.full_render_scene()
    wf = build_lane(ctx, sd, paths, variant=variant)
    wf.set_launch(variant, trace_block, queue_blocks)
    c = sd["camera"]
    wf.initialize(_lib.camera_look_at(c["eye"], c["target"], c["up"], c["fov"], c["width"], c["height"],
                                      c["lens_radius"]))
    gd = None
    if guided:
        from . import guide as G

        lo = sd["points"].min(0) - 1e-3
        hi = sd["points"].max(0) + 1e-3
        gd = G.Guide(ctx, lo, hi, capacity=1 << 18, minibatch=16384, seed=1)
        G.attach(wf, gd, sampling=True, training=True)
    for k in range(warmup):
        wf.step(1)
        if gd is not None and k % train_every == train_every - 1:
            G.train_pass(wf, 1)
    ctx.synchronize()
    out = profile_lane(ctx, wf, wf._keep[1], steps, variant, trace_block, queue_blocks, paths,
                       (lambda: G.train_pass(wf, 1)) if gd is not None else None, train_every)
    out["workload"] = ("1920x1080, depth 8, 2017294 tris, every material, medium, 2048x1024 sun sky; "
                       f"{paths} slots, {'guided' if guided else 'plain'}")
    if gd is not None:
        G.attach(wf, None)
        gd.close()
    wf.close()
    if own:
        ctx.close()
    return out


def profile_lane(ctx, wf, scene, steps, variant, trace_block, queue_blocks, paths, train=None, train_every=4):
    """Stage times and bytes of `steps` bounces of an initialised, warmed-up lane (train: called
    every `train_every` bounces, outside the timed stages)."""
    from . import _lib

    names = _lib.STAGE_NAMES
    ms = {k: 0.0 for k in names}
    items = {k: 0 for k in names}
    for k in range(steps):
        counts_before, _ = wf.queue_counts()  # the queues this bounce consumes
        t = wf.time_stages(1)
        for name in names:
            ms[name] += t[name]
        for name, qs in QUEUE_OF.items():
            items[name] += int(sum(int(counts_before[q]) for q in qs))
        for name in ("ray_cast", "logic", "queues"):
            items[name] += paths
        if train is not None and k % train_every == train_every - 1:
            train()
    st = wf.state()
    tb_ray, counts = traversal_bytes_per_ray(ctx, scene, st, variant)
    out = {"launch": [variant, trace_block, queue_blocks], "steps": steps, "stages": {}}
    total_ms = 0.0
    for name in names:
        t_ms = ms[name] / steps
        total_ms += t_ms
        per = BYTES_PER_ITEM.get(name)
        if name == "ray_cast":
            per_item = per + tb_ray
        else:
            per_item = per
        row = {"ms": round(t_ms, 4)}
        if per_item is not None and items[name]:
            b = per_item * items[name] / steps
            row.update(items=int(items[name] / steps), bytes_per_item=round(per_item, 1), bytes=int(b),
                       gbs=round(b / (t_ms * 1e-3) / 1e9, 1) if t_ms > 0 else None)
            row["frac"] = round(row["gbs"] / HBM_PEAK_GBS, 4) if row["gbs"] else None
        out["stages"][name] = row
    out["ms_per_bounce"] = round(total_ms, 4)
    out["traversal"] = {"bytes_per_ray": round(tb_ray, 1), "per_ray": {k: round(v, 3) for k, v in counts.items()}}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--guided", action="store_true")
    ap.add_argument("--trace-block", type=int, default=64, help="0: persistent ray cast; 64/128/256: one thread per slot")
    ap.add_argument("--queue-blocks", type=int, default=4096)
    ap.add_argument("--variant", choices=["stackless", "threaded", "threaded_single"], default="stackless")
    a = ap.parse_args()
    out = profile(a.steps, a.warmup, a.guided, a.trace_block, a.queue_blocks, a.variant)
    for name, row in out["stages"].items():
        print(f"[stage] {name:18s} {row['ms']:8.4f} ms  {row.get('gbs') or 0:8.1f} GB/s", file=sys.stderr)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
