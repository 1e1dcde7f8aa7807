"""Builders shared by the wavefront tests: the same scene on the GPU and in the oracle."""
import numpy as np

import oracle as O
from raccoongpurender_amd import _lib


def camera(c):
    return _lib.camera_look_at(c["eye"], c["target"], c["up"], c["fov"], c["width"], c["height"],
                               c.get("lens_radius", 0.0))


def oracle_lane(sd, n, **kw):
    bvh = _lib.BVH(sd["points"], sd["indices"])
    lane = O.OracleWavefront(sd["points"], sd["indices"], bvh.export(), sd["materials"], sd["env"],
                             camera(sd["camera"]), n, **kw)
    lane.initialize()
    return lane


def gpu_lane(ctx, sd, n, **kw):
    bvh = _lib.BVH(sd["points"], sd["indices"])
    scene = _lib.Scene(ctx, bvh)
    mats = _lib.Materials(ctx, sd["materials"])
    env = _lib.Envmap(ctx, sd["env"])
    quirks = kw.pop("quirks", True)
    flags = kw.pop("flags", 0) | (0 if quirks else _lib.WF_NO_REFERENCE_QUIRKS)
    if kw.pop("consistent", False):
        flags |= _lib.WF_CONSISTENT_NORMALS
    wf = _lib.Wavefront(ctx, scene, mats, env, n, flags=flags, **kw)
    wf.initialize(camera(sd["camera"]))
    wf._keep = (bvh, scene, mats, env)
    return wf


def mean_image(lane):
    rgb, cnt = lane.accumulation()
    with np.errstate(invalid="ignore", divide="ignore"):
        m = rgb / cnt[:, None]
    return m, cnt


def state_mismatches(a, b):
    """Per-field counts of slots whose state differs bit for bit."""
    out = {}
    out["rng"] = int(np.any(a["rng"] != b["rng"], axis=1).sum())
    for k in ("ro", "rd"):
        out["ray_" + k] = int(np.any(a["rays"][k].view(np.uint32) != b["rays"][k].view(np.uint32), axis=1).sum())
    out["pixel"] = int((a["rays"]["pixel"] != b["rays"]["pixel"]).sum())
    for k in ("T", "L"):
        out["path_" + k] = int(np.any(a["paths"][k].view(np.uint32) != b["paths"][k].view(np.uint32), axis=1).sum())
    out["logic_i"] = int((a["paths"]["logic_i"] != b["paths"]["logic_i"]).sum())
    ea, eb = a["ext"], b["ext"]
    out["hit_surface"] = int((ea["hit_surface"] != eb["hit_surface"]).sum())
    out["hit_volume"] = int((ea["hit_volume"] != eb["hit_volume"]).sum())
    out["tmin"] = int((ea["tmin"].view(np.uint32) != eb["tmin"].view(np.uint32)).sum())
    hit = ea["hit_surface"] >= 0
    out["Ng"] = int(np.any(ea["Ng"].view(np.uint32) != eb["Ng"].view(np.uint32), axis=1)[hit].sum())
    out["shade"] = int((np.any(a["shade"]["Le"].view(np.uint32) != b["shade"]["Le"].view(np.uint32), axis=1)
                        | np.any(a["shade"]["T"].view(np.uint32) != b["shade"]["T"].view(np.uint32), axis=1)).sum())
    va, vb = a["vol"], b["vol"]
    cnt_ne = va["count"] != vb["count"]
    lists_ne = np.array([not np.array_equal(x["volumes"][:x["count"]], y["volumes"][:y["count"]])
                         for x, y in zip(va, vb)]) if len(va) else np.zeros(0, bool)
    out["vol"] = int((cnt_ne | lists_ne).sum())
    return out
