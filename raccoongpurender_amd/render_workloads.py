"""The full-render workload: displaced icospheres with random transforms and materials,
a ground plane, a homogeneous-medium box and a synthetic sky with a sharp sun.

64 objects: 24 icospheres subdivided 6 times (81,920 triangles each) and 40 subdivided 3
times (1,280 each), every one radially displaced by 4-octave value noise (amplitude 0.15,
its own seed), plus a 2-triangle ground plane and a 12-triangle 1 x 1 x 1 medium box:
2,017,294 triangles, rendered to depth 8. Materials are drawn per object, uniformly among lambertian (albedo
U[0.1, 0.9]), anisotropic Ward (alpha_x, alpha_y U[0.05, 0.5]), mirror and dielectric
(IOR 1.5). The medium has sigma_s = 0.8, sigma_a = 0.1, i.e. C = sigma_t = 0.9 and
R = sigma_s / sigma_t. The 2048 x 1024 float environment is a sky gradient plus a Gaussian
sun lobe of peak 5e4 and 0.5 degree width.
"""
import numpy as np

from . import scenes
from ._lib import MAT_DIELECTRIC, MAT_LAMBERTIAN, MAT_SPECULAR, MAT_VOLUME, MAT_WARD, MaterialSet
from .render_scenes import MeshBuilder, box, quad, sky_envmap


def _rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def full_render_scene(width=1920, height=1080, seed=5, n_large=24, n_small=40, large_level=6, small_level=3,
                      env_size=(2048, 1024), sun_peak=5.0e4, sun_width_deg=0.5, n_mid=0, mid_level=5):
    rng = np.random.default_rng(seed)
    mb = MeshBuilder()
    lam = [((0, 0, 0), (0.5, 0.5, 0.5))]
    wards, diel = [], [1.5]
    g = quad((-12, 0, 12), (12, 0, 12), (12, 0, -12), (-12, 0, -12))
    mb.add(g[0], g[1], MAT_LAMBERTIAN, 0)
    base = {lv: scenes.icosphere(lv) for lv in {large_level, mid_level, small_level}}
    levels = [large_level] * n_large + [mid_level] * n_mid + [small_level] * n_small
    rng.shuffle(levels)
    for i, lv in enumerate(levels):
        v, f = base[lv]
        r = 1.0 + 0.15 * scenes.fbm(v, octaves=4, seed=7 + i)
        v = v * r[:, None]
        s = rng.uniform(0.4, 1.2)
        c = np.array([rng.uniform(-8, 8), rng.uniform(0.5, 2.5), rng.uniform(-8, 8)])
        v = (v @ _rotation(rng).T) * s + c
        kind = rng.integers(0, 4)
        if kind == 0:
            lam.append(((0, 0, 0), tuple(rng.uniform(0.1, 0.9, 3))))
            mt, mi = MAT_LAMBERTIAN, len(lam) - 1
        elif kind == 1:
            wards.append((rng.uniform(0.05, 0.5), rng.uniform(0.05, 0.5), tuple(rng.uniform(0.1, 0.9, 3)),
                          (1.0, 1.0, 1.0), 0.5))
            mt, mi = MAT_WARD, len(wards) - 1
        elif kind == 2:
            mt, mi = MAT_SPECULAR, 0
        else:
            mt, mi = MAT_DIELECTRIC, 0
        mb.add(v, f[:, [0, 2, 1]], mt, mi)
    bv, bf = box((-0.5, 0.0, 3.0), (0.5, 1.0, 4.0))
    mb.add(bv, bf, MAT_VOLUME, 0)
    v, f, types, idx = mb.build()
    ms = MaterialSet(types, idx, lambertians=lam, wards=wards, dielectrics=diel,
                     volumes=[(0.9, (0.8 / 0.9, 0.8 / 0.9, 0.8 / 0.9))])
    env = sky_envmap(env_size[0], env_size[1], sun_peak=sun_peak, sun_width_deg=sun_width_deg)
    cam = dict(eye=(0.0, 4.0, 14.0), target=(0.0, 0.8, 0.0), up=(0.0, 1.0, 0.0), fov=45.0, width=width,
               height=height, lens_radius=0.0045)
    # the workload renders at most 8 path vertices (depth 8)
    return dict(points=v, indices=f, materials=ms, env=env, camera=cam, max_depth=8)


# ---------------------------------------------------------------- batched multi-job workload
BATCH_SPP = (1, 4, 16, 64, 256)


def batch_scene_variants(count=8, lo=100_000, hi=2_000_000, seed=99):
    """`count` variants of the full-render scene with triangle counts spaced geometrically from
    about `lo` to `hi`: n_large level-6 (81,920 triangles) and n_mid level-5 (20,480) icospheres
    with the 40 small ones, ground and medium box; each variant its own placement and materials."""
    out = []
    small = 40 * 1280 + 2 + 12
    for k, target in enumerate(np.geomspace(lo, hi, count)):
        rest = max(0.0, target - small)
        n_large = int(rest // 81920)
        n_mid = int(round((rest - n_large * 81920) / 20480))
        out.append(dict(n_large=n_large, n_mid=n_mid, seed=seed * 100 + k,
                        triangles=n_large * 81920 + n_mid * 20480 + small))
    return out


def batch_jobs(count=256, variants=8, seed=99, res_lo=64, res_hi=1024):
    """Render jobs of the batched workload: scene variant uniform over `variants`, width and
    height each drawn log-uniformly from [res_lo, res_hi], spp uniform over BATCH_SPP."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(count):
        v = int(rng.integers(0, variants))
        w = int(round(np.exp(rng.uniform(np.log(res_lo), np.log(res_hi)))))
        h = int(round(np.exp(rng.uniform(np.log(res_lo), np.log(res_hi)))))
        spp = int(BATCH_SPP[int(rng.integers(0, len(BATCH_SPP)))])
        jobs.append(dict(variant=v, width=w, height=h, spp=spp, seed=int(rng.integers(0, 2**62))))
    return jobs


def batch_slots(jobs, total=1 << 22, minimum=256):
    """Path slots per job, proportional to its sample budget (width * height * spp), at least
    `minimum`, multiples of 256, summing to about `total`."""
    work = np.array([j["width"] * j["height"] * j["spp"] for j in jobs], np.float64)
    raw = np.maximum(minimum, total * work / work.sum())
    slots = (np.ceil(raw / 256) * 256).astype(np.int64)
    cap = np.array([max(minimum, (j["width"] * j["height"] * j["spp"] + 255) // 256 * 256) for j in jobs])
    return [int(x) for x in np.minimum(slots, cap)]


def device_batch_jobs(ctx, total_slots=1 << 22, count=256, variants=8, seed=99):
    """The batched workload on the device: 8 scene variants (BVH, scene, materials) sharing one
    envmap, and the 256 jobs (camera of the variant at the job's resolution, spp, seed, slots).
    Returns (jobs, specs, variant descriptions); each job keeps its device objects alive."""
    from . import _lib

    vdesc = batch_scene_variants(variants, seed=seed)
    specs = batch_jobs(count, variants, seed=seed)
    slots = batch_slots(specs, total_slots)
    dev = []
    env = None
    for v in vdesc:
        sd = full_render_scene(64, 64, seed=v["seed"], n_large=v["n_large"], n_mid=v["n_mid"])
        bvh = _lib.BVH(sd["points"], sd["indices"])
        if env is None:
            env = _lib.Envmap(ctx, sd["env"])
        dev.append(dict(bvh=bvh, scene=_lib.Scene(ctx, bvh), mats=_lib.Materials(ctx, sd["materials"]),
                        cam=sd["camera"], env=env))
    jobs = []
    for spec, n in zip(specs, slots):
        d = dev[spec["variant"]]
        c = d["cam"]
        cam = _lib.camera_look_at(c["eye"], c["target"], c["up"], c["fov"], spec["width"], spec["height"],
                                  c["lens_radius"])
        j = _lib.BatchJob(d["scene"], d["mats"], env, cam, n, spp=spec["spp"], seed=spec["seed"])
        j.keep = d
        jobs.append(j)
    return jobs, specs, vdesc
