"""ORACLE — test infrastructure only.

ctypes front end of the CPU restatement in oracle/*.cpp (built by oracle/Makefile into
oracle/_build/liboracle.so). Imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker; the product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])
    return LIB_PATH


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    return C.CDLL(LIB_PATH)


lib = _load()
vp = C.c_void_p
u32 = C.c_uint32
u64 = C.c_uint64


def _sig(name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


_sig("orc_splitmix64_seq", None, u64, u32, vp)
_sig("orc_random_initialize", None, u64, u32, vp)
_sig("orc_xoshiro_next", None, vp, u32, vp)
_sig("orc_xoshiro_uniform", None, vp, u32, vp)
_sig("orc_xoshiro_uniform_integer", None, vp, u32, vp)
_sig("orc_xoshiro_jump", None, vp)
_sig("orc_slabs_with_hit", C.c_int, vp, vp, vp, vp, C.c_float, vp)
_sig("orc_intersect_triangle", C.c_int, vp, vp, vp, vp, vp, vp)
_sig("orc_stackless_build", C.c_int, u32, vp, vp, vp, vp, vp, vp, vp, vp)
_sig("orc_stackless_intersect", C.c_int, vp, vp, vp, u32, vp, C.c_int32, vp, u32, vp, vp, vp, C.c_int)
_sig("orc_mtbvh_build", C.c_int, u32, vp, vp, vp, vp, vp, vp, vp, vp, vp)
_sig("orc_mtbvh_intersect", C.c_int, vp, vp, u32, vp, vp, u32, vp, C.c_int32, vp, u32, vp, vp, vp, C.c_int)
_sig("orc_tbvh_intersect", C.c_int, vp, vp, u32, vp, vp, u32, vp, C.c_int32, vp, u32, vp, vp, vp, C.c_int)
_sig("orc_bruteforce_intersect", C.c_int, u32, vp, u32, vp, C.c_int32, vp, u32, vp, vp, C.c_int)
_sig("orc_hit_normals", None, u32, vp, vp, vp, u32, vp, C.c_int32, vp, vp)


def p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


HIT_DTYPE = np.dtype([("prim", np.int32), ("t", np.float32), ("u", np.float32), ("v", np.float32)])
REF_SL_NODE = np.dtype([
    ("L_lower", np.float32, 4), ("L_upper", np.float32, 4),
    ("R_lower", np.float32, 4), ("R_upper", np.float32, 4),
    ("link_parent", np.uint32), ("link_L", np.uint32), ("link_R", np.uint32), ("link_sibling", np.uint32),
    ("primitive_indices_beg", np.uint32), ("primitive_indices_end", np.uint32), ("pad", np.uint32, 2),
])
REF_MT_NODE = np.dtype([
    ("lower", np.float32, 4), ("upper", np.float32, 4),
    ("primitive_indices_beg", np.int32), ("primitive_indices_end", np.int32), ("pad", np.int32, 2),
])
assert REF_SL_NODE.itemsize == 96 and REF_MT_NODE.itemsize == 48
COUNTER_NAMES = ("branch_visits", "leaf_visits", "parent_steps", "prim_tests", "node_visits")
# Worker threads of the restatement: the CPUs this process may use, at most 16 (the host's CPU
# share for one GPU, whatever nproc shows there).
try:
    NTHREADS = max(1, min(16, len(os.sched_getaffinity(0))))
except (AttributeError, OSError):
    NTHREADS = max(1, min(16, os.cpu_count() or 1))


# ------------------------------------------------------------------ RNG
def splitmix64_seq(seed, n):
    out = np.zeros(n, np.uint64)
    lib.orc_splitmix64_seq(seed, n, p(out))
    return out


def random_initialize(seed_offset, n):
    s = np.zeros((n, 4), np.uint32)
    lib.orc_random_initialize(seed_offset, n, p(s))
    return s


def xoshiro_next(state, n):
    out = np.zeros(n, np.uint32)
    lib.orc_xoshiro_next(p(state), n, p(out))
    return out


def xoshiro_uniform(state, n):
    out = np.zeros(n, np.float32)
    lib.orc_xoshiro_uniform(p(state), n, p(out))
    return out


def xoshiro_uniform_integer(state, n):
    out = np.zeros(n, np.uint64)
    lib.orc_xoshiro_uniform_integer(p(state), n, p(out))
    return out


def xoshiro_jump(state):
    lib.orc_xoshiro_jump(p(state))


# ------------------------------------------------------------------ geometry
def slabs_with_hit(p0, p1, ro, rd, farclip):
    a = [np.ascontiguousarray(x, np.float32) for x in (p0, p1, ro, rd)]
    h = np.zeros(2, np.float32)
    r = lib.orc_slabs_with_hit(p(a[0]), p(a[1]), p(a[2]), p(a[3]), np.float32(farclip), p(h))
    return bool(r), float(h[0]), float(h[1])


def intersect_triangle(ro, rd, v0, v1, v2, tmin):
    a = [np.ascontiguousarray(x, np.float32) for x in (ro, rd, v0, v1, v2)]
    tuv = np.array([tmin, 0, 0], np.float32)
    r = lib.orc_intersect_triangle(*[p(x) for x in a], p(tuv))
    return bool(r), tuv


class RefStackless:
    """Reference 96-byte stackless layout built from a tree (common/stackless_bvh.hpp:283-341)."""

    def __init__(self, tree):
        nn = len(tree["left"])
        self.nodes = np.zeros(nn, REF_SL_NODE)
        self.prim_ids = np.zeros(max(len(tree["prim_ids"]), 1), np.uint32)
        args = [np.ascontiguousarray(tree[k]) for k in ("left", "right", "bounds", "prim_beg", "prim_cnt", "prim_ids")]
        r = lib.orc_stackless_build(nn, *[p(a) for a in args], p(self.nodes), p(self.prim_ids))
        if r < 0:
            raise RuntimeError("orc_stackless_build failed")


class RefMTBVH:
    """Reference threaded layout: MTBVHNode + links[order][node][hit, miss]: orders 0..5 the
    multi-threaded BVH's direction classes (common/threaded_bvh.hpp:398-525), order 6 the
    single-table threaded BVH (common/threaded_bvh.hpp:269-364)."""

    def __init__(self, tree):
        nn = len(tree["left"])
        self.nn = nn
        self.nodes = np.zeros(nn, REF_MT_NODE)
        self.links = np.zeros((7, nn, 2), np.int32)
        self.prim_ids = np.zeros(max(len(tree["prim_ids"]), 1), np.uint32)
        args = [np.ascontiguousarray(tree[k]) for k in ("left", "right", "bounds", "prim_beg", "prim_cnt", "prim_ids")]
        r = lib.orc_mtbvh_build(nn, *[p(a) for a in args], p(self.nodes), p(self.links), p(self.prim_ids))
        if r < 0:
            raise RuntimeError("orc_mtbvh_build failed")


def _geom(points, indices, spheres):
    points = np.ascontiguousarray(points, np.float32)
    indices = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    ntris = indices.size // 3
    spheres = None if spheres is None or len(spheres) == 0 else np.ascontiguousarray(spheres, np.float32)
    return points, indices, ntris, spheres


def intersect_stackless(ref, points, indices, rays, spheres=None, nthreads=NTHREADS):
    points, indices, ntris, spheres = _geom(points, indices, spheres)
    rays = np.ascontiguousarray(rays)
    hits = np.zeros(len(rays), HIT_DTYPE)
    cnt = np.zeros(5, np.uint64)
    lib.orc_stackless_intersect(p(ref.nodes), p(ref.prim_ids), p(points), points.shape[1], p(indices), ntris,
                                p(spheres), len(rays), p(rays), p(hits), p(cnt), nthreads)
    return hits, dict(zip(COUNTER_NAMES, (int(x) for x in cnt)))


def intersect_mtbvh(ref, points, indices, rays, spheres=None, nthreads=NTHREADS):
    points, indices, ntris, spheres = _geom(points, indices, spheres)
    rays = np.ascontiguousarray(rays)
    hits = np.zeros(len(rays), HIT_DTYPE)
    cnt = np.zeros(5, np.uint64)
    lib.orc_mtbvh_intersect(p(ref.nodes), p(ref.links), ref.nn, p(ref.prim_ids), p(points), points.shape[1],
                            p(indices), ntris, p(spheres), len(rays), p(rays), p(hits), p(cnt), nthreads)
    return hits, dict(zip(COUNTER_NAMES, (int(x) for x in cnt)))


def intersect_tbvh(ref, points, indices, rays, spheres=None, nthreads=NTHREADS):
    """The single-table threaded walk (common/threaded_bvh.hpp:366-395) over ref.links[6]."""
    points, indices, ntris, spheres = _geom(points, indices, spheres)
    rays = np.ascontiguousarray(rays)
    hits = np.zeros(len(rays), HIT_DTYPE)
    cnt = np.zeros(5, np.uint64)
    lib.orc_tbvh_intersect(p(ref.nodes), p(ref.links), ref.nn, p(ref.prim_ids), p(points), points.shape[1],
                           p(indices), ntris, p(spheres), len(rays), p(rays), p(hits), p(cnt), nthreads)
    return hits, dict(zip(COUNTER_NAMES, (int(x) for x in cnt)))


def intersect_bruteforce(points, indices, rays, spheres=None, nthreads=NTHREADS):
    points, indices, ntris, spheres = _geom(points, indices, spheres)
    rays = np.ascontiguousarray(rays)
    hits = np.zeros(len(rays), HIT_DTYPE)
    nprims = ntris + (0 if spheres is None else len(spheres))
    lib.orc_bruteforce_intersect(nprims, p(points), points.shape[1], p(indices), ntris, p(spheres), len(rays),
                                 p(rays), p(hits), nthreads)
    return hits


def hit_normals(rays, hits, points, indices, spheres=None):
    points, indices, ntris, spheres = _geom(points, indices, spheres)
    out = np.zeros((len(hits), 3), np.float32)
    lib.orc_hit_normals(len(hits), p(np.ascontiguousarray(rays)), p(np.ascontiguousarray(hits)), p(points),
                        points.shape[1], p(indices), ntris, p(spheres), p(out))
    return out


# ====================================================================== wavefront step (restated)
class orc_wf_desc(C.Structure):
    _fields_ = [
        ("n", u32), ("rr_start", u32), ("max_depth", u32), ("quirks", u32), ("variant", u32), ("nthreads", u32),
        ("clamp", C.c_float), ("bxdf_p", C.c_float), ("seed", u64),
        ("points", vp), ("stride", u32), ("indices", vp), ("sphere_begin", C.c_int32), ("spheres", vp),
        ("sl_nodes", vp), ("sl_prims", vp), ("mt_nodes", vp), ("mt_links", vp), ("mt_nn", u32), ("mt_prims", vp),
        ("nprims", u32), ("mat_types", vp), ("mat_indices", vp),
        ("lambertians", vp), ("nl", u32), ("wards", vp), ("nw", u32), ("iors", vp), ("nd", u32),
        ("volumes", vp), ("nv", u32),
        ("env_rgb", vp), ("env_channels", u32), ("env_w", u32), ("env_h", u32),
        ("camera", vp), ("width", u32), ("height", u32), ("lens_radius", C.c_float), ("consistent", u32),
    ]


_sig("orc_wf_create", vp, C.POINTER(orc_wf_desc))
_sig("orc_wf_destroy", None, vp)
_sig("orc_wf_initialize", None, vp, u32, u32)
_sig("orc_wf_step", None, vp, u32)
_sig("orc_wf_queue_counts", None, vp, vp, C.POINTER(u64))
_sig("orc_wf_read_accumulation", None, vp, vp, vp)
_sig("orc_wf_read_state", None, vp, vp, vp, vp, vp, vp, vp, vp)
_sig("orc_envmap_build", vp, vp, u32, u32, u32)
_sig("orc_envmap_export", None, vp, C.c_int, vp, vp, vp, vp)
_sig("orc_envmap_destroy", None, vp)
_sig("orc_envmap_lookup", None, vp, u32, vp, C.c_int, vp)
_sig("orc_envmap_sample", None, vp, vp, u32, vp, vp, vp)
_sig("orc_envmap_pdf", None, vp, u32, vp, vp, vp)
_sig("orc_envmap_sample_each", None, vp, vp, u32, vp, vp, vp)
_sig("orc_envmap_pdf_each", None, vp, u32, vp, vp, vp)

EXT_DTYPE = np.dtype([("hit_surface", np.int32), ("hit_volume", np.int32), ("tmin", np.float32), ("Ng", np.float32, 3),
                      ("pad", np.float32, 2)])
PATH_DTYPE = np.dtype([("T", np.float32, 3), ("logic_i", np.uint32), ("L", np.float32, 3), ("pad", np.uint32)])
SHADE_DTYPE = np.dtype([("Le", np.float32, 3), ("T", np.float32, 3), ("pad", np.float32, 2)])
VOL_DTYPE = np.dtype([("volumes", np.int32, 7), ("count", np.int32)])
STATE_RAY_DTYPE = np.dtype([("ro", np.float32, 3), ("tmax", np.float32), ("rd", np.float32, 3), ("pixel", np.uint32)])


class OracleEnvmap:
    def __init__(self, rgb):
        rgb = np.ascontiguousarray(rgb, np.float32)
        self.rgb = rgb
        self.height, self.width = rgb.shape[:2]
        self.h = lib.orc_envmap_build(p(rgb), rgb.shape[2], self.width, self.height)

    def export(self, table):
        P = self.width * self.height
        out = dict(pdf=np.zeros(P, np.float32), height=np.zeros(P, np.float32), alias=np.zeros(P, np.int32),
                   fragments=np.zeros((P, 4), np.float32))
        lib.orc_envmap_export(self.h, table, p(out["pdf"]), p(out["height"]), p(out["alias"]), p(out["fragments"]))
        return out

    def lookup(self, dirs, filtered):
        dirs = np.ascontiguousarray(dirs, np.float32)
        out = np.zeros((len(dirs), 3), np.float32)
        lib.orc_envmap_lookup(self.h, len(dirs), p(dirs), 1 if filtered else 0, p(out))
        return out

    def sample(self, state, n, normal=None):
        nrm = None if normal is None else np.ascontiguousarray(normal, np.float32)
        dirs = np.zeros((n, 3), np.float32)
        pdfs = np.zeros(n, np.float32)
        lib.orc_envmap_sample(self.h, p(state), n, p(nrm), p(dirs), p(pdfs))
        return dirs, pdfs

    def pdf(self, dirs, normal=None):
        dirs = np.ascontiguousarray(dirs, np.float32)
        nrm = None if normal is None else np.ascontiguousarray(normal, np.float32)
        out = np.zeros(len(dirs), np.float32)
        lib.orc_envmap_pdf(self.h, len(dirs), p(dirs), p(nrm), p(out))
        return out

    def sample_each(self, states, normals=None):
        """One sample per state row (n x 4 u32): (directions, pdfs, advanced states)."""
        st = np.array(states, np.uint32).reshape(-1, 4)
        n = len(st)
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(n, 3)
        dirs = np.zeros((n, 3), np.float32)
        pdfs = np.zeros(n, np.float32)
        lib.orc_envmap_sample_each(self.h, p(st), n, p(nrm), p(dirs), p(pdfs))
        return dirs, pdfs, st

    def texel_index(self, dirs):
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros(len(dirs), np.int32)
        lib.orc_envmap_index(self.h, len(dirs), p(dirs), p(out))
        return out

    def pdf_each(self, dirs, normals=None):
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(len(dirs), 3)
        out = np.zeros(len(dirs), np.float32)
        lib.orc_envmap_pdf_each(self.h, len(dirs), p(dirs), p(nrm), p(out))
        return out

    def __del__(self):
        try:
            lib.orc_envmap_destroy(self.h)
        except Exception:
            pass


_sig("orc_ward_sample", None, vp, u32, vp, vp, vp, vp)
_sig("orc_ward_eval", None, vp, u32, vp, vp, vp, vp, vp)
_sig("orc_camera_rays", None, vp, u32, u32, C.c_float, u32, vp, vp, vp, vp)
_sig("orc_free_flight", None, vp, u32, u32, vp, vp, vp, vp)
_sig("orc_envmap_index", None, vp, u32, vp, vp)
_sig("orc_math_f64", None, C.c_int, u32, vp, vp, vp)
_sig("orc_pow_f32", None, u32, vp, C.c_float, vp)
MATH_FUNCS = ("exp", "log", "sin", "cos", "atan", "atan2", "pow")


def math_f64(name, x, y=None):
    """The oracle's double-precision exp/log/sin/cos/atan/atan2/pow (oracle_math.h) over arrays."""
    x = np.ascontiguousarray(x, np.float64)
    yy = None if y is None else np.ascontiguousarray(np.broadcast_to(y, x.shape), np.float64)
    out = np.zeros_like(x)
    lib.orc_math_f64(MATH_FUNCS.index(name), x.size, p(x), p(yy), p(out))
    return out


def pow_f32(x, y):
    """(float) pow((double) x, (double) y) with the oracle's pow, elementwise over float32 x."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    lib.orc_pow_f32(x.size, p(x), np.float32(y), p(out))
    return out


def _f32(a, cols=None):
    a = np.ascontiguousarray(a, np.float32)
    return a if cols is None else a.reshape(-1, cols)


def _ward9(ax, ay, refl=(1.0, 1.0, 1.0), edge=None, falloff=1.0):
    return np.array([ax, ay, *refl, *(refl if edge is None else edge), falloff], np.float32)


def ward_sample(ward, wo, n, u):
    """The lane's Ward half-vector sampler (kernels/ward.cl:40-74 + the anisotropic extension)
    for rows of wo, normal and (u0, u1)."""
    wo, n, u = _f32(wo, 3), _f32(n, 3), _f32(u, 2)
    wi = np.zeros((len(wo), 3), np.float32)
    lib.orc_ward_sample(p(_f32(ward)), len(wo), p(wo), p(n), p(u), p(wi))
    return wi


def ward_eval(ward, wo, wi, n):
    """(pdf, f): the lane's Ward pdf (kernels/ward.cl:155-176) and BRDF (ward_stage)."""
    wo, wi, n = _f32(wo, 3), _f32(wi, 3), _f32(n, 3)
    pdf = np.zeros(len(wo), np.float32)
    f = np.zeros((len(wo), 3), np.float32)
    lib.orc_ward_eval(p(_f32(ward)), len(wo), p(wo), p(wi), p(n), p(pdf), p(f))
    return pdf, f


def camera_rays(camera, pixels, states):
    """new_path's primary rays (kernels/new_path.cl:44-71) for pixel indices, one state row each
    (advanced in place): (ro, rd)."""
    cam = camera.as_array()
    pixels = np.ascontiguousarray(pixels, np.uint32)
    assert states.dtype == np.uint32 and states.flags["C_CONTIGUOUS"] and states.shape == (len(pixels), 4)
    ro = np.zeros((len(pixels), 3), np.float32)
    rd = np.zeros((len(pixels), 3), np.float32)
    lib.orc_camera_rays(p(cam), camera.width, camera.height, np.float32(camera.lens_radius), len(pixels), p(pixels),
                        p(states), p(ro), p(rd))
    return ro, rd


def free_flight(C_by_volume, lists, counts, states):
    """homogeneous_volume_through (kernels/homogeneous_volume_through.cl:24-40) per row: medium
    ids (up to 7, indexing C_by_volume) and their count; states advanced in place.
    Returns (chosen medium or -1, distance)."""
    Cv = _f32(C_by_volume)
    n = len(counts)
    l8 = np.zeros((n, 8), np.int32)
    l8[:, :7] = np.asarray(lists, np.int32).reshape(n, 7)
    l8[:, 7] = counts
    assert states.dtype == np.uint32 and states.flags["C_CONTIGUOUS"] and states.shape == (n, 4)
    hv = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    lib.orc_free_flight(p(Cv), len(Cv), n, p(l8), p(states), p(hv), p(t))
    return hv, t


class OracleWavefront:
    """The reference's WavefrontLane step restated on the CPU (stable queues)."""

    def __init__(self, points, indices, tree, materials, env_rgb, camera, path_count, spheres=None, seed=0,
                 rr_start=0, max_depth=0, clamp=0.0, bxdf_p=0.0, quirks=True, variant="stackless",
                 nthreads=NTHREADS, consistent=False):
        self.points, self.indices, self.ntris, self.spheres = _geom(points, indices, spheres)
        self.keep = []
        d = orc_wf_desc()
        d.n, d.rr_start, d.max_depth, d.quirks = path_count, rr_start, max_depth, 1 if quirks else 0
        d.variant = {"stackless": 0, "threaded": 1, "threaded_single": 2}[variant]
        d.nthreads, d.clamp, d.bxdf_p, d.seed = nthreads, clamp, bxdf_p, seed
        d.points, d.stride, d.indices = p(self.points), self.points.shape[1], p(self.indices)
        d.sphere_begin, d.spheres = self.ntris, p(self.spheres)
        if len(tree["left"]):
            self.sl = RefStackless(tree)
            self.mt = RefMTBVH(tree)
            d.sl_nodes, d.sl_prims = p(self.sl.nodes), p(self.sl.prim_ids)
            d.mt_nodes, d.mt_links, d.mt_nn, d.mt_prims = p(self.mt.nodes), p(self.mt.links), self.mt.nn, p(self.mt.prim_ids)
        lam, war, ior, vol = materials.flat()
        self.keep += [lam, war, ior, vol, materials.types, materials.indices]
        d.nprims, d.mat_types, d.mat_indices = len(materials.types), p(materials.types), p(materials.indices)
        d.lambertians, d.nl = p(lam), len(materials.lambertians)
        d.wards, d.nw = p(war), len(materials.wards)
        d.iors, d.nd = p(ior), len(materials.dielectrics)
        d.volumes, d.nv = p(vol), len(materials.volumes)
        env = np.ascontiguousarray(env_rgb, np.float32)
        cam = camera.as_array()
        self.keep += [env, cam]
        d.env_rgb, d.env_channels, d.env_w, d.env_h = p(env), env.shape[2], env.shape[1], env.shape[0]
        d.camera, d.width, d.height, d.lens_radius = p(cam), camera.width, camera.height, camera.lens_radius
        d.consistent = 1 if consistent else 0
        self.desc = d
        self.n = path_count
        self.npixels = camera.width * camera.height
        self.h = lib.orc_wf_create(C.byref(d))

    def initialize(self, lane_index=0, num_lanes=1):
        lib.orc_wf_initialize(self.h, lane_index, num_lanes)

    def step(self, steps=1):
        lib.orc_wf_step(self.h, steps)

    def queue_counts(self):
        c = np.zeros(7, np.uint32)
        npx = u64(0)
        lib.orc_wf_queue_counts(self.h, p(c), C.byref(npx))
        return c, npx.value

    def accumulation(self):
        rgb = np.zeros((self.npixels, 3), np.float64)
        cnt = np.zeros(self.npixels, np.uint32)
        lib.orc_wf_read_accumulation(self.h, p(rgb), p(cnt))
        return rgb, cnt

    def state(self):
        n = self.n
        st = dict(rng=np.zeros((n, 4), np.uint32), rays=np.zeros(n, STATE_RAY_DTYPE), paths=np.zeros(n, PATH_DTYPE),
                  ext=np.zeros(n, EXT_DTYPE), shade=np.zeros(n, SHADE_DTYPE), vol=np.zeros(n, VOL_DTYPE),
                  items=np.zeros((7, n), np.uint32))
        lib.orc_wf_read_state(self.h, *[p(st[k]) for k in ("rng", "rays", "paths", "ext", "shade", "vol", "items")])
        return st

    def __del__(self):
        try:
            lib.orc_wf_destroy(self.h)
        except Exception:
            pass


# ------------------------------------------------------------------ output path (numpy)
# Intermediate image: half-float means and a sample count per pixel (RGB16IntermediateValueType,
# kernels/types.cl), the count widened to 32 bits.
RGB16_DTYPE = np.dtype([("r", "<u2"), ("g", "<u2"), ("b", "<u2"), ("pad", "<u2"), ("count", "<u4")])


def _half_bits(x):
    return np.asarray(x, np.float32).astype(np.float16).view(np.uint16)


def _half_value(bits):
    return np.asarray(bits, np.uint16).view(np.float16).astype(np.float32)


def to_intermediate(rgb_sums, counts):
    """accumlation_to_intermediate (kernels/accumlation.cl:6-25): div = 1 / count in float, 0 when
    not finite; each channel sum (as float) times div, rounded to half (nearest even)."""
    rgb = np.asarray(rgb_sums).astype(np.float32)
    cnt = np.asarray(counts, np.uint32)
    with np.errstate(divide="ignore", invalid="ignore"):
        div = np.float32(1.0) / cnt.astype(np.float32)
    div = np.where(np.isfinite(div), div, np.float32(0.0)).astype(np.float32)
    out = np.zeros(len(cnt), RGB16_DTYPE)
    for k, name in enumerate("rgb"):
        out[name] = _half_bits(rgb[:, k] * div)
    out["count"] = cnt
    return out


def merge_intermediate(a, b):
    """merge_intermediate (kernels/accumlation.cl:27-50): (a * sa + b * sb) * (1 / (sa + sb)) in
    float, rounded to half; zero when both counts are zero."""
    sa = a["count"].astype(np.uint32)
    sb = b["count"].astype(np.uint32)
    sab = (sa + sb).astype(np.uint32)
    out = np.zeros(len(a), RGB16_DTYPE)
    with np.errstate(divide="ignore", invalid="ignore"):
        one = np.float32(1.0) / sab.astype(np.float32)
    fa, fb = sa.astype(np.float32), sb.astype(np.float32)
    for name in "rgb":
        with np.errstate(invalid="ignore"):  # 0 * inf where both counts are zero; replaced below
            v = (_half_value(a[name]) * fa + _half_value(b[name]) * fb) * one
        out[name] = np.where(sab == 0, np.uint16(0), _half_bits(v))
    out["count"] = sab
    return out


def tonemap_intermediate(x):
    """tonemap (kernels/accumlation.cl:68-79): int(pow(mean, 1/2.2) * 256) clamped to 0..255, alpha
    255; the power by the oracle's double pow, rounded to float, as the device computes it."""
    out = np.zeros((len(x), 4), np.uint8)
    e = np.float32(1.0) / np.float32(2.2)
    for k, name in enumerate("rgb"):
        m = _half_value(x[name]).astype(np.float32)
        p = pow_f32(m, e) * np.float32(256.0)
        out[:, k] = np.clip(p.astype(np.int64), 0, 255)
    out[:, 3] = 255
    return out
