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
NTHREADS = min(16, os.cpu_count() or 1)


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
    """Reference threaded layout: MTBVHNode + links[6][node][hit, miss] (common/threaded_bvh.hpp:398-525)."""

    def __init__(self, tree):
        nn = len(tree["left"])
        self.nn = nn
        self.nodes = np.zeros(nn, REF_MT_NODE)
        self.links = np.zeros((6, nn, 2), np.int32)
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
