"""ctypes binding of the C ABI declared in include/raccoon_gpu.h.

The HIP library must have been built in-tree (raccoongpurender_amd/libraccoongpu.so);
there is no fallback path: a missing library raises at import of this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libraccoongpu.so")

RGPU_ABI_VERSION = 1
RGPU_OK = 0
ERRORS = {
    -1: "invalid argument",
    -2: "out of memory",
    -3: "HIP error",
    -4: "no device",
    -5: "no code object for this GPU",
    -6: "unsupported",
}
VARIANT_STACKLESS = 0
VARIANT_THREADED = 1
VARIANTS = {"stackless": VARIANT_STACKLESS, "threaded": VARIANT_THREADED}

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python -m raccoongpurender_amd.build_lib` "
        "(or __graft_entry__.build()); there is no non-HIP fallback."
    )
lib = C.CDLL(LIB_PATH)


class RgpuError(RuntimeError):
    def __init__(self, code, what, detail=""):
        self.code = code
        msg = f"{what} failed: {ERRORS.get(code, code)}"
        if detail:
            msg += f" ({detail})"
        super().__init__(msg)


class rgpu_config(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("flags", C.c_uint32)]


class rgpu_bvh_options(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("max_leaf", C.c_uint32), ("bins", C.c_uint32), ("max_depth", C.c_uint32)]


class rgpu_bvh_stats(C.Structure):
    _fields_ = [("nodes", C.c_uint32), ("leaves", C.c_uint32), ("depth", C.c_uint32), ("slots", C.c_uint32),
                ("sah_cost", C.c_double)]


vp = C.c_void_p
u32 = C.c_uint32
i32 = C.c_int32
PF = C.POINTER(C.c_float)


def _sig(name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


_sig("rgpu_device_count", C.c_int, C.POINTER(C.c_int))
_sig("rgpu_create", C.c_int, C.c_int, C.POINTER(rgpu_config), C.POINTER(vp))
_sig("rgpu_destroy", C.c_int, vp)
_sig("rgpu_last_error_string", C.c_char_p, vp)
_sig("rgpu_alloc", C.c_int, vp, C.c_size_t, C.POINTER(vp))
_sig("rgpu_free", C.c_int, vp, vp)
_sig("rgpu_upload", C.c_int, vp, vp, vp, vp, C.c_size_t)
_sig("rgpu_download", C.c_int, vp, vp, vp, vp, C.c_size_t)
_sig("rgpu_fill_u32", C.c_int, vp, vp, vp, u32, C.c_size_t)
_sig("rgpu_synchronize", C.c_int, vp, vp)
_sig("rgpu_copy", C.c_int, vp, vp, vp, vp, C.c_size_t)
_sig("rgpu_time_extension_ray_cast", C.c_int, vp, vp, C.c_int, vp, vp, u32, u32, C.c_int, C.POINTER(C.c_double))
_sig("rgpu_bvh_build", C.c_int, vp, u32, u32, vp, u32, vp, u32, C.POINTER(rgpu_bvh_options), C.POINTER(vp))
_sig("rgpu_bvh_destroy", C.c_int, vp)
_sig("rgpu_bvh_get_stats", C.c_int, vp, C.POINTER(rgpu_bvh_stats))
_sig("rgpu_bvh_export", C.c_int, vp, vp, vp, vp, vp, vp, vp)
_sig("rgpu_scene_create", C.c_int, vp, vp, vp, u32, u32, vp, u32, vp, u32, C.POINTER(vp))
_sig("rgpu_scene_destroy", C.c_int, vp)
_sig("rgpu_scene_bytes", C.c_int, vp, C.POINTER(C.c_uint64))
_sig("rgpu_extension_ray_cast", C.c_int, vp, vp, vp, C.c_int, vp, vp, u32, u32)
_sig("rgpu_extension_ray_cast_stats", C.c_int, vp, vp, C.c_int, vp, vp, u32, C.POINTER(C.c_uint64))

# Every symbol the public header declares; tests check the library exports them all.
EXPORTED = [
    "rgpu_create", "rgpu_destroy", "rgpu_last_error_string", "rgpu_device_count",
    "rgpu_alloc", "rgpu_free", "rgpu_upload", "rgpu_download", "rgpu_fill_u32", "rgpu_synchronize", "rgpu_copy",
    "rgpu_time_extension_ray_cast",
    "rgpu_bvh_build", "rgpu_bvh_destroy", "rgpu_bvh_get_stats", "rgpu_bvh_export",
    "rgpu_scene_create", "rgpu_scene_destroy", "rgpu_scene_bytes",
    "rgpu_extension_ray_cast", "rgpu_extension_ray_cast_stats",
]
COUNTER_NAMES = ("branch_visits", "leaf_visits", "parent_steps", "prim_tests", "node_visits")


def ptr(a):
    """Host pointer of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return C.c_void_p(a.ctypes.data)


def device_count():
    n = C.c_int(0)
    rc = lib.rgpu_device_count(C.byref(n))
    return n.value if rc == RGPU_OK else 0


class Context:
    """One GPU (rgpu_create). Owns a default-priority HIP stream."""

    def __init__(self, device_index=0):
        h = vp()
        cfg = rgpu_config(RGPU_ABI_VERSION, 0)
        rc = lib.rgpu_create(device_index, C.byref(cfg), C.byref(h))
        if rc != RGPU_OK:
            raise RgpuError(rc, "rgpu_create")
        self.h = h

    def check(self, rc, what):
        if rc != RGPU_OK:
            detail = lib.rgpu_last_error_string(self.h)
            raise RgpuError(rc, what, detail.decode() if detail else "")

    def close(self):
        if self.h:
            lib.rgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        self.check(lib.rgpu_synchronize(self.h, None), "rgpu_synchronize")

    def copy(self, dst, src, nbytes):
        self.check(lib.rgpu_copy(self.h, None, dst.ptr, src.ptr, nbytes), "rgpu_copy")

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        b = DeviceBuffer(self, arr.nbytes)
        b.upload(arr)
        return b


class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        h = vp()
        ctx.check(lib.rgpu_alloc(ctx.h, self.nbytes, C.byref(h)), "rgpu_alloc")
        self.ptr = h

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.ctx.check(lib.rgpu_upload(self.ctx.h, None, self.ptr, ptr(arr), arr.nbytes), "rgpu_upload")

    def download(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        self.ctx.check(lib.rgpu_download(self.ctx.h, None, ptr(out), self.ptr, out.nbytes), "rgpu_download")
        return out

    def fill_u32(self, value, count=None):
        count = self.nbytes // 4 if count is None else count
        self.ctx.check(lib.rgpu_fill_u32(self.ctx.h, None, self.ptr, value, count), "rgpu_fill_u32")

    def free(self):
        if self.ptr:
            lib.rgpu_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ctx.h:
                self.free()
        except Exception:
            pass


class BVH:
    """Host-built binned-SAH BVH2 (rgpu_bvh_build). Runs without a GPU."""

    def __init__(self, points, indices, spheres=None, max_leaf=4, bins=16, max_depth=63):
        points = np.ascontiguousarray(points, dtype=np.float32)
        indices = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
        assert points.ndim == 2 and points.shape[1] in (3, 4)
        spheres = None if spheres is None or len(spheres) == 0 else np.ascontiguousarray(spheres, dtype=np.float32)
        self.points, self.indices, self.spheres = points, indices, spheres
        self.ntris = indices.size // 3
        self.nspheres = 0 if spheres is None else spheres.shape[0]
        opt = rgpu_bvh_options(RGPU_ABI_VERSION, max_leaf, bins, max_depth)
        h = vp()
        rc = lib.rgpu_bvh_build(ptr(points), points.shape[1], points.shape[0], ptr(indices), self.ntris,
                                ptr(spheres), self.nspheres, C.byref(opt), C.byref(h))
        if rc != RGPU_OK:
            raise RgpuError(rc, "rgpu_bvh_build")
        self.h = h
        st = rgpu_bvh_stats()
        lib.rgpu_bvh_get_stats(h, C.byref(st))
        self.nodes, self.leaves, self.depth, self.slots, self.sah_cost = (
            st.nodes, st.leaves, st.depth, st.slots, st.sah_cost)

    def export(self):
        n, s = self.nodes, self.slots
        bounds = np.zeros((n, 6), np.float32)
        left = np.zeros(n, np.int32)
        right = np.zeros(n, np.int32)
        beg = np.zeros(n, np.uint32)
        cnt = np.zeros(n, np.uint32)
        pids = np.zeros(max(s, 1), np.uint32)
        lib.rgpu_bvh_export(self.h, ptr(bounds), ptr(left), ptr(right), ptr(beg), ptr(cnt), ptr(pids))
        return dict(bounds=bounds, left=left, right=right, prim_beg=beg, prim_cnt=cnt, prim_ids=pids[:s])

    def __del__(self):
        try:
            if self.h:
                lib.rgpu_bvh_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Scene:
    """Device-resident scene (rgpu_scene_create): packed node tables, primitive records, vertices."""

    def __init__(self, ctx, bvh):
        self.ctx = ctx
        self.bvh = bvh
        h = vp()
        p = bvh.points
        ctx.check(lib.rgpu_scene_create(ctx.h, bvh.h, ptr(p), p.shape[1], p.shape[0], ptr(bvh.indices), bvh.ntris,
                                        ptr(bvh.spheres), bvh.nspheres, C.byref(h)), "rgpu_scene_create")
        self.h = h

    def bytes(self):
        out = (C.c_uint64 * 4)()
        self.ctx.check(lib.rgpu_scene_bytes(self.h, out), "rgpu_scene_bytes")
        return dict(stackless_nodes=out[0], threaded_tables=out[1], prim_records=out[2], vertices=out[3])

    def extension_ray_cast(self, variant, rays_dev, hits_dev, count, block=0, stream=None):
        v = VARIANTS[variant] if isinstance(variant, str) else variant
        self.ctx.check(lib.rgpu_extension_ray_cast(self.ctx.h, stream, self.h, v, rays_dev.ptr, hits_dev.ptr,
                                                   count, block), "rgpu_extension_ray_cast")

    def extension_ray_cast_stats(self, variant, rays_dev, hits_dev, count):
        v = VARIANTS[variant] if isinstance(variant, str) else variant
        out = (C.c_uint64 * 5)()
        self.ctx.check(lib.rgpu_extension_ray_cast_stats(self.ctx.h, self.h, v, rays_dev.ptr, hits_dev.ptr, count,
                                                         out), "rgpu_extension_ray_cast_stats")
        return dict(zip(COUNTER_NAMES, (int(x) for x in out)))

    def time_extension_ray_cast(self, variant, rays_dev, hits_dev, count, iters, block=0):
        v = VARIANTS[variant] if isinstance(variant, str) else variant
        ms = C.c_double(0)
        self.ctx.check(lib.rgpu_time_extension_ray_cast(self.ctx.h, self.h, v, rays_dev.ptr, hits_dev.ptr, count,
                                                        block, iters, C.byref(ms)), "rgpu_time_extension_ray_cast")
        return ms.value

    def close(self):
        if self.h:
            lib.rgpu_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


RAY_DTYPE = np.dtype([("ro", np.float32, 3), ("tmax", np.float32), ("rd", np.float32, 3), ("pad", np.uint32)])
HIT_DTYPE = np.dtype([("prim", np.int32), ("t", np.float32), ("u", np.float32), ("v", np.float32)])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16


def trace(ctx, scene, rays, variant="stackless", block=0):
    """Upload rays, cast them on the GPU and return the hits (host arrays)."""
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    n = rays.shape[0]
    rd = ctx.to_device(rays)
    hd = ctx.alloc(max(n, 1) * HIT_DTYPE.itemsize)
    scene.extension_ray_cast(variant, rd, hd, n, block)
    ctx.synchronize()
    out = hd.download(HIT_DTYPE, (n,))
    rd.free()
    hd.free()
    return out
