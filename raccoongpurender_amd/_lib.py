"""ctypes binding of the C ABI declared in include/raccoon_gpu.h.

The HIP library must have been built in-tree (raccoongpurender_amd/libraccoongpu.so);
there is no fallback path: a missing library raises at import of this module.
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libraccoongpu.so")

RGPU_ABI_VERSION = 3
RGPU_OK = 0
RGPU_ERR_INVALID_ARG = -1
RGB16_DTYPE = np.dtype([("r", "<u2"), ("g", "<u2"), ("b", "<u2"), ("pad", "<u2"), ("count", "<u4")])
TREELET_OFF = 0xFFFFFFFF
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
VARIANT_THREADED_SINGLE = 2
VARIANTS = {"stackless": VARIANT_STACKLESS, "threaded": VARIANT_THREADED, "threaded_single": VARIANT_THREADED_SINGLE}

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
    _fields_ = [("abi_version", C.c_uint32), ("max_leaf", C.c_uint32), ("bins", C.c_uint32), ("max_depth", C.c_uint32),
                ("traversal_cost", C.c_float)]


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
_sig("rgpu_bvh_build_device", C.c_int, vp, vp, u32, u32, vp, u32, vp, u32, C.POINTER(rgpu_bvh_options),
     C.POINTER(vp), C.POINTER(C.c_double))
_sig("rgpu_bvh_destroy", C.c_int, vp)
_sig("rgpu_bvh_get_stats", C.c_int, vp, C.POINTER(rgpu_bvh_stats))
_sig("rgpu_bvh_export", C.c_int, vp, vp, vp, vp, vp, vp, vp)
_sig("rgpu_scene_create", C.c_int, vp, vp, vp, u32, u32, vp, u32, vp, u32, C.POINTER(vp))
_sig("rgpu_scene_destroy", C.c_int, vp)
_sig("rgpu_set_trace_tuning", C.c_int, vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32)
_sig("rgpu_scene_bytes", C.c_int, vp, C.POINTER(C.c_uint64))
_sig("rgpu_extension_ray_cast", C.c_int, vp, vp, vp, C.c_int, vp, vp, u32, u32)
_sig("rgpu_extension_ray_cast_stats", C.c_int, vp, vp, C.c_int, vp, vp, u32, C.POINTER(C.c_uint64))

# Every symbol the public header declares; tests check the library exports them all.
EXPORTED = [
    "rgpu_create", "rgpu_destroy", "rgpu_last_error_string", "rgpu_device_count",
    "rgpu_alloc", "rgpu_free", "rgpu_upload", "rgpu_download", "rgpu_fill_u32", "rgpu_synchronize", "rgpu_copy",
    "rgpu_time_extension_ray_cast",
    "rgpu_bvh_build", "rgpu_bvh_destroy", "rgpu_bvh_get_stats", "rgpu_bvh_export",
    "rgpu_scene_create", "rgpu_scene_destroy", "rgpu_scene_bytes", "rgpu_set_trace_tuning",
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
        self._children = []  # weak references, creation order

    def check(self, rc, what):
        if rc != RGPU_OK:
            detail = lib.rgpu_last_error_string(self.h)
            raise RgpuError(rc, what, detail.decode() if detail else "")

    def adopt(self, obj):
        """Track an object whose C handle refers to this context: close() releases it first,
        because every rgpu_*_destroy of a device object reads its context."""
        if len(self._children) >= 64 and len(self._children) % 64 == 0:
            self._children = [r for r in self._children if r() is not None]
        self._children.append(weakref.ref(obj))
        return obj

    def close(self):
        if self.h:
            alive = [o for o in (r() for r in reversed(self._children)) if o is not None]
            # wavefronts first: a guide refuses to close while a wavefront has it attached
            alive.sort(key=lambda o: 0 if isinstance(o, Wavefront) else 1)
            for o in alive:
                o.close()
            self._children = []
            lib.rgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        self.check(lib.rgpu_synchronize(self.h, None), "rgpu_synchronize")

    def set_trace_tuning(self, waves_per_simd=0, refill=0, leaf_batch=0, treelet=0):
        """treelet: stackless top-of-tree records per workgroup in LDS (0 default, TREELET_OFF none)."""
        self.check(lib.rgpu_set_trace_tuning(self.h, waves_per_simd, refill, leaf_batch, treelet),
                   "rgpu_set_trace_tuning")

    def merge_intermediate(self, a, b):
        """Host intermediates in, their count-weighted merge out (rgpu_merge_intermediate)."""
        n = len(a)
        da, db = self.to_device(np.ascontiguousarray(a, RGB16_DTYPE)), self.to_device(np.ascontiguousarray(b, RGB16_DTYPE))
        self.check(lib.rgpu_merge_intermediate(self.h, None, da.ptr, db.ptr, n), "rgpu_merge_intermediate")
        return da.download(RGB16_DTYPE, n)

    def tonemap_intermediate(self, x):
        n = len(x)
        dx = self.to_device(np.ascontiguousarray(x, RGB16_DTYPE))
        out = self.alloc(max(n, 1) * 4)
        self.check(lib.rgpu_tonemap_intermediate(self.h, None, dx.ptr, n, out.ptr), "rgpu_tonemap_intermediate")
        return out.download(np.uint8, (n, 4))

    def copy(self, dst, src, nbytes):
        self.check(lib.rgpu_copy(self.h, None, dst.ptr, src.ptr, nbytes), "rgpu_copy")

    def stream(self):
        """A second stream of this device at the default priority (Stream.close releases it)."""
        return Stream(self)

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        b = DeviceBuffer(self, arr.nbytes)
        b.upload(arr)
        return b


class Stream:
    """A HIP stream of a context (rgpu_stream_create), for work that overlaps the context's own."""

    def __init__(self, ctx):
        h = vp()
        ctx.check(lib.rgpu_stream_create(ctx.h, C.byref(h)), "rgpu_stream_create")
        self.ctx, self.h = ctx, h
        ctx.adopt(self)

    def synchronize(self):
        self.ctx.check(lib.rgpu_synchronize(self.ctx.h, self.h), "rgpu_synchronize")

    def close(self):
        if self.h and self.ctx.h:
            lib.rgpu_stream_destroy(self.ctx.h, self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        h = vp()
        ctx.check(lib.rgpu_alloc(ctx.h, self.nbytes, C.byref(h)), "rgpu_alloc")
        self.ptr = h
        ctx.adopt(self)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.ctx.check(lib.rgpu_upload(self.ctx.h, None, self.ptr, ptr(arr), arr.nbytes), "rgpu_upload")

    def download(self, dtype, shape, stream=None):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        self.ctx.check(lib.rgpu_download(self.ctx.h, stream.h if stream is not None else None, ptr(out), self.ptr,
                                         out.nbytes), "rgpu_download")
        return out

    def fill_u32(self, value, count=None):
        count = self.nbytes // 4 if count is None else count
        self.ctx.check(lib.rgpu_fill_u32(self.ctx.h, None, self.ptr, value, count), "rgpu_fill_u32")

    def free(self):
        if self.ptr and self.ctx.h:
            lib.rgpu_free(self.ctx.h, self.ptr)
        self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BVH:
    """Host-built binned-SAH BVH2 (rgpu_bvh_build). Runs without a GPU."""

    def __init__(self, points, indices, spheres=None, max_leaf=4, bins=16, max_depth=63, traversal_cost=0.0,
                 device=None):
        """device: a Context to build on its GPU (rgpu_bvh_build_device; the same tree), else the host."""
        points = np.ascontiguousarray(points, dtype=np.float32)
        indices = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
        assert points.ndim == 2 and points.shape[1] in (3, 4)
        spheres = None if spheres is None or len(spheres) == 0 else np.ascontiguousarray(spheres, dtype=np.float32)
        self.points, self.indices, self.spheres = points, indices, spheres
        self.ntris = indices.size // 3
        self.nspheres = 0 if spheres is None else spheres.shape[0]
        opt = rgpu_bvh_options(RGPU_ABI_VERSION, max_leaf, bins, max_depth, traversal_cost)
        h = vp()
        self.device_ms = None
        if device is not None:
            ms = C.c_double(0.0)
            device.check(lib.rgpu_bvh_build_device(device.h, ptr(points), points.shape[1], points.shape[0],
                                                   ptr(indices), self.ntris, ptr(spheres), self.nspheres,
                                                   C.byref(opt), C.byref(h), C.byref(ms)), "rgpu_bvh_build_device")
            self.device_ms = ms.value
        else:
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
        ctx.adopt(self)

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
        if self.h and self.ctx.h:
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


# ====================================================================== wavefront path tracing
MAT_LAMBERTIAN, MAT_SPECULAR, MAT_DIELECTRIC, MAT_WARD, MAT_VOLUME = 1, 2, 3, 4, 5
WF_ACCUM64 = 1
WF_NO_REFERENCE_QUIRKS = 2
WF_CONSISTENT_NORMALS = 4
QUEUE_NAMES = ("new_path", "lambertian", "specular", "dielectric", "ward", "volume_surface", "volume_inside")


class rgpu_lambertian(C.Structure):
    _fields_ = [("Le", C.c_float * 3), ("R", C.c_float * 3), ("back_emission", C.c_int32)]


class rgpu_ward(C.Structure):
    _fields_ = [("alpha_x", C.c_float), ("alpha_y", C.c_float), ("reflectance", C.c_float * 3),
                ("edgetint", C.c_float * 3), ("falloff", C.c_float)]


class rgpu_dielectric(C.Structure):
    _fields_ = [("ior", C.c_float)]


class rgpu_volume(C.Structure):
    _fields_ = [("C", C.c_float), ("R", C.c_float * 3)]


class rgpu_camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("forward", C.c_float * 3), ("up", C.c_float * 3), ("right", C.c_float * 3),
                ("imageplane_o", C.c_float * 3), ("imageplane_r", C.c_float * 3), ("imageplane_b", C.c_float * 3),
                ("width", C.c_uint32), ("height", C.c_uint32), ("lens_radius", C.c_float)]

    def as_array(self):
        return np.array([*self.eye, *self.forward, *self.up, *self.right, *self.imageplane_o, *self.imageplane_r,
                         *self.imageplane_b], np.float32)


class rgpu_wavefront_config(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("path_count", C.c_uint32), ("seed", C.c_uint64),
                ("rr_start", C.c_uint32), ("max_depth", C.c_uint32), ("radiance_clamp", C.c_float),
                ("bxdf_select_p", C.c_float), ("flags", C.c_uint32), ("variant", C.c_uint32),
                ("queue_blocks", C.c_uint32)]


_sig("rgpu_materials_create", C.c_int, vp, u32, vp, vp, vp, u32, vp, u32, vp, u32, vp, u32, C.POINTER(vp))
_sig("rgpu_materials_destroy", C.c_int, vp)
_sig("rgpu_envmap_create", C.c_int, vp, vp, u32, u32, u32, C.POINTER(vp))
_sig("rgpu_envmap_destroy", C.c_int, vp)
_sig("rgpu_envmap_export", C.c_int, vp, C.c_int, vp, vp, vp, vp, vp)
_sig("rgpu_envmap_sample", C.c_int, vp, vp, vp, vp, vp, u32, vp)
_sig("rgpu_envmap_pdf", C.c_int, vp, vp, vp, vp, vp, u32, vp)
_sig("rgpu_camera_look_at", C.c_int, vp, vp, vp, C.c_float, C.c_float, u32, u32, C.c_float, C.POINTER(rgpu_camera))
_sig("rgpu_wavefront_create", C.c_int, vp, vp, vp, vp, C.POINTER(rgpu_wavefront_config), C.POINTER(vp))
_sig("rgpu_wavefront_destroy", C.c_int, vp)
_sig("rgpu_wavefront_initialize", C.c_int, vp, C.POINTER(rgpu_camera), u32, u32)
_sig("rgpu_wavefront_step", C.c_int, vp, vp, u32)
_sig("rgpu_wavefront_time_steps", C.c_int, vp, u32, C.POINTER(C.c_double))
_sig("rgpu_wavefront_queue_counts", C.c_int, vp, vp, C.POINTER(C.c_uint64))
_sig("rgpu_wavefront_read_accumulation", C.c_int, vp, vp, vp)
_sig("rgpu_wavefront_tonemap", C.c_int, vp, vp)
_sig("rgpu_wavefront_stat", C.c_int, vp, C.POINTER(C.c_uint64))
_sig("rgpu_wavefront_read_state", C.c_int, vp, vp, vp, vp, vp, vp, vp, vp)
_sig("rgpu_wavefront_intermediate", C.c_int, vp, vp, vp)
_sig("rgpu_merge_intermediate", C.c_int, vp, vp, vp, vp, u32)
_sig("rgpu_tonemap_intermediate", C.c_int, vp, vp, vp, u32, vp)
_sig("rgpu_wavefront_checkpoint_size", C.c_int, vp, C.POINTER(C.c_uint64))
_sig("rgpu_wavefront_checkpoint", C.c_int, vp, vp, C.c_uint64)
_sig("rgpu_wavefront_resume", C.c_int, vp, vp, C.c_uint64)
EXPORTED += ["rgpu_bvh_build_device", "rgpu_wavefront_intermediate", "rgpu_merge_intermediate",
             "rgpu_tonemap_intermediate"]
EXPORTED += ["rgpu_wavefront_checkpoint_size", "rgpu_wavefront_checkpoint", "rgpu_wavefront_resume"]
_sig("rgpu_wavefront_time_stages", C.c_int, vp, u32, C.POINTER(C.c_double))
EXPORTED += ["rgpu_wavefront_time_stages"]
STAGE_NAMES = ("new_path", "guide_infer", "lambertian", "ward", "mirror_dielectric", "medium", "ray_cast", "logic",
               "queues")
_sig("rgpu_stream_create", C.c_int, vp, C.POINTER(vp))
_sig("rgpu_stream_destroy", C.c_int, vp, vp)
_sig("rgpu_wavefront_set_publish", C.c_int, vp, u32)
_sig("rgpu_wavefront_update_intermediate", C.c_int, vp, vp, vp, vp)
EXPORTED += ["rgpu_stream_create", "rgpu_stream_destroy", "rgpu_wavefront_set_publish",
             "rgpu_wavefront_update_intermediate"]
_sig("rgpu_wavefront_graph_capture", C.c_int, vp, u32)
_sig("rgpu_wavefront_graph_launch", C.c_int, vp, vp, u32)
_sig("rgpu_wavefront_set_launch", C.c_int, vp, C.c_int, u32, u32)
_sig("rgpu_wavefront_get_launch", C.c_int, vp, C.POINTER(C.c_int), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32))
_sig("rgpu_wavefront_autotune", C.c_int, vp, C.POINTER(rgpu_camera), u32, u32, C.POINTER(C.c_double))
EXPORTED += ["rgpu_wavefront_graph_capture", "rgpu_wavefront_graph_launch", "rgpu_wavefront_set_launch",
             "rgpu_wavefront_get_launch", "rgpu_wavefront_autotune"]


class rgpu_batch_job(C.Structure):
    _fields_ = [("scene", vp), ("materials", vp), ("envmap", vp), ("camera", rgpu_camera), ("path_count", u32),
                ("spp", u32), ("seed", C.c_uint64)]


_sig("rgpu_wavefront_create_batch", C.c_int, vp, C.POINTER(rgpu_batch_job), u32, C.POINTER(rgpu_wavefront_config),
     C.POINTER(vp))
_sig("rgpu_wavefront_job_count", C.c_int, vp, C.POINTER(u32))
_sig("rgpu_wavefront_job_queue_counts", C.c_int, vp, u32, vp, C.POINTER(C.c_uint64))
_sig("rgpu_wavefront_job_accumulation", C.c_int, vp, u32, vp, vp)
_sig("rgpu_wavefront_job_stat", C.c_int, vp, u32, C.POINTER(C.c_uint64))
_sig("rgpu_wavefront_job_tonemap", C.c_int, vp, u32, vp)
_sig("rgpu_wavefront_job_tonemap_noalloc", C.c_int, vp, u32, vp)
_sig("rgpu_wavefront_run_batch", C.c_int, vp, u32, u32, C.POINTER(u32))
EXPORTED += ["rgpu_wavefront_create_batch", "rgpu_wavefront_job_count", "rgpu_wavefront_job_queue_counts",
             "rgpu_wavefront_job_accumulation", "rgpu_wavefront_job_stat", "rgpu_wavefront_job_tonemap",
             "rgpu_wavefront_run_batch", "rgpu_wavefront_job_tonemap_noalloc"]
EXPORTED += [
    "rgpu_materials_create", "rgpu_materials_destroy", "rgpu_envmap_create", "rgpu_envmap_destroy",
    "rgpu_envmap_export", "rgpu_envmap_sample", "rgpu_envmap_pdf", "rgpu_camera_look_at", "rgpu_wavefront_create", "rgpu_wavefront_destroy",
    "rgpu_wavefront_initialize", "rgpu_wavefront_step", "rgpu_wavefront_time_steps", "rgpu_wavefront_queue_counts",
    "rgpu_wavefront_read_accumulation", "rgpu_wavefront_tonemap", "rgpu_wavefront_stat", "rgpu_wavefront_read_state",
]

EXT_DTYPE = np.dtype([("hit_surface", np.int32), ("hit_volume", np.int32), ("tmin", np.float32), ("Ng", np.float32, 3),
                      ("pad", np.float32, 2)])
PATH_DTYPE = np.dtype([("T", np.float32, 3), ("logic_i", np.uint32), ("L", np.float32, 3), ("pad", np.uint32)])
SHADE_DTYPE = np.dtype([("Le", np.float32, 3), ("T", np.float32, 3), ("pad", np.float32, 2)])
VOL_DTYPE = np.dtype([("volumes", np.int32, 7), ("count", np.int32)])
STATE_RAY_DTYPE = np.dtype([("ro", np.float32, 3), ("tmax", np.float32), ("rd", np.float32, 3), ("pixel", np.uint32)])


def camera_look_at(eye, target, up, fov_y_deg, width, height, lens_radius=0.0, focus_distance=0.0):
    cam = rgpu_camera()
    a = [np.ascontiguousarray(x, np.float32) for x in (eye, target, up)]
    rc = lib.rgpu_camera_look_at(ptr(a[0]), ptr(a[1]), ptr(a[2]), fov_y_deg, focus_distance, width, height,
                                 lens_radius, C.byref(cam))
    if rc != RGPU_OK:
        raise RgpuError(rc, "rgpu_camera_look_at")
    return cam


class MaterialSet:
    """Per-primitive material assignment plus the per-type parameter tables (host side)."""

    def __init__(self, types, indices, lambertians=(), wards=(), dielectrics=(), volumes=()):
        self.types = np.ascontiguousarray(types, np.int32)
        self.indices = np.ascontiguousarray(indices, np.int32)
        self.lambertians = list(lambertians)  # (Le3, R3)
        self.wards = list(wards)              # (ax, ay, refl3, edge3, falloff)
        self.dielectrics = list(dielectrics)  # ior
        self.volumes = list(volumes)          # (C, R3)

    def c_arrays(self):
        L = (rgpu_lambertian * max(1, len(self.lambertians)))()
        for i, (Le, R) in enumerate(self.lambertians):
            L[i].Le[:] = list(map(float, Le))
            L[i].R[:] = list(map(float, R))
            L[i].back_emission = 0
        W = (rgpu_ward * max(1, len(self.wards)))()
        for i, (ax, ay, refl, edge, fo) in enumerate(self.wards):
            W[i].alpha_x, W[i].alpha_y, W[i].falloff = ax, ay, fo
            W[i].reflectance[:] = list(map(float, refl))
            W[i].edgetint[:] = list(map(float, edge))
        D = (rgpu_dielectric * max(1, len(self.dielectrics)))()
        for i, ior in enumerate(self.dielectrics):
            D[i].ior = ior
        V = (rgpu_volume * max(1, len(self.volumes)))()
        for i, (c, R) in enumerate(self.volumes):
            V[i].C = c
            V[i].R[:] = list(map(float, R))
        return L, W, D, V

    def flat(self):
        """Float tables in the oracle's layout."""
        lam = np.array([[*Le, *R, 0.0] for Le, R in self.lambertians] or [[0.0] * 7], np.float32)
        war = np.array([[ax, ay, *r, *e, fo] for ax, ay, r, e, fo in self.wards] or [[0.0] * 9], np.float32)
        ior = np.array(self.dielectrics or [1.5], np.float32)
        vol = np.array([[c, *R] for c, R in self.volumes] or [[1.0, 0, 0, 0]], np.float32)
        return lam, war, ior, vol


class Materials:
    def __init__(self, ctx, ms):
        self.ctx = ctx
        L, W, D, V = ms.c_arrays()
        h = vp()
        ctx.check(lib.rgpu_materials_create(ctx.h, len(ms.types), ptr(ms.types), ptr(ms.indices), L,
                                            len(ms.lambertians), W, len(ms.wards), D, len(ms.dielectrics), V,
                                            len(ms.volumes), C.byref(h)), "rgpu_materials_create")
        self.h = h
        ctx.adopt(self)

    def close(self):
        if self.h and self.ctx.h:
            lib.rgpu_materials_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Envmap:
    """Environment map + alias tables. ctx=None builds the host tables only."""

    def __init__(self, ctx, rgb):
        rgb = np.ascontiguousarray(rgb, np.float32)
        assert rgb.ndim == 3 and rgb.shape[2] in (3, 4)
        self.ctx, self.rgb = ctx, rgb
        self.height, self.width = rgb.shape[:2]
        h = vp()
        rc = lib.rgpu_envmap_create(ctx.h if ctx else None, ptr(rgb), rgb.shape[2], self.width, self.height,
                                    C.byref(h))
        if rc != RGPU_OK:
            raise RgpuError(rc, "rgpu_envmap_create")
        self.h = h
        if ctx is not None:
            ctx.adopt(self)

    def export(self, table):
        P = self.width * self.height
        pdf = np.zeros(P, np.float32)
        hts = np.zeros(P, np.float32)
        al = np.zeros(P, np.int32)
        fr = np.zeros((P, 4), np.float32)
        deg = C.c_int32(0)
        lib.rgpu_envmap_export(self.h, table, ptr(pdf), ptr(hts), ptr(al), ptr(fr), C.byref(deg))
        return dict(pdf=pdf, height=hts, alias=al, fragments=fr, degenerate=bool(deg.value))

    def sample_device(self, states, out, count, normals=None, stream=None):
        """Device buffers: states (count x 4 u32, advanced in place), out (count x 4 f32:
        direction, pdf), normals (count x 3 f32, six-axis tables) or None (uniform table)."""
        self.ctx.check(lib.rgpu_envmap_sample(self.ctx.h, stream, self.h, states.ptr,
                                              normals.ptr if normals is not None else None, count, out.ptr),
                       "rgpu_envmap_sample")

    def pdf_device(self, dirs, pdfs, count, normals=None, stream=None):
        self.ctx.check(lib.rgpu_envmap_pdf(self.ctx.h, stream, self.h, dirs.ptr,
                                           normals.ptr if normals is not None else None, count, pdfs.ptr),
                       "rgpu_envmap_pdf")

    def sample(self, states, normals=None):
        """Host arrays in, host arrays out: (directions, pdfs, advanced states)."""
        states = np.ascontiguousarray(states, np.uint32).reshape(-1, 4)
        n = len(states)
        s_dev = self.ctx.to_device(states)
        out = self.ctx.alloc(max(n, 1) * 16)
        n_dev = None if normals is None else self.ctx.to_device(np.ascontiguousarray(normals, np.float32))
        self.sample_device(s_dev, out, n, n_dev)
        res = out.download(np.float32, (n, 4))
        return res[:, :3].copy(), res[:, 3].copy(), s_dev.download(np.uint32, (n, 4))

    def pdf(self, dirs, normals=None):
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = len(dirs)
        d_dev = self.ctx.to_device(dirs)
        out = self.ctx.alloc(max(n, 1) * 4)
        n_dev = None if normals is None else self.ctx.to_device(np.ascontiguousarray(normals, np.float32))
        self.pdf_device(d_dev, out, n, n_dev)
        return out.download(np.float32, n)

    def close(self):
        if self.h and (self.ctx is None or self.ctx.h):
            lib.rgpu_envmap_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Wavefront:
    """One resident wavefront (WavefrontLane) on the GPU."""

    def __init__(self, ctx, scene, materials, envmap, path_count, seed=0, rr_start=0, max_depth=0, clamp=0.0,
                 bxdf_p=0.0, flags=0, variant="stackless", queue_blocks=0):
        self.ctx = ctx
        self.refs = (scene, materials, envmap)
        self.n = path_count
        cfg = rgpu_wavefront_config(RGPU_ABI_VERSION, path_count, seed, rr_start, max_depth, clamp, bxdf_p, flags,
                                    VARIANTS[variant] if isinstance(variant, str) else variant, queue_blocks)
        h = vp()
        ctx.check(lib.rgpu_wavefront_create(ctx.h, scene.h, materials.h, envmap.h, C.byref(cfg), C.byref(h)),
                  "rgpu_wavefront_create")
        self.h = h
        self.npixels = 0
        ctx.adopt(self)

    def initialize(self, camera, lane_index=0, num_lanes=1):
        self.ctx.check(lib.rgpu_wavefront_initialize(self.h, C.byref(camera), lane_index, num_lanes),
                       "rgpu_wavefront_initialize")
        self.npixels = camera.width * camera.height
        self.camera = camera

    def step(self, steps=1):
        self.ctx.check(lib.rgpu_wavefront_step(self.h, None, steps), "rgpu_wavefront_step")

    def time_steps(self, steps):
        ms = C.c_double(0)
        self.ctx.check(lib.rgpu_wavefront_time_steps(self.h, steps, C.byref(ms)), "rgpu_wavefront_time_steps")
        return ms.value

    def queue_counts(self):
        c = np.zeros(7, np.uint32)
        npx = C.c_uint64(0)
        self.ctx.check(lib.rgpu_wavefront_queue_counts(self.h, ptr(c), C.byref(npx)), "rgpu_wavefront_queue_counts")
        return c, npx.value

    def accumulation(self):
        rgb = np.zeros((self.npixels, 3), np.float64)
        cnt = np.zeros(self.npixels, np.uint32)
        self.ctx.check(lib.rgpu_wavefront_read_accumulation(self.h, ptr(rgb), ptr(cnt)),
                       "rgpu_wavefront_read_accumulation")
        return rgb, cnt

    def tonemap(self):
        out = np.zeros((self.npixels, 4), np.uint8)
        self.ctx.check(lib.rgpu_wavefront_tonemap(self.h, ptr(out)), "rgpu_wavefront_tonemap")
        return out

    def stat(self):
        t = C.c_uint64(0)
        self.ctx.check(lib.rgpu_wavefront_stat(self.h, C.byref(t)), "rgpu_wavefront_stat")
        return t.value

    def state(self):
        n = self.n
        st = dict(rng=np.zeros((n, 4), np.uint32), rays=np.zeros(n, STATE_RAY_DTYPE), paths=np.zeros(n, PATH_DTYPE),
                  ext=np.zeros(n, EXT_DTYPE), shade=np.zeros(n, SHADE_DTYPE), vol=np.zeros(n, VOL_DTYPE),
                  items=np.zeros((7, n), np.uint32))
        self.ctx.check(lib.rgpu_wavefront_read_state(self.h, *[ptr(st[k]) for k in
                                                                ("rng", "rays", "paths", "ext", "shade", "vol",
                                                                 "items")]), "rgpu_wavefront_read_state")
        return st

    def intermediate(self, dev=None):
        """The lane's intermediate image (RGB16_DTYPE, one per pixel); into device buffer dev if given."""
        out = dev if dev is not None else self.ctx.alloc(max(self.npixels, 1) * 12)
        self.ctx.check(lib.rgpu_wavefront_intermediate(self.h, None, out.ptr), "rgpu_wavefront_intermediate")
        return out if dev is not None else out.download(RGB16_DTYPE, self.npixels)

    def time_stages(self, steps):
        """Mean ms per step of each stage of the bounce (STAGE_NAMES), HIP-event timed."""
        out = (C.c_double * len(STAGE_NAMES))()
        self.ctx.check(lib.rgpu_wavefront_time_stages(self.h, steps, out), "rgpu_wavefront_time_stages")
        return dict(zip(STAGE_NAMES, (float(x) for x in out)))

    def set_publish(self, every_steps=1):
        """Publish the intermediate image for previews every `every_steps` steps (0: off)."""
        self.ctx.check(lib.rgpu_wavefront_set_publish(self.h, every_steps), "rgpu_wavefront_set_publish")

    def update_intermediate(self, dst, stream=None, took=None):
        """Preview side of the handoff: copy the published image into device buffer dst if the
        lock is free (dst unchanged otherwise); `took` (a 4-byte device buffer) receives 1 or 0.
        Asynchronous on `stream` (a Stream, or None for the context's stream)."""
        self.ctx.check(lib.rgpu_wavefront_update_intermediate(self.h, stream.h if stream is not None else None,
                                                              dst.ptr, took.ptr if took is not None else None),
                       "rgpu_wavefront_update_intermediate")

    def checkpoint(self):
        """Host snapshot of the lane (numpy uint8 array) that resume() continues exactly."""
        nb = C.c_uint64(0)
        self.ctx.check(lib.rgpu_wavefront_checkpoint_size(self.h, C.byref(nb)), "rgpu_wavefront_checkpoint_size")
        blob = np.zeros(nb.value, np.uint8)
        self.ctx.check(lib.rgpu_wavefront_checkpoint(self.h, ptr(blob), nb.value), "rgpu_wavefront_checkpoint")
        return blob

    def resume(self, blob):
        blob = np.ascontiguousarray(blob, np.uint8)
        self.ctx.check(lib.rgpu_wavefront_resume(self.h, ptr(blob), blob.nbytes), "rgpu_wavefront_resume")
        self.npixels = int(blob[16:20].view(np.uint32)[0])  # header: magic, version, path count, pixels

    def graph_capture(self, steps):
        self.ctx.check(lib.rgpu_wavefront_graph_capture(self.h, steps), "rgpu_wavefront_graph_capture")

    def graph_launch(self, launches=1):
        self.ctx.check(lib.rgpu_wavefront_graph_launch(self.h, None, launches), "rgpu_wavefront_graph_launch")

    def set_launch(self, variant, trace_block, queue_blocks):
        v = VARIANTS[variant] if isinstance(variant, str) else variant
        self.ctx.check(lib.rgpu_wavefront_set_launch(self.h, v, trace_block, queue_blocks),
                       "rgpu_wavefront_set_launch")

    def get_launch(self):
        v, b, q = C.c_int(0), C.c_uint32(0), C.c_uint32(0)
        self.ctx.check(lib.rgpu_wavefront_get_launch(self.h, C.byref(v), C.byref(b), C.byref(q)),
                       "rgpu_wavefront_get_launch")
        return {v: k for k, v in VARIANTS.items()}[v.value], b.value, q.value

    def autotune(self, camera, warmup=4, steps=8):
        ms = C.c_double(0)
        cam = C.byref(camera) if camera is not None else None
        self.ctx.check(lib.rgpu_wavefront_autotune(self.h, cam, warmup, steps, C.byref(ms)), "rgpu_wavefront_autotune")
        self.camera = camera
        return ms.value

    def close(self):
        if self.h and self.ctx.h:
            lib.rgpu_wavefront_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchJob:
    """One render job of a batch: device scene, materials, envmap, camera, slots, spp, seed."""

    def __init__(self, scene, materials, envmap, camera, path_count, spp=0, seed=0):
        self.scene, self.materials, self.envmap, self.camera = scene, materials, envmap, camera
        self.path_count, self.spp, self.seed = int(path_count), int(spp), int(seed)
        self.slots = (self.path_count + 255) // 256 * 256
        self.npixels = camera.width * camera.height


class Batch(Wavefront):
    """Several render jobs in one persistent wavefront (rgpu_wavefront_create_batch)."""

    def __init__(self, ctx, jobs, rr_start=0, max_depth=0, clamp=0.0, bxdf_p=0.0, flags=0, variant="stackless",
                 queue_blocks=0):
        self.ctx = ctx
        self.jobs = list(jobs)
        self.refs = tuple((j.scene, j.materials, j.envmap) for j in self.jobs)
        self.n = sum(j.slots for j in self.jobs)
        self.first_slot = np.cumsum([0] + [j.slots for j in self.jobs])[:-1]
        arr = (rgpu_batch_job * len(self.jobs))()
        for k, j in enumerate(self.jobs):
            arr[k] = rgpu_batch_job(j.scene.h, j.materials.h, j.envmap.h, j.camera, j.path_count, j.spp, j.seed)
        cfg = rgpu_wavefront_config(RGPU_ABI_VERSION, 0, 0, rr_start, max_depth, clamp, bxdf_p, flags,
                                    VARIANTS[variant] if isinstance(variant, str) else variant, queue_blocks)
        h = vp()
        ctx.check(lib.rgpu_wavefront_create_batch(ctx.h, arr, len(self.jobs), C.byref(cfg), C.byref(h)),
                  "rgpu_wavefront_create_batch")
        self.h = h
        self.npixels = 0
        ctx.adopt(self)

    def job_queue_counts(self, job):
        c = np.zeros(7, np.uint32)
        started = C.c_uint64(0)
        self.ctx.check(lib.rgpu_wavefront_job_queue_counts(self.h, job, ptr(c), C.byref(started)),
                       "rgpu_wavefront_job_queue_counts")
        return c, started.value

    def job_accumulation(self, job):
        P = self.jobs[job].npixels
        rgb = np.zeros((P, 3), np.float32)
        cnt = np.zeros(P, np.uint32)
        self.ctx.check(lib.rgpu_wavefront_job_accumulation(self.h, job, ptr(rgb), ptr(cnt)),
                       "rgpu_wavefront_job_accumulation")
        return rgb, cnt

    def job_stat(self, job):
        t = C.c_uint64(0)
        self.ctx.check(lib.rgpu_wavefront_job_stat(self.h, job, C.byref(t)), "rgpu_wavefront_job_stat")
        return t.value

    def job_tonemap(self, job):
        out = np.zeros((self.jobs[job].npixels, 4), np.uint8)
        self.ctx.check(lib.rgpu_wavefront_job_tonemap(self.h, job, ptr(out)), "rgpu_wavefront_job_tonemap")
        return out

    def job_tonemap_noalloc(self, job):
        """The same image through the entry point that allocates no device memory (jobs of at
        most 1024 x 1024 pixels)."""
        out = np.zeros((self.jobs[job].npixels, 4), np.uint8)
        self.ctx.check(lib.rgpu_wavefront_job_tonemap_noalloc(self.h, job, ptr(out)),
                       "rgpu_wavefront_job_tonemap_noalloc")
        return out

    def job_slots(self, job):
        a = int(self.first_slot[job])
        return slice(a, a + self.jobs[job].slots)

    def run(self, max_steps=100000, check_every=16):
        done = u32(0)
        self.ctx.check(lib.rgpu_wavefront_run_batch(self.h, max_steps, check_every, C.byref(done)),
                       "rgpu_wavefront_run_batch")
        return done.value

