"""Path-guiding network bindings (rgpu_guide_* in include/raccoon_gpu.h)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import RGPU_ABI_VERSION, ptr

vp, u32 = C.c_void_p, C.c_uint32
SAMPLE_DTYPE = np.dtype([("p", np.float32, 3), ("n", np.float32, 3), ("wo", np.float32, 3), ("wi", np.float32, 3),
                         ("weight", np.float32), ("scene_class", np.float32), ("other_pdf", np.float32),
                         ("pad", np.float32)])
assert SAMPLE_DTYPE.itemsize == 64
N_OUT = 64
LOBES = 8


class rgpu_guide_config(C.Structure):
    _fields_ = [("abi_version", u32), ("select_p", C.c_float), ("capacity", u32), ("minibatch", u32),
                ("learning_rate", C.c_float), ("seed", u32), ("scene_lo", C.c_float * 3),
                ("scene_hi", C.c_float * 3), ("max_depth", u32), ("weight_cap", C.c_float),
                ("log_kappa_max", C.c_float), ("learn_select", u32)]


_sig = _lib._sig
_sig("rgpu_guide_create", C.c_int, vp, C.POINTER(rgpu_guide_config), C.POINTER(vp))
_sig("rgpu_guide_destroy", C.c_int, vp)
_sig("rgpu_guide_param_count", C.c_int)
_sig("rgpu_guide_params", C.c_int, vp, vp, vp)
_sig("rgpu_guide_forward", C.c_int, vp, vp, u32, vp)
_sig("rgpu_guide_gradients", C.c_int, vp, vp, u32, vp)
_sig("rgpu_guide_train_samples", C.c_int, vp, vp, u32)
_sig("rgpu_guide_read_samples", C.c_int, vp, vp, u32, C.POINTER(u32))
_sig("rgpu_wavefront_attach_guide", C.c_int, vp, vp, C.c_int, C.c_int)
_sig("rgpu_wavefront_guide_pending", C.c_int, vp, C.POINTER(u32))
_sig("rgpu_wavefront_guide_train", C.c_int, vp, u32, C.POINTER(u32))
_sig("rgpu_wavefront_guide_lobes", C.c_int, vp, u32, u32, vp)
_lib.EXPORTED += ["rgpu_guide_create", "rgpu_guide_destroy", "rgpu_guide_param_count", "rgpu_guide_params",
                  "rgpu_guide_forward", "rgpu_guide_gradients", "rgpu_guide_train_samples", "rgpu_guide_read_samples",
                  "rgpu_wavefront_attach_guide", "rgpu_wavefront_guide_pending", "rgpu_wavefront_guide_train",
                  "rgpu_wavefront_guide_lobes"]
PARAM_COUNT = _lib.lib.rgpu_guide_param_count()


class Guide:
    def __init__(self, ctx, scene_lo=(0, 0, 0), scene_hi=(1, 1, 1), select_p=0.5, capacity=1 << 18,
                 minibatch=16384, lr=1e-3, seed=1, max_depth=0, weight_cap=0.0, log_kappa_max=0.0,
                 learn_select=False):
        self.ctx = ctx
        cfg = rgpu_guide_config(RGPU_ABI_VERSION, select_p, capacity, minibatch, lr, seed,
                                (C.c_float * 3)(*scene_lo), (C.c_float * 3)(*scene_hi), max_depth, weight_cap,
                                log_kappa_max, 1 if learn_select else 0)
        h = vp()
        ctx.check(_lib.lib.rgpu_guide_create(ctx.h, C.byref(cfg), C.byref(h)), "rgpu_guide_create")
        self.h = h
        ctx.adopt(self)

    def params(self):
        out = np.zeros(PARAM_COUNT, np.float32)
        self.ctx.check(_lib.lib.rgpu_guide_params(self.h, ptr(out), None), "rgpu_guide_params")
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == PARAM_COUNT
        self.ctx.check(_lib.lib.rgpu_guide_params(self.h, None, ptr(w)), "rgpu_guide_params")

    def forward(self, samples):
        samples = np.ascontiguousarray(samples, SAMPLE_DTYPE)
        out = np.zeros((len(samples), N_OUT), np.float32)
        self.ctx.check(_lib.lib.rgpu_guide_forward(self.h, ptr(samples), len(samples), ptr(out)), "rgpu_guide_forward")
        return out

    def gradients(self, samples):
        samples = np.ascontiguousarray(samples, SAMPLE_DTYPE)
        g = np.zeros(PARAM_COUNT, np.float32)
        self.ctx.check(_lib.lib.rgpu_guide_gradients(self.h, ptr(samples), len(samples), ptr(g)),
                       "rgpu_guide_gradients")
        return g

    def train(self, samples):
        samples = np.ascontiguousarray(samples, SAMPLE_DTYPE)
        self.ctx.check(_lib.lib.rgpu_guide_train_samples(self.h, ptr(samples), len(samples)),
                       "rgpu_guide_train_samples")

    def buffered_samples(self, rows=None):
        """The training samples a wavefront has buffered (all of them, or the first `rows`)."""
        n = u32(0)
        self.ctx.check(_lib.lib.rgpu_guide_read_samples(self.h, None, 0, C.byref(n)), "rgpu_guide_read_samples")
        rows = n.value if rows is None else int(rows)
        out = np.zeros(rows, SAMPLE_DTYPE)
        self.ctx.check(_lib.lib.rgpu_guide_read_samples(self.h, ptr(out), rows, C.byref(n)),
                       "rgpu_guide_read_samples")
        return out

    def close(self):
        """Releases the network. Raises (and keeps the guide) while a wavefront has it attached."""
        if self.h and self.ctx.h:
            self.ctx.check(_lib.lib.rgpu_guide_destroy(self.h), "rgpu_guide_destroy")
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def attach(wavefront, guide, sampling=True, training=True):
    wavefront.ctx.check(_lib.lib.rgpu_wavefront_attach_guide(wavefront.h, guide.h if guide else None,
                                                             1 if sampling else 0, 1 if training else 0),
                        "rgpu_wavefront_attach_guide")
    wavefront._guide = guide


def pending(wavefront):
    n = C.c_uint32(0)
    wavefront.ctx.check(_lib.lib.rgpu_wavefront_guide_pending(wavefront.h, C.byref(n)), "rgpu_wavefront_guide_pending")
    return n.value


def vertex_lobes(wavefront, first, count):
    """Lobe records [first, first + count) as the last inference wrote them, one per vertex of
    that bounce in queue order (the lambertian queue, then the Ward queue): an array
    (count, 8, 6) of (mean xyz, kappa, weight, log(weight C(kappa)))."""
    out = np.zeros((count, LOBES, 6), np.float32)
    wavefront.ctx.check(_lib.lib.rgpu_wavefront_guide_lobes(wavefront.h, first, count, ptr(out)),
                        "rgpu_wavefront_guide_lobes")
    return out


def train_pass(wavefront, min_samples):
    n = C.c_uint32(0)
    wavefront.ctx.check(_lib.lib.rgpu_wavefront_guide_train(wavefront.h, min_samples, C.byref(n)),
                        "rgpu_wavefront_guide_train")
    return n.value


def raw_to_lobes(out):
    """Raw head outputs (rows x 64) -> lobe parameters (mu, kappa, weight), as the kernels do."""
    out = np.asarray(out, np.float64)
    o = out[:, :40].reshape(-1, LOBES, 5)
    m = o[..., :3]
    mx = np.abs(m).max(-1, keepdims=True)
    # a mean whose largest component is below 1e-15 (or not finite) has no direction: the
    # kernels give that lobe +z
    ok = (mx > 1e-15) & (mx <= 3.0e38)
    n = np.sqrt(np.where(ok, (m * m).sum(-1, keepdims=True), 1.0))
    mu = np.where(ok, m / n, np.array([0.0, 0.0, 1.0]))
    kappa = np.exp(np.clip(o[..., 3], -3.0, 7.5))
    lg = o[..., 4] - o[..., 4].max(-1, keepdims=True)
    w = np.exp(lg) / np.exp(lg).sum(-1, keepdims=True)
    return mu, kappa, w
