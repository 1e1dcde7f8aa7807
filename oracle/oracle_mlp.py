"""ORACLE — test infrastructure only. ctypes front end of oracle/oracle_mlp.cpp (FP64 model of
the path-guiding network)."""
import ctypes as C

import numpy as np

import oracle as O

lib = O.lib
vp, u32 = C.c_void_p, C.c_uint32
O._sig("orc_mlp_param_count", C.c_int)
O._sig("orc_mlp_encode", None, u32, vp, vp, vp, vp, vp)
O._sig("orc_mlp_forward", None, vp, u32, vp, vp, C.c_int)
O._sig("orc_mlp_pdf", None, u32, vp, vp, vp)
O._sig("orc_mlp_loss_grad", C.c_double, vp, u32, vp, vp, vp, vp, C.c_int)
O._sig("orc_mlp_loss_grad_sel", C.c_double, vp, u32, vp, vp, vp, vp, vp, C.c_int)
O._sig("orc_adam", None, u32, vp, vp, vp, vp, C.c_double, C.c_double, C.c_double, C.c_double, u32)
O._sig("orc_mlp_sample", None, vp, u32, vp, vp)
O._sig("orc_bf16", C.c_double, C.c_double)
P = O.p
PARAM_COUNT = lib.orc_mlp_param_count()
# parameter offsets of the four layers ([out][in] weights, then biases)
_IN, _HID, _OUT = 96, 64, 64
OFFSETS = {"W0": 0}
OFFSETS["B0"] = OFFSETS["W0"] + _HID * _IN
OFFSETS["W1"] = OFFSETS["B0"] + _HID
OFFSETS["B1"] = OFFSETS["W1"] + _HID * _HID
OFFSETS["W2"] = OFFSETS["B1"] + _HID
OFFSETS["B2"] = OFFSETS["W2"] + _HID * _HID
OFFSETS["W3"] = OFFSETS["B2"] + _HID
OFFSETS["B3"] = OFFSETS["W3"] + _OUT * _HID
assert OFFSETS["B3"] + _OUT == PARAM_COUNT
N_IN, N_OUT = 96, 64


def encode(p01, n, wo, cls=None):
    """cls: per-row scene class codes (0 = none, 1..8 = one-hot feature 77 + c), or None."""
    p01, n, wo = [np.ascontiguousarray(a, np.float64) for a in (p01, n, wo)]
    x = np.zeros((len(p01), N_IN))
    c = None if cls is None else np.ascontiguousarray(cls, np.uint32)
    lib.orc_mlp_encode(len(p01), P(p01), P(n), P(wo), None if c is None else c.ctypes.data, P(x))
    return x


def forward(params, x, emulate=False):
    params = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros((len(x), N_OUT))
    lib.orc_mlp_forward(P(params), len(x), P(x), P(out), 1 if emulate else 0)
    return out


def pdf(out, w):
    out = np.ascontiguousarray(out, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    r = np.zeros(len(w))
    lib.orc_mlp_pdf(len(w), P(out), P(w), P(r))
    return r


def loss_grad(params, x, wi, weight, emulate=False, other_pdf=None):
    """Mean loss and gradient; with other_pdf (the other strategies' density at wi per row), rows
    with other_pdf > 0 add the one-sample mixture term of the learned selection probability."""
    params = np.ascontiguousarray(params, np.float64)
    x, wi, weight = [np.ascontiguousarray(a, np.float64) for a in (x, wi, weight)]
    g = np.zeros(PARAM_COUNT)
    if other_pdf is None:
        loss = lib.orc_mlp_loss_grad(P(params), len(x), P(x), P(wi), P(weight), P(g), 1 if emulate else 0)
    else:
        po = np.ascontiguousarray(other_pdf, np.float64)
        loss = lib.orc_mlp_loss_grad_sel(P(params), len(x), P(x), P(wi), P(weight), P(po), P(g), 1 if emulate else 0)
    return loss, g


def adam(w, m, v, g, lr, t, b1=0.9, b2=0.999, eps=1e-8):
    lib.orc_adam(len(w), P(w), P(m), P(v), P(np.ascontiguousarray(g, np.float64)), lr, b1, b2, eps, t)


def sample(out_row, u):
    out_row = np.ascontiguousarray(out_row, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    w = np.zeros((len(u), 3))
    lib.orc_mlp_sample(P(out_row), len(u), P(u), P(w))
    return w


def bf16(x):
    return lib.orc_bf16(float(x))
