"""Learned region prefilter: the 64 -> 256 -> 256 -> 1 MLP of the C ABI
(saber_hip_mlp_*), bf16 feature helpers and the region feature vector.

The MLP runs on hand-written bf16 MFMA kernels; there is no CPU path here.
"""
import ctypes as C

import numpy as np

from . import _ffi, _live_batches

N_IN, N_HID = 64, 256
NPARAMS = N_IN * N_HID + N_HID + N_HID * N_HID + N_HID + N_HID + 1
# offsets of the parameter blocks in the flat array
W1, B1, W2 = 0, N_IN * N_HID, N_IN * N_HID + N_HID
B2 = W2 + N_HID * N_HID
W3 = B2 + N_HID
B3 = W3 + N_HID


def bf16_bits(x):
    """float32 array -> bf16 bit patterns (uint16), round to nearest even."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    r = (u + np.uint32(0x7FFF) + ((u >> np.uint32(16)) & np.uint32(1))) >> np.uint32(16)
    return r.astype(np.uint16)


def bf16_float(b):
    """bf16 bit patterns -> float32."""
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << np.uint32(16)).view(np.float32)


def unpack(params):
    """Flat parameter array -> (W1 [64,256], b1, W2 [256,256], b2, w3, b3)."""
    p = np.asarray(params, dtype=np.float32)
    return (p[W1:B1].reshape(N_IN, N_HID), p[B1:W2], p[W2:B2].reshape(N_HID, N_HID), p[B2:W3], p[W3:B3],
            float(p[B3]))


def _x_bits(x):
    x = np.asarray(x)
    if x.dtype != np.uint16:
        x = bf16_bits(x)
    x = np.ascontiguousarray(x, dtype=np.uint16)
    if x.ndim != 2 or x.shape[1] != N_IN:
        raise ValueError(f"features must be [n, {N_IN}]")
    return x


class Mlp:
    """Device MLP of one context (saber_hip_mlp_*)."""

    def __init__(self, ctx, params=None, seed=0):
        self.ctx = ctx
        h = C.c_void_p()
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        if p is not None and p.size != NPARAMS:
            raise ValueError(f"params must hold {NPARAMS} floats")
        _ffi.check(_ffi.lib().saber_hip_mlp_create(ctx.h, _ffi.ptr(p, C.c_float), int(seed), C.byref(h)))
        self.h = h
        _live_batches.add(self)
        ctx._children.add(self)

    def params(self):
        out = np.zeros(NPARAMS, np.float32)
        _ffi.check(_ffi.lib().saber_hip_mlp_get_params(self.h, _ffi.ptr(out, C.c_float)))
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        _ffi.check(_ffi.lib().saber_hip_mlp_set_params(self.h, _ffi.ptr(p, C.c_float)))

    def forward(self, x):
        x = _x_bits(x)
        out = np.zeros(len(x), np.float32)
        _ffi.check(_ffi.lib().saber_hip_mlp_forward(self.h, _ffi.ptr(x, C.c_uint16), len(x),
                                                     _ffi.ptr(out, C.c_float)))
        return out

    def backward(self, x, labels, pos_weight=1.0):
        """(gradient of the mean weighted logistic loss, loss)."""
        x = _x_bits(x)
        lab = np.ascontiguousarray(labels, dtype=np.float32)
        g = np.zeros(NPARAMS, np.float32)
        loss = C.c_float()
        _ffi.check(_ffi.lib().saber_hip_mlp_backward(self.h, _ffi.ptr(x, C.c_uint16), _ffi.ptr(lab, C.c_float),
                                                      len(x), float(pos_weight), _ffi.ptr(g, C.c_float),
                                                      C.byref(loss)))
        return g, loss.value

    def train_step(self, x, labels, lr, pos_weight=1.0):
        x = _x_bits(x)
        lab = np.ascontiguousarray(labels, dtype=np.float32)
        loss = C.c_float()
        _ffi.check(_ffi.lib().saber_hip_mlp_train_step(self.h, _ffi.ptr(x, C.c_uint16), _ffi.ptr(lab, C.c_float),
                                                        len(x), float(pos_weight), float(lr), C.byref(loss)))
        return loss.value

    def upload(self, x, labels=None):
        x = _x_bits(x)
        lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.float32)
        _ffi.check(_ffi.lib().saber_hip_mlp_upload(self.h, _ffi.ptr(x, C.c_uint16), _ffi.ptr(lab, C.c_float),
                                                    len(x)))
        self.n_resident = len(x)

    def train_resident(self, row0, n, lr, pos_weight=1.0, want_loss=True):
        loss = C.c_float()
        _ffi.check(_ffi.lib().saber_hip_mlp_train_resident(self.h, int(row0), int(n), float(pos_weight), float(lr),
                                                            C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def forward_resident(self, row0, n):
        out = np.zeros(int(n), np.float32)
        _ffi.check(_ffi.lib().saber_hip_mlp_forward_resident(self.h, int(row0), int(n), _ffi.ptr(out, C.c_float)))
        return out

    def time(self, n, warmup=2, iters=5):
        """(forward ms, training step ms) over rows [0, n) of the resident data."""
        f, s = C.c_double(), C.c_double()
        _ffi.check(_ffi.lib().saber_hip_mlp_time(self.h, int(n), int(warmup), int(iters), C.byref(f), C.byref(s)))
        return f.value, s.value

    def close(self):
        if self.h:
            _ffi.lib().saber_hip_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def region_features(ctx, qres, qoff, sres, soff, jobs, hits, xscore, xq_lo, xq_hi, hits_per_region,
                    hsp_threshold=40, mean=None, std=None, time_iters=0):
    """bf16 feature vectors [n, 64] (uint16 bit patterns) of the regions
    `jobs` (band-job structured array) with the X-drop results of their hits
    (hits: [n * hits_per_region, 4] u32 query, subject, qp, sp); with mean and
    std, each feature is standardised on the device before rounding. With
    time_iters > 0, returns (features, mean ms of the feature kernel)."""
    mu = inv = None
    if mean is not None:
        mu = np.ascontiguousarray(mean, np.float32)
        s = np.asarray(std, np.float32)
        inv = np.ascontiguousarray(np.where(s > 0, 1.0 / np.where(s > 0, s, 1), 0.0), np.float32)
    qres = np.ascontiguousarray(qres, np.uint8)
    qoff = np.ascontiguousarray(qoff, np.uint64)
    sres = np.ascontiguousarray(sres, np.uint8)
    soff = np.ascontiguousarray(soff, np.uint64)
    qs = _ffi.SeqSet(_ffi.ptr(qres, C.c_uint8), _ffi.ptr(qoff, C.c_uint64), len(qoff) - 1, None)
    ss = _ffi.SeqSet(_ffi.ptr(sres, C.c_uint8), _ffi.ptr(soff, C.c_uint64), len(soff) - 1, None)
    jobs = np.ascontiguousarray(jobs)
    h = np.ascontiguousarray(hits, np.uint32).reshape(-1, 4)
    xs = np.ascontiguousarray(xscore, np.int32)
    lo = np.ascontiguousarray(xq_lo, np.uint32)
    hi = np.ascontiguousarray(xq_hi, np.uint32)
    n = len(jobs)
    out = np.zeros((max(n, 1), N_IN), np.uint16)
    args = (ctx.h, C.byref(qs), C.byref(ss), jobs.ctypes.data_as(C.POINTER(_ffi.BandJob)), n,
            _ffi.ptr(h, C.c_uint32), _ffi.ptr(xs, C.c_int32), _ffi.ptr(lo, C.c_uint32), _ffi.ptr(hi, C.c_uint32),
            int(hits_per_region), int(hsp_threshold), _ffi.ptr(mu, C.c_float), _ffi.ptr(inv, C.c_float),
            _ffi.ptr(out, C.c_uint16))
    if time_iters > 0:
        ms = C.c_double()
        _ffi.check(_ffi.lib().saber_hip_region_features_time(*args, int(time_iters), C.byref(ms)))
        return out[:n], ms.value
    _ffi.check(_ffi.lib().saber_hip_region_features(*args))
    return out[:n]


def fold_normalization(params, mean, std):
    """Parameters that give, on raw features x, the outputs the model
    `params` gives on standardised features (x - mean) / std."""
    p = np.array(params, dtype=np.float64)
    w1 = p[W1:B1].reshape(N_IN, N_HID)
    s = np.where(np.asarray(std, np.float64) > 0, np.asarray(std, np.float64), 1.0)
    mu = np.asarray(mean, np.float64)
    p[B1:W2] = p[B1:W2] - (mu / s) @ w1
    p[W1:B1] = (w1 / s[:, None]).ravel()
    return p.astype(np.float32)
