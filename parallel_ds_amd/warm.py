"""Learned warm start over the C-ABI: the per-variable MLP (BF16 MFMA kernels on the
device) and the train -> predict -> solve pipeline on a batch of LPs."""
import ctypes as C

import numpy as np

from ._native import DsError, lib, ptr

N_FEATURES = 17  # the 16 LP features of ds_warm.hip and the family prior
WIDTH = 256
KPAD = 32


def param_count():
    return int(lib().ds_mlp_param_count())


def layout():
    """Offsets of the FP32 parameter vector: name -> (offset, shape)."""
    W, K = WIDTH, KPAD
    out, o = {}, 0
    for name, shape in (("W1", (W, K)), ("b1", (W,)), ("W2", (W, W)), ("b2", (W,)), ("W3", (W, W)), ("b3", (W,)),
                        ("w4", (W,)), ("b4", (1,))):
        out[name] = (o, shape)
        o += int(np.prod(shape))
    return out


def unpack(p):
    return {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in layout().items()}


class Mlp:
    """Per-variable network: n_features -> 256 -> 256 -> 256 -> 1 (ReLU), logit of
    'basic at the optimum'."""

    def __init__(self, n_features=N_FEATURES, seed=1, device=0):
        self._h = C.c_void_p()
        rc = lib().ds_mlp_create(device, n_features, seed, C.byref(self._h))
        if rc != 0:
            raise DsError(rc, "ds_mlp_create", "(no usable GPU?)")
        self.n_features = n_features

    def _check(self, rc, where):
        if rc != 0:
            d = lib().ds_mlp_last_error(self._h)
            raise DsError(rc, where, d.decode() if d else "")

    def close(self):
        if self._h:
            lib().ds_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def get_params(self):
        p = np.zeros(param_count(), np.float32)
        self._check(lib().ds_mlp_get_params(self._h, ptr(p)), "ds_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.shape != (param_count(),):
            raise ValueError("parameter vector of %d floats expected" % param_count())
        self._check(lib().ds_mlp_set_params(self._h, ptr(p)), "ds_mlp_set_params")

    def set_norm(self, mu, sd):
        mu = np.ascontiguousarray(mu, np.float32)
        sd = np.ascontiguousarray(sd, np.float32)
        self._check(lib().ds_mlp_set_norm(self._h, ptr(mu), ptr(sd)), "ds_mlp_set_norm")

    def get_norm(self):
        mu = np.zeros(self.n_features, np.float32)
        sd = np.zeros(self.n_features, np.float32)
        self._check(lib().ds_mlp_get_norm(self._h, ptr(mu), ptr(sd)), "ds_mlp_get_norm")
        return mu, sd

    def prior(self):
        """Family prior found by the last fit (feature 16), or an empty array."""
        n = C.c_int()
        self._check(lib().ds_mlp_get_prior(self._h, None, 0, C.byref(n)), "ds_mlp_get_prior")
        p = np.zeros(n.value, np.float32)
        if n.value:
            self._check(lib().ds_mlp_get_prior(self._h, ptr(p), n.value, C.byref(n)), "ds_mlp_get_prior")
        return p

    def infer(self, F):
        F = np.ascontiguousarray(F, np.float32)
        z = np.zeros(F.shape[0], np.float32)
        self._check(lib().ds_mlp_infer(self._h, F.shape[0], ptr(F), ptr(z)), "ds_mlp_infer")
        return z

    def grad(self, F, y, w=None):
        """(mean weighted BCE loss, gradient w.r.t. the parameter vector)."""
        F = np.ascontiguousarray(F, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        w = None if w is None else np.ascontiguousarray(w, np.float32)
        g = np.zeros(param_count(), np.float32)
        loss = C.c_double()
        self._check(lib().ds_mlp_grad(self._h, F.shape[0], ptr(F), ptr(y), ptr(w), C.byref(loss), ptr(g)),
                    "ds_mlp_grad")
        return loss.value, g

    def adam_step(self, lr):
        self._check(lib().ds_mlp_adam_step(self._h, C.c_float(lr)), "ds_mlp_adam_step")

    def fit(self, batch, epochs=4, minibatch=8192, lr=1e-3, seed=7):
        """Train on `batch`'s last solved bases; returns (per-epoch losses, device ms)."""
        losses = np.zeros(epochs)
        ms = C.c_double()
        self._check(lib().ds_mlp_fit(self._h, batch._ctx, epochs, minibatch, C.c_float(lr), seed, ptr(losses),
                                     C.byref(ms)), "ds_mlp_fit")
        return losses, ms.value

    def train_stats(self):
        ms, fl = C.c_double(), C.c_longlong()
        self._check(lib().ds_mlp_train_stats(self._h, C.byref(ms), C.byref(fl)), "ds_mlp_train_stats")
        return ms.value, fl.value
