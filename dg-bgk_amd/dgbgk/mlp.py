"""Python handle on the warm-start MLP kernels (include/dgbgk.h, dgbgk_mlp_*)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import lib


def _check(code: int):
    if code != 0:
        raise _lib.DgbgkError(code, lib().dgbgk_mlp_last_error().decode())


class MLP:
    """fin -> width (SiLU) x hidden -> fout; BF16 MFMA kernels, Adam."""

    def __init__(self, fin=5, width=256, hidden=6, fout=4, batch=65536):
        h = C.c_void_p()
        _check(lib().dgbgk_mlp_create(fin, width, hidden, fout, batch, C.byref(h)))
        self.h = h
        self.fin, self.width, self.hidden, self.fout, self.batch = fin, width, hidden, fout, batch
        n = C.c_long()
        _check(lib().dgbgk_mlp_nparams(h, C.byref(n)))
        self.nparams = n.value

    def close(self):
        if getattr(self, "h", None):
            lib().dgbgk_mlp_destroy(self.h)
            self.h = None

    __del__ = close

    def set_params(self, p):
        _check(lib().dgbgk_mlp_set_params(self.h, np.ascontiguousarray(p, dtype=np.float32)))

    def get_params(self) -> np.ndarray:
        p = np.empty(self.nparams, np.float32)
        _check(lib().dgbgk_mlp_get_params(self.h, p))
        return p

    def init(self, seed=0):
        """He-normal weights, zero biases."""
        rng = np.random.default_rng(seed)
        parts = []
        for l in range(self.hidden + 1):
            di = self.fin if l == 0 else self.width
            do = self.fout if l == self.hidden else self.width
            parts.append(rng.standard_normal(do * di) * np.sqrt(2.0 / di))
            parts.append(np.zeros(do))
        self.set_params(np.concatenate(parts).astype(np.float32))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((len(x), self.fout), np.float32)
        _check(lib().dgbgk_mlp_forward(self.h, len(x), x, y))
        return y

    def loss_grad(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = C.c_float()
        g = np.empty(self.nparams, np.float32)
        _check(lib().dgbgk_mlp_loss_grad(self.h, len(x), x, y, C.byref(loss), g))
        return loss.value, g

    def set_data(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        _check(lib().dgbgk_mlp_set_data(self.h, len(x), x, y))

    def train(self, steps, lr=1e-3, seed=0):
        losses = np.empty(steps, np.float32)
        _check(lib().dgbgk_mlp_train(self.h, steps, lr, seed, losses))
        return losses
