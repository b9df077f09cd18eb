"""The learned CG warm start: a per-vertex network 15 -> 128 -> 128 -> 128 -> 3
(GELU) whose output seeds the PCG of each Newton step (sdx_mlp_* C ABI).

Feature rows, targets and parameter layout are documented in
include/sdx_fem.h.  Inference runs in BF16 on the matrix cores; training runs
in FP32 on the device with Adam.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi
from ._lib import check, lib

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "warm_start_mlp.npz")


def _f(a):
    return a.ctypes.data_as(_fp)


class WarmStartModel:
    def __init__(self, device: int = 0, seed: int = 0):
        m = C.c_void_p()
        check(lib().sdx_mlp_create(device, seed, C.byref(m)))
        self._m = m

    def close(self):
        if getattr(self, "_m", None):
            lib().sdx_mlp_destroy(self._m)
            self._m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(_abi.MLP_PARAMS, dtype=np.float32)
        check(lib().sdx_mlp_get_params(self._m, _f(p), p.size))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().sdx_mlp_set_params(self._m, _f(p), p.size))

    def forward(self, X, bf16: bool = True) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], _abi.MLP_OUT), dtype=np.float32)
        check(lib().sdx_mlp_forward(self._m, _f(X), X.shape[0], int(bf16), _f(Y)))
        return Y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(_abi.MLP_PARAMS, dtype=np.float32)
        loss = C.c_double()
        check(lib().sdx_mlp_loss_grad(self._m, _f(X), _f(T), X.shape[0], _f(g), C.byref(loss)))
        return loss.value, g

    @staticmethod
    def train_params(epochs=10, batch=4096, lr=1e-3, seed=1):
        tp = _abi.TrainParams()
        lib().sdx_default_train_params(C.byref(tp))
        tp.epochs, tp.batch, tp.lr, tp.seed = int(epochs), int(batch), float(lr), int(seed)
        return tp

    def train(self, X, T, epochs=10, batch=4096, lr=1e-3, seed=1) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        losses = np.zeros(epochs)
        check(lib().sdx_mlp_train(self._m, _f(X), _f(T), X.shape[0], C.byref(self.train_params(epochs, batch, lr, seed)),
                                  losses.ctypes.data_as(_dp)))
        return losses

    def train_collected(self, scene, epochs=10, batch=4096, lr=1e-3, seed=1) -> np.ndarray:
        losses = np.zeros(epochs)
        check(lib().sdx_mlp_train_collected(self._m, scene._w, C.byref(self.train_params(epochs, batch, lr, seed)),
                                            losses.ctypes.data_as(_dp)))
        return losses

    def save(self, path: str = DEFAULT_WEIGHTS):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.params)

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS, device: int = 0) -> "WarmStartModel":
        m = cls(device)
        with np.load(path, allow_pickle=False) as f:
            m.params = f["params"]
        return m
