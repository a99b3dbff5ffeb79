"""Learned cache (hash grid + MLP on MFMA) over the C ABI."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import check, lib

N_WEIGHTS = 15360  # the default 4 x 64 MLP


def n_weights(n_neurons: int = 64, n_hidden_layers: int = 4) -> int:
    """Weight-buffer length of the MLP: A1 [H x 32], the hidden layers
    [H x H], the output block [16 x H] (row 0 live)."""
    h = n_neurons
    return 32 * h + (n_hidden_layers - 1) * h * h + 16 * h


class CacheConfigC(C.Structure):
    _fields_ = [("n_levels", C.c_int), ("log2_table_size", C.c_int), ("base_resolution", C.c_int),
                ("growth", C.c_float), ("origin", C.c_float * 3), ("extent", C.c_float * 3),
                ("learning_rate", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float),
                ("loss_scale", C.c_float), ("seed", C.c_uint64), ("max_batch", C.c_int), ("n_neurons", C.c_int),
                ("n_hidden_layers", C.c_int)]


@dataclass
class CacheConfig:
    """Defaults: the learned-cache problem (16 levels x 2 features, T = 2^19,
    N_min = 16, b = 1.5; Adam lr 1e-2, beta 0.9 / 0.99, eps 1e-15)."""
    n_levels: int = 16
    log2_table_size: int = 19
    base_resolution: int = 16
    growth: float = 1.5
    origin: tuple = (-1.1, -1.1, -1.1)
    extent: tuple = (2.2, 2.2, 2.2)
    learning_rate: float = 1e-2
    beta1: float = 0.9
    beta2: float = 0.99
    epsilon: float = 1e-15
    loss_scale: float = 128.0
    seed: int = 1
    max_batch: int = 1 << 18  # largest training batch (activation store allocated at creation)
    n_neurons: int = 64       # MLP width: 16, 32 or 64
    n_hidden_layers: int = 4  # MLP depth: 1 to 4

    @classmethod
    def for_bounds(cls, lo, hi, pad: float = 0.05, **kw):
        """Input box = mesh bounds padded by 5% per side (eval_tcnn_nc_wos.cu:522-523)."""
        lo = np.asarray(lo, np.float64)
        hi = np.asarray(hi, np.float64)
        ext = hi - lo
        return cls(origin=tuple(lo - pad * ext), extent=tuple(ext * (1 + 2 * pad)), **kw)

    def to_c(self) -> CacheConfigC:
        c = CacheConfigC()
        c.n_levels, c.log2_table_size, c.base_resolution, c.growth = (self.n_levels, self.log2_table_size,
                                                                      self.base_resolution, self.growth)
        for a in range(3):
            c.origin[a] = self.origin[a]
            c.extent[a] = self.extent[a]
        c.learning_rate, c.beta1, c.beta2, c.epsilon = self.learning_rate, self.beta1, self.beta2, self.epsilon
        c.loss_scale, c.seed, c.max_batch = self.loss_scale, self.seed, self.max_batch
        c.n_neurons, c.n_hidden_layers = self.n_neurons, self.n_hidden_layers
        return c


class Cache:
    def __init__(self, config: CacheConfig | None = None):
        self.config = config or CacheConfig()
        self.output_offset = 0.0
        h = C.c_void_p()
        self._cfg = self.config.to_c()
        check(lib().n2wos_cache_create(C.byref(self._cfg), C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().n2wos_cache_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def layout(self) -> dict:
        nl = C.c_int()
        t = C.c_uint32()
        tot = C.c_uint32()
        off = (C.c_uint32 * 16)()
        sz = (C.c_uint32 * 16)()
        res = (C.c_int * 16)()
        dense = (C.c_int * 16)()
        check(lib().n2wos_cache_layout(self.handle, C.byref(nl), C.byref(t), off, sz, res, dense, C.byref(tot)))
        return {"n_levels": nl.value, "table_size": t.value, "offsets": list(off), "sizes": list(sz),
                "res": list(res), "dense": list(dense), "total_entries": tot.value}

    def set_output_offset(self, offset: float) -> None:
        """Constant added to the network output in inference and training
        (normalise the targets by their mean before training)."""
        check(lib().n2wos_cache_set_output_offset(self.handle, float(offset)))
        self.output_offset = float(offset)

    def copy_from(self, other: "Cache", stream=None) -> None:
        """Queue a device copy of other's parameters into this cache (same
        layout) on a stream: a frozen snapshot while other keeps training."""
        import torch
        s = torch.cuda.current_stream().cuda_stream if stream is None else getattr(stream, "cuda_stream", stream)
        check(lib().n2wos_cache_copy_params(self.handle, other.handle, s))
        self.output_offset = getattr(other, "output_offset", 0.0)

    def infer(self, points, out=None, stream=None):
        import torch
        n = points.shape[0]
        if out is None:
            out = torch.empty(n, dtype=torch.float32, device=points.device)
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        check(lib().n2wos_cache_infer(self.handle, points.data_ptr(), n, out.data_ptr(), s))
        return out

    def train_step(self, points, targets, loss_out=None, stream=None):
        import torch
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        check(lib().n2wos_cache_train_step(self.handle, points.data_ptr(), targets.data_ptr(), points.shape[0],
                                           loss_out.data_ptr() if loss_out is not None else None, s))

    def get_params(self):
        lay = self.layout()
        grid = np.zeros(2 * lay["total_entries"], np.float32)
        w = np.zeros(n_weights(self.config.n_neurons, self.config.n_hidden_layers), np.float32)
        check(lib().n2wos_cache_get_params(self.handle, grid.ctypes.data, w.ctypes.data))
        return grid, w

    def set_params(self, grid=None, weights=None):
        g = np.ascontiguousarray(grid, np.float32) if grid is not None else None
        w = np.ascontiguousarray(weights, np.float32) if weights is not None else None
        check(lib().n2wos_cache_set_params(self.handle, g.ctypes.data if g is not None else None,
                                           w.ctypes.data if w is not None else None))

    def gradients(self, points, targets):
        lay = self.layout()
        gg = np.zeros(2 * lay["total_entries"], np.float32)
        wg = np.zeros(n_weights(self.config.n_neurons, self.config.n_hidden_layers), np.float32)
        loss = np.zeros(1, np.float32)
        check(lib().n2wos_cache_gradients(self.handle, points.data_ptr(), targets.data_ptr(), points.shape[0],
                                          gg.ctypes.data, wg.ctypes.data, loss.ctypes.data))
        return gg, wg, float(loss[0])
