"""ctypes binding of the surrogate MLP (include/brams_surrogate.h).

Device arrays are passed as raw pointers; torch tensors serve only as device
memory.  Every product runs in the library's BF16 MFMA kernels.
"""
import ctypes as C

import numpy as np

from . import abi

NIN, NIN_PAD, NOUT, NOUT_PAD = 68, 128, 47, 64


def _bind():
    L = abi.lib()
    if getattr(L, "_mlp_bound", False):
        return L
    L.brams_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]
    L.brams_mlp_destroy.argtypes = [C.c_void_p]
    L.brams_mlp_destroy.restype = None
    L.brams_mlp_num_params.argtypes = [C.c_void_p]
    L.brams_mlp_num_params.restype = C.c_int64
    L.brams_mlp_get_params.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_int64]
    L.brams_mlp_set_params.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_int64]
    L.brams_mlp_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    L.brams_mlp_gradients.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                      C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.brams_mlp_train.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int,
                                  C.c_float, C.c_uint64, C.POINTER(C.c_float)]
    L.brams_mlp_train_range.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_float, C.c_uint64, C.POINTER(C.c_float)]
    L._mlp_bound = True
    return L


class MLP:
    """One network of the ensemble (GELU hidden layers, FP32 master weights)."""

    def __init__(self, nhidden=256, nlayers=4, max_batch=8192, seed=0):
        L = _bind()
        h = C.c_void_p()
        abi.check(L.brams_mlp_create(nhidden, nlayers, max_batch, seed, C.byref(h)), "brams_mlp_create")
        self.h = h
        self.nhidden, self.nlayers, self.max_batch = nhidden, nlayers, max_batch

    def close(self):
        if self.h:
            _bind().brams_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self):
        return int(_bind().brams_mlp_num_params(self.h))

    def get_params(self):
        n = self.num_params
        out = np.zeros(n, np.float32)
        abi.check(_bind().brams_mlp_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float)), n), "get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        abi.check(_bind().brams_mlp_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float)), len(p)),
                  "set_params")

    def forward(self, X, out, stream=0):
        """X: device bf16 [M][128] (M multiple of 64), out: device f32 [M][64]."""
        abi.check(_bind().brams_mlp_forward(self.h, C.c_void_p(X.data_ptr()), X.shape[0],
                                            C.c_void_p(out.data_ptr()), C.c_void_p(stream)), "forward")

    def gradients(self, X, Y):
        n = self.num_params
        g = np.zeros(n, np.float32)
        loss = np.zeros(1, np.float32)
        abi.check(_bind().brams_mlp_gradients(self.h, C.c_void_p(X.data_ptr()), C.c_void_p(Y.data_ptr()),
                                              X.shape[0], g.ctypes.data_as(C.POINTER(C.c_float)),
                                              loss.ctypes.data_as(C.POINTER(C.c_float))), "gradients")
        return float(loss[0]), g

    def train(self, X, Y, batch=8192, epochs=1, lr0=1e-3, seed=0, chunk=None, log=None):
        """Adam with a cosine schedule over `epochs`; with `chunk`, the run is
        split into calls of that many epochs and `log` reports each."""
        losses = np.zeros(epochs, np.float32)
        chunk = chunk or epochs
        for e0 in range(0, epochs, chunk):
            n = min(chunk, epochs - e0)
            out = np.zeros(n, np.float32)
            abi.check(_bind().brams_mlp_train_range(
                self.h, C.c_void_p(X.data_ptr()), C.c_void_p(Y.data_ptr()), X.shape[0], batch, e0, n,
                epochs, lr0, seed, out.ctypes.data_as(C.POINTER(C.c_float))), "train")
            losses[e0:e0 + n] = out
            if log:
                log("  epochs %d-%d of %d: loss %.5g" % (e0 + 1, e0 + n, epochs, out[-1]))
        return losses
