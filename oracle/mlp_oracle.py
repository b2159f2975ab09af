"""ORACLE -- test infrastructure only.

FP64 numpy restatement of the surrogate MLP's forward pass and of the
gradient of its training loss, used to check the GPU kernels
(brams_amd/csrc/mlp_kernels.h, surrogate.hip).  It rounds to bfloat16 at the
same points as the GPU (weights, inputs, hidden activations and the
back-propagated deltas that feed the matrix cores) and accumulates in FP64,
so the remaining GPU/oracle difference is the GPU's FP32 accumulation.
With exact=True the rounding is off and the functions are the plain FP64
network and its exact loss gradient, which tests/test_mlp_oracle.py pins
with central finite differences.  The surrogate has no counterpart in the
reference.
"""
import numpy as np
from scipy.special import erf

NIN, NIN_PAD, NOUT, NOUT_PAD = 68, 128, 47, 64


def bf16(x):
    """Round float values to bfloat16 (round to nearest even), returned as float64."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def layer_dims(nhidden, nlayers):
    dims = []
    for l in range(nlayers + 1):
        din = NIN_PAD if l == 0 else nhidden
        dout = NOUT_PAD if l == nlayers else nhidden
        dims.append((din, dout))
    return dims


def unpack(flat, nhidden, nlayers):
    """Flat parameter vector (per layer W[out][in] then b[out]) -> [(W, b)]."""
    out, off = [], 0
    for din, dout in layer_dims(nhidden, nlayers):
        W = flat[off:off + din * dout].reshape(dout, din)
        off += din * dout
        b = flat[off:off + dout]
        off += dout
        out.append((np.asarray(W, np.float64), np.asarray(b, np.float64)))
    assert off == len(flat)
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def _identity(x):
    return np.array(x, dtype=np.float64)


def forward(layers, X, exact=False):
    """Returns the output and the per-layer (input activation, pre-activation)."""
    rnd = _identity if exact else bf16
    nl = len(layers)
    a = rnd(X)
    cache = []
    for l, (W, b) in enumerate(layers):
        Wb = rnd(W)
        if l == nl - 1:
            Wb[NOUT:] = 0.0
        z = a @ Wb.T + b
        cache.append((a, z))
        if l < nl - 1:
            a = rnd(gelu(z))
    return z, cache


def loss(layers, X, Y, exact=False):
    out, _ = forward(layers, X, exact)
    e = out[:, :NOUT] - Y[:, :NOUT]
    return np.sum(e * e) / (X.shape[0] * NOUT)


def loss_and_grads(layers, X, Y, exact=False):
    """Mean squared error over the first NOUT outputs and its gradient."""
    rnd = _identity if exact else bf16
    out, cache = forward(layers, X, exact)
    B = X.shape[0]
    e = out[:, :NOUT] - Y[:, :NOUT]
    loss = np.sum(e * e) / (B * NOUT)
    d = np.zeros_like(out)
    d[:, :NOUT] = 2.0 * e / (B * NOUT)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        a, z = cache[l]
        db_ = rnd(d)
        dW = db_.T @ a
        dbias = d.sum(axis=0)
        grads[l] = (dW, dbias)
        if l == 0:
            break
        W = rnd(layers[l][0])
        if l == len(layers) - 1:
            W[NOUT:] = 0.0
        da = db_ @ W
        d = da * gelu_grad(cache[l - 1][1])
    return loss, grads
