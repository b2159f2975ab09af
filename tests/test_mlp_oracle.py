"""Pins of the MLP oracle (oracle/mlp_oracle.py) on the CPU.

With its bf16 rounding off, the oracle's loss gradient must equal central
finite differences of its own FP64 loss (a dropped term, a transposed
product or a wrong GELU derivative fails this), and GELU must meet its
closed forms.  The GPU kernels are then checked against this oracle in
tests/test_gpu_surrogate.py."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as MO  # noqa: E402


def _net(nh, nl, seed):
    rng = np.random.default_rng(seed)
    layers = []
    for din, dout in MO.layer_dims(nh, nl):
        W = rng.normal(size=(dout, din)) / np.sqrt(din)
        b = rng.normal(size=dout) * 0.1
        if dout == MO.NOUT_PAD:
            W[MO.NOUT:] = 0.0
            b[MO.NOUT:] = 0.0
        layers.append((W, b))
    X = np.zeros((6, MO.NIN_PAD))
    X[:, :MO.NIN] = rng.normal(size=(6, MO.NIN))
    Y = np.zeros((6, MO.NOUT_PAD))
    Y[:, :MO.NOUT] = rng.normal(size=(6, MO.NOUT))
    return layers, X, Y


def test_gradient_matches_central_differences():
    nh, nl = 16, 3
    layers, X, Y = _net(nh, nl, 0)
    _, grads = MO.loss_and_grads(layers, X, Y, exact=True)
    flat = MO.pack(layers)
    gflat = MO.pack(grads)
    rng = np.random.default_rng(1)
    # entries of every layer's W (valid rows, real inputs) and b
    idx, off = [], 0
    for din, dout in MO.layer_dims(nh, nl):
        nin = MO.NIN if din == MO.NIN_PAD else din
        nout = MO.NOUT if dout == MO.NOUT_PAD else dout
        for _ in range(12):
            idx.append(off + rng.integers(nout) * din + rng.integers(nin))
        off += din * dout
        for _ in range(4):
            idx.append(off + rng.integers(nout))
        off += dout
    h = 1e-6
    for i in idx:
        p = flat.copy()
        p[i] += h
        lp = MO.loss(MO.unpack(p, nh, nl), X, Y, exact=True)
        p[i] -= 2 * h
        lm = MO.loss(MO.unpack(p, nh, nl), X, Y, exact=True)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gflat[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, gflat[i])


def test_padded_outputs_carry_no_gradient():
    layers, X, Y = _net(16, 2, 2)
    _, grads = MO.loss_and_grads(layers, X, Y, exact=True)
    W, b = grads[-1]
    assert np.all(W[MO.NOUT:] == 0.0) and np.all(b[MO.NOUT:] == 0.0)


def test_gelu_closed_forms():
    z = np.array([-8.0, -1.0, 0.0, 1.0, 8.0])
    g = MO.gelu(z)
    assert g[2] == 0.0
    assert abs(g[4] - 8.0) < 1e-12 and abs(g[0]) < 1e-12
    # GELU(x) - GELU(-x) = x (Phi(x) + Phi(-x) = 1)
    np.testing.assert_allclose(g[3] - g[1], 1.0, rtol=1e-15)
    # Phi(1) = 0.841344746068543 (normal table)
    np.testing.assert_allclose(g[3], 0.841344746068543, rtol=1e-12)
    h = 1e-6
    zz = np.linspace(-4, 4, 17)
    fd = (MO.gelu(zz + h) - MO.gelu(zz - h)) / (2 * h)
    np.testing.assert_allclose(MO.gelu_grad(zz), fd, rtol=1e-7, atol=1e-9)


def test_bf16_rounding_is_nearest_even():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to the even mantissa (1)
    assert MO.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0
    assert MO.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6
    assert MO.bf16(np.array([1.0 + 2.0 ** -7]))[0] == 1.0 + 2.0 ** -7
