"""GPU checks of the surrogate's BF16 MFMA kernels against the CPU oracle
(oracle/mlp_oracle.py): forward values (which also pins the MFMA lane
layout), the loss gradient of every parameter, and that training reduces the
loss."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as MO  # noqa: E402

pytestmark = pytest.mark.gpu


def _dev(a, dtype):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def _data(B, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((B, MO.NIN_PAD), np.float32)
    X[:, :MO.NIN] = rng.normal(size=(B, MO.NIN))
    Y = np.zeros((B, MO.NOUT_PAD), np.float32)
    Y[:, :MO.NOUT] = rng.normal(size=(B, MO.NOUT))
    return MO.bf16(X).astype(np.float32), Y


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("nhidden,nlayers", [(64, 1), (256, 4)])
def test_forward_matches_oracle(nhidden, nlayers):
    import torch
    from brams_amd.mlp import MLP
    print("case mlp forward %d x %d" % (nhidden, nlayers), flush=True)
    net = MLP(nhidden, nlayers, max_batch=256, seed=3)
    X, _ = _data(192, 0)
    out = torch.zeros((192, MO.NOUT_PAD), dtype=torch.float32, device="cuda")
    net.forward(_dev(X, torch.bfloat16), out)
    ref, _ = MO.forward(MO.unpack(net.get_params().astype(np.float64), nhidden, nlayers), X)
    got = out.cpu().numpy()
    assert _rel(got[:, :MO.NOUT], ref[:, :MO.NOUT]) < 1e-2
    assert np.all(got[:, MO.NOUT:] == 0.0)


def test_gradients_match_oracle():
    import torch
    from brams_amd.mlp import MLP
    nh, nl = 256, 4
    net = MLP(nh, nl, max_batch=512, seed=5)
    X, Y = _data(512, 1)
    loss, g = net.gradients(_dev(X, torch.bfloat16), _dev(Y, torch.float32))
    layers = MO.unpack(net.get_params().astype(np.float64), nh, nl)
    ref_loss, ref_g = MO.loss_and_grads(layers, X.astype(np.float64), Y.astype(np.float64))
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    got = MO.unpack(g.astype(np.float64), nh, nl)
    for l, ((gw, gb), (rw, rb)) in enumerate(zip(got, ref_g)):
        print("layer %d dW rel %.2e db rel %.2e" % (l, _rel(gw, rw), _rel(gb, rb)))
        assert _rel(gw, rw) < 2e-2, l
        assert _rel(gb, rb) < 2e-2, l


def test_training_reduces_loss():
    import torch
    from brams_amd.mlp import MLP
    rng = np.random.default_rng(2)
    N = 8192
    X = np.zeros((N, MO.NIN_PAD), np.float32)
    X[:, :MO.NIN] = rng.uniform(-1, 1, size=(N, MO.NIN))
    Y = np.zeros((N, MO.NOUT_PAD), np.float32)
    W = rng.normal(size=(MO.NIN, MO.NOUT)) / np.sqrt(MO.NIN)
    Y[:, :MO.NOUT] = np.tanh(X[:, :MO.NIN] @ W)
    net = MLP(256, 4, max_batch=1024, seed=9)
    losses = net.train(_dev(MO.bf16(X).astype(np.float32), torch.bfloat16), _dev(Y, torch.float32),
                       batch=1024, epochs=20, lr0=2e-3, seed=1)
    print("epoch losses", losses[:3], losses[-3:])
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.1 * losses[0]
