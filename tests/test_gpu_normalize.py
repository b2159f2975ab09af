"""brams_sur_normalize against numpy on a dataset of more than 2^28
elements (2.2M rows x 128 columns): the launch caps its grid at 2^20
workgroups and loops, which must cover every row, including the zeroed
padding columns, in both output types.  (One thread per element over 4.8e9
work-items once wrapped the 32-bit grid size and left rows unnormalised.)"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NCOL, LDO = 2_200_000, 68, 128


@pytest.mark.parametrize("bf16", [False, True])
def test_normalize_large_matches_numpy(bf16):
    import torch
    from brams_amd import abi
    from brams_amd.surrogate import _bind
    print("case normalize n=%d ldo=%d bf16=%d (%.3g elements)" % (N, LDO, bf16, N * LDO), flush=True)
    assert N * LDO > 2 ** 28
    L = _bind()
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(7)
    raw = torch.randn((N, NCOL), generator=g, device=dev, dtype=torch.float32) * 3.0 + 1.5
    mu = torch.linspace(-1.0, 2.0, NCOL, device=dev, dtype=torch.float32)
    sd = torch.linspace(0.5, 4.0, NCOL, device=dev, dtype=torch.float32)
    out = torch.full((N, LDO), 7.0, device=dev, dtype=torch.bfloat16 if bf16 else torch.float32)
    abi.check(L.brams_sur_normalize(raw.data_ptr(), N, NCOL, mu.data_ptr(), sd.data_ptr(),
                                    out.data_ptr(), LDO, 1 if bf16 else 0), "normalize")
    torch.cuda.synchronize()
    rows = np.unique(np.concatenate([np.arange(0, 2048), np.arange(N // 2 - 1024, N // 2 + 1024),
                                     np.arange(N - 2048, N),
                                     np.random.default_rng(0).integers(0, N, 4096)]))
    idx = torch.from_numpy(rows).to(dev)
    got = out.index_select(0, idx).float().cpu().numpy()
    x = raw.index_select(0, idx).cpu().numpy().astype(np.float64)
    want = (x - mu.cpu().numpy()) / sd.cpu().numpy()
    tol = 8e-3 if bf16 else 1e-6
    np.testing.assert_allclose(got[:, :NCOL], want, rtol=tol, atol=tol)
    assert np.all(got[:, NCOL:] == 0.0)
    # every row was written: no sentinel value survives anywhere
    assert not bool((out[:, :NCOL].float() == 7.0).all(dim=1).any())
