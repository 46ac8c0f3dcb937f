"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_fixtures.py): the oracle must still reproduce them on
the CPU, and the HIP library must reproduce them on the GPU, bit for bit
(any NaN matching any NaN)."""
import os

import numpy as np
import pytest

from oracle import oracle
from strata_amd.engine import ProcessN
from strata_amd.raster import Float32Raster

from golden.make_fixtures import compute, kernels, window

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KERNELS = kernels()


def load_inputs():
    z = np.load(os.path.join(HERE, "inputs.npz"))
    return z["data"], z["valid"]


def same(a, b):
    return ((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))).all()


@pytest.mark.parametrize("name", sorted(KERNELS))
def test_oracle_reproduces_fixture(name):
    data, valid = load_inputs()
    want = np.load(os.path.join(HERE, f"{name}.npz"))
    outs = compute(KERNELS[name], data, valid)
    for i, o in enumerate(outs):
        assert same(o.Data, want[f"data{i}"])
        assert np.array_equal(o.Valid, want[f"valid{i}"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(KERNELS))
def test_gpu_reproduces_fixture(gpu, name):
    import torch
    data, valid = load_inputs()
    want = np.load(os.path.join(HERE, f"{name}.npz"))
    src = window(torch.from_numpy(data).to(gpu), torch.from_numpy(valid.view(np.int64)).to(gpu))
    k = KERNELS[name]
    n = src.Width * src.Height
    outs = [Float32Raster(torch.zeros(n, dtype=torch.float32, device=gpu), src.Width, src.Height, src.Width,
                          torch.zeros((n + 63) // 64, dtype=torch.int64, device=gpu), 0) for _ in range(k.nout)]
    assert ProcessN(None, outs, [src], k) is None
    torch.cuda.synchronize()
    for i, o in enumerate(outs):
        assert same(o.Data.cpu().numpy(), want[f"data{i}"])
        assert np.array_equal(o.Valid.cpu().numpy().view(np.uint64), want[f"valid{i}"])
