"""spi_mlp_forward_noalloc (the actor MLP with its hidden activations in the library's static
device scratch instead of a caller workspace) against spi_mlp_forward with a caller workspace:
bit-identical outputs on two calls in a row, the second smaller than the first, and a clear
error for a size beyond the scratch."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _forward_with_workspace(L, x, M, dims_c, n_layers, params, out_dim):
    from spi_amd import _lib
    wf = int(L.spi_mlp_work_floats(M, n_layers, dims_c))
    work = torch.empty(max(wf, 1), dtype=torch.float32, device="cuda")
    y = torch.empty((M, out_dim), dtype=torch.float32, device="cuda")
    rc = L.spi_mlp_forward(_lib.ptr(x), M, n_layers, dims_c, _lib.ptr(params), _lib.ptr(work), wf, _lib.ptr(y),
                           _lib.stream())
    _lib.check(rc, "spi_mlp_forward")
    return y


def test_noalloc_equals_workspace_forward_twice(gpu):
    from spi_amd import _lib
    from spi_amd.agents.policy import PolicyMLP
    L = _lib.lib()
    dims = [900, 512, 256, 128, 12]
    net = PolicyMLP.random(dims[0], dims[-1], hidden=dims[1:-1], seed=11)
    dims_c = (C.c_int * len(dims))(*dims)
    n_layers = len(dims) - 1
    gen = torch.Generator().manual_seed(5)
    for M in (4096, 300):           # the largest the scratch holds, then a smaller call after it
        x = (torch.randn(M, dims[0], generator=gen) * 2).cuda()
        ref = _forward_with_workspace(L, x, M, dims_c, n_layers, net.params, dims[-1])
        y = torch.full((M, dims[-1]), float("nan"), dtype=torch.float32, device="cuda")
        rc = L.spi_mlp_forward_noalloc(_lib.ptr(x), M, n_layers, dims_c, _lib.ptr(net.params), _lib.ptr(y),
                                       _lib.stream())
        _lib.check(rc, "spi_mlp_forward_noalloc")
        torch.cuda.synchronize()
        assert torch.equal(y, ref), (M, (y - ref).abs().max().item())
    # beyond the scratch: refused on the host before any device work
    M = _lib.SPI_MLP_SCRATCH_FLOATS // (2 * 512) + 1
    x = torch.zeros((M, dims[0]), dtype=torch.float32, device="cuda")
    y = torch.zeros((M, dims[-1]), dtype=torch.float32, device="cuda")
    rc = L.spi_mlp_forward_noalloc(_lib.ptr(x), M, n_layers, dims_c, _lib.ptr(net.params), _lib.ptr(y), _lib.stream())
    assert rc == 3 and b"SPI_MLP_SCRATCH_FLOATS" in L.spi_last_error()
