"""The torch.library operators without a GPU: registration, the shape-only
implementations (meta tensors) and the batching rules' dimension handling.
No kernel runs here; meta inputs dispatch to the shape-only functions."""
import pytest
import torch

import pfcuda_amd  # noqa: F401  (registers the operators)

META = "meta"
OPS = ("pfaffian", "slog_pfaffian", "slog_pfaffian_grad", "slog_pfaffian_ragged")


def test_operators_registered():
    for name in OPS:
        assert hasattr(torch.ops.pfcuda, name), name


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.complex64, torch.complex128])
def test_shapes_and_dtypes(dtype):
    A = torch.empty(2, 3, 40, 40, dtype=dtype, device=META)
    la, ph = torch.ops.pfcuda.slog_pfaffian(A, 0)
    assert la.shape == (2, 3) and la.dtype == torch.float64
    assert ph.shape == (2, 3) and ph.dtype == dtype
    p = torch.ops.pfcuda.pfaffian(A)
    assert p.shape == (2, 3) and p.dtype == dtype
    G = torch.ops.pfcuda.slog_pfaffian_grad(A)
    assert G.shape == A.shape and G.dtype == dtype
    data = torch.empty(1000, dtype=dtype, device=META)
    off = torch.empty(7, dtype=torch.int64, device=META)
    siz = torch.empty(7, dtype=torch.int32, device=META)
    rl, rp = torch.ops.pfcuda.slog_pfaffian_ragged(data, off, siz, 12)
    assert rl.shape == (7,) and rl.dtype == torch.float64 and rp.shape == (7,) and rp.dtype == dtype


@pytest.mark.parametrize("in_dim", [0, 1, 2])
def test_batching_rules_move_the_mapped_dim(in_dim):
    # map over a dimension of size 3 placed at in_dim of a (5, n, n) batch
    X = torch.empty(3, 5, 40, 40, dtype=torch.float64, device=META).movedim(0, in_dim)
    la, ph = torch.func.vmap(lambda a: torch.ops.pfcuda.slog_pfaffian(a, 0), in_dims=in_dim)(X)
    assert la.shape == (3, 5) and ph.shape == (3, 5)
    assert torch.func.vmap(torch.ops.pfcuda.pfaffian, in_dims=in_dim)(X).shape == (3, 5)
    assert torch.func.vmap(torch.ops.pfcuda.slog_pfaffian_grad, in_dims=in_dim)(X).shape == (3, 5, 40, 40)
    # nested maps fold into one batch
    Y = torch.empty(2, 3, 40, 40, dtype=torch.float64, device=META)
    assert torch.func.vmap(torch.func.vmap(torch.ops.pfcuda.pfaffian))(Y).shape == (2, 3)


def test_unmapped_input_passes_through():
    A = torch.empty(4, 40, 40, dtype=torch.float64, device=META)
    other = torch.empty(6, device=META)
    la, ph = torch.func.vmap(lambda o: torch.ops.pfcuda.slog_pfaffian(A, 0))(other)
    assert la.shape == (6, 4)
