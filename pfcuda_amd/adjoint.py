"""Reverse-mode rule of slog_pfaffian: grad log|Pf| = 1/2 A^-T.

The reference differentiates through jnp.linalg.inv (pfcuda/cuda_api.py:101);
this module computes the same quantity.
"""
from __future__ import annotations

import torch


def slog_grad(A: torch.Tensor) -> torch.Tensor:
    """1/2 A^-T for a batch (..., n, n)."""
    return 0.5 * torch.linalg.inv(A).mT
