"""Batched small-matrix helpers of the reference's public utils
(qmctorch/utils/algebra_utils.py:7-55)."""
from typing import List

import numpy as np
import torch


def btrace(M: torch.Tensor) -> torch.Tensor:
    """Trace over the last two dimensions: (..., N, N) -> (...)."""
    return M.diagonal(dim1=-2, dim2=-1).sum(-1)


def bproj(M: torch.Tensor, P: torch.Tensor) -> torch.Tensor:
    """P^T M P for batched matrices."""
    return P.transpose(-1, -2) @ M @ P


def bdet2(M: torch.Tensor) -> torch.Tensor:
    """Determinants of batched 2 x 2 matrices."""
    return M[..., 0, 0] * M[..., 1, 1] - M[..., 0, 1] * M[..., 1, 0]


def double_factorial(values: List) -> np.ndarray:
    """n!! elementwise, with (-1)!! = 0!! = 1."""
    out = []
    for n in np.asarray(values, dtype=np.int64).ravel():
        r = 1
        while n > 1:
            r *= int(n)
            n -= 2
        out.append(r)
    return np.asarray(out).reshape(np.shape(values))
