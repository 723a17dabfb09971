"""Precision switches and the minibatch loader used by the solver.

References: qmctorch/utils/torch_utils.py:8-21 (precision switches, TF32 off),
:156-207 (DataLoader over a position tensor).
"""
from math import ceil

import torch


def set_torch_double_precision() -> None:
    """Default dtype float64, TF32 off (torch_utils.py:8-13)."""
    torch.set_default_dtype(torch.float64)
    torch.backends.cuda.matmul.allow_tf32 = False
    torch.backends.cudnn.allow_tf32 = False


def set_torch_single_precision() -> None:
    """Default dtype float32, TF32 off (torch_utils.py:16-21)."""
    torch.set_default_dtype(torch.float32)
    torch.backends.cuda.matmul.allow_tf32 = False
    torch.backends.cudnn.allow_tf32 = False


class DataLoader:
    """Iterates a position tensor in contiguous minibatches.

    Same contract as qmctorch/utils/torch_utils.py:156-207: ``dataset`` may be
    re-assigned between epochs (the solver does so after resampling).
    """

    def __init__(self, data: torch.Tensor, batch_size: int, pin_memory: bool = False):
        self.dataset = data
        self.batch_size = batch_size
        self.pin_memory = pin_memory

    def __len__(self):
        return ceil(len(self.dataset) / self.batch_size)

    def __iter__(self):
        self.count = 0
        return self

    def __next__(self):
        if self.count * self.batch_size >= len(self.dataset):
            raise StopIteration
        lo = self.count * self.batch_size
        hi = min(lo + self.batch_size, len(self.dataset))
        self.count += 1
        return self.dataset[lo:hi]
