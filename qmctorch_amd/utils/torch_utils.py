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


def fast_power(x: torch.Tensor, k: torch.Tensor, mask0=None, mask2=None) -> torch.Tensor:
    """x ** k elementwise for integer exponents (torch_utils.py:24-59), with the
    exponents 0, 1, 2 taken without pow: 1, x and x * x.  mask0 / mask2 may
    carry precomputed ``k == 0`` / ``k == 2`` masks."""
    if int(k.max()) > 2:
        return x ** k
    out = x.clone()
    out.masked_fill_(k == 0 if mask0 is None else mask0, 1)
    sq = k == 2 if mask2 is None else mask2
    out[..., sq] = out[..., sq] * out[..., sq]
    return out


def gradients(out: torch.Tensor, inp: torch.Tensor) -> torch.Tensor:
    """d(sum out)/d inp, detached (torch_utils.py:62-74)."""
    return torch.autograd.grad(out, inp, grad_outputs=torch.ones_like(out))[0].detach()


def hessian(out: torch.Tensor, inp: torch.Tensor):
    """(H, g): the gradient g of sum(out) with respect to inp and the gradient
    of sum(g) -- the row sums of the Hessian (torch_utils.py:77-91)."""
    g = torch.autograd.grad(out, inp, grad_outputs=torch.ones_like(out), create_graph=True)[0]
    h = torch.autograd.grad(g, inp, grad_outputs=torch.ones_like(g))[0]
    return h, g.detach()


def diagonal_hessian(out: torch.Tensor, inp: torch.Tensor):
    """(diag H, g) for a batch inp of shape (N, D): d^2 out / d inp_d^2 per
    column d and the gradient, detached (torch_utils.py:94-125)."""
    g = torch.autograd.grad(out, inp, grad_outputs=torch.ones_like(out), create_graph=True)[0]
    diag = torch.zeros_like(g)
    for d in range(g.shape[1]):
        col = torch.autograd.grad(g[:, d], inp, grad_outputs=torch.ones_like(g[:, d]), create_graph=True)[0]
        diag[:, d] = col[:, d]
    return diag, g.detach()


class DataSet(torch.utils.data.Dataset):
    """Rows of a tensor as a torch Dataset (torch_utils.py:128-153)."""

    def __init__(self, data: torch.Tensor) -> None:
        self.data = data

    def __len__(self) -> int:
        return self.data.shape[0]

    def __getitem__(self, index: int) -> torch.Tensor:
        return self.data[index, :]
