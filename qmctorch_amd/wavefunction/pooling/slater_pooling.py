"""Slater determinant pooling (pooling/slater_pooling.py:11-849).

Keeps the configuration bookkeeping of the reference (``configs``, ``nconfs``,
unique configurations per spin).  Inside a wavefunction the determinants,
inverses and traces are computed by the fused kernels; ``forward`` and
``operator`` evaluate them on their own from caller-given MO matrices
(``qmc_slater_pool``: row-per-lane FP64 Gauss-Jordan), with the reference's
output shapes.  Every configuration method gives the explicit determinants and
traces, which the reference's single/double-excitation shortcuts equal
(tests/wavefunction/pooling/test_slater.py, test_trace_trick.py).
"""
import operator as op
from typing import Callable, Optional, Tuple

import numpy as np
import torch
from torch import nn

from ... import _native as N
from ..param_views import f64
from ..plan import component_plan


class SlaterPooling(nn.Module):
    def __init__(self, config_method, configs, mol, cuda: bool = False) -> None:
        super().__init__()
        self.config_method = config_method
        self.configs = configs
        self.nconfs = len(configs[0])
        self.nmo = int(mol.basis.nmo)
        self.nup, self.ndown = mol.nup, mol.ndown
        self.nelec = self.nup + self.ndown
        self.index_max_orb_up = int(configs[0].max()) + 1 if configs[0].numel() else 0
        self.index_max_orb_down = int(configs[1].max()) + 1 if configs[1].numel() else 0
        self.use_explicit_operator = False
        self.mol = mol
        self.device = torch.device("cuda" if cuda else "cpu")
        self._plan = None

    def _device_plan(self):
        if self._plan is None:
            cfg = (np.asarray(self.configs[0], dtype=np.int64), np.asarray(self.configs[1], dtype=np.int64))
            self._plan = component_plan(self.mol, configs=cfg)
        return self._plan

    def _pool(self, mo: torch.Tensor, bop: Optional[torch.Tensor], op_squared: bool, need_det: bool):
        N.require_device(mo, "MO matrix")
        if mo.dim() != 3 or mo.shape[1] != self.nelec:
            raise ValueError("MO matrices must have shape (Nbatch, %d, Nmo)" % self.nelec)
        plan = self._device_plan()
        m = f64(mo)
        nw, ld = m.shape[0], m.shape[2]
        if ld < plan.sizes().nmo:
            raise ValueError("the configurations use MO columns up to %d" % (plan.sizes().nmo - 1))
        nop = 1
        b = None
        if bop is not None:
            b = f64(bop)
            if b.dim() == 4:
                nop = b.shape[0]
            if b.shape[-3:] != m.shape[-3:] or (b.dim() == 4 and b.shape[1] != nw) or b.dim() not in (3, 4):
                raise ValueError("the operator must have the MO shape, or (Nop,) + that shape")
        dev = m.device
        det = torch.empty(nw, self.nconfs, dtype=torch.float64, device=dev) if need_det else None
        up = torch.empty(nop, nw, self.nconfs, dtype=torch.float64, device=dev) if b is not None else None
        dn = torch.empty_like(up) if b is not None else None
        if nw:
            N.check(N.load().qmc_slater_pool(plan.handle, m.data_ptr(), ld, N.ptr(b), nop, 1 if op_squared else 0,
                                             nw, N.ptr(det), N.ptr(up), N.ptr(dn), N.current_stream()))
        return det, up, dn, (b is not None and b.dim() == 4)

    def forward(self, input: torch.Tensor) -> torch.Tensor:
        """Determinant products D_up D_down of every configuration, (Nbatch, nconfs)
        (slater_pooling.py:64-111)."""
        det, _, _, _ = self._pool(input, None, False, True)
        return det.to(torch.get_default_dtype())

    def operator(self, mo: torch.Tensor, bop: torch.Tensor, op: Optional[Callable] = op.add,
                 op_squared: bool = False, inv_mo: Optional[Tuple[torch.Tensor, torch.Tensor]] = None):
        """tr(A^-1 B) (or tr((A^-1 B)^2)) per spin and configuration, combined by
        op (slater_pooling.py:262-409): (Nbatch, nconfs) for one operator,
        (Nop, Nbatch, nconfs) for a stack of operators (Nop, Nbatch, Nelec, Nmo).
        The ground-state method returns (Nbatch, 1).  inv_mo is accepted for
        signature compatibility; the inverses are recomputed on the device."""
        _, up, dn, multi = self._pool(mo, bop, op_squared, False)
        dt = torch.get_default_dtype()
        up, dn = up.to(dt), dn.to(dt)
        if not multi:
            up, dn = up[0], dn[0]
        if op is not None:
            return op(up, dn)
        return up, dn
