"""Device evaluation of Pade Jastrow factors for the module-level API.

``evaluate`` runs ``qmc_jastrow_eval`` for an e-e factor, an e-n factor or
their product and returns J (Nbatch, 1), dJ (Nbatch, 3, Nelec) and lap J
(Nbatch, Nelec) in the reference's layouts
(jastrow_factor_electron_electron.py:124-260, jastrow_factor_electron_nuclei.py:60-161,
combine_jastrow.py:33-195).  The values are not differentiable through
autograd: parameter gradients of the wavefunction come from its adjoint kernels.
"""
import numpy as np
import torch

from ... import _native as N
from ..param_views import f64
from ..plan import component_plan


def _plan_for(owner, mol, use_ee: bool, use_en: bool):
    key = (use_ee, use_en)
    if getattr(owner, "_jas_plan_key", None) != key:
        owner._jas_plan = component_plan(mol, use_ee=use_ee, use_en=use_en)
        owner._jas_plan_key = key
    return owner._jas_plan


def evaluate(owner, mol, ee, en, pos: torch.Tensor, need_grad: bool, need_lap: bool):
    """(J, dJ or None, lap J or None) of the factors ee and/or en (modules
    holding a ``jastrow_kernel.weight``), float64 device tensors."""
    N.require_device(pos, "positions")
    nel = mol.nup + mol.ndown
    if pos.dim() != 2 or pos.shape[1] != 3 * nel:
        raise ValueError("positions must have shape (Nbatch, %d)" % (3 * nel))
    x = f64(pos)
    nw = x.shape[0]
    dev = x.device
    plan = _plan_for(owner, mol, ee is not None, en is not None)
    keep = {
        "w_ee": f64(ee.jastrow_kernel.weight).to(dev) if ee is not None else None,
        "w_en": f64(en.jastrow_kernel.weight).to(dev) if en is not None else None,
        "atoms": torch.as_tensor(np.asarray(mol.atom_coords, dtype=np.float64), device=dev).contiguous(),
    }
    prm = N.QmcParams()
    prm.w_ee, prm.w_en, prm.atom_coords = N.ptr(keep["w_ee"]), N.ptr(keep["w_en"]), keep["atoms"].data_ptr()
    j = torch.empty(nw, dtype=torch.float64, device=dev)
    dj = torch.empty(nw, 3, nel, dtype=torch.float64, device=dev) if need_grad else None
    d2j = torch.empty(nw, nel, dtype=torch.float64, device=dev) if need_lap else None
    if nw:
        o = N.QmcJastrowOut()
        o.value, o.grad, o.lap = j.data_ptr(), N.ptr(dj), N.ptr(d2j)
        terms = (1 if ee is not None else 0) | (2 if en is not None else 0)
        N.check(N.load().qmc_jastrow_eval(plan.handle, prm, terms, x.data_ptr(), nw, o, N.current_stream()))
    return j.view(-1, 1), dj, d2j


def shaped(j, dj, d2j, derivative, sum_grad, full_grad_in_tuple=False):
    """Outputs in the reference's forms for derivative 0, 1, 2 or [0, 1, 2]."""
    dt = torch.get_default_dtype()

    def grad_out(force_full=False):
        return (dj if (force_full or not sum_grad) else dj.sum(1)).to(dt)

    if derivative == 0:
        return j.to(dt)
    if derivative == 1:
        return grad_out()
    if derivative == 2:
        return d2j.to(dt)
    if derivative == [0, 1, 2]:
        return j.to(dt), grad_out(full_grad_in_tuple), d2j.to(dt)
    raise ValueError("derivative not understood")
