"""Device-resident parameter views of one plan (the QmcParams layouts).

One ``qmc_params_gather`` launch per evaluation rebuilds W = mo_scf *
mo_modifier (FP64 and FP32, plain and in AO-entry order) and the exponents in
group order from the live parameters (molecular_orbitals.py:65-95,
atomic_orbitals.py:53-55), so optimiser updates -- in place, through ``.data``
or by ``load_state_dict`` -- are always seen, without host work or a cache
that could go stale.  The buffers are reused from call to call; calls of one
wavefunction are ordered on torch's current stream.
"""
import ctypes

import torch

from .. import _native as N


class ParamViews:
    def __init__(self, plan, device: torch.device):
        self.plan = plan
        sz = plan.sizes()
        f64, f32 = torch.float64, torch.float32
        self.mo_w64 = torch.empty(sz.nao * sz.nmo, dtype=f64, device=device)
        self.mo_w32 = torch.empty(sz.nao * sz.nmo, dtype=f32, device=device)
        self.entry_w64 = torch.empty(sz.nentry * sz.bucket, dtype=f64, device=device)
        self.entry_w32 = torch.empty(sz.nentry * sz.bucket, dtype=f32, device=device)
        self.grp_exp = torch.empty(sz.nexp_grp, dtype=f64, device=device)
        self.grp_expf = torch.empty(sz.nexp_grp, dtype=f32, device=device)
        v = N.QmcParamViews()
        for name in ("mo_w64", "mo_w32", "entry_w64", "entry_w32", "grp_exp", "grp_expf"):
            setattr(v, name, getattr(self, name).data_ptr())
        self.views = v

    def gather(self, mo_scf: torch.Tensor, mo_modifier: torch.Tensor, bas_exp: torch.Tensor) -> None:
        N.check(N.load().qmc_params_gather(self.plan.handle, mo_scf.data_ptr(), mo_modifier.data_ptr(),
                                           mo_scf.shape[1], bas_exp.data_ptr(), ctypes.byref(self.views),
                                           N.current_stream()))


def f64(t: torch.Tensor) -> torch.Tensor:
    """float64 contiguous view of a parameter (no copy in the default FP64 setting)."""
    t = t.detach()
    if t.dtype != torch.float64:
        t = t.to(torch.float64)
    return t.contiguous()
