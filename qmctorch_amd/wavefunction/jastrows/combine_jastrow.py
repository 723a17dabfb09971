"""Product of Jastrow factors (jastrows/combine_jastrow.py:7-195).

State-dict keys follow the reference (``jastrow.jastrow_terms.<i>.jastrow_kernel.weight``);
the product rule is applied by summing the log-terms, inside the fused
wavefunction kernels or, for ``forward``, in ``qmc_jastrow_eval``.
"""
from typing import List

from torch import nn


class CombineJastrow(nn.Module):
    def __init__(self, jastrow: List[nn.Module]) -> None:
        super().__init__()
        self.jastrow_terms = nn.ModuleList()
        kinds = set()
        for j in jastrow:
            kind = getattr(j, "kind", None)
            if kind not in ("ee", "en") or kind in kinds:
                raise NotImplementedError("the device path combines at most one Pade e-e and one Pade e-n factor")
            kinds.add(kind)
            self.jastrow_terms.append(j)
        self.requires_autograd = True
        self.nterms = len(self.jastrow_terms)

    def __repr__(self):
        return " + ".join(t.__repr__() for t in self.jastrow_terms)

    def terms(self):
        return list(self.jastrow_terms)

    def forward(self, pos, derivative=0, sum_grad=True):
        """Product of the factors and its derivatives (combine_jastrow.py:33-106):
        J (Nbatch, 1); derivative=1 as the single factors; derivative=2 the
        Laplacian of the product; [0, 1, 2] returns the gradient per axis
        (Nbatch, 3, Nelec) whatever sum_grad is, as the reference does."""
        from ._device import evaluate, shaped

        ee = next((t for t in self.jastrow_terms if t.kind == "ee"), None)
        en = next((t for t in self.jastrow_terms if t.kind == "en"), None)
        mol = (ee or en).mol
        need = derivative if isinstance(derivative, list) else [derivative]
        j, dj, d2j = evaluate(self, mol, ee, en, pos, 1 in need, 2 in need)
        return shaped(j, dj, d2j, derivative, sum_grad, full_grad_in_tuple=True)
