"""Product of Jastrow factors (jastrows/combine_jastrow.py:7-195).

State-dict keys follow the reference (``jastrow.jastrow_terms.<i>.jastrow_kernel.weight``);
the product rule is applied inside the fused kernels by summing the log-terms.
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
        raise NotImplementedError("combined Jastrow factors are evaluated inside the fused wavefunction kernels")
