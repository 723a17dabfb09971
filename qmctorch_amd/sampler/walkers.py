"""Initial walker positions, drawn on the device (sampler/walkers.py:8-150).

``Walkers.initialize`` accepts the domains of ``Molecule.domain`` (``center``,
``uniform``: ``min``/``max``, ``normal``: ``mean``/``sigma`` (a covariance),
``atomic``: ``atom_coords``/``atom_num``/``atom_nelec``) or explicit positions.
The draw runs in one ``qmc_walkers_init`` launch, one thread per walker, with
Philox4x32-10 keyed by a seed taken from torch's global generator (so
``torch.manual_seed`` makes it reproducible) and counted by walker id, so a
walker's start does not depend on how many walkers are drawn with it.

The "atomic" domain places every electron near an atom: the electron-to-atom
assignment is the atom list (atom A repeated atom_nelec[A] times) in a random
order per walker, and each coordinate gets a Gaussian offset whose width
depends on how many electrons of the same atom came before it (in electron
order):

    rank on its atom   width
    first              1 / Z
    second to fifth    2 / (Z - 2)
    sixth and later    3 / (Z - 3)

Where a width would not be finite and positive (a second electron on H or He,
where the reference's formula divides by zero or goes negative) it is 1 bohr.
"""
import math
from typing import Dict, Optional, Union

import numpy as np
import torch

from .. import _native as N
from .. import log


def shell_widths(z: float):
    """Gaussian widths (first, second-fifth, sixth+) of the electrons placed on
    an atom of charge z."""
    out = []
    for num, shift in ((1.0, 0.0), (2.0, 2.0), (3.0, 3.0)):
        d = z - shift
        w = num / d if d > 0 else float("inf")
        out.append(w if math.isfinite(w) and w > 0 else 1.0)
    return out


class Walkers:
    def __init__(self, nwalkers: int = 100, nelec: int = 1, ndim: int = 3, init: Optional[Dict] = None,
                 cuda: bool = False) -> None:
        self.nwalkers = nwalkers
        self.ndim = ndim
        self.nelec = nelec
        self.init_domain = init
        self.pos = None
        self.status = None
        self.cuda = cuda
        self.device = torch.device("cuda" if cuda else "cpu")

    def initialize(self, pos: Union[None, torch.Tensor] = None, seed: Optional[int] = None) -> None:
        """Explicit positions (the last nwalkers rows when more are given), or a
        draw from the initialisation domain on the device."""
        if self.cuda:
            self.device = torch.device("cuda")
        if pos is not None:
            if len(pos) > self.nwalkers:
                pos = pos[-self.nwalkers:, :]
            self.pos = pos
            return
        log.debug("  Initialize walkers")
        self.pos = self.draw(self.init_domain, self.nwalkers, self.nelec, seed=seed)

    @staticmethod
    def describe(domain: Dict, nelec: int) -> N.QmcWalkerInit:
        """The kernel's description of an initialisation domain."""
        d = N.QmcWalkerInit()
        if "center" in domain:
            d.mode, d.lo, d.hi = N.QMC_INIT_CENTER, -1e-3, 1e-3
        elif "min" in domain:
            d.mode, d.lo, d.hi = N.QMC_INIT_UNIFORM, float(domain["min"]), float(domain["max"])
        elif "mean" in domain:
            d.mode = N.QMC_INIT_NORMAL
            mean = np.broadcast_to(np.asarray(domain["mean"], dtype=np.float64), (3,))
            cov = np.asarray(domain["sigma"], dtype=np.float64) * np.ones((3, 3)) * np.eye(3) \
                if np.ndim(domain["sigma"]) == 0 else np.asarray(domain["sigma"], dtype=np.float64)
            chol = np.linalg.cholesky(cov)
            for i in range(3):
                d.mean[i] = mean[i]
            for i in range(9):
                d.chol[i] = chol.reshape(-1)[i]
        elif "atom_coords" in domain:
            d.mode = N.QMC_INIT_ATOMIC
            xyz = np.asarray(domain["atom_coords"], dtype=np.float64).reshape(-1, 3)
            z = np.asarray(domain["atom_num"], dtype=np.float64)
            ne = np.asarray(domain["atom_nelec"], dtype=np.int64)
            if len(z) > N.QMC_INIT_MAXATOM:
                raise ValueError("at most %d atoms in an atomic initialisation domain" % N.QMC_INIT_MAXATOM)
            if int(ne.sum()) != nelec:
                raise ValueError("atom_nelec sums to %d, not to %d electrons" % (int(ne.sum()), nelec))
            d.natom = len(z)
            for a in range(len(z)):
                for k in range(3):
                    d.atom_xyz[3 * a + k] = xyz[a, k]
                for k, w in enumerate(shell_widths(z[a])):
                    d.atom_width[3 * a + k] = w
                d.atom_nelec[a] = int(ne[a])
        else:
            raise ValueError("Init walkers not recognized")
        return d

    def draw(self, domain: Dict, nwalkers: int, nelec: int, seed: Optional[int] = None,
             walker0: int = 0) -> torch.Tensor:
        d = self.describe(domain, nelec)
        d.seed = int(torch.randint(0, 2 ** 62, (1,)).item()) if seed is None else int(seed)
        d.walker0 = int(walker0)
        dev = torch.device("cuda")
        pos = torch.empty(nwalkers, 3 * nelec, dtype=torch.float64, device=dev)
        if nwalkers:
            N.check(N.load().qmc_walkers_init(d, nelec, nwalkers, pos.data_ptr(), N.current_stream()))
        out = pos.to(torch.get_default_dtype())
        return out if self.cuda else out.cpu()
