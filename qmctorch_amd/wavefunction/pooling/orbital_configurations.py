"""CI configurations (pooling/orbital_configurations.py:6-402 of the reference).

ground_state, single(n,m), single_double(n,m), cas(n,m) and explicit tuples,
with the reference's ordering and its in-place "replace" excitation
convention (:255-290), so configuration lists and therefore ``fc.weight``
indices match the reference one for one.
"""
from itertools import combinations, product
from typing import List, Tuple

import torch


class OrbitalConfigurations:
    def __init__(self, mol) -> None:
        self.nup = mol.nup
        self.ndown = mol.ndown
        self.nelec = self.nup + self.ndown
        self.spin = getattr(mol, "spin", self.nup - self.ndown)
        self.norb = int(mol.basis.nmo)

    def get_configs(self, configs) -> Tuple[torch.LongTensor, torch.LongTensor]:
        if isinstance(configs, tuple):
            assert len(configs) == 2 and configs[0].shape == configs[1].shape
            return configs
        configs = configs.lower()
        if configs == "ground_state":
            return self._get_ground_state_config()
        for name, fn in (("cas(", self._cas), ("single_double(", self._single_double), ("single(", self._single)):
            if configs.startswith(name):
                nel, norb = [int(v) for v in configs[len(name) : -1].split(",")]
                self.sanity_check(nel, norb)
                nocc, nvirt = self._get_orb_number(nel, norb)
                return fn(nocc, nvirt, nel)
        raise ValueError("Config error: %s not recognized" % configs)

    def sanity_check(self, nelec: int, norb: int) -> None:
        if nelec > self.nelec:
            raise ValueError("required number of electron in config too large")
        if norb > self.norb:
            raise ValueError("required number of orbitals in config too large")

    def _get_ground_state_config(self):
        return torch.LongTensor([list(range(self.nup))]), torch.LongTensor([list(range(self.ndown))])

    @staticmethod
    def _get_orb_number(nelec: int, norb: int):
        nocc = (nelec // 2, nelec // 2) if nelec % 2 == 0 else (nelec // 2 + 1, nelec // 2)
        return nocc, (norb - nocc[0], norb - nocc[1])

    @staticmethod
    def _excite(conf: List[int], iocc: int, ivirt: int) -> List[int]:
        conf = list(conf)
        conf[iocc] = ivirt
        return conf

    def _single(self, nocc, nvirt, nel=None):
        gu, gd = list(range(self.nup)), list(range(self.ndown))
        cup, cdn = [gu], [gd]
        for io in range(self.nup - 1, self.nup - 1 - nocc[0], -1):
            for iv in range(self.nup, self.nup + nvirt[0]):
                cup.append(self._excite(gu, io, iv))
                cdn.append(gd)
        for io in range(self.ndown - 1, self.ndown - 1 - nocc[1], -1):
            for iv in range(self.ndown, self.ndown + nvirt[1]):
                cup.append(gu)
                cdn.append(self._excite(gd, io, iv))
        return torch.LongTensor(cup), torch.LongTensor(cdn)

    def _single_double(self, nocc, nvirt, nel=None):
        gu, gd = list(range(self.nup)), list(range(self.ndown))
        cu, cd = self._single(nocc, nvirt)
        cup, cdn = cu.tolist(), cd.tolist()
        ou = list(range(self.nup - 1, self.nup - 1 - nocc[0], -1))
        vu = list(range(self.nup, self.nup + nvirt[0]))
        od = list(range(self.ndown - 1, self.ndown - 1 - nocc[1], -1))
        vd = list(range(self.ndown, self.ndown + nvirt[1]))
        for a in ou:
            for b in vu:
                for c in od:
                    for d in vd:
                        cup.append(self._excite(gu, a, b))
                        cdn.append(self._excite(gd, c, d))
        for o1, o2 in combinations(ou, 2):
            for v1, v2 in combinations(vu, 2):
                cup.append(self._excite(self._excite(gu, o1, v2), o2, v1))
                cdn.append(gd)
        for o1, o2 in combinations(od, 2):
            for v1, v2 in combinations(vd, 2):
                cup.append(gu)
                cdn.append(self._excite(self._excite(gd, o1, v2), o2, v1))
        return torch.LongTensor(cup), torch.LongTensor(cdn)

    def _cas(self, nocc, nvirt, nelec):
        if self.spin != 0:
            raise ValueError("CAS active space not possible with spin polarized calculation")
        lo, hi = self.nup - nocc[0], self.nup + nvirt[0]
        frz = list(range(lo))
        cu = [frz + list(c) for c in combinations(range(lo, hi), nelec // 2)]
        cd = [frz + list(c) for c in combinations(range(lo, hi), nelec // 2)]
        pairs = list(product(cu, cd))
        return torch.LongTensor([p[0] for p in pairs]), torch.LongTensor([p[1] for p in pairs])
