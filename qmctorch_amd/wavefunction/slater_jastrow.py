"""Slater-Jastrow wavefunction on MI355X (wavefunction/slater_jastrow.py:30-482).

Same constructor, attributes, sub-modules and state_dict keys as the reference
(``ao.atom_coords``, ``ao.bas_exp``, ``mo.mo_modifier``, ``fc.weight``,
``jastrow.jastrow_kernel.weight`` / ``jastrow.jastrow_terms.<i>...``).  The
numerical work -- AO -> MO -> Slater determinants -> CI sum -> Jastrow, the
Jacobi kinetic energy and the potentials -- runs in the fused kernels of
libqmcmi355x (``qmc_wf_eval``); parameter gradients of ``forward`` come from
the adjoint kernel ``qmc_logpsi_grad_rows`` instead of autograd.

Extra keyword ``precision``: ``"fp64"`` (default, the reference's double
precision end to end) or ``"mixed"`` (psi, pdf, sweeps and drift with AO, MO and
Jastrow in FP32, Slater elimination and acceptance in FP64).  Kinetic and local
energies are FP64 in both.
"""
from typing import Optional, Union

import numpy as np
import torch
from torch import nn

from .. import _native as N
from .. import log
from .jastrows.combine_jastrow import CombineJastrow
from .jastrows.elec_elec.jastrow_factor_electron_electron import JastrowFactorElectronElectron
from .jastrows.elec_elec.kernels.pade_jastrow_kernel import PadeJastrowKernel
from .orbitals.atomic_orbitals import AtomicOrbitals
from .orbitals.molecular_orbitals import MolecularOrbitals
from .param_views import ParamViews, f64
from .plan import DevicePlan, SystemTables, nuclear_repulsion
from .pooling.orbital_configurations import OrbitalConfigurations
from .pooling.slater_pooling import SlaterPooling
from .wf_base import WaveFunction

_PREC = {"fp64": N.QMC_PREC_FP64, "mixed": N.QMC_PREC_MIXED}


class SlaterJastrow(WaveFunction):
    def __init__(
        self,
        mol,
        jastrow: Optional[Union[str, nn.Module, list]] = "default",
        backflow=None,
        configs: str = "ground_state",
        kinetic: str = "jacobi",
        cuda: bool = False,
        include_all_mo: bool = True,
        mix_mo: bool = False,
        orthogonalize_mo: bool = False,
        precision: str = "fp64",
        eloc_precision: Optional[str] = None,
    ) -> None:
        super().__init__(mol.nelec, 3, kinetic, cuda)
        if backflow is not None:
            raise NotImplementedError("backflow orbitals are not part of the MI355X path")
        if kinetic not in ("jacobi", "auto"):
            raise ValueError("kinetic must be 'jacobi' or 'auto'")
        if not include_all_mo and isinstance(configs, str) and configs.startswith("cas("):
            raise ValueError("CAS calculation only possible with include_all_mo=True")
        if precision not in _PREC:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if eloc_precision not in (None, "fp64"):
            raise ValueError("the kinetic and local energies are evaluated in FP64: FP32 atomic orbitals left "
                             "errors above 1e-6 Ha/electron near nodes of psi, at no speed gain")
        # precision: psi, pdf, the Metropolis / HMC sweeps and the drift
        self.precision = precision
        self.eloc_precision = "fp64"
        self.mol = mol
        self.atoms = mol.atoms
        self.natom = mol.natom

        # configurations (slater_jastrow.py:158-169)
        self.orb_confs = OrbitalConfigurations(mol)
        self.configs_method = configs if isinstance(configs, str) else "explicit"
        self.configs = self.orb_confs.get_configs(configs)
        self.nci = len(self.configs[0])
        mx = [int(c.max()) for c in self.configs if c.numel()]
        self.highest_occ_mo = max(mx) + 1

        self.use_backflow = False
        self.ao = AtomicOrbitals(mol, cuda)
        self.include_all_mo = include_all_mo
        self.nmo_opt = int(mol.basis.nmo) if include_all_mo else self.highest_occ_mo
        self.mo = MolecularOrbitals(mol, include_all_mo, self.highest_occ_mo, mix_mo, orthogonalize_mo, cuda)
        self.pool = SlaterPooling(self.configs_method, self.configs, mol, cuda)

        # CI layer, ground state weight 1 (slater_jastrow.py:179-191)
        self.fc = nn.Linear(self.nci, 1, bias=False)
        self.fc.weight.data.fill_(0.0)
        self.fc.weight.data[0][0] = 1.0
        if cuda:
            self.fc = self.fc.to(self.device)

        # Jastrow (slater_jastrow.py:193-222)
        if jastrow is None:
            self.jastrow = None
            self.use_jastrow = False
        else:
            self.use_jastrow = True
            if isinstance(jastrow, str) and jastrow == "default":
                self.jastrow = JastrowFactorElectronElectron(mol, PadeJastrowKernel, cuda=cuda)
            elif isinstance(jastrow, list):
                self.jastrow = CombineJastrow(jastrow)
            elif isinstance(jastrow, nn.Module):
                self.jastrow = jastrow
            else:
                raise TypeError("Jastrow factor not supported.")
            self.jastrow_type = self.jastrow.__repr__()
            if cuda:
                self.jastrow = self.jastrow.to(self.device)

        # kinetic='auto' (wf_base.py:118-182: -1/2 lap psi / psi by double
        # autograd) asks for the same quantity; the device evaluates it with
        # the exact Jacobi kernel in either case
        self.kinetic_method = kinetic
        self.kinetic_energy = self.kinetic_energy_jacobi
        self.gradients = self.gradients_jacobi
        self._plan = None
        self._plan_key = None
        self._plan_tables = None
        self.log_data()

    # ------------------------------------------------------------------ plan
    def _jastrow_terms(self):
        if not self.use_jastrow:
            return None, None
        terms = self.jastrow.terms() if isinstance(self.jastrow, CombineJastrow) else [self.jastrow]
        ee = en = None
        for t in terms:
            kind = getattr(t, "kind", None)
            if kind == "ee":
                ee = t
            elif kind == "en":
                en = t
            else:
                raise NotImplementedError("Jastrow term %r is not supported on the device" % t)
        return ee, en

    def device_plan(self) -> DevicePlan:
        """Plan for the current exponents and atomic positions: rebuilt when
        exponent sharing breaks or when ao.atom_coords moves (V_nn follows the
        AO layer's coordinates, wf_base.py:97-116)."""
        key = (self.ao.bas_exp._version, self.ao.atom_coords._version)
        if self._plan is not None and self._plan_key == key:
            return self._plan
        ex = self.ao.bas_exp.detach().to("cpu", torch.float64).numpy()
        coords = self.ao.atom_coords.detach().to("cpu", torch.float64).numpy()
        vnn = nuclear_repulsion(coords, self.mol.atomic_number)
        if self._plan is not None and self._plan_tables.grouping_valid(ex) and vnn == self._plan_tables.vnn:
            self._plan_key = key
            return self._plan
        ee, en = self._jastrow_terms()
        tab = SystemTables(
            self.mol.basis, self.mol.atom_coords, self.mol.atomic_number, self.mol.nup, self.mol.ndown,
            (self.configs[0].numpy(), self.configs[1].numpy()), self.highest_occ_mo,
            ee is not None, en is not None, self.ao.norm_cst.cpu().numpy(), ex, vnn_coords=coords,
        )
        self._plan = DevicePlan(tab)
        self._plan_tables = tab
        self._plan_key = key
        return self._plan

    def _params(self, precision: str):
        """Parameter views handed to the kernels: one gather launch refreshes W
        and the grouped exponents from the live parameters; the remaining
        pointers are the parameters themselves (float64)."""
        ee, en = self._jastrow_terms()
        plan = self.device_plan()
        mod = f64(self.mo.mo_modifier)
        dev = mod.device
        pv = getattr(self, "_pviews", None)
        if pv is None or pv.plan is not plan or pv.mo_w64.device != dev:
            pv = self._pviews = ParamViews(plan, dev)
        mo_scf = self.mo.mo_scf  # a constant of the MO layer (not a parameter)
        if mo_scf.device != dev or mo_scf.dtype != torch.float64 or not mo_scf.is_contiguous():
            c = getattr(self, "_mo_scf64", None)
            if c is None or c.device != dev or c.shape != mo_scf.shape:
                c = self._mo_scf64 = mo_scf.to(dev, torch.float64).contiguous()
            mo_scf = c
        bas_exp = f64(self.ao.bas_exp)
        pv.gather(mo_scf, mod, bas_exp)
        mixed = precision == "mixed"
        keep = {
            "mo_modifier": mod,
            "bas_exp": bas_exp,
            "atom_coords": f64(self.ao.atom_coords),
            "ci": f64(self.fc.weight).reshape(-1),
            "w_ee": f64(ee.jastrow_kernel.weight) if ee is not None else None,
            "w_en": f64(en.jastrow_kernel.weight) if en is not None else None,
            "views": pv,
        }
        p = N.QmcParams()
        p.bas_exp, p.atom_coords = bas_exp.data_ptr(), keep["atom_coords"].data_ptr()
        p.ci, p.w_ee, p.w_en = keep["ci"].data_ptr(), N.ptr(keep["w_ee"]), N.ptr(keep["w_en"])
        p.mo_w = (pv.mo_w32 if mixed else pv.mo_w64).data_ptr()
        p.entry_w = (pv.entry_w32 if mixed else pv.entry_w64).data_ptr()
        p.grp_exp, p.grp_expf = pv.grp_exp.data_ptr(), pv.grp_expf.data_ptr()
        return p, keep

    def _check_pos(self, pos: torch.Tensor) -> torch.Tensor:
        N.require_device(pos, "positions")
        if pos.dim() != 2 or pos.shape[1] != self.nelec * 3:
            raise ValueError("positions must have shape (Nbatch, %d)" % (self.nelec * 3))
        return pos.detach().to(torch.float64).contiguous()

    def _eval(self, pos: torch.Tensor, mode: int, precision: Optional[str] = None):
        precision = precision or (self.precision if mode == N.QMC_EVAL_PSI else self.eloc_precision)
        x = self._check_pos(pos)
        nw = x.shape[0]
        plan = self.device_plan()
        # the local-energy kernel reads W in FP64 in both precisions
        prm, keep = self._params(precision if mode == N.QMC_EVAL_PSI else "fp64")
        out = {k: torch.empty(nw, dtype=torch.float64, device=x.device)
               for k in (("psi", "logabs", "sign") if mode == N.QMC_EVAL_PSI
                         else ("psi", "logabs", "sign", "kinetic", "eloc", "epot"))}
        if nw == 0:  # an empty batch (its tensors have no storage to hand over)
            return out
        o = N.QmcWfOut()
        for k, v in out.items():
            setattr(o, k, v.data_ptr())
        N.check(N.load().qmc_wf_eval(plan.handle, prm, _PREC[precision], mode, x.data_ptr(), nw, o,
                                     N.current_stream()))
        return out

    # ---------------------------------------------------------- public API
    def forward(self, x: torch.Tensor, ao: Optional[torch.Tensor] = None) -> torch.Tensor:
        """psi = J * sum_n c_n D_up D_dn, shape (Nbatch, 1) (slater_jastrow.py:243-286).

        Differentiable with respect to ``mo.mo_modifier``, ``fc.weight``,
        the Jastrow weights and ``ao.bas_exp`` through the adjoint kernel."""
        if ao is not None:
            raise NotImplementedError("forward(x, ao=...) is not supported by the fused device path")
        ee, en = self._jastrow_terms()
        params = [self.mo.mo_modifier, self.fc.weight, self.ao.bas_exp,
                  ee.jastrow_kernel.weight if ee is not None else None,
                  en.jastrow_kernel.weight if en is not None else None]
        if self.ao.atom_coords.requires_grad and torch.is_grad_enabled() and not getattr(self, "_warned_geo", False):
            log.warning("ao.atom_coords requires grad: the device adjoint does not produce gradients with "
                        "respect to atomic positions (forces are not part of this path)")
            self._warned_geo = True
        return _PsiFunction.apply(self, x, *params)

    def pdf(self, pos: torch.Tensor, return_grad: Optional[bool] = False) -> torch.Tensor:
        if return_grad:
            return self.gradients(pos, pdf=True)
        psi = self._eval(pos, N.QMC_EVAL_PSI)["psi"]
        return (psi * psi).to(torch.get_default_dtype())

    def log_psi(self, pos: torch.Tensor, precision: Optional[str] = None):
        """(log|psi|, sign(psi)) in float64, shape (Nbatch,)."""
        out = self._eval(pos, N.QMC_EVAL_PSI, precision)
        return out["logabs"], out["sign"]

    def kinetic_energy_autograd(self, x: torch.Tensor) -> torch.Tensor:
        """-1/2 lap(psi)/psi (wf_base.py:118-182), evaluated exactly by the Jacobi kernel."""
        return self.kinetic_energy_jacobi(x)

    def kinetic_energy_jacobi(self, x: torch.Tensor, **kwargs) -> torch.Tensor:
        """-1/2 lap(psi)/psi by the Jacobi formula, (Nbatch, 1) (slater_jastrow.py:312-344)."""
        return self._eval(x, N.QMC_EVAL_ELOC)["kinetic"].to(torch.get_default_dtype()).view(-1, 1)

    def local_energy(self, pos: torch.Tensor) -> torch.Tensor:
        """T + V_en + V_ee + V_nn, (Nbatch, 1) (wf_base.py:184-215)."""
        return self._eval(pos, N.QMC_EVAL_ELOC)["eloc"].to(torch.get_default_dtype()).view(-1, 1)

    def local_energy_parts(self, pos: torch.Tensor, precision: Optional[str] = None):
        """All outputs of one local-energy evaluation (float64 dict)."""
        return self._eval(pos, N.QMC_EVAL_ELOC, precision)

    def nuclear_repulsion(self) -> torch.Tensor:
        """V_nn of the AO layer's atomic positions (wf_base.py:97-116), a 0-d tensor."""
        return torch.as_tensor(self.device_plan().tables.vnn, dtype=torch.get_default_dtype())

    def _potentials(self, pos: torch.Tensor):
        x = self._check_pos(pos)
        nw = x.shape[0]
        vee = torch.empty(nw, dtype=torch.float64, device=x.device)
        ven = torch.empty(nw, dtype=torch.float64, device=x.device)
        if nw:
            prm, keep = self._params("fp64")
            o = N.QmcJastrowOut()
            o.v_ee, o.v_en = vee.data_ptr(), ven.data_ptr()
            N.check(N.load().qmc_jastrow_eval(self.device_plan().handle, prm, 0, x.data_ptr(), nw, o,
                                              N.current_stream()))
        dt = torch.get_default_dtype()
        return vee.to(dt).view(-1, 1), ven.to(dt).view(-1, 1)

    def electronic_potential(self, pos: torch.Tensor) -> torch.Tensor:
        """sum_{i<j} 1/r_ij per walker, (Nbatch, 1) (wf_base.py:49-70)."""
        return self._potentials(pos)[0]

    def nuclear_potential(self, pos: torch.Tensor) -> torch.Tensor:
        """-sum_{i,A} Z_A / r_iA per walker with the AO layer's atomic positions,
        (Nbatch, 1) (wf_base.py:72-95)."""
        return self._potentials(pos)[1]

    def local_energy_autodiff(self, pos: torch.Tensor) -> torch.Tensor:
        """Local energy differentiable with respect to the parameters through the
        hand-written adjoint (qmc_eloc_grad_rows); used by grad='auto'."""
        ee, en = self._jastrow_terms()
        params = [self.mo.mo_modifier, self.fc.weight, self.ao.bas_exp,
                  ee.jastrow_kernel.weight if ee is not None else None,
                  en.jastrow_kernel.weight if en is not None else None]
        return _ELocFunction.apply(self, pos, *params)

    def drift(self, pos: torch.Tensor, precision: Optional[str] = None):
        """(grad_i ln|psi| as (Nbatch, Nelec, 3), log|psi|, sign(psi)) in float64
        from qmc_wf_drift; ``precision`` defaults to the sampling precision."""
        precision = precision or self.precision
        x = self._check_pos(pos)
        nw = x.shape[0]
        plan = self.device_plan()
        prm, keep = self._params(precision)
        dr = torch.empty(nw, self.nelec, 3, dtype=torch.float64, device=x.device)
        la = torch.empty(nw, dtype=torch.float64, device=x.device)
        sg = torch.empty(nw, dtype=torch.float64, device=x.device)
        if nw == 0:
            return dr, la, sg
        N.check(N.load().qmc_wf_drift(plan.handle, prm, _PREC[precision], x.data_ptr(), nw, dr.data_ptr(),
                                      la.data_ptr(), sg.data_ptr(), N.current_stream()))
        return dr, la, sg

    def gradients_jacobi(self, x: torch.Tensor, sum_grad: Optional[bool] = False,
                         pdf: Optional[bool] = False) -> torch.Tensor:
        """grad psi (or grad psi^2 with pdf=True), shape (Nbatch, 3 Nelec)
        (slater_jastrow.py:346-447): psi * (grad ln J + grad S / S), FP64."""
        dr, la, sg = self.drift(x, "fp64")
        psi = sg * torch.exp(la)
        out = dr.view(x.shape[0], -1) * psi[:, None]
        if pdf:
            out = 2.0 * out * psi[:, None]
        return out.to(torch.get_default_dtype())

    def log_data(self) -> None:
        log.info("")
        log.info(" Wave Function")
        log.info("  Jastrow factor      : {0}", self.use_jastrow)
        if self.use_jastrow:
            log.info("  Jastrow kernel      : {0}", self.jastrow_type)
        log.info("  Highest MO included : {0}", self.nmo_opt)
        log.info("  Configurations      : {0}", self.configs_method)
        log.info("  Number of confs     : {0}", self.nci)
        log.info("  Kinetic energy      : {0}", self.kinetic_method)
        log.info("  Precision           : {0}", self.precision)
        log.info("  Number var  param   : {0}", self.get_number_parameters())
        log.info("  Cuda support        : {0}", self.cuda)

    def geometry(self, pos=None, convert_to_angs: Optional[bool] = False):
        from ..utils.constants import BOHR2ANGS

        conv = BOHR2ANGS if convert_to_angs else 1.0
        return [(self.ao.atom_coords[i].detach().cpu().numpy() * conv).tolist() for i in range(self.natom)]


class _PsiFunction(torch.autograd.Function):
    """psi(pos; theta) with a hand-written adjoint for theta.

    backward: with g_w = dL/dpsi_w * psi_w, one qmc_logpsi_grad launch pair
    (AO values recomputed, per-wavefront partial sums, one slot reduction)
    gives dL/dW = sum_w g_w AO_w^T Gamma_w, dL/dc, dL/dw_ee, dL/dw_en and
    dL/dbas_exp; then dL/dmo_modifier = dL/dW * mo_scf (molecular_orbitals.py:92).
    Nothing is stored by the forward.  With respect to the positions:
    dL/dx = dL/dpsi psi grad ln|psi| (the FP64 drift kernel, the derivative of
    gradients_jacobi)."""

    @staticmethod
    def forward(ctx, wf, x, mo_mod, ci, bas_exp, w_ee, w_en):
        out = wf._eval(x, N.QMC_EVAL_PSI)
        psi = out["psi"]
        ctx.wf = wf
        ctx.save_for_backward(x.detach(), psi)
        return psi.to(torch.get_default_dtype()).view(-1, 1)

    @staticmethod
    def backward(ctx, grad_out):
        wf = ctx.wf
        x, psi = ctx.saved_tensors
        g = grad_out.reshape(-1).to(torch.float64) * psi
        d_x = None
        if ctx.needs_input_grad[1]:
            dr, _, _ = wf.drift(x, "fp64")
            d_x = (g[:, None] * dr.view(x.shape[0], -1)).to(x.dtype)
        if not any(ctx.needs_input_grad[2:]):
            return None, d_x, None, None, None, None, None
        r = adjoint_reduce(wf, x, g, "logpsi", need_exp=ctx.needs_input_grad[4])
        return (None, d_x) + _param_grads(wf, r, ctx.needs_input_grad)


def _param_grads(wf, r, needs):
    """Gradients of (mo_modifier, fc.weight, bas_exp, w_ee, w_en) from the
    reduced adjoint outputs."""
    dt = torch.get_default_dtype()
    d_mod = d_ci = d_exp = d_ee = d_en = None
    if needs[2]:
        full = torch.zeros_like(wf.mo.mo_modifier, dtype=torch.float64)
        full[:, : r["w"].shape[1]] = r["w"]
        d_mod = (full * wf.mo.mo_scf.to(full)).to(wf.mo.mo_modifier.dtype)
    if needs[3]:
        d_ci = r["ci"].view(1, -1).to(wf.fc.weight.dtype)
    if needs[4] and r.get("exp") is not None:
        d_exp = r["exp"].to(wf.ao.bas_exp.dtype)
    if needs[5]:
        d_ee = r["jas"][0:1].to(dt)
    if needs[6]:
        d_en = r["jas"][1:2].to(dt)
    return d_mod, d_ci, d_exp, d_ee, d_en


def adjoint_reduce(wf: SlaterJastrow, pos: torch.Tensor, weights: torch.Tensor, which: str,
                   need_exp: bool = False, workspace_bytes: Optional[int] = None):
    """sum_w g_w d f_w / d theta for f = log|psi| (which="logpsi", qmc_logpsi_grad)
    or f = E_L (which="eloc", qmc_eloc_grad): dict of w (nao, nmo), ci (nci,),
    jas (2,), exp (nprim) or None, and eloc (nw,) for E_L.  The workspace is
    the size that uses the most partial-sum slots unless workspace_bytes is given."""
    x = wf._check_pos(pos)
    nw = x.shape[0]
    plan = wf.device_plan()
    prm, keep = wf._params("fp64")
    dev, f64_ = x.device, torch.float64
    sz = plan.sizes()
    lib = N.load()
    with_exp = need_exp and which == "logpsi"
    out = {
        "w": torch.empty(sz.nao, sz.nmo, dtype=f64_, device=dev),
        "ci": torch.empty(plan.tables.nci, dtype=f64_, device=dev),
        "jas": torch.empty(2, dtype=f64_, device=dev),
        "exp": torch.empty(plan.tables.nprim_ref, dtype=f64_, device=dev) if with_exp else None,
        "eloc": torch.empty(nw, dtype=f64_, device=dev) if which == "eloc" else None,
    }
    if workspace_bytes is None:
        workspace_bytes = int(lib.qmc_adjoint_workspace_size(plan.handle, nw, 1 if with_exp else 0))
    ws = torch.empty(max(workspace_bytes, 8), dtype=torch.uint8, device=dev)
    g = weights.detach().to(dev, f64_).contiguous()
    o = N.QmcGradOut()
    o.grad_w, o.grad_ci, o.grad_jas, o.grad_exp = (out["w"].data_ptr(), out["ci"].data_ptr(),
                                                    out["jas"].data_ptr(), N.ptr(out["exp"]))
    xp = x.data_ptr() if nw else 0
    gp = g.data_ptr() if nw else 0
    if which == "logpsi":
        N.check(lib.qmc_logpsi_grad(plan.handle, prm, xp, nw, gp, o, ws.data_ptr(), workspace_bytes,
                                    N.current_stream()))
    else:
        N.check(lib.qmc_eloc_grad(plan.handle, prm, xp, nw, gp, N.ptr(out["eloc"]) if nw else 0, o,
                                  ws.data_ptr(), workspace_bytes, N.current_stream()))
    return out


def wf_adjoint_rows(wf: SlaterJastrow, pos: torch.Tensor, need_exp: bool = False):
    """Per-walker/per-electron adjoint rows of log|psi| (qmc_logpsi_grad_rows)."""
    x = wf._check_pos(pos)
    nw, nel = x.shape[0], wf.nelec
    plan = wf.device_plan()
    prm, keep = wf._params("fp64")
    dev = x.device
    f64 = torch.float64
    nao = plan.tables.nao
    rows = {
        "ao": torch.empty(nw * nel, nao, dtype=f64, device=dev),
        "gamma": torch.empty(nw * nel, plan.tables.nmo, dtype=f64, device=dev),
        "dci": torch.empty(nw, plan.tables.nci, dtype=f64, device=dev),
        "djas": torch.empty(nw, 2, dtype=f64, device=dev),
        "dexp": torch.empty(nw * nel, plan.tables.nprim_ref, dtype=f64, device=dev) if need_exp else None,
    }
    if nw == 0:  # empty rows reduce to zero gradients
        return rows
    N.check(N.load().qmc_logpsi_grad_rows(plan.handle, prm, x.data_ptr(), nw, rows["ao"].data_ptr(),
                                          rows["gamma"].data_ptr(), rows["dci"].data_ptr(),
                                          rows["djas"].data_ptr(), N.ptr(rows["dexp"]), N.current_stream()))
    return rows


class _ELocFunction(torch.autograd.Function):
    """E_L(pos; theta) with the hand-written adjoint of qmc_eloc_grad: for
    g = dL/dE_L per walker, dL/dW = sum_w g_w (AO_w^T GA_w + b_w^T GB_w),
    dL/dc = g^T dE_L/dc, dL/dw_ee,en = g^T dE_L/dw (reduced on the device).
    Exponent and position gradients of E_L are not provided."""

    @staticmethod
    def forward(ctx, wf, x, mo_mod, ci, bas_exp, w_ee, w_en):
        el = wf._eval(x, N.QMC_EVAL_ELOC)["eloc"]
        ctx.wf = wf
        ctx.save_for_backward(x.detach())
        return el.to(torch.get_default_dtype()).view(-1, 1)

    @staticmethod
    def backward(ctx, grad_out):
        wf = ctx.wf
        (x,) = ctx.saved_tensors
        if ctx.needs_input_grad[4]:
            raise NotImplementedError("d E_L / d bas_exp is not provided by the device adjoint: "
                                      "freeze 'ao' (Solver.configure(freeze=['ao']))")
        # no gradient with respect to the positions: sampled positions arrive
        # with requires_grad set (metropolis.py:177) but are constants here
        g = grad_out.reshape(-1).to(torch.float64)
        r = adjoint_reduce(wf, x, g, "eloc")
        return (None, None) + _param_grads(wf, r, ctx.needs_input_grad)


def eloc_adjoint_rows(wf: SlaterJastrow, pos: torch.Tensor, need_rows: bool = True):
    """Rows of the local-energy adjoint (qmc_eloc_grad_rows), FP64."""
    x = wf._check_pos(pos)
    nw, nel = x.shape[0], wf.nelec
    plan = wf.device_plan()
    prm, keep = wf._params("fp64")
    dev, f64 = x.device, torch.float64
    nao, nmo = plan.tables.nao, plan.tables.nmo
    r = {
        "eloc": torch.empty(nw, dtype=f64, device=dev),
        "ao": torch.empty(nw * nel, nao, dtype=f64, device=dev) if need_rows else None,
        "b": torch.empty(nw * nel, nao, dtype=f64, device=dev) if need_rows else None,
        "ga": torch.empty(nw * nel, nmo, dtype=f64, device=dev) if need_rows else None,
        "gb": torch.empty(nw * nel, nmo, dtype=f64, device=dev) if need_rows else None,
        "dci": torch.empty(nw, plan.tables.nci, dtype=f64, device=dev),
        "djas": torch.empty(nw, 2, dtype=f64, device=dev),
        "drift": torch.empty(nw, nel, 3, dtype=f64, device=dev),
    }
    if nw == 0:
        return r
    N.check(N.load().qmc_eloc_grad_rows(plan.handle, prm, x.data_ptr(), nw, N.ptr(r["eloc"]), N.ptr(r["ao"]),
                                        N.ptr(r["b"]), N.ptr(r["ga"]), N.ptr(r["gb"]), N.ptr(r["dci"]),
                                        N.ptr(r["djas"]), N.ptr(r["drift"]), N.current_stream()))
    return r
