"""qmc_logpsi_grad_noalloc (static device scratch, no allocation) against
qmc_logpsi_grad (caller workspace) through the C ABI, on two calls in a row,
the second with fewer walkers: with a workspace of QMC_NOALLOC_SCRATCH_BYTES
both use the same slot count, so the gradients agree bit for bit; with the
default workspace they agree within 1e-12 relative (summation order)."""
import ctypes
import sys

import pytest
import torch

from qmctorch_amd import _native as N
from tests import helpers

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _grads(wf, x, g, noalloc, workspace_bytes=None):
    plan = wf.device_plan()
    prm, keep = wf._params("fp64")
    sz = plan.sizes()
    out = {"w": torch.full((sz.nao * sz.nmo,), float("nan"), dtype=torch.float64, device=DEV),
           "ci": torch.full((wf.nci,), float("nan"), dtype=torch.float64, device=DEV),
           "jas": torch.full((2,), float("nan"), dtype=torch.float64, device=DEV),
           "exp": torch.full((plan.tables.nprim_ref,), float("nan"), dtype=torch.float64, device=DEV)}
    o = N.QmcGradOut()
    o.grad_w, o.grad_ci, o.grad_jas, o.grad_exp = (out["w"].data_ptr(), out["ci"].data_ptr(), out["jas"].data_ptr(),
                                                    out["exp"].data_ptr())
    lib = N.load()
    nw = x.shape[0]
    if noalloc:
        N.check(lib.qmc_logpsi_grad_noalloc(plan.handle, ctypes.byref(prm), x.data_ptr(), nw, g.data_ptr(),
                                            ctypes.byref(o), N.current_stream()))
    else:
        wsb = workspace_bytes or int(lib.qmc_adjoint_workspace_size(plan.handle, nw, 1))
        ws = torch.empty(wsb, dtype=torch.uint8, device=DEV)
        N.check(lib.qmc_logpsi_grad(plan.handle, ctypes.byref(prm), x.data_ptr(), nw, g.data_ptr(),
                                    ctypes.byref(o), ws.data_ptr(), wsb, N.current_stream()))
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("case", ["h2o_ccpvdz_ee_en", "lih_631g_cas22_ee_en", "c6h6_ccpvdz_ee_en"])
def test_noalloc_logpsi_grad_matches_workspace_entry_point(case):
    torch.set_default_dtype(torch.float64)
    _say("case noalloc", case)
    mol, kw = helpers.CASES[case]()
    wf, _ = helpers.make_pair(mol, **kw)
    sizes = (64, 24) if mol.nelec > 20 else (3000, 700)
    for nw in sizes:  # two calls in a row, the second smaller
        x = helpers.positions(mol, nw, seed=nw).to(DEV).contiguous()
        g = torch.linspace(-1.0, 2.0, nw, dtype=torch.float64, device=DEV)
        a = _grads(wf, x, g, noalloc=True)
        b = _grads(wf, x, g, noalloc=False, workspace_bytes=N.QMC_NOALLOC_SCRATCH_BYTES)
        c = _grads(wf, x, g, noalloc=False)
        for k in a:
            assert torch.isfinite(a[k]).all(), k
            assert torch.equal(a[k], b[k]), (nw, k)
            scale = float(c[k].abs().max())
            torch.testing.assert_close(a[k], c[k], rtol=1e-12, atol=1e-13 * scale)
        _say("  nw %d: bit-identical to the same-size workspace, within 1e-12 of the default" % nw)
