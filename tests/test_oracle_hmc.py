"""The oracle's restatement of the HMC sampler (sampler/hamiltonian.py) pinned
by the mathematics of the leapfrog integrator.

For a Gaussian target pdf(x) = exp(-sum_i a_i x_i^2 / 2) the potential is
U = sum_i a_i x_i^2 / 2 and one leapfrog step of size eps is the linear map
    (q, p) -> M_i (q, p),  M_i = [[1 - eps^2 a_i/2, eps],
                                  [-eps a_i (1 - eps^2 a_i/4), 1 - eps^2 a_i/2]]
per coordinate, so L steps are M_i^L; the energy change follows from the end
point.  A dropped half step, a swapped update order, a sign error in grad U
or in the acceptance test each fail one of these checks.  The autograd
gradient of U = -log psi^2 is pinned against the Jacobi gradient and central
differences.
"""
import math

import numpy as np
import pytest
import torch

from oracle import vmc_oracle as O
from tests import helpers, systems

A = torch.tensor([0.5, 1.0, 2.0, 3.5, 0.8, 1.7], dtype=torch.float64)


def _pdf(x):
    return torch.exp(-0.5 * (A * x * x).sum(1))


def _leapfrog_closed_form(q, p, eps, L):
    c = 1.0 - 0.5 * eps * eps * A
    M = torch.stack([torch.stack([c, torch.full_like(A, eps)], -1),
                     torch.stack([-eps * A * (1.0 - 0.25 * eps * eps * A), c], -1)], -2)  # (6, 2, 2)
    ML = torch.linalg.matrix_power(M, L)
    v = torch.stack([q, p], -1).unsqueeze(-1)  # (nw, 6, 2, 1)
    out = (ML @ v).squeeze(-1)
    return out[..., 0], out[..., 1]


@pytest.mark.parametrize("L", [1, 2, 7])
def test_leapfrog_matches_closed_form_map(L):
    g = torch.Generator().manual_seed(L)
    q = torch.randn(16, 6, generator=g, dtype=torch.float64)
    p = torch.randn(16, 6, generator=g, dtype=torch.float64)
    eps = 0.37
    qL, pL = _leapfrog_closed_form(q, p, eps, L)
    got, acc = O.hmc_step(_pdf, q, p, eps, L, torch.zeros(16, dtype=torch.float64))  # u = 0: always kept
    assert bool(acc.all())
    torch.testing.assert_close(got, qL, rtol=1e-13, atol=1e-13)
    # acceptance boundary: exp(E0 - E1) against u just above / below it
    e0 = 0.5 * (A * q * q).sum(1) + 0.5 * (p * p).sum(1)
    e1 = 0.5 * (A * qL * qL).sum(1) + 0.5 * (pL * pL).sum(1)
    ratio = torch.exp(e0 - e1)
    up, acc_up = O.hmc_step(_pdf, q, p, eps, L, ratio * (1 + 1e-9))
    assert not bool(acc_up[ratio < 0.999].any())
    torch.testing.assert_close(up[~acc_up], q[~acc_up], rtol=0, atol=0)
    _, acc_dn = O.hmc_step(_pdf, q, p, eps, L, ratio * (1 - 1e-9))
    assert bool(acc_dn.all())


def test_leapfrog_is_reversible_and_second_order():
    g = torch.Generator().manual_seed(3)
    q = torch.randn(8, 6, generator=g, dtype=torch.float64)
    p = torch.randn(8, 6, generator=g, dtype=torch.float64)
    # energy error of the integrator is O(eps^2) at fixed trajectory time
    errs = []
    for eps, L in ((0.2, 5), (0.1, 10), (0.05, 20)):
        qL, pL = _leapfrog_closed_form(q, p, eps, L)
        e0 = 0.5 * (A * q * q).sum(1) + 0.5 * (p * p).sum(1)
        e1 = 0.5 * (A * qL * qL).sum(1) + 0.5 * (pL * pL).sum(1)
        errs.append(float((e1 - e0).abs().max()))
        # running the map back from (qL, -pL) returns to (q, -p)
        qb, pb = _leapfrog_closed_form(qL, -pL, eps, L)
        torch.testing.assert_close(qb, q, rtol=1e-12, atol=1e-12)
    assert errs[0] / errs[1] > 3.5 and errs[1] / errs[2] > 3.5


def test_hamiltonian_gaussian_moments():
    """<x_i^2> = 1 / a_i under exp(-sum a x^2 / 2) sampled by the oracle HMC."""
    nw = 2000
    torch.manual_seed(0)
    x0 = torch.randn(nw, 6, dtype=torch.float64)
    kept, rate = O.hamiltonian(_pdf, x0, nstep=60, step_size=0.3, L=5, ntherm=20, ndecor=10, seed=1)
    assert 0.5 < rate <= 1.0
    m = (kept * kept).mean(0)
    err = (kept * kept).std(0) / math.sqrt(nw)  # kept steps are 10 HMC steps apart: ~independent per walker
    assert torch.all((m - 1.0 / A).abs() < 6 * err), (m, 1.0 / A, err)


def test_grad_potential_against_jacobi_and_differences():
    mol, kw = helpers.CASES["lih_631g_cas22_ee_en"]()
    _, orc = helpers.make_pair(mol, cuda=False, **kw)
    pos = helpers.positions(mol, 6, seed=2)
    gu = O.grad_potential(orc.pdf, pos)
    psi = orc.forward(pos).detach().view(-1, 1)
    gj = orc.gradients_jacobi(pos.clone().requires_grad_()).detach()
    torch.testing.assert_close(gu, -2.0 * gj / psi, rtol=1e-9, atol=1e-10)
    h = 1e-5
    for k in (0, 4, 11):
        e = torch.zeros_like(pos)
        e[:, k] = h
        with torch.no_grad():
            fd = (-torch.log(orc.pdf(pos + e)) + torch.log(orc.pdf(pos - e))) / (2 * h)
        torch.testing.assert_close(gu[:, k], fd, rtol=1e-6, atol=1e-6)
