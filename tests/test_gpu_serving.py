"""Batched many-system launches and the continuous-batching request loop (MI355X).

Bar: a batched launch reproduces the per-system launches bit for bit (the
kernels run the same per-walker arithmetic), every request's result is
identical alone and in a mix, and the FP64 batched local energy matches the
oracle within 1e-10 relative.
"""
import sys

import numpy as np
import pytest
import torch

from oracle import vmc_oracle as O
from qmctorch_amd import _native as N
from qmctorch_amd.serving import (BatchRunner, SystemContext, VMCRequest, VMCServer, WalkerSet, catalog,
                                  latency_summary, poisson_requests, serve_per_request, sweep_each)
from tests import helpers

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _double():
    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(old)


@pytest.fixture(scope="module")
def contexts():
    torch.set_default_dtype(torch.float64)
    return {name: SystemContext(catalog.wavefunction(name, "mixed"), name) for name in catalog.SYSTEM_NAMES}


JOBS = [("h2", 300, 11), ("lih", 517, 12), ("h2o", 1000, 13), ("n2", 77, 14), ("c6h6", 40, 15), ("h2o", 64, 16)]


def _walker_sets(contexts, jobs):
    out = []
    for name, nw, seed in jobs:
        ctx = contexts[name]
        out.append(WalkerSet(ctx, catalog.initial_positions(ctx.wf.mol, nw, seed), seed))
    return out


def _same(a, b):
    return (torch.equal(a.pos, b.pos) and torch.equal(a.logabs, b.logabs) and torch.equal(a.sign, b.sign)
            and torch.equal(a.naccept, b.naccept))


def test_batch_sweep_equals_per_system_launches(contexts):
    _say("case batch sweep vs per-system launches", [j[:2] for j in JOBS])
    runner = BatchRunner(64)
    wb = _walker_sets(contexts, JOBS)
    ws = _walker_sets(contexts, JOBS)
    runner.sweep([(w, 17) for w in wb])
    runner.sweep([(w, 13) for w in wb])
    sweep_each([(w, 30) for w in ws])
    torch.cuda.synchronize()
    for (name, nw, _), a, b in zip(JOBS, wb, ws):
        assert _same(a, b), name
        assert a.step == 30 and int(a.naccept.sum()) > 0, name


def test_batch_alone_equals_mix(contexts):
    _say("case batch alone vs mix")
    runner = BatchRunner(64)
    mix = _walker_sets(contexts, JOBS)
    runner.sweep([(w, 9) for w in mix])
    for k, job in enumerate(JOBS):
        alone = _walker_sets(contexts, [job])[0]
        runner.sweep([(alone, 9)])
        torch.cuda.synchronize()
        assert _same(alone, mix[k]), job


def test_batch_uneven_step_counts(contexts):
    _say("case batch with different step counts per job")
    runner = BatchRunner(64)
    steps = [3, 0, 7, 1, 2, 5]
    wb = _walker_sets(contexts, JOBS)
    ws = _walker_sets(contexts, JOBS)
    runner.sweep(list(zip(wb, steps)))
    sweep_each(list(zip(ws, steps)))
    torch.cuda.synchronize()
    for job, a, b, n in zip(JOBS, wb, ws, steps):
        assert a.step == n, job
        if n == 0:  # untouched: log|psi| was never evaluated
            assert torch.equal(a.pos, b.pos) and int(a.naccept.sum()) == 0, job
        else:
            assert _same(a, b), job


def test_batch_local_energy_equals_single_eval(contexts):
    _say("case batch local energy vs qmc_wf_eval")
    runner = BatchRunner(64)
    ws = _walker_sets(contexts, JOBS)
    outs = [torch.empty(w.nw, dtype=torch.float64, device="cuda") for w in ws]
    runner.local_energy(list(zip(ws, outs)))
    for w, e in zip(ws, outs):
        ref = w.ctx.wf.local_energy_parts(w.pos)["eloc"]
        assert torch.equal(e, ref), w.ctx.name


def test_batch_excitation_plans_equal_per_system_launches(contexts):
    """Jobs of excitation plans (more than 8 unique configurations per spin)
    run through the batched excitation kernels, beside ordinary jobs: sweeps
    bit-identical to their per-system launches, local energies identical to
    the single-plan evaluation."""
    _say("case batch with excitation plans (LiH and H2O single_double(4,6)) vs per-system launches")
    exc = {}
    for case, name in (("lih_631g_sd46_ee_en", "lih_sd46"), ("h2o_ccpvdz_sd46_ee", "h2o_sd46")):
        mol, kw = helpers.CASES[case]()
        wf, _ = helpers.make_pair(mol, precision="mixed", **kw)
        exc[name] = (SystemContext(wf, name), mol)
    jobs = [("h2o", 200, 21), ("lih_sd46", 150, 22), ("h2o_sd46", 61, 23), ("lih_sd46", 37, 24), ("lih", 90, 25)]

    def sets():
        out = []
        for name, nw, seed in jobs:
            if name in exc:
                ctx, mol = exc[name]
                out.append(WalkerSet(ctx, helpers.positions(mol, nw, seed=seed).cuda(), seed))
            else:
                ctx = contexts[name]
                out.append(WalkerSet(ctx, catalog.initial_positions(ctx.wf.mol, nw, seed), seed))
        return out

    runner = BatchRunner(8)
    wb, ws = sets(), sets()
    runner.sweep([(w, 11) for w in wb])
    runner.sweep([(w, 6) for w in wb])
    sweep_each([(w, 17) for w in ws])
    torch.cuda.synchronize()
    for (name, nw, _), a, b in zip(jobs, wb, ws):
        assert _same(a, b), name
        assert int(a.naccept.sum()) > 0, name
    outs = [torch.empty(w.nw, dtype=torch.float64, device="cuda") for w in wb]
    runner.local_energy(list(zip(wb, outs)))
    for (name, _, _), w, e in zip(jobs, wb, outs):
        ref = w.ctx.wf.local_energy_parts(w.pos)["eloc"]
        assert torch.equal(e, ref), name


def test_batch_local_energy_fp64_vs_oracle():
    _say("case batch fp64 local energy vs oracle (LiH CAS(2,2), H2O)")
    runner = BatchRunner(8)
    sets, oracles = [], []
    for case in ("lih_631g_cas22_ee_en", "h2o_ccpvdz_ee_en"):
        mol, kw = helpers.CASES[case]()
        wf, orc = helpers.make_pair(mol, **kw)
        ctx = SystemContext(wf, case)
        pos = helpers.positions(mol, 50, seed=4).cuda()
        sets.append(WalkerSet(ctx, pos, 0))
        oracles.append((orc, pos.cpu()))
    outs = [torch.empty(w.nw, dtype=torch.float64, device="cuda") for w in sets]
    runner.local_energy(list(zip(sets, outs)))
    for e, (orc, pos) in zip(outs, oracles):
        ref = orc.local_energy(pos).detach().reshape(-1)
        rel = ((e.cpu() - ref).abs() / ref.abs().clamp_min(1e-300)).max().item()
        _say("  max rel dE_L", rel)
        assert rel < 1e-10


def test_batch_rejects_bad_jobs(contexts):
    runner = BatchRunner(2)
    ws = _walker_sets(contexts, JOBS[:3])
    with pytest.raises(ValueError):
        runner.sweep([(w, 1) for w in ws])  # 3 jobs > capacity 2
    lib = N.load()
    arr = (N.QmcSweepJob * 1)()
    arr[0].plan = contexts["h2"].plan.handle
    arr[0].params = contexts["h2"].prm_sweep
    arr[0].nw = 4
    arr[0].sweep.nstep = 1
    rc = lib.qmc_metropolis_sweep_batch(runner.handle, N.QMC_PREC_MIXED, arr, 1, N.current_stream())
    assert rc != 0 and "walker" in lib.qmc_last_error().decode()


def test_server_each_result_same_alone_and_in_mix(contexts):
    reqs = poisson_requests(rate=2000.0, n=14, seed=3, wmax=2048)
    _say("case server alone vs mix", [(r.system, r.nwalkers, r.nsweeps) for r in reqs])
    server = VMCServer(precision="mixed", steps_per_tick=3, contexts=contexts)
    mix = {r.rid: r for r in server.run(reqs, realtime=False)}
    assert sorted(mix) == [r.rid for r in reqs]
    assert server.peak_active > 1
    for req in reqs:
        alone = server.run([req], realtime=False)[0]
        m = mix[req.rid]
        assert np.array_equal(alone.eloc, m.eloc), req
        assert alone.energy == m.energy and alone.acceptance == m.acceptance
        assert np.isfinite(m.energy) and m.eloc.shape == (req.nwalkers,)
    plain = {r.rid: r for r in serve_per_request(reqs, contexts, realtime=False)}
    worst = max(float(np.max(np.abs(plain[k].eloc - mix[k].eloc) / np.abs(mix[k].eloc))) for k in mix)
    _say("  plain per-request path vs server: max rel |dE_L|", worst)
    assert worst < 1e-12


def test_server_memory_budget_admission(contexts):
    reqs = [VMCRequest(i, s, w, k, 100 + i) for i, (s, w, k) in
            enumerate([("h2o", 900, 5), ("n2", 700, 4), ("lih", 1500, 6), ("h2", 2000, 2), ("h2o", 300, 7)])]
    _say("case server memory budget")
    free = VMCServer(precision="mixed", steps_per_tick=2, contexts=contexts)
    ref = {r.rid: r for r in free.run(reqs, realtime=False)}
    nbytes = sorted(free.request_bytes(r) for r in reqs)
    budget = nbytes[-1] + nbytes[-2]
    tight = VMCServer(precision="mixed", steps_per_tick=2, contexts=contexts, hbm_budget_bytes=budget)
    got = {r.rid: r for r in tight.run(reqs, realtime=False)}
    assert tight.peak_bytes <= budget < free.peak_bytes
    for k in ref:
        assert np.array_equal(ref[k].eloc, got[k].eloc)


def test_server_realtime_poisson(contexts):
    reqs = poisson_requests(rate=400.0, n=40, seed=5, wmax=4096)
    _say("case server realtime, 40 requests at 400 req/s")
    server = VMCServer(precision="mixed", steps_per_tick=4, contexts=contexts)
    server.run(reqs[:3], realtime=False)  # warm
    import time

    t0 = time.perf_counter()
    res = server.run(reqs, realtime=True, profile_ticks=True)
    wall = time.perf_counter() - t0
    s = latency_summary(res, wall)
    _say("  ", s)
    assert s["requests"] == 40 and all(r.done >= r.arrival for r in res)
    assert all(np.isfinite(r.energy) for r in res)
    tp = server.tick_profile
    _say("   ticks", tp)
    assert tp["ticks"] > 0
    assert tp["device_tick_us_p50"] > 0.0 and tp["host_submit_us_p50"] > 0.0
    assert 0.0 < tp["device_busy_s"] < wall


def test_config4_sixty_four_uneven_systems(contexts):
    """The seeded 64-system draw of the batched-uneven configuration (systems from
    {H2, LiH, H2O, N2, C6H6}, walkers log-uniform 256-8192, 172 650 walkers):
    one batched sweep per chunk equals one launch per system bit for bit, and
    the FP64 batched local energy matches the oracle on the first walkers of
    every system within 1e-10 relative."""
    draw = catalog.uneven_batch(64, seed=0)
    assert min(n for _, n in draw) >= 256 and max(n for _, n in draw) <= 8192 and len(draw) == 64
    _say("case config 4: 64 systems, %d walkers" % sum(n for _, n in draw))
    jobs = [(name, nw, 500 + i) for i, (name, nw) in enumerate(draw)]
    a = _walker_sets(contexts, jobs)
    b = _walker_sets(contexts, jobs)
    runner = BatchRunner(64)
    for nstep in (7, 13):
        runner.sweep([(w, nstep) for w in a])
        sweep_each([(w, nstep) for w in b])
    torch.cuda.synchronize()
    same = sum(_same(x, y) for x, y in zip(a, b))
    _say("  identical walker sets after 20 steps: %d / 64" % same)
    assert same == 64
    el = [torch.empty(w.nw, dtype=torch.float64, device="cuda") for w in a]
    runner.local_energy(list(zip(a, el)))
    torch.cuda.synchronize()
    orcs = {}
    worst = 0.0
    for (name, nw, _), w, e in zip(jobs, a, el):
        mol = contexts[name].wf.mol
        if name not in orcs:
            wf = contexts[name].wf
            ee, en = wf._jastrow_terms()
            orcs[name] = O.OracleSlaterJastrow(mol.basis, mol.atom_coords, mol.atomic_number, mol.nup, mol.ndown,
                                               jastrow_ee=float(ee.jastrow_kernel.weight),
                                               jastrow_en=float(en.jastrow_kernel.weight))
        k = 2 if name == "c6h6" else 8
        ref = orcs[name].local_energy(w.pos[:k].cpu()).detach().view(-1)
        got = e[:k].cpu()
        worst = max(worst, float(((got - ref).abs() / ref.abs()).max()))
    _say("  FP64 batched E_L vs oracle: max rel err %.2e" % worst)
    assert worst < 1e-10


def test_config5_serving_at_full_load(contexts):
    """Poisson arrivals at 2000 req/s of the serving configuration (molecules of
    the config-4 set, walkers log-uniform 128-16384, 1-20 sweeps + one E_L
    batch), in real time: every request completes, a sample of requests gives
    the same result alone as in the mix, and the plain per-request path agrees
    within 1e-12 relative."""
    import time

    reqs = poisson_requests(rate=2000.0, n=120, seed=11, wmax=16384)
    assert max(r.nwalkers for r in reqs) > 8192
    _say("case config 5: 120 requests at 2000 req/s, walkers up to %d" % max(r.nwalkers for r in reqs))
    server = VMCServer(precision="mixed", steps_per_tick=4, contexts=contexts)
    server.run(reqs[:2], realtime=False)  # warm every instance used
    t0 = time.perf_counter()
    res = server.run(reqs, realtime=True)
    s = latency_summary(res, time.perf_counter() - t0)
    _say("  ", {k: round(v, 3) if isinstance(v, float) else v for k, v in s.items()})
    assert s["requests"] == 120 and server.peak_active > 1
    mix = {r.rid: r for r in res}
    for req in reqs[::17]:
        alone = server.run([req], realtime=False)[0]
        assert np.array_equal(alone.eloc, mix[req.rid].eloc), req
    sample = reqs[::10]
    plain = {r.rid: r for r in serve_per_request(sample, contexts, realtime=False)}
    worst = max(float(np.max(np.abs(plain[k].eloc - mix[k].eloc) / np.abs(mix[k].eloc))) for k in plain)
    _say("  plain per-request path vs server on %d requests: max rel |dE_L| %.2e" % (len(plain), worst))
    assert worst < 1e-12
