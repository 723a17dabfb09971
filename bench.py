"""VMC benchmark: walker-steps/s of the Metropolis hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--walkers 8192]
                    [--system h2o] [--precision mixed] [--no-cpu-baseline]

Workload (one "step" = one all-electron Metropolis move of every walker:
proposal, psi^2 at the proposal, accept/reject): H2O at equilibrium
(O-H 0.9572 A, 104.52 deg), cc-pVDZ (25 Cartesian AOs), 5 up + 5 down
electrons, single determinant x Pade e-e x Pade e-n Jastrow, 8192 walkers
initialised around the nuclei.  MO coefficients from the package's own RHF
(no SCF program offline) -- the cost of a step does not depend on their values.

Prints one JSON line (rank 0).  Besides the metric it reports the local-energy
throughput (FP64), the effect of the mixed-precision sweep on the sampled
chains against the FP64 sweep, the roofline position of the sweep kernel
(HIP events on the launch stream) and two baselines of the oracle restatement
of the reference algorithm: on the host cores (bounded sample) and through
torch-ROCm on the same GPU at the same walker count (the reference's own
GPU-PyTorch path).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "VMC walker-steps/s at <=1e-6 Ha/elec error vs FP64; p50/p99 request latency"
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 (vector with packed FP32 / matrix)


def algorithmic_cost(mol, wf):
    """Per walker-step flops and transcendentals of the reference algorithm
    (F_step / T_step of the benchmark definition), from the real basis."""
    ne, nat = mol.nelec, mol.natom
    npr = len(mol.basis.bas_exp)
    nao = mol.basis.nao
    nocc = wf.highest_occ_mo
    nup, ndn = mol.nup, mol.ndown
    u = 1
    F = (15 * ne * npr + 2 * (nup + ndn) * nao * nocc + u * (2.0 / 3.0) * (nup ** 3 + ndn ** 3)
         + 12 * ne * (ne - 1) / 2 + 12 * ne * nat + 25)
    T = ne * npr + ne * (ne - 1) + 2 * ne * nat + 6 * ne + 2
    return F, T


def build_system(name):
    from qmctorch_amd.serving import catalog

    return catalog.molecule(name)


def sj_wavefunction(mol, configs="ground_state", w_ee=1.0, w_en=1.0, ci=None, precision="mixed"):
    """SlaterJastrow through the package's public API (the reference's module
    paths): Pade e-e (and e-n) Jastrow, HF orbitals of the molecule, CI weights."""
    import torch

    from qmctorch_amd.wavefunction import SlaterJastrow
    from qmctorch_amd.wavefunction.jastrows.elec_elec import JastrowFactor as JEE
    from qmctorch_amd.wavefunction.jastrows.elec_elec import PadeJastrowKernel as PEE
    from qmctorch_amd.wavefunction.jastrows.elec_nuclei import JastrowFactor as JEN
    from qmctorch_amd.wavefunction.jastrows.elec_nuclei import PadeJastrowKernel as PEN

    jas = [JEE(mol, PEE, kernel_kwargs={"w": w_ee}, cuda=True)]
    if w_en is not None:
        jas.append(JEN(mol, PEN, kernel_kwargs={"w": w_en}, cuda=True))
    wf = SlaterJastrow(mol, jastrow=jas, configs=configs, cuda=True, precision=precision)
    if ci is not None:
        with torch.no_grad():
            wf.fc.weight.copy_(torch.as_tensor(ci, dtype=torch.float64).view(1, -1))
    return wf


def start_positions(mol, nw, seed):
    """Walker starts from the package's device initialiser (atomic domain)."""
    from qmctorch_amd.sampler.walkers import Walkers

    return Walkers(nw, mol.nelec, cuda=True).draw(mol.domain("atomic"), nw, mol.nelec, seed=seed).contiguous()


def _restatement(mol, wf_dev):
    from oracle import vmc_oracle as O

    ee, en = wf_dev._jastrow_terms()
    return O, O.OracleSlaterJastrow(mol.basis, mol.atom_coords, mol.atomic_number, mol.nup, mol.ndown,
                                    jastrow_ee=1.0 if ee is not None else None,
                                    jastrow_en=1.0 if en is not None else None)


def measured_traffic(args, nw, steps_per_launch):
    """HBM bytes per sweep launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (scaled by a plain copy of known size) on this workload
    and launch length; None when no pass was taken for them."""
    path = os.path.join(ROOT, "profiles", "sweep_kernel_%s_%d_traffic.json" % (args.system, nw))
    if not os.path.isfile(path) or args.precision != "mixed":
        return None, None
    with open(path) as f:
        d = json.load(f)
    per = d.get("traffic_bytes_per_launch_by_steps", {})
    v = per.get(str(steps_per_launch))
    if v is None and steps_per_launch == d.get("steps_per_launch", 50):
        v = d.get("traffic_bytes_per_launch")
    return v, (os.path.relpath(path, ROOT) if v is not None else None)


def cpu_baseline(mol, wf_dev, sample_walkers=1024, min_seconds=12.0, max_steps=3000, warm=2):
    """Reference baselines: the oracle restatement of the reference Metropolis
    step (host proposal RNG, per-step host<->device copies, FP64 ATen ops) timed
    on the host CPU cores and, on the same GPU, through torch-ROCm -- the
    reference's own GPU-PyTorch path (gpu_torch_baseline)."""
    import torch

    O, orc = _restatement(mol, wf_dev)
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    pos = start_positions(mol, sample_walkers, seed=3).cpu()
    O.metropolis(orc.pdf, pos.clone(), warm, step_size=0.2, seed=0)
    # chunks of 50 steps until at least min_seconds of CPU work
    sample_steps, dt = 0, 0.0
    x = pos.clone()
    while dt < min_seconds and sample_steps < max_steps:
        t0 = time.perf_counter()
        kept, _ = O.metropolis(orc.pdf, x, 50, step_size=0.2, seed=sample_steps)
        dt += time.perf_counter() - t0
        x = kept[-sample_walkers:].clone()
        sample_steps += 50
    cpu = {"value": sample_walkers * sample_steps / dt, "unit": "walker-steps/s", "cores": cores,
           "kind": "port",
           "sample": "%d walkers x %d Metropolis steps of the same system, torch CPU float64 (%.1f s)"
                     % (sample_walkers, sample_steps, dt)}
    return cpu


def gpu_torch_baseline(mol, wf_dev, nw, sample_steps=200, warm=3):
    """The same restatement on torch-ROCm (FP64, same walker count as the
    workload): the reference's GPU-PyTorch walker-steps/s on this GPU."""
    import torch

    O, orc = _restatement(mol, wf_dev)
    orc.to_device("cuda")
    pos = start_positions(mol, nw, seed=3)
    O.metropolis(orc.pdf, pos.clone(), warm, step_size=0.2, seed=0, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    O.metropolis(orc.pdf, pos.clone(), sample_steps, step_size=0.2, seed=0, device="cuda")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"value": nw * sample_steps / dt, "unit": "walker-steps/s", "kind": "port",
            "sample": "%d walkers x %d Metropolis steps, torch-ROCm float64 on the same GPU (%.1f s)"
                      % (nw, sample_steps, dt)}


def hmc_leg(mol, wf, nw, steps=200, L=10, eps=0.05, ref_steps=3, reference=True):
    """Hamiltonian Monte Carlo (qmc_hmc_sweep) on the same system and walkers:
    HMC steps/s and drift evaluations/s (L per step), acceptance, and the
    reference algorithm (autograd leapfrog, hamiltonian.py:143-201) on
    torch-ROCm FP64 on the same GPU."""
    import numpy as np
    import torch

    from qmctorch_amd import _native as N

    dev = torch.device("cuda")
    x = start_positions(mol, nw, seed=4)
    logabs = torch.empty(nw, dtype=torch.float64, device=dev)
    sign = torch.empty_like(logabs)
    nacc = torch.zeros(nw, dtype=torch.int32, device=dev)
    wk = N.QmcWalkers()
    wk.pos, wk.logabs, wk.sign, wk.naccept, wk.walker_id = x.data_ptr(), logabs.data_ptr(), sign.data_ptr(), \
        nacc.data_ptr(), 0
    hm = N.QmcHmcDesc()
    hm.traj_length, hm.step_size, hm.seed, hm.keep, hm.keep_every = L, eps, 2024, 0, 1
    plan = wf.device_plan()
    prm, keep = wf._params(wf.precision)
    prec = N.QMC_PREC_MIXED if wf.precision == "mixed" else N.QMC_PREC_FP64
    lib, stream = N.load(), N.current_stream()
    per_launch = 10

    def run(step0, nstep, events=None):
        for s0 in range(step0, step0 + nstep, per_launch):
            hm.step0, hm.nstep = s0, min(per_launch, step0 + nstep - s0)
            if events is not None:
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
            N.check(lib.qmc_hmc_sweep(plan.handle, prm, prec, wk, nw, hm, stream))
            if events is not None:
                e1.record()
                events.append((e0, e1))

    run(0, 20)
    torch.cuda.synchronize()
    nacc.zero_()
    events = []
    t0 = time.perf_counter()
    run(20, steps, events)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    launch_ms = float(np.mean([a.elapsed_time(b) for a, b in events]))
    out = {"walkers": nw, "traj_length": L, "step_size": eps, "steps_per_launch": per_launch,
           "hmc_steps_per_s": nw * steps / dt, "drift_evals_per_s": nw * steps * L / dt,
           "launch_ms": launch_ms, "acceptance": float(nacc.sum().item()) / (nw * steps),
           "precision": wf.precision}
    if not reference:
        return out
    # baseline: the reference algorithm on torch-ROCm (FP64), same walkers, a few steps
    O, orc = _restatement(mol, wf)
    orc.to_device("cuda")
    pos = start_positions(mol, nw, seed=4)
    O.hamiltonian(orc.pdf, pos.clone(), 1, step_size=eps, L=L, seed=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    O.hamiltonian(orc.pdf, pos.clone(), ref_steps, step_size=eps, L=L, seed=0)
    torch.cuda.synchronize()
    rdt = time.perf_counter() - t0
    out["reference_gpu_torch"] = {"hmc_steps_per_s": nw * ref_steps / rdt,
                                  "sample": "%d walkers x %d HMC steps (L=%d), autograd leapfrog, torch-ROCm "
                                            "float64 (%.1f s)" % (nw, ref_steps, L, rdt)}
    out["speedup_vs_reference_gpu_torch"] = out["hmc_steps_per_s"] / out["reference_gpu_torch"]["hmc_steps_per_s"]
    return out


def sweep_precision_effect(mol, wf, nw, steps=2000, seed=77):
    """Effect of the FP32-bulk sweep on the sampled energy: the same chains
    (same start, seed and walker ids) run by the FP32-bulk and the FP64 sweep;
    a walker's trajectory differs only after an acceptance decision flips.
    Reports the fraction of walkers whose final positions differ and the
    difference of the FP64 energy means per electron."""
    import torch

    from qmctorch_amd.sampler import Metropolis
    from qmctorch_amd.serving import catalog

    start = catalog.initial_positions(mol, nw, seed=seed)
    out = {}
    for prec in ("mixed", "fp64"):
        old = wf.precision
        wf.precision = prec
        try:
            s = Metropolis(nwalkers=nw, nstep=steps, step_size=0.2, nelec=mol.nelec, init=mol.domain("atomic"),
                           move={"type": "all-elec", "proba": "normal"}, cuda=True, seed=seed)
            out[prec] = s(wf.pdf, pos=start, with_tqdm=False).detach()
        finally:
            wf.precision = old
    diff = (out["mixed"] != out["fp64"]).any(dim=1)
    e32 = wf.local_energy_parts(out["mixed"], "fp64")["eloc"]
    e64 = wf.local_energy_parts(out["fp64"], "fp64")["eloc"]
    err = float(torch.sqrt(e64.var() / nw)) / mol.nelec
    return {"walkers": nw, "steps": steps, "fraction_of_walkers_diverged": float(diff.double().mean()),
            "energy_difference_per_electron": float((e32.mean() - e64.mean()).abs()) / mol.nelec,
            "statistical_error_per_electron": err}


def batched_uneven(nsys=64, steps=100, seed=0, precision="mixed"):
    """64 systems drawn from {H2, LiH, H2O, N2, C6H6}, walkers log-uniform
    256-8192: sweeps of all systems in one batched call per chunk of steps
    (one launch per register bucket) vs the same sweeps one system per launch."""
    import torch

    from qmctorch_amd.serving import BatchRunner, SystemContext, WalkerSet, catalog, sweep_each

    draw = catalog.uneven_batch(nsys, seed=seed)
    ctx = {n: SystemContext(catalog.wavefunction(n, precision), n) for n in sorted({d[0] for d in draw})}

    def walkers():
        return [WalkerSet(ctx[n], catalog.initial_positions(ctx[n].wf.mol, nw, 1000 + i), 1000 + i)
                for i, (n, nw) in enumerate(draw)]

    runner = BatchRunner(max(nsys, 1))
    total = sum(nw for _, nw in draw) * steps
    chunk = 50
    res = {}
    for mode in ("batched", "per_system"):
        ws = walkers()
        warm = [(w, 5) for w in ws]
        runner.sweep(warm) if mode == "batched" else sweep_each(warm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            n = min(chunk, steps - done)
            jobs = [(w, n) for w in ws]
            runner.sweep(jobs) if mode == "batched" else sweep_each(jobs)
            done += n
        torch.cuda.synchronize()
        res[mode] = total / (time.perf_counter() - t0)
    mix = {}
    for n, _ in draw:
        mix[n] = mix.get(n, 0) + 1
    return {"systems": nsys, "mix": mix, "walkers": sum(nw for _, nw in draw), "steps": steps,
            "walker_steps_per_s_batched": res["batched"], "walker_steps_per_s_per_system": res["per_system"],
            "speedup": res["batched"] / res["per_system"]}


def serving(rates, nreq, precision="mixed", steps_per_tick=4, plain=True, seed=0, lookahead=1):
    """Poisson request streams (config-4 molecule set, walkers 128-16384,
    1-20 sweeps + one E_L batch): continuous batching vs the plain
    per-request path on the same GPU."""
    from qmctorch_amd.serving import (SystemContext, VMCServer, catalog, latency_summary, poisson_requests,
                                      serve_per_request)

    ctx = {n: SystemContext(catalog.wavefunction(n, precision), n) for n in catalog.SYSTEM_NAMES}
    server = VMCServer(precision=precision, steps_per_tick=steps_per_tick, contexts=ctx, lookahead=lookahead)
    server.run(poisson_requests(1000.0, 10, seed=99), realtime=False)  # warm every kernel
    out = []
    for rate in rates:
        reqs = poisson_requests(rate, nreq, seed=seed)
        t0 = time.perf_counter()
        res = server.run(reqs, realtime=True, profile_ticks=True)
        row = {"rate_req_per_s": rate, "batched": latency_summary(res, time.perf_counter() - t0),
               "peak_requests_in_flight": server.peak_active, "ticks": server.tick_profile}
        if plain:
            t0 = time.perf_counter()
            res_p = serve_per_request(reqs, ctx, realtime=True)
            row["per_request"] = latency_summary(res_p, time.perf_counter() - t0)
        out.append(row)
    return out


def config_legs(precision="mixed"):
    """The other benchmark configurations at their stated sizes (one short run each)."""
    import numpy as np
    import torch

    from qmctorch_amd.sampler import Metropolis
    from qmctorch_amd.serving import catalog
    from qmctorch_amd.solver import Solver
    out = {}
    # H2/STO-3G FP64, e-e Pade, 1000 walkers x 2000 steps, then E_L of the final positions
    mol = build_system("h2")
    wf = sj_wavefunction(mol, w_ee=1.0, w_en=None, precision="fp64")
    pos0 = catalog.initial_positions(mol, 1000, seed=0)
    sampler = Metropolis(nwalkers=1000, nstep=2000, step_size=0.2, nelec=mol.nelec, init=mol.domain("atomic"),
                         move={"type": "all-elec", "proba": "normal"}, cuda=True, seed=0)
    # warm-up outside the timed region: plan upload and first launch of this kernel instance
    Metropolis(nwalkers=1000, nstep=50, step_size=0.2, nelec=mol.nelec, init=mol.domain("atomic"),
               move={"type": "all-elec", "proba": "normal"}, cuda=True, seed=1)(wf.pdf, pos=pos0, with_tqdm=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x = sampler(wf.pdf, pos=pos0, with_tqdm=False).detach()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    e = wf.local_energy(x).reshape(-1)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    out["h2_sto3g_fp64"] = {"walkers": 1000, "steps": 2000, "walker_steps_per_s": 1000 * 2000 / (t1 - t0),
                            "local_energy_s": t2 - t1, "energy": float(e.mean()),
                            "error": float(e.std() / np.sqrt(e.numel())), "acceptance": sampler.last_acceptance}

    # LiH/6-31G CAS(2,2), CI [1, N(0, 0.1) x 3], e-e + e-n Pade.  Local energies are FP64 in both
    # precisions (error 0); the mixed path's error enters through log|psi| of the sweep: its
    # difference from FP64 on 10^5 Gaussian positions (sigma 1 bohr around the nuclei, seed 1) and
    # on 10^5 positions of an FP64 chain, and the acceptance decisions it flips along a chain
    mol = build_system("lih")
    rng = np.random.default_rng(0)
    ci = [1.0] + list(0.1 * rng.standard_normal(3))
    wf64 = sj_wavefunction(mol, configs="cas(2,2)", ci=ci, precision="fp64")
    wfmx = sj_wavefunction(mol, configs="cas(2,2)", ci=ci, precision="mixed")
    gauss = catalog.initial_positions(mol, 100000, seed=1)
    chain = Metropolis(nwalkers=10000, nstep=510, ntherm=500, step_size=0.2, nelec=mol.nelec,
                       init=mol.domain("atomic"), move={"type": "all-elec", "proba": "normal"}, cuda=True, seed=1)
    cpos = chain(wf64.pdf, pos=catalog.initial_positions(mol, 10000, seed=2), with_tqdm=False).detach()
    errs = {}
    for name, p in (("gaussian", gauss), ("fp64_chain", cpos)):
        l64, s64 = wf64.log_psi(p, "fp64")
        lmx, smx = wfmx.log_psi(p, "mixed")
        d = (lmx - l64).abs()
        ok = torch.isfinite(d)
        errs[name] = {"positions": int(p.shape[0]), "max_abs_dlogpsi": float(d[ok].max()),
                      "mean_abs_dlogpsi": float(d[ok].mean()), "sign_flips": int((smx != s64).sum())}
    errs["local_energy_error_per_electron"] = 0.0
    errs["chain"] = sweep_precision_effect(mol, wfmx, 4096, steps=1000)
    nw = 16384
    st = catalog.initial_positions(mol, nw, seed=3)
    sm = Metropolis(nwalkers=nw, nstep=500, step_size=0.2, nelec=mol.nelec, init=mol.domain("atomic"),
                    move={"type": "all-elec", "proba": "normal"}, cuda=True, seed=3)
    sm(wfmx.pdf, pos=st, with_tqdm=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sm(wfmx.pdf, pos=st, with_tqdm=False)
    torch.cuda.synchronize()
    out["lih_631g_cas22_mixed"] = {"walkers": nw, "walker_steps_per_s": nw * 500 / (time.perf_counter() - t0),
                                   "mixed_vs_fp64": errs}

    # H2O/cc-pVDZ, 8192 walkers, energy minimisation (manual gradient through the adjoint kernel,
    # freeze ao and ci, Adam lr 1e-3, 25 resampling steps per epoch)
    mol = build_system("h2o")
    wf = sj_wavefunction(mol, precision=precision)
    samp = Metropolis(nwalkers=8192, nstep=100, step_size=0.2, nelec=mol.nelec, init=mol.domain("atomic"),
                      move={"type": "all-elec", "proba": "normal"}, cuda=True, seed=4)
    samp.walkers.initialize(pos=catalog.initial_positions(mol, 8192, seed=4))
    opt = torch.optim.Adam(wf.parameters(), lr=1e-3)
    solver = Solver(wf=wf, sampler=samp, optimizer=opt)
    solver.configure(freeze=["ao", "ci"], track=["local_energy"])
    solver.prepare_optimization(None, None)  # initial sampling: 100 steps from the seeded positions
    solver.run_epochs(1, verbose=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nep = 10
    solver.run_epochs(nep, verbose=False)
    torch.cuda.synchronize()
    ep = (time.perf_counter() - t0) / nep
    obs = solver.observable
    en = [float(v) for v in obs.energy] if hasattr(obs, "energy") else []
    out["h2o_ccpvdz_optimisation"] = {"walkers": 8192, "epochs": nep, "s_per_epoch": ep,
                                      "energy_first_last": en[:1] + en[-1:]}
    return out


def walker_scan(mol, wf, precision, counts=(16384, 32768, 65536), launches=8, spl=50):
    """The same sweep kernel at larger walker counts (one launch = spl steps):
    how far the 8192-walker launch is from filling the chip (1 368 waves on
    1 024 SIMDs there)."""
    import numpy as np
    import torch

    from qmctorch_amd import _native as N

    F, T = algorithmic_cost(mol, wf)
    plan = wf.device_plan()
    prm, keep = wf._params(precision)
    prec = N.QMC_PREC_MIXED if precision == "mixed" else N.QMC_PREC_FP64
    lib, stream = N.load(), N.current_stream()
    rows = []
    for nw in counts:
        x = start_positions(mol, nw, seed=7)
        la = torch.empty(nw, dtype=torch.float64, device="cuda")
        sg = torch.empty_like(la)
        nacc = torch.zeros(nw, dtype=torch.int32, device="cuda")
        wk = N.QmcWalkers()
        wk.pos, wk.logabs, wk.sign, wk.naccept, wk.walker_id = x.data_ptr(), la.data_ptr(), sg.data_ptr(), \
            nacc.data_ptr(), 0
        sw = N.QmcSweepDesc()
        sw.sigma2, sw.seed, sw.keep, sw.keep_every, sw.step_size = 0.2 / (2 * math.sqrt(2 * math.log(2.0))), 99, 0, 1, 0.2
        times = []
        for i in range(launches + 2):
            sw.step0, sw.nstep, sw.init_logpsi = i * spl, spl, 1 if i == 0 else 0
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            N.check(lib.qmc_metropolis_sweep(plan.handle, prm, prec, wk, nw, sw, stream))
            e1.record()
            if i >= 2:
                times.append((e0, e1))
        torch.cuda.synchronize()
        ms = float(np.mean([a.elapsed_time(b) for a, b in times]))
        tf = (F + 4 * T) * nw * spl / (ms * 1e-3) / 1e12
        rows.append({"walkers": nw, "launch_ms": ms, "walker_steps_per_s": nw * spl / (ms * 1e-3),
                     "achieved_tflops": tf, "frac": tf / PEAK_FP32_TFLOPS})
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--walkers", type=int, default=8192)
    ap.add_argument("--system", default="h2o")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--steps-per-launch", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the H2 / LiH / H2O-optimisation legs")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-system batched-launch leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the request-loop leg")
    ap.add_argument("--no-hmc", action="store_true", help="skip the Hamiltonian Monte Carlo leg")
    ap.add_argument("--no-scan", action="store_true", help="skip the larger-walker-count sweep scan")
    ap.add_argument("--serving-rates", default="50,500,2000", help="comma-separated Poisson rates (req/s)")
    ap.add_argument("--serving-requests", type=int, default=300)
    ap.add_argument("--serving-lookahead", type=int, default=1, help="ticks the server loop may run ahead")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (known bytes) to scale HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the VMC hot path on a single MI355X "
              "(walkers shard trivially across GPUs); --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from qmctorch_amd import _native as N

    torch.set_default_dtype(torch.float64)
    mol = build_system(args.system)
    wf = sj_wavefunction(mol, precision=args.precision)
    nw = args.walkers
    dev = torch.device("cuda")
    x = start_positions(mol, nw, seed=0)
    logabs = torch.empty(nw, dtype=torch.float64, device=dev)
    sign = torch.empty(nw, dtype=torch.float64, device=dev)
    nacc = torch.zeros(nw, dtype=torch.int32, device=dev)
    plan = wf.device_plan()
    prm, keep = wf._params(args.precision)
    prec = N.QMC_PREC_MIXED if args.precision == "mixed" else N.QMC_PREC_FP64
    lib = N.load()
    wk = N.QmcWalkers()
    wk.pos, wk.logabs, wk.sign, wk.naccept, wk.walker_id = x.data_ptr(), logabs.data_ptr(), sign.data_ptr(), \
        nacc.data_ptr(), 0
    sw = N.QmcSweepDesc()
    sw.sigma2 = 0.2 / (2 * math.sqrt(2 * math.log(2.0)))
    sw.seed = 12345
    sw.keep = 0
    sw.keep_every = 1
    sw.step_size = 0.2
    stream = N.current_stream()
    spl = args.steps_per_launch
    # launch arguments built once: the timed region is ctypes calls of the sweep only
    sweep_fn = lib.qmc_metropolis_sweep
    plan_h, prm_ref, wk_ref, sw_ref = plan.handle, ctypes.byref(prm), ctypes.byref(wk), ctypes.byref(sw)

    def launches(step0, nstep):
        out, s = [], step0
        while s < step0 + nstep:
            n = min(spl, step0 + nstep - s)
            out.append((s, n))
            s += n
        return out

    def run(plan_launches, init, events=None):
        for k, (s0, n) in enumerate(plan_launches):
            sw.step0, sw.nstep, sw.init_logpsi = s0, n, 1 if (init and k == 0) else 0
            if events is not None:
                events[k][0].record()
            rc = sweep_fn(plan_h, prm_ref, prec, wk_ref, nw, sw_ref, stream)
            if events is not None:
                events[k][1].record()
            if rc:
                N.check(rc)

    # set-up (untimed): the walkers are thermalised from their atomic start, as
    # the sampled state a VMC run works on, then the W warm-up steps
    therm = 200
    run(launches(0, therm), True)
    run(launches(therm, max(args.warmup, 1)), False)
    torch.cuda.synchronize()
    nacc.zero_()
    timed = launches(therm + max(args.warmup, 1), args.steps)
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in timed]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(timed, False)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = nw * args.steps / elapsed
    acc = float(nacc.sum().item()) / (nw * args.steps)
    # device time: launches of the same lengths, queued behind a short untimed
    # launch so that every event pair brackets its kernel alone (on an idle GPU
    # a start event completes before the host has submitted the launch, and
    # the pair would also hold the submission latency)
    s_dev = timed[-1][0] + timed[-1][1]
    run([(s_dev, 5)], False)
    run([(s_dev + 5 + s0 - timed[0][0], n) for s0, n in timed], False, events)
    torch.cuda.synchronize()
    launch_s = [a.elapsed_time(b) * 1e-3 for a, b in events]
    device_s = float(np.sum(launch_s))
    steps_per_launch = timed[0][1]

    F, T = algorithmic_cost(mol, wf)
    flops = (F + 4 * T) * nw * args.steps
    achieved = flops / device_s / 1e12

    # the table-driven kernels (any basis layout) on the same workload and launch length
    table = None
    spec_name = plan.specialization
    if spec_name is not None:
        plan.set_specialized(False)
        try:
            tl = launches(s_dev, args.steps)
            tev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in tl]
            run(launches(0, steps_per_launch), False)  # first launch of this instance
            torch.cuda.synchronize()
            run([(s_dev, 5)], False)  # the timed launches queue behind it, as above
            run(tl, False, tev)
            torch.cuda.synchronize()
            tdev = float(np.sum([a.elapsed_time(b) * 1e-3 for a, b in tev]))
            nst = sum(n for _, n in tl)
            table = {"walker_steps_per_s_device": nw * nst / tdev, "launch_ms": tdev / len(tl) * 1e3,
                     "steps_per_launch": tl[0][1], "frac": (F + 4 * T) * nw * nst / tdev / 1e12 / PEAK_FP32_TFLOPS,
                     "note": "table-driven AO rows (any basis layout); the headline uses the rows generated "
                             "for this layout (%s)" % spec_name}
        finally:
            plan.set_specialized(True)

    # local-energy throughput on the sampled positions: the FP64 kernel the
    # wavefunction uses for E_L, and the FP32-bulk kernel with its error
    pos = x.clone()

    def el_rate_of(prec_name):
        wf.local_energy_parts(pos, prec_name)
        torch.cuda.synchronize()
        reps = 5
        t1 = time.perf_counter()
        for _ in range(reps):
            wf.local_energy_parts(pos, prec_name)
        torch.cuda.synchronize()
        return reps * nw / (time.perf_counter() - t1)

    el_rate = el_rate_of("fp64")
    sweep_vs64 = sweep_precision_effect(mol, wf, nw) if args.precision == "mixed" else None

    traffic, traffic_src = measured_traffic(args, nw, steps_per_launch)
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "device_ms_per_step": device_s / args.steps * 1e3,
        "walker_steps_per_s_device": nw * args.steps / device_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": "%s Metropolis all-electron sweep, %d walkers, step 0.2 bohr, SJ ee+en Pade, %s"
                        % (args.system, nw, args.precision),
            "system": args.system, "walkers": nw, "nelec": mol.nelec, "nao": mol.basis.nao,
            "nprim": len(mol.basis.bas_exp), "steps_per_launch": steps_per_launch,
            "parallelism": "single GPU",
        },
        "acceptance": acc,
        "local_energy_per_s": el_rate,
        "local_energy_precision": "fp64",
        "sweep_fp32_bulk_vs_fp64_chain": sweep_vs64,
        "roofline": {
            "bound": "valu_fp32",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": "sweep_kernel_wlds" if spec_name else "sweep_kernel",
            "launch_ms": float(np.mean(launch_s)) * 1e3,
            "launches": len(launch_s),
            "flops_per_walker_step": F + 4 * T,
        },
        "table_driven": table,
    }
    if not args.no_scan:
        res["walker_scan"] = walker_scan(mol, wf, args.precision)
    if not args.no_hmc:
        res["hmc"] = hmc_leg(mol, wf, nw, reference=not args.no_cpu_baseline)
    if not args.no_configs:
        res["configs"] = config_legs(args.precision)
    if not args.no_batched:
        res["batched_uneven"] = batched_uneven(precision=args.precision)
    if not args.no_serving:
        rows = serving([float(r) for r in args.serving_rates.split(",")], args.serving_requests,
                       precision=args.precision, lookahead=args.serving_lookahead)
        res["serving"] = rows
        head = next((r for r in rows if r["rate_req_per_s"] == 500.0), rows[0])
        res["p50_ms"] = head["batched"]["p50_ms"]
        res["p99_ms"] = head["batched"]["p99_ms"]
        res["p50_p99_at_rate_req_per_s"] = head["rate_req_per_s"]
    if args.copy_calibration:
        src = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device=dev).uniform_()
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        res["copy_calibration_bytes"] = {"read": src.numel() * 8, "write": dst.numel() * 8}
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(mol, wf)
        g = gpu_torch_baseline(mol, wf, nw)
        res["reference_gpu_torch"] = g
        res["speedup_vs_reference_gpu_torch"] = value / g["value"]
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
