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
throughput, the local-energy error of the mixed-precision path against the
FP64 path on the sampled positions, the roofline position of the sweep kernel
(HIP events on the launch stream) and two baselines of the oracle restatement
of the reference algorithm: on the host cores (bounded sample) and through
torch-ROCm on the same GPU at the same walker count (the reference's own
GPU-PyTorch path).
"""
import argparse
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
    from tests import systems

    return {"h2o": systems.h2o, "lih": systems.lih, "h2": systems.h2, "n2": systems.n2,
            "c6h6": systems.benzene}[name]()


def _restatement(mol, wf_dev):
    from oracle import vmc_oracle as O

    ee, en = wf_dev._jastrow_terms()
    return O, O.OracleSlaterJastrow(mol.basis, mol.atom_coords, mol.atomic_number, mol.nup, mol.ndown,
                                    jastrow_ee=1.0 if ee is not None else None,
                                    jastrow_en=1.0 if en is not None else None)


def measured_traffic(args, nw, steps_per_launch):
    """HBM bytes per sweep launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (scaled by a plain copy of known size), when they were
    taken on this workload; None otherwise."""
    path = os.path.join(ROOT, "profiles", "sweep_kernel_%s_%d_traffic.json" % (args.system, nw))
    if not os.path.isfile(path) or args.precision != "mixed" or steps_per_launch != 50:
        return None, None
    with open(path) as f:
        d = json.load(f)
    return d["traffic_bytes_per_launch"], os.path.relpath(path, ROOT)


def cpu_baseline(mol, wf_dev, sample_walkers=1024, min_seconds=12.0, max_steps=3000, warm=2):
    """Reference baselines: the oracle restatement of the reference Metropolis
    step (host proposal RNG, per-step host<->device copies, FP64 ATen ops) timed
    on the host CPU cores and, on the same GPU, through torch-ROCm -- the
    reference's own GPU-PyTorch path (gpu_torch_baseline)."""
    import torch

    from tests import helpers

    O, orc = _restatement(mol, wf_dev)
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    pos = helpers.positions(mol, sample_walkers, seed=3)
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

    from tests import helpers

    O, orc = _restatement(mol, wf_dev)
    orc.to_device("cuda")
    pos = helpers.positions(mol, nw, seed=3).cuda()
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
    from tests import helpers

    dev = torch.device("cuda")
    x = helpers.positions(mol, nw, seed=4).to(dev).contiguous()
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
    pos = helpers.positions(mol, nw, seed=4).cuda()
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


def serving(rates, nreq, precision="mixed", steps_per_tick=4, plain=True, seed=0):
    """Poisson request streams (config-4 molecule set, walkers 128-16384,
    1-20 sweeps + one E_L batch): continuous batching vs the plain
    per-request path on the same GPU."""
    from qmctorch_amd.serving import (SystemContext, VMCServer, catalog, latency_summary, poisson_requests,
                                      serve_per_request)

    ctx = {n: SystemContext(catalog.wavefunction(n, precision), n) for n in catalog.SYSTEM_NAMES}
    server = VMCServer(precision=precision, steps_per_tick=steps_per_tick, contexts=ctx)
    server.run(poisson_requests(1000.0, 10, seed=99), realtime=False)  # warm every kernel
    out = []
    for rate in rates:
        reqs = poisson_requests(rate, nreq, seed=seed)
        t0 = time.perf_counter()
        res = server.run(reqs, realtime=True)
        row = {"rate_req_per_s": rate, "batched": latency_summary(res, time.perf_counter() - t0),
               "peak_requests_in_flight": server.peak_active}
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
    from tests import helpers, systems

    out = {}
    # H2/STO-3G FP64, e-e Pade, 1000 walkers x 2000 steps, then E_L of the final positions
    mol = systems.h2()
    wf, _ = helpers.make_pair(mol, w_ee=1.0, perturb=0.0, precision="fp64")
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

    # LiH/6-31G CAS(2,2), CI [1, N(0, 0.1) x 3], e-e + e-n Pade: mixed vs FP64 local energy on 10^5
    # Gaussian positions (sigma 1 bohr around the nuclei, seed 1) and 10^5 positions of an FP64 chain
    mol = systems.lih()
    rng = np.random.default_rng(0)
    ci = [1.0] + list(0.1 * rng.standard_normal(3))
    wf64, _ = helpers.make_pair(mol, configs="cas(2,2)", w_ee=1.0, w_en=1.0, ci=ci, perturb=0.0, precision="fp64")
    wfmx, _ = helpers.make_pair(mol, configs="cas(2,2)", w_ee=1.0, w_en=1.0, ci=ci, perturb=0.0, precision="mixed")
    gauss = catalog.initial_positions(mol, 100000, seed=1)
    chain = Metropolis(nwalkers=10000, nstep=510, ntherm=500, step_size=0.2, nelec=mol.nelec,
                       init=mol.domain("atomic"), move={"type": "all-elec", "proba": "normal"}, cuda=True, seed=1)
    cpos = chain(wf64.pdf, pos=catalog.initial_positions(mol, 10000, seed=2), with_tqdm=False).detach()
    errs = {}
    for name, p in (("gaussian", gauss), ("fp64_chain", cpos)):
        e64 = wf64.local_energy_parts(p, "fp64")["eloc"]
        emx = wfmx.local_energy_parts(p, "mixed")["eloc"]
        d = (emx - e64).abs() / mol.nelec
        ok = torch.isfinite(d)
        errs[name] = {"positions": int(p.shape[0]), "max": float(d[ok].max()), "mean": float(d[ok].mean()),
                      "median": float(d[ok].median()),
                      "bias_of_mean": float((emx[ok].mean() - e64[ok].mean()).abs() / mol.nelec)}
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
                                   "eloc_error_per_electron_vs_fp64": errs}

    # H2O/cc-pVDZ, 8192 walkers, energy minimisation (manual gradient through the adjoint kernel,
    # freeze ao and ci, Adam lr 1e-3, 25 resampling steps per epoch)
    mol = systems.h2o()
    wf, _ = helpers.make_pair(mol, w_ee=1.0, w_en=1.0, perturb=0.0, precision=precision)
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
    from tests import helpers

    F, T = algorithmic_cost(mol, wf)
    plan = wf.device_plan()
    prm, keep = wf._params(precision)
    prec = N.QMC_PREC_MIXED if precision == "mixed" else N.QMC_PREC_FP64
    lib, stream = N.load(), N.current_stream()
    rows = []
    for nw in counts:
        x = helpers.positions(mol, nw, seed=7).cuda().contiguous()
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
    from tests import helpers

    torch.set_default_dtype(torch.float64)
    mol = build_system(args.system)
    wf, _ = helpers.make_pair(mol, w_ee=1.0, w_en=1.0, perturb=0.0, precision=args.precision)
    nw = args.walkers
    dev = torch.device("cuda")
    x = helpers.positions(mol, nw, seed=0).to(dev).contiguous()
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

    def run(step0, nstep, init, events=None):
        s = step0
        first = init
        while s < step0 + nstep:
            sw.step0 = s
            sw.nstep = min(spl, step0 + nstep - s)
            sw.init_logpsi = 1 if first else 0
            if events is not None:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
            N.check(lib.qmc_metropolis_sweep(plan.handle, prm, prec, wk, nw, sw, stream))
            if events is not None:
                e1.record()
                events.append((e0, e1, sw.nstep))
            first = False
            s += sw.nstep

    run(0, max(args.warmup, 1), True)
    torch.cuda.synchronize()
    nacc.zero_()
    events = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(args.warmup, args.steps, False, events)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = nw * args.steps / elapsed
    acc = float(nacc.sum().item()) / (nw * args.steps)
    full = [(a.elapsed_time(b) * 1e-3, n) for a, b, n in events if n == spl]
    launch_s = float(np.mean([t for t, _ in full])) if full else elapsed / max(len(events), 1)
    steps_per_launch = spl if full else args.steps

    F, T = algorithmic_cost(mol, wf)
    flops_launch = (F + 4 * T) * nw * steps_per_launch
    achieved = flops_launch / launch_s / 1e12

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

    el_rate = el_rate_of(wf.eloc_precision)
    el_rate_fp32 = el_rate_of("mixed")
    e64 = wf.local_energy_parts(pos, "fp64")["eloc"]
    emx = wf.local_energy_parts(pos, "mixed")["eloc"]
    d = ((emx - e64).abs() / mol.nelec)
    err32 = {"max": float(d.max()), "median": float(d.median()), "mean": float(d.mean()),
             "bias_of_mean": float((emx.mean() - e64.mean()).abs() / mol.nelec)}
    e_used = wf.local_energy_parts(pos)["eloc"]
    du = (e_used - e64).abs() / mol.nelec
    err = {"max": float(du.max()), "mean": float(du.mean()), "eloc_precision": wf.eloc_precision}
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
        "eloc_error_per_electron_vs_fp64": err,
        "local_energy_fp32_bulk": {"per_s": el_rate_fp32, "error_per_electron_vs_fp64": err32},
        "sweep_fp32_bulk_vs_fp64_chain": sweep_vs64,
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": "sweep_kernel",
            "launch_ms": launch_s * 1e3,
            "flops_per_walker_step": F + 4 * T,
        },
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
                       precision=args.precision)
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
