"""Build libqmcmi355x.so (HIP, gfx950) in-tree.

    python -m qmctorch_amd.csrc.build [--verbose] [--resource-usage] [--incremental]

Each kernel family is its own translation unit, compiled in parallel and
linked into qmctorch_amd/libqmcmi355x.so (next to the package, so the library
travels with the source tree).  Packed FP32 instructions are disabled
(-packed-fp32-ops); the compiler would otherwise form v_pk_* from float code.
"""
import argparse
import os
import re
import subprocess
import time
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libqmcmi355x.so")
# (source, object tag, extra defines): kernel families are compiled once per radial type
_FAMILIES = ["qmc_wf_f32.hip", "qmc_wf_f64.hip", "qmc_sweep_f32.hip", "qmc_sweep_f64.hip", "qmc_grad.hip",
             "qmc_batch_sweep_f32.hip", "qmc_batch_sweep_f64.hip", "qmc_batch_wf_f32.hip", "qmc_batch_wf_f64.hip",
             "qmc_exc.hip", "qmc_exc_hmc.hip",
]
# HMC / drift kernels: single-configuration (MC=1) and multi-configuration (MC=8) instances
_HMC = [("qmc_hmc_f32.hip", mc) for mc in (1, 8)] + [("qmc_hmc_f64.hip", mc) for mc in (1, 8)]
UNITS = [("qmc_abi.hip", "qmc_abi", []), ("qmc_api.hip", "qmc_api", []), ("qmc_noalloc.hip", "qmc_noalloc", [])] + [
    (f, f.replace(".hip", "_r%d" % r), ["-DQMC_RAD=%d" % r]) for f in _FAMILIES for r in (0, 1)
] + [(f, f.replace(".hip", "_mc%d_r%d" % (mc, r)), ["-DQMC_RAD=%d" % r, "-DQMC_HMC_MC=%d" % mc])
     for f, mc in _HMC for r in (0, 1)]
SOURCES = [u[0] for u in UNITS]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-mcode-object-version=5",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-I", HERE,
]


def _obj_dir() -> str:
    d = os.path.join(HERE, "_build")
    os.makedirs(d, exist_ok=True)
    return d


def compile_one(unit, extra=()) -> subprocess.CompletedProcess:
    """Compile one unit into a temporary object, moved into place on success
    and dated to the start of the compile: a source edited while hipcc ran
    (its host and device passes read the sources at different times) leaves
    the object older than the edit, so an incremental build recompiles it."""
    src, tag, defs = unit
    obj = os.path.join(_obj_dir(), tag + ".o")
    tmp = os.path.join(_obj_dir(), "%s.tmp%d.o" % (tag, os.getpid()))
    t0 = time.time()
    cmd = [HIPCC, *FLAGS, *defs, *extra, "-c", src if os.path.isabs(src) else os.path.join(HERE, src), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode == 0:
        os.replace(tmp, obj)
        os.utime(obj, (t0, t0))
    elif os.path.exists(tmp):
        os.remove(tmp)
    return r


def spec_units():
    """Units of the kernels generated for the benchmark molecules' basis layouts
    (written into _build/spec by specialize.py) plus their registry."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from qmctorch_amd.csrc import specialize

    files = specialize.write_sources(os.path.join(_obj_dir(), "spec"))
    return [(f, "spec_" + os.path.basename(f)[:-4], []) for f in files]


_INCLUDE = re.compile(r'^\s*#\s*include\s+"([^"]+)"', re.M)


def _deps(path: str, seen=None) -> set:
    """The file and every quoted include it reaches in csrc/ and include/."""
    seen = set() if seen is None else seen
    if path in seen or not os.path.isfile(path):
        return seen
    seen.add(path)
    with open(path) as f:
        for name in _INCLUDE.findall(f.read()):
            for d in (os.path.dirname(path), HERE, os.path.join(ROOT, "include")):
                cand = os.path.join(d, name)
                if os.path.isfile(cand):
                    _deps(cand, seen)
                    break
    return seen


def _stale(unit) -> bool:
    obj = os.path.join(_obj_dir(), unit[1] + ".o")
    src = unit[0] if os.path.isabs(unit[0]) else os.path.join(HERE, unit[0])
    if not os.path.isfile(obj):
        return True
    return os.path.getmtime(obj) < max(os.path.getmtime(d) for d in _deps(src))


def build(verbose: bool = False, resource_usage: bool = False, jobs: int = 0, incremental: bool = False) -> str:
    """Compile every unit and link the library.  With incremental=True only the
    units older than their source or a header they include are recompiled (a
    developer shortcut; the default always compiles everything)."""
    extra = ["-Rpass-analysis=kernel-resource-usage"] if resource_usage else []
    units = UNITS + spec_units()
    if incremental:
        units = [u for u in units if _stale(u)]
    # the largest units first, so the parallel build ends sooner
    units = sorted(units, key=lambda u: 0 if u[0].startswith(os.path.join(_obj_dir(), "spec")) else 1)
    jobs = jobs or max(1, min(len(units), os.cpu_count() or 4, 16))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda u: compile_one(u, extra), units))
    for unit, r in zip(units, results):
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed on %s" % unit[1])
        if verbose or resource_usage:
            sys.stderr.write(r.stderr)
    objs = [os.path.join(_obj_dir(), u[1] + ".o") for u in UNITS + spec_units()]
    tmp = LIB + ".tmp%d" % os.getpid()
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libqmcmi355x.so failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--resource-usage", action="store_true")
    ap.add_argument("-j", type=int, default=0)
    ap.add_argument("--incremental", action="store_true", help="recompile only out-of-date units")
    a = ap.parse_args()
    print(build(a.verbose, a.resource_usage, a.j, a.incremental))
