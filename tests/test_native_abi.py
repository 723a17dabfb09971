"""CPU checks of the C-ABI boundary: the library loads, exports every entry
point declared in include/qmc_mi355x.h, and the ctypes structures match the
C layout (checked by compiling the header with gcc)."""
import ctypes
import os
import re
import subprocess

import pytest

from qmctorch_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qmc_mi355x.h")


def _declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|size_t)\s+(qmc_\w+)\s*\(", src, flags=re.M)))


def test_library_exports_every_declared_symbol():
    if not os.path.isfile(N.LIBPATH):
        pytest.skip("libqmcmi355x.so not built")
    lib = N.load()
    names = _declared_functions()
    assert len(names) >= 8
    for name in names:
        assert hasattr(lib, name), name
        assert name in N.EXPORTS, "binding missing for %s" % name
    assert lib.qmc_version().decode().startswith("libqmcmi355x")


def test_error_path_without_gpu_calls():
    if not os.path.isfile(N.LIBPATH):
        pytest.skip("libqmcmi355x.so not built")
    lib = N.load()
    rc = lib.qmc_plan_create(None, None)
    assert rc == 1 and "null" in lib.qmc_last_error().decode()


def test_ctypes_struct_layout_matches_header(tmp_path):
    structs = {
        "QmcSystemDesc": N.QmcSystemDesc,
        "QmcParams": N.QmcParams,
        "QmcWfOut": N.QmcWfOut,
        "QmcWalkers": N.QmcWalkers,
        "QmcSweepDesc": N.QmcSweepDesc,
        "QmcHmcDesc": N.QmcHmcDesc,
        "QmcSweepJob": N.QmcSweepJob,
        "QmcEvalJob": N.QmcEvalJob,
        "QmcParamViews": N.QmcParamViews,
        "QmcPlanSizes": N.QmcPlanSizes,
        "QmcJastrowOut": N.QmcJastrowOut,
        "QmcAoOut": N.QmcAoOut,
        "QmcWalkerInit": N.QmcWalkerInit,
        "QmcGradOut": N.QmcGradOut,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "qmc_mi355x.h"', "int main(void) {"]
    for sname, cls in structs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (sname, sname))
        for fname, _ in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (sname, fname, sname, fname))
    lines.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(c), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n") if l)
    for sname, cls in structs.items():
        assert int(out[sname]) == ctypes.sizeof(cls), sname
        for fname, _ in cls._fields_:
            assert int(out["%s.%s" % (sname, fname)]) == getattr(cls, fname).offset, (sname, fname)


# ---------------------------------------------------------------- Philox
M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox_host(ctr, key):
    """Philox4x32-10 from Salmon et al., SC'11 (independent host restatement)."""
    c = list(ctr)
    k0, k1 = key
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k0) & MASK, p1 & MASK, ((p0 >> 32) ^ c[3] ^ k1) & MASK, p0 & MASK]
        k0 = (k0 + W0) & MASK
        k1 = (k1 + W1) & MASK
    return c


KAT = [  # Random123 known-answer vectors for philox4x32-10
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((MASK, MASK, MASK, MASK), (MASK, MASK), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,ref", KAT)
def test_philox_host_known_answers(ctr, key, ref):
    assert tuple(philox_host(ctr, key)) == ref


def test_batch_entry_points_reject_bad_arguments_without_gpu_calls():
    if not os.path.isfile(N.LIBPATH):
        pytest.skip("libqmcmi355x.so not built")
    lib = N.load()
    h = ctypes.c_void_p()
    assert lib.qmc_batch_create(0, ctypes.byref(h)) != 0 and not h.value
    assert "batch" in lib.qmc_last_error().decode()
    assert lib.qmc_metropolis_sweep_batch(None, N.QMC_PREC_MIXED, None, 1, None) != 0
    assert lib.qmc_wf_eval_batch(None, N.QMC_PREC_MIXED, N.QMC_EVAL_ELOC, None, 1, None) != 0
    assert "null" in lib.qmc_last_error().decode()


def test_acceptance_bits_uniform_and_independent():
    """The sweep's acceptance uniform is assembled from the low bytes of the four
    Philox words of electron 0 (Box-Muller consumes the high 24 bits of each):
    on 20000 counters it is uniform (Kolmogorov-Smirnov) and uncorrelated with
    the four proposal uniforms."""
    import numpy as np

    n = 20000
    key = (0x9E3779B9, 12345)
    words = np.array([philox_host((w, 0, 7, 0), key) for w in range(n)], dtype=np.uint64)
    u = ((words[:, 0] & 0xFF) | ((words[:, 1] & 0xFF) << 8) | ((words[:, 2] & 0xFF) << 16)
         | ((words[:, 3] & 0xFF) << 24)).astype(np.float64) * 2.0 ** -32
    s = np.sort(u)
    ks = np.max(np.maximum(np.arange(1, n + 1) / n - s, s - np.arange(n) / n))
    assert ks < 1.63 / np.sqrt(n)  # 1 % level
    for k in range(4):
        hi = (words[:, k] >> 8).astype(np.float64) * 2.0 ** -24
        assert abs(np.corrcoef(u, hi)[0, 1]) < 4.0 / np.sqrt(n)


def _lih_tables():
    import numpy as np

    from qmctorch_amd.wavefunction.plan import SystemTables, atomic_orbital_norm
    from tests import systems

    mol = systems.lih()
    b = mol.basis
    cfg = (np.array([[0, 1], [0, 2]]), np.array([[0, 1], [0, 1]]))
    return SystemTables(b, mol.atom_coords, mol.atomic_number, mol.nup, mol.ndown, cfg, 3, True, True,
                        atomic_orbital_norm(b), np.asarray(b.bas_exp, dtype=np.float64))


def _corrupt(tab, field, index, value):
    arr = getattr(tab, field).copy()
    arr.reshape(-1)[index] = value
    setattr(tab, field, arr)


@pytest.mark.parametrize("field,index,value,reason", [
    ("grp_atom", 0, 2, "grp_atom"),
    ("grp_atom", 1, -1, "grp_atom"),
    ("grp_prim_src", 3, 10 ** 6, "grp_prim_src"),
    ("grp_nao", 0, 10 ** 4, "AO entries"),
    ("grp_ao0", 1, -3, "AO entries"),
    ("grp_ao0", 0, 1, "AO entries"),
    ("gao_mu", 2, 11, "gao_mu"),
    ("gao_src", 5, -1, "gao_src"),
    ("gao_k", 0, 4, "gao_k"),
    ("conf_up", 3, 11, "conf_up"),
    ("conf_dn", 0, -1, "conf_dn"),
    ("ci_up", 1, 2, "ci_up"),
    ("ci_dn", 0, 1, "ci_dn"),
])
def test_plan_create_rejects_out_of_range_tables(field, index, value, reason):
    """qmc_plan_create range-checks every index table on the host and returns
    QMC_ERR_ARG before it allocates or copies anything on the device."""
    if not os.path.isfile(N.LIBPATH):
        pytest.skip("libqmcmi355x.so not built")
    lib = N.load()
    tab = _lih_tables()
    _corrupt(tab, field, index, value)
    h = ctypes.c_void_p()
    desc = tab.desc()
    assert lib.qmc_plan_create(ctypes.byref(desc), ctypes.byref(h)) == N.QMC_ERR_ARG
    assert not h.value
    assert reason in lib.qmc_last_error().decode()


def test_plan_create_rejects_bad_counts_and_missing_tables():
    if not os.path.isfile(N.LIBPATH):
        pytest.skip("libqmcmi355x.so not built")
    lib = N.load()
    tab = _lih_tables()
    h = ctypes.c_void_p()
    d = tab.desc()
    d.nci = 0
    assert lib.qmc_plan_create(ctypes.byref(d), ctypes.byref(h)) == N.QMC_ERR_ARG
    d = tab.desc()
    d.gao_src = None
    assert lib.qmc_plan_create(ctypes.byref(d), ctypes.byref(h)) == N.QMC_ERR_ARG
    assert "missing" in lib.qmc_last_error().decode()
    d = tab.desc()
    d.nentry = d.nentry - 1  # the last group's entries then run past nentry
    assert lib.qmc_plan_create(ctypes.byref(d), ctypes.byref(h)) == N.QMC_ERR_ARG
    assert not h.value
