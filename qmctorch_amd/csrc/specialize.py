"""Generate AO rows specialised to one basis layout, and their kernels.

The table-driven AO row of qmc_device.h reads the group layout (atom,
primitive count and offset, AO entries, Cartesian powers) from the plan on
every group; each read is a scalar load whose result the next instruction
waits for.  For the layouts of the benchmark molecules this module writes a
row in which every one of those values, and the contraction coefficients,
are compile-time constants (only exponents, MO rows and positions are still
read, at constant offsets the compiler can issue together), and instantiates the sweep, the
batched sweep, the (FP64) local-energy and the (mixed) HMC kernels with it.  At plan creation
the library matches a plan to a generated layout by the signature computed
here and in qmc_abi.hip over the same fields; other plans use the generic row.

    python -m qmctorch_amd.csrc.specialize OUTDIR     (also run by build.py)
"""
import os
import sys
from typing import Dict, List, Tuple

import numpy as np

QMC_MAXCONF = 8  # unique configurations per spin handled by the kernels (include/qmc_mi355x.h)
FNV_OFFSET = 0xCBF29CE484222325
FNV_PRIME = 0x100000001B3
MASK = (1 << 64) - 1


def _fnv(h: int, values) -> int:
    for v in np.asarray(values, dtype=np.int64).ravel().tolist():
        for b in int(v & 0xFFFFFFFF).to_bytes(4, "little"):
            h ^= b
            h = (h * FNV_PRIME) & MASK
    return h


def layout_signature(tab) -> int:
    """FNV-1a over the layout fields baked into a generated row (same order
    and 32-bit little-endian encoding as layout_signature in qmc_abi.hip)."""
    h = FNV_OFFSET
    h = _fnv(h, [tab.radial, tab.natom, len(tab.grp_atom)])
    for arr in (tab.grp_atom, tab.grp_nprim, tab.grp_prim0, tab.grp_ao0, tab.grp_nao, tab.grp_tight):
        h = _fnv(h, arr)
    h = _fnv(h, [len(tab.gao_mu)])
    h = _fnv(h, tab.gao_mu)
    h = _fnv(h, tab.gao_k)
    # contraction coefficients and r powers are literals of the generated row:
    # their float64 bit patterns, as pairs of 32-bit words
    for arr in (tab.gao_K, tab.gao_n):
        h = _fnv(h, np.ascontiguousarray(arr, dtype=np.float64).view(np.uint32).astype(np.int64))
    return h


def row_blocks(tab) -> int:
    """16-lane row blocks of a wavefront holding the walkers of ``tab``'s
    molecule (one lane per electron, 64 // Ne walkers per wavefront): the rows
    the generated sweeps' matrix tiles compute."""
    ne = int(tab.nelec)
    return (64 // ne * ne + 15) // 16


def nm_bucket_fine(nmo: int, nup: int, ndown: int) -> int:
    need = max(nmo, nup, ndown, 1)
    if need <= 8:
        return need
    for b in (12, 16, 24):
        if need <= b:
            return b
    return -1


def row_source(tab) -> str:
    """Body of AoSpec::row for the layout of ``tab``."""
    lines = ["        const double* ac = prm.atom_coords;"]
    ngroup = len(tab.grp_atom)
    g = 0
    while g < ngroup:
        at = int(tab.grp_atom[g])
        lines.append("        {  // atom %d" % at)
        lines.append("            const double* acp = sink.pace(ac);")
        lines.append("            const double dxd = ex - ldc(acp, %d), dyd = ey - ldc(acp, %d), dzd = ez - ldc(acp, %d);"
                     % (3 * at, 3 * at + 1, 3 * at + 2))
        lines.append("            const AtomGeo<T> G = atom_geo<T, RAD>(dxd, dyd, dzd);")
        while g < ngroup and int(tab.grp_atom[g]) == at:
            npr = int(tab.grp_nprim[g])
            p0 = int(tab.grp_prim0[g])
            i0 = int(tab.grp_ao0[g])
            nao = int(tab.grp_nao[g])
            tight = int(tab.grp_tight[g]) if len(tab.grp_tight) else 0

            def block(T, geo):
                out = ["%s E[%d], A[%d];" % (T, npr, npr),
                       "group_prims_from<%s, RAD, %d>(sink.pace(group_exponents<%s>(prm)), %d, %s.r2, %s.r, E, A);"
                       % (T, npr, T, p0, geo, geo)]
                for i in range(i0, i0 + nao):
                    k = int(tab.gao_k[i])
                    kc = ", ".join(repr(float(v)) for v in tab.gao_K[i, :npr])
                    nc = ", ".join(repr(float(v)) for v in tab.gao_n[i, :npr])
                    out.append("{")
                    out.append("    constexpr double K[%d] = {%s};" % (npr, kc))
                    out.append("    constexpr double N[%d] = {%s};" % (npr, nc))
                    out.append("    ao_entry_with<%s, RAD, DERIV, %d>([&](int p) { return (%s)K[p]; }, "
                               "[&](int p) { return N[p]; }, %d, %d, %d, %d, %d, E, A, %s.dx, %s.dy, %s.dz, "
                               "%s.r2, %s.r, %s.ir, sink);"
                               % (T, npr, T, i, int(tab.gao_mu[i]), k & 15, (k >> 4) & 15, (k >> 8) & 15,
                                  geo, geo, geo, geo, geo, geo))
                    out.append("}")
                return out

            lines.append("            sink.fence();")
            lines.append("            // group %d: %d primitives from %d, entries %d..%d%s"
                         % (g, npr, p0, i0, i0 + nao - 1, ", FP64 in the mixed local energy" if tight else ""))
            if tight:
                lines.append("            if (TIGHT) {")
                lines.append("                const AtomGeo<double> D = atom_geo<double, RAD>(dxd, dyd, dzd);")
                lines += ["                " + x for x in block("double", "D")]
                lines.append("            } else {")
                lines += ["                " + x for x in block("T", "G")]
                lines.append("            }")
            else:
                lines.append("            {")
                lines += ["                " + x for x in block("T", "G")]
                lines.append("            }")
            g += 1
        lines.append("        }")
    return "\n".join(lines)


ROW = '''// Generated by qmctorch_amd/csrc/specialize.py for the basis layout of %(name)s
// (%(ngroup)d exponent groups, %(nentry)d AO entries, register bucket %(nm)d).  Do not edit.
#include "qmc_launch.h"

namespace qmc {
namespace {

struct AoSpec {
    static constexpr int kEntries = %(nentry)d;
    static constexpr int kRowBlocks = %(rowblocks)d;  // 16-lane row blocks holding the catalog molecule's walkers
    template <typename T, int RAD, bool DERIV, bool TIGHT, typename Sink>
    static __device__ __forceinline__ void row(const Tables& t, const QmcParams& prm, double ex, double ey,
                                               double ez, Sink&& sink) {
%(body)s
    }
};

constexpr int kNM = %(nm)d;
constexpr int kRAD = %(rad)d;

}  // namespace
'''

SWEEP = '''
int spec_%(name)s_sweep(const SArgs& a, int prec, hipStream_t s) {
    if (prec != QMC_PREC_MIXED) return -1;
%(launch)s
    return 0;
}

int spec_%(name)s_sweep_batch(const SJob* jobs, int njobs, int nwaves, int prec, hipStream_t s) {
    if (prec != QMC_PREC_MIXED) return -1;
%(launch_batch)s
    return 0;
}

}  // namespace qmc
'''

HMC = '''
constexpr int kMC = %(mc)d;

int spec_%(name)s_hmc(const HArgs& a, int prec, hipStream_t s) {
    if (prec != QMC_PREC_MIXED || a.t.nconf_max > kMC) return -1;
    hipLaunchKernelGGL((hmc_kernel<float, kNM, kRAD, kMC, AoSpec>), grid_for(a.t, a.nw), dim3(kBlock), 0, s, a);
    return 0;
}

}  // namespace qmc
'''

ELOC = '''
int spec_%(name)s_wf(const KArgs& a, int prec, int mode, hipStream_t s) {
    if (prec != QMC_PREC_FP64 || mode != QMC_EVAL_ELOC) return -1;
    hipLaunchKernelGGL((wf_kernel<double, kNM, kRAD, true, AoSpec>), grid_for(a.t, a.nw), dim3(kBlock), 0, s, a);
    return 0;
}

int spec_%(name)s_wf_batch(const WJob* jobs, int njobs, int nwaves, int prec, int mode, hipStream_t s) {
    if (prec != QMC_PREC_FP64 || mode != QMC_EVAL_ELOC) return -1;
    hipLaunchKernelGGL((wf_batch_kernel<double, kNM, kRAD, true, AoSpec>), grid_for_waves(nwaves), dim3(kBlock), 0,
                       s, jobs, njobs, nwaves);
    return 0;
}

}  // namespace qmc
'''

REGISTRY = '''// Generated by qmctorch_amd/csrc/specialize.py: the specialised layouts built into the library.
#include "qmc_launch.h"

namespace qmc {

%(decls)s
// host-side table (the device pass of this unit needs none of it)
#if !defined(__HIP_DEVICE_COMPILE__)
const SpecEntry* find_spec(uint64_t signature, int nm, int rows) {
    static const SpecEntry table[] = {
%(rows)s        {0, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr}};
    for (const SpecEntry* e = table; e->name; ++e)
        if (e->signature == signature && e->nm == nm && rows <= e->max_rows) return e;
    return nullptr;
}
#endif

}  // namespace qmc
'''


# W rows of at most this many bytes are staged in LDS by the single-system sweep
WLDS_MAX_BYTES = 8192


def sweep_launch(nentry: int, nm: int) -> Tuple[str, str]:
    """Launches of the single-system and batched sweeps: both with the W rows
    staged in LDS when they fit (the two must run the same arithmetic, so that
    a walker's chain is the same alone and in a batch), else both without."""
    if nentry * nm * 4 <= WLDS_MAX_BYTES:
        single = ("    hipLaunchKernelGGL((sweep_kernel_wlds<float, kNM, kRAD, AoSpec>), grid_for(a.t, a.nw), dim3(kBlock),\n"
                  "                       (size_t)a.t.nentry * kNM * sizeof(float), s, a);")
        batch = ("    hipLaunchKernelGGL((sweep_batch_kernel_wlds<float, kNM, kRAD, AoSpec>), grid_for_waves(nwaves),\n"
                 "                       dim3(kBlock), (kBlock / 64) * %d * sizeof(float), s, jobs, njobs, nwaves, %d);"
                 % (nentry * nm, nentry * nm))
        return single, batch
    single = ("    hipLaunchKernelGGL((sweep_kernel<float, kNM, kRAD, AoSpec>), grid_for(a.t, a.nw), dim3(kBlock), 0, s, "
              "a);")
    batch = ("    hipLaunchKernelGGL((sweep_batch_kernel<float, kNM, kRAD, AoSpec>), grid_for_waves(nwaves), dim3(kBlock), "
             "0, s,\n                       jobs, njobs, nwaves);")
    return single, batch


def unit_sources(name: str, tab, nm: int, mc: int = 1) -> Dict[str, str]:
    """Three units per layout (sweeps; local energy; HMC) so that they compile
    in parallel.  mc bounds the unique configurations per spin of the HMC
    instance (1, or QMC_MAXCONF)."""
    head = ROW % dict(name=name, ngroup=len(tab.grp_atom), nentry=len(tab.gao_mu), nm=nm, rad=int(tab.radial),
                      rowblocks=row_blocks(tab), body=row_source(tab))
    single, batch = sweep_launch(len(tab.gao_mu), nm)
    sweep = SWEEP % dict(name=name, launch=single, launch_batch=batch)
    return {"qmc_spec_%s_sweep.hip" % name: head + sweep,
            "qmc_spec_%s_eloc.hip" % name: head + ELOC % dict(name=name),
            "qmc_spec_%s_hmc.hip" % name: head + HMC % dict(name=name, mc=mc)}


def registry_source(layouts) -> str:
    decls, rows = "", ""
    for name, tab, nm, _ in layouts:
        decls += ("int spec_%(n)s_sweep(const SArgs&, int, hipStream_t);\n"
                  "int spec_%(n)s_sweep_batch(const SJob*, int, int, int, hipStream_t);\n"
                  "int spec_%(n)s_wf(const KArgs&, int, int, hipStream_t);\n"
                  "int spec_%(n)s_wf_batch(const WJob*, int, int, int, int, hipStream_t);\n"
                  "int spec_%(n)s_hmc(const HArgs&, int, hipStream_t);\n" % dict(n=name))
        rows += ("        {0x%016xull, %d, %d, \"%s\", &spec_%s_sweep, &spec_%s_sweep_batch, &spec_%s_wf, "
                 "&spec_%s_wf_batch, &spec_%s_hmc},\n" % (layout_signature(tab), nm, 16 * row_blocks(tab), name,
                                                           name, name, name, name, name))
    return REGISTRY % dict(decls=decls, rows=rows)


def catalog_layouts() -> List[Tuple[str, object, int, int]]:
    """(name, SystemTables, register bucket, HMC configuration bound) of the benchmark molecules:
    ground states of H2/STO-3G, LiH/6-31G, H2O, N2, C6H6/cc-pVDZ and
    LiH/6-31G CAS(2,2)."""
    from ..serving import catalog
    from ..wavefunction.plan import SystemTables, atomic_orbital_norm
    from ..wavefunction.pooling.orbital_configurations import OrbitalConfigurations

    out = []
    for name, configs in (("h2", "ground_state"), ("lih", "ground_state"), ("lih_cas22", "cas(2,2)"),
                          ("h2o", "ground_state"), ("n2", "ground_state"), ("c6h6", "ground_state")):
        mol = catalog.molecule(name.split("_")[0], mo_coeffs="random")
        cu, cd = OrbitalConfigurations(mol).get_configs(configs)
        cu, cd = np.asarray(cu), np.asarray(cd)
        nmo = int(max(cu.max(initial=-1), cd.max(initial=-1))) + 1
        tab = SystemTables(mol.basis, mol.atom_coords, mol.atomic_number, mol.nup, mol.ndown, (cu, cd), nmo,
                           True, True, atomic_orbital_norm(mol.basis), np.asarray(mol.basis.bas_exp))
        nconf = max(len({tuple(r) for r in cu.tolist()}), len({tuple(r) for r in cd.tolist()}))
        out.append((name, tab, nm_bucket_fine(nmo, mol.nup, mol.ndown), 1 if nconf == 1 else QMC_MAXCONF))
    return out


def write_sources(outdir: str, layouts=None) -> List[str]:
    """Write one unit per layout plus the registry; returns the file paths
    (files whose content is unchanged are left untouched)."""
    os.makedirs(outdir, exist_ok=True)
    layouts = catalog_layouts() if layouts is None else layouts
    files: Dict[str, str] = {}
    for name, tab, nm, mc in layouts:
        for fname, text in unit_sources(name, tab, nm, mc).items():
            files[os.path.join(outdir, fname)] = text
    files[os.path.join(outdir, "qmc_spec_registry.hip")] = registry_source(layouts)
    for path, text in files.items():
        if not os.path.isfile(path) or open(path).read() != text:
            with open(path, "w") as f:
                f.write(text)
    return sorted(files)


if __name__ == "__main__":
    for p in write_sources(sys.argv[1] if len(sys.argv) > 1 else "."):
        print(p)
