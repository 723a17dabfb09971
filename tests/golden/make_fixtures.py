"""Convert fixtures held by the reference's own tests into small .npz files.

Run with an interpreter that has h5py (this image: /opt/conda/bin/python3.9):

    /opt/conda/bin/python3.9 tests/golden/make_fixtures.py /root/reference

Only data is read (HDF5 datasets, text cube grids); nothing from the reference
is imported or executed.  Outputs (committed, small):

* ``LiH_adf_dz.npz``  -- molecule cache of tests/hdf5/LiH_adf_dz.hdf5 (STO basis + MOs)
* ``C_adf_dzp.npz``   -- molecule cache of tests/hdf5/C_adf_dzp.hdf5
* ``H2_adf_dzp.npz``  -- molecule cache of tests/hdf5/H2_adf_dzp.hdf5
* ``LiH_golden.npz``  -- tests/hdf5/LiH_adf_dz_QMCTorch.hdf5: the wavefunction
  states dumped by the solver groups and the (pos, local_energy) pairs of the
  single_point groups
* ``C_cube.npz``      -- tests/cube/C_AO_*.cub and C_MO_*.cub grids (21x21 plane)

Also committed, copied byte for byte from the reference's tests/hdf5 (data,
not generated): ``LiH_adf_dz.hdf5``, ``H2_adf_dzp.hdf5``, ``C_adf_dzp.hdf5``,
the HDF5 files the .npz molecule caches above were decoded from; the package's
own HDF5 reader is checked against both.
"""
import glob
import os
import sys

import h5py
import numpy as np

OUT = os.path.dirname(os.path.abspath(__file__))


def _read_group(g, prefix, out):
    for k, v in g.items():
        name = prefix + "/" + k
        if isinstance(v, h5py.Group):
            _read_group(v, name, out)
        else:
            val = v[()]
            if isinstance(val, bytes):
                val = np.array(val.decode())
            elif isinstance(val, np.ndarray) and val.dtype.kind in ("S", "O"):
                val = np.array([x.decode() if isinstance(x, bytes) else str(x) for x in val.ravel()]).reshape(val.shape)
            elif isinstance(val, np.ndarray) and val.dtype == bool:
                val = val.astype(np.int8)
            out[name] = np.asarray(val)


def molecule_cache(ref, name):
    out = {}
    with h5py.File(os.path.join(ref, "tests", "hdf5", name + ".hdf5"), "r") as f:
        _read_group(f["molecule"], "molecule", out)
    out = {k: v for k, v in out.items() if not k.startswith("molecule/calculator/")}
    np.savez(os.path.join(OUT, name + ".npz"), **out)


def lih_golden(ref):
    out = {}
    with h5py.File(os.path.join(ref, "tests", "hdf5", "LiH_adf_dz_QMCTorch.hdf5"), "r") as f:
        for grp in f.keys():
            if grp.startswith("single_point"):
                for k in ("pos", "local_energy", "energy", "variance", "error"):
                    out[grp + "/" + k] = np.asarray(f[grp][k][()])
            if grp.startswith("SolverSlaterJastrow"):
                wf = f[grp]["wf"]
                for k in (
                    "ao/bas_exp", "ao/norm_cst", "ao/bas_n", "ao/atom_coords",
                    "mo_scf/weight", "mo/weight", "fc/weight", "jastrow/weight",
                    "configs_0", "configs_1",
                ):
                    if k in wf:
                        out[grp + "/wf/" + k] = np.asarray(wf[k][()])
                s = f[grp]["sampler"]
                out[grp + "/sampler/step_size"] = np.asarray(s["step_size"][()])
                out[grp + "/sampler/nstep"] = np.asarray(s["nstep"][()])
            if grp.startswith("wf_opt"):
                out[grp + "/energy"] = np.asarray(f[grp]["energy"][()])
                out[grp + "/local_energy"] = np.asarray(f[grp]["local_energy"][()])
                for m in ("best", "last"):
                    for k, v in f[grp]["models"][m].items():
                        out[grp + "/models/" + m + "/" + k] = np.asarray(v[()])
    np.savez_compressed(os.path.join(OUT, "LiH_golden.npz"), **out)


def read_cube(path):
    """Gaussian cube: header, atoms, then values (x slowest)."""
    with open(path) as f:
        lines = f.readlines()
    natom, *orig = lines[2].split()
    natom = int(natom)
    axes = []
    for i in range(3):
        t = lines[3 + i].split()
        axes.append((int(t[0]), [float(x) for x in t[1:4]]))
    vals = np.array([float(x) for l in lines[6 + abs(natom):] for x in l.split()])
    n = [a[0] for a in axes]
    return np.array([float(x) for x in orig[:3]]), axes, vals.reshape(n)


def cubes(ref):
    out = {}
    for kind in ("AO", "MO"):
        files = glob.glob(os.path.join(ref, "tests", "cube", "C_%s_*.cub" % kind))
        for fn in sorted(files):
            idx = int(os.path.basename(fn).split("%")[-1].replace("AO", "").replace(".cub", "").lstrip("AO"))
            origin, axes, vals = read_cube(fn)
            out["%s/%d" % (kind, idx)] = vals
            out["%s/%d/origin" % (kind, idx)] = origin
            out["%s/%d/axes" % (kind, idx)] = np.array([a[1] for a in axes])
            out["%s/%d/n" % (kind, idx)] = np.array([a[0] for a in axes])
    np.savez_compressed(os.path.join(OUT, "C_cube.npz"), **out)


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for n in ("LiH_adf_dz", "C_adf_dzp", "H2_adf_dzp"):
        molecule_cache(ref, n)
    lih_golden(ref)
    cubes(ref)
    print("fixtures written to", OUT)
