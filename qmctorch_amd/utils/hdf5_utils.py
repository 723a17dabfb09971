"""Observable dumps (utils/hdf5_utils.py:130-160 of the reference).

Writes the reference's HDF5 layout when h5py is importable; this image has no
h5py for the torch interpreter, in which case dumps are skipped with one
informational log line.
"""
from types import SimpleNamespace

import numpy as np
import torch

from .logger import log

try:  # pragma: no cover - depends on the environment
    import h5py
except ImportError:  # pragma: no cover
    h5py = None

_warned = False


def _skip():
    global _warned
    if not _warned:
        log.info("  h5py not available: HDF5 output skipped")
        _warned = True


def _to_np(v):
    if isinstance(v, torch.Tensor):
        return v.detach().cpu().numpy()
    if isinstance(v, (list, tuple)):
        try:
            return np.asarray([_to_np(x) for x in v])
        except Exception:
            return None
    if isinstance(v, (int, float, np.ndarray, np.number, bool)):
        return np.asarray(v)
    if isinstance(v, str):
        return np.bytes_(v)
    return None


def _write(grp, name, v):
    if isinstance(v, (SimpleNamespace, dict)):
        sub = grp.require_group(name)
        items = vars(v).items() if isinstance(v, SimpleNamespace) else v.items()
        for k, x in items:
            _write(sub, str(k), x)
        return
    a = _to_np(v)
    if a is not None and name not in grp:
        grp.create_dataset(name, data=a)


def dump_to_hdf5(obj, fname: str, root_name: str = None) -> str:
    if h5py is None:
        _skip()
        return root_name
    with h5py.File(fname, "a") as f:
        if root_name is None:
            root_name = obj.__class__.__name__
        name, i = root_name, 1
        while name in f:
            i += 1
            name = "%s_%d" % (root_name, i)
        grp = f.create_group(name)
        items = vars(obj).items() if hasattr(obj, "__dict__") else []
        for k, v in items:
            if k.startswith("_"):
                continue
            try:
                _write(grp, k, v)
            except Exception:
                pass
    return name


def add_group_attr(fname: str, grp_name: str, attr: dict) -> None:
    if h5py is None or grp_name is None:
        return
    with h5py.File(fname, "a") as f:
        for k, v in attr.items():
            f[grp_name].attrs[k] = v
