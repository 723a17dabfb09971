"""Object <-> HDF5 dumps (the reference's utils/hdf5_utils.py:27-483).

Same public functions and the same file layout: an object becomes a group
whose members are its attributes (``__dict__``, mapping keys, registered extra
attributes and ``state_dict`` entries), recursively; leaves become datasets
(tensors and parameters as arrays, unicode arrays as byte strings, lists as
one array when they stack, else as ``<name>_<i>`` members); names starting
with "_" and ``None`` / ``torch.device`` values are skipped.  Loading walks a
group back into a ``SimpleNamespace`` tree, decoding byte strings.

h5py is used when it is importable; otherwise this package's pure-Python
HDF5 reader/writer (``h5lite``), which writes files h5py reads and reads the
files h5py writes with its default format.
"""
from types import SimpleNamespace

import numpy as np
import torch

from . import h5lite
from .logger import log

try:  # pragma: no cover - depends on the environment
    import h5py as _h5
except ImportError:  # pragma: no cover
    _h5 = h5lite


def open_file(fname: str, mode: str = "r"):
    """An h5py.File, or an h5lite.File where h5py is absent."""
    return _h5.File(fname, mode)


# ------------------------------------------------------------------ loading
def load_from_hdf5(obj, fname: str, obj_name: str) -> None:
    """Fill ``obj`` with the content of group ``obj_name`` (hdf5_utils.py:27-40)."""
    with open_file(fname, "r") as h5:
        load_object(h5[obj_name], obj, obj_name)


def load_object(grp, parent_obj, grp_name: str) -> None:
    for name, child in grp.items():
        (load_group if isgroup(child) else load_data)(child, parent_obj, name)


def load_group(grp, parent_obj, grp_name: str) -> None:
    try:
        if not hasattr(parent_obj, grp_name):
            setattr(parent_obj, grp_name, SimpleNamespace())
        load_object(grp, getattr(parent_obj, grp_name), grp_name)
    except Exception as err:  # the reference logs and keeps going
        log.critical("Issue loading {grp}: {err}", grp=grp_name, err=err)


def load_data(grp, parent_obj, grp_name: str) -> None:
    try:
        setattr(parent_obj, grp_name, cast_loaded_data(grp[()]))
    except Exception as err:
        log.critical("Issue loading {grp}: {err}", grp=grp_name, err=err)


def cast_loaded_data(data):
    """Scalar byte strings come back as str (hdf5_utils.py:95-104)."""
    return bytes2str(data) if type(data) is bytes else data


def bytes2str(bstr):
    if isinstance(bstr, bytes):
        return bstr.decode("utf-8")
    if isinstance(bstr, str):
        return bstr
    raise TypeError("%r should be a bytes or str but got %s instead" % (bstr, type(bstr)))


def isgroup(grp) -> bool:
    return isinstance(grp, (_h5.Group, _h5.File)) if _h5 is not h5lite else isinstance(grp, h5lite.Group)


# ------------------------------------------------------------------ dumping
def dump_to_hdf5(obj, fname: str, root_name: str = None) -> str:
    """Append ``obj`` as a new top-level group; a taken name becomes
    ``<name>_<k>`` with k = 1 + the number of members starting with it
    (hdf5_utils.py:132-160).  Returns the group name used."""
    with open_file(fname, "a") as h5:
        if root_name is None:
            root_name = obj.__class__.__name__
        if root_name in h5:
            taken = sum(1 for n in h5 if n.startswith(root_name))
            new = "%s_%d" % (root_name, taken + 1)
            log.info(" Warning : dump to hdf5: {0} exists in {1}, stored as {2}", root_name, fname, new)
            root_name = new
        insert_object(obj, h5, root_name)
    return root_name


def insert_object(obj, parent_grp, obj_name: str) -> None:
    (insert_group if haschildren(obj) else insert_data)(obj, parent_grp, obj_name)


def insert_group(obj, parent_grp, obj_name: str) -> None:
    if obj_name.startswith("_"):  # torch internals and private state
        return
    if obj_name in parent_grp:
        log.critical(" Warning : {0} already exists, keeping the stored version", obj_name)
        return
    grp = parent_grp.create_group(obj_name)
    for name in get_children_names(obj):
        try:
            insert_object(get_child_object(obj, name), grp, name)
        except Exception as err:
            log.critical("Issue inserting {0} ({1}): {2}", name, type(obj).__name__, err)


def _leaf_array(obj):
    """The array a leaf is stored as, or None for values that are not stored."""
    if obj is None or isinstance(obj, torch.device):
        return None
    if isinstance(obj, torch.Tensor):  # tensors and parameters
        obj = obj.detach().cpu().numpy()
    if isinstance(obj, tuple):
        obj = [o.detach().cpu().numpy() if isinstance(o, torch.Tensor) else o for o in obj]
    if isinstance(obj, list) and obj and all(isinstance(o, torch.Tensor) for o in obj):
        obj = [o.detach().cpu().numpy() for o in obj]
    a = np.array(obj) if isinstance(obj, list) else np.asarray(obj)
    if a.dtype.kind == "U":
        a = a.astype("S")
    return a


def insert_data(obj, parent_grp, obj_name: str) -> None:
    if obj_name.startswith("_"):
        return
    try:
        a = _leaf_array(obj)
        if a is None:
            return
        if a.dtype.kind == "O":
            raise TypeError("ragged or object data")
        parent_grp.create_dataset(obj_name, data=a)
    except Exception:
        if isinstance(obj, (list, tuple)):  # elements one by one (hdf5_utils.py:293-312)
            for i, el in enumerate(obj):
                try:
                    insert_object(el, parent_grp, "%s_%d" % (obj_name, i))
                except Exception as err:
                    log.critical("Issue inserting {0}_{1}: {2}", obj_name, i, err)
        else:
            log.critical("Issue inserting data {0} of type {1}", obj_name, type(obj).__name__)


def haschildren(obj) -> bool:
    if isinstance(obj, torch.Tensor):
        return False
    return hasattr(obj, "__dict__") or hasattr(obj, "keys")


def children(obj):
    return obj.items() if isinstance(obj, dict) or not hasattr(obj, "__dict__") else obj.__dict__.items()


def get_children_names(obj):
    # mappings by their keys (an OrderedDict also has an empty __dict__)
    names = set(obj.keys()) if isinstance(obj, dict) or not hasattr(obj, "__dict__") else set(obj.__dict__)
    names.update(getattr(obj, "__extra_attr__", []))
    if hasattr(obj, "state_dict"):
        names.update(obj.state_dict().keys())
    return list(names)


def get_child_object(obj, child_name: str):
    if isinstance(obj, dict):
        return obj.get(child_name)
    for getter in ("__getattr__", "__getattribute__", "__getitem__"):
        if hasattr(obj, getter):
            try:
                return getattr(obj, getter)(child_name)
            except (AttributeError, KeyError, TypeError):
                continue
    return None


def add_group_attr(fname: str, grp_name: str, attr: dict) -> None:
    """Set attributes of group ``grp_name`` (hdf5_utils.py:455-468)."""
    if grp_name is None:
        return
    with open_file(fname, "a") as h5:
        for k, v in attr.items():
            h5[grp_name].attrs[k] = v


def register_extra_attributes(obj, attr_names) -> None:
    """Names dumped besides ``__dict__`` (hdf5_utils.py:471-483)."""
    if not hasattr(obj, "__extra_attr__"):
        obj.__extra_attr__ = []
    obj.__extra_attr__ += list(attr_names)
