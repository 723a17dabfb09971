from .algebra_utils import bdet2, bproj, btrace
from .constants import ANGS2BOHR, BOHR2ANGS
from .hdf5_utils import add_group_attr, bytes2str, dump_to_hdf5, load_from_hdf5, register_extra_attributes
from .logger import log
from .provenance import get_git_tag
from .torch_utils import (DataLoader, DataSet, diagonal_hessian, fast_power, gradients, hessian,
                          set_torch_double_precision, set_torch_single_precision)

__all__ = [
    "set_torch_double_precision",
    "set_torch_single_precision",
    "DataSet",
    "DataLoader",
    "add_group_attr",
    "dump_to_hdf5",
    "load_from_hdf5",
    "bytes2str",
    "register_extra_attributes",
    "fast_power",
    "btrace",
    "bdet2",
    "bproj",
    "diagonal_hessian",
    "gradients",
    "hessian",
    "get_git_tag",
    "ANGS2BOHR",
    "BOHR2ANGS",
    "log",
]
