"""ctypes binding of libqmcmi355x.so (C ABI declared in include/qmc_mi355x.h).

The library is loaded lazily on first use and the import fails loudly when it
is missing: there is no CPU or PyTorch fallback for the kernels.
"""
import ctypes
import os
import threading

import torch

LIBNAME = "libqmcmi355x.so"
_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, LIBNAME)

QMC_OK = 0
QMC_ERR_ARG = 1
QMC_ERR_HIP = 2
QMC_ERR_UNSUPPORTED = 3
QMC_MAXP = 12
QMC_MAXCONF = 8
QMC_RADIAL_GTO_PURE = 0
QMC_RADIAL_STO = 1
QMC_PREC_FP64 = 0
QMC_PREC_MIXED = 1
QMC_EVAL_PSI = 0
QMC_EVAL_ELOC = 1

c_i32p = ctypes.POINTER(ctypes.c_int32)
c_f64p = ctypes.POINTER(ctypes.c_double)


class QmcSystemDesc(ctypes.Structure):
    _fields_ = [
        ("nelec", ctypes.c_int32), ("nup", ctypes.c_int32), ("ndown", ctypes.c_int32), ("natom", ctypes.c_int32),
        ("nao", ctypes.c_int32),
        ("nprim_ref", ctypes.c_int32),
        ("radial_type", ctypes.c_int32),
        ("ngroup", ctypes.c_int32),
        ("grp_atom", c_i32p), ("grp_nprim", c_i32p), ("grp_prim0", c_i32p), ("grp_prim_src", c_i32p),
        ("grp_ao0", c_i32p), ("grp_nao", c_i32p), ("grp_tight", c_i32p),
        ("nentry", ctypes.c_int32),
        ("gao_mu", c_i32p), ("gao_k", c_i32p), ("gao_K", c_f64p), ("gao_n", c_f64p), ("gao_src", c_i32p),
        ("atom_z", c_f64p), ("jastrow_atoms", c_f64p),
        ("nmo", ctypes.c_int32), ("nci", ctypes.c_int32),
        ("nconf_up", ctypes.c_int32), ("nconf_dn", ctypes.c_int32),
        ("conf_up", c_i32p), ("conf_dn", c_i32p), ("ci_up", c_i32p), ("ci_dn", c_i32p),
        ("use_jastrow_ee", ctypes.c_int32), ("use_jastrow_en", ctypes.c_int32),
        ("vnn", ctypes.c_double),
    ]


class QmcParams(ctypes.Structure):
    _fields_ = [
        ("bas_exp", ctypes.c_void_p), ("atom_coords", ctypes.c_void_p), ("mo_w", ctypes.c_void_p),
        ("ci", ctypes.c_void_p), ("w_ee", ctypes.c_void_p), ("w_en", ctypes.c_void_p),
        ("grp_exp", ctypes.c_void_p), ("grp_expf", ctypes.c_void_p), ("entry_w", ctypes.c_void_p),
    ]


class QmcWfOut(ctypes.Structure):
    _fields_ = [
        ("psi", ctypes.c_void_p), ("logabs", ctypes.c_void_p), ("sign", ctypes.c_void_p),
        ("kinetic", ctypes.c_void_p), ("eloc", ctypes.c_void_p), ("epot", ctypes.c_void_p),
        ("v_ee", ctypes.c_void_p), ("v_en", ctypes.c_void_p),
    ]


class QmcWalkers(ctypes.Structure):
    _fields_ = [
        ("pos", ctypes.c_void_p), ("logabs", ctypes.c_void_p), ("sign", ctypes.c_void_p),
        ("naccept", ctypes.c_void_p), ("walker_id", ctypes.c_void_p),
    ]


class QmcSweepDesc(ctypes.Structure):
    _fields_ = [
        ("step0", ctypes.c_int64), ("nstep", ctypes.c_int32), ("init_logpsi", ctypes.c_int32),
        ("sigma2", ctypes.c_double), ("seed", ctypes.c_uint64), ("keep_from", ctypes.c_int64),
        ("keep_every", ctypes.c_int32), ("uniform_move", ctypes.c_int32), ("step_size", ctypes.c_double),
        ("keep", ctypes.c_void_p), ("one_electron", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


class QmcHmcDesc(ctypes.Structure):
    _fields_ = [
        ("step0", ctypes.c_int64), ("nstep", ctypes.c_int32), ("traj_length", ctypes.c_int32),
        ("step_size", ctypes.c_double), ("seed", ctypes.c_uint64), ("keep_from", ctypes.c_int64),
        ("keep_every", ctypes.c_int32), ("reserved", ctypes.c_int32), ("keep", ctypes.c_void_p),
    ]


class QmcSweepJob(ctypes.Structure):
    _fields_ = [
        ("plan", ctypes.c_void_p), ("params", QmcParams), ("walkers", QmcWalkers), ("nw", ctypes.c_int64),
        ("sweep", QmcSweepDesc),
    ]


class QmcEvalJob(ctypes.Structure):
    _fields_ = [
        ("plan", ctypes.c_void_p), ("params", QmcParams), ("pos", ctypes.c_void_p), ("nw", ctypes.c_int64),
        ("out", QmcWfOut),
    ]


class QmcParamViews(ctypes.Structure):
    _fields_ = [
        ("mo_w64", ctypes.c_void_p), ("mo_w32", ctypes.c_void_p), ("entry_w64", ctypes.c_void_p),
        ("entry_w32", ctypes.c_void_p), ("grp_exp", ctypes.c_void_p), ("grp_expf", ctypes.c_void_p),
    ]


class QmcPlanSizes(ctypes.Structure):
    _fields_ = [("nexp_grp", ctypes.c_int32), ("nentry", ctypes.c_int32), ("bucket", ctypes.c_int32),
                ("nao", ctypes.c_int32), ("nmo", ctypes.c_int32)]


class QmcJastrowOut(ctypes.Structure):
    _fields_ = [("value", ctypes.c_void_p), ("grad", ctypes.c_void_p), ("lap", ctypes.c_void_p),
                ("v_ee", ctypes.c_void_p), ("v_en", ctypes.c_void_p)]


class QmcAoOut(ctypes.Structure):
    _fields_ = [("ao", ctypes.c_void_p), ("grad", ctypes.c_void_p), ("lap", ctypes.c_void_p),
                ("hess_diag", ctypes.c_void_p), ("hess_mixed", ctypes.c_void_p)]


class QmcGradOut(ctypes.Structure):
    _fields_ = [("grad_w", ctypes.c_void_p), ("grad_ci", ctypes.c_void_p), ("grad_jas", ctypes.c_void_p),
                ("grad_exp", ctypes.c_void_p)]


QMC_INIT_CENTER, QMC_INIT_UNIFORM, QMC_INIT_NORMAL, QMC_INIT_ATOMIC = 0, 1, 2, 3
QMC_INIT_MAXATOM = 32
QMC_NOALLOC_SCRATCH_BYTES = 64 << 20


class QmcWalkerInit(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int32), ("natom", ctypes.c_int32), ("seed", ctypes.c_uint64), ("walker0", ctypes.c_int64),
        ("lo", ctypes.c_double), ("hi", ctypes.c_double), ("mean", ctypes.c_double * 3),
        ("chol", ctypes.c_double * 9), ("atom_xyz", ctypes.c_double * (3 * QMC_INIT_MAXATOM)),
        ("atom_width", ctypes.c_double * (3 * QMC_INIT_MAXATOM)), ("atom_nelec", ctypes.c_int32 * QMC_INIT_MAXATOM),
    ]


EXPORTS = {
    "qmc_plan_create": ([ctypes.POINTER(QmcSystemDesc), ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "qmc_plan_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "qmc_plan_set_specialized": ([ctypes.c_void_p, ctypes.c_int32], ctypes.c_int),
    "qmc_plan_specialization": ([ctypes.c_void_p], ctypes.c_char_p),
    "qmc_plan_bucket": ([ctypes.c_void_p], ctypes.c_int),
    "qmc_wf_eval": ([ctypes.c_void_p, ctypes.POINTER(QmcParams), ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                     ctypes.c_int64, ctypes.POINTER(QmcWfOut), ctypes.c_void_p], ctypes.c_int),
    "qmc_ao_eval": ([ctypes.c_void_p, ctypes.POINTER(QmcParams), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "qmc_metropolis_sweep": ([ctypes.c_void_p, ctypes.POINTER(QmcParams), ctypes.c_int32, ctypes.POINTER(QmcWalkers),
                              ctypes.c_int64, ctypes.POINTER(QmcSweepDesc), ctypes.c_void_p], ctypes.c_int),
    "qmc_hmc_sweep": ([ctypes.c_void_p, ctypes.POINTER(QmcParams), ctypes.c_int32, ctypes.POINTER(QmcWalkers),
                       ctypes.c_int64, ctypes.POINTER(QmcHmcDesc), ctypes.c_void_p], ctypes.c_int),
    "qmc_wf_drift": ([ctypes.c_void_p, ctypes.POINTER(QmcParams), ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64,
                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "qmc_logpsi_grad_rows": ([ctypes.c_void_p, ctypes.POINTER(QmcParams), ctypes.c_void_p, ctypes.c_int64,
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_void_p],
                             ctypes.c_int),
    "qmc_eloc_grad_rows": ([ctypes.c_void_p, ctypes.POINTER(QmcParams), ctypes.c_void_p, ctypes.c_int64]
                           + [ctypes.c_void_p] * 8 + [ctypes.c_void_p], ctypes.c_int),
    "qmc_batch_create": ([ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "qmc_batch_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "qmc_metropolis_sweep_batch": ([ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(QmcSweepJob), ctypes.c_int32,
                                    ctypes.c_void_p], ctypes.c_int),
    "qmc_wf_eval_batch": ([ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(QmcEvalJob),
                           ctypes.c_int32, ctypes.c_void_p], ctypes.c_int),
    "qmc_philox4x32_10": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p], ctypes.c_int),
    "qmc_plan_sizes": ([ctypes.c_void_p, ctypes.POINTER(QmcPlanSizes)], ctypes.c_int),
    "qmc_params_gather": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                           ctypes.POINTER(QmcParamViews), ctypes.c_void_p], ctypes.c_int),
    "qmc_jastrow_eval": ([ctypes.c_void_p, ctypes.POINTER(QmcParams), ctypes.c_int32, ctypes.c_void_p,
                          ctypes.c_int64, ctypes.POINTER(QmcJastrowOut), ctypes.c_void_p], ctypes.c_int),
    "qmc_slater_pool": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                         ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                         ctypes.c_void_p], ctypes.c_int),
    "qmc_ao_eval_full": ([ctypes.c_void_p, ctypes.POINTER(QmcParams), ctypes.c_void_p, ctypes.c_int64,
                          ctypes.c_int32, ctypes.POINTER(QmcAoOut), ctypes.c_void_p], ctypes.c_int),
    "qmc_adjoint_workspace_size": ([ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32], ctypes.c_size_t),
    "qmc_logpsi_grad": ([ctypes.c_void_p, ctypes.POINTER(QmcParams), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                         ctypes.POINTER(QmcGradOut), ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p], ctypes.c_int),
    "qmc_eloc_grad": ([ctypes.c_void_p, ctypes.POINTER(QmcParams), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                       ctypes.c_void_p, ctypes.POINTER(QmcGradOut), ctypes.c_void_p, ctypes.c_size_t,
                       ctypes.c_void_p], ctypes.c_int),
    "qmc_logpsi_grad_noalloc": ([ctypes.c_void_p, ctypes.POINTER(QmcParams), ctypes.c_void_p, ctypes.c_int64,
                                 ctypes.c_void_p, ctypes.POINTER(QmcGradOut), ctypes.c_void_p], ctypes.c_int),
    "qmc_walkers_init": ([ctypes.POINTER(QmcWalkerInit), ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p,
                          ctypes.c_void_p], ctypes.c_int),
    "qmc_last_error": ([], ctypes.c_char_p),
    "qmc_version": ([], ctypes.c_char_p),
}

_lib = None
_lock = threading.Lock()


class NativeLibraryMissing(RuntimeError):
    pass


def load(path: str = LIBPATH):
    """Load the library (once).  Raises if it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.isfile(path):
            raise NativeLibraryMissing(
                "%s not found: build it with `python -m qmctorch_amd.csrc.build` (there is no CPU fallback)" % path
            )
        lib = ctypes.CDLL(path)
        for name, (args, res) in EXPORTS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
        return lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().qmc_last_error().decode()
        raise RuntimeError("libqmcmi355x error %d: %s" % (rc, msg))


def ptr(t) -> int:
    """Device pointer of a tensor (or 0 for None)."""
    if t is None:
        return 0
    return t.data_ptr()


def current_stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def require_device(t: torch.Tensor, what: str = "input") -> None:
    if not t.is_cuda:
        raise RuntimeError(
            "%s is on %s: the MI355X kernels need HIP device tensors (use cuda=True / .cuda()); "
            "there is no CPU path" % (what, t.device)
        )
