from .jastrow_factor_electron_electron import JastrowFactorElectronElectron as JastrowFactor
from .kernels.pade_jastrow_kernel import PadeJastrowKernel

__all__ = ["JastrowFactor", "PadeJastrowKernel"]
