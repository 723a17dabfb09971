from .jastrow_factor_electron_nuclei import JastrowFactorElectronNuclei as JastrowFactor
from .kernels.pade_jastrow_kernel import PadeJastrowKernel

__all__ = ["JastrowFactor", "PadeJastrowKernel"]
