from .pade_jastrow_kernel import PadeJastrowKernel

__all__ = ["PadeJastrowKernel"]
