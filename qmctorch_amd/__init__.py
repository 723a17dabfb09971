"""MI355X-native implementation of the QMCTorch VMC hot path.

The package mirrors the public surface of ``qmctorch`` (Molecule, SlaterJastrow,
the Pade Jastrow factors, Metropolis, Solver) so that scripts written for the
reference run unchanged; the numerical work is done by hand-written HIP kernels
for gfx950 reached through the C-ABI library ``libqmcmi355x.so``
(see ``include/qmc_mi355x.h``).

Reference: qmctorch/__init__.py:1-21 (logger + version banner).
"""
from .__version__ import __version__
from .utils.logger import log

__all__ = ["__version__", "log"]
