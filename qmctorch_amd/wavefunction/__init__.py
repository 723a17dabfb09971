from .wf_base import WaveFunction
from .slater_jastrow import SlaterJastrow

__all__ = ["WaveFunction", "SlaterJastrow"]
