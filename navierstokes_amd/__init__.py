"""MI355X (gfx950) time-advance hot path of the mathe1s/NavierStokes solver.

libnsamd.so (HIP kernels + C ABI, include/ns_api.h) is the compute path;
this package holds the host-side set-up and the Python mirror of the
reference's routine interface.
"""
from .config import Case, acoustic_pulse_case, pulse_vortex_case, random_case, reference_case  # noqa: F401

__all__ = ["Case", "acoustic_pulse_case", "pulse_vortex_case", "random_case", "reference_case", "Solver"]


def __getattr__(name):
    if name == "Solver":
        from .solver import Solver
        return Solver
    raise AttributeError(name)
