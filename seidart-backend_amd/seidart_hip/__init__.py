"""seidart_hip -- gfx950 HIP CPML-FDTD solvers with the inputs and outputs of
SeidarT-Backend (src/fortran/cpmlfdtd.f08).

    Solver        in-memory handle over libseidart_hip.so (seidart_hip.solver)
    cpmlfdtd      file-based drop-in of seismic2/seismic3/electromag2/electromag3
    models        seeded synthetic models, CPML profiles, Ricker sources
"""
from ._lib import SdtError, UnstableError  # noqa: F401  (fails loudly without the library)
from .solver import Solver, run_config  # noqa: F401

__all__ = ["Solver", "run_config", "SdtError", "UnstableError"]
