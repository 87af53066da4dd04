"""MI355X (gfx950) single-device time-step hot path of DistributedNavierStokes.

Finite-volume API (``fv``) and pseudo-spectral API (``spectral``) over the C
ABI of ``libdnsrocm.so`` (include/dns_rocm.h).  Mutating functions of the
reference (``step!``) are spelled with a trailing underscore (``step_``).
"""
from .problem import DNSError, FVProblem, SpectralProblem, tanhstretch  # noqa: F401
from .fv import (  # noqa: F401
    ROCmBackend, setup, scalarfield, vectorfield, velocityfield_, exchange_halo_, project_, momentum_, step_,
    cfl_timestep, advance_cfl_, solve_, maxdiv, kinetic_energy, channelstats, channelprofiles, spectrumstats, energyspectrum, rk3,
)
from .fv import VectorField  # noqa: F401
from .spectral import (  # noqa: F401
    spectral_setup, specvelocity, spectral_velocityfield_, taylorgreen_, spectral_randomfield_, spec_to_phys_,
    phys_to_spec_, spectral_project_, spectral_rhs_, spectral_step_, spectral_solve_, shell_energies, shellforcing,
    spectral_energy, spectral_dissipation, spectral_maxdiv, spectral_stats, spectral_spectrum, wray3,
)
from .spectral_io import (  # noqa: F401
    spectral_save, spectral_load_, checkpointlist, spectral_latest, checkpointer, snapshotsaver, statswriter,
)
