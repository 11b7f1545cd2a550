"""dna_amd — MI355X-native Geant4-DNA chain (transport + prechemistry, IRT chemistry,
SSB/DSB scoring) behind the reference's host API (runAtEnergy, runChemistryWorker,
runValidation).  The compute runs in libdna_amd.so (HIP, gfx950); this package is the
host-side mirror of the reference's TypeScript host."""
from .abi import DnaParams  # noqa: F401
from .geometry import DNATarget, default_params  # noqa: F401
from .lib import Context, DnaError, Surrogate, lib  # noqa: F401
