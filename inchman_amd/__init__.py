"""MI355X-native GPGMP: the Inchman spatial stochastic reaction-drift-diffusion hot
path in hand-written gfx950 HIP behind a C ABI (include/gpgmp_hip.h)."""
from ._lib import GpgmpError, lib  # noqa: F401
from .gpgmp import (  # noqa: F401
    DiffusionModel,
    Ensemble,
    HomogeneousSolver,
    InhomogeneousSolver,
    Reaction,
    Solver,
    Species,
    STOCHASTIC_HOMOGENEOUS,
    STOCHASTIC_INHOMOGENEOUS,
)
