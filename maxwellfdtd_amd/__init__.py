"""maxwellfdtd_amd -- MI355X-native per-time-step engine for MaxwellFDTD models.

The constructors mirror the reference package's exported names
(src/MaxwellFDTD.jl:12-14); `solve` runs the model on the HIP engine.
"""
from .constants import C0, EPS0, HBAR, ME, MU0, QE
from .grids import Grid, Grid1D, Grid2D, Grid3D
from .materials import (DebyeSusceptibility, DrudeSusceptibility, IonizationRate,
                        LorentzSusceptibility, Material, Plasma, PowerLawRate, susceptibility)
from .monitors import FieldMonitor, SpectralMonitor
from .models import MI355X, Model, Solution, solve
from .pml import CPML
from .sources import HardSource, SoftSource, TFSFSource

solve_ = solve

__all__ = [
    "C0", "EPS0", "MU0", "QE", "ME", "HBAR", "Grid", "Grid1D", "Grid2D", "Grid3D",
    "Material", "DebyeSusceptibility", "DrudeSusceptibility", "LorentzSusceptibility",
    "Plasma", "PowerLawRate", "IonizationRate", "susceptibility", "FieldMonitor",
    "SpectralMonitor", "MI355X", "Model", "Solution", "solve", "CPML", "SoftSource",
    "HardSource", "TFSFSource",
]
