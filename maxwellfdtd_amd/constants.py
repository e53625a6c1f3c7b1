"""CODATA 2018 constants, the values PhysicalConstants.CODATA2018 gives the reference
(src/MaxwellFDTD.jl:17-23)."""

C0 = 299792458.0            # speed of light in vacuum (exact)
EPS0 = 8.8541878128e-12     # vacuum electric permittivity
MU0 = 1.25663706212e-6      # vacuum magnetic permeability
QE = 1.602176634e-19        # elementary charge (exact)
ME = 9.1093837015e-31       # electron mass
HBAR = 1.054571817e-34      # reduced Planck constant
