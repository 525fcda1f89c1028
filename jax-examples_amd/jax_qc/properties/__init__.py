"""Post-SCF properties (jax_qc/properties/*): dipole moment with exact
dipole integrals, Mulliken populations and charges, orbital analysis."""

from jax_qc.properties.dipole import (
    AU_TO_DEBYE,
    dipole_moment,
    dipole_moment_debye,
    electronic_dipole,
    nuclear_dipole,
)
from jax_qc.properties.mulliken import mulliken_analysis, mulliken_charges, mulliken_populations
from jax_qc.properties.orbital_analysis import OrbitalInfo, orbital_analysis

__all__ = ["AU_TO_DEBYE", "dipole_moment", "dipole_moment_debye", "electronic_dipole",
           "nuclear_dipole", "mulliken_analysis", "mulliken_charges", "mulliken_populations",
           "OrbitalInfo", "orbital_analysis"]
