from jax_qc.core.types import (  # noqa: F401
    BasisSet,
    CalcConfig,
    IntegralSet,
    Molecule,
    SCFResult,
    SCFState,
    Shell,
    make_molecule,
)
