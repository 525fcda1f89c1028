from jax_qc.basis.build import (  # noqa: F401
    build_basis_set,
    build_shell_indices,
    bse_dict_to_shells,
    even_tempered_dict,
    parse_electron_shell,
    shell_arrays,
)
from jax_qc.basis.normalize import contraction_norm, normalize_shell, primitive_norm  # noqa: F401
