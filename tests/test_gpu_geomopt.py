"""BFGS on the GPU path: H2 / STO-3G converges to the known bond length
(energy decreases monotonically along accepted steps) and the final
gradient is below grad_tol."""

import numpy as np
import pytest

from jax_qc import molecules as M

pytestmark = pytest.mark.gpu


def test_h2_bfgs_converges():
    from jax_qc.geomopt import optimize_geometry

    r = optimize_geometry(M.h2(), "sto-3g", "rhf", max_steps=20, verbose=0, tau=0.0)
    assert r.converged
    E = [e for _, e in r.trajectory]
    assert all(b <= a + 1e-10 for a, b in zip(E, E[1:]))
    d = np.linalg.norm(r.molecule.coords[0] - r.molecule.coords[1])
    assert d == pytest.approx(1.3459, abs=2e-3)  # RHF/STO-3G H2 equilibrium (Bohr)
    assert np.abs(r.gradient).max() < 1e-4


def test_uhf_bfgs_energy_decreases():
    from jax_qc.geomopt import optimize_geometry

    r = optimize_geometry(M.oh_radical(), "sto-3g", "uhf", max_steps=4, verbose=0, tau=0.0)
    E = [e for _, e in r.trajectory]
    assert E[-1] < E[0]
