"""CPU FP64 oracle for the McMurchie-Davidson integral and Fock-build path.

TEST INFRASTRUCTURE ONLY.  This package restates the reference algorithm
(jax_qc, pure Python/NumPy) on the CPU so that the HIP kernels can be
checked against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package never
imports, links or falls back to it.

Contents:
* ``mmd``   -- ctypes wrapper of ``mmd_oracle.c`` (Boys, Hermite E/R,
  primitive and contracted S/T/V/ERI blocks, dense ERI tensor, Schwarz Q).
* ``scf``   -- NumPy restatement of jax_qc/scf (fock, density, energy,
  orthogonalize, guess, diis, damping, rhf, uhf) and jax_qc/grad
  (rhf_gradient FD-of-integrals, numerical_gradient).
* ``os_check`` -- an independent Obara-Saika/HRR primitive ERI, used only to
  pin the MMD restatement.

Pinning status per function (what checks each one in tests/):
* Boys (both modes): mpmath 1F1 golden table, F_n(0)=1/(2n+1), F_1
  recurrence, erf closed form, reference test values (test_boys.py).
* hermite_e / hermite_r: base cases (test_integrals_general_l.py:98-121),
  Gauss-Hermite quadrature of the 1-D overlap, finite-difference derivative
  identity of R.
* primitive/contracted blocks: analytic s forms at rel 1e-14, independent
  Obara-Saika ERIs at 1e-13, H2 STO-3G example numbers, 8-fold symmetry,
  Schwarz bound, S diagonal = 1.
* SCF restatement: the reference's hard-coded PySCF energies
  (examples/07_rhf_scf.py:31-36, examples/08_rhf_tier2.py:33-41).
* rhf_gradient restatement: full-SCF central differences and invariants.
* UHF guess/DIIS path: parity unpinned beyond the energy of the converged
  state (only an import of the reference could pin the iteration path).
"""
