"""The two precision modes of the energy + gradient path, as keyword sets
for ``jax_qc.geomopt.optimizer.energy_and_gradient`` (and
``optimize_bfgs``).

* ``STRICT`` -- the plain FP64 drop-in: Schwarz screening at 1e-12 and
  nothing else; FP64 ERIs, FP64 supermatrices, full Fock builds, FP64
  derivative integrals at Schwarz 1e-12.
* ``MIXED`` -- ERI primitive sums in FP32 for quartets with Schwarz bound
  below 1e-4; supermatrix tiles by max |element| m: below 3e-7 int8 and
  below 1e-4 int16 multiples of a per-tile scale (each element off by at
  most m/254 or m/65534, i.e. < 1.2e-9 and < 1.6e-9, 2-3x the 6e-10 an FP32
  tile at the 1e-2 bound allows), below 1e-2 FP32, FP64 above; primitive pairs whose Cauchy-Schwarz bound is below 1e-12 left out
  (1e-10 for the derivative integrals); incremental Fock builds
  G(D) = G(D_prev) + G(D - D_prev) with tiles screened at 1e-13 (full
  builds at the start and before convergence is accepted); gradient with
  Schwarz 1e-11, density-weighted screening 1e-10 and FP32 primitive sums
  for quartets with Schwarz bound below 1e-2.  Fock accumulation,
  S^-1/2, the eigensolve / purification and DIIS stay in FP64.

Measured on the (H2O)_16 UHF triplet (480 functions): MIXED holds
|dE| ~ 2.5e-8 Eh and max |dgrad| ~ 2.8e-7 Eh/Bohr against STRICT (integer
thresholds 3e-5 / 1e-7, where no element is worse than an FP32 tile: dE
1.9e-8 at 5 % fewer full Fock builds/s; at 3e-4 / 1e-6 dE reaches 5e-7); the density
screening sets most of the gradient difference (8.5e-7 at 3e-10, 3.0e-6 at
1e-9), the FP32 bound of the derivative sums about 3e-8 at 1e-2.
"""

from __future__ import annotations

STRICT = dict(tau=1e-12, grad_tau=1e-12)

MIXED_GRAD = dict(grad_tau=1e-11, grad_tau_density=1e-10, grad_tau32=1e-2)

MIXED = dict(tau=1e-12, tau32=1e-4, tile_f32=1e-2, tile_i16=1e-4, tile_i8=3e-7, prim_eps=1e-12,
             fock_tau=1e-13, grad_prim_eps=1e-10, **MIXED_GRAD)
