"""Many molecules per launch: batched integrals, Fock builds and RHF.

The reference runs one molecule at a time (jax_qc/scf/rhf.py:52-169); this
package runs a list of molecules of uneven size together: one ERI launch per
quartet class for all of them, one Fock build (four kernels, captured as one
graph) per SCF iteration for all of them, and the SCF algebra batched over
molecules on padded matrices."""

from jax_qc.batch.rhf import BatchPairTable, BatchedRHF, plan_quartets_batched

__all__ = ["BatchPairTable", "BatchedRHF", "plan_quartets_batched"]
