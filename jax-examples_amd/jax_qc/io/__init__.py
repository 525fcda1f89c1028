"""Molecule file formats (replaces jax_qc/io/)."""

from jax_qc.io.xyz import parse_xyz_string, read_xyz, write_xyz  # noqa: F401
