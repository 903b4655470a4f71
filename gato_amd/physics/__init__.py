"""Molecule drivers over the RHF hot path."""
