"""Propagators of the hot path (ADFWI/propagator)."""
from .acoustic_propagator import AcousticPropagator
from .boundary_condition import bc_gerjan, bc_pml, bc_pml_xz, bc_sincos
from .elastic_propagator import ElasticPropagator

__all__ = ["AcousticPropagator", "ElasticPropagator", "bc_pml", "bc_pml_xz", "bc_sincos", "bc_gerjan"]
