"""lulesh_amd: LULESH.jl's Lagrange leap-frog cycle on MI355X (gfx950).

Python mirror of the reference's Domain/driver API over the C ABI of
liblulesh_hip (include/lulesh_hip.h).
"""
from .domain import (ELEM_FIELDS, KERNELS, NODE_FIELDS, QUIRKS_REFERENCE, SCALARS, SHADOW_FIELDS,
                     Batch, Domain, LuleshError, LuleshProblem, c5_instances,
                     calcTimeConstraintsForElems, get_shadow, grad_run, grad_run_revolve,
                     lagrangeElements,
                     lagrangeLeapFrog, lagrangeNodal, main, replay, run, set_shadow,
                     timeIncrement, total_internal_energy, verify, verifyAndWriteFinalOutput,
                     vjp_cycle)

__all__ = ["Domain", "Batch", "c5_instances", "grad_run", "grad_run_revolve", "replay", "vjp_cycle", "set_shadow",
           "get_shadow", "total_internal_energy", "LuleshProblem", "LuleshError",
           "timeIncrement", "lagrangeLeapFrog", "lagrangeNodal", "lagrangeElements",
           "calcTimeConstraintsForElems", "run", "verify", "verifyAndWriteFinalOutput", "main",
           "QUIRKS_REFERENCE", "ELEM_FIELDS", "NODE_FIELDS", "SCALARS", "SHADOW_FIELDS",
           "KERNELS"]
