"""lulesh_amd: LULESH.jl's Lagrange leap-frog cycle on MI355X (gfx950).

Python mirror of the reference's Domain/driver API over the C ABI of
liblulesh_hip (include/lulesh_hip.h).
"""
from .domain import (ELEM_FIELDS, KERNELS, NODE_FIELDS, QUIRKS_REFERENCE, SCALARS, Batch, Domain,
                     c5_instances, grad_run, replay, total_internal_energy,
                     LuleshError, LuleshProblem, calcTimeConstraintsForElems, lagrangeElements,
                     lagrangeLeapFrog, lagrangeNodal, main, run, timeIncrement, verify,
                     verifyAndWriteFinalOutput)

__all__ = ["Domain", "Batch", "c5_instances", "grad_run", "replay", "total_internal_energy", "LuleshProblem", "LuleshError", "timeIncrement", "lagrangeLeapFrog",
           "lagrangeNodal", "lagrangeElements", "calcTimeConstraintsForElems", "run", "verify",
           "verifyAndWriteFinalOutput", "main", "QUIRKS_REFERENCE", "ELEM_FIELDS",
           "NODE_FIELDS", "SCALARS", "KERNELS"]
