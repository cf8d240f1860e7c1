"""registration_amd: MI355X (gfx950) implementation of py-lddmm's point-set LDDMM hot path.

Modules mirror the reference's (py-lddmm): kernelFunctions (Kernel), pointEvolution (landmark
flow / adjoint / EPDiff), surfaces and curves (currents data term), surfaceMatching and
curveMatching (drivers), conjugateGradient (cg).  ``serving`` adds continuous batching of many
registrations.  All kernel sums run in liblddmm_hip.so (hand-written HIP for gfx950).
"""
from ._lib import LIB_PATH, load  # noqa: F401

__all__ = ["kernelFunctions", "pointEvolution", "surfaces", "curves", "surfaceMatching",
           "curveMatching", "conjugateGradient", "matchingParam", "serving", "synthetic"]
