"""bsamgpu -- MI355X-native adaptive FAS multigrid V-cycle (BSAM) host layer.

The solve path (V-cycle kernels) lives in the HIP library ``libbsamgpu.so``
behind the C ABI declared in ``include/bsamgpu.h``; this package is the host
side above it, mirroring the reference's MATLAB host layer
(ebProblem / ebOptions / ebSetup / ebTagCluster / ebAddLevel / ebSolve).
"""
from .problem import Options, Problem, BsamError  # noqa: F401
from . import hierarchy, tagcluster  # noqa: F401
