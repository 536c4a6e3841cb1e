"""MI355X (gfx950) helical cone-beam FDK reconstruction.

A from-scratch HIP implementation of the hot path of
RuiLiu1217/AnalyticalReconBigPatient: curved-detector fan-to-parallel rebinning,
cone/row preweighting with the ramp filter, and 3D-weighted helical FDK
backprojection, exported through the reference's C entry points
(``include/arecon.h``) and mirrored here at the MATLAB level
(:mod:`.mexapi`) and as a device-resident pipeline (:mod:`.pipeline`).

Modules:
  _lib       ctypes binding of libarecon.so (raises if it is missing)
  geometry   the MATLAB drivers' geometry derivations (FP64)
  mexapi     ParallelRebinningCBCurve, reWeigAdFiltr_GPU,
             Parallel_FDK_Helical_3DWeighting, ... on host arrays
  pipeline   HelicalFDK: rebin -> filter -> backproject in HBM (needs torch)
  phantom    analytic ellipsoid phantoms and their projections
  build      hipcc/gcc build recipe
"""
from ._lib import AreconError, lib  # noqa: F401

__version__ = "0.1.0"
