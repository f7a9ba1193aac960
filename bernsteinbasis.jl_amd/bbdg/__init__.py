"""bbdg -- Bernstein-Bezier DG operators on MI355X (gfx950).

Host-side mirror of BernsteinBasis.jl's operator interface (src/BernsteinBasis.jl:19-23)
over the C-ABI library ``libbbdg.so`` built from ``csrc/``.  Importing the package needs
no GPU; creating a Context (or any device array / operator call) does, and raises if the
library is not built or no device is visible.
"""
from . import basis, mesh  # noqa: F401  (host set-up code, numpy)
from ._lib import BBError, lib  # noqa: F401
from .device import Context, DeviceArray, default_context  # noqa: F401
from .operators import (  # noqa: F401
    BernsteinDerivativeMatrix_3D_r,
    BernsteinDerivativeMatrix_3D_s,
    BernsteinDerivativeMatrix_3D_t,
    BernsteinLift,
    ElevationMatrix,
    MultithreadedBernsteinLift,
    bernstein_2d_scalar_multiindex_lookup,
    face_gather_,
    face_lift_multiply_,
    fmask,
    grad_,
    ij_to_linear,
    ijk_to_linear,
    interface_flux_,
    l_j,
    linear_to_ijkl_lookup,
    mul_,
    tet_offsets,
    threaded_mul_,
    tri_offsets,
)
from .solver import Solver  # noqa: F401
