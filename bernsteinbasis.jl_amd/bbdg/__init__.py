"""bbdg -- Bernstein-Bezier DG operators on MI355X (gfx950).

Host-side mirror of BernsteinBasis.jl's operator interface (src/BernsteinBasis.jl:19-23)
over the C-ABI library ``libbbdg.so`` built from ``csrc/``.
"""
from . import mesh, basis  # noqa: F401  (host set-up code, pure numpy)
