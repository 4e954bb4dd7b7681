"""MI355X (gfx950) HIP implementation of PolyOrder's periodic pseudospectral
SCFT hot path: RQM4 / OS2 propagators, Simpson / Quad4 density quadrature and
Anderson mixing, behind the C ABI of include/polystring_amd.h."""
from . import capi, workloads  # noqa: F401
from .capi import Context, ModelAB, Propagator, density  # noqa: F401
