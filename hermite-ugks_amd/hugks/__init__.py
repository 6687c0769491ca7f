"""Host side of the MI355X Hermite-UGKS time-step library.

``capi``        ctypes binding of libhugks.so (include/hugks.h)
``cases``       &params -> descriptor (the reference's initialize())
``namelist``    reader for the reference's namelist files
``quadrature``  Gauss-Hermite / Newton-Cotes velocity grids
"""
from . import capi, cases, namelist, quadrature  # noqa: F401
