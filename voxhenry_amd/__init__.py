"""MI355X-native VoxHenry solve path: circulant-tensor FFT matvec, Schur
preconditioner and GMRES as hand-written HIP kernels behind a C ABI
(include/voxhenry.h), with a thin Python layer mirroring the reference's
operator interface."""
from ._lib import VH_C128, VH_C64, VoxHenryError, lib  # noqa: F401
from .core import Context, Geometry, Precon, MU0  # noqa: F401
