"""ctypes binding of libvoxhenry.so (the C ABI declared in include/voxhenry.h).

The library is loaded from this package directory; if it is missing the import
fails loudly -- there is no CPU fallback for the solve path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, 'libvoxhenry.so')

VH_C128 = 0
VH_C64 = 1
VH_PC_SCHUR_INVERT = 0
VH_PC_SCHUR_GMRES = 1

_lib = None


class VoxHenryError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__('voxhenry error %d: %s' % (code, msg))
        self.code = code


_P = C.c_void_p
_I64 = C.c_int64
_D = C.c_double
_I = C.c_int
_PD = C.POINTER(C.c_double)
_PI64 = C.POINTER(C.c_int64)

# name -> argtypes (all return int status unless listed in _RESTYPES)
SIGNATURES = {
    'vh_ctx_create': [_I, C.POINTER(_P)],
    'vh_ctx_destroy': [_P],
    'vh_ctx_set_stream': [_P, _P],
    'vh_ctx_synchronize': [_P],
    'vh_last_error': [],
    'vh_version': [],
    'vh_geom_create': [_P, _I64, _I64, _I64, _D, _I64, _P, _I, C.POINTER(_P)],
    'vh_geom_destroy': [_P],
    'vh_geom_info': [_P, _P],
    'vh_geom_generate_green': [_P, _P],
    'vh_geom_set_green_spectrum': [_P, _P, _I64, _I64, _I64, _D, _D, _P],
    'vh_geom_set_green_realspace': [_P, _P, _P],
    'vh_geom_get_green_realspace': [_P, _P],
    'vh_geom_autotune': [_P, _I, _I64, _I, _P, _P, _P],
    'vh_precon_prepare_many': [_P, _I64, _P, _P, _I],
    'vh_geom_set_zshape': [_P, _I, _I],
    'vh_geom_set_xyshape': [_P, _I, _I],
    'vh_geom_get_shapes': [_P, _P],
    'vh_geom_set_incidence': [_P, _I64, _I64, _P, _P, _P],
    'vh_geom_set_impedance': [_P, _P],
    'vh_geom_build_topology': [_P, _I64, _P, _P, _P, _P, _I64, _P, _P],
    'vh_geom_get_rhs': [_P, _P],
    'vh_matvec': [_P, _D, _D, _I, _I, _I64, _P, _I64, _P, _I64, _I],
    'vh_matvec_noalloc': [_P, _D, _D, _I, _I, _I64, _P, _I64, _P, _I64],
    'vh_matvec_timed': [_P, _D, _D, _I, _I, _I64, _P, _I64, _P, _I64, _I, _P, _P, _P],
    'vh_fft3d': [_P, _P, _I, _I, _I],
    'vh_solve_batch': [_P, _I64, _P, _I64, _P, _P, _P, _P, _I, _I, _I, _D, _P, _I, _I, _I, _D, _D, _P, _P, _P, _P,
                       _P],
    'vh_solve_batch_bytes': [_P, _I, _I, _I, _P],
    'vh_geom_currents_on_grid': [_P, _I64, _P, _I64, _P, _I64, _P, _I],
    'vh_precon_create': [_P, C.POINTER(_P)],
    'vh_precon_destroy': [_P],
    'vh_precon_prepare': [_P, _D, _D, _P],
    'vh_precon_apply': [_P, _I, _I64, _P, _I64, _P, _I64, _I],
    'vh_precon_get_ainv': [_P, _P],
    'vh_precon_set_type': [_P, _I, _D],
    'vh_precon_inner_steps': [_P, _P],
    'vh_precon_schur_multiply': [_P, _I64, _P, _I64, _P, _I64, _I],
    'vh_gmres': [_P, _P, _D, _D, _I, _I, _I64, _P, _P, _I, _I, _D, _I, _P, _P, _P, _P, _P, _I64, _P],
    'vh_ldl_solve_host': [_I64, _P, _P, _P, _I64, _P, _P, _P, _P, _P],
    'vh_solve_mixed': [_P, _P, _D, _D, _I64, _P, _P, _I, _D, _I, _I, _I, _D, _P, _P, _P, _P],
}
_RESTYPES = {'vh_last_error': C.c_char_p, 'vh_version': C.c_int}


def lib():
    global _lib
    if _lib is None:
        # torch ships its own HIP runtime (SONAME libamdhip64.so.7) whose users
        # reference the unversioned name; load it first so this library binds to
        # the same runtime instead of pulling a second copy into the process.
        import torch  # noqa: F401
        if not os.path.exists(LIBPATH):
            raise ImportError('libvoxhenry.so not built (%s); run `python -m voxhenry_amd.build` '
                              '-- the solve path has no CPU fallback' % LIBPATH)
        L = C.CDLL(LIBPATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPES.get(name, C.c_int)
        _lib = L
    return _lib


def declared_symbols():
    return list(SIGNATURES)


def check(rc):
    if rc != 0:
        raise VoxHenryError(rc, lib().vh_last_error().decode(errors='replace'))


def ptr(a):
    """Raw pointer of a numpy array or a torch tensor."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    return C.c_void_p(a.data_ptr())
