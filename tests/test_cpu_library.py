"""CPU checks of the C ABI library: it loads without a GPU, exports every
symbol include/voxhenry.h declares, and its host-only sparse LDL^T
(vh_ldl_solve_host, the ldlchol/ldlsolve stand-in) solves Schur systems."""
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import topology
from tests import problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, 'include', 'voxhenry.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(vh_[a-z0-9_]+)\s*\(', txt)))


def test_library_exports_every_declared_symbol():
    from voxhenry_amd._lib import lib, SIGNATURES
    L = lib()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
        assert n in SIGNATURES, 'ctypes signature missing for ' + n
    assert L.vh_version() >= 100


def test_error_path_without_device():
    """Argument errors come back as status codes with a message."""
    import ctypes as C
    from voxhenry_amd._lib import lib
    L = lib()
    rc = L.vh_ctx_create(0, None)
    assert rc == 1 and b'null' in L.vh_last_error()


def face_coords(s):
    L, M, N = s['dims']
    nv = s['nv']
    ArT = s['Ar'].T.tocsr()
    nn = s['Ar'].shape[0]
    xyz = np.zeros((nn, 3), np.float32)
    typ = -np.ones(nn, np.int32)
    sub = np.array(np.unravel_index(s['idxS'], (L, M, N), order='F')).T
    for d in range(3):
        for v in range(nv):
            r = d * nv + v
            for q in range(ArT.indptr[r], ArT.indptr[r + 1]):
                c = sub[v] + 0.5
                c[d] = sub[v][d] + (1 if ArT.data[q] > 0 else 0)
                xyz[ArT.indices[q]] = c
                typ[ArT.indices[q]] = d
    return xyz, typ


@pytest.mark.parametrize('geometric', [False, True])
@pytest.mark.parametrize('name', ['wire', 'bar', 'hollow'])
def test_host_ldl_solve(name, geometric):
    from voxhenry_amd._lib import lib, ptr, check
    geo = {'wire': problems.wire_vhr, 'bar': lambda: problems.bar(14, 5, 4),
           'hollow': lambda: problems.bar(9, 6, 5)}[name]()
    if name == 'hollow':
        geo['sigma'][3:6, 2:4, 1:4] = 0.0
    s = topology.build_system(geo)
    rng = np.random.default_rng(1)
    Ainv = rng.uniform(0.5, 2.0, s['num_curr'])
    S = (s['Ar'] @ sp.diags(Ainv) @ s['Ar'].T).tocsc()
    S.sort_indices()
    n = S.shape[0]
    B = np.asfortranarray(rng.standard_normal((n, 3)))
    X = np.zeros((n, 3), order='F')
    st = np.zeros(3, np.int64)
    xyz, typ = face_coords(s) if geometric else (None, None)
    check(lib().vh_ldl_solve_host(n, ptr(S.indptr.astype(np.int64)), ptr(S.indices.astype(np.int64)), ptr(S.data),
                                  3, ptr(B), ptr(X), ptr(xyz), ptr(typ), ptr(st)))
    ref = spla.spsolve(S, B)
    assert np.linalg.norm(X - ref) / np.linalg.norm(ref) < 1e-12
    assert st[0] > 0 and st[1] >= 1


def test_host_ldl_rejects_indefinite():
    from voxhenry_amd._lib import lib, ptr
    S = sp.csc_matrix(np.array([[1.0, 2.0], [2.0, 1.0]]))
    S.sort_indices()
    b = np.ones(2)
    x = np.zeros(2)
    rc = lib().vh_ldl_solve_host(2, ptr(S.indptr.astype(np.int64)), ptr(S.indices.astype(np.int64)), ptr(S.data),
                                 1, ptr(b), ptr(x), None, None, None)
    assert rc == 4 and b'pivot' in lib().vh_last_error()
