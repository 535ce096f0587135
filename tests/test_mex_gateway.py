"""The MATLAB drop-in boundary: the MEX gateway voxhenry_amd/mex/vh_mex.c,
compiled against the test-only fake MEX host (tests/fake_mex), driven the way
the shadow .m files drive it from VoxHenry_executer.m.

CPU tests: argument checking, state errors and memory hygiene (nothing here
touches a device).  GPU tests: the gateway's matvec, preconditioner, Schur
product and pgmres/fpgmres against the FP64 oracle (lse_matvect_mult.m,
lse_sparse_precon_multiply.m, lse_sparse_schur_compl_multiply.m, pgmres.m),
MATLAB data conventions included: 1-based idxS as doubles, full L x M x N
impedance grids, a real-typed right-hand side, CSC sparse Ae."""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import topology, green, precon as opc, gmres as ogm, driver, matvec as omv
from tests import problems

HERE = os.path.dirname(os.path.abspath(__file__))
MU0 = 4 * np.pi * 1e-7


class MexError(Exception):
    def __init__(self, ident, msg):
        super().__init__('%s: %s' % (ident, msg))
        self.ident, self.msg = ident, msg


class FakeMatlab:
    """mxArray conversion and fm_call through ctypes."""

    def __init__(self):
        import voxhenry_amd._lib as vl
        vl.lib()    # torch's HIP runtime and libvoxhenry.so first: the gateway links to the same copy
        path = os.path.join(HERE, 'fake_mex', 'libvh_mex_fake.so')
        if not os.path.exists(path):
            from tests.fake_mex import build as fb
            fb.build(verbose=False)
        L = C.CDLL(path)
        P, I64 = C.c_void_p, C.c_int64
        sig = {'fm_dense': ([C.c_int, P, C.c_int, P], P), 'fm_sparse': ([I64, I64, P, P, P], P),
               'fm_string': ([C.c_char_p], P), 'fm_destroy': ([P], None),
               'fm_call': ([C.c_int, P, C.c_int, P], C.c_int), 'fm_error_id': ([], C.c_char_p),
               'fm_error_msg': ([], C.c_char_p), 'fm_shape': ([P, P, P], C.c_int), 'fm_read': ([P, P], None),
               'fm_leaked': ([], C.c_longlong), 'fm_locked': ([], C.c_int), 'fm_exit': ([], C.c_int)}
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes, f.restype = args, res
        self.L = L

    def _arr(self, v):
        L = self.L
        if isinstance(v, str):
            return L.fm_string(v.encode())
        if sp.issparse(v):
            m = sp.csc_matrix(v, dtype=np.float64)
            m.sort_indices()
            jc = np.ascontiguousarray(m.indptr, dtype=np.int64)
            ir = np.ascontiguousarray(m.indices, dtype=np.int64)
            pr = np.ascontiguousarray(m.data, dtype=np.float64)
            return L.fm_sparse(m.shape[0], m.shape[1], jc.ctypes.data, ir.ctypes.data, pr.ctypes.data)
        if v is None:
            a = np.zeros((0, 0))
        else:
            a = np.asarray(v)
            if a.ndim == 0:
                a = a.reshape(1, 1)
            elif a.ndim == 1:
                a = a.reshape(-1, 1)    # MATLAB column vector
        cplx = np.iscomplexobj(a)
        flat = np.ascontiguousarray(a.ravel(order='F'), dtype=np.complex128 if cplx else np.float64)
        if cplx:
            flat = flat.view(np.float64)
        dims = np.array(a.shape, dtype=np.int64)
        return L.fm_dense(a.ndim, dims.ctypes.data, int(cplx), flat.ctypes.data if flat.size else None)

    def _out(self, p):
        dims = np.zeros(8, np.int64)
        cplx = C.c_int(0)
        nd = self.L.fm_shape(p, dims.ctypes.data, C.byref(cplx))
        shape = tuple(int(d) for d in dims[:nd])
        n = int(np.prod(shape))
        buf = np.empty(2 * n if cplx.value else n, np.float64)
        if n:
            self.L.fm_read(p, buf.ctypes.data)
        out = buf.view(np.complex128) if cplx.value else buf
        return out.reshape(shape, order='F')

    def __call__(self, cmd, *args, nout=0):
        ins = [self._arr(cmd)] + [self._arr(a) for a in args]
        prhs = (C.c_void_p * len(ins))(*ins)
        plhs = (C.c_void_p * max(nout, 1))()
        try:
            failed = self.L.fm_call(nout, plhs, len(ins), prhs)
            if failed:
                raise MexError(self.L.fm_error_id().decode(), self.L.fm_error_msg().decode())
            outs = []
            for i in range(max(nout, 1)):
                if plhs[i]:
                    outs.append(self._out(plhs[i]))
                    self.L.fm_destroy(plhs[i])
                else:
                    outs.append(None)
            return outs[0] if nout <= 1 else outs
        finally:
            for a in ins:
                self.L.fm_destroy(a)

    def leaked(self):
        return self.L.fm_leaked()

    def locked(self):
        return self.L.fm_locked()

    def exit(self):
        return self.L.fm_exit()


_ML = None


def matlab():
    global _ML
    if _ML is None:
        _ML = FakeMatlab()
    return _ML


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def expect(ident, fn, *args, **kw):
    with pytest.raises(MexError) as e:
        fn(*args, **kw)
    assert e.value.ident == ident, e.value
    return e.value.msg


# ---------------------------------------------------------------------------
# CPU: argument and state checks happen before the device is touched
# ---------------------------------------------------------------------------

def test_usage_and_state_errors():
    m = matlab()
    m('reset')
    expect('voxhenry:usage', m, 'no_such_command')
    expect('voxhenry:usage', m, 'geom', 4, 3)                         # too few arguments
    expect('voxhenry:usage', m, 'geom', 4, 3, 2.5, 1e-6, [1.0, 2.0])  # N not an integer
    expect('voxhenry:usage', m, 'geom', 4, 3, 2, -1e-6, [1.0, 2.0])   # dx <= 0
    msg = expect('voxhenry:usage', m, 'geom', 4, 3, 2, 1e-6, [1.0, 0.0])
    assert 'idxS(2)' in msg                                           # 1-based: 0 is not a voxel
    expect('voxhenry:usage', m, 'geom', 4, 3, 2, 1e-6, [2.0, 1.0])    # not increasing
    expect('voxhenry:usage', m, 'geom', 4, 3, 2, 1e-6, [1.0, 25.0])   # beyond L*M*N
    expect('voxhenry:usage', m, 'geom', 4, 3, 2, 1e-6, [1.5])         # not integral
    expect('voxhenry:state', m, 'matvec', np.ones(10))
    expect('voxhenry:state', m, 'precon_apply', np.ones(10))
    expect('voxhenry:state', m, 'incidence', sp.eye(3, format='csc'))
    expect('voxhenry:usage', m, 'pgmres', np.ones(10), 100, -1.0, 10)  # tol <= 0 is a usage error first
    expect('voxhenry:state', m, 'pgmres', np.ones(10), 100, 1e-8, 10)
    expect('voxhenry:state', m, 'solve_batch', np.ones((10, 1)), np.array([1j]), [1.0], 'fp64', 1e-8)
    # prepare: Ae must have 5*numel(idxS) columns; the type must be supported
    zr = np.ones((2, 2, 1))
    idx = np.array([1.0, 2.0, 3.0, 4.0])
    st = np.array([1 + 1j, 2j, 3j])
    expect('voxhenry:usage', m, 'prepare', 1e-6, zr, None, idx, st, sp.csc_matrix((3, 19)))
    expect('voxhenry:usage', m, 'prepare', 1e-6, zr, None, idx, st, sp.csc_matrix((3, 20)), 'schur_approx')
    expect('voxhenry:usage', m, 'prepare', 1e-6, zr, np.ones(3), idx, st, sp.csc_matrix((3, 20)))
    expect('voxhenry:usage', m, 'matvec', np.ones(3), nout=2)
    assert m.locked() == 0          # no context was created by any of these
    assert m.leaked() == 0          # every mxMalloc'ed temporary was released by the gateway


@pytest.mark.skipif(__import__('torch').cuda.device_count() > 0, reason='checks the no-device error path')
def test_library_error_without_device():
    m = matlab()
    m('reset')
    msg = expect('voxhenry:lib', m, 'geom', 4, 3, 2, 1e-6, np.arange(1.0, 25.0))
    assert 'libvoxhenry error' in msg
    assert m.locked() == 0 and m.leaked() == 0


# ---------------------------------------------------------------------------
# GPU: the drop-in calls against the oracle
# ---------------------------------------------------------------------------

def _grid(geo, s, f):
    zr, zi = topology.impedance(geo, s['idxS'], f)
    return zr, zi


@pytest.mark.gpu
@pytest.mark.parametrize('kind', ['hollow', 'superconductor'])
def test_gateway_matvec(kind):
    m = matlab()
    m('reset')
    if kind == 'hollow':
        geo = problems.bar(9, 6, 5)
        geo['sigma'][3:6, 2:4, 1:4] = 0.0
        f = 3e8
    else:
        geo = problems.bar(8, 4, 3, lam=np.full((8, 4, 3), 8e-8))
        f = 2e9
    s = topology.build_system(geo)
    L, M, N = s['dims']
    fN2, st2o = green.circulant_tensor(L, M, N, geo['dx'])
    st2 = m('geom', L, M, N, geo['dx'], s['idxS'] + 1.0, nout=1)
    assert st2.shape == (1, 3)
    assert rel(st2.ravel(), np.asarray(st2o).ravel()) < 1e-13
    assert m.locked() == 1
    m('incidence', s['Ar'])
    zr, zi = _grid(geo, s, f)
    m('impedance', zr, zi) if zi is not None else m('impedance', zr)
    jwmu = 1j * 2 * np.pi * f * MU0
    n = s['num_curr'] + s['num_node']
    rng = np.random.default_rng(3)
    x_real = rng.standard_normal(n)                          # real-typed, as the first Arnoldi vector
    X = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    y1 = m('matvec', x_real, jwmu, nout=1)
    Y = m('matvec', X, jwmu, nout=1)
    assert y1.shape == (n, 1) and Y.shape == (n, 2)
    A = lambda v: omv.lse_matvect_mult(v, jwmu * fN2, s['Ar'], zr, zi, geo['dx'], s['idxS5'])
    assert rel(y1[:, 0], A(x_real.astype(complex))) < 1e-12
    for k in range(2):
        assert rel(Y[:, k], A(X[:, k])) < 1e-12
    expect('voxhenry:usage', m, 'matvec', X[:-1], jwmu)      # wrong row count
    if zi is not None:
        # the batched tier uses one impedance for every frequency: a two-fluid
        # (frequency-dependent) impedance is refused, not silently reused
        expect('voxhenry:usage', m, 'solve_batch', X, np.array([jwmu]), np.ones(2), 'fp64', 1e-8)
    m('reset')
    assert m.locked() == 0 and m.leaked() == 0
    expect('voxhenry:state', m, 'matvec', X, jwmu)


@pytest.mark.gpu
def test_gateway_executer_flow():
    """The executer's per-frequency sequence through the shadow files:
    lse_sparse_precon_prepare -> 'prepare', then pgmres with the matvec and
    preconditioner handles -> 'pgmres'; a second frequency reuses the geometry."""
    m = matlab()
    m('reset')
    geo = problems.bar(12, 5, 4)
    prob = driver.Problem(geo)
    s = prob.sys
    b = s['rhs'][:, 0]                                         # real, as the executer builds it
    for f in (1e3, 2.5e9):
        jwmu = 1j * 2 * np.pi * f * MU0
        zr, zi = topology.impedance(geo, s['idxS'], f)
        st = jwmu * np.asarray(prob.st2)                       # st_sparse_precon, executer :267
        m('prepare', geo['dx'], zr, zi, s['idxS'] + 1.0, st, s['Ar'])
        A, Pm, _, _, ps = prob.operators(f)
        rng = np.random.default_rng(int(f) % 97)
        v = rng.standard_normal(len(b)) + 1j * rng.standard_normal(len(b))
        assert rel(m('matvec', v, nout=1)[:, 0], A(v)) < 1e-12            # jwmu recovered from st / st2
        assert rel(m('precon_apply', v, nout=1)[:, 0], Pm(v)) < 1e-11
        w = v[s['num_curr']:]
        assert rel(m('schur_multiply', w, nout=1)[:, 0],
                   opc.lse_sparse_schur_compl_multiply(w, s['Ar'], ps.Ainv)) < 1e-12
        tol = driver.reference_tol(f)
        x, flag, relres, it, resvec = m('pgmres', b, 100, tol, 10, nout=5)
        xo, flo, rro, ito, rvo, _ = ogm.pgmres(A, b, 100, tol, 10, Pm)
        assert flag[0, 0] == 0 and flo == 0 and relres[0, 0] <= tol
        assert it.shape == (1, 2) and it[0, 1] == ito[1]                  # iter = [p outit], pgmres.m:178
        assert resvec.shape[1] == 1 and abs(len(resvec) - len(rvo)) <= 2
        assert rel(x[:, 0], xo) < 1e-7
        Yg, Yo = b @ x[:, 0], b @ xo
        assert abs(Yg - Yo) / abs(Yo) < 1e-9
    # schur_gmres: flexible outer solve (executer :297) with prectol 0.1
    m('prepare', geo['dx'], zr, zi, s['idxS'] + 1.0, st, s['Ar'], 'schur_gmres')
    x2, flag2, relres2 = m('fpgmres', b, 100, 1e-8, 10, nout=3)
    assert flag2[0, 0] == 0 and relres2[0, 0] <= 1e-8
    assert abs(b @ x2[:, 0] - Yg) / abs(Yg) < 1e-7
    m('reset')
    assert m.locked() == 0 and m.leaked() == 0


@pytest.mark.gpu
def test_gateway_exit_handler_releases():
    m = matlab()
    m('reset')
    geo = problems.bar(4, 3, 2)
    s = topology.build_system(geo)
    L, M, N = s['dims']
    m('geom', L, M, N, geo['dx'], s['idxS'] + 1.0)
    assert m.locked() == 1
    assert m.exit() == 1            # MATLAB clearing the MEX: the handler releases the device objects
    expect('voxhenry:state', m, 'incidence', s['Ar'])
    assert m.leaked() == 0


@pytest.mark.gpu
def test_gateway_solve_batch():
    """The batched tier from MATLAB: every (port, frequency) column in one
    device solve, a preconditioner per frequency, the reference's tolerance
    per column; Y of each column against the oracle's pgmres."""
    m = matlab()
    m('reset')
    geo = problems.bar(12, 5, 4)
    prob = driver.Problem(geo)
    s = prob.sys
    L, M, N = s['dims']
    m('geom', L, M, N, geo['dx'], s['idxS'] + 1.0)
    m('incidence', s['Ar'])
    zr, zi = topology.impedance(geo, s['idxS'], 1e9)
    m('impedance', zr, zi)
    freqs = np.array([1e3, 1e8, 2.5e9])
    jw = 1j * 2 * np.pi * freqs * MU0
    b = s['rhs'][:, 0]
    B = np.repeat(b[:, None], 3, axis=1)                   # real-typed, as the executer builds it
    pidx = np.array([1.0, 2.0, 3.0])
    tols = np.array([driver.reference_tol(f) for f in freqs])
    expect('voxhenry:usage', m, 'solve_batch', B, jw, np.array([1.0, 2.0, 4.0]), 'fp64', tols)
    expect('voxhenry:usage', m, 'solve_batch', B, jw, pidx, 'fp16', tols)
    X, flag, relres, its, inner = m('solve_batch', B, jw, pidx, 'fp64', tols, nout=5)
    assert X.shape == (len(b), 3) and flag.shape == (1, 3)
    assert np.all(flag == 0) and np.all(relres[0] <= tols)
    for k, f in enumerate(freqs):
        A, Pm, *_ = prob.operators(f)
        xo = ogm.pgmres(A, b, 100, tols[k], 10, Pm)[0]
        Yg, Yo = b @ X[:, k], b @ xo
        assert abs(Yg - Yo) / abs(Yo) < 1e-9, (f, abs(Yg - Yo) / abs(Yo))
    Xm, flagm, relm = m('solve_batch', B, jw, pidx, 'mixed', 1e-8, nout=3)
    assert np.all(flagm == 0) and np.all(relm[0] <= 1e-8)
    for k in range(1, 3):
        assert abs(b @ Xm[:, k] - b @ X[:, k]) / abs(b @ X[:, k]) < 1e-6
    # fp64_true: every column of this bar converges in the pgmres pass, so the
    # solutions are those of 'fp64' bit for bit
    Xt, flagt = m('solve_batch', B, jw, pidx, 'fp64_true', tols, nout=2)
    assert np.all(flagt == 0) and np.array_equal(Xt, X)
    m('reset')
    assert m.locked() == 0 and m.leaked() == 0
