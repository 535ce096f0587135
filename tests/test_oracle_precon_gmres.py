"""Pins of the oracle preconditioner (exact Schur inverse) and GMRES semantics."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import gmres, precon, topology, green
from tests import problems


def _prep(geo, f):
    s = topology.build_system(geo)
    L, M, N = s['dims']
    _, st2 = green.generate_G(L, M, N, geo['dx'])
    jwmu = 1j * 2 * np.pi * f * 4e-7 * np.pi
    zr, zi = topology.impedance(geo, s['idxS'], f)
    return s, precon.precon_prepare(geo['dx'], zr, zi, s['idxS'], jwmu * st2, s['Ar'])


@pytest.mark.parametrize('f', [1.0, 1e9])
def test_precon_is_exact_inverse_of_sparse_block(f):
    geo = problems.bar(5, 3, 3)
    s, P = _prep(geo, f)
    nc = s['num_curr']
    rng = np.random.default_rng(0)
    x = rng.standard_normal(nc) + 1j * rng.standard_normal(nc)
    y = rng.standard_normal(s['num_node']) + 1j * rng.standard_normal(s['num_node'])
    v = np.concatenate([x / P.Ainv - s['Ar'].T @ y, s['Ar'] @ x])
    out = precon.lse_sparse_precon_multiply(v, P)
    assert np.linalg.norm(out[:nc] - x) / np.linalg.norm(x) < 1e-12
    assert np.linalg.norm(out[nc:] - y) / np.linalg.norm(y) < 1e-12


def test_schur_matrix_spd():
    geo = problems.bar(4, 3, 2)
    s, P = _prep(geo, 1e6)
    S = P.S.toarray()
    assert np.abs(S - S.T).max() == 0
    assert np.linalg.eigvalsh(S).min() > 0
    v = np.random.default_rng(1).standard_normal(S.shape[0])
    assert np.allclose(precon.lse_sparse_schur_compl_multiply(v, s['Ar'], P.Ainv), S @ v, rtol=1e-13)


def test_schur_gmres_variant_pins():
    """schur_gmres (lse_sparse_precon_multiply.m:26-46): with a tight inner
    tolerance it is the exact inverse (the schur_invert apply); at the
    reference's prectol = 0.1 the Jacobi-preconditioned inner residual meets
    the tolerance, and diag(S) is the diagonal of Ar diag(A_inv) Ar'."""
    geo = problems.bar(6, 3, 3)
    s, P = _prep(geo, 1e9)
    nc = s['num_curr']
    rng = np.random.default_rng(3)
    v = rng.standard_normal(nc + s['num_node']) + 1j * rng.standard_normal(nc + s['num_node'])
    exact = precon.lse_sparse_precon_multiply(v, P)
    tight, it_tight = precon.lse_sparse_precon_multiply_schur_gmres(v, P, 1e-13)
    assert np.linalg.norm(tight - exact) / np.linalg.norm(exact) < 1e-10
    dg = precon.schur_jacobi_diag(P)
    Ad = abs(s['Ar']).multiply(abs(s['Ar'])) @ P.Ainv          # sum_k Ar(v,k)^2 A_inv(k)
    assert np.allclose(dg, Ad, rtol=1e-14)
    loose, it_loose = precon.lse_sparse_precon_multiply_schur_gmres(v, P, 0.1)
    bb = v[nc:] - P.Ar @ (P.Ainv * v[:nc])
    d = loose[nc:]
    r = (bb - P.S @ d) / dg
    assert np.linalg.norm(r) / np.linalg.norm(bb / dg) <= 0.1
    assert 0 < it_loose < it_tight


def _dense_case(n=30, seed=0, diag=4.0):
    rng = np.random.default_rng(seed)
    A = np.eye(n) * diag + rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return A, b


def test_pgmres_full_krylov_is_exact():
    A, b = _dense_case()
    x, flag, relres, it, rv, _ = gmres.pgmres(lambda v: A @ v, b, restart=30, tol=1e-13, maxit=1)
    assert flag == 0 and relres < 1e-12
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10)


def test_pgmres_left_preconditioned_and_restarted():
    A, b = _dense_case(40, 3, diag=20.0)
    Minv = np.linalg.inv(np.diag(np.diag(A)))
    x, flag, relres, it, rv, _ = gmres.pgmres(lambda v: A @ v, b, restart=7, tol=1e-10, maxit=50,
                                              L1=lambda v: Minv @ v)
    # flag reports the TRUE residual (pgmres.m:181-186) while the loop stops on
    # the preconditioned one (pgmres_iter.m:126)
    assert flag == int(relres > 1e-10)
    assert relres < 1e-8
    assert rv[-1] < 1e-10
    assert it[1] >= 2          # needed restarts
    # an exact preconditioner converges in one step
    Ainv = np.linalg.inv(A)
    x1, flag1, rr1, it1, rv1, _ = gmres.pgmres(lambda v: A @ v, b, restart=5, tol=1e-10, maxit=2,
                                               L1=lambda v: Ainv @ v)
    assert it1 == (1, 1) and rr1 < 1e-12


def test_pgmres_maxit_flag():
    A, b = _dense_case(40, 4)
    x, flag, relres, it, rv, _ = gmres.pgmres(lambda v: A @ v, b, restart=2, tol=1e-14, maxit=1)
    assert flag == 1 and it == (2, 1)


def test_fpgmres_right_preconditioned():
    A, b = _dense_case(40, 5)
    Minv = np.linalg.inv(np.diag(np.diag(A)))
    x, flag, relres, it, rv = gmres.fpgmres(lambda v: A @ v, b, restart=40, tol=1e-11, maxit=3,
                                            R1=lambda v: Minv @ v)
    assert flag == 0 and relres < 1e-10
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-8)
