"""Schur-complement preconditioner, restated from
src_lse_formation/lse_sparse_precon_prepare.m:28-104,149-159 (schur_invert,
ldlt_decomp, CHOLMOD) and lse_sparse_precon_multiply.m:61-97 (apply),
lse_sparse_schur_compl_multiply.m:1-5, and the schur_gmres variant
(lse_sparse_precon_prepare.m:83-88, lse_sparse_precon_multiply.m:26-46).

CHOLMOD 3.0.12 ``ldlchol``/``ldlsolve`` (third party, not vendored) compute an
exact sparse LDL^T solve of the SPD Schur matrix S = Ar*A_inv*Ar'; this oracle
uses SuperLU (scipy) for the same exact solve -- the ordering changes only
rounding.  Pinned by the exact-inverse identity: applying the preconditioner to
[A_inv^-1 x - Ar' y; Ar x] returns (x, y) (tests/test_oracle_precon.py).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


class PreconState:
    pass


def precon_prepare(dx, z_real, z_imag, idxS, st, Ar):
    """A_inv (real diagonal, :36-64), S = Ar*A_inv*Ar' (:102) and its factor (:154)."""
    z = z_real.ravel(order='F')[idxS].astype(complex)
    if z_imag is not None:
        z = z + z_imag.ravel(order='F')[idxS]
    dp = 1.0 / (z / dx + st[0])
    d2 = 1.0 / (z / (6 * dx) + st[1])
    d3 = 1.0 / (z / (2 * dx) + st[2])
    Ainv = np.concatenate([np.abs(dp), np.abs(dp), np.abs(dp), np.abs(d2), np.abs(d3)])
    S = (Ar @ sp.diags(Ainv) @ Ar.T).tocsc()
    s = PreconState()
    s.Ainv, s.S, s.Ar = Ainv, S, Ar.tocsr()
    s.ArT = Ar.T.tocsr()
    s.lu = spla.splu(S, permc_spec='MMD_AT_PLUS_A', diag_pivot_thresh=0.0,
                     options=dict(SymmetricMode=True))
    return s


def schur_solve(s, bb):
    bb = np.asarray(bb)
    if np.iscomplexobj(bb):
        return s.lu.solve(np.ascontiguousarray(bb.real)) + 1j * s.lu.solve(np.ascontiguousarray(bb.imag))
    return s.lu.solve(bb)


def lse_sparse_precon_multiply(v, s):
    """lse_sparse_precon_multiply.m:84-88:
    bb = b - Ae*A_inv*a;  y = S^-1 bb;  x = (a + Ae'*y) .* diag(A_inv)."""
    nc = s.Ainv.shape[0]
    a, b = v[:nc], v[nc:]
    bb = b - s.Ar @ (s.Ainv * a)
    y = schur_solve(s, bb)
    out = np.empty(v.shape, dtype=complex)
    out[nc:] = y
    out[:nc] = (a + s.ArT @ y) * s.Ainv
    return out


def lse_sparse_schur_compl_multiply(schur_in, Ar, Ainv):
    """S*v = Ae*A_inv*Ae'*v, lse_sparse_schur_compl_multiply.m:3."""
    return Ar @ (Ainv * (Ar.T @ schur_in))


def schur_jacobi_diag(s):
    """diag(S) of the schur_gmres Jacobi preconditioner, lse_sparse_precon_prepare.m:85-86."""
    return np.asarray(s.S.diagonal()).ravel()


def lse_sparse_precon_multiply_schur_gmres(v, s, prectol=0.1):
    """schur_gmres branch, lse_sparse_precon_multiply.m:26-46: bb = b - Ae*A_inv*a;
    d = gmres(S, bb, min(100, num_curr-1), prectol, 10, diag(S)) (MATLAB gmres:
    left preconditioned, x0 = 0, stopping on the preconditioned residual -- the
    semantics pgmres.m keeps); x = (a + Ae'*d) .* diag(A_inv).  Returns (x, inner
    Arnoldi steps)."""
    from . import gmres as og
    nc = s.Ainv.shape[0]
    a, b = v[:nc], v[nc:]
    bb = b - s.Ar @ (s.Ainv * a)
    dg = schur_jacobi_diag(s)
    d, flag, relres, it, resvec, nmv = og.pgmres(lambda u: lse_sparse_schur_compl_multiply(u, s.Ar, s.Ainv), bb,
                                                 min(100, nc - 1), prectol, 10, lambda u: u / dg)
    out = np.empty(v.shape, dtype=complex)
    out[nc:] = d
    out[:nc] = (a + s.ArT @ d) * s.Ainv
    return out, len(resvec) - 1
