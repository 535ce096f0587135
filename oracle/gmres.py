"""Restarted GMRES with the reference's semantics.

* ``pgmres``  -- src_iterative_solver/pgmres.m:70-197 + pgmres_iter.m:22-153:
  left preconditioner L1 (applied as a function, iterapp.m:57), one-pass
  classical Gram-Schmidt (:107-108), least squares re-solved every step
  (:114-121), stop when the preconditioned residual < tol*||L1 b|| (:126,
  pgmres.m:142,159), restart with r = V*res (:153), final true relres (:181).
* ``fpgmres`` -- fpgmres.m:58-152 + fpgmres_iter.m:26-121: flexible right
  preconditioning (Z basis), bnorm = ||b||.

Pinned by (tests/test_oracle_gmres.py): small dense systems against
numpy.linalg.solve, full-Krylov exactness (n steps solve an n x n system), the
stopping rule and the restart semantics.
"""
import numpy as np


def _lstsq(Hc, rhs):
    y = np.linalg.lstsq(Hc, rhs, rcond=None)[0]
    return y, rhs - Hc @ y


def pgmres_iter(A, x, r, m, L1, tol):
    n = r.shape[0]
    V = np.zeros((n, m + 1), dtype=complex)
    rn = np.linalg.norm(r)
    V[:, 0] = r / rn
    H = np.zeros((m + 1, m), dtype=complex)
    resvec = np.zeros(m)
    for k in range(1, m + 1):
        w = A(V[:, k - 1])
        if L1 is not None:
            w = L1(w)
        H[:k, k - 1] = V[:, :k].conj().T @ w
        w = w - V[:, :k] @ H[:k, k - 1]
        H[k, k - 1] = np.linalg.norm(w)
        V[:, k] = w / H[k, k - 1]                      # no breakdown guard (:111)
        rhs = np.zeros(k + 1, dtype=complex)
        rhs[0] = rn
        y, res = _lstsq(H[:k + 1, :k], rhs)
        resvec[k - 1] = np.linalg.norm(res)
        if resvec[k - 1] < tol:
            break
    x = x + V[:, :k] @ y
    r = V[:, :k + 1] @ res
    return x, r, k, resvec[:k]


def pgmres(A, b, restart=50, tol=1e-3, maxit=10, L1=None):
    """Returns x, flag, relres, iter=(p, outit), resvec, n_matvec."""
    b = np.asarray(b, dtype=complex)
    r = b.copy()
    bp = b.copy()
    if L1 is not None:
        r = L1(r)
        bp = L1(bp)
    bnorm = np.linalg.norm(bp)
    resvec = [np.linalg.norm(r) / bnorm]
    x = np.zeros_like(b)
    outit = 0
    p = 0
    nmv = 0
    while resvec[-1] > tol and outit < maxit:
        outit += 1
        x, r, p, rv = pgmres_iter(A, x, r, restart, L1, tol * bnorm)
        nmv += p
        resvec += list(rv / bnorm)
    relres = np.linalg.norm(b - A(x)) / np.linalg.norm(b)
    flag = 0 if relres <= tol else 1
    return x, flag, relres, (p, outit), np.array(resvec), nmv + 1


def fpgmres_iter(A, x, r, m, R1, tol):
    n = r.shape[0]
    V = np.zeros((n, m + 1), dtype=complex)
    Z = np.zeros((n, m), dtype=complex)
    rn = np.linalg.norm(r)
    V[:, 0] = r / rn
    H = np.zeros((m + 1, m), dtype=complex)
    resvec = np.zeros(m)
    for k in range(1, m + 1):
        Z[:, k - 1] = R1(V[:, k - 1]) if R1 is not None else V[:, k - 1]
        w = A(Z[:, k - 1])
        H[:k, k - 1] = V[:, :k].conj().T @ w
        w = w - V[:, :k] @ H[:k, k - 1]
        H[k, k - 1] = np.linalg.norm(w)
        V[:, k] = w / H[k, k - 1]
        rhs = np.zeros(k + 1, dtype=complex)
        rhs[0] = rn
        y, res = _lstsq(H[:k + 1, :k], rhs)
        resvec[k - 1] = np.linalg.norm(res)
        if resvec[k - 1] < tol:
            break
    x = x + Z[:, :k] @ y
    r = V[:, :k + 1] @ res
    return x, r, k, resvec[:k]


def fpgmres(A, b, restart=50, tol=1e-3, maxit=10, R1=None):
    b = np.asarray(b, dtype=complex)
    r = b.copy()
    bnorm = np.linalg.norm(b)
    resvec = [np.linalg.norm(r) / bnorm]
    x = np.zeros_like(b)
    outit = 0
    p = 0
    while resvec[-1] > tol and outit < maxit:
        outit += 1
        x, r, p, rv = fpgmres_iter(A, x, r, restart, R1, tol * bnorm)
        resvec += list(rv / bnorm)
    relres = np.linalg.norm(b - A(x)) / np.linalg.norm(b)
    flag = 0 if relres <= tol else 1
    return x, flag, relres, (p, outit), np.array(resvec)
