"""Per-frequency solve loop and Y -> Z -> R, L, restated from
VoxHenry_executer.m:215-328.

Pinned by: exact DC resistance of a full-face-port bar (R = l/(sigma A)), the
committed R/L table of results_numex1_straight_conductor/data_R_jL_mat.mat and
the numex1 golden solution (tests/test_oracle_solve.py).
"""
import numpy as np
from . import topology, green, matvec, precon, gmres

MU0 = 4 * np.pi * 1e-7          # EMconstants.m:13


def reference_tol(f):
    """VoxHenry_executer.m:219-223."""
    return 1e-12 if f < 1e6 else 1e-8


class Problem:
    """Geometry + operator data shared by all frequencies (executer :84-193)."""

    def __init__(self, geo, pad=None):
        self.geo = geo
        self.sys = topology.build_system(geo)
        L, M, N = self.sys['dims']
        self.fN2, self.st2 = green.circulant_tensor(L, M, N, geo['dx'], pad)

    def operators(self, f):
        """(A, P, z_real, z_imag, state) for frequency f (executer :231-290)."""
        s = self.sys
        z_real, z_imag = topology.impedance(self.geo, s['idxS'], f)
        jwmu = 1j * 2 * np.pi * f * MU0
        fN = jwmu * self.fN2                                       # :266
        st = jwmu * self.st2                                       # :267
        state = precon.precon_prepare(s['dx'], z_real, z_imag, s['idxS'], st, s['Ar'])
        A = lambda J: matvec.lse_matvect_mult(J, fN, s['Ar'], z_real, z_imag, s['dx'], s['idxS5'])
        P = lambda v: precon.lse_sparse_precon_multiply(v, state)
        return A, P, z_real, z_imag, state


def port_admittance(rhs, x):
    """Y(:,port) = [sum(rhs(:,kk).*x)]_kk, unconjugated (executer :311-316)."""
    return rhs.T @ x


def z_to_rl(Y, f):
    """Z = inv(Y); R = |Re Z|; L = |Im Z|/(2 pi f) (executer :327-328)."""
    Z = np.linalg.inv(Y)
    return Z, np.abs(Z.real), np.abs(Z.imag) / (2 * np.pi * f)


def solve_sweep(prob, freqs, tol=None, restart=100, maxit=10, flexible=False):
    """Returns dict(Y, Z, R, L, x (last), flags, relres, iters) over freqs."""
    s = prob.sys
    nports = s['rhs'].shape[1]
    out = dict(Y=[], Z=[], R=[], L=[], flags=[], relres=[], iters=[], x=[])
    for f in freqs:
        t = reference_tol(f) if tol is None else tol
        A, P, *_ = prob.operators(f)
        Y = np.zeros((nports, nports), dtype=complex)
        xs = []
        for p in range(nports):
            if flexible:
                x, flag, rr, it, rv = gmres.fpgmres(A, s['rhs'][:, p], restart, t, maxit, P)
            else:
                x, flag, rr, it, rv, _ = gmres.pgmres(A, s['rhs'][:, p], restart, t, maxit, P)
            Y[:, p] = port_admittance(s['rhs'], x)
            out['flags'].append(flag)
            out['relres'].append(rr)
            out['iters'].append(len(rv) - 1)
            xs.append(x)
        Z, R, Lh = z_to_rl(Y, f)
        out['Y'].append(Y); out['Z'].append(Z); out['R'].append(R); out['L'].append(Lh)
        out['x'].append(np.stack(xs, axis=1))
    for k in ('Y', 'Z', 'R', 'L'):
        out[k] = np.array(out[k])
    return out
