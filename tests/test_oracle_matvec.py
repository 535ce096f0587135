"""Pins of the oracle matvec (oracle/matvec.py, lse_matvect_mult.m)."""
import numpy as np
import pytest

from oracle import green, matvec, topology
from tests import problems

# 5x5 block structure of lse_matvect_mult.m:57-86 as (out, in, tensor comp, sign)
BLOCKS = [(0, 0, 0, 1), (0, 3, 1, 1), (0, 4, 1, 1),
          (1, 1, 0, 1), (1, 3, 2, 1), (1, 4, 2, -1),
          (2, 2, 0, 1), (2, 4, 4, 1),
          (3, 0, 1, -1), (3, 1, 2, -1), (3, 3, 3, 1), (3, 4, 5, 1),
          (4, 0, 1, -1), (4, 1, 2, 1), (4, 2, 4, -1), (4, 3, 5, 1), (4, 4, 6, 1)]
SX = green.SIGN_X; SY = green.SIGN_Y; SZ = green.SIGN_Z


def brute_conv(G, J):
    """sum_n G~(m-n) J(n) with G~ the signed even/odd extension of G (direct sum)."""
    L, M, N, _ = G.shape
    out = np.zeros((L, M, N, 5), dtype=complex)
    m = np.stack(np.meshgrid(np.arange(L), np.arange(M), np.arange(N), indexing='ij'), -1).reshape(-1, 3)
    for a in range(len(m)):
        d = m[a] - m                       # m - n for all n
        ad = np.abs(d)
        for (o, i, k, s) in BLOCKS:
            sgn = s * np.where(d[:, 0] < 0, SX[k], 1) * np.where(d[:, 1] < 0, SY[k], 1) * np.where(d[:, 2] < 0, SZ[k], 1)
            g = G[ad[:, 0], ad[:, 1], ad[:, 2], k]
            out[tuple(m[a]) + (o,)] += np.sum(sgn * g * J[m[:, 0], m[:, 1], m[:, 2], i])
    return out


@pytest.mark.parametrize('dims', [(2, 4, 4), (3, 3, 3), (4, 3, 2)])
def test_fft_block_equals_direct_convolution(dims):
    L, M, N = dims
    G, _ = green.generate_G(L, M, N, 1e-6)
    fN = np.fft.fftn(green.embed(G), axes=(0, 1, 2))
    rng = np.random.default_rng(2)
    J = rng.standard_normal((L, M, N, 5)) + 1j * rng.standard_normal((L, M, N, 5))
    a = matvec.fft_block(J, fN, dims)
    b = brute_conv(G, J)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-13


def dense_operator(geo, f):
    s = topology.build_system(geo)
    L, M, N = s['dims']
    fN2, st2 = green.circulant_tensor(L, M, N, geo['dx'])
    jwmu = 1j * 2 * np.pi * f * 4e-7 * np.pi
    zr, zi = topology.impedance(geo, s['idxS'], f)
    n = s['num_curr'] + s['num_node']
    A = np.zeros((n, n), dtype=complex)
    for j in range(n):
        e = np.zeros(n, dtype=complex)
        e[j] = 1
        A[:, j] = matvec.lse_matvect_mult(e, jwmu * fN2, s['Ar'], zr, zi, geo['dx'], s['idxS5'])
    return A, s


def test_dense_operator_structure():
    """Unit-vector probing (the method of lse_generate_Zmat.m:22-60): the current
    block Z is complex symmetric and the off-diagonal blocks are -Ar' and Ar."""
    geo = problems.bar(3, 2, 2)
    A, s = dense_operator(geo, 1e9)
    nc = s['num_curr']
    Z = A[:nc, :nc]
    assert np.abs(Z - Z.T).max() < 1e-12 * np.abs(Z).max()
    assert np.abs(A[:nc, nc:] + s['Ar'].T.toarray()).max() == 0
    assert np.abs(A[nc:, :nc] - s['Ar'].toarray()).max() == 0
    assert np.abs(A[nc:, nc:]).max() == 0


def test_superconductor_diagonal():
    """Two-fluid z_real + z_imag enters the diagonal as (z_real+z_imag) w_c / dx."""
    geo = problems.bar(3, 2, 2, sigma=1e7, lam=np.full((3, 2, 2), 90e-9))
    f = 1e9
    s = topology.build_system(geo)
    zr, zi = topology.impedance(geo, s['idxS'], f)
    assert zi is not None and np.all(np.abs(zi) > 0)
    # normal-conductor limit lambda -> 0 (flagged by lambdaL == 0) restores 1/sigma
    geo2 = problems.bar(3, 2, 2, sigma=1e7, lam=np.zeros((3, 2, 2)))
    zr2, zi2 = topology.impedance(geo2, s['idxS'], f)
    assert np.allclose(zr2, 1e-7) and np.all(zi2 == 0)


@pytest.mark.parametrize('sigma,lam,f', [(1e7, 90e-9, 1e6), (1e7, 90e-9, 1e10), (5.8e7, 50e-9, 3e8),
                                         (1e6, 200e-9, 1e11), (0.0, 90e-9, 1e9)])
def test_two_fluid_impedance_identity(sigma, lam, f):
    """Two-fluid conductor (VoxHenry_executer.m:243-254): the normal channel
    sigma and the superfluid channel 1/(j w mu lam^2) conduct in parallel, so
    z_real + z_imag = 1 / (sigma + 1/(j w mu lam^2)); at sigma = 0 the
    impedance is the kinetic inductance j w mu lam^2 alone.  A dropped square
    or sign in the reference's den = (sigma w mu lam^2)^2 + 1 fails this.
    The product's per-voxel impedance (solver.impedance) obeys it too."""
    from voxhenry_amd import solver as vsol
    mu = 4 * np.pi * 1e-7
    jwml = 1j * 2 * np.pi * f * mu * lam ** 2
    expect = jwml if sigma == 0 else 1.0 / (sigma + 1.0 / jwml)
    geo = problems.bar(3, 2, 2, sigma=sigma, lam=np.full((3, 2, 2), lam))
    zr, zi = topology.impedance(geo, np.arange(12), f)
    z = zr + zi
    assert np.all(np.abs(z - expect) <= 1e-14 * abs(expect)), (z.ravel()[0], expect)
    st = dict(sigma=np.full((3, 2, 2), float(sigma)), lambdaL=np.full((3, 2, 2), lam))
    zp = vsol.impedance(st, np.arange(12), f)
    assert np.all(np.abs(zp - expect) <= 1e-14 * abs(expect))
