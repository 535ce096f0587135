"""Pins of the oracle's Green's-tensor producer (oracle/green.py)."""
import numpy as np
import pytest

from oracle import green

# gauss_1d.m literal tables, orders 3 and 5
GL3 = ([-0.774596669241483377035853079956, 0.0, 0.774596669241483377035853079956], [5 / 9, 8 / 9, 5 / 9])
GL5 = ([-0.906179845938663992797626878299, -0.538469310105683091036314420700, 0.0,
        0.538469310105683091036314420700, 0.906179845938663992797626878299],
       [0.236926885056189087514264040720, 0.478628670499366468041291514836, 0.568888888888888888888888888889,
        0.478628670499366468041291514836, 0.236926885056189087514264040720])


@pytest.mark.parametrize('n,tab', [(3, GL3), (5, GL5)])
def test_gauss_legendre_matches_reference_table(n, tab):
    x, w = green.gl_nodes(n)
    assert np.allclose(x, tab[0], atol=1e-15, rtol=0)
    assert np.allclose(w, tab[1], atol=1e-15, rtol=0)


def unit_cube_coulomb():
    """int_{[0,1]^3} int_{[0,1]^3} 1/|r-r'| by reduction to 8 int (1-u)(1-v)(1-w)/|u| over
    [0,1]^3, split into 3 pyramids (largest coordinate) where the integrand is smooth."""
    x, w = np.polynomial.legendre.leggauss(40)
    x = 0.5 * (x + 1); w = 0.5 * w
    u, s, t = np.meshgrid(x, x, x, indexing='ij')
    W = w[:, None, None] * w[None, :, None] * w[None, None, :]
    f = u * (1 - u) * (1 - u * s) * (1 - u * t) / np.sqrt(1 + s * s + t * t)
    return 8 * 3 * np.sum(W * f)


def test_ktable_self_term_closed_form():
    K = green.load_ktable()
    ref = unit_cube_coulomb() / (4 * np.pi) * (1e-3) ** 5
    assert abs(unit_cube_coulomb() - 1.882312644390) < 1e-11
    assert abs(K[0, 0].real - ref) / ref < 1e-11


def test_ktable_symmetries():
    K = green.load_ktable()
    # rows (1,2,1) and (2,1,1) are the same distance -> same pulse-pulse value
    assert abs(K[1, 0] - K[3, 0]) / abs(K[1, 0]) < 1e-13
    # linear(x)-pulse vanishes for offsets without an x component (rows 0..2)
    assert np.all(np.abs(K[:3, 1].real) < 1e-30 * 1e6)
    assert np.all(np.abs(K[3:, 1].real) > 1e-22)


def test_far_field_point_source_limit():
    dx = 1e-6
    R = 40
    K = green._vv_sums(np.array([[R * dx, 0, 0]]), dx, 3, green.K0_ONE_OVER_R)
    I = green._assemble7(K)[0]
    # pulse-pulse -> dx^6 / (4 pi R)
    ref = dx ** 6 / (4 * np.pi * R * dx)
    assert abs(I[0].real - ref) / ref < 1e-3


def test_medium_symmetric_scatter_equals_direct_quadrature():
    """The rotation scatter of the symmetric medium path reproduces the direct
    7-component quadrature at every medium offset except the singular self term."""
    dx = 1e-6
    n = 5
    G = np.zeros((n, n, n, 7), dtype=complex)
    rows = green._sym_rows(n, n, n)
    Ks = green._vv_sums(np.array(rows, float) * dx, dx, green.NP_MEDIUM, green.K0_ONE_OVER_R)
    green._scatter_sym(G, rows, Ks[:, [0, 1, 4]])
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing='ij'), -1).reshape(-1, 3)
    D = green._assemble7(green._vv_sums(g * dx, dx, green.NP_MEDIUM, green.K0_ONE_OVER_R)).reshape(n, n, n, 7)
    mask = np.ones((n, n, n), bool)
    mask[0, 0, 0] = False
    for c in range(7):
        a, b = G[..., c][mask], D[..., c][mask]
        scale = np.abs(D[..., 0][mask]).max() / dx ** (0 if c == 0 else 1 if c in (1, 2, 4) else 2)
        assert np.abs(a - b).max() < 1e-12 * scale, c


def test_dx_scaling():
    G1, _ = green.generate_G(4, 3, 3, 1e-6)
    G2, _ = green.generate_G(4, 3, 3, 2e-6)
    assert np.abs(G2.real - 2 * G1.real).max() < 1e-12 * np.abs(G1.real).max()


def test_embedding_and_spectral_parity():
    L, M, N = 5, 4, 3
    G, _ = green.generate_G(L, M, N, 1e-6)
    Gp = green.embed(G)
    assert Gp.shape == (10, 8, 6, 7)
    # reflection signs: comp 2 odd in x, comp 3 odd in y, comp 5 odd in z
    assert np.allclose(Gp[10 - 2, 1, 1, 1], -G[2, 1, 1, 1])
    assert np.allclose(Gp[1, 8 - 3, 2, 2], -G[1, 3, 2, 2])
    assert np.allclose(Gp[1, 1, 6 - 2, 4], -G[1, 1, 2, 4])
    assert np.allclose(Gp[10 - 3, 8 - 1, 6 - 1, 0], G[3, 1, 1, 0])
    assert np.all(Gp[L, :, :, :] == 0)
    F = np.fft.fftn(Gp, axes=(0, 1, 2))
    for c in range(7):
        odd = c in (1, 2, 4)
        big = np.abs(F[..., c]).max()
        small = np.abs(F[..., c].real if odd else F[..., c].imag).max()
        assert small < 1e-9 * big, c


def test_smooth_padding_gives_same_convolution():
    from oracle import matvec
    L, M, N = 5, 4, 3
    G, _ = green.generate_G(L, M, N, 1e-6)
    rng = np.random.default_rng(0)
    J = rng.standard_normal((L, M, N, 5)) + 1j * rng.standard_normal((L, M, N, 5))
    a = matvec.fft_block(J, np.fft.fftn(green.embed(G), axes=(0, 1, 2)), (L, M, N))
    b = matvec.fft_block(J, np.fft.fftn(green.embed(G, (9, 7, 5)), axes=(0, 1, 2)), (L, M, N))
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-14
