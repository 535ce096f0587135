"""Real-space Green's tensor, circulant embedding and its FFT.

Restated from src_lse_formation/lse_generate_circulant_tensor.m (default
branches: lse_enhanced_near=3, lse_exploit_symmetry=1, fl_oneoverr_kernel=1,
:16-26), with the quadrature of src_lin_vie/cubatures/volume_volume.m:1-117,
volume_volume_sym.m:1-121, weights_points.m:141-214, gauss_1d.m:80-115, the
sign table of src_lin_vie/gperiodic_coeff_nop_linJVIE.m:6-42 and
src_lin_vie/fft_operator.m:1-7.

Pinned by (tests/test_oracle_green.py):
* Gauss-Legendre nodes/weights equal the literal tables of gauss_1d.m;
* the K table's self term equals the closed-form unit-cube Coulomb integral;
* far-field terms tend to dx^6/(4 pi R) (point-source limit);
* even/odd spectral symmetry of the 7 components;
* the whole operator through the numex1 golden-solution residual
  (tests/test_oracle_solve.py).
"""
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
KTABLE = os.path.join(HERE, '..', 'tests', 'golden', 'K_near_interactions_1mm.txt')

K0_ONE_OVER_R = 1e-30          # lse_generate_circulant_tensor.m:23
NP_FAR, NP_MEDIUM, N_MEDIUM = 3, 5, 5   # :34-43

# gperiodic_coeff_nop_linJVIE.m:13-23 -- per-axis reflection signs of the 7 components
SIGN_X = np.array([1, -1, 1, 1, 1, 1, 1], dtype=float)
SIGN_Y = np.array([1, 1, -1, 1, 1, 1, 1], dtype=float)
SIGN_Z = np.array([1, 1, 1, 1, -1, 1, 1], dtype=float)


def gl_nodes(n):
    """Gauss-Legendre nodes/weights on [-1,1] (gauss_1d.m)."""
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def load_ktable(path=KTABLE):
    """K_near_interactions_1mm.txt: 6 rows x 3 complex (lse_generate_circulant_tensor.m:363)."""
    rows = []
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if line:
                rows.append([complex(t.replace('i', 'j')) for t in line.split(',')])
    return np.array(rows, dtype=complex)


def _vv_sums(off, dx, npts, k0):
    """The 7 kernel integrals of volume_volume.m:54-97 for a batch of offsets.

    off: (C,3) offsets r_m - r_n in metres.  Returns (C,7) with columns
    K1, K3, K5, K7, K8, K9, K10 (pulse-pulse, linear-pulse x/y/z,
    linear-linear x/y/z).  Uses the separable 6-D Gauss-Legendre product rule of
    weights_points.m, grouped as (x,x') pairs per axis.
    """
    x, w = gl_nodes(npts)
    h = dx / 2
    xi, xj = np.meshgrid(x, x, indexing='ij')      # (test, basis)
    wij = np.outer(w, w).ravel()
    d = (xi - xj).ravel()                           # test - basis
    t = (h * xi).ravel()                            # linear test function
    tt = (h * xj * h * xi).ravel()                  # linear-linear
    J = h ** 6
    out = np.zeros((off.shape[0], 7), dtype=complex)
    C = max(1, 200000 // (len(d) ** 3))
    for s in range(0, off.shape[0], C):
        o = off[s:s + C]
        ax = o[:, 0, None] + h * d[None, :]
        ay = o[:, 1, None] + h * d[None, :]
        az = o[:, 2, None] + h * d[None, :]
        R = np.sqrt(ax[:, :, None, None] ** 2 + ay[:, None, :, None] ** 2 + az[:, None, None, :] ** 2)
        with np.errstate(divide='ignore', invalid='ignore'):
            ker = np.exp(-1j * k0 * R) / R / (4 * np.pi)
        W = wij[:, None, None] * wij[None, :, None] * wij[None, None, :]
        kw = ker * W[None]
        s_x = kw.sum(axis=(2, 3))                   # (c, px)
        s_y = kw.sum(axis=(1, 3))
        s_z = kw.sum(axis=(1, 2))
        out[s:s + C, 0] = kw.sum(axis=(1, 2, 3)) * J
        out[s:s + C, 1] = (s_x * t).sum(1) * J
        out[s:s + C, 2] = (s_y * t).sum(1) * J
        out[s:s + C, 3] = (s_z * t).sum(1) * J
        out[s:s + C, 4] = (s_x * tt).sum(1) * J
        out[s:s + C, 5] = (s_y * tt).sum(1) * J
        out[s:s + C, 6] = (s_z * tt).sum(1) * J
    return out


def _assemble7(K):
    """volume_volume.m:100-115: I = [K1, -K3, K5, K8+K9, 2K7, K8-K9, K8+K9+4K10]."""
    K1, K3, K5, K7, K8, K9, K10 = K.T
    return np.stack([K1, -K3, K5, K8 + K9, 2 * K7, K8 - K9, K8 + K9 + 4 * K10], axis=-1)


def _sym_rows(n1, n2, n3):
    """inter_idx order of lse_generate_circulant_tensor.m:151-161 / :342-352."""
    return [(mx, my, mz) for mx in range(n1) for my in range(n2) for mz in range(my + 1)]


def _scatter_sym(G, rows, K):
    """Rotational/transposition scatter of lse_generate_circulant_tensor.m:184-305
    (identical code at :405-490).  rows are 0-based (mx,my,mz); K is (n,3) =
    [pulse-pulse, linear(x)-pulse, linear(x)-linear(x')]."""
    for (mx, my, mz), (k1, k2, k3) in zip(rows, K):
        G[mx, my, mz, 0] = k1
        G[mx, my, mz, 1] = -k2
        G[mx, my, mz, 3] += k3; G[mx, my, mz, 5] += k3; G[mx, my, mz, 6] += k3
        if my != mz:
            G[mx, mz, my, 0] = k1
            G[mx, mz, my, 1] = -k2
            G[mx, mz, my, 3] += k3; G[mx, mz, my, 5] += k3; G[mx, mz, my, 6] += k3
        G[mz, mx, my, 2] = k2
        G[mz, mx, my, 3] += k3; G[mz, mx, my, 5] -= k3; G[mz, mx, my, 6] += k3
        if my != mz:
            G[my, mx, mz, 2] = k2
            G[my, mx, mz, 3] += k3; G[my, mx, mz, 5] -= k3; G[my, mx, mz, 6] += k3
        G[my, mz, mx, 4] = 2 * k2
        G[my, mz, mx, 6] += 4 * k3
        if my != mz:
            G[mz, my, mx, 4] = 2 * k2
            G[mz, my, mx, 6] += 4 * k3


def generate_G(L, M, N, dx, k0=K0_ONE_OVER_R, ktable=None):
    """Normalised real-space tensor G_mn (L,M,N,7) complex and st_sparse_precon (3,),
    lse_generate_circulant_tensor.m:52-772 with ko=1 (the driver passes ko=1,
    VoxHenry_executer.m:141)."""
    if min(L, M, N) < 2:
        raise ValueError('grid dims must be >= 2 (near-field fast path, :318-328)')
    G = np.zeros((L, M, N, 7), dtype=complex)
    # far, :56-74
    g = np.stack(np.meshgrid(np.arange(L), np.arange(M), np.arange(N), indexing='ij'), -1).reshape(-1, 3)
    G[...] = _assemble7(_vv_sums(g * dx, dx, NP_FAR, k0)).reshape(L, M, N, 7)
    # medium, :79-310
    nL, nM, nN = min(L, N_MEDIUM), min(M, N_MEDIUM), min(N, N_MEDIUM)
    if nL != nM or nL != nN:
        gm = np.stack(np.meshgrid(np.arange(nL), np.arange(nM), np.arange(nN), indexing='ij'), -1).reshape(-1, 3)
        G[:nL, :nM, :nN, :] = _assemble7(_vv_sums(gm * dx, dx, NP_MEDIUM, k0)).reshape(nL, nM, nN, 7)
    else:
        rows = _sym_rows(nL, nM, nN)
        Ks = _vv_sums(np.array(rows, dtype=float) * dx, dx, NP_MEDIUM, k0)
        G[:nL, :nM, :nN, :] = 0
        _scatter_sym(G, rows, Ks[:, [0, 1, 4]])      # volume_volume_sym dir=1: K1, K3, K8
    # near from the K table, :328-490
    Kt = load_ktable() if ktable is None else np.array(ktable, dtype=complex)
    rows = _sym_rows(2, 2, 2)
    if len(rows) != Kt.shape[0]:
        raise ValueError('K table has %d rows, expected %d' % (Kt.shape[0], len(rows)))
    s = dx / 1e-3                                    # :397-400
    Kt = Kt * np.array([s ** 5, s ** 6, s ** 7])
    G[:2, :2, :2, :] = 0
    _scatter_sym(G, rows, Kt)
    G = _normalise(G, dx)
    st = np.array([G[0, 0, 0, 0], G[0, 0, 0, 3], G[0, 0, 0, 6]])      # :772
    return G, st


def _normalise(G, dx):
    """Basis normalisation of the 7 components, :761-768."""
    G = G / dx ** 4
    G[..., 1:3] /= dx
    G[..., 3] /= dx ** 2
    G[..., 4] /= dx
    G[..., 5:7] /= dx ** 2
    return G


def far_entries(m, dx, k0=K0_ONE_OVER_R):
    """Normalised tensor entries G_mn (C,7) at integer offsets m (C,3) from the
    far-field rule alone (:56-74, then :761-768): the values generate_G holds
    at every offset outside the medium zone [0, 5)^3, on any grid.  Lets a
    test check single entries of a large grid's tensor without computing the
    whole grid."""
    m = np.asarray(m, dtype=float).reshape(-1, 3)
    return _normalise(_assemble7(_vv_sums(m * dx, dx, NP_FAR, k0)), dx)


def embed(G, pad=None):
    """Signed circulant embedding, lse_generate_circulant_tensor.m:776-816.

    ``pad`` = padded sizes (default 2L,2M,2N as the reference); any size >= 2n-1
    gives the same linear convolution on the cropped grid.
    """
    L, M, N, _ = G.shape
    Px, Py, Pz = pad if pad is not None else (2 * L, 2 * M, 2 * N)
    Gp = np.zeros((Px, Py, Pz, 7), dtype=complex)

    def idx(n, P):
        # source index for each padded index along one axis (-1 = zero) and reflection flag
        src = np.full(P, -1)
        ref = np.zeros(P, dtype=bool)
        src[:n] = np.arange(n)
        for i in range(1, n):
            src[P - i] = i
            ref[P - i] = True
        return src, ref
    sx, rx = idx(L, Px)
    sy, ry = idx(M, Py)
    sz, rz = idx(N, Pz)
    vx, vy, vz = sx >= 0, sy >= 0, sz >= 0
    sub = G[np.ix_(sx[vx], sy[vy], sz[vz])]           # (px,py,pz,7)
    sign = (np.where(rx[vx][:, None], SIGN_X[None], 1.0)[:, None, None, :] *
            np.where(ry[vy][:, None], SIGN_Y[None], 1.0)[None, :, None, :] *
            np.where(rz[vz][:, None], SIGN_Z[None], 1.0)[None, None, :, :])
    Gp[np.ix_(np.flatnonzero(vx), np.flatnonzero(vy), np.flatnonzero(vz))] = sub * sign
    return Gp


def circulant_tensor(L, M, N, dx, pad=None):
    """(fN_all2, st_sparse_precon2) as returned by lse_generate_circulant_tensor(dx,1,L,M,N,0)."""
    G, st = generate_G(L, M, N, dx)
    Gp = embed(G, pad)
    fN = np.fft.fftn(Gp, axes=(0, 1, 2))              # fft_operator.m:6-7
    return fN, st
