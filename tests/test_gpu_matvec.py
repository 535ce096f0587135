"""GPU parity of the FFT engine, the Green's tensor producer and the matvec
against the FP64 oracle (lse_matvect_mult.m / lse_generate_circulant_tensor.m)."""
import sys

import numpy as np
import pytest

from oracle import topology, green, matvec as omv
from tests import problems

pytestmark = pytest.mark.gpu

MU0 = 4 * np.pi * 1e-7
TOL64 = 1e-12     # FP64 path vs oracle, relative 2-norm
TOL32 = 2e-5      # FP32 path vs oracle, relative 2-norm
NXY = 3           # compiled X/Y-pass shapes per precision


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize('dims,pad_mode', [((2, 4, 4), 0), ((30, 10, 10), 0), ((7, 5, 3), 1), ((64, 12, 6), 1)])
@pytest.mark.parametrize('prec', [0, 1])
def test_fft3d(ctx, dims, pad_mode, prec):
    import torch
    import voxhenry_amd as vh
    say('fft3d', dims, pad_mode, prec)
    L, M, N = dims
    idx = np.arange(L * M * N)
    G = vh.Geometry(ctx, dims, 1e-6, idx, pad_mode=pad_mode)
    P = G.pad
    rng = np.random.default_rng(1)
    a = rng.standard_normal((2,) + P[::-1]) + 1j * rng.standard_normal((2,) + P[::-1])   # [c][z][y][x]
    dt = torch.complex128 if prec == 0 else torch.complex64
    t = torch.tensor(a, dtype=dt, device='cuda').contiguous()
    G.fft3d(t, prec, -1)
    torch.cuda.synchronize()
    ref = np.fft.fftn(a, axes=(1, 2, 3))
    err = rel(t.cpu().numpy(), ref)
    assert err < (1e-14 if prec == 0 else 2e-6), err
    G.fft3d(t, prec, +1)
    torch.cuda.synchronize()
    back = t.cpu().numpy() / np.prod(P)
    assert rel(back, a) < (1e-14 if prec == 0 else 2e-6)


@pytest.mark.parametrize('dims', [(2, 4, 4), (3, 3, 3), (6, 5, 4), (12, 7, 9)])
def test_green_realspace(ctx, dims):
    import voxhenry_amd as vh
    say('green', dims)
    L, M, N = dims
    dx = 1e-6
    G = vh.Geometry(ctx, dims, dx, np.arange(L * M * N), pad_mode=1)
    st = G.generate_green()
    Gd = G.get_green_realspace()
    Go, sto = green.generate_G(L, M, N, dx)
    for c in range(7):
        e = np.abs(Gd[..., c] - Go[..., c]).max() / np.abs(Go[..., c]).max()
        assert e < 1e-13, (c, e)
    assert np.allclose(st, sto, rtol=1e-14, atol=0)


def _problem(name):
    if name == 'wire':
        geo = problems.wire_vhr()
    elif name == 'cube3':
        geo = problems.bar(3, 3, 3)
    elif name == 'bar432':
        geo = problems.bar(4, 3, 2)
    elif name == 'cfg1':
        geo = problems.bar(30, 10, 10)
    elif name == 'hollow':
        geo = problems.bar(9, 6, 5)
        geo['sigma'][3:6, 2:4, 1:4] = 0.0
    return geo


@pytest.mark.parametrize('name', ['wire', 'cube3', 'bar432', 'hollow', 'cfg1'])
@pytest.mark.parametrize('tensor', ['oracle', 'gpu'])
def test_matvec_fp64_vs_oracle(ctx, name, tensor):
    import voxhenry_amd as vh
    say('matvec64', name, tensor)
    geo = _problem(name)
    s = topology.build_system(geo)
    L, M, N = s['dims']
    f = 2.5e9
    jwmu = 1j * 2 * np.pi * f * MU0
    fN2, st2 = green.circulant_tensor(L, M, N, geo['dx'])
    pad_mode = 0 if tensor == 'oracle' else 1
    G = vh.Geometry(ctx, s['dims'], geo['dx'], s['idxS'], pad_mode=pad_mode)
    if tensor == 'oracle':
        G.set_green_spectrum(fN2, 1.0, st2)
    else:
        G.generate_green()
    G.set_incidence(s['Ar'])
    z = problems.voxel_z(geo, s, f)
    G.set_impedance(z)
    zr, zi = topology.impedance(geo, s['idxS'], f)
    rng = np.random.default_rng(7)
    n = s['num_curr'] + s['num_node']
    X = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    Y = G.matvec(X, jwmu, vh.VH_C128, 0)
    for k in range(3):
        ref = omv.lse_matvect_mult(X[k], jwmu * fN2, s['Ar'], zr, zi, geo['dx'], s['idxS5'])
        assert rel(Y[k], ref) < TOL64, (k, rel(Y[k], ref))
    # FP32 path
    Y32 = G.matvec(X.astype(np.complex64), jwmu, vh.VH_C64, 0)
    ref = omv.lse_matvect_mult(X[0], jwmu * fN2, s['Ar'], zr, zi, geo['dx'], s['idxS5'])
    assert rel(Y32[0].astype(complex), ref) < TOL32


def test_matvec_device_tensors_and_modes(ctx):
    import torch
    import voxhenry_amd as vh
    geo = problems.bar(10, 6, 4)
    s = topology.build_system(geo)
    f = 1e8
    jwmu = 1j * 2 * np.pi * f * MU0
    G = problems.device_geometry(ctx, geo, s, pad_mode=1, f=f)
    rng = np.random.default_rng(3)
    n = G.n
    X = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    Yh = G.matvec(X, jwmu, vh.VH_C128, 0)
    Xd = torch.tensor(X, device='cuda')
    Yd = G.matvec(Xd, jwmu, vh.VH_C128, 0)
    torch.cuda.synchronize()
    assert np.abs(Yd.cpu().numpy() - Yh).max() == 0.0
    # mode 1 (FFT block + diagonal) equals mode 0 minus the incidence terms
    Y1 = G.matvec(X[:, :s['num_curr']].copy(), jwmu, vh.VH_C128, 1)
    inc = -(s['Ar'].T @ X[0, s['num_curr']:])
    assert rel(Y1[0] + inc, Yh[0, :s['num_curr']]) < 1e-13


def test_golden_solution_residual(ctx):
    """numex1 golden x (reference FP64 solution at 2.5 GHz) through the GPU operator."""
    import os
    import voxhenry_amd as vh
    say('golden residual')
    geo = problems.numex1()
    s = topology.build_system(geo)
    d = np.load(os.path.join(problems.GOLDEN, 'numex1_x.npz'))
    x = d['x']
    nc = s['num_curr']
    xr = np.concatenate([x[:nc], x[nc:][s['keep']]])
    f = float(d['freq'])
    G = problems.device_geometry(ctx, geo, s, pad_mode=1, f=f)
    jwmu = 1j * 2 * np.pi * f * MU0
    b = s['rhs'][:, 0]
    r = b - G.matvec(xr, jwmu, vh.VH_C128, 0)
    relres = np.linalg.norm(r) / np.linalg.norm(b)
    say('golden relres', relres)
    assert 1e-10 < relres < 3e-9


# Grids whose padded lengths exercise every radix of the pass kernels (2, 3,
# 4, 5, 7, 8, 16), one- to four-stage plans, long x lines (Px = 800 needs 64
# threads per line) and sparse voxel sets (the voxel-map gather and scatter).
# Thin z (Pz = 3, 4, 5, 7, 8 from N = 2..4 under both paddings) takes the
# one-radix pass-Z kernels, in FP64 the two-lanes-per-pencil one from Pz = 4.
RADIX_SHAPES = [((400, 4, 4), False, 1), ((9, 33, 17), True, 1), ((64, 3, 40), False, 1), ((5, 6, 2), True, 1),
                ((128, 20, 3), True, 1), ((16, 16, 16), False, 1), ((4, 4, 128), True, 1), ((6, 5, 64), False, 1),
                ((3, 4, 100), False, 1), ((9, 5, 2), False, 0), ((12, 6, 4), True, 0), ((70, 3, 3), True, 0)]


@pytest.mark.parametrize('dims,hollow,pad_mode', RADIX_SHAPES)
def test_matvec_radix_shapes(ctx, dims, hollow, pad_mode):
    import voxhenry_amd as vh
    say('matvec shapes', dims, hollow, pad_mode)
    L, M, N = dims
    geo = problems.bar(L, M, N)
    if hollow:
        rng = np.random.default_rng(11)
        mask = rng.random((L, M, N)) < 0.3
        mask[0, :, :] = False
        mask[-1, :, :] = False
        geo['sigma'][mask] = 0.0
    s = topology.build_system(geo)
    f = 1e9
    jwmu = 1j * 2 * np.pi * f * MU0
    fN2, _ = green.circulant_tensor(L, M, N, geo['dx'])
    G = problems.device_geometry(ctx, geo, s, pad_mode=pad_mode, f=f)
    zr, zi = topology.impedance(geo, s['idxS'], f)
    rng = np.random.default_rng(5)
    n = s['num_curr'] + s['num_node']
    X = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    Y = G.matvec(X, jwmu, vh.VH_C128, 0)
    Y32 = G.matvec(X.astype(np.complex64), jwmu, vh.VH_C64, 0)
    for k in range(3):
        ref = omv.lse_matvect_mult(X[k], jwmu * fN2, s['Ar'], zr, zi, geo['dx'], s['idxS5'])
        assert rel(Y[k], ref) < TOL64, (k, rel(Y[k], ref))
        assert rel(Y32[k].astype(complex), ref) < TOL32, (k, rel(Y32[k].astype(complex), ref))
    G.close()


@pytest.mark.parametrize('dims', [(16, 16, 16), (400, 4, 4), (9, 33, 17), (4, 4, 128), (5, 4, 100), (600, 6, 3)])
def test_matvec_every_zshape(ctx, dims):
    """Every compiled pass-Z shape and every X/Y-pass shape (the
    autotuner's choices) gives the same operator as the oracle; the
    autotuner itself leaves a valid choice.  (600 x 6 x 3: Px = 1200 is too
    long for the 128-thread X shape in FP64, which is then skipped.)"""
    import voxhenry_amd as vh
    say('zshapes', dims)
    L, M, N = dims
    geo = problems.bar(L, M, N)
    s = topology.build_system(geo)
    f = 3e9
    jwmu = 1j * 2 * np.pi * f * MU0
    fN2, _ = green.circulant_tensor(L, M, N, geo['dx'])
    G = problems.device_geometry(ctx, geo, s, pad_mode=1, f=f)
    zr, zi = topology.impedance(geo, s['idxS'], f)
    rng = np.random.default_rng(9)
    n = s['num_curr'] + s['num_node']
    X = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    ref = [omv.lse_matvect_mult(X[k], jwmu * fN2, s['Ar'], zr, zi, geo['dx'], s['idxS5']) for k in range(2)]
    for prec, nsh, tol in ((vh.VH_C128, 2, TOL64), (vh.VH_C64, 4, TOL32)):
        for sh in range(nsh):
            G.set_zshape(prec, sh)
            Y = G.matvec(X if prec == vh.VH_C128 else X.astype(np.complex64), jwmu, prec, 0)
            for k in range(2):
                e = rel(Y[k].astype(complex), ref[k])
                assert e < tol, (prec, sh, k, e)
        # every X/Y-pass shape (threads x element bound of the x and y passes)
        G.set_zshape(prec, 0)
        for sh in range(NXY):
            try:
                G.set_xyshape(prec, sh)
            except vh.VoxHenryError:
                continue          # cannot hold this grid's pencils
            Y = G.matvec(X if prec == vh.VH_C128 else X.astype(np.complex64), jwmu, prec, 0)
            for k in range(2):
                e = rel(Y[k].astype(complex), ref[k])
                assert e < tol, (prec, 'xy', sh, k, e)
        best, ms = G.autotune(prec, nrhs=2, reps=1)
        assert 0 <= best < nsh and len(ms) == nsh + NXY and np.all(ms > 0)
        shp = G.shapes()
        assert shp['z64' if prec == vh.VH_C128 else 'z32'] == best
        assert np.isfinite(ms[nsh + shp['xy64' if prec == vh.VH_C128 else 'xy32']])
    G.close()


@pytest.mark.parametrize('prec', ['c128', 'c64'])
def test_matvec_line_runs(ctx, prec):
    """x lines of every kind the X passes distinguish: one run starting past
    x = 0, one run ending before x = L, a full line, two runs, a single voxel
    at the far end, empty lines (run descriptors and the voxel-map path)."""
    import voxhenry_amd as vh
    say('line runs', prec)
    L, M, N = 20, 6, 3
    geo = problems.bar(L, M, N)
    s0 = geo['sigma'].copy()
    geo['sigma'][:] = 0.0
    geo['sigma'][4:15, 0, 0] = s0[4:15, 0, 0]          # one run inside the line
    geo['sigma'][:, 1, 0] = s0[:, 1, 0]                # full line
    geo['sigma'][0:5, 2, 0] = s0[0:5, 2, 0]            # two runs
    geo['sigma'][8:, 2, 0] = s0[8:, 2, 0]
    geo['sigma'][19, 3, 1] = s0[19, 3, 1]              # one voxel at x = L
    geo['sigma'][:12, 4, 2] = s0[:12, 4, 2]            # run from x = 0
    geo['sigma'][3:, 5, 2] = s0[3:, 5, 2]              # run to x = L
    geo['ports_P'] = [np.array([[20, 2, 1, 2]])]
    geo['ports_N'] = [np.array([[1, 2, 1, 1]])]
    s = topology.build_system(geo)
    f = 1e9
    jwmu = 1j * 2 * np.pi * f * MU0
    fN2, _ = green.circulant_tensor(L, M, N, geo['dx'])
    G = problems.device_geometry(ctx, geo, s, pad_mode=1, f=f)
    zr, zi = topology.impedance(geo, s['idxS'], f)
    rng = np.random.default_rng(21)
    n = s['num_curr'] + s['num_node']
    X = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    if prec == 'c128':
        Y, tol = G.matvec(X, jwmu, vh.VH_C128, 0), TOL64
    else:
        Y, tol = G.matvec(X.astype(np.complex64), jwmu, vh.VH_C64, 0), TOL32
    for k in range(2):
        ref = omv.lse_matvect_mult(X[k], jwmu * fN2, s['Ar'], zr, zi, geo['dx'], s['idxS5'])
        assert rel(Y[k].astype(complex), ref) < tol, (k, rel(Y[k].astype(complex), ref))
    G.close()
