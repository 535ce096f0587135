"""The allocation-free matvec entry point (vh_matvec_noalloc: pass
intermediates in the library's fixed per-GPU scratch array) against
vh_matvec on the same device vectors: identical results on two calls in a
row, the second with a smaller geometry whose empty x lines rely on the
scratch being cleared of the first call's data; a request beyond the fixed
scratch size is refused with VH_ERR_ARG."""
import sys

import numpy as np
import pytest

from oracle import topology
from tests import problems

pytestmark = pytest.mark.gpu
MU0 = 4 * np.pi * 1e-7


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def hollow(L, M, N, seed):
    geo = problems.bar(L, M, N)
    rng = np.random.default_rng(seed)
    mask = rng.random((L, M, N)) < 0.35
    mask[0] = mask[-1] = False
    geo['sigma'][mask] = 0.0
    geo['sigma'][1:-1, 0, 0] = 0.0            # some x lines without voxels
    return geo


def noalloc(G, jwmu, prec, mode, X, Y):
    from voxhenry_amd._lib import lib, ptr
    nrhs, ld = X.shape
    return lib().vh_matvec_noalloc(G.h, float(np.real(jwmu)), float(np.imag(jwmu)), prec, mode, nrhs, ptr(X), ld,
                                   ptr(Y), Y.shape[1])


def test_noalloc_matvec_matches_vh_matvec(ctx):
    import torch
    import voxhenry_amd as vh
    from voxhenry_amd._lib import check
    f = 2e9
    jwmu = 1j * 2 * np.pi * f * MU0
    # first a larger geometry with several vectors, then a smaller one
    for (dims, nrhs, seed) in (((30, 10, 10), 3, 1), ((9, 6, 5), 1, 2)):
        geo = hollow(*dims, seed)
        s = topology.build_system(geo)
        G = problems.device_geometry(ctx, geo, s, pad_mode=1, f=f)
        rng = np.random.default_rng(seed)
        for prec, dt in ((vh.VH_C128, torch.complex128), (vh.VH_C64, torch.complex64)):
            for mode in (0, 2):
                n = G.n if mode == 0 else G.num_curr
                Xh = rng.standard_normal((nrhs, n)) + 1j * rng.standard_normal((nrhs, n))
                X = torch.tensor(Xh, dtype=dt, device='cuda')
                ref = G.matvec(X, jwmu, prec, mode)          # vh_matvec, on_device = 1
                Y = torch.full_like(X, float('nan'))
                ctx.use_torch_stream()
                check(noalloc(G, jwmu, prec, mode, X, Y))
                torch.cuda.synchronize()
                say('noalloc', dims, nrhs, prec, mode, 'max |diff|', (Y - ref).abs().max().item())
                assert torch.equal(Y, ref)
        G.close()


def test_noalloc_refuses_beyond_scratch(ctx):
    import torch
    import voxhenry_amd as vh
    from voxhenry_amd._lib import lib
    geo = problems.bar(30, 10, 10)
    s = topology.build_system(geo)
    G = problems.device_geometry(ctx, geo, s, pad_mode=1, f=1e9)
    nrhs = 60      # Px, Py = 60, 20: 60 x 5 x (Px M N + Px Py N) x 16 B = 86.4 MB > 64 MiB
    X = torch.zeros((nrhs, G.n), dtype=torch.complex128, device='cuda')
    Y = torch.zeros_like(X)
    rc = noalloc(G, 1j, vh.VH_C128, 0, X, Y)
    assert rc == 1 and b'scratch' in lib().vh_last_error()
    G.close()
