"""The oracle's post-processing (post_obtain_curr_coefs_on_grid.m) pinned on
the reference's own solution: current conservation along the numex1 bar and
the port current of the executer's Y assembly."""
import os

import numpy as np

from oracle import post, topology
from tests import problems


def _numex1():
    geo = problems.numex1()
    d = np.load(os.path.join(problems.GOLDEN, 'numex1_x.npz'))
    return geo, d['x'], float(d['freq'])


def test_grid_carries_the_port_current_through_every_slice():
    """numex1 (60x20x20, ports on the x = 1 and x = 60 faces): the x-directed
    current coefficients of every x-slice sum to the same total, the port
    current sum(rhs .* x) of VoxHenry_executer.m:311-316, and to the
    admittance 1/Z of the reference's golden R/L table at 2.5 GHz.  A grid
    filled in row-major order, a swapped component block or a shifted voxel
    order breaks the equality slice by slice."""
    geo, x, f = _numex1()
    s = topology.build_system(geo)
    Mc = post.susceptibility(geo['sigma'], f)
    grids, cmin, cmax = post.curr_coefs_on_grid(x, Mc)
    L, M, N = geo['sigma'].shape
    assert all(g.shape == (L, M, N) for g in grids)
    Jx = grids[0]
    slices = Jx.sum(axis=(1, 2))
    Iport = s['rhs'][:, 0] @ x[:len(s['rhs'])]
    assert np.max(np.abs(slices - Iport)) / abs(Iport) < 1e-9
    # the other directions carry no net current through an x-slice
    assert np.max(np.abs(grids[1].sum(axis=(1, 2)))) < 1e-6 * abs(Iport)
    # against the golden table: Z = R + j*omega*L of the 1x1 port
    rl = np.load(os.path.join(problems.GOLDEN, 'numex_RL.npz'))
    k = int(np.argmin(np.abs(rl['freq1'] - f)))
    Z = rl['RL1'][k].real + 1j * 2 * np.pi * f * rl['RL1'][k].imag
    assert abs(Iport - 1 / Z) / abs(1 / Z) < 1e-6
    jxyz = np.abs(x[:3 * s['nv']])
    assert cmin == jxyz.min() and cmax == jxyz.max()


def test_grid_places_hand_worked_voxels():
    """3 x 2 x 2 grid with voxels (2,1,1) and (3,2,2) (1-based) empty: the
    non-empty voxels in column-major order are (1,1,1), (3,1,1), (1,2,1), ...,
    so x entry v of block c lands at those positions."""
    sigma = np.ones((3, 2, 2))
    sigma[1, 0, 0] = 0.0
    sigma[2, 1, 1] = 0.0
    Mc = post.susceptibility(sigma * 5.8e7, 1e9)
    order = [(0, 0, 0), (2, 0, 0), (0, 1, 0), (1, 1, 0), (2, 1, 0), (0, 0, 1), (1, 0, 1), (2, 0, 1),
             (0, 1, 1), (1, 1, 1)]
    nv = len(order)
    x = np.arange(5 * nv + 7, dtype=float) + 1.0           # node potentials after the currents are ignored
    grids, cmin, cmax = post.curr_coefs_on_grid(x, Mc)
    for c in range(5):
        g = grids[c]
        for v, (i, j, k) in enumerate(order):
            assert g[i, j, k] == x[c * nv + v]
        assert g[1, 0, 0] == 0 and g[2, 1, 1] == 0
    assert cmin == 1.0 and cmax == 3 * nv
