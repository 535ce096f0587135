"""Pins of the oracle topology (panel IDs, Ae, ports, rhs) and of the full
solve against the reference's committed golden data."""
import os

import numpy as np
import pytest

from oracle import topology, driver, vhr, matvec
from tests import problems


def test_panel_counts_full_bar():
    L, M, N = 4, 3, 2
    geo = problems.bar(L, M, N)
    s = topology.build_system(geo)
    # full bar: (L+1)MN x-panels, L(M+1)N y-panels, LM(N+1) z-panels
    assert s['Ae'].shape == ((L + 1) * M * N + L * (M + 1) * N + L * M * (N + 1), 5 * L * M * N)
    assert s['Ae'].nnz == 16 * L * M * N


def test_ae_column_structure():
    geo = problems.wire_vhr()
    s = topology.build_system(geo)
    Ae = s['Ae'].tocsc()
    nv = s['nv']
    sums = np.asarray(Ae.sum(axis=0)).ravel()
    # Jx,Jy,Jz columns: -1 on the entering panel, +1 on the leaving one -> sum 0
    assert np.all(sums[:3 * nv] == 0)
    # J2d: +1/2 +1/2 -1/2 -1/2 ; J3d: 4*(+1/2) - 2*1
    assert np.all(sums[3 * nv:] == 0)
    assert np.all(np.diff(Ae.indptr) == np.repeat([2, 2, 2, 4, 6], nv))
    # every node touches at least one voxel face
    assert np.all(np.diff(Ae.tocsr().indptr) > 0)


def test_panel_sharing_between_neighbours():
    """Neighbouring voxels share the panel between them (lse_generate_panelIDs.m)."""
    geo = problems.wire_vhr()
    s = topology.build_system(geo)
    ids, idxS = s['ids'], s['idxS']
    L, M, N = s['dims']
    pos = {int(v): k for k, v in enumerate(idxS)}
    for k, v in enumerate(idxS):
        x, y, z = np.unravel_index(v, (L, M, N), order='F')
        for ax, (a, b) in enumerate([(0, 1), (2, 3), (4, 5)]):
            nb = [x, y, z]
            nb[ax] += 1
            if nb[ax] < (L, M, N)[ax]:
                w = nb[0] + L * (nb[1] + M * nb[2])
                if w in pos:
                    assert ids[k, b] == ids[pos[w], a]


def test_wire_input_file_and_ports():
    geo = problems.wire_vhr()
    assert geo['sigma'].shape == (2, 4, 4)
    assert (geo['sigma'] > 0).sum() == 24
    assert len(geo['freq']) == 41
    s = topology.build_system(geo)
    assert len(s['inject']) == 12 and len(s['grnd']) == 12
    assert s['Ar'].shape[0] == s['Ae'].shape[0] - 24
    # rhs = -sum of the excitation rows of Ae: only Jx / J2d / J3d of the x=2 voxels
    assert np.count_nonzero(s['rhs'][:, 0]) == 3 * 12


@pytest.mark.slow
def test_numex1_golden_solution_residual():
    geo = problems.numex1()
    prob = driver.Problem(geo)
    s = prob.sys
    d = np.load(os.path.join(problems.GOLDEN, 'numex1_x.npz'))
    nc = s['num_curr']
    xr = np.concatenate([d['x'][:nc], d['x'][nc:][s['keep']]])
    assert np.all(d['x'][nc:][~s['keep']] == 0)       # port potentials are zero in the golden x
    f = float(d['freq'])
    A, P, zr, zi, st = prob.operators(f)
    b = s['rhs'][:, 0]
    relres = np.linalg.norm(b - A(xr)) / np.linalg.norm(b)
    assert 1e-10 < relres < 3e-9
    # Y from the golden x reproduces the committed R, L at 2.5 GHz
    Z, R, Lh = driver.z_to_rl(np.array([[b @ xr]]), f)
    rl = np.load(os.path.join(problems.GOLDEN, 'numex_RL.npz'))
    k = np.argmin(np.abs(rl['freq1'] - f))
    assert abs(R[0, 0] - rl['RL1'][k].real) / rl['RL1'][k].real < 1e-13
    assert abs(Lh[0, 0] - rl['RL1'][k].imag) / rl['RL1'][k].imag < 1e-13


@pytest.mark.slow
def test_numex1_RL_table():
    geo = problems.numex1()
    prob = driver.Problem(geo)
    rl = np.load(os.path.join(problems.GOLDEN, 'numex_RL.npz'))
    for k in (20, 40):     # 1e5 Hz (tol 1e-12) and 1e10 Hz (tol 1e-8)
        f = rl['freq1'][k]
        out = driver.solve_sweep(prob, [f])
        R, Lh = out['R'][0, 0, 0], out['L'][0, 0, 0]
        assert abs(R - rl['RL1'][k].real) / rl['RL1'][k].real < 1e-8, (f, R)
        assert abs(Lh - rl['RL1'][k].imag) / rl['RL1'][k].imag < 1e-8, (f, Lh)


def test_dc_resistance_closed_form():
    """Full-face ports on a bar: R(DC) = l / (sigma A) (numex1 golden: 5.1724137931e-3)."""
    L, M, N = 12, 3, 3
    dx = 1e-6
    geo = problems.bar(L, M, N, dx=dx)
    prob = driver.Problem(geo)
    out = driver.solve_sweep(prob, [1.0])
    R = out['R'][0, 0, 0]
    ref = (L * dx) / (5.8e7 * (M * dx) * (N * dx))
    assert abs(R - ref) / ref < 1e-9
