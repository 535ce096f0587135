"""Result files in the reference's .mat layout (VoxHenry_executer.m:351-365):
variable names, shapes and values, checked against the layout of the
reference's committed numex1 results (R_jL_mat 1 x 1 x 41 complex, freq_all
1 x 41) with its golden table as the data."""
import os

import numpy as np
import scipy.io as sio

from tests import problems
from voxhenry_amd import results


def test_r_jl_file_layout_and_values(tmp_path):
    rl = np.load(os.path.join(problems.GOLDEN, 'numex_RL.npz'))
    f, RL = rl['freq1'], rl['RL1']
    out = {'freq': f, 'R': RL.real.reshape(-1, 1, 1), 'L': RL.imag.reshape(-1, 1, 1)}
    path = results.save_r_jl(str(tmp_path / 'numex1'), out)
    assert path.endswith('numex1-data_R_jL_mat.mat')
    d = sio.loadmat(path)
    assert d['R_jL_mat'].shape == (1, 1, 41) and np.iscomplexobj(d['R_jL_mat'])
    assert d['freq_all'].shape == (1, 41)
    assert int(d['num_freq'][0, 0]) == 41 and int(d['num_ports'][0, 0]) == 1
    assert np.array_equal(d['R_jL_mat'].ravel(), RL)
    assert np.array_equal(d['freq_all'].ravel(), f)


def test_r_jl_port_order():
    """R_jL_mat(kk, port, f): row = measured port, column = driven port."""
    R = np.arange(2 * 3 * 3, dtype=float).reshape(2, 3, 3)
    out = {'freq': [1e6, 1e9], 'R': R, 'L': -R}
    m = results.r_jl_mat(out)
    assert m.shape == (3, 3, 2)
    assert m[1, 2, 0] == R[0, 1, 2] - 1j * R[0, 1, 2]
    assert m[2, 0, 1] == R[1, 2, 0] - 1j * R[1, 2, 0]


def test_voxel_centres():
    r = results.voxel_centres((3, 2, 4), 1e-6)
    assert r.shape == (3, 2, 4, 3)
    assert np.allclose(r[0, 0, 0], [0.5e-6] * 3, rtol=0, atol=1e-20)
    assert np.allclose(r[2, 1, 3], [2.5e-6, 1.5e-6, 3.5e-6], rtol=0, atol=1e-20)
