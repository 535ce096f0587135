"""Write the small golden fixtures under tests/golden/ from the reference's own
committed result files (data only; nothing of the reference is executed).

Run in the build container, where the reference is mounted:
    python tests/golden/make_fixtures.py /root/reference

Outputs:
  numex_RL.npz        -- R_jL_mat tables of results_numex{1..4}*/data_R_jL_mat.mat
  numex_timings.npz   -- sim_CPU_lse / sim_CPU_pre of results_numex*/data_CPU_timings.mat
  numex1_x.npz        -- the FP64 solution x of results_numex1_straight_conductor/
                         data_curr_plot.mat (60x20x20 bar @0.5um, 2.5 GHz, port 1)
  numex1_geometry.npz -- the numex1 .vhr reduced to (dims, dx, sigma, freq, ports);
                         the grid is full, so the voxel list is stored as a check sum
  numex3_geometry.npz -- the numex3 circular coil .vhr (circular_coil_loop150.0u_rad5.0u,
                         310x310x10 @1um) reduced to (dims, dx, sigma, the 0-based
                         column-major indices of its voxels, ports); its 49-value
                         frequency list is the one of VoxHenry_generator_numex3_circular_coil.m:39-44
K_near_interactions_1mm.txt and wire_len2.0u_dia4.0u.vhr are plain copies of
reference data files (the near-field table and the smallest input).
"""
import os
import sys
import numpy as np
import scipy.io as sio

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    ex = ['results_numex1_straight_conductor', 'results_numex2_wire',
          'results_numex3_circular_coil', 'results_numex4_square_coil']
    rl, tm = {}, {}
    for i, e in enumerate(ex, 1):
        d = sio.loadmat(os.path.join(ref, e, 'data_R_jL_mat.mat'))
        rl['freq%d' % i] = d['freq_all'].ravel()
        rl['RL%d' % i] = d['R_jL_mat'].ravel()
        t = sio.loadmat(os.path.join(ref, e, 'data_CPU_timings.mat'))
        tm['lse%d' % i] = t['sim_CPU_lse']
        tm['pre%d' % i] = t['sim_CPU_pre'].ravel()
    np.savez_compressed(os.path.join(HERE, 'numex_RL.npz'), **rl)
    np.savez_compressed(os.path.join(HERE, 'numex_timings.npz'), **tm)
    d = sio.loadmat(os.path.join(ref, ex[0], 'data_curr_plot.mat'))
    np.savez_compressed(os.path.join(HERE, 'numex1_x.npz'), x=d['x'].ravel(), dx=float(d['dx']),
                        freq=2.5e9)
    sys.path.insert(0, os.path.join(HERE, '..', '..'))
    from oracle import vhr
    g = vhr.read_vhr(os.path.join(ref, 'Input_files', 'straight_cond1_len30.0u_wid10.0u_dist20.0u.vhr'))
    s = g['sigma']
    assert (s == 5.8e7).all(), 'numex1 grid expected full'
    np.savez_compressed(os.path.join(HERE, 'numex1_geometry.npz'), dims=np.array(s.shape), dx=g['dx'],
                        sigma=5.8e7, freq=g['freq'], P=g['ports_P'][0], N=g['ports_N'][0])
    g = vhr.read_vhr(os.path.join(ref, 'Input_files', 'circular_coil_loop150.0u_rad5.0u.vhr'))
    s = g['sigma']
    vals = np.unique(s[s != 0])
    assert len(vals) == 1 and g['lambdaL'] is None and len(g['ports_P']) == 1
    idx = np.flatnonzero(s.ravel(order='F') != 0).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, 'numex3_geometry.npz'), dims=np.array(s.shape), dx=g['dx'],
                        sigma=float(vals[0]), idx=idx, freq=g['freq'], P=g['ports_P'][0], N=g['ports_N'][0],
                        well=g['well_cond'])


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
