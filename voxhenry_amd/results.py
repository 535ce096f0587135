"""Result files in the reference's MATLAB layout (VoxHenry_executer.m:351-365).

save_r_jl writes <prefix>-data_R_jL_mat.mat: num_freq, num_ports, freq_all
(1 x nfreq) and R_jL_mat (nports x nports x nfreq complex, R + j L with
R = |Re Z|, L = |Im Z| / (2 pi f), :325-326).

save_currents writes <prefix>-data_curr_plot.mat for one solution x: x
(n x 1), Mc (L x M x N, the susceptibility the plots use to find the empty
voxels, :347-349), dx, plot_option and r (L x M x N x 3 voxel centres,
generategridfrombbox.m), plus the current coefficients on the grid computed
on the GPU (post_obtain_curr_coefs_on_grid.m:1-45): Jx_currs_grid ..
J3d_currs_grid, cmin, cmax.  The node-current plot inputs Ae_only_leaving /
Ae_only_entering_bndry are not written.
"""
import numpy as np
import scipy.io as sio

MU0 = 4 * np.pi * 1e-7
EO = 1.0 / (299792458.0 ** 2 * MU0)      # EMconstants.m:13-15
GRID_NAMES = ('Jx_currs_grid', 'Jy_currs_grid', 'Jz_currs_grid', 'J2d_currs_grid', 'J3d_currs_grid')


def r_jl_mat(out):
    """R_jL_mat (nports x nports x nfreq) from a sweep result (R, L of shape nfreq x nports x nports)."""
    R = np.asarray(out['R'], dtype=float)
    L = np.asarray(out['L'], dtype=float)
    return np.ascontiguousarray(np.transpose(R + 1j * L, (1, 2, 0)))


def save_r_jl(prefix, out):
    freq = np.asarray(out['freq'], dtype=float)
    rjl = r_jl_mat(out)
    path = prefix + '-data_R_jL_mat.mat'
    sio.savemat(path, {'num_freq': float(len(freq)), 'num_ports': float(rjl.shape[0]),
                       'freq_all': freq.reshape(1, -1), 'R_jL_mat': rjl})
    return path


def susceptibility(sigma, freq):
    """Mc = (epsilon_r - j sigma/(eo omega)) - 1 with epsilon_r = 1 (pre_define_structure_params.m:23)."""
    return (1.0 - 1j * np.asarray(sigma, dtype=float) / (EO * 2 * np.pi * freq)) - 1.0


def voxel_centres(dims, dx):
    """r(mx, my, mz, :) = ((mx - 1/2) dx, (my - 1/2) dx, (mz - 1/2) dx), generategridfrombbox.m."""
    L, M, N = dims
    ax = [(np.arange(n) + 0.5) * dx for n in (L, M, N)]
    X, Y, Z = np.meshgrid(*ax, indexing='ij')
    return np.stack([X, Y, Z], axis=3)


def save_currents(prefix, solver, x, freq, plot_option=1):
    """Write <prefix>-data_curr_plot.mat for solution x of solver (a solver.Solver) at freq."""
    G = solver.G
    x = np.asarray(x, dtype=np.complex128)
    grids, cmin, cmax = G.currents_on_grid(x)
    sigma = solver.structure['sigma']
    d = {'x': x.reshape(-1, 1), 'Mc': susceptibility(sigma, freq), 'dx': float(G.dx),
         'plot_option': float(plot_option), 'r': voxel_centres(G.dims, G.dx), 'cmin': float(cmin), 'cmax': float(cmax)}
    for name, g in zip(GRID_NAMES, grids):
        d[name] = np.asfortranarray(g)
    path = prefix + '-data_curr_plot.mat'
    sio.savemat(path, d)
    return path
