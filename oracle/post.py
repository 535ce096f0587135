"""Post-processing restated from the reference (test infrastructure: the
checker of the GPU path, never called by the product).

curr_coefs_on_grid restates src_post_process/post_obtain_curr_coefs_on_grid.m:1-45.
It is pinned in tests/test_oracle_post.py by current conservation on the
reference's own solution (numex1, 2.5 GHz): every x-slice of the Jx grid
carries the same total current, equal to the port admittance of the
reference's golden R/L table.  A transposed grid, a swapped component or a
shifted voxel order breaks that.
"""
import numpy as np

TOLA = 1e-12     # post_obtain_curr_coefs_on_grid.m:5


def empty_voxels(Mc):
    """ignore_unk: |Mc| < tola in column-major order (:7-17)."""
    return (np.abs(np.asarray(Mc)) < TOLA).ravel(order='F')


def curr_coefs_on_grid(x, Mc):
    """(Jx, Jy, Jz, J2d, J3d grids of shape (L, M, N), cmin, cmax), :1-45.

    x: solution vector, its first 5*nv entries the voxel currents in the
    column-major order of the non-empty voxels; Mc: susceptibility grid
    (empty where |Mc| < 1e-12).
    """
    Mc = np.asarray(Mc)
    L, M, N = Mc.shape
    keep = ~empty_voxels(Mc)
    nv = int(keep.sum())                                          # num_cube_new (:19-20)
    x = np.asarray(x)
    blocks = [x[c * nv:(c + 1) * nv] for c in range(5)]           # :22-23
    jxyz = np.concatenate(blocks[:3])
    cmin = float(np.min(np.abs(jxyz)))                            # :25
    cmax = float(np.max(np.abs(jxyz)))                            # :26
    grids = []
    for b in blocks:                                              # :28-45
        flat = np.zeros(L * M * N, dtype=np.result_type(b, np.float64))
        flat[keep] = b
        grids.append(flat.reshape((L, M, N), order='F'))
    return grids, cmin, cmax


def susceptibility(sigma, freq):
    """Mc = Mr - 1, Mr = epsilon_r - j*sigma/(eo*omega) with epsilon_r = 1
    (pre_define_structure_params.m:23) and eo = 1/(co^2 mu) (EMconstants.m:13-15);
    VoxHenry_executer.m:347-349."""
    mu = 4 * np.pi * 1e-7
    eo = 1.0 / (299792458.0 ** 2 * mu)
    omega = 2 * np.pi * freq
    return (1.0 - 1j * np.asarray(sigma, dtype=float) / (eo * omega)) - 1.0
