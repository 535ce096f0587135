"""FP64 restatement of src_lse_formation/lse_matvect_mult.m:1-153 (default
``schur_invert`` layout: the identity rows of :146-153 are not used).

Pinned by (tests/test_oracle_matvec.py, tests/test_oracle_solve.py):
* brute-force direct convolution sum_n G(m-n) J(n) on tiny grids;
* complex symmetry of the dense operator built by unit-vector probing
  (the method of lse_generate_Zmat.m:22-60);
* the numex1 golden solution x (results_numex1_straight_conductor/
  data_curr_plot.mat, 2.5 GHz) gives ||b - A x|| / ||b|| ~ 1e-9.
"""
import numpy as np

# FFT threads of the oracle: 1 uses numpy.fft; more use scipy.fft's pocketfft
# with that many workers (the CPU baseline's multi-core timing).
FFT_WORKERS = 1


def _fftn(a, s=None, axes=None):
    if FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.fftn(a, s=s, axes=axes, workers=FFT_WORKERS)
    return np.fft.fftn(a, s=s, axes=axes)


def _ifftn(a):
    if FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.ifftn(a, workers=FFT_WORKERS)
    return np.fft.ifftn(a)

# lse_matvect_mult.m:99-103 -- diagonal weights per component, times 1/dx
DIAG_W = np.array([1.0, 1.0, 1.0, 1.0 / 6.0, 1.0 / 2.0])


def spectral_apply(fJ, fN):
    """The 5x5 block product of lse_matvect_mult.m:57-86 on spectra
    fJ (5,...) with fN (...,7).  Returns (5,...)."""
    G = [fN[..., k] for k in range(7)]
    o1 = G[0] * fJ[0] + G[1] * fJ[3] + G[1] * fJ[4]
    o2 = G[0] * fJ[1] + G[2] * fJ[3] - G[2] * fJ[4]
    o3 = G[0] * fJ[2] + G[4] * fJ[4]
    o4 = -G[1] * fJ[0] - G[2] * fJ[1] + G[3] * fJ[3] + G[5] * fJ[4]
    o5 = -G[1] * fJ[0] + G[2] * fJ[1] - G[4] * fJ[2] + G[5] * fJ[3] + G[6] * fJ[4]
    return np.stack([o1, o2, o3, o4, o5])


def fft_block(Jc, fN_all, dims):
    """Convolution part only: crop(ifftn(fN * fftn(pad(J)))) on the (L,M,N,5) grid."""
    L, M, N = dims
    P = fN_all.shape[:3]
    fJ = np.stack([_fftn(Jc[..., c], s=P, axes=(0, 1, 2)) for c in range(5)])
    out = spectral_apply(fJ, fN_all)
    return np.stack([_ifftn(out[c])[:L, :M, :N] for c in range(5)], axis=-1)


def lse_matvect_mult(JIn0, fN_all, Ae, z_real, z_imag, dx, idx5):
    """y = A x for the full system [Z, -Ae'; Ae, 0] (lse_matvect_mult.m)."""
    num_node, num_curr = Ae.shape
    L, M, N = z_real.shape
    V = L * M * N
    JIn0 = np.asarray(JIn0).ravel()
    Jfull = np.zeros(5 * V, dtype=complex)
    Jfull[idx5] = JIn0[:num_curr]                                       # :38
    Jc = Jfull.reshape((L, M, N, 5), order='F')
    conv = fft_block(Jc, fN_all, (L, M, N)).reshape(-1, order='F')[idx5]   # :57-93, :123
    vox = idx5[:num_curr // 5]
    z = z_real.ravel(order='F')[vox].astype(complex)
    if z_imag is not None:
        z = z + z_imag.ravel(order='F')[vox]
    w = np.repeat(DIAG_W / dx, num_curr // 5)
    out = np.zeros(num_curr + num_node, dtype=complex)
    out[:num_curr] = w * np.tile(z, 5) * JIn0[:num_curr] + conv          # :98-110
    out[:num_curr] -= Ae.T @ JIn0[num_curr:num_curr + num_node]         # :138
    out[num_curr:] = Ae @ JIn0[:num_curr]                               # :140
    return out
