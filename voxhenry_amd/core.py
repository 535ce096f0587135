"""Object layer over the C ABI: a device context and a device-resident geometry
(grid, voxel map, Green's tensor spectrum, impedance, incidence matrix).

Host vectors are numpy arrays (interleaved complex); device vectors are torch
tensors on the HIP device (torch provides device memory and streams only).
"""
import ctypes as C

import numpy as np

from ._lib import lib, check, ptr, VH_C128, VH_C64, VH_PC_SCHUR_INVERT, VH_PC_SCHUR_GMRES

MU0 = 4 * np.pi * 1e-7        # EMconstants.m:13
# solve_batch methods (the method argument of vh_solve_batch)
METHODS = {'fp64': 0, 'mixed': 1, 'fp64_true': 2}


def _torch():
    import torch
    return torch


class Context:
    """One HIP device context (include/voxhenry.h: vh_ctx_*)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().vh_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def use_torch_stream(self):
        """Launch on torch's current HIP stream so torch events time our kernels."""
        torch = _torch()
        s = torch.cuda.current_stream(self.device)
        check(lib().vh_ctx_set_stream(self.h, C.c_void_p(s.cuda_stream)))

    def use_stream(self, stream):
        """Launch on the given torch stream (one stream per context lets the
        geometries of several contexts run concurrently)."""
        check(lib().vh_ctx_set_stream(self.h, C.c_void_p(stream.cuda_stream)))

    def synchronize(self):
        check(lib().vh_ctx_synchronize(self.h))

    def close(self):
        if self.h:
            check(lib().vh_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Geometry:
    """Device geometry (vh_geom_*).

    idx: 0-based column-major linear indices of the non-empty voxels (idxS).
    pad_mode 0 = (2L,2M,2N) like the reference; 1 = smallest smooth >= 2n-1.
    """

    def __init__(self, ctx, dims, dx, idx, pad_mode=1):
        self.ctx = ctx
        self.dims = tuple(int(d) for d in dims)
        self.dx = float(dx)
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        h = C.c_void_p()
        check(lib().vh_geom_create(ctx.h, *self.dims, self.dx, len(idx), ptr(idx), int(pad_mode), C.byref(h)))
        self.h = h
        self.st2 = None
        self._refresh()

    def _refresh(self):
        info = np.zeros(10, dtype=np.int64)
        check(lib().vh_geom_info(self.h, ptr(info)))
        self.nv, self.num_curr, self.num_node = int(info[0]), int(info[1]), int(info[2])
        self.pad = (int(info[3]), int(info[4]), int(info[5]))
        self.nlines = int(info[6])

    @property
    def n(self):
        return self.num_curr + self.num_node

    # -- tensor -------------------------------------------------------------
    def generate_green(self):
        st = np.zeros(6)
        check(lib().vh_geom_generate_green(self.h, ptr(st)))
        self.st2 = st[0::2] + 1j * st[1::2]
        return self.st2

    def set_green_spectrum(self, fN, scale=1.0, st2=None):
        fN = np.asfortranarray(fN, dtype=np.complex128)
        P = fN.shape[:3]
        st = None if st2 is None else np.ascontiguousarray(np.stack([np.real(st2), np.imag(st2)], 1).ravel())
        flat = np.ascontiguousarray(fN.ravel(order='F'))
        check(lib().vh_geom_set_green_spectrum(self.h, ptr(flat), P[0], P[1], P[2], float(np.real(scale)),
                                               float(np.imag(scale)), ptr(st)))
        if st2 is not None:
            self.st2 = np.asarray(st2)

    def set_green_realspace(self, G, st2=None):
        flat = np.ascontiguousarray(np.asarray(G, dtype=np.complex128).ravel(order='F'))
        st = None if st2 is None else np.ascontiguousarray(np.stack([np.real(st2), np.imag(st2)], 1).ravel())
        check(lib().vh_geom_set_green_realspace(self.h, ptr(flat), ptr(st)))
        if st2 is not None:
            self.st2 = np.asarray(st2)

    def get_green_realspace(self):
        L, M, N = self.dims
        out = np.zeros(L * M * N * 7, dtype=np.complex128)
        check(lib().vh_geom_get_green_realspace(self.h, ptr(out)))
        return out.reshape((L, M, N, 7), order='F')

    # -- operator data --------------------------------------------------------
    def set_incidence(self, Ar):
        """Ar: scipy.sparse matrix num_node x num_curr (port rows removed)."""
        A = Ar.tocsc()
        A.sort_indices()
        cp = np.ascontiguousarray(A.indptr, dtype=np.int64)
        ri = np.ascontiguousarray(A.indices, dtype=np.int64)
        v = np.ascontiguousarray(A.data, dtype=np.float64)
        check(lib().vh_geom_set_incidence(self.h, A.shape[0], A.nnz, ptr(cp), ptr(ri), ptr(v)))
        self._refresh()

    def build_topology(self, ports_P, ports_N, well=None):
        """Panel ids, Ar and port right-hand sides built on the host by the library.

        ports_P / ports_N: per port an int array (k,4) of 1-based (x, y, z, side).
        Returns rhs (nports x n, real)."""
        def pack(lst):
            ptrs = np.zeros(len(lst) + 1, dtype=np.int64)
            rows = []
            for i, a in enumerate(lst):
                a = np.asarray(a, dtype=np.int64).reshape(-1, 4)
                ptrs[i + 1] = ptrs[i] + len(a)
                rows.append(a)
            flat = np.ascontiguousarray(np.concatenate(rows) if rows else np.zeros((0, 4), np.int64), dtype=np.int64)
            return ptrs, flat
        pp, P = pack(ports_P)
        pn, N = pack(ports_N)
        w = np.ascontiguousarray(np.zeros((0, 4), np.int64) if well is None else np.asarray(well, np.int64).reshape(-1, 4))
        nn = np.zeros(1, dtype=np.int64)
        check(lib().vh_geom_build_topology(self.h, len(ports_P), ptr(pp), ptr(P), ptr(pn), ptr(N), len(w),
                                           ptr(w) if len(w) else None, ptr(nn)))
        self._refresh()
        rhs = np.zeros((len(ports_P), self.n))
        if len(ports_P):
            check(lib().vh_geom_get_rhs(self.h, ptr(rhs)))
        self.rhs = rhs
        return rhs

    def set_impedance(self, z):
        """z = z_real + z_imag per non-empty voxel (complex, nv)."""
        z = np.ascontiguousarray(np.asarray(z, dtype=np.complex128))
        assert z.shape == (self.nv,)
        check(lib().vh_geom_set_impedance(self.h, ptr(z)))

    # -- post-processing --------------------------------------------------------
    def currents_on_grid(self, x):
        """Current coefficients on the voxel grid, post_obtain_curr_coefs_on_grid.m:1-45,
        computed on the GPU.

        x: solution(s) of shape (n,) or (nrhs, n) (only the first 5*nv entries,
        Jx, Jy, Jz, J2d, J3d per voxel, are read), numpy complex128.  Returns
        (grids, cmin, cmax): grids has shape (5, L, M, N) (or (nrhs, 5, L, M, N)),
        component c of voxel (i, j, k) at grids[c, i, j, k], zero at empty voxels;
        cmin / cmax = min / max of |J| over Jx, Jy, Jz (:24-25).
        """
        xx = np.ascontiguousarray(np.atleast_2d(x), dtype=np.complex128)
        nrhs, ld = xx.shape
        L, M, N = self.dims
        ncell = L * M * N
        g = np.zeros((nrhs, 5 * ncell), dtype=np.complex128)
        mm = np.zeros(2 * nrhs)
        check(lib().vh_geom_currents_on_grid(self.h, nrhs, ptr(xx), ld, ptr(g), 5 * ncell, ptr(mm), 0))
        grids = g.reshape(nrhs, 5, N, M, L).transpose(0, 1, 4, 3, 2)     # column-major cells -> [x][y][z]
        if np.ndim(x) == 1:
            return grids[0], mm[0], mm[1]
        return grids, mm[0::2], mm[1::2]

    # -- matvec ---------------------------------------------------------------
    def matvec(self, x, jwmu, prec=VH_C128, mode=0):
        """y = A x for a host (numpy) or device (torch) vector/stack of vectors.

        x: shape (n,) or (nrhs, n); complex128 for VH_C128, complex64 for VH_C64.
        """
        n = self.n if mode == 0 else self.num_curr
        if isinstance(x, np.ndarray):
            dt = np.complex128 if prec == VH_C128 else np.complex64
            xx = np.ascontiguousarray(np.atleast_2d(x), dtype=dt)
            nrhs, ld = xx.shape
            y = np.zeros_like(xx)
            check(lib().vh_matvec(self.h, float(np.real(jwmu)), float(np.imag(jwmu)), prec, mode, nrhs, ptr(xx), ld,
                                  ptr(y), ld, 0))
            return y.reshape(np.shape(x))
        torch = _torch()
        xx = x if x.dim() == 2 else x.unsqueeze(0)
        assert xx.is_contiguous()
        nrhs, ld = xx.shape
        y = torch.zeros_like(xx)
        self.ctx.use_torch_stream()
        check(lib().vh_matvec(self.h, float(np.real(jwmu)), float(np.imag(jwmu)), prec, mode, nrhs, ptr(xx), ld,
                              ptr(y), ld, 1))
        return y.reshape(x.shape)

    def matvec_timed(self, x, y, jwmu, prec, mode, reps):
        """Average ms of one matvec call on device-resident torch tensors."""
        xx = x if x.dim() == 2 else x.unsqueeze(0)
        yy = y if y.dim() == 2 else y.unsqueeze(0)
        nrhs, ld = xx.shape
        ms = np.zeros(1)
        self.ctx.use_torch_stream()
        check(lib().vh_matvec_timed(self.h, float(np.real(jwmu)), float(np.imag(jwmu)), prec, mode, nrhs, ptr(xx), ld,
                                    ptr(yy), yy.shape[1], reps, ptr(ms), None, None))
        return float(ms[0])

    def matvec_timed_passes(self, x, y, jwmu, prec, mode, reps):
        """(ms per call, [ms per pass X, Y, Z, Yi, Xi, node]) from HIP events on the launch stream."""
        xx = x if x.dim() == 2 else x.unsqueeze(0)
        yy = y if y.dim() == 2 else y.unsqueeze(0)
        nrhs, ld = xx.shape
        ms = np.zeros(1)
        mp = np.zeros(6)
        npass = np.zeros(1, np.int32)
        self.ctx.use_torch_stream()
        check(lib().vh_matvec_timed(self.h, float(np.real(jwmu)), float(np.imag(jwmu)), prec, mode, nrhs, ptr(xx), ld,
                                    ptr(yy), yy.shape[1], reps, ptr(ms), ptr(mp), ptr(npass)))
        return float(ms[0]), [float(v) for v in mp]

    # -- solvers ------------------------------------------------------------
    def gmres(self, b, jwmu, precon=None, restart=100, tol=1e-8, maxit=10, prec=VH_C128, flexible=False,
              resvec_cap=0):
        """pgmres (left precon) / fpgmres (flexible) for one or more rhs sharing the operator.

        b: numpy (n,) or (nrhs, n).  Returns x and a dict of per-rhs statistics."""
        dt = np.complex128 if prec == VH_C128 else np.complex64
        bb = np.ascontiguousarray(np.atleast_2d(b), dtype=dt)
        nrhs, n = bb.shape
        assert n == self.n
        x = np.zeros_like(bb)
        flag = np.zeros(nrhs, np.int32)
        relres = np.zeros(nrhs)
        it = np.zeros(2 * nrhs, np.int32)
        nit = np.zeros(nrhs, np.int32)
        nres = np.zeros(nrhs, np.int32)
        cap = int(resvec_cap)
        rv = np.zeros(max(1, nrhs * cap))
        check(lib().vh_gmres(self.h, precon.h if precon is not None else None, float(np.real(jwmu)),
                             float(np.imag(jwmu)), prec, int(flexible), nrhs, ptr(bb), ptr(x), 0, restart, tol, maxit,
                             ptr(flag), ptr(relres), ptr(it), ptr(nit), ptr(rv) if cap else None, cap, ptr(nres)))
        stats = dict(flag=flag, relres=relres, iter=it.reshape(nrhs, 2), niters=nit, nresvec=nres)
        if cap:
            stats['resvec'] = [rv[i * cap:i * cap + min(cap, nres[i])] for i in range(nrhs)]
        return x.reshape(np.shape(b)), stats

    def solve_mixed(self, b, jwmu, precon, tol=1e-8, outer_restart=8, outer_maxit=4, inner_restart=40,
                    inner_tol=1e-5):
        bb = np.ascontiguousarray(np.atleast_2d(b), dtype=np.complex128)
        nrhs, n = bb.shape
        x = np.zeros_like(bb)
        flag = np.zeros(nrhs, np.int32)
        relres = np.zeros(nrhs)
        oit = np.zeros(nrhs, np.int32)
        iit = np.zeros(nrhs, np.int32)
        check(lib().vh_solve_mixed(self.h, precon.h, float(np.real(jwmu)), float(np.imag(jwmu)), nrhs, ptr(bb),
                                   ptr(x), 0, tol, outer_restart, outer_maxit, inner_restart, inner_tol, ptr(flag),
                                   ptr(relres), ptr(oit), ptr(iit)))
        return x.reshape(np.shape(b)), dict(flag=flag, relres=relres, outer_iters=oit, inner_iters=iit)

    def solve_batch(self, B, jwmu, precons, pidx, method='mixed', tol=1e-8, restart=None, maxit=None,
                    inner_restart=30, inner_tol=1e-5, hbm_budget=0.0, x0=None):
        """Many (port, frequency) systems in one batched solve (vh_solve_batch).

        B: (nsys, n) rhs, a numpy array (host) or a complex128 torch tensor on
        the device (then X is a device tensor too); jwmu: (nsys,) complex;
        precons: list of Precon; pidx: (nsys,) index into precons.  method
        'fp64' (pgmres semantics, default restart 100 / maxit 10), 'fp64_true'
        (the same, then the systems it leaves above tol -- it stops on the
        preconditioned residual -- continue in FP64 flexible GMRES until the
        true residual meets tol) or 'mixed' (FP32-GMRES iterative refinement
        on the FP64 residual, up to restart * maxit steps, default 8 x 4).
        tol: one tolerance, or one per
        system (e.g. reference_tol of each system's frequency).  x0 (mixed
        only): initial guesses of the shape of B; the refinement starts
        from B - A x0 (default: zero, as pgmres.m)."""
        m = METHODS[method]
        if restart is None:
            restart = 8 if m == 1 else 100
        if maxit is None:
            maxit = 4 if m == 1 else 10
        on_dev = not isinstance(B, np.ndarray)
        if on_dev:
            torch = _torch()
            assert B.dtype == torch.complex128 and B.is_cuda and B.dim() == 2
            BB = B.contiguous()
            self.ctx.use_torch_stream()
        else:
            BB = np.ascontiguousarray(np.atleast_2d(B), dtype=np.complex128)
        nsys, n = BB.shape
        assert n == self.n
        jw = np.ascontiguousarray(np.asarray(jwmu, dtype=np.complex128).ravel())
        pi = np.ascontiguousarray(np.asarray(pidx, dtype=np.int32).ravel())
        hs = (C.c_void_p * len(precons))(*[p.h.value for p in precons])
        tols = None
        if np.ndim(tol) > 0:
            tols = np.ascontiguousarray(np.asarray(tol, dtype=np.float64).ravel())
            assert tols.size == nsys, 'one tolerance per system'
            tol = float(tols.max())
        if x0 is None:
            X = _torch().zeros_like(BB) if on_dev else np.zeros_like(BB)
        else:
            X = x0.clone().contiguous() if on_dev else np.ascontiguousarray(x0, dtype=np.complex128).copy()
            assert tuple(X.shape) == tuple(BB.shape), 'x0: one guess per system'
        flag = np.zeros(nsys, np.int32)
        relres = np.zeros(nsys)
        its = np.zeros(nsys, np.int32)
        iits = np.zeros(nsys, np.int32)
        rest = np.zeros(nsys)
        check(lib().vh_solve_batch(self.h, len(precons), C.cast(hs, C.c_void_p), nsys, ptr(jw), ptr(pi), ptr(BB),
                                   ptr(X), int(on_dev), int(x0 is not None), m, float(tol),
                                   ptr(tols) if tols is not None else None,
                                   restart, maxit, inner_restart, inner_tol,
                                   float(hbm_budget),
                                   ptr(flag), ptr(relres), ptr(its), ptr(iits), ptr(rest)))
        # resid_est: fp64 the reference-style preconditioned residual (pgmres.m:126), mixed the outer estimate
        return X, dict(flag=flag, relres=relres, iters=its, inner_iters=iits, resid_est=rest)

    def batch_bytes(self, method='mixed', restart=None, inner_restart=30):
        """Device bytes per system that solve_batch budgets on this geometry
        (Krylov workspace + the system's share of the matvec workspaces)."""
        m = METHODS[method]
        if restart is None:
            restart = 8 if m == 1 else 100
        b = np.zeros(1)
        check(lib().vh_solve_batch_bytes(self.h, m, int(restart), int(inner_restart), ptr(b)))
        return float(b[0])

    def autotune(self, prec, nrhs=1, reps=3):
        """Time every compiled pass-Z shape on nrhs vectors of this geometry
        and keep the fastest, then every X/Y-pass shape (vh_geom_autotune);
        returns (best pass-Z index, ms per pass-Z shape then per X/Y shape)."""
        self.ctx.use_torch_stream()
        best = C.c_int(0)
        ms = np.zeros(16)
        nsh = C.c_int(16)
        check(lib().vh_geom_autotune(self.h, prec, int(nrhs), int(reps), C.byref(best), ptr(ms), C.byref(nsh)))
        return best.value, ms[:nsh.value]

    def set_zshape(self, prec, shape):
        check(lib().vh_geom_set_zshape(self.h, prec, int(shape)))

    def set_xyshape(self, prec, shape):
        check(lib().vh_geom_set_xyshape(self.h, prec, int(shape)))

    def shapes(self):
        """Pass shapes in use: dict(z32, z64, xy32, xy64)."""
        v = np.zeros(4, np.int32)
        check(lib().vh_geom_get_shapes(self.h, ptr(v)))
        return dict(z32=int(v[0]), z64=int(v[1]), xy32=int(v[2]), xy64=int(v[3]))

    def fft3d(self, data, prec, direction):
        self.ctx.use_torch_stream()
        ncomp = data.numel() // (self.pad[0] * self.pad[1] * self.pad[2])
        check(lib().vh_fft3d(self.h, ptr(data), ncomp, prec, direction))

    def close(self):
        if getattr(self, 'h', None):
            check(lib().vh_geom_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Precon:
    """Schur-complement preconditioner of one geometry (vh_precon_*)."""

    def __init__(self, geom):
        h = C.c_void_p()
        check(lib().vh_precon_create(geom.h, C.byref(h)))
        self.h = h
        self.geom = geom
        self.stats = None

    def prepare(self, jwmu):
        st = np.zeros(4, dtype=np.int64)
        check(lib().vh_precon_prepare(self.h, float(np.real(jwmu)), float(np.imag(jwmu)), ptr(st)))
        self.stats = dict(nnzL=int(st[0]), levels_fwd=int(st[1]), levels_bwd=int(st[2]), n_node=int(st[3]))
        return self.stats

    @staticmethod
    def prepare_many(precons, jwmus, z=None, nthreads=0):
        """Prepare one preconditioner per frequency of one geometry on host
        threads (vh_precon_prepare_many); z: optional (n, nv) complex
        impedances (default: the geometry's)."""
        n = len(precons)
        hs = (C.c_void_p * n)(*[p.h.value for p in precons])
        jw = np.ascontiguousarray(np.asarray(jwmus, dtype=np.complex128).ravel())
        zz = None if z is None else np.ascontiguousarray(np.asarray(z, dtype=np.complex128))
        check(lib().vh_precon_prepare_many(C.cast(hs, C.c_void_p), n, ptr(jw), ptr(zz), int(nthreads)))

    def ainv(self):
        out = np.zeros(self.geom.num_curr)
        check(lib().vh_precon_get_ainv(self.h, ptr(out)))
        return out

    def set_type(self, kind, prectol=0.1):
        """'schur_invert' (exact LDL^T inverse, default) or 'schur_gmres'
        (inner GMRES on the Schur complement to prectol, FP64;
        lse_sparse_precon_multiply.m:26-46)."""
        t = {'schur_invert': VH_PC_SCHUR_INVERT, 'schur_gmres': VH_PC_SCHUR_GMRES}[kind]
        check(lib().vh_precon_set_type(self.h, t, float(prectol)))

    @property
    def inner_steps(self):
        """Inner Arnoldi steps (all systems) of the last schur_gmres apply."""
        v = C.c_int64(0)
        check(lib().vh_precon_inner_steps(self.h, C.byref(v)))
        return v.value

    def schur_multiply(self, x):
        """S x = Ar diag(A_inv) Ar' x for complex node vectors (rows of x)."""
        xx = np.ascontiguousarray(np.atleast_2d(x), dtype=np.complex128)
        nrhs, ld = xx.shape
        y = np.zeros_like(xx)
        check(lib().vh_precon_schur_multiply(self.h, nrhs, ptr(xx), ld, ptr(y), ld, 0))
        return y.reshape(np.shape(x))

    def apply(self, x, prec=VH_C128):
        if isinstance(x, np.ndarray):
            dt = np.complex128 if prec == VH_C128 else np.complex64
            xx = np.ascontiguousarray(np.atleast_2d(x), dtype=dt)
            nrhs, ld = xx.shape
            y = np.zeros_like(xx)
            check(lib().vh_precon_apply(self.h, prec, nrhs, ptr(xx), ld, ptr(y), ld, 0))
            return y.reshape(np.shape(x))
        xx = x if x.dim() == 2 else x.unsqueeze(0)
        nrhs, ld = xx.shape
        import torch
        y = torch.zeros_like(xx)
        self.geom.ctx.use_torch_stream()
        check(lib().vh_precon_apply(self.h, prec, nrhs, ptr(xx), ld, ptr(y), ld, 1))
        return y.reshape(x.shape)

    def close(self):
        if getattr(self, 'h', None):
            check(lib().vh_precon_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
