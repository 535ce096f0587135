"""Frequency sweep of one structure on the device, mirroring the solve part of
VoxHenry_executer.m:84-328:

* set-up (once): non-empty voxels (pre_define_structure_params.m:5-16), the
  Green's tensor (GPU, lse_generate_circulant_tensor), panel ids / Ar / port
  right-hand sides (library host code);
* per frequency: tolerance rule (:219-223), impedance z_real + z_imag (:231-259),
  preconditioner prepare (:273), one batched solve over all ports (:276-298),
  Y(:,port) = sum(rhs(:,kk).*x) (:311-316), Z = inv(Y), R = |Re Z|,
  L = |Im Z|/(2 pi f) (:327-328).

method 'fp64' runs pgmres with the reference semantics in complex128 on the
GPU; 'mixed' runs the FP64 flexible GMRES with the FP32 inner solver.
"""
import time

import numpy as np

from .core import Geometry, Precon, MU0
from ._lib import VH_C128, VH_C64


def reference_tol(f):
    """VoxHenry_executer.m:219-223."""
    return 1e-12 if f < 1e6 else 1e-8


def nonempty(structure):
    """0-based column-major indices of the non-empty voxels (idxS)."""
    s = np.abs(structure['sigma'].ravel(order='F')) > 1e-12
    if structure.get('lambdaL') is not None:
        s |= np.abs(structure['lambdaL'].ravel(order='F')) > 1e-12
    return np.flatnonzero(s)


def impedance(structure, idx, f):
    """z = z_real + z_imag per non-empty voxel (VoxHenry_executer.m:231-259)."""
    sig = structure['sigma'].ravel(order='F')[idx]
    lam = structure.get('lambdaL')
    with np.errstate(divide='ignore', invalid='ignore'):
        if lam is None:
            return (1.0 / sig).astype(complex)
        lam = lam.ravel(order='F')[idx]
        w = 2 * np.pi * f
        den = (w * MU0 * sig * lam ** 2) ** 2 + 1
        z = sig * (w * MU0 * lam ** 2) ** 2 / den + 1j * w * MU0 * lam ** 2 / den
        normal = (np.abs(lam) <= 1e-12) & (np.abs(sig) > 1e-12)
        z[normal] = 1.0 / sig[normal]
    return z


class Solver:
    def __init__(self, ctx, structure, pad_mode=1, precon_type='schur_invert', prectol=0.1, autotune=True):
        """precon_type: the executer's fl_precon_type (VoxHenry_executer.m:37-38):
        'schur_invert' (exact Schur inverse, pgmres) or 'schur_gmres' (inner
        GMRES on the Schur complement to prectol, with the flexible fpgmres
        outer solver, :295-297).  autotune: time every compiled pass-Z shape on
        this grid in both precisions and keep the fastest (the best shape
        depends on the padded grid: 1.6-2x on a 200x200x9 or 60x20x20 grid)."""
        self.ctx = ctx
        self.precon_type = precon_type
        self.prectol = prectol
        self.structure = structure
        self.idx = nonempty(structure)
        dims = structure['sigma'].shape
        t0 = time.perf_counter()
        self.G = Geometry(ctx, dims, structure['dx'], self.idx, pad_mode=pad_mode)
        self.G.generate_green()
        if autotune:
            for p in (VH_C64, VH_C128):
                self.G.autotune(p, nrhs=8, reps=2)
        t1 = time.perf_counter()
        self.rhs = self.G.build_topology(structure['ports_P'], structure['ports_N'], structure.get('well'))
        t2 = time.perf_counter()
        self.P = Precon(self.G)
        if precon_type != 'schur_invert':
            self.P.set_type(precon_type, prectol)
        self.nports = len(structure['ports_P'])
        self.times = dict(tensor=t1 - t0, topology=t2 - t1)
        self._z_f = None

    def prepare(self, f):
        """Impedance and preconditioner for frequency f (executer :225-273)."""
        if self.structure.get('lambdaL') is not None or self._z_f is None:
            self.G.set_impedance(impedance(self.structure, self.idx, f))
            self._z_f = f
        jwmu = 1j * 2 * np.pi * f * MU0
        t = time.perf_counter()
        self.P.prepare(jwmu)
        self.times['precon'] = self.times.get('precon', 0.0) + time.perf_counter() - t
        return jwmu

    def solve(self, f, method='fp64', tol=None, restart=100, maxit=10, **kw):
        """All ports at frequency f; returns X (nports x n) and statistics."""
        tol = reference_tol(f) if tol is None else tol
        jwmu = self.prepare(f)
        t = time.perf_counter()
        if method == 'fp64':
            X, st = self.G.gmres(self.rhs, jwmu, self.P, restart=restart, tol=tol, maxit=maxit, prec=VH_C128,
                                 flexible=self.precon_type == 'schur_gmres')
        elif method == 'mixed':
            X, st = self.G.solve_mixed(self.rhs, jwmu, self.P, tol=tol, **kw)
        else:
            raise ValueError(method)
        st['solve_time'] = time.perf_counter() - t
        return X, st

    def precons_for(self, freqs, nthreads=0):
        """One prepared preconditioner per frequency (symbolic analysis shared,
        host factorisations on parallel threads); each keeps the impedance of
        its frequency, which the batched solve uses for its systems."""
        freqs = np.asarray(freqs, float)
        if self._z_f is None:
            self.G.set_impedance(impedance(self.structure, self.idx, freqs[0]))
            self._z_f = freqs[0]
        out = [Precon(self.G) for _ in freqs]
        z = None
        if self.structure.get('lambdaL') is not None:
            z = np.array([impedance(self.structure, self.idx, f) for f in freqs])
        t = time.perf_counter()
        Precon.prepare_many(out, 1j * 2 * np.pi * freqs * MU0, z, nthreads)
        self.times['precon'] = self.times.get('precon', 0.0) + time.perf_counter() - t
        return out

    def sweep_batched(self, freqs, method='mixed', tol='reference', precons=None, pidx_of_freq=None, device=False,
                      **kw):
        """Every (port, frequency) system in one batched solve (vh_solve_batch).

        tol 'reference': each system gets the tolerance the reference's driver
        uses at its frequency (reference_tol, VoxHenry_executer.m:219-223);
        a number applies to every system.

        precons[pidx_of_freq[k]] preconditions frequency k (default: one
        preconditioner per frequency).  Every system uses the impedance its
        preconditioner was prepared with, so two-fluid conductors need one
        preconditioner per frequency (their impedance depends on f)."""
        freqs = np.asarray(freqs, float)
        if self.structure.get('lambdaL') is not None and precons is not None and \
                len(set(np.asarray(pidx_of_freq).tolist())) < len(freqs):
            raise ValueError('two-fluid conductors need one preconditioner per frequency')
        if self._z_f is None:
            self.G.set_impedance(impedance(self.structure, self.idx, freqs[0]))
            self._z_f = freqs[0]
        if precons is None:
            precons = self.precons_for(freqs)
            pidx_of_freq = np.arange(len(freqs))
        npo = self.nports
        jw = np.repeat(1j * 2 * np.pi * freqs * MU0, npo)
        pidx = np.repeat(np.asarray(pidx_of_freq), npo)
        if isinstance(tol, str):
            assert tol == 'reference', tol
            tol = np.repeat([reference_tol(f) for f in freqs], npo)
        if device:
            # right-hand sides and solutions stay in device memory; only the
            # port rows of the solutions (where the port vectors are nonzero)
            # come back, for Y = P^T X
            import torch
            nz = np.flatnonzero(np.any(self.rhs != 0, axis=0))
            rhs = torch.tensor(self.rhs, dtype=torch.complex128, device='cuda')
            B = rhs.repeat(len(freqs), 1).contiguous()
            torch.cuda.current_stream().synchronize()   # this stream only: other geometries may run beside
            t = time.perf_counter()
            X, st = self.G.solve_batch(B, jw, precons, pidx, method=method, tol=tol, **kw)
            Xp = X.index_select(1, torch.tensor(nz, device='cuda')).cpu().numpy()
            st['solve_time'] = time.perf_counter() - t
            Xp = Xp.reshape(len(freqs), npo, -1)
            Pnz = self.rhs[:, nz]
            Ys = [Pnz @ Xp[k].T for k in range(len(freqs))]
        else:
            B = np.repeat(self.rhs[None], len(freqs), axis=0).reshape(-1, self.G.n)
            t = time.perf_counter()
            X, st = self.G.solve_batch(B, jw, precons, pidx, method=method, tol=tol, **kw)
            st['solve_time'] = time.perf_counter() - t
            X = X.reshape(len(freqs), npo, -1)
            Ys = [self.rhs @ X[k].T for k in range(len(freqs))]
        out = dict(freq=freqs, Y=[], Z=[], R=[], L=[], stats=st)
        for k, f in enumerate(freqs):
            Y = Ys[k]
            Z = np.linalg.inv(Y)
            out['Y'].append(Y); out['Z'].append(Z)
            out['R'].append(np.abs(Z.real)); out['L'].append(np.abs(Z.imag) / (2 * np.pi * f))
        for k in ('Y', 'Z', 'R', 'L'):
            out[k] = np.array(out[k])
        return out

    def sweep(self, freqs, method='fp64', tol=None, **kw):
        out = dict(freq=np.asarray(freqs, float), Y=[], Z=[], R=[], L=[], stats=[])
        for f in freqs:
            X, st = self.solve(f, method, tol, **kw)
            Y = self.rhs @ X.T                                   # Y[kk, port] = sum(rhs_kk .* x_port)
            Z = np.linalg.inv(Y)
            out['Y'].append(Y)
            out['Z'].append(Z)
            out['R'].append(np.abs(Z.real))
            out['L'].append(np.abs(Z.imag) / (2 * np.pi * f))
            out['stats'].append(st)
        for k in ('Y', 'Z', 'R', 'L'):
            out[k] = np.array(out[k])
        return out


def workspace_bytes(S, nsys, methods=('mixed', 'fp64'), inner_restart=40):
    """Device bytes the batched solves of `nsys` systems of solver S allocate
    when every listed method runs on it: the Krylov arena grows to the
    largest method's layout, the matvec keeps a pass-X grid per precision
    and one shared B grid (vh_solve_batch_bytes per method)."""
    per = max(S.G.batch_bytes(m, inner_restart=inner_restart) for m in methods)
    if 'mixed' in methods and 'fp64' in methods:
        Px = S.G.pad[0]
        L, M, N = S.G.dims
        per += 5.0 * Px * M * N * 8     # the FP32 pass-X grid beside the FP64 one
    return nsys * per


def sweep_concurrent(jobs, method='mixed', workers=None, **kw):
    """Batched sweeps of several geometries as one run (the uneven batch).

    jobs: list of (solver, freqs, precons, pidx_of_freq, stream) or of
    (solver, freqs, precons, pidx_of_freq, stream, hbm_budget), where every
    solver has a Context of its own.  workers (default: one per job) host
    threads take the jobs from one queue in list order, each running its
    geometry's whole sweep (one vh_solve_batch over its port x frequency
    systems) on the job's stream; the library calls release the interpreter
    lock, so the geometries overlap on the GPU.  A job's hbm_budget (bytes)
    caps its solve's workspaces (vh_solve_batch chunks the systems to fit),
    so that geometries that run together cannot oversubscribe HBM.  Returns
    the sweeps' outputs in job order and the wall time."""
    import threading
    import torch
    out = [None] * len(jobs)
    errs = []
    nw = len(jobs) if workers is None else max(1, min(int(workers), len(jobs)))
    go = threading.Barrier(nw + 1)
    lock = threading.Lock()
    nxt = [0]

    def run():
        try:
            go.wait()
            while not errs:
                with lock:
                    i = nxt[0]
                    nxt[0] += 1
                if i >= len(jobs):
                    return
                S, freqs, precons, pidx, stream = jobs[i][:5]
                opts = dict(kw)
                if len(jobs[i]) > 5 and jobs[i][5]:
                    opts['hbm_budget'] = float(jobs[i][5])
                with torch.cuda.stream(stream):
                    out[i] = S.sweep_batched(freqs, method=method, precons=precons, pidx_of_freq=pidx, device=True,
                                             **opts)
                    stream.synchronize()
        except Exception as e:
            errs.append(e)

    th = [threading.Thread(target=run) for _ in range(nw)]
    for t in th:
        t.start()
    go.wait()
    t0 = time.perf_counter()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    if errs:
        raise errs[0]
    return out, wall
