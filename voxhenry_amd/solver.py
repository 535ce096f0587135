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
GPU ('fp64_true' carries the systems it leaves above tolerance on to the true
residual in FP64 flexible GMRES); 'mixed' runs the FP64 flexible GMRES with
the FP32 inner solver.
"""
import time

import numpy as np

from .core import Geometry, Precon, MU0
from ._lib import VH_C128, VH_C64


# accuracy-driven tolerance of sweep_batched(rl_rtol=...): the R error of a
# solve to a relative residual tol was measured at up to ~20 tol |Z| / R (the
# 5-turn spiral at 100 GHz; up to ~25 on the uneven batch's bars), so the
# tolerance is rl_rtol * (R / |Z|) / 50.
# Below 1e-12 (the tightest residual the driver asks for) the target is not
# pursued: such a system (R ~ 1e-9 |Z|, a superconductor at low frequency)
# cannot have R to rl_rtol from any affordable solve; it is counted as
# unresolved instead.
RL_MARGIN = 50.0
RL_TOL_FLOOR = 1e-12


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
                      rl_rtol=None, **kw):
        """Every (port, frequency) system in one batched solve (vh_solve_batch).

        tol 'reference': each system gets the tolerance the reference's driver
        uses at its frequency (reference_tol, VoxHenry_executer.m:219-223);
        a number applies to every system.

        precons[pidx_of_freq[k]] preconditions frequency k (default: one
        preconditioner per frequency).  Every system uses the impedance its
        preconditioner was prepared with, so two-fluid conductors need one
        preconditioner per frequency (their impedance depends on f).

        rl_rtol (mixed only): relative accuracy wanted of R and L.  Where R
        (or omega L) is a small part of |Z| a 1e-8 residual leaves it
        inaccurate (its error grows like tol * |Z| / R), so after the first
        solve the systems of every frequency whose worst port has
        min(R, omega L) / |Z| = rho get their
        tolerance tightened to rl_rtol * rho / RL_MARGIN when that is below
        their tolerance and not below RL_TOL_FLOOR, and continue from their
        solutions (warm-started refinement).  stats['refined'] counts the
        systems that continued, stats['rl_unresolved'] those whose target
        lies below the floor (their R is not fixed to rl_rtol); flag /
        relres stay relative to the original tolerance."""
        freqs = np.asarray(freqs, float)
        if self.structure.get('lambdaL') is not None and precons is not None and \
                len(set(np.asarray(pidx_of_freq).tolist())) < len(freqs):
            raise ValueError('two-fluid conductors need one preconditioner per frequency')
        if rl_rtol is not None and method != 'mixed':
            raise ValueError('rl_rtol needs the mixed method (warm-started refinement)')
        if self._z_f is None:
            self.G.set_impedance(impedance(self.structure, self.idx, freqs[0]))
            self._z_f = freqs[0]
        if precons is None:
            precons = self.precons_for(freqs)
            pidx_of_freq = np.arange(len(freqs))
        npo = self.nports
        nf = len(freqs)
        jw = np.repeat(1j * 2 * np.pi * freqs * MU0, npo)
        pidx = np.repeat(np.asarray(pidx_of_freq), npo)
        if isinstance(tol, str):
            assert tol == 'reference', tol
            tol = np.repeat([reference_tol(f) for f in freqs], npo)
        tols = np.broadcast_to(np.asarray(tol, float), (nf * npo,)).copy()
        nz = np.flatnonzero(np.any(self.rhs != 0, axis=0))
        Pnz = self.rhs[:, nz]
        if device:
            # right-hand sides and solutions stay in device memory; only the
            # port rows of the solutions (where the port vectors are nonzero)
            # come back, for Y = P^T X
            import torch
            rhs = torch.tensor(self.rhs, dtype=torch.complex128, device='cuda')
            B = rhs.repeat(nf, 1).contiguous()
            nzd = torch.tensor(nz, device='cuda')
            torch.cuda.current_stream().synchronize()   # this stream only: other geometries may run beside
            rows = lambda X: X.index_select(1, nzd).cpu().numpy()
        else:
            B = np.repeat(self.rhs[None], nf, axis=0).reshape(-1, self.G.n)
            rows = lambda X: X[:, nz]
        t = time.perf_counter()
        X, st = self.G.solve_batch(B, jw, precons, pidx, method=method, tol=tols, **kw)
        Xp = rows(X).reshape(nf, npo, -1)
        Ys = np.array([Pnz @ Xp[k].T for k in range(nf)])     # Y[kk, port] = sum(rhs_kk .* x_port)
        st['refined'] = 0
        st['rl_unresolved'] = 0
        if rl_rtol is not None:
            Zs = np.linalg.inv(Ys)
            d = np.diagonal(Zs, axis1=1, axis2=2)
            # the smaller of R / |Z| and omega L / |Z| over the ports: L is
            # the small part of Z at low frequency as R is at high frequency
            rho = np.min(np.minimum(np.abs(d.real), np.abs(d.imag)) / np.maximum(np.abs(d), 1e-300), axis=1)
            target = np.repeat(rl_rtol * rho / RL_MARGIN, npo)
            st['rl_unresolved'] = int(np.sum(target < RL_TOL_FLOOR))
            want = np.minimum(tols, target)
            sel = np.flatnonzero((want < tols * 0.999) & (target >= RL_TOL_FLOOR))
            if len(sel):
                if device:
                    import torch
                    idx = torch.tensor(sel, device='cuda')
                    Bs, X0 = B.index_select(0, idx).contiguous(), X.index_select(0, idx).contiguous()
                else:
                    Bs, X0 = B[sel], X[sel]
                X2, st2 = self.G.solve_batch(Bs, jw[sel], precons, pidx[sel], method=method, tol=want[sel], x0=X0,
                                             **kw)
                if device:
                    X.index_copy_(0, idx, X2)
                else:
                    X[sel] = X2
                for key in ('iters', 'inner_iters'):
                    st[key][sel] += st2[key]
                st['relres'][sel] = st2['relres']
                st['resid_est'][sel] = st2['resid_est']
                st['flag'][sel] = (st2['relres'] > tols[sel]).astype(st['flag'].dtype)
                st['refined'] = len(sel)
                ks = np.unique(sel // npo)
                Xp2 = rows(X2).reshape(len(sel), -1)
                for q, s_ in enumerate(sel):
                    Xp[s_ // npo, s_ % npo] = Xp2[q]
                for k in ks:
                    Ys[k] = Pnz @ Xp[k].T
        st['solve_time'] = time.perf_counter() - t
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


def workspace_bytes(S, nsys, methods=('mixed', 'fp64'), inner_restart=30):
    """Device bytes the batched solves of `nsys` systems of solver S allocate
    when every listed method runs on it: the Krylov arena grows to the
    largest method's layout, the matvec keeps a pass-X grid per precision
    and one shared B grid (vh_solve_batch_bytes per method)."""
    per = max(S.G.batch_bytes(m, inner_restart=inner_restart) for m in methods)
    if 'mixed' in methods and ('fp64' in methods or 'fp64_true' in methods):
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
