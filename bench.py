#!/usr/bin/env python3
"""Benchmark of the MI355X VoxHenry solve path (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--detail FILE]

Workload (a "step"): the batched multi-port bus -- 8 parallel copper bars of
400x4x4 um at 8 um pitch, 1 um voxels in a 400x64x4 grid, 8 ports x 64
log-spaced frequencies 1e5..1e10 Hz = 512 right-hand sides, every (port,
frequency) system solved to a true FP64 relative residual <= 1e-8, or 1e-12
below 1 MHz (the reference driver's rule, VoxHenry_executer.m:219-223).  The
mixed path (complex64 inner FFT matvec + preconditioner under FP64 flexible
GMRES) is the headline `value`; the plain FP64 path (pgmres semantics in
complex128) is timed on the same step for the ratio.  A solve counts only
when its true FP64 residual ||b - A x|| / ||b|| meets its tolerance: solves/s
= converged systems / seconds.  R and L of both methods are compared with an
FP64 solve of the same systems to 1e-12 (untimed).  Inputs are synthetic
(generated geometry), resident in HBM before timing; per-frequency
preconditioner factors are prepared before the timed region.

Also measured: the bare FFT-matvec microbench -- a fully conductive 128^3
copper grid (padded to 256^3), 1 um voxels at 1 GHz, currents N(0,1) -- in
complex128 and complex64, with per-pass kernel times from HIP events on the
launch stream; `roofline` reports its dominant kernel (pass Z: z-FFT, Green's
product, inverse z-FFT).  Further lines: the 30x10x10 bar (10 f), the 5-turn
spiral (32 f), the numex1 bar (60x20x20 at 0.5 um, the reference's 41
frequencies) and the 96-geometry uneven batch (1-4 ports each, 16 f),
admitted into concurrent groups by HBM budget.  In these lines the mixed
sweeps also continue the systems whose R is a small part of |Z| to the
tolerance that fixes R and L to 1e-6 (sweep_batched rl_rtol; counted in the
timed sweep).

The CPU baseline is the FP64 numpy/scipy oracle (oracle/) timed on a bounded
sample of the same bus workload on this host.

stdout receives exactly one compact JSON line (the last line); progress
goes to stderr and the full record (per-geometry figures, per-pass times)
to the --detail file.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "batched port x freq solves/s at 1e-8 resid, mixed vs FP64; FFT-matvec HBM GB/s"
PASS_NAMES = ['X', 'Y', 'Z', 'Yi', 'Xi', 'node']
# pass-Z shapes (threads, elements per thread) in the library's FP32 table order
ZSHAPES32 = [(512, 8), (256, 8), (512, 4), (256, 4)]
TIGHT = 1e-12          # tolerance of the FP64 reference solves the R/L errors are measured against
# the reference's numex1 frequencies (1, 2.5, 5, 7.5 per decade, 1 Hz .. 10 GHz; results_numex1_*)
NUMEX1_FREQS = np.array([m * 10.0 ** e for e in range(10) for m in (1, 2.5, 5, 7.5)] + [1e10])


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--nfreq', type=int, default=64)
    ap.add_argument('--matvec-reps', type=int, default=20)
    ap.add_argument('--cube', type=int, default=128)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-fp64', action='store_true')
    ap.add_argument('--cpu-solves', type=int, default=2)
    ap.add_argument('--no-configs', action='store_true', help='skip the bar / spiral / numex1 / uneven-batch lines')
    ap.add_argument('--uneven-count', type=int, default=96, help='geometries of the uneven batch to run (all 96)')
    ap.add_argument('--uneven-nfreq', type=int, default=16)
    ap.add_argument('--uneven-workers', type=int, default=8)
    ap.add_argument('--detail', default=os.path.join(ROOT, 'gpurun_out', 'bench_detail.json'),
                    help='file that receives the full record (JSON)')
    ap.add_argument('--probe-pass-z', action='store_true',
                    help='only the 128^3 matvec (3 calls in complex64, then 3 in complex128) and a reference '
                         'elementwise kernel of known traffic (1 GiB read + 1 GiB written), for hardware-counter runs')
    ap.add_argument('--probe-bus-step', action='store_true',
                    help='only the bus set-up and one mixed step (512 systems), for hardware-counter runs')
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'pass_z_cube128_c64_traffic.md')


def recorded_traffic():
    """HBM bytes per launch of the roofline kernel, from the committed
    FETCH_SIZE / WRITE_SIZE summary (None when absent)."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith('traffic_bytes_per_launch:'):
                return int(line.split(':')[1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def r3(v):
    """Round for the compact line (3 significant digits)."""
    if v is None:
        return None
    return float('%.3g' % v)


def probe_pass_z(ctx, n=128):
    import torch
    import voxhenry_amd as vh
    G = vh.Geometry(ctx, (n, n, n), 1e-6, np.arange(n ** 3), pad_mode=1)
    G.generate_green()
    for prec, dt in ((vh.VH_C64, torch.complex64), (vh.VH_C128, torch.complex128)):
        best, ms = G.autotune(prec, nrhs=1, reps=3)
        log('[probe] prec %d pass-Z shape %d (ms per shape %s)' % (prec, best, np.round(ms, 3)))
        x = torch.zeros((1, G.num_curr), dtype=dt, device='cuda')
        y = torch.zeros_like(x)
        for _ in range(3):
            G.matvec_timed(x, y, 1j * 1e4, prec, 2, 1)
        del x, y
    a = torch.empty(1 << 28, dtype=torch.float32, device='cuda')
    a.fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    log('[probe] done')


def pass_bytes(g, nv, c, g_t, octant, nrhs=1):
    """Algorithmic HBM bytes of each fused pass for nrhs vectors (c bytes per
    complex, g_t bytes per stored tensor element)."""
    L, M, N, Px, Py, Pz = g
    a = Px * M * N
    b = Px * Py * N
    return {
        'X': nrhs * 5 * c * (nv + a),
        'Y': nrhs * 5 * c * (a + b),
        'Z': nrhs * 5 * c * (2 * b) + 7 * octant * g_t,
        'Yi': nrhs * 5 * c * (b + a),
        'Xi': nrhs * 5 * c * (a + 2 * nv) + nrhs * c * nv,
    }


def b_fft(V, nv, c, tensor_bytes, nrhs=1):
    """Canonical 3-pass byte model of one FFT-block matvec: b*c*(80V + 15nv) +
    tensor (stored size) + 8nv."""
    return nrhs * c * (80 * V + 15 * nv) + tensor_bytes + 8 * nv


def matvec_bench(ctx, n, reps):
    import torch
    import voxhenry_amd as vh
    from voxhenry_amd import structures
    st = structures.cube(n)
    L = M = N = n
    G = vh.Geometry(ctx, (L, M, N), st['dx'], np.arange(L * M * N), pad_mode=1)
    G.generate_green()
    G.set_impedance(np.full(L * M * N, 1.0 / 5.8e7, dtype=complex))
    tuned = {name: G.autotune(p, nrhs=1, reps=3)[0] for name, p in (('c128', vh.VH_C128), ('c64', vh.VH_C64))}
    f = 1e9
    jwmu = 1j * 2 * np.pi * f * vh.MU0
    nc = G.num_curr
    rng = np.random.default_rng(0)
    xh = rng.standard_normal(nc) + 1j * rng.standard_normal(nc)
    out = {}
    Px, Py, Pz = G.pad
    octant = (Px // 2 + 1) * (Py // 2 + 1) * (Pz // 2 + 1)
    res = {}
    for prec, name, c, gt, dt in ((vh.VH_C128, 'c128', 16, 16, torch.complex128), (vh.VH_C64, 'c64', 8, 4, torch.complex64)):
        x = torch.tensor(xh, dtype=dt, device='cuda').reshape(1, -1).contiguous()
        y = torch.zeros_like(x)
        G.matvec_timed(x, y, jwmu, prec, 1, 3)       # warm-up (allocates the workspaces)
        ms, ms_pass = G.matvec_timed_passes(x, y, jwmu, prec, 1, reps)
        res[name] = y.cpu().numpy().ravel().astype(np.complex128)
        pb = pass_bytes((L, M, N, Px, Py, Pz), L * M * N, c, gt, octant)
        bf = b_fft(L * M * N, L * M * N, c, 7 * octant * gt)
        out[name] = dict(ms=ms, matvecs_per_s=1e3 / ms, B_fft_GB=bf / 1e9, GBs_B_fft=bf / ms / 1e6,
                         frac_B_fft=bf / ms / 1e6 / HBM_PEAK_GBS,
                         passes={k: dict(ms=ms_pass[i], GB=pb.get(k, 0) / 1e9,
                                         GBs=(pb[k] / ms_pass[i] / 1e6) if k in pb and ms_pass[i] > 0 else None)
                                 for i, k in enumerate(PASS_NAMES)})
        del x, y
    out['c64_vs_c128_rel'] = float(np.linalg.norm(res['c64'] - res['c128']) / np.linalg.norm(res['c128']))
    out['zshape'] = tuned
    out['shapes'] = G.shapes()
    out['grid'] = [L, M, N]
    out['pad'] = [Px, Py, Pz]
    G.close()
    torch.cuda.empty_cache()
    return out


def bus_setup(ctx, nfreq):
    import voxhenry_amd as vh
    from voxhenry_amd import structures, solver as vsol
    freqs = np.logspace(5, 10, nfreq)
    st = structures.bus(freq=freqs)
    t0 = time.perf_counter()
    S = vsol.Solver(ctx, st)
    S.G.set_impedance(vsol.impedance(st, S.idx, freqs[0]))
    # pass shapes tuned for the batch widths of the solve (untimed set-up)
    S.zshape = {p: S.G.autotune(p, nrhs=64, reps=2)[0] for p in (vh.VH_C64, vh.VH_C128)}
    precs = S.precons_for(freqs)      # host LDL^T factorisations on a thread pool
    return S, freqs, precs, time.perf_counter() - t0


def bus_rhs(S, freqs, tol=None):
    """Device-resident right-hand sides of one step: the port vectors repeated
    for every frequency (system s = frequency s // nports, port s % nports),
    with each system's j*omega*mu, preconditioner index and tolerance (the
    reference driver's rule, or `tol` for every system)."""
    import torch
    import voxhenry_amd as vh
    from voxhenry_amd import solver as vsol
    npo = S.nports
    rhs = torch.tensor(S.rhs, dtype=torch.complex128, device='cuda')
    B = rhs.repeat(len(freqs), 1).contiguous()
    jw = np.repeat(1j * 2 * np.pi * np.asarray(freqs) * vh.MU0, npo)
    pidx = np.repeat(np.arange(len(freqs)), npo)
    if tol is None:
        tols = np.repeat([vsol.reference_tol(f) for f in freqs], npo)
    else:
        tols = np.full(len(freqs) * npo, float(tol))
    # the port vectors are nonzero on the port node rows only
    nz = torch.tensor(np.flatnonzero(np.any(S.rhs != 0, axis=0)), device='cuda')
    Pnz = rhs.index_select(1, nz).transpose(0, 1).contiguous()     # (nnz, nports)
    return (nz, Pnz), B, jw, pidx, tols


def bus_step(S, freqs, precs, method, data, **opts):
    """One step: every (port, frequency) system of the bus in one batched solve
    (per-system j*omega*mu, the preconditioner of its own frequency), then the
    port admittances Y(f) = P^T X(f) on the device; returns the port
    impedance matrices Z(f) and the solver statistics."""
    import torch
    (nz, Pnz), B, jw, pidx, tols = data
    npo = S.nports
    t = time.perf_counter()
    X, st = S.G.solve_batch(B, jw, precs, pidx, method=method, tol=tols, **opts)
    Xp = X.index_select(1, nz).reshape(len(freqs), npo, -1)      # port rows of every solution
    Yh = torch.matmul(Xp, Pnz).cpu().numpy()                    # Y[k] = X_k P^T (nports x nports)
    log('[bench]   %s step solve %.3f s, iters mean %.1f, inner mean %.1f, max relres %.2e, converged %d/%d'
        % (method, time.perf_counter() - t, st['iters'].mean(), st['inner_iters'].mean(), st['relres'].max(),
           int(np.sum(st['flag'] == 0)), len(st['flag'])))
    Z = np.stack([np.linalg.inv(Yh[k].T) for k in range(len(freqs))])
    return Z, st


def rl_errors(Z, Zref, freqs):
    """Largest relative error over frequencies and ports of the self
    resistance R = |Re Z_pp| and inductance L = |Im Z_pp| / (2 pi f), of R
    relative to |Z_pp|, and of Z normwise per frequency, against Zref."""
    d, dr = np.diagonal(Z, axis1=-2, axis2=-1), np.diagonal(Zref, axis1=-2, axis2=-1)
    R, Rr = np.abs(d.real), np.abs(dr.real)
    L, Lr = np.abs(d.imag), np.abs(dr.imag)
    return dict(R=float(np.max(np.abs(R - Rr) / Rr)), L=float(np.max(np.abs(L - Lr) / Lr)),
                R_over_absZ=float(np.max(np.abs(R - Rr) / np.abs(dr))),
                Z=float(max(np.linalg.norm(Z[k] - Zref[k]) / np.linalg.norm(Zref[k]) for k in range(len(Z)))))


def method_summary(nsys, dt, st):
    conv = int(np.sum(st['flag'] == 0))
    return dict(solves_per_s=conv / dt, s=dt, converged=conv, systems=nsys,
                relres_max=float(np.max(st['relres'])), iters_mean=float(np.mean(st['iters'])),
                iters_max=int(np.max(st['iters'])), inner_iters_mean=float(np.mean(st['inner_iters'])))


RL_RTOL = 1e-6     # accuracy of R and L the mixed sweeps of the configuration lines refine to (sweep_batched)
# The timed passes of a configuration line: label -> (solve_batch method, extra
# sweep arguments).  'mixed' and 'fp64_true' both carry every system to its
# true-residual tolerance (the solves/s ratio compares them); 'mixed_rl'
# continues the systems whose R or omega L is a small part of |Z| until both
# are resolved to RL_RTOL; 'fp64' is the reference's pgmres semantics (stops
# on the preconditioned residual: its unconverged systems are not counted).
CONFIG_PASSES = (('mixed', 'mixed', {}), ('mixed_rl', 'mixed', dict(rl_rtol=RL_RTOL)),
                 ('fp64_true', 'fp64_true', {}), ('fp64', 'fp64', {}))


def sweep_config(ctx, st, freqs, passes=CONFIG_PASSES, reps=1, keep=None, tight=True, **solve_kw):
    """Batched (port x frequency) solve of one geometry, device-resident:
    set-up (tensor, topology, preconditioners) untimed, then one warm-up and
    `reps` timed sweeps per method (the reference driver's tolerance per
    system); the R/L of each method are compared with an untimed FP64 sweep
    to 1e-12.  keep (a list): the solver and its preconditioners are
    appended to it and left open (for a later concurrent pass) instead of
    being released."""
    import torch
    from voxhenry_amd import solver as vsol
    t0 = time.perf_counter()
    S = vsol.Solver(ctx, st)
    precs = S.precons_for(freqs)
    pidx = np.arange(len(freqs))
    setup = time.perf_counter() - t0
    out = dict(nv=int(len(S.idx)), n=int(S.G.n), nports=S.nports, nfreq=len(freqs), setup_s=setup,
               grid=list(st['sigma'].shape))
    nsys = len(freqs) * S.nports
    res = {}
    for label, m, extra in passes:
        kw = dict(solve_kw, **extra)
        S.sweep_batched(freqs, method=m, precons=precs, pidx_of_freq=pidx, device=True, **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            r = S.sweep_batched(freqs, method=m, precons=precs, pidx_of_freq=pidx, device=True, **kw)
        torch.cuda.synchronize()
        res[label] = r
        out[label] = method_summary(nsys, (time.perf_counter() - t) / reps, r['stats'])
        out[label]['refined'] = int(r['stats'].get('refined', 0))
    if tight:
        ref = S.sweep_batched(freqs, method='fp64_true', tol=TIGHT, precons=precs, pidx_of_freq=pidx, device=True,
                              maxit=20, **solve_kw)
        out['ref_relres_max'] = float(ref['stats']['relres'].max())
        for label in res:
            out[label]['err_vs_tight_fp64'] = rl_errors(res[label]['Z'], ref['Z'], freqs)
        out['R_ohm'] = [float(v) for v in ref['R'][:, 0, 0]]
        out['L_H'] = [float(v) for v in ref['L'][:, 0, 0]]
    if 'mixed' in out and 'fp64_true' in out:
        out['speedup_mixed_over_fp64'] = out['mixed']['solves_per_s'] / max(out['fp64_true']['solves_per_s'], 1e-300)
    if keep is not None:
        keep.append((S, precs, pidx))
        return out
    for P in precs:
        P.close()
    S.G.close()
    torch.cuda.empty_cache()
    return out


def compact_config(c):
    """The few figures of a configuration line that go into the stdout line:
    solves/s of each pass, the mixed / fp64_true ratio, converged counts and
    the [R, L] errors against the FP64 sweep to 1e-12."""
    labels = [k for k, _, _ in CONFIG_PASSES if k in c]
    o = {k: r3(c[k]['solves_per_s']) for k in labels}
    o['ratio_mixed_over_fp64_true'] = r3(c.get('speedup_mixed_over_fp64'))
    o['converged'] = ', '.join('%s %d/%d' % (k, c[k]['converged'], c[k]['systems']) for k in labels)
    if 'mixed_rl' in c:
        o['mixed_rl_refined'] = c['mixed_rl'].get('refined', 0)
    if all('err_vs_tight_fp64' in c[k] for k in labels):
        o['RL_err_vs_fp64_1e-12'] = {k: [r3(c[k]['err_vs_tight_fp64']['R']), r3(c[k]['err_vs_tight_fp64']['L'])]
                                     for k in labels}
    return o


def uneven_config(count, nfreq, seed=0, workers=8, budget_frac=0.7):
    """The seeded uneven batch (bars, round wires, loops, two-fluid strips,
    1-4 ports each): every geometry's port x frequency systems in one
    batched solve, the geometries run as one job queue served by `workers`
    host threads (a stream each), admitted in groups whose device memory
    (set-up plus the solves' workspaces, vh_solve_batch_bytes) fits
    budget_frac of the free HBM; a geometry that does not fit beside its
    group opens the next one (its set-up is kept).  Per group, through the
    job queue: a mixed warm-up sweep and the FP64 sweep to 1e-12 that serves
    as the R/L reference (untimed; together they allocate the workspaces of
    both methods), then one timed pass per method.  solves/s = converged
    systems / summed timed wall."""
    import torch
    import voxhenry_amd as vh
    from voxhenry_amd import structures, solver as vsol
    freqs = np.logspace(6, 10, nfreq)
    pidx = np.arange(nfreq)
    geos = list(structures.uneven_batch(seed, count))
    # admission order: the largest remaining geometry, then the two smallest,
    # and so on, so that every group mixes a few large geometries (which set
    # its memory) with small ones that keep the other workers busy
    desc = sorted(range(len(geos)), key=lambda i: -np.count_nonzero(geos[i][1]['sigma']) * len(geos[i][1]['ports_P']))
    order, lo, hi = [], 0, len(desc) - 1
    while lo <= hi:
        order.append(desc[lo])
        lo += 1
        for _ in range(2):
            if lo <= hi:
                order.append(desc[hi])
                hi -= 1
    free0 = torch.cuda.mem_get_info()[0]
    budget = budget_frac * free0
    cap = 0.5 * budget                 # one geometry's workspaces: at most half the budget (it chunks beyond)
    labels = [k for k, _, _ in CONFIG_PASSES]
    tot = {k: dict(wall=0.0, conv=0) for k in labels}
    nsys_all = 0
    worst = {}
    kinds = {}
    groups = []
    per = []
    pos = 0
    setup_s = 0.0
    t_all = time.perf_counter()
    pending = None                     # a set-up geometry that opens the next group

    def setup(i):
        kind, st = geos[i]
        gctx = vh.Context(0)
        S = vsol.Solver(gctx, st)
        precs = S.precons_for(freqs)
        nsys = nfreq * S.nports
        return dict(i=i, kind=kind, S=S, precs=precs, ctx=gctx, stream=torch.cuda.Stream(), nsys=nsys,
                    ws=min(vsol.workspace_bytes(S, nsys), cap))

    while pos < len(order) or pending is not None:
        jobs = []
        need = 0.0
        while pending is not None or pos < len(order):
            if pending is None:
                t0 = time.perf_counter()
                j = setup(order[pos])
                pos += 1
                setup_s += time.perf_counter() - t0
            else:
                j, pending = pending, None
            free = torch.cuda.mem_get_info()[0]
            if jobs and need + j['ws'] > free - (1 - budget_frac) * free0:
                pending = j            # does not fit beside the group: it opens the next group
                break
            jobs.append(j)
            need += j['ws']
        groups.append(len(jobs))

        def pool(method, **kw):
            torch.cuda.synchronize()
            outs, wall = vsol.sweep_concurrent([(j['S'], freqs, j['precs'], pidx, j['stream'], j['ws']) for j in jobs],
                                               method=method, workers=workers, **kw)
            torch.cuda.synchronize()
            return outs, wall

        pool('mixed')                                        # warm-up
        refs, _ = pool('fp64_true', tol=TIGHT, maxit=20)     # the R/L reference (and the FP64 warm-up)
        res = {}
        for label, m, extra in CONFIG_PASSES:
            outs, wall = pool(m, **extra)
            res[label] = outs
            conv = sum(int(np.sum(o['stats']['flag'] == 0)) for o in outs)
            tot[label]['wall'] += wall
            tot[label]['conv'] += conv
            tot[label]['refined'] = tot[label].get('refined', 0) + sum(int(o['stats'].get('refined', 0))
                                                                        for o in outs)
            tot[label]['rl_unresolved'] = tot[label].get('rl_unresolved', 0) + \
                sum(int(o['stats'].get('rl_unresolved', 0)) for o in outs)
            log('[bench]   uneven group of %d, %s: %d/%d converged in %.3f s'
                % (len(jobs), label, conv, sum(j['nsys'] for j in jobs), wall))
        for k, j in enumerate(jobs):
            ref = refs[k]
            nsys_all += j['nsys']
            kd = kinds.setdefault(j['kind'], dict(geometries=0, systems=0, **{k + '_conv': 0 for k in labels}))
            kd['geometries'] += 1
            kd['systems'] += j['nsys']
            rec = dict(kind=j['kind'], grid=list(geos[j['i']][1]['sigma'].shape), ports=j['S'].nports,
                       nv=int(len(j['S'].idx)), ref_relres_max=float(ref['stats']['relres'].max()))
            for m in labels:
                o = res[m][k]
                c = int(np.sum(o['stats']['flag'] == 0))
                kd[m + '_conv'] += c
                e = rl_errors(o['Z'], ref['Z'], freqs)
                rec[m] = dict(converged=c, relres_max=float(o['stats']['relres'].max()), err=e)
                w = worst.setdefault(m, {})
                for q in ('R', 'L', 'R_over_absZ', 'Z'):
                    w[q] = max(w.get(q, 0.0), e[q])
                    if j['kind'] != 'strip':
                        w[q + '_normal'] = max(w.get(q + '_normal', 0.0), e[q])
            per.append(rec)
        for j in jobs:
            for P in j['precs']:
                P.close()
            j['S'].G.close()
            j['ctx'].close()
        del jobs
        torch.cuda.empty_cache()
    sps = {m: tot[m]['conv'] / tot[m]['wall'] for m in tot}
    out = dict(seed=seed, geometries=count, nfreq=nfreq, systems=nsys_all, workers=workers, groups=groups,
               budget_GB=budget / 1e9, setup_s=setup_s, wall_s=time.perf_counter() - t_all,
               speedup_mixed_over_fp64=sps['mixed'] / max(sps['fp64_true'], 1e-300),
               by_kind=kinds, err_vs_tight_fp64_max=worst, per_geometry=per)
    for m in labels:
        out[m] = dict(solves_per_s=sps[m], s=tot[m]['wall'], converged=tot[m]['conv'], systems=nsys_all,
                      refined=tot[m].get('refined', 0), rl_unresolved=tot[m].get('rl_unresolved', 0))
    return out


def timed_steps(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_matvec_baseline(n=128, reps=2):
    """Oracle FFT block (lse_matvect_mult.m:57-93: 5 padded forward FFTs, the
    5x5 spectral product, 5 inverse FFTs, crop) on the n^3 cube, timed on the
    host's cores (at most 16).  The tensor and currents are random of the
    right shapes (the timing does not depend on their values; generating the
    real tensor on the CPU takes minutes).  Reported as matvecs/s and as GB/s
    under the same byte model as the GPU figure."""
    from oracle import matvec as omv
    rng = np.random.default_rng(0)
    P = (2 * n, 2 * n, 2 * n)
    fN = np.empty(P + (7,), np.complex128)
    fN.real = rng.standard_normal(fN.shape)
    fN.imag = rng.standard_normal(fN.shape)
    J = rng.standard_normal((n, n, n, 5)) + 1j * rng.standard_normal((n, n, n, 5))
    cores = max(1, min(16, os.cpu_count() or 1))
    omv.FFT_WORKERS = cores
    try:
        omv.fft_block(J, fN, (n, n, n))                      # warm-up (FFT plans)
        t = time.perf_counter()
        for _ in range(reps):
            omv.fft_block(J, fN, (n, n, n))
        dt = (time.perf_counter() - t) / reps
    finally:
        omv.FFT_WORKERS = 1
    V = n ** 3
    gb = b_fft(V, V, 16, 7 * (2 * n) ** 3 * 16) / 1e9
    return dict(value=1.0 / dt, unit='matvecs/s', cores=cores, kind='port', GBs=gb / dt, cpu=cpu_model(),
                sample='%d FFT-block matvecs (complex128) of the %d^3 cube, FFTs on %d threads, spectral '
                       'product in numpy; random tensor and currents of the real shapes' % (reps, n, cores))


def cpu_baseline(nsolves):
    """Oracle (FP64 numpy/scipy restatement of the reference's algorithm) on a
    bounded sample of the bus workload: set-up untimed, then `nsolves`
    (port, frequency) solves timed with one core and with the host's cores
    (FFT workers; at most 16, the share of one GPU's host)."""
    from oracle import driver, gmres, matvec as omv
    from voxhenry_amd import structures
    st = structures.bus()
    geo = dict(st, well_cond=st['well'])
    t0 = time.perf_counter()
    prob = driver.Problem(geo)
    setup = time.perf_counter() - t0
    fs = np.logspace(5, 10, 64)[[40, 63]][:max(1, nsolves)]
    ops = [prob.operators(f) for f in fs]
    cores = max(1, min(16, os.cpu_count() or 1))

    def run(workers):
        omv.FFT_WORKERS = workers
        t = 0.0
        for A, P, *_ in ops:
            t1 = time.perf_counter()
            gmres.pgmres(A, prob.sys['rhs'][:, 0], 100, 1e-8, 10, P)
            t += time.perf_counter() - t1
        return len(ops) / t

    one = run(1)
    many = run(cores) if cores > 1 else one
    omv.FFT_WORKERS = 1
    return dict(value=many, unit='solves/s', cores=cores, kind='port', one_core=one, cpu=cpu_model(),
                sample='%d bus solves (port 1 at f=%s Hz), oracle pgmres + SuperLU Schur solve, FFTs on %d '
                       'threads (one-core figure in one_core); set-up %.1f s untimed'
                       % (len(ops), ','.join('%.3g' % f for f in fs), cores, setup))


def bus_kernels(S, precs, freqs):
    """Dominant kernels of the bus step at its real batch width (64 systems per
    launch, the widest active set of one preconditioner group): per-pass
    matvec times from HIP events, and one preconditioner apply (one
    frequency's factors, 64 vectors) timed with events on the launch
    stream.  Bytes per launch from the pass model and B_pc = b*c*(4*n_node +
    10*nv) + factor bytes (L in CSR and CSC, D, block inverses)."""
    import torch
    import voxhenry_amd as vh
    G = S.G
    L, M, N = G.dims
    Px, Py, Pz = G.pad
    octant = (Px // 2 + 1) * (Py // 2 + 1) * (Pz // 2 + 1)
    nb = 64
    jw = 1j * 2 * np.pi * 1e9 * vh.MU0
    out = dict(batch=nb)
    rng = np.random.default_rng(0)
    for prec, name, c, gt, dt in ((vh.VH_C128, 'c128', 16, 16, torch.complex128), (vh.VH_C64, 'c64', 8, 4, torch.complex64)):
        xh = rng.standard_normal((nb, G.n)) + 1j * rng.standard_normal((nb, G.n))
        x = torch.tensor(xh, dtype=dt, device='cuda')
        y = torch.zeros_like(x)
        G.matvec_timed(x, y, jw, prec, 0, 2)
        ms, ms_pass = G.matvec_timed_passes(x, y, jw, prec, 0, 5)
        pb = pass_bytes((L, M, N, Px, Py, Pz), G.nv, c, gt, octant, nrhs=nb)
        zi = PASS_NAMES.index('Z')
        out[name] = dict(matvec_ms=ms, pass_z_ms=ms_pass[zi], pass_z_GB=pb['Z'] / 1e9,
                         pass_z_GBs=pb['Z'] / ms_pass[zi] / 1e6)
        # preconditioner apply of one frequency on the same 64 vectors
        k = len(precs) // 2
        P = precs[k]
        if P.stats is None:
            P.prepare(1j * 2 * np.pi * freqs[k] * vh.MU0)     # same factors again, with their sizes
        P.apply(x, prec)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            P.apply(x, prec)
        e1.record()
        e1.synchronize()
        pms = e0.elapsed_time(e1) / 5
        st = P.stats
        nnzL = st['nnzL']
        fb = 8 if prec == vh.VH_C128 else 4
        bpc = nb * c * (4 * st['n_node'] + 10 * G.nv) + 2 * nnzL * (4 + fb) + st['n_node'] * fb
        out[name].update(precon_ms=pms, precon_GB=bpc / 1e9, precon_GBs=bpc / pms / 1e6)
        del x, y
    torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: there is no multi-GPU path.  Every (port, frequency) system and '
              'every geometry is independent, so several GPUs would each run their own share of the batch.',
              file=sys.stderr)
        sys.exit(2)
    import torch
    import voxhenry_amd as vh
    ctx = vh.Context(0)
    ctx.use_torch_stream()
    if args.probe_pass_z:
        probe_pass_z(ctx)
        return
    if args.probe_bus_step:
        S, freqs, precs, _ = bus_setup(ctx, args.nfreq)
        data = bus_rhs(S, freqs)
        log('[probe] bus set-up done')
        bus_step(S, freqs, precs, 'mixed', data)
        torch.cuda.synchronize()
        log('[probe] done')
        return
    detail = {}

    log('[bench] FFT-matvec microbench %d^3' % args.cube)
    mv = matvec_bench(ctx, args.cube, args.matvec_reps)
    detail['fft_matvec'] = mv
    log('[bench] matvec c128 %.3f ms  c64 %.3f ms' % (mv['c128']['ms'], mv['c64']['ms']))

    log('[bench] bus set-up (%d frequencies)' % args.nfreq)
    S, freqs, precs, t_setup = bus_setup(ctx, args.nfreq)
    nsol = len(freqs) * S.nports
    log('[bench] set-up %.1f s, %d systems per step' % (t_setup, nsol))

    data = bus_rhs(S, freqs)
    s_mixed, (Zm, stm) = timed_steps(lambda: bus_step(S, freqs, precs, 'mixed', data), args.steps, args.warmup)
    log('[bench] mixed step %.3f s' % s_mixed)
    mixed = method_summary(nsol, s_mixed, stm)
    fp64 = None
    if not args.no_fp64:
        s64, (Z64, st64) = timed_steps(lambda: bus_step(S, freqs, precs, 'fp64_true', data), max(1, args.steps // 2),
                                       max(1, args.warmup))
        log('[bench] fp64 step %.3f s' % s64)
        fp64 = method_summary(nsol, s64, st64)
        # the reference-style stopping residual ||M r|| / ||M b|| (pgmres.m:126)
        fp64['precond_resid_max'] = float(st64['resid_est'].max())
    log('[bench] bus FP64 reference solve to %g' % TIGHT)
    Zt, stt = bus_step(S, freqs, precs, 'fp64_true', bus_rhs(S, freqs, TIGHT), maxit=20)
    ref_relres = float(stt['relres'].max())
    mixed['err_vs_tight_fp64'] = rl_errors(Zm, Zt, freqs)
    if fp64:
        fp64['err_vs_tight_fp64'] = rl_errors(Z64, Zt, freqs)
    log('[bench] bus kernels at the solve batch width')
    bk = bus_kernels(S, precs, freqs)
    detail['bus'] = dict(mixed=mixed, fp64=fp64, setup_s=t_setup, ref_relres_max=ref_relres, kernels=bk,
                         L_self_nH_at_fmin=float(np.abs(np.diagonal(Zt[0]).imag).mean() / (2 * np.pi * freqs[0]) * 1e9))

    z = mv['c64']['passes']['Z']
    nt, e = ZSHAPES32[mv['zshape']['c64']]
    kname = 'k_pass_z_p2<float, 256, %d, %d, true>' % (nt, e)   # compile-time kernel of the power-of-two Pz = 256
    roof = dict(kernel=kname + ' (z-FFT + Green product + inverse z-FFT), 128^3 grid, complex64, 1 vector per launch',
                bound='hbm', achieved=z['GBs'], peak=HBM_PEAK_GBS, unit='GB/s',
                frac=(z['GBs'] / HBM_PEAK_GBS) if z['GBs'] else None, traffic=recorded_traffic(),
                bytes_per_launch=z['GB'] * 1e9, launch_ms=z['ms'])

    # the bus geometry holds a Krylov workspace sized from most of the free
    # memory: release it before the configuration sweeps
    del data
    for P in precs:
        P.close()
    S.G.close()
    torch.cuda.empty_cache()

    configs = {}
    if not args.no_configs:
        from voxhenry_amd import structures
        log('[bench] bar 30x10x10, 10 frequencies')
        configs['bar_30x10x10'] = sweep_config(ctx, structures.bar(30, 10, 10), np.logspace(4, 10, 10))
        log('[bench] spiral, 32 frequencies')
        configs['spiral_5turn'] = sweep_config(ctx, structures.spiral(), np.logspace(6, 11, 32))
        log('[bench] numex1 bar 60x20x20 at 0.5 um, 41 frequencies')
        configs['numex1_bar'] = sweep_config(ctx, structures.bar(60, 20, 20, dx=0.5e-6), NUMEX1_FREQS)
        if args.uneven_count > 0:
            log('[bench] uneven batch, %d geometries x %d frequencies' % (args.uneven_count, args.uneven_nfreq))
            configs['uneven_batch'] = uneven_config(args.uneven_count, args.uneven_nfreq,
                                                    workers=args.uneven_workers)
    detail['configs'] = configs

    cpu = cpu_mv = None
    if not args.no_cpu:
        log('[bench] CPU baseline (oracle)')
        cpu = cpu_baseline(args.cpu_solves)
        log('[bench] CPU baseline of the 128^3 FFT-block matvec (oracle)')
        cpu_mv = cpu_matvec_baseline(128, 2)
    detail['cpu_baseline'] = cpu
    detail['cpu_baseline_matvec'] = cpu_mv
    detail['roofline'] = roof

    # ---- the compact stdout line -------------------------------------------------
    def msum(m):
        if m is None:
            return None
        o = dict(solves_per_s=r3(m['solves_per_s']), converged='%d/%d' % (m['converged'], m['systems']),
                 s_per_step=r3(m['s']), iters_mean=r3(m['iters_mean']), relres_max=r3(m['relres_max']))
        if m.get('inner_iters_mean'):
            o['inner_iters_mean'] = r3(m['inner_iters_mean'])
        if 'err_vs_tight_fp64' in m:
            e = m['err_vs_tight_fp64']
            o['err_vs_fp64_1e-12'] = dict(R=r3(e['R']), L=r3(e['L']), Z=r3(e['Z']))
        return o

    cfg_line = {}
    for k, c in configs.items():
        cfg_line[k] = compact_config(c)
    if 'uneven_batch' in configs:
        u = configs['uneven_batch']
        cfg_line['uneven_batch'].update(geometries=u['geometries'], systems=u['systems'], groups=len(u['groups']))
    line = dict(metric=METRIC, value=r3(mixed['solves_per_s']), unit='solves/s', n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=r3(s_mixed * 1e3), higher_is_better=True, scaling='weak',
                vs_baseline=None, dtype='c64 inner / f64 outer', data='synthetic',
                config=dict(workload='bus: 8 copper bars 400x4x4 um, 400x64x4 grid, 8 ports x %d freqs 1e5-1e10 Hz '
                                     '= %d systems per step, tol 1e-8 (1e-12 below 1 MHz), true FP64 residual; '
                                     'value counts converged systems only' % (len(freqs), nsol),
                            parallelism='1 GPU'),
                mixed=msum(mixed), fp64=msum(fp64),
                fp64_speedup=r3(mixed['solves_per_s'] / fp64['solves_per_s']) if fp64 else None,
                matvec_128cube=dict(c128_ms=r3(mv['c128']['ms']), c64_ms=r3(mv['c64']['ms']),
                                    c128_GBs_B_fft=r3(mv['c128']['GBs_B_fft']), c64_GBs_B_fft=r3(mv['c64']['GBs_B_fft']),
                                    c128_pass_z_GBs=r3(mv['c128']['passes']['Z']['GBs'])),
                roofline=dict(bound='hbm', achieved=r3(roof['achieved']), peak=HBM_PEAK_GBS, unit='GB/s',
                              frac=r3(roof['frac']), traffic=roof['traffic'], kernel=roof['kernel']),
                roofline_bus=dict(batch=bk['batch'],
                                  pass_z_c128=dict(GBs=r3(bk['c128']['pass_z_GBs']), ms=r3(bk['c128']['pass_z_ms'])),
                                  pass_z_c64=dict(GBs=r3(bk['c64']['pass_z_GBs']), ms=r3(bk['c64']['pass_z_ms'])),
                                  precon_c128=dict(GBs=r3(bk['c128']['precon_GBs']), ms=r3(bk['c128']['precon_ms'])),
                                  precon_c64=dict(GBs=r3(bk['c64']['precon_GBs']), ms=r3(bk['c64']['precon_ms']))),
                cpu_baseline=None if cpu is None else dict(value=r3(cpu['value']), unit=cpu['unit'], cores=cpu['cores'],
                                                           kind=cpu['kind'], sample=cpu['sample'],
                                                           one_core=r3(cpu['one_core']), cpu=cpu['cpu']),
                cpu_baseline_matvec=None if cpu_mv is None else dict(value=r3(cpu_mv['value']), unit=cpu_mv['unit'],
                                                                     cores=cpu_mv['cores'], GBs=r3(cpu_mv['GBs'])),
                configs=cfg_line, setup_s=r3(t_setup))
    if args.detail:
        os.makedirs(os.path.dirname(os.path.abspath(args.detail)), exist_ok=True)
        with open(args.detail, 'w') as fh:
            json.dump(dict(line=line, detail=detail), fh, default=float)
            fh.write('\n')
        log('[bench] full record in %s' % args.detail)
    sys.stderr.flush()
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
