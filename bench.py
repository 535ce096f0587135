#!/usr/bin/env python3
"""Benchmark of the MI355X VoxHenry solve path (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (a "step"): the batched multi-port bus -- 8 parallel copper bars of
400x4x4 um at 8 um pitch, 1 um voxels in a 400x64x4 grid, 8 ports x 64
log-spaced frequencies 1e5..1e10 Hz = 512 right-hand sides, every (port,
frequency) system solved to a true FP64 relative residual <= 1e-8, or 1e-12
below 1 MHz (the reference driver's rule, VoxHenry_executer.m:219-223: at
low frequency L comes from the small imaginary part of Z, which a 1e-8
residual does not fix to 1e-6).  The mixed
path (complex64 inner FFT matvec + preconditioner under FP64 flexible GMRES) is
the headline `value` (solves/s); the plain FP64 path (pgmres semantics in
complex128) is timed on the same step for the ratio.  Inputs are synthetic
(generated geometry), resident in HBM before timing; per-frequency
preconditioner factors are prepared before the timed region.

Also measured: the bare FFT-matvec microbench -- a fully conductive 128^3
copper grid (padded to 256^3), 1 um voxels at 1 GHz, currents N(0,1) -- in
complex128 and complex64, with the per-pass kernel times from HIP events on the
launch stream.  `roofline` reports its dominant kernel (pass Z: z-FFT, Green's
product, inverse z-FFT).

The CPU baseline is the FP64 numpy/scipy oracle (oracle/) timed on a bounded
sample of the same bus workload on this host; cpu_baseline_matvec times the
oracle's FFT-block matvec on the 128^3 cube next to the GPU figure.
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
    ap.add_argument('--no-configs', action='store_true', help='skip the bar / spiral / uneven-batch lines')
    ap.add_argument('--uneven-count', type=int, default=96, help='geometries of the uneven batch to run (all 96)')
    ap.add_argument('--uneven-nfreq', type=int, default=16)
    ap.add_argument('--probe-pass-z', action='store_true',
                    help='only the 128^3 complex64 matvec (3 calls) and a reference elementwise kernel of known '
                         'traffic (1 GiB read + 1 GiB written), for hardware-counter runs')
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


def probe_pass_z(ctx, n=128):
    import torch
    import voxhenry_amd as vh
    G = vh.Geometry(ctx, (n, n, n), 1e-6, np.arange(n ** 3), pad_mode=1)
    G.generate_green()
    best, ms = G.autotune(vh.VH_C64, nrhs=1, reps=3)
    log('[probe] pass-Z shape %d (ms per shape %s)' % (best, np.round(ms, 3)))
    x = torch.zeros((1, G.num_curr), dtype=torch.complex64, device='cuda')
    y = torch.zeros_like(x)
    for _ in range(3):
        G.matvec_timed(x, y, 1j * 1e4, vh.VH_C64, 2, 1)
    a = torch.empty(1 << 28, dtype=torch.float32, device='cuda')
    a.fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    log('[probe] done')


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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
                         passes={k: dict(ms=ms_pass[i], GB=pb.get(k, 0) / 1e9,
                                         GBs=(pb[k] / ms_pass[i] / 1e6) if k in pb and ms_pass[i] > 0 else None)
                                 for i, k in enumerate(PASS_NAMES)})
        del x, y
    out['c64_vs_c128_rel'] = float(np.linalg.norm(res['c64'] - res['c128']) / np.linalg.norm(res['c128']))
    out['zshape'] = tuned
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


def bus_rhs(S, freqs):
    """Device-resident right-hand sides of one step: the port vectors repeated
    for every frequency (system s = frequency s // nports, port s % nports),
    with each system's j*omega*mu, preconditioner index and tolerance."""
    import torch
    import voxhenry_amd as vh
    from voxhenry_amd import solver as vsol
    npo = S.nports
    rhs = torch.tensor(S.rhs, dtype=torch.complex128, device='cuda')
    B = rhs.repeat(len(freqs), 1).contiguous()
    jw = np.repeat(1j * 2 * np.pi * np.asarray(freqs) * vh.MU0, npo)
    pidx = np.repeat(np.arange(len(freqs)), npo)
    # the reference driver's tolerance at each system's frequency
    tols = np.repeat([vsol.reference_tol(f) for f in freqs], npo)
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
    log('[bench]   %s step solve %.3f s, iters mean %.1f, inner mean %.1f, max relres %.2e'
        % (method, time.perf_counter() - t, st['iters'].mean(), st['inner_iters'].mean(), st['relres'].max()))
    Z = np.stack([np.linalg.inv(Yh[k].T) for k in range(len(freqs))])
    return Z, st


def sweep_config(ctx, st, freqs, tol='reference', methods=('mixed', 'fp64'), reps=1, keep=None):
    """Batched (port x frequency) solve of one geometry, device-resident:
    set-up (tensor, topology, preconditioners) untimed, then one warm-up and
    `reps` timed sweeps per method; returns solves/s per method, R/L of the
    FP64 sweep and the mixed-vs-FP64 errors (self terms, and normwise Z).
    keep (a list): the solver and its preconditioners are appended to it and
    left open (for a later concurrent pass) instead of being released."""
    import torch
    from voxhenry_amd import solver as vsol
    t0 = time.perf_counter()
    S = vsol.Solver(ctx, st)
    precs = S.precons_for(freqs)
    pidx = np.arange(len(freqs))
    setup = time.perf_counter() - t0
    out = dict(nv=int(len(S.idx)), n=int(S.G.n), nports=S.nports, nfreq=len(freqs), setup_s=setup,
               grid=list(st['sigma'].shape))
    res = {}
    for m in methods:
        S.sweep_batched(freqs, method=m, tol=tol, precons=precs, pidx_of_freq=pidx, device=True)
        t = time.perf_counter()
        for _ in range(reps):
            r = S.sweep_batched(freqs, method=m, tol=tol, precons=precs, pidx_of_freq=pidx, device=True)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / reps
        res[m] = r
        nsys = len(freqs) * S.nports
        out[m] = dict(solves_per_s=nsys / dt, s=dt, all_converged=bool(np.all(r['stats']['flag'] == 0)),
                      relres_max=float(r['stats']['relres'].max()), iters_max=int(r['stats']['iters'].max()))
    if 'mixed' in res and 'fp64' in res:
        a, b = res['mixed'], res['fp64']
        ii = np.arange(S.nports)
        out['mixed_vs_fp64'] = dict(
            R_self=float(np.max(np.abs(a['R'][:, ii, ii] - b['R'][:, ii, ii]) / np.abs(b['R'][:, ii, ii]))),
            # R relative to |Z|: where R << |Z| (skin-effect limit, or the
            # tiny loss of a superconductor) a solve to a 1e-8 residual
            # fixes R only to about 1e-8 |Z| / R
            R_over_absZ=float(np.max(np.abs(a['R'][:, ii, ii] - b['R'][:, ii, ii]) / np.abs(b['Z'][:, ii, ii]))),
            L_self=float(np.max(np.abs(a['L'][:, ii, ii] - b['L'][:, ii, ii]) / np.abs(b['L'][:, ii, ii]))),
            Z_norm=float(max(np.linalg.norm(a['Z'][k] - b['Z'][k]) / np.linalg.norm(b['Z'][k])
                             for k in range(len(freqs)))))
        out['speedup_mixed_over_fp64'] = out['mixed']['solves_per_s'] / out['fp64']['solves_per_s']
    if 'fp64' in res:
        out['R_ohm'] = [float(v) for v in res['fp64']['R'][:, 0, 0]]
        out['L_H'] = [float(v) for v in res['fp64']['L'][:, 0, 0]]
    if keep is not None:
        keep.append((S, precs, pidx))
        return out
    for P in precs:
        P.close()
    S.G.close()
    torch.cuda.empty_cache()
    return out


def uneven_config(ctx, count, nfreq, seed=0, wave=8):
    """A prefix of the seeded uneven batch (bars, round wires, loops, two-fluid
    strips), each geometry's port x frequency systems in one batched solve;
    aggregate solves/s over the geometries (set-up untimed), one after
    another and `wave` geometries at a time concurrently (a context and
    stream each; geometries are released wave by wave, since every one holds
    its Krylov workspace -- restart 100 in FP64)."""
    import torch
    import voxhenry_amd as vh
    from voxhenry_amd import structures, solver as vsol
    freqs = np.logspace(6, 10, nfreq)
    tot = {'mixed': 0.0, 'fp64': 0.0}
    conc_wall = {'mixed': 0.0, 'fp64': 0.0}
    nsys = 0
    per = []
    # R relative to itself over the normal conductors; R relative to |Z| over
    # all (a two-fluid strip's R is ~1e-9 of |Z|, below any solve's resolution)
    worst = dict(R_self_normal=0.0, L_self_normal=0.0, Z_norm_normal=0.0, R_over_absZ=0.0, L_self=0.0, Z_norm=0.0)
    geos = list(structures.uneven_batch(seed, count))
    for w0 in range(0, len(geos), wave):
        jobs = []      # (solver, preconditioners, pidx, context, stream) per geometry
        wsys = 0
        for kind, st in geos[w0:w0 + wave]:
            gctx = vh.Context(0)
            stream = torch.cuda.Stream()
            keep = []
            with torch.cuda.stream(stream):
                r = sweep_config(gctx, st, freqs, keep=keep)
            jobs.append(keep[0] + (gctx, stream))
            wsys += r['nfreq'] * r['nports']
            for m in tot:
                tot[m] += r[m]['s']
            e = r['mixed_vs_fp64']
            if kind != 'strip':
                worst['R_self_normal'] = max(worst['R_self_normal'], e['R_self'])
                worst['L_self_normal'] = max(worst['L_self_normal'], e['L_self'])
                worst['Z_norm_normal'] = max(worst['Z_norm_normal'], e['Z_norm'])
            for k in ('R_over_absZ', 'L_self', 'Z_norm'):
                worst[k] = max(worst[k], e[k])
            per.append(dict(kind=kind, grid=r['grid'], nv=r['nv'], setup_s=round(r['setup_s'], 2),
                            mixed_solves_per_s=round(r['mixed']['solves_per_s'], 1),
                            fp64_solves_per_s=round(r['fp64']['solves_per_s'], 1),
                            mixed_converged=r['mixed']['all_converged'], fp64_converged=r['fp64']['all_converged'],
                            relres_max=dict(mixed=r['mixed']['relres_max'], fp64=r['fp64']['relres_max'])))
            log('[bench]   uneven %s %s: mixed %.1f fp64 %.1f solves/s' % (kind, r['grid'], r['mixed']['solves_per_s'],
                                                                         r['fp64']['solves_per_s']))
        nsys += wsys
        # concurrent pass over the wave: every geometry's whole sweep on its
        # own stream from its own host thread, started together
        for m in ('mixed', 'fp64'):
            torch.cuda.synchronize()
            _, wall = vsol.sweep_concurrent([(S, freqs, precs, pidx, stream)
                                             for S, precs, pidx, gctx, stream in jobs], method=m)
            torch.cuda.synchronize()
            conc_wall[m] += wall
            log('[bench]   uneven wave of %d concurrent %s: %.1f solves/s (%.3f s)' % (len(jobs), m, wsys / wall, wall))
        for S, precs, pidx, gctx, stream in jobs:
            for P in precs:
                P.close()
            S.G.close()
        del jobs
        torch.cuda.empty_cache()
    conc = {m: nsys / conc_wall[m] for m in conc_wall}
    return dict(seed=seed, geometries=count, nfreq=nfreq, systems=nsys,
                mixed_solves_per_s=nsys / tot['mixed'], fp64_solves_per_s=nsys / tot['fp64'],
                speedup_mixed_over_fp64=tot['fp64'] / tot['mixed'],
                concurrent=dict(wave=wave, mixed_solves_per_s=conc['mixed'], fp64_solves_per_s=conc['fp64'],
                                speedup_mixed_over_fp64=conc['mixed'] / conc['fp64']),
                mixed_vs_fp64_max=worst, per_geometry=per)


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

    log('[bench] FFT-matvec microbench %d^3' % args.cube)
    mv = matvec_bench(ctx, args.cube, args.matvec_reps)
    log('[bench] matvec c128 %.3f ms  c64 %.3f ms' % (mv['c128']['ms'], mv['c64']['ms']))

    log('[bench] bus set-up (%d frequencies)' % args.nfreq)
    S, freqs, precs, t_setup = bus_setup(ctx, args.nfreq)
    nsol = len(freqs) * S.nports
    log('[bench] set-up %.1f s, %d systems per step' % (t_setup, nsol))

    data = bus_rhs(S, freqs)
    ms_mixed, (Zm, stm) = timed_steps(lambda: bus_step(S, freqs, precs, 'mixed', data), args.steps, args.warmup)
    log('[bench] mixed step %.3f s' % ms_mixed)
    fp64 = None
    if not args.no_fp64:
        s64, (Z64, st64) = timed_steps(lambda: bus_step(S, freqs, precs, 'fp64', data), max(1, args.steps // 2),
                                       max(1, args.warmup))
        log('[bench] fp64 step %.3f s' % s64)
        w = 2 * np.pi * np.asarray(freqs)[:, None]
        dm, d64 = np.diagonal(Zm, axis1=1, axis2=2), np.diagonal(Z64, axis1=1, axis2=2)
        fp64 = dict(solves_per_s=nsol / s64, s_per_step=s64,
                    # self resistance / inductance of every port, mixed vs FP64, relative
                    max_rel_err_R_mixed_vs_fp64=float(np.max(np.abs(dm.real - d64.real) / np.abs(d64.real))),
                    max_rel_err_L_mixed_vs_fp64=float(np.max(np.abs(dm.imag - d64.imag) / np.abs(d64.imag))),
                    # whole port impedance matrix, normwise per frequency
                    max_rel_err_Z_mixed_vs_fp64=float(max(np.linalg.norm(Zm[k] - Z64[k]) / np.linalg.norm(Z64[k])
                                                          for k in range(len(freqs)))),
                    L_self_nH_at_fmin=float(d64[0].imag.mean() / w[0, 0] * 1e9),
                    iters_mean=float(st64['iters'].mean()), iters_max=int(st64['iters'].max()),
                    relres_max=float(st64['relres'].max()),
                    # the reference-style stopping residual ||M r|| / ||M b|| (pgmres.m:126)
                    precond_resid_max=float(st64['resid_est'].max()),
                    all_converged=bool(np.all(st64['flag'] == 0)))
    mixed = dict(solves_per_s=nsol / ms_mixed, s_per_step=ms_mixed, outer_iters_mean=float(stm['iters'].mean()),
                 inner_iters_mean=float(stm['inner_iters'].mean()), relres_max=float(stm['relres'].max()),
                 outer_resid_est_max=float(stm['resid_est'].max()),
                 all_converged=bool(np.all(stm['flag'] == 0)))

    z = mv['c64']['passes']['Z']
    nt, e = ZSHAPES32[mv['zshape']['c64']]
    kname = ('k_pass_z_p2<float, 256, %d, true>' % nt) if e == 8 else ('k_pass_z<float, %d, %d, 4>' % (e, nt))
    roof = dict(kernel=kname + ' (z-FFT + Green product + inverse z-FFT), 128^3 grid, complex64',
                bound='hbm', achieved=z['GBs'], peak=HBM_PEAK_GBS, unit='GB/s',
                frac=(z['GBs'] / HBM_PEAK_GBS) if z['GBs'] else None, traffic=recorded_traffic())

    # the bus geometry holds a Krylov workspace sized from most of the free
    # memory: release it before the configuration sweeps
    import torch
    del data
    for P in precs:
        P.close()
    S.G.close()
    torch.cuda.empty_cache()

    configs = None
    if not args.no_configs:
        from voxhenry_amd import structures
        configs = {}
        log('[bench] bar 30x10x10, 10 frequencies')
        configs['bar_30x10x10'] = sweep_config(ctx, structures.bar(30, 10, 10), np.logspace(4, 10, 10))
        log('[bench] spiral, 32 frequencies')
        configs['spiral_5turn'] = sweep_config(ctx, structures.spiral(), np.logspace(6, 11, 32))
        if args.uneven_count > 0:
            log('[bench] uneven batch, %d geometries x %d frequencies' % (args.uneven_count, args.uneven_nfreq))
            configs['uneven_batch'] = uneven_config(ctx, args.uneven_count, args.uneven_nfreq)

    cpu = cpu_mv = None
    if not args.no_cpu:
        log('[bench] CPU baseline (oracle)')
        cpu = cpu_baseline(args.cpu_solves)
        log('[bench] CPU baseline of the 128^3 FFT-block matvec (oracle)')
        cpu_mv = cpu_matvec_baseline(128, 2)

    line = dict(metric=METRIC, value=mixed['solves_per_s'], unit='solves/s', n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=ms_mixed * 1e3, higher_is_better=True, scaling='weak',
                vs_baseline=None, dtype='c64 inner / f64 outer', data='synthetic',
                config=dict(workload='bus: 8 bars 400x4x4 um, 400x64x4 grid, 8 ports x %d freqs 1e5-1e10 Hz '
                                     '= %d systems per step, tol 1e-8 (1e-12 below 1 MHz: %d freqs), true FP64 residual'
                                     % (len(freqs), nsol, int(np.sum(np.asarray(freqs) < 1e6))),
                            parallelism='1 GPU'),
                mixed=mixed, fp64=fp64,
                fp64_speedup=(mixed['solves_per_s'] / fp64['solves_per_s']) if fp64 else None,
                fft_matvec=mv, roofline=roof, cpu_baseline=cpu, cpu_baseline_matvec=cpu_mv, configs=configs,
                setup_s=t_setup)
    print(json.dumps(line))


if __name__ == '__main__':
    main()
