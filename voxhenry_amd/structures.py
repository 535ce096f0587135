"""Voxel structures: the `.vhr` input format and seeded synthetic generators.

A structure is a dict with the fields of the reference's input reader
(src_pre_process/pre_input_file.m:1-133): sigma (L,M,N), lambdaL (L,M,N or
None), freq, dx, ports_P / ports_N (per port an int array (k,4) of 1-based
x, y, z, side with side 1..6 = -x,+x,-y,+y,-z,+z) and well (extra grounded
faces).
"""
import re

import numpy as np

SIDES = {'-x': 1, '+x': 2, '-y': 3, '+y': 4, '-z': 5, '+z': 6}


def read_vhr(path):
    """Parse a `.vhr` file (keywords freq, dx, LMN, superconductor,
    StartVoxelList, N <port> <P|N|other> x y z side), pre_input_file.m:36-114."""
    freq, dx = [], 1e-6
    sigma = lam = None
    supercond = False
    names, P, N, well = [], [], [], []
    with open(path) as fh:
        lines = fh.read().splitlines()
    i = 0
    while i < len(lines):
        tok = [t for t in re.split(r'[\s=,]+', lines[i].strip()) if t]
        i += 1
        if not tok:
            continue
        key = tok[0].lower()
        if key == 'freq':
            freq.extend(float(t) for t in tok[1:])
        elif key == 'dx' and len(tok) > 1:
            dx = float(tok[1])
        elif key == 'lmn' and len(tok) > 3:
            dims = tuple(int(t) for t in tok[1:4])
            sigma = np.zeros(dims)
            lam = np.zeros(dims)
        elif key == 'superconductor':
            supercond = True
        elif key == 'startvoxellist':
            while i < len(lines):
                t = lines[i].split()
                if not t or t[0] != 'V':
                    break
                x, y, z = int(t[1]) - 1, int(t[2]) - 1, int(t[3]) - 1
                sigma[x, y, z] = float(t[4])
                if supercond:
                    lam[x, y, z] = float(t[5])
                i += 1
        elif key == 'n' and len(tok) > 6:
            name = tok[1].lower()
            if name not in names:
                names.append(name)
                P.append([])
                N.append([])
            k = names.index(name)
            row = [int(tok[3]), int(tok[4]), int(tok[5]), SIDES[tok[6]]]
            kind = tok[2].upper()
            (P[k] if kind == 'P' else N[k] if kind == 'N' else well).append(row)
    arr = lambda l: np.array(l, dtype=np.int64).reshape(-1, 4)
    return dict(sigma=sigma, lambdaL=lam if supercond else None, freq=np.array(freq or [1.0]), dx=dx,
                ports_P=[arr(p) for p in P], ports_N=[arr(n) for n in N], well=arr(well))


def _faces(mask_slice_coords, side):
    return np.array([[x + 1, y + 1, z + 1, side] for (x, y, z) in mask_slice_coords], dtype=np.int64)


def end_faces(sigma, axis, at_end, box=None):
    """Faces of the voxels on the first (at_end False, side -axis) or last
    (side +axis) occupied layer along `axis`, restricted to `box` slices."""
    occ = np.abs(sigma) > 0
    if box is not None:
        m = np.zeros_like(occ)
        m[box] = occ[box]
        occ = m
    idx = np.argwhere(occ)
    layer = idx[:, axis].max() if at_end else idx[:, axis].min()
    sel = idx[idx[:, axis] == layer]
    side = 2 * axis + (2 if at_end else 1)
    return _faces(sel, side)


def bar(L, M, N, dx=1e-6, sigma=5.8e7, freq=(1e9,), lam=None):
    """Fully filled bar; one port: P on the +x face of x = L, N on the -x face of x = 1."""
    s = np.full((L, M, N), float(sigma))
    return dict(sigma=s, lambdaL=None if lam is None else np.full((L, M, N), float(lam)), freq=np.asarray(freq, float),
                dx=dx, ports_P=[end_faces(s, 0, True)], ports_N=[end_faces(s, 0, False)],
                well=np.zeros((0, 4), np.int64))


def bus(nbars=8, length=400, section=4, pitch=8, width_grid=64, dx=1e-6, sigma=5.8e7, freq=(1e9,)):
    """Parallel copper bars along x (cross-section section^2, y pitch `pitch`),
    one port per bar on its end faces (the multi-port bus config)."""
    s = np.zeros((length, width_grid, section))
    P, N = [], []
    for i in range(nbars):
        y0 = pitch * i
        s[:, y0:y0 + section, :] = sigma
        box = (slice(None), slice(y0, y0 + section), slice(None))
        P.append(end_faces(s, 0, True, box))
        N.append(end_faces(s, 0, False, box))
    return dict(sigma=s, lambdaL=None, freq=np.asarray(freq, float), dx=dx, ports_P=P, ports_N=N,
                well=np.zeros((0, 4), np.int64))


def spiral(turns=5, outer=100, width=5, height=5, spacing=2, dx=1e-6, sigma=5.8e7, freq=(1e9,)):
    """Square spiral coil traced inward from the outer bottom-left corner; one
    port across its two ends (start: -x faces of the first column, end: -y
    faces of the last segment's bottom row)."""
    s = np.zeros((outer, outer, height))
    pitch = width + spacing
    last = None
    for t in range(turns):
        o = pitch * t
        x0 = 0 if t == 0 else o - pitch
        s[x0:outer - o, o:o + width, :] = sigma                          # bottom, +x
        s[outer - o - width:outer - o, o:outer - o, :] = sigma           # right, +y
        s[o:outer - o, outer - o - width:outer - o, :] = sigma           # top, -x
        s[o:o + width, o + pitch:outer - o, :] = sigma                   # left, -y
        last = (o, o + pitch)
    P = np.array([[1, y + 1, z + 1, 1] for z in range(height) for y in range(width)], dtype=np.int64)
    ox, oy = last
    N = np.array([[ox + x + 1, oy + 1, z + 1, 3] for z in range(height) for x in range(width)], dtype=np.int64)
    return dict(sigma=s, lambdaL=None, freq=np.asarray(freq, float), dx=dx, ports_P=[P], ports_N=[N],
                well=np.zeros((0, 4), np.int64))


def wire(length, diameter, dx=1e-6, sigma=5.8e7, freq=(1e9,)):
    """Round wire along x: voxel centres inside the circle of the given diameter
    (the disc test of VoxHenry_generator_numex2_wire.m:106-122)."""
    d = int(diameter)
    c = (d - 1) / 2.0
    yy, zz = np.meshgrid(np.arange(d), np.arange(d), indexing='ij')
    disc = (yy - c) ** 2 + (zz - c) ** 2 <= (d / 2.0) ** 2
    s = np.zeros((length, d, d))
    s[:, disc] = sigma
    return dict(sigma=s, lambdaL=None, freq=np.asarray(freq, float), dx=dx, ports_P=[end_faces(s, 0, True)],
                ports_N=[end_faces(s, 0, False)], well=np.zeros((0, 4), np.int64))


def loop(radius, section=5, gap=2, dx=1e-6, sigma=5.8e7, freq=(1e9,)):
    """Circular loop in the xy plane (square section `section`), cut by a gap of
    `gap` voxels on the +x axis; the port spans the gap (torus test of
    VoxHenry_generator_numex3_circular_coil.m:97-124 on a square section)."""
    R = float(radius)
    n = int(2 * R + section + 2)
    c = (n - 1) / 2.0
    xx, yy = np.meshgrid(np.arange(n), np.arange(n), indexing='ij')
    r = np.hypot(xx - c, yy - c)
    ring = np.abs(r - R) < section / 2.0
    cut = (np.abs(yy - c) < gap / 2.0 + 0.5) & (xx > c)
    ring &= ~cut
    s = np.zeros((n, n, section))
    s[ring, :] = sigma
    # the two ends of the cut: voxels adjacent to the gap on either side (faces normal to y)
    ys = np.flatnonzero(cut.any(axis=0))
    y_lo, y_hi = ys.min() - 1, ys.max() + 1
    P = [[x + 1, y_hi + 1, z + 1, 3] for x in range(n) for z in range(section) if s[x, y_hi, z] > 0 and x > c]
    Nn = [[x + 1, y_lo + 1, z + 1, 4] for x in range(n) for z in range(section) if s[x, y_lo, z] > 0 and x > c]
    return dict(sigma=s, lambdaL=None, freq=np.asarray(freq, float), dx=dx,
                ports_P=[np.array(P, np.int64)], ports_N=[np.array(Nn, np.int64)], well=np.zeros((0, 4), np.int64))


def strip(length, width, thickness=2, dx=1e-6, sigma=1e7, lam=90e-9, freq=(1e9,)):
    """Two-fluid superconducting strip along x (normal sigma plus London depth
    lam, VoxHenry_executer.m:243-254)."""
    s = np.full((length, width, thickness), float(sigma))
    return dict(sigma=s, lambdaL=np.full(s.shape, float(lam)), freq=np.asarray(freq, float), dx=dx,
                ports_P=[end_faces(s, 0, True)], ports_N=[end_faces(s, 0, False)], well=np.zeros((0, 4), np.int64))


def cube(n, dx=1e-6, sigma=5.8e7):
    """Fully filled n^3 grid without ports (the FFT-matvec microbench)."""
    s = np.full((n, n, n), float(sigma))
    return dict(sigma=s, lambdaL=None, freq=np.array([1e9]), dx=dx, ports_P=[], ports_N=[],
                well=np.zeros((0, 4), np.int64))


def uneven_batch(seed, count=96):
    """Seeded mixture of bars, round wires, loops and two-fluid strips (the
    uneven-batch config); returns a list of (kind, structure)."""
    rng = np.random.default_rng(seed)
    out = []
    kinds = ['bar', 'wire', 'loop', 'strip']
    for i in range(count):
        k = kinds[i % 4]
        if k == 'bar':
            out.append((k, bar(int(rng.integers(16, 513)), int(rng.integers(2, 17)), int(rng.integers(2, 17)))))
        elif k == 'wire':
            out.append((k, wire(int(rng.integers(16, 257)), int(rng.integers(4, 21)))))
        elif k == 'loop':
            out.append((k, loop(int(rng.integers(20, 151)))))
        else:
            out.append((k, strip(int(rng.integers(16, 513)), int(rng.integers(4, 33)))))
    return out
