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


def write_vhr(path, structure, title='synthetic geometry'):
    """Write a structure as a `.vhr` input deck in the layout of
    src_pre_process/pre_output_file.m:44-130: header comment, freq lines of at
    most ten values, dx, LMN, the Superconductor keyword when lambdaL is
    given, the V voxel lines between StartVoxelList / EndVoxelList in
    column-major voxel order, one N line per excitation (P) and ground (N)
    face of every port (portK names), and the grounded faces (N portGND G).
    Numbers are written with 17 significant digits (the reference uses %g),
    so read_vhr gives back the same structure."""
    inv = {v: k for k, v in SIDES.items()}
    sig = np.asarray(structure['sigma'], float)
    lam = structure.get('lambdaL')
    L, M, N = sig.shape
    occ = np.abs(sig) > 0
    if lam is not None:
        lam = np.asarray(lam, float)
        occ |= np.abs(lam) > 0
    g = lambda v: '%.17g' % v
    with open(path, 'w') as fh:
        fh.write('* VoxHenry input file\n* %s\n\n' % title)
        fh.write('* Frequency points (Hz)\n')
        fr = [float(f) for f in np.atleast_1d(structure['freq'])]
        for k in range(0, len(fr), 10):
            fh.write('freq= ' + ' '.join(g(f) for f in fr[k:k + 10]) + '\n')
        fh.write('\n* Voxel size (m)\ndx=%s\n\n' % g(structure['dx']))
        fh.write('* Voxel grid dimension in voxel units; x, y, z\nLMN=%d,%d,%d\n\n' % (L, M, N))
        if lam is not None:
            fh.write('* Specify there are superconductors\nSuperconductor\n\n')
        fh.write('* Voxel list\n* Format is:\n* V <index_x> <index_y> <index_z> <conductivity S/m>\n*\n')
        fh.write('StartVoxelList\n')
        for li in np.flatnonzero(occ.ravel(order='F')):
            x, y, z = np.unravel_index(li, (L, M, N), order='F')
            if lam is None:
                fh.write('V %d %d %d %s\n' % (x + 1, y + 1, z + 1, g(sig[x, y, z])))
            else:
                fh.write('V %d %d %d %s %s\n' % (x + 1, y + 1, z + 1, g(sig[x, y, z]), g(lam[x, y, z])))
        fh.write('EndVoxelList\n\n* Port nodes list\n')
        for p, (PP, NN) in enumerate(zip(structure['ports_P'], structure['ports_N'])):
            fh.write('*\n')
            for kind, rows in (('P', PP), ('N', NN)):
                for r in np.asarray(rows, np.int64).reshape(-1, 4):
                    fh.write('N port%d %s %d %d %d %s\n' % (p + 1, kind, r[0], r[1], r[2], inv[int(r[3])]))
        well = np.asarray(structure.get('well', np.zeros((0, 4))), np.int64).reshape(-1, 4)
        if len(well):
            fh.write('\n* Grounded nodes list\n*\n')
            for r in well:
                fh.write('N portGND G %d %d %d %s\n' % (r[0], r[1], r[2], inv[int(r[3])]))


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
    return loops([radius], section, gap, dx, sigma, freq)


def loops(radii, section=5, gap=2, dx=1e-6, sigma=5.8e7, freq=(1e9,)):
    """Concentric circular loops in the xy plane (square section `section`,
    radii must differ by more than `section`), all cut by one gap of `gap`
    voxels on the +x axis; loop i has its own port across the gap (P on the
    +y side of the cut, N on the -y side), so k loops give k ports."""
    radii = [float(r) for r in radii]
    R = max(radii)
    n = int(2 * R + section + 2)
    c = (n - 1) / 2.0
    xx, yy = np.meshgrid(np.arange(n), np.arange(n), indexing='ij')
    r = np.hypot(xx - c, yy - c)
    cut = (np.abs(yy - c) < gap / 2.0 + 0.5) & (xx > c)
    ys = np.flatnonzero(cut.any(axis=0))
    y_lo, y_hi = ys.min() - 1, ys.max() + 1
    s = np.zeros((n, n, section))
    P, Nn = [], []
    for Ri in radii:
        ring = (np.abs(r - Ri) < section / 2.0) & ~cut
        s[ring, :] = sigma
        # the two ends of this ring at the cut: its voxels adjacent to the gap (faces normal to y)
        P.append(np.array([[x + 1, y_hi + 1, z + 1, 3] for x in range(n) for z in range(section)
                           if ring[x, y_hi] and x > c], np.int64).reshape(-1, 4))
        Nn.append(np.array([[x + 1, y_lo + 1, z + 1, 4] for x in range(n) for z in range(section)
                            if ring[x, y_lo] and x > c], np.int64).reshape(-1, 4))
    return dict(sigma=s, lambdaL=None, freq=np.asarray(freq, float), dx=dx,
                ports_P=P, ports_N=Nn, well=np.zeros((0, 4), np.int64))


def parallel(kind, count, length, section, dx=1e-6, freq=(1e9,), **kw):
    """`count` parallel conductors along x, one port each (P on the +x end
    faces, N on the -x end faces of that conductor): kind 'bar' (square
    section x section, y pitch 2*section), 'wire' (round, diameter section,
    pitch section + 2) or 'strip' (two-fluid, width section, thickness
    kw.get('thickness', 2), pitch section + 4)."""
    if kind == 'bar':
        one, pitch = bar(length, section, section, dx=dx, freq=freq), 2 * section
    elif kind == 'wire':
        one, pitch = wire(length, section, dx=dx, freq=freq), section + 2
    elif kind == 'strip':
        one, pitch = strip(length, section, kw.get('thickness', 2), dx=dx, freq=freq), section + 4
    else:
        raise ValueError(kind)
    s1 = one['sigma']
    L, M1, N = s1.shape
    M = pitch * (count - 1) + M1
    s = np.zeros((L, M, N))
    lam = None if one['lambdaL'] is None else np.zeros((L, M, N))
    P, Nn = [], []
    for i in range(count):
        y0 = pitch * i
        s[:, y0:y0 + M1, :] = s1
        if lam is not None:
            lam[:, y0:y0 + M1, :] = one['lambdaL']
        box = (slice(None), slice(y0, y0 + M1), slice(None))
        P.append(end_faces(s, 0, True, box))
        Nn.append(end_faces(s, 0, False, box))
    return dict(sigma=s, lambdaL=lam, freq=np.asarray(freq, float), dx=dx, ports_P=P, ports_N=Nn,
                well=np.zeros((0, 4), np.int64))


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


def split_ports(structure, k):
    """The structure's one port split into k terminal pairs: its excitation
    faces and its ground faces, each ordered by (z, y, x), are cut into k
    contiguous bands, and band i of both forms port i (a multi-terminal
    conductor; every other terminal is held at zero potential while one is
    driven, as for any port's nodes, VoxHenry_executer.m:165-193).  k is
    limited by the number of faces at either end."""
    P, N = structure['ports_P'][0], structure['ports_N'][0]
    k = max(1, min(int(k), len(P), len(N)))
    order = lambda a: a[np.lexsort((a[:, 0], a[:, 1], a[:, 2]))]
    out = dict(structure)
    out['ports_P'] = [np.ascontiguousarray(b) for b in np.array_split(order(P), k)]
    out['ports_N'] = [np.ascontiguousarray(b) for b in np.array_split(order(N), k)]
    return out


def uneven_batch(seed, count=96):
    """Seeded mixture of the four conductor families of the uneven-batch
    workload, cycled bar, wire, loop, strip; every geometry has 1-4 ports,
    its end (or gap) faces split into that many terminal pairs
    (split_ports):

    * copper bars, length 16-512 voxels, square section 2-16;
    * round copper wires, length 16-256, diameter 4-20;
    * circular copper loops, radius 20-150, 5 x 5 section;
    * two-fluid strips (sigma 1e7 S/m, London depth 90 nm), length 16-512,
      width 4-32, thickness 2.

    Returns a list of (kind, structure)."""
    rng = np.random.default_rng(seed)
    out = []
    kinds = ['bar', 'wire', 'loop', 'strip']
    for i in range(count):
        k = kinds[i % 4]
        nports = int(rng.integers(1, 5))
        if k == 'bar':
            L, sec = int(rng.integers(16, 513)), int(rng.integers(2, 17))
            st = bar(L, sec, sec)
        elif k == 'wire':
            st = wire(int(rng.integers(16, 257)), int(rng.integers(4, 21)))
        elif k == 'loop':
            st = loop(int(rng.integers(20, 151)))
        else:
            st = strip(int(rng.integers(16, 513)), int(rng.integers(4, 33)))
        out.append((k, split_ports(st, nports)))
    return out
